"""The reference's protocol phases on a pcg.Context (thread_main.cpp:84-142).

A trajectory starts either with pulling then release (afm type "pull, then
release": velocity tips for pull-release time, then AFM off) or with an
equilibration phase of equil_time without AFM; after equilibration, AFM is
switched on ("perform") with tips placed at the positions of that moment
(state::setup_afm, state.cpp:548-568) and the run goes on to the total
time. A phase lasts while t < until, with t accumulated step by step as the
library accumulates it.

The context must have AFM enabled in its parameters (afm_on) and its model
must reserve two tip slots (pcg_model n_tips + n_ftips >= 2), for example
with model.add_pulling.

Without AFM after equilibration the reference hands the trajectory to the
simulation-box protocol (squeezing, force-minimum search, oscillations,
thread_main.cpp:130-135 and 144-330). Here that protocol runs on the device
inside every step once pcg_box_start has been called (the box itself comes
from Context.box_setup before the trajectory starts); the host only counts
steps to the total time.

Deviation: after pull-then-release and equilibration the reference turns
AFM back on and appends a resting tip to the two old ones (setup_afm only
adds, state.cpp:555); here that last free phase runs without tips.
"""
from __future__ import annotations

import numpy as np


def setup_afm_tips(pos: np.ndarray, chain: np.ndarray, kind: str, magnitude: float) -> dict:
    """Tips of state::setup_afm at positions pos: the first residue of chain 0
    held by a resting velocity tip, the last one pulled along the direction
    from the first to the last, at speed `magnitude` ("const velocity") or
    by a force of that size ("const force")."""
    i0 = 0
    i1 = int(np.nonzero(np.asarray(chain) == chain[0])[0][-1])
    d = pos[i1] - pos[i0]
    u = d * (1.0 / np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))  # unit(): v * norm_inv(v)
    if kind == "const velocity":
        return {"tip_res": np.array([i0, i1], np.int32),
                "tip_orig": np.array([pos[i0], pos[i1]], np.float64),
                "tip_vel": np.array([np.zeros(3), magnitude * u], np.float64)}
    if kind == "const force":
        return {"tip_res": np.array([i0], np.int32),
                "tip_orig": np.array([pos[i0]], np.float64),
                "tip_vel": np.zeros((1, 3), np.float64),
                "ftip_res": np.array([i1], np.int32),
                "ftip_force": (magnitude * u)[None].astype(np.float64)}
    raise ValueError(f"unknown AFM tip type {kind!r}")


def steps_until(t: float, until: float, dt: float) -> int:
    """Steps the loop `while t < until: advance` makes from time t."""
    k = 0
    while t < until:
        t += dt
        k += 1
    return k


def run_phases(ctx, chain, dt: float, total_time: float, equil_time: float = 0.0,
               afm: str | None = None, magnitude: float = 0.0,
               pull_release_time: float | None = None, pull_speed: float = 0.0,
               box: bool = False) -> list:
    """Runs one trajectory through the reference's phases (traj_init_step,
    pull_release_step, equil_step): returns [(phase, steps)].

    afm: None (equilibrate, then run freely), "const velocity" or "const
    force" (equilibrate, then pull with tips placed at that moment), or
    "pull, then release" (pull at pull_speed for pull_release_time, then
    equilibrate without AFM, then run freely).

    box: the context has a simulation box (Context.box_setup); without AFM
    pulling the run after equilibration is the box protocol ("BOX": the
    device's phases from SQUEEZING to FREEFORM, Context.box_state tells
    which one it is in)."""
    phases = []

    def advance(name, until):
        t, _ = ctx.time()
        k = steps_until(t, until, dt)
        if k:
            ctx.step(k)
        phases.append((name, k))

    if afm == "pull, then release":
        ctx.set_tips(setup_afm_tips(ctx.r(), chain, "const velocity", pull_speed))
        advance("PULL_RELEASE", min(total_time, pull_release_time))
        ctx.set_tips({})
        t, _ = ctx.time()
        advance("EQUIL", t + equil_time)
    else:
        ctx.set_tips({})
        advance("EQUIL", min(total_time, equil_time))
    if afm in ("const velocity", "const force"):
        ctx.set_tips(setup_afm_tips(ctx.r(), chain, afm, magnitude))
    elif box:
        ctx.box_start()
        advance("BOX", total_time)
        return phases
    advance("FREEFORM", total_time)
    return phases
