"""Walls and the simulation-box protocol of the oracle (CPU).

Pinning:
  - the wall kernels (wall::harmonic::eval_free / eval_connected,
    wall::lj::sift_free / eval_connected / process_candidates, and
    wall::log_forces) against the reference's own sources compiled into
    oracle/_ref, call by call on the oracle's own state;
  - the harmonic connections made at set-up (state.cpp:653-670, not
    compilable alone) against a brute-force selection;
  - the box protocol (thread_main.cpp:144-330, not compilable alone) against
    closed forms: the squeezed box side as an arithmetic series of the
    accelerating wall speed, the oscillation as a sum of sines, phase
    durations in steps, the shear of the z normals, the PBC cell following the
    box extent.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import ref_ffi as R
from oracle_ffi import Oracle
from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd.protocol import steps_until

D, I, U8 = C.c_double, C.c_int32, C.c_uint8
A = P.ANGSTROM
needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (reference tree absent)")


def ptr(a, ct):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def protein(n=90, seed=5):
    m = M.synthetic_protein(n, M.job_seed(seed))
    rng = np.random.default_rng(seed)
    m["pos"] = m["pos"] + rng.normal(scale=0.3 * A, size=m["pos"].shape)
    return m


def walled(axes, n, limits=None, cycle_dur=0.0, **kw):
    prm = P.default_params()
    prm.update(qa_on=0, pid_on=0, afm_on=0)
    prm.update(kw)
    pbc = P.set_box_walls(prm, axes, n, limits=limits)
    P.set_lj_wall_timing(prm, 1.5, cycle_dur)
    return prm, pbc


def bbox_box(m, gap_A, pbc, **kw):
    """The model's bounding box widened by gap_A plus a different few
    hundredths of an angstrom on each face, so that no two residues sit at
    the same distance from their walls (the reference sorts LJ candidates
    with an unstable sort)."""
    lo = m["pos"].min(axis=0) - (gap_A + np.array([0.00, 0.03, 0.06])) * A
    hi = m["pos"].max(axis=0) + (gap_A + np.array([0.015, 0.045, 0.075])) * A
    return P.box_params(lo, hi, pbc, **kw)


def planes_now(prm, o):
    """The parameters with the oracle's current planes (after box moves)."""
    st = o.box_state()
    q = P.clone(prm)
    q["wall_origin"] = st[15:33].reshape(6, 3).tolist()
    q["wall_normal"] = st[33:51].reshape(6, 3).tolist()
    return q


def wall_F(o):
    return o.box_state()[51:69].reshape(6, 3)


def close(a, b, tol=1e-13):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-12)
    assert np.abs(a - b).max() <= tol * scale, (np.abs(a - b).max(), scale)


# ------------------------------------------------------------ set-up ----
def test_harmonic_connections_follow_the_setup():
    """state::setup_walls' harmonic connections: after each axis, every
    harmonic plane so far takes up to its limit of the unconnected residues
    nearest to it (ties by index), the planes still through the origin; the
    bead offset is the residue's projection on the plane."""
    m = protein(120, seed=3)
    prm, _ = walled({"x": "harmonic", "y": "solid", "z": "harmonic"}, m["n"],
                    limits={"harmonic": 7}, hw_on=1)
    o = Oracle(prm, m)
    got = o.wall_conns(0)
    pos = m["pos"]
    conn = np.zeros(m["n"], bool)
    want = []
    planes = [w for w in range(prm["n_walls"]) if prm["wall_kind"][w] == 1]
    for w_end in planes:
        if prm["wall_side"][w_end] % 2 == 0:
            continue
        for w in [q for q in planes if q <= w_end]:
            nrm = np.array(prm["wall_normal"][w])
            d = np.abs(pos @ nrm)
            order = sorted(range(m["n"]), key=lambda i: (d[i], i))
            made = 0
            for i in order:
                if made == prm["wall_limit"][w]:
                    break
                if conn[i]:
                    continue
                want.append((w, i, pos[i] - (pos[i] @ nrm) * nrm))
                conn[i] = True
                made += 1
    # x planes take 7 + 7 on the x pass and 7 + 7 more on the z pass
    assert len(want) == 7 * 6
    assert [(int(w), int(r)) for w, r in zip(got["wall"], got["res"])] == [(w, i) for w, i, _ in want]
    np.testing.assert_allclose(got["off"], np.array([x for _, _, x in want]), rtol=0, atol=1e-15)
    assert np.all(got["sat"] == 0.0)  # the natural distance: the saturation 0 passed in its place


# ------------------------------------------------------ wall kernels ----
@needs_ref
def test_harmonic_wall_terms_match_reference():
    m = protein()
    prm, pbc = walled({"x": "harmonic", "y": "harmonic", "z": "solid"}, m["n"],
                      limits={"harmonic": 5}, hw_on=1)
    o = Oracle(prm, m)
    o.box_setup(bbox_box(m, 1.5, pbc))
    q = planes_now(prm, o)
    ps = R.params_struct(q)
    L = R.lib()
    r = np.ascontiguousarray(o.r())
    n = m["n"]
    conns = o.wall_conns(0)
    is_conn = np.zeros(n, np.uint8)
    is_conn[conns["res"]] = 1

    F1, V1 = o.eval_term("hw_free")
    wf1 = wall_F(o)
    F2, wf2 = np.zeros((n, 3)), np.zeros((6, 3))
    V2 = L.ref_hw_free(n, ptr(r, D), C.byref(ps), ptr(is_conn, U8), ptr(F2, D), ptr(wf2, D))
    assert (np.abs(F2).max(axis=1) > 0).sum() >= 3 and V1 == 0.0 and V2 == 0.0
    close(F1, F2)
    close(wf1, wf2)

    F1, V1 = o.eval_term("hw_conn")
    wf1 = wall_F(o)
    k = len(conns["res"])
    res = np.ascontiguousarray(conns["res"], np.int32)
    wl = np.ascontiguousarray(conns["wall"], np.int32)
    off = np.ascontiguousarray(conns["off"])
    nat = np.ascontiguousarray(conns["sat"])
    F2, wf2 = np.zeros((n, 3)), np.zeros((6, 3))
    V2 = L.ref_hw_conn(n, ptr(r, D), C.byref(ps), k, ptr(res, I), ptr(wl, I), ptr(off, D),
                       ptr(nat, D), ptr(F2, D), ptr(wf2, D))
    assert k == 5 * (2 + 4) and V2 > 0  # the y pass connects the x planes again
    close(F1, F2)
    close(wf1, wf2)
    assert abs(V1 - V2) <= 1e-13 * abs(V2)


def _lj_ref_cycle(o, prm, L, n):
    """One reference call per stage on the oracle's state; returns what the
    reference computes."""
    ps = R.params_struct(planes_now(prm, o))
    r = np.ascontiguousarray(o.r())
    pend = o.ljw_pending()
    F2, wf2 = np.zeros((n, 3)), np.zeros((6, 3))
    cap = 4 * n
    cr, cw, cd = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap)
    nc = I(0)
    V2 = L.ref_ljw_sift(n, ptr(r, D), C.byref(ps), ptr(np.ascontiguousarray(pend["is_conn"]), U8),
                        ptr(np.ascontiguousarray(pend["num_conns"]), I), ptr(F2, D), ptr(wf2, D),
                        cap, ptr(cr, I), ptr(cw, I), ptr(cd, D), C.byref(nc))
    return ps, r, F2, wf2, V2, (cr[:nc.value], cw[:nc.value], cd[:nc.value])


@needs_ref
def test_lj_wall_cycle_matches_reference():
    """sift_free -> process_candidates -> eval_connected -> process_candidates
    on the same state: forces, energies, reaction forces, candidates,
    saturations, releases and the connection slots equal the reference's."""
    m = protein()
    n = m["n"]
    prm, pbc = walled({"x": "lj", "y": "solid", "z": "lj"}, n, limits={"lj": 3}, cycle_dur=0.02,
                      ljw_on=1)
    o = Oracle(prm, m)
    o.box_setup(bbox_box(m, 1.0, pbc))
    L = R.lib()

    # sift_free
    ps, r, F2, wf2, V2, cands = _lj_ref_cycle(o, prm, L, n)
    F1, V1 = o.eval_term("ljw_free")
    close(F1, F2)
    close(wall_F(o), wf2)
    assert abs(V1 - V2) <= 1e-13 * abs(V2) and V2 > 0
    pend = o.ljw_pending()
    assert len(cands[0]) >= 6
    np.testing.assert_array_equal(pend["cand_res"], cands[0])
    np.testing.assert_array_equal(pend["cand_wall"], cands[1])
    np.testing.assert_array_equal(pend["cand_d"], cands[2])

    def ref_process(o):
        s = o.ljw_slots()
        pend = o.ljw_pending()
        k = len(s["res"])
        cap = k + len(pend["cand_res"]) + 1
        outs = [np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros((cap, 3)),
                np.zeros(cap), np.zeros(cap, np.uint8)]
        conn = np.ascontiguousarray(pend["is_conn"])
        num = np.ascontiguousarray(pend["num_conns"])
        ns = I(0)
        ar = lambda x, t: np.ascontiguousarray(x, t)  # noqa: E731
        L.ref_ljw_process(n, ptr(np.ascontiguousarray(o.r()), D), C.byref(R.params_struct(planes_now(prm, o))),
                          k, ptr(ar(s["res"], np.int32), I), ptr(ar(s["wall"], np.int32), I),
                          ptr(ar(s["off"], np.float64), D), ptr(ar(s["sat"], np.float64), D),
                          ptr(ar(s["vac"], np.uint8), U8), len(pend["removed"]),
                          ptr(ar(pend["removed"], np.int32), I), len(pend["cand_res"]),
                          ptr(ar(pend["cand_res"], np.int32), I), ptr(ar(pend["cand_wall"], np.int32), I),
                          ptr(ar(pend["cand_d"], np.float64), D), ptr(conn, U8), ptr(num, I), cap,
                          C.byref(ns), ptr(outs[0], I), ptr(outs[1], I), ptr(outs[2], D),
                          ptr(outs[3], D), ptr(outs[4], U8))
        k2 = ns.value
        return [x[:k2] for x in outs], conn, num

    def check_process(o):
        (res, wl, off, sat, vac), conn, num = ref_process(o)
        o.ljw_process()
        s = o.ljw_slots()
        pend = o.ljw_pending()
        np.testing.assert_array_equal(s["res"], res)
        np.testing.assert_array_equal(s["wall"], wl)
        np.testing.assert_array_equal(s["vac"], vac)
        np.testing.assert_array_equal(s["off"], off)
        np.testing.assert_array_equal(s["sat"], sat)
        np.testing.assert_array_equal(pend["is_conn"], conn)
        np.testing.assert_array_equal(pend["num_conns"], num)
        return s

    s = check_process(o)
    assert (s["vac"] == 0).sum() == 12  # four LJ planes at their limit of 3

    # eval_connected after pulling some connected residues off their beads
    r = o.r()
    for j, i in enumerate(s["res"][:4]):
        r[i] += (5.0 + 0.3 * j) * A * np.array([1.0, -1.0, 0.5])
    o.set_r(r)
    s = o.ljw_slots()
    k = len(s["res"])
    sat = np.ascontiguousarray(s["sat"])
    F2, wf2 = np.zeros((n, 3)), np.zeros((6, 3))
    rem = np.zeros(k + 1, np.int32)
    nrem = I(0)
    ps = R.params_struct(planes_now(prm, o))
    V2 = L.ref_ljw_conn(n, ptr(np.ascontiguousarray(o.r()), D), C.byref(ps), k,
                        ptr(np.ascontiguousarray(s["res"]), I), ptr(np.ascontiguousarray(s["wall"]), I),
                        ptr(np.ascontiguousarray(s["off"]), D), ptr(sat, D),
                        ptr(np.ascontiguousarray(s["vac"]), U8), ptr(F2, D), ptr(wf2, D), k + 1,
                        ptr(rem, I), C.byref(nrem))
    F1, V1 = o.eval_term("ljw_conn")
    close(F1, F2)
    close(wall_F(o), wf2)
    assert V1 == V2 == 0.0  # new connections carry saturation 0 on their first evaluation
    np.testing.assert_array_equal(o.ljw_slots()["sat"], sat)
    np.testing.assert_array_equal(o.ljw_pending()["removed"], rem[:nrem.value])
    assert 0 < (sat > 0).sum() < k  # some saturate, the pulled ones do not

    # a second evaluation weighs by the new saturations, then the releases go
    for _ in range(3):
        s = o.ljw_slots()
        F1, V1 = o.eval_term("ljw_conn")
    assert V1 != 0.0
    s = check_process(o)


@needs_ref
def test_log_forces_match_reference():
    """wall::log_forces over steps of a walled trajectory: the averaged
    reaction forces, the work and the averaged z force (with an LJ axis, whose
    -axis plane the reference logs twice)."""
    m = protein(60, seed=8)
    n = m["n"]
    prm, pbc = walled({"x": "lj", "y": "solid", "z": "solid"}, n, limits={"lj": 2},
                      cycle_dur=0.05, ljw_on=1)
    o = Oracle(prm, m, seed=3)
    window = 7 * prm["dt"]
    o.box_setup(bbox_box(m, 1.2, pbc, avg_forces_over=window))
    st0 = o.box_state()
    ts, Fs = [], []
    for _ in range(40):
        ts.append(o.t)
        o.step(1)
        Fs.append(wall_F(o).copy())
    st = o.box_state()
    present = np.ones(6, np.int32)
    extra = np.array([0], np.int32)  # the -x LJ plane
    out = [np.zeros(18), np.zeros(6), np.zeros(6), D(0), D(0)]
    R.lib().ref_log_forces(len(ts), ptr(np.array(ts), D), ptr(np.ascontiguousarray(np.array(Fs)), D),
                           ptr(np.ascontiguousarray(st0[33:51]), D),
                           ptr(np.ascontiguousarray(st0[69:87]), D), ptr(present, I), 1,
                           ptr(extra, I), window, ptr(out[0], D), ptr(out[1], D), ptr(out[2], D),
                           C.byref(out[3]), C.byref(out[4]))
    assert np.abs(np.array(Fs)).max() > 0
    np.testing.assert_array_equal(st[93:99], out[1])
    close(st[99:117], out[0], 1e-14)
    close(st[87:93], out[2], 1e-14)
    assert st[13] == out[4].value == 1.0
    assert abs(st[14] - out[3].value) <= 1e-14 * max(abs(out[3].value), 1e-12)


# ------------------------------------------------------ box protocol ----
def _run_until(o, dt, total, stop):
    """Steps one at a time until stop(state) or the total time; returns the
    (phase, t, state) of every step start."""
    trace = []
    while o.t < total:
        st = o.box_state()
        trace.append((int(st[0]), o.t, st))
        if stop(st):
            break
        o.step(1)
    return trace


def test_squeezing_follows_the_accelerating_walls():
    """Squeezing a cubic box with solid walls: every step both faces of every
    axis move in by shift_k = vel_k dt / 2, vel_k = min(k dt / T, 1) v
    (T = accel_dist / v), so the side after K steps is the start side less
    the closed sum; the phase turns to REST_AFTER_SQUEEZING at the first step
    the volume is at most n / density, with until = t + rest."""
    m = protein(40, seed=2)
    n = m["n"]
    prm, pbc = walled({"x": "solid", "y": "solid", "z": "solid"}, n)
    dt = prm["dt"]
    half = 40.0 * A
    v = 0.5 * A / P.TAU
    T = 0.3 * A / v  # 60 steps of acceleration
    density = n / (2 * 33.0 * A) ** 3
    box = P.box_params([-half] * 3, [half] * 3, pbc, total_time=1e9, squeezing=1,
                       target_density=density, vel_above_2V=v, target_vel=v, accel_dist=0.3 * A,
                       rest_period=25 * dt)
    o = Oracle(prm, m, seed=1)
    o.set_noise_scale(0.0)
    o.box_setup(box)
    o.box_start()
    trace = _run_until(o, dt, 1e9, lambda st: int(st[0]) != 1)
    K = len(trace) - 2  # wall moves made (the step that sees the target volume moves none)
    last = trace[-1][2]
    assert int(last[0]) == 2
    # closed form of sum_{k<K} 0.5 dt min(k dt / T, 1) v
    ka = math.ceil(T / dt)  # first step at full speed
    k1 = min(K, ka)
    s = 0.5 * dt * v * (dt / T * k1 * (k1 - 1) / 2.0 + max(K - ka, 0))
    side = 2 * half - 2 * s
    ext = last[6:9] - last[3:6]
    np.testing.assert_allclose(ext, side, rtol=1e-11)
    assert ext[0] ** 3 <= n / density < (ext[0] + 2 * 0.5 * dt * v) ** 3
    # walls on the faces, the box centred
    org = last[15:33].reshape(6, 3)
    np.testing.assert_allclose(org[0], [last[3], 0, 0], atol=1e-13)
    np.testing.assert_allclose(org[5], [0, 0, last[8]], atol=1e-13)
    # the rest ends rest_period later, then the phase is over (no force-min)
    assert abs(last[2] - (trace[-2][1] + 25 * dt)) < 1e-12  # set by the step before
    o.step(steps_until(o.t, last[2], dt) + 1)
    assert int(o.box_state()[0]) == 9


def _osc_setup(shear=False, relative=False):
    m = protein(30, seed=4)
    n = m["n"]
    prm, pbc = walled({"x": "solid", "y": "solid", "z": "solid"}, n)
    dt = prm["dt"]
    omega = 2 * math.pi / (160 * dt)
    box = P.box_params([-30 * A] * 3, [30 * A] * 3, pbc, total_time=1e9, force_min=1,
                       oscillations=1, avg_forces_over=0.0, rest_period=10 * dt,
                       target_vel=0.2 * A / P.TAU, accel_dist=0.01 * A, angular_freq=omega,
                       num_of_cycles=1, amp_abs=1.0 * A, amp_relative=int(relative),
                       amp_rel=0.02, osc_shear=int(shear))
    o = Oracle(prm, m, seed=1)
    o.set_noise_scale(0.0)
    o.box_setup(box)
    o.box_start()
    return o, prm, box


def test_phases_and_oscillation_closed_forms():
    """Far walls (no force): force-min finds |<F_z>| = 0 below its threshold
    after one averaged step, rests, the walls close in until half the amplitude
    is covered, rest, then oscillate for (2 pi cycles + pi/2) / omega with the
    z extent following 2 sum_k 0.5 A w sin(w k dt) dt (a closed sum of
    sines), then rest and FREEFORM."""
    o, prm, box = _osc_setup()
    dt = prm["dt"]
    trace = _run_until(o, dt, 1e9, lambda st: int(st[0]) == 9)
    phases = [p for p, _, _ in trace]
    order = [p for i, p in enumerate(phases) if i == 0 or phases[i - 1] != p]
    assert order == [1, 3, 4, 5, 6, 7, 8, 9]  # squeezing off passes straight to force-min
    count = {p: phases.count(p) for p in set(phases)}
    # a step starts in a phase once the step before entered it, so each phase
    # shows as many step starts as it makes steps
    t_in = {}
    for i in range(1, len(trace)):
        if phases[i] != phases[i - 1]:
            t_in[phases[i]] = trace[i - 1][1]  # the step whose phase logic entered it
    assert count[1] == 1 and count[3] == 1  # one step to get the first average
    for p in (4, 6, 8):  # rests of rest_period
        assert count[p] == steps_until(t_in[p], t_in[p] + box["rest_period"], dt)
    # the max-amplitude steps: displacement sums shift_k = min(k dt / T, 1) v dt
    A_ = box["amp_abs"]
    v, T = box["target_vel"], box["accel_dist"] / box["target_vel"]
    disp, K = 0.0, 0
    while disp <= 0.5 * A_:
        disp += min(K * dt / T, 1.0) * v * dt
        K += 1
    assert count[5] == K
    st5 = [s for p, _, s in trace if p == 6][0]
    assert abs(st5[11] - disp) < 1e-15 and abs(st5[9] - A_) < 1e-15
    # oscillations: K steps of sin
    w = box["angular_freq"]
    st_a = [s for p, _, s in trace if p == 7][0]
    st_b = [s for p, _, s in trace if p == 8][0]
    Ko = count[7]
    assert abs(st_a[2] - (st_a[1] + (2 * math.pi + math.pi / 2) / w)) < 1e-12
    assert Ko == steps_until(st_a[1], st_a[2], dt)
    closed = 0.5 * A_ * w * dt * math.sin(Ko * w * dt / 2) * math.sin((Ko - 1) * w * dt / 2) / math.sin(w * dt / 2)
    ext_a, ext_b = st_a[8] - st_a[5], st_b[8] - st_b[5]
    assert abs((ext_b - ext_a) - 2 * closed) < 1e-9 * A_
    assert abs(st_a[10] - w) == 0.0
    # the x and y faces never move
    np.testing.assert_array_equal(st_b[3:5], st_a[3:5])


def test_shear_oscillations_tilt_the_z_normals():
    """Shear (relative amplitude of the x extent): the z planes' normals gain
    every shift in their x component, their origins stay; the box keeps its
    size."""
    o, prm, box = _osc_setup(shear=True, relative=True)
    dt = prm["dt"]
    trace = _run_until(o, dt, 1e9, lambda st: int(st[0]) == 9)
    first = [s for p, _, s in trace if p == 5][0]
    last = trace[-1][2]
    assert last[12] == 1.0
    np.testing.assert_allclose(last[9], box["amp_rel"] * (first[6] - first[3]), rtol=1e-15)
    nz = last[33:51].reshape(6, 3)
    total = 0.0
    for (p, _, s), (_, _, s2) in zip(trace[:-1], trace[1:]):
        if p in (5, 7):
            total += s2[33 + 3 * 4] - s[33 + 3 * 4]
    assert abs(nz[4, 0] - total) < 1e-15 and abs(nz[5, 0] - total) < 1e-15
    assert abs(nz[4, 0]) > 0
    np.testing.assert_array_equal(last[3:9], first[3:9])
    np.testing.assert_array_equal(last[15:33], first[15:33])


def test_periodic_squeezing_shrinks_the_cell():
    """Periodic axes: the PBC cell is the box extent after every move, and
    the neighbour list is rebuilt when the cell has shrunk by more than the
    pad (nl::verify's box displacement)."""
    m = protein(40, seed=6)
    n = m["n"]
    prm, pbc = walled({"x": "periodic", "y": "periodic", "z": "periodic"}, n, dh_on=1)
    assert pbc == [1, 1, 1] and prm["n_walls"] == 0
    dt = prm["dt"]
    v = 2.0 * A / P.TAU
    box = P.box_params([-45 * A] * 3, [45 * A] * 3, pbc, total_time=1e9, squeezing=1,
                       target_density=n / (70 * A) ** 3, vel_above_2V=v, target_vel=v,
                       accel_dist=0.5 * A, rest_period=5 * dt)
    o = Oracle(prm, m, seed=1)
    o.set_noise_scale(0.0)
    o.box_setup(box)
    st = o.box_state()
    np.testing.assert_array_equal(st[117:120], [90 * A] * 3)
    o.box_start()
    rebuilds0 = o.nl_rebuilds()
    trace = _run_until(o, dt, 1e9, lambda s: int(s[0]) != 1)
    for _, _, s in trace:
        np.testing.assert_array_equal(s[117:120], s[6:9] - s[3:6])
    shrink = 90 * A - (trace[-1][2][6] - trace[-1][2][3])
    assert shrink > 10 * A
    # every rebuild follows a cell change of more than the pad (3 l1 per step)
    assert o.nl_rebuilds() - rebuilds0 >= int(3 * shrink / prm["nl_pad"]) - 1
