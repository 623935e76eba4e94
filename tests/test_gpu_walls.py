"""GPU parity of the walls and the simulation-box protocol (HIP path through
the C ABI against the FP64 oracle, whose wall kernels are pinned against the
reference's own sources and whose box protocol against closed forms in
tests/test_walls.py).

* every wall term (solid, harmonic free and connected, LJ free and
  connected) on identical positions: forces, energies and the planes'
  reaction forces, FP64 to 1e-9 of the largest component, mixed (FP32 force
  accumulation) to 1e-5;
* LJ candidate processing: the same connections, bead offsets and
  saturations;
* whole noise-free FP64 trajectories through the box protocol (squeezing,
  force-minimum search, oscillations, shear; periodic squeezing of the PBC
  cell; LJ walls) with the step captured as a hipGraph or not: positions to
  1e-8, the box, planes, averaged forces and work, phases step for step;
* a checkpoint taken mid-protocol resumes the same trajectory.
"""
import math

import numpy as np
import pytest

from oracle_ffi import Oracle
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G
from test_walls import A, bbox_box, protein, walled

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-9
MIXED_F_TOL = 1e-5


def pair(prm, m, box=None, precision=G.FP64, seed=17, graph=False):
    o = Oracle(prm, m, seed=seed)
    o.set_noise_scale(0.0)
    c = G.Context(prm, m, precision=precision, noise=G.RAN2, seed=seed)
    c.set_noise_scale(0.0)
    c.use_graph(graph)
    if box is not None:
        o.box_setup(box)
        c.box_setup(box)
    return o, c


def same_positions(o, c):
    c.set(0, o.r())
    c.set_time(o.t)


def box_close(so, sg, tol=1e-10):
    """The pcg_box_state arrays of the oracle and the library."""
    assert int(so[0]) == int(sg[0]), ("phase", so[0], sg[0])
    scale = max(np.abs(so[3:9]).max(), 1.0)
    assert np.abs(so[1:13] - sg[1:13]).max() <= tol * max(scale, np.abs(so[1:3]).max()), (so[1:13], sg[1:13])
    assert so[13] == sg[13]
    assert abs(so[14] - sg[14]) <= 1e-8 * max(abs(so[14]), 1e-3)
    np.testing.assert_allclose(sg[15:51], so[15:51], rtol=0, atol=tol)  # origins, normals
    fs = max(np.abs(so[51:69]).max(), 1e-6)
    assert np.abs(sg[51:69] - so[51:69]).max() <= 1e-8 * fs  # reaction forces
    np.testing.assert_allclose(sg[69:87], so[69:87], rtol=0, atol=tol)  # shifts
    ws = max(np.abs(so[87:93]).max(), 1e-6)
    assert np.abs(sg[87:93] - so[87:93]).max() <= 1e-8 * ws  # work
    np.testing.assert_array_equal(sg[93:99], so[93:99])
    avs = max(np.abs(so[99:117]).max(), 1e-6)
    assert np.abs(sg[99:117] - so[99:117]).max() <= 1e-8 * avs
    np.testing.assert_allclose(sg[117:120], so[117:120], rtol=0, atol=tol)


def term_close(o, c, term, ftol):
    F1, V1 = o.eval_term(term)
    F2, V2 = c.eval_term(term)
    w1 = o.box_state()[51:69]
    w2 = c.box_state(raw=True)[51:69]
    scale = np.abs(F1).max()
    assert scale > 0, term
    assert np.abs(F2 - F1).max() <= ftol * scale, (term, np.abs(F2 - F1).max() / scale)
    assert abs(V2 - V1) <= ftol * max(abs(V1), 1e-300) + 1e-12, (term, V1, V2)
    assert np.abs(w2 - w1).max() <= ftol * max(np.abs(w1).max(), 1e-300), term
    return V1


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
def test_harmonic_and_solid_wall_terms(precision):
    m = protein()
    prm, pbc = walled({"x": "harmonic", "y": "harmonic", "z": "solid"}, m["n"],
                      limits={"harmonic": 5}, hw_on=1)
    o, c = pair(prm, m, bbox_box(m, 1.5, pbc), precision)
    tol = FP64_TOL if precision == G.FP64 else MIXED_F_TOL
    hc_o, hc_g = o.wall_conns(0), c.wall_conns(0)
    np.testing.assert_array_equal(hc_g["res"], hc_o["res"])
    np.testing.assert_array_equal(hc_g["wall"], hc_o["wall"])
    np.testing.assert_allclose(hc_g["off"], hc_o["off"], rtol=0, atol=1e-15)
    for term in ("solid_wall", "hw_free", "hw_conn"):
        term_close(o, c, term, tol)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
def test_lj_wall_terms_and_processing(precision):
    m = protein()
    prm, pbc = walled({"x": "lj", "y": "solid", "z": "lj"}, m["n"], limits={"lj": 3},
                      cycle_dur=0.02, ljw_on=1)
    o, c = pair(prm, m, bbox_box(m, 1.0, pbc), precision)
    tol = FP64_TOL if precision == G.FP64 else MIXED_F_TOL
    V = term_close(o, c, "ljw_free", tol)
    assert V > 0
    o.ljw_process()
    c.ljw_process()
    lo, lg = o.wall_conns(1), c.wall_conns(1)
    assert len(lo["res"]) == 12
    np.testing.assert_array_equal(lg["res"], lo["res"])
    np.testing.assert_array_equal(lg["wall"], lo["wall"])
    np.testing.assert_allclose(lg["off"], lo["off"], rtol=0, atol=1e-15)
    # pull four connected residues beyond the breaking distance
    r = o.r()
    for j, i in enumerate(lo["res"][:4]):
        r[i] += (5.0 + 0.3 * j) * A * np.array([1.0, -1.0, 0.5])
    o.set_r(r)
    c.set(0, r)
    # evaluate and process in turn, as each step does: saturations move,
    # the pulled connections are released
    for k in range(5):
        F1, V1 = o.eval_term("ljw_conn")
        F2, V2 = c.eval_term("ljw_conn")
        assert np.abs(F2 - F1).max() <= tol * max(np.abs(F1).max(), 1e-300), k
        assert abs(V2 - V1) <= tol * abs(V1) + 1e-12, k
        np.testing.assert_array_equal(c.wall_conns(1)["sat"], o.wall_conns(1)["sat"])
        o.ljw_process()
        c.ljw_process()
        lo, lg = o.wall_conns(1), c.wall_conns(1)
        np.testing.assert_array_equal(lg["res"], lo["res"])
        np.testing.assert_array_equal(lg["sat"], lo["sat"])
    assert len(lo["res"]) == 8


def protocol_box(m, pbc, dt, **kw):
    """A box that squeezes fast, finds the force minimum over short averages,
    then oscillates with a short period: every phase within a few hundred
    steps. The z walls start far enough out that no residue reaches them
    (the force-minimum search moves walls that feel a force further in)."""
    lo = m["pos"].min(axis=0) - np.array([6.0, 6.5, 14.0]) * A
    hi = m["pos"].max(axis=0) + np.array([6.2, 6.7, 14.2]) * A
    vol = float(np.prod(hi - lo))
    base = dict(total_time=1e9, squeezing=1, target_density=m["n"] / (0.8 * vol),
                vel_above_2V=10.0 * A / P.TAU, target_vel=10.0 * A / P.TAU, accel_dist=0.05 * A,
                force_min=1, force_for_max_vel=0.4 * P.EPS / A, avg_forces_over=5 * dt,
                oscillations=1, angular_freq=2 * math.pi / (120 * dt), num_of_cycles=1,
                amp_abs=1.0 * A, rest_period=15 * dt)
    base.update(kw)
    return P.box_params(lo, hi, pbc, **base)


def run_lockstep(o, c, steps, chunk, check_conns=False):
    phases = set()
    for _ in range(steps // chunk):
        o.step(chunk)
        c.step(chunk)
        err = np.abs(c.r() - o.r()).max()
        assert err < 1e-8, err
        so, sg = o.box_state(), c.box_state(raw=True)
        box_close(so, sg)
        phases.add(int(so[0]))
        if check_conns:
            lo, lg = o.wall_conns(1), c.wall_conns(1)
            np.testing.assert_array_equal(lg["res"], lo["res"])
            np.testing.assert_allclose(lg["sat"], lo["sat"], rtol=0, atol=1e-12)
        assert abs(c.energy()[0] - o.V) <= 1e-9 * max(abs(o.V), 1.0)
    return phases


@pytest.mark.parametrize("graph", [False, True], ids=["stages", "graph"])
@pytest.mark.parametrize("variant", ["squeeze", "shear"])
def test_box_protocol_trajectory(variant, graph):
    m = protein(60, seed=12)
    prm, pbc = walled({"x": "solid", "y": "solid", "z": "solid"}, m["n"])
    dt = prm["dt"]
    kw = {} if variant == "squeeze" else dict(amp_relative=1, amp_rel=0.05, osc_shear=1)
    box = protocol_box(m, pbc, dt, **kw)
    o, c = pair(prm, m, box, graph=graph)
    o.box_start()
    c.box_start()
    phases = run_lockstep(o, c, 600, 25)
    assert 7 in phases, phases  # sampled mid-oscillation
    st = o.box_state()
    assert int(st[0]) == 9, G.PHASES[int(st[0])]  # through every phase
    if variant == "shear":
        assert st[12] == 1.0 and st[33 + 12] != 0.0


def test_lj_walls_under_squeezing():
    """LJ walls closing in: connections form and saturate slowly (a long
    cycle: with the reference's sign a saturated connection pulls its residue
    onto the bead, which then blows up, so a lockstep comparison stays short
    of saturation)."""
    m = protein(60, seed=14)
    prm, pbc = walled({"x": "lj", "y": "solid", "z": "lj"}, m["n"], limits={"lj": 4},
                      cycle_dur=5.0, ljw_on=1)
    dt = prm["dt"]
    box = protocol_box(m, pbc, dt, force_min=0, oscillations=0)
    o, c = pair(prm, m, box)
    o.box_start()
    c.box_start()
    run_lockstep(o, c, 300, 20, check_conns=True)
    assert len(o.wall_conns(1)["res"]) > 0


def test_periodic_squeezing():
    m = protein(60, seed=16)
    prm, pbc = walled({"x": "periodic", "y": "periodic", "z": "periodic"}, m["n"], dh_on=1)
    dt = prm["dt"]
    lo, hi = [-30 * A] * 3, [30 * A] * 3
    box = P.box_params(lo, hi, pbc, total_time=1e9, squeezing=1,
                       target_density=m["n"] / (50 * A) ** 3, vel_above_2V=20.0 * A / P.TAU,
                       target_vel=20.0 * A / P.TAU, accel_dist=0.5 * A, rest_period=20 * dt)
    o, c = pair(prm, m, box, graph=True)
    o.box_start()
    c.box_start()
    run_lockstep(o, c, 300, 25)
    assert o.box_state()[117] < 55 * A
    assert c.nl_rebuilds() == o.nl_rebuilds()


def test_harmonic_walls_in_the_step():
    """Harmonic walls evaluated by pcg_step (their connections push away
    from the bead, as the reference's sign has it: a short run)."""
    m = protein(60, seed=18)
    prm, pbc = walled({"x": "harmonic", "y": "solid", "z": "solid"}, m["n"],
                      limits={"harmonic": 4}, hw_on=1, wall_hh1=0.05 * P.EPS / A ** 2)
    o, c = pair(prm, m, bbox_box(m, 6.0, pbc))
    run_lockstep(o, c, 40, 10)
    assert np.abs(o.box_state()[51:57]).max() > 0  # the x planes felt their connections


def test_checkpoint_mid_protocol_resumes():
    m = protein(60, seed=12)
    prm, pbc = walled({"x": "solid", "y": "lj", "z": "solid"}, m["n"], limits={"lj": 3},
                      cycle_dur=5.0, ljw_on=1)
    dt = prm["dt"]
    box = protocol_box(m, pbc, dt)
    o, c = pair(prm, m, box, graph=True)
    o.box_start()
    c.box_start()
    o.step(150)
    c.step(150)
    ck = c.checkpoint()
    c2 = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=99)
    c2.restore(ck)
    c2.use_graph(True)
    for _ in range(4):
        o.step(50)
        c.step(50)
        c2.step(50)
        # the plain path sums forces with atomics: equal to rounding, not bits
        assert np.isfinite(c.r()).all() and np.abs(c2.r() - c.r()).max() < 1e-10
        box_close(c.box_state(raw=True), c2.box_state(raw=True))
        assert np.abs(c.r() - o.r()).max() < 1e-8
        box_close(o.box_state(), c.box_state(raw=True))


def test_run_phases_hands_over_to_the_box():
    """protocol.run_phases without AFM: equilibration, then the box
    protocol on the device to the total time, as the oracle does it."""
    from pas_cg_amd import protocol as X
    m = protein(60, seed=12)
    prm, pbc = walled({"x": "solid", "y": "solid", "z": "solid"}, m["n"])
    dt = prm["dt"]
    o, c = pair(prm, m, protocol_box(m, pbc, dt))
    ph = X.run_phases(c, m["chain"], dt, 300 * dt, equil_time=40 * dt, box=True)
    assert [p for p, _ in ph] == ["EQUIL", "BOX"]
    o.step(ph[0][1])
    o.box_start()
    o.step(ph[1][1])
    assert ph[0][1] + ph[1][1] == X.steps_until(0.0, 300 * dt, dt)
    assert np.abs(c.r() - o.r()).max() < 1e-8
    box_close(o.box_state(), c.box_state(raw=True))
    assert int(o.box_state()[0]) >= 5
