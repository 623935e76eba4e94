"""The report scalars (make_report.cpp:132-247).

CPU: the oracle's restatement (oracle/report.py) is pinned by closed forms and
an independent method: RMSD of a rigidly moved copy is 0, RMSD equals Horn's
quaternion solution (an eigenvector of a 4x4 matrix, no SVD), Rg of a cube's
corners, asphericity of three collinear points, and a mirror image of a
chiral set keeps a non-zero RMSD (no reflections, d = sign det).
GPU: pcg_report and pcg_batch_report against the oracle on live states.

PID activity counts (pid/count_active.cpp:4-164): the oracle's count is
pinned against the reference's own count_active in test_oracle_vs_ref.py.
The device counts every pair of a system that the list rules admit instead
of walking the PID list; the CPU test here pins the fact that makes the two
equal (the count over a list rebuilt some steps ago equals the count over a
list rebuilt at the current positions), and the GPU tests compare the
device's counts with the oracle's on the same positions.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import report as R  # noqa: E402

from pas_cg_amd import model as M  # noqa: E402
from pas_cg_amd import params as P  # noqa: E402


def _rot(seed):
    q = np.random.default_rng(seed).normal(size=4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def _horn_rmsd(x, y):
    """Optimal-superposition RMSD by Horn's quaternion method (J. Opt. Soc.
    Am. A 4, 629 (1987)): the largest eigenvalue of a symmetric 4x4 matrix."""
    x = x - x.mean(0)
    y = y - y.mean(0)
    S = x.T @ y
    Sxx, Sxy, Sxz, Syx, Syy, Syz, Szx, Szy, Szz = S.ravel()
    N = np.array([[Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx],
                  [Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz],
                  [Szx - Sxz, Sxy + Syx, -Sxx + Syy - Szz, Syz + Szy],
                  [Sxy - Syx, Szx + Sxz, Syz + Szy, -Sxx - Syy + Szz]])
    lam = np.linalg.eigvalsh(N)[-1]
    return np.sqrt(max((x * x).sum() + (y * y).sum() - 2 * lam, 0.0) / len(x))


def test_rmsd_of_a_rigidly_moved_copy_is_zero():
    x = np.random.default_rng(1).normal(size=(50, 3)) * 10
    y = x @ _rot(2).T + np.array([3.0, -7.0, 11.0])
    assert R.rmsd(x, y) < 1e-12


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_rmsd_matches_horn_quaternion_method(seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(80, 3)) * 8
    y = x @ _rot(seed + 10).T + rng.normal(size=(80, 3)) * 0.7
    assert abs(R.rmsd(x, y) - _horn_rmsd(x, y)) < 1e-10


def test_rmsd_does_not_reflect():
    x = np.random.default_rng(6).normal(size=(30, 3)) * 5
    y = x * np.array([-1.0, 1.0, 1.0])  # mirror image
    assert R.rmsd(x, y) > 0.1
    assert abs(R.rmsd(x, y) - _horn_rmsd(x, y)) < 1e-10


def test_gyration_radius_and_asphericity_closed_forms():
    a = 2.0
    cube = np.array([[i, j, k] for i in (0, a) for j in (0, a) for k in (0, a)], float)
    assert abs(R.gyration_radius(cube) - a * np.sqrt(3) / 2) < 1e-14
    line = np.array([[-1.0, 0, 0], [0, 0, 0], [1.0, 0, 0]]) + 5.0
    # singular values of the centred line: sqrt(2), 0, 0 -> W = sqrt(2)
    assert abs(R.asphericity(line) - np.sqrt(2.0)) < 1e-14
    # a regular tetrahedron's centred positions have equal singular values -> W = 0
    tet = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    assert abs(R.asphericity(tet)) < 1e-13


def _oracle_of_context(c, m):
    from pas_cg_amd import pcg as G
    r = c.r()
    v = c.get(7)
    t, _ = c.time()
    V, _ = c.energy()
    mass = np.array([P.default_params()["mass"][a] for a in m["atype"]])
    act, _, _ = c.nat_cont_state(len(m["cont"]))
    i1, i2, _, vac = c.list(6)
    pairs = [(a, b) for a, b, f in zip(i1, i2, vac) if not f]
    return R.scalars(t, V, r, m.get("orig_pos", m["pos"]), v, mass, nc_active=act,
                     nc_type=np.asarray(m["cont_type"]), qa_pairs=pairs, chain=m["chain"])


@pytest.mark.gpu
def test_context_report_matches_oracle():
    """pcg_report of the drop-in context, on a periodic three-chain box with
    native contacts and QA on, after 400 noisy steps, against the oracle's
    scalars of the same state."""
    from pas_cg_amd import pcg as G
    one = M.synthetic_protein(40, M.job_seed(77))
    m = M.replicate(one, 3, 45.0 * P.ANGSTROM, 77)
    prm = P.default_params()
    prm.update(chir_on=1, qa_on=1)
    c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=7)
    c.step(400)
    got = c.report()
    want = _oracle_of_context(c, m)
    for k, w in want.items():
        assert abs(got[k] - w) <= 1e-9 * max(abs(w), 1.0), (k, got[k], w)
    assert want["ICN"] > 0
    c.close()


@pytest.mark.gpu
def test_batch_report_matches_oracle():
    """pcg_batch_report of every system of a ragged batch after a mixed run
    against the oracle's geometric and energy scalars of the same state."""
    from pas_cg_amd import pcg as G
    prm = P.default_params()
    prm.update(chir_on=1, qa_on=1, afm_on=1)
    models = [M.synthetic_protein(n, M.job_seed(90 + n)) for n in (25, 140, 61)]
    b = G.Batch(prm, models, precision=G.MIXED, steps_per_launch=150)
    b.run(300)
    reps = b.report()
    for s, m in enumerate(models):
        r, v = b.get(s, 0), b.get(s, 7)
        V, _ = b.energy(s)
        mass = np.array([prm["mass"][a] for a in m["atype"]])
        want = R.scalars(reps[s]["TIME"], V, r, m["orig_pos"], v, mass)
        for k, w in want.items():
            assert abs(reps[s][k] - w) <= 1e-9 * max(abs(w), 1.0), (s, k, reps[s][k], w)
        assert 0 <= reps[s]["ICN"] <= len(m["cont"])
    b.close()


def _pid_params(**kw):
    """PID on with the sink-LJ forces and a side-chain depth table, which stay
    stable on compact synthetic chains; QA off (its free pairs would take the
    PID pairs)."""
    prm = P.default_params()
    prm.update(pid_on=1, qa_on=0, pid_lambda_variant=0)
    P.set_pid_forces(prm, "sink lj", bbp_depth=0.3, bbm_depth=0.25,
                     ss_depth_table=[[0.1 + 0.01 * ((a * 7 + b) % 13) for b in range(21)]
                                     for a in range(21)])
    prm.update(kw)
    return prm


def _three_chains(seed=5):
    one = M.synthetic_protein(40, M.job_seed(seed))
    return M.replicate(one, 3, 45.0 * P.ANGSTROM, seed)


@pytest.mark.parametrize("nat_cont_on", [1, 0])
def test_pid_counts_do_not_depend_on_list_age(nat_cont_on):
    from oracle_ffi import Oracle
    o = Oracle(_pid_params(nat_cont_on=nat_cont_on), _three_chains(), seed=3)
    seen = np.zeros(6, np.int64)
    for _ in range(4):
        o.step(80)
        aged = o.pid_counts()
        o.fix_nl()                     # rebuild at the current positions
        fresh = o.pid_counts()
        assert (aged == fresh).all(), (aged, fresh)
        seen += fresh
    assert seen[0] > 0 and seen[3] > 0 and seen[2] > 0


def _oracle_pid_counts(prm, m, r):
    from oracle_ffi import Oracle
    o = Oracle(prm, m, seed=1)
    o.set_r(r)
    o.fix_nl()
    return o.pid_counts()


PID_KEYS = ["PID-INTRHC", "PID-B1-B1", "PID-B1-S1", "PID-INTEHC", "PID-B1-B2", "PID-S1-S2"]


@pytest.mark.gpu
@pytest.mark.parametrize("nat_cont_on", [1, 0])
def test_context_pid_counts_match_oracle(nat_cont_on):
    """The six PID counts of pcg_report on a periodic three-chain box, at
    several points of an FP64 trajectory, equal the oracle's count_active on
    the same positions."""
    from pas_cg_amd import pcg as G
    m = _three_chains()
    prm = _pid_params(nat_cont_on=nat_cont_on)
    c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=7)
    seen = np.zeros(6)
    for _ in range(3):
        c.step(150)
        got = np.array([c.report()[k] for k in PID_KEYS])
        want = _oracle_pid_counts(prm, m, c.r())
        assert (got == want).all(), (got, want)
        seen += got
    assert seen[0] > 0 and seen[3] > 0
    c.close()


@pytest.mark.gpu
def test_batch_pid_counts_match_oracle():
    """pcg_batch_report's PID counts of every system of a ragged mixed batch
    (one periodic multichain system among single chains) equal the oracle's
    on the same positions; PID off gives zeros."""
    from pas_cg_amd import pcg as G
    prm = _pid_params()
    models = [M.synthetic_protein(n, M.job_seed(60 + n)) for n in (33, 150)] + [_three_chains(8)]
    b = G.Batch(prm, models, precision=G.MIXED, steps_per_launch=100)
    b.run(300)
    reps = b.report()
    for s, m in enumerate(models):
        got = np.array([reps[s][k] for k in PID_KEYS])
        want = _oracle_pid_counts(prm, m, b.get(s, 0))
        assert (got == want).all(), (s, got, want)
    assert sum(reps[s]["PID-INTRHC"] for s in range(3)) > 0
    assert reps[2]["PID-INTEHC"] > 0
    b.close()
    off = dict(prm, pid_on=0)
    b = G.Batch(off, models[:1], precision=G.MIXED, steps_per_launch=50)
    b.run(50)
    assert all(b.report()[0][k] == 0 for k in PID_KEYS)
    b.close()
