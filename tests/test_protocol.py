"""Protocol phases (pas_cg_amd.protocol, thread_main.cpp:84-142) and the AFM
tips placed when a phase turns AFM on (pcg_set_tips, state.cpp:548-568).

CPU: the step count of `while t < until` with the accumulated time, the tip
geometry of setup_afm, and the oracle's tips reset: a velocity tip placed
at its residue pulls nothing at the moment it is placed, then pulls along
its velocity. GPU: the drop-in context run through equilibration and then
pulling (constant velocity or constant force, tips placed at the positions
of that moment) follows the oracle driven through the same phases, in FP64
with the reference's ran2 noise; pull-then-release likewise; the batched
form places tips per system.
"""
import numpy as np
import pytest

from oracle_ffi import Oracle
from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import protocol as X


def test_steps_until_counts_the_accumulated_time():
    dt = P.default_params()["dt"]
    assert X.steps_until(0.0, 0.0, dt) == 0
    assert X.steps_until(0.0, dt, dt) == 1
    t, k = 0.0, 0
    while t < 37 * dt:
        t += dt
        k += 1
    assert X.steps_until(0.0, 37 * dt, dt) == k
    assert X.steps_until(5 * dt, 5 * dt, dt) == 0


def test_setup_afm_tips_geometry():
    m = M.synthetic_protein(30, M.job_seed(2))
    pos = m["pos"]
    tv = X.setup_afm_tips(pos, m["chain"], "const velocity", 0.01)
    u = (pos[-1] - pos[0]) / np.linalg.norm(pos[-1] - pos[0])
    assert tv["tip_res"].tolist() == [0, 29]
    assert np.allclose(tv["tip_vel"][1], 0.01 * u, rtol=1e-14) and not tv["tip_vel"][0].any()
    assert np.array_equal(tv["tip_orig"], pos[[0, 29]])
    tf = X.setup_afm_tips(pos, m["chain"], "const force", 2.0)
    assert tf["ftip_res"].tolist() == [29] and np.allclose(tf["ftip_force"][0], 2.0 * u, rtol=1e-14)
    with pytest.raises(ValueError):
        X.setup_afm_tips(pos, m["chain"], "walls", 1.0)


def _afm_params():
    p = P.default_params()
    p.update(qa_on=0, afm_on=1)
    return p


def test_oracle_tips_placed_now_start_at_their_residue():
    m = M.add_pulling(M.synthetic_protein(40, M.job_seed(4)))
    o = Oracle(_afm_params(), m, seed=4)
    o.step(30)
    tips = X.setup_afm_tips(o.r(), m["chain"], "const velocity", 0.5)
    o.set_tips(tips)
    F, V = o.eval_term("vel_afm")
    assert not F.any() and V == 0.0          # every tip sits on its residue
    o.set_noise_scale(0.0)
    o.step(20)
    F, V = o.eval_term("vel_afm")
    assert V > 0 and np.abs(F[-1]).max() > 0  # the pulled tip has moved on


def _run_both(kind, steps_equil, steps_pull, magnitude):
    from pas_cg_amd import pcg as G
    p = _afm_params()
    m = M.add_pulling(M.synthetic_protein(48, M.job_seed(6)))
    dt = p["dt"]
    c = G.Context(p, m, precision=G.FP64, noise=G.RAN2, seed=6)
    o = Oracle(p, m, seed=6)
    total = (steps_equil + steps_pull) * dt
    if kind == "pull, then release":
        ph = X.run_phases(c, m["chain"], dt, total, equil_time=steps_pull * dt, afm=kind,
                          pull_release_time=steps_equil * dt, pull_speed=magnitude)
        o.set_tips(X.setup_afm_tips(o.r(), m["chain"], "const velocity", magnitude))
        o.step(ph[0][1])
        o.set_tips({})
        o.step(ph[1][1])
        o.step(ph[2][1])
    else:
        ph = X.run_phases(c, m["chain"], dt, total, equil_time=steps_equil * dt, afm=kind,
                          magnitude=magnitude)
        o.set_tips({})
        o.step(ph[0][1])
        o.set_tips(X.setup_afm_tips(o.r(), m["chain"], kind, magnitude))
        o.step(ph[1][1])
    return c, o, ph


@pytest.mark.gpu
@pytest.mark.parametrize("kind,mag", [("const velocity", 0.05), ("const force", 3.0),
                                      ("pull, then release", 0.05)])
def test_context_phases_follow_oracle(kind, mag):
    c, o, ph = _run_both(kind, 30, 40, mag)
    assert sum(k for _, k in ph) == 70
    assert np.abs(c.r() - o.r()).max() < 1e-9
    V = c.energy()[0]
    assert abs(V - o.V) <= 1e-9 * abs(o.V), (V, o.V)
    c.close()


@pytest.mark.gpu
def test_batch_set_tips_per_system():
    """Tips placed on one system of a batch mid-run pull it as the oracle's
    do; the other system, without tips, runs as before."""
    from pas_cg_amd import pcg as G
    p = _afm_params()
    models = [M.add_pulling(M.synthetic_protein(n, M.job_seed(70 + n))) for n in (40, 57)]
    b = G.Batch(p, models, seeds=[1, 2], precision=G.FP64, steps_per_launch=10)
    b.set_noise_scale(0.0)
    oo = [Oracle(p, m, seed=s) for m, s in zip(models, (1, 2))]
    for s, o in enumerate(oo):
        o.set_noise_scale(0.0)
        b.set_tips(s, {})
        o.set_tips({})
    b.run(20)
    for o in oo:
        o.step(20)
    tips = X.setup_afm_tips(b.get(0, 0), models[0]["chain"], "const force", 4.0)
    b.set_tips(0, tips)
    oo[0].set_tips(X.setup_afm_tips(oo[0].r(), models[0]["chain"], "const force", 4.0))
    b.run(30)
    for s, o in enumerate(oo):
        o.step(30)
        assert np.abs(b.get(s, 0) - o.r()).max() < 1e-8, s
    with pytest.raises(G.PcgError):
        b.set_tips(1, X.setup_afm_tips(b.get(1, 0), models[1]["chain"], "const velocity", 0.1)
                   | {"ftip_res": np.array([3], np.int32), "ftip_force": np.ones((1, 3))})
    b.close()
