"""The AFM tips' statistics (pcg_afm_stats): the moving averages that
afm::vel::eval_forces and afm::force::eval_forces keep per tip
(afm/vel/eval_forces.cpp:23-29, afm/force/eval_forces.cpp:12,
cg/include/cg/types/avg.h:14-29) and make_report::add_afm_stuff prints.

CPU: the oracle's averages after a run equal the reference's own
eval_forces sources fed the same samples (time, positions and velocities
of every force evaluation of that run), for a resting tip (whose
perpendicular force is NaN in both), a pulling tip and a force tip; with a
zero window every average is its last sample. GPU: the plain context's
averages follow the oracle's through a ran2 FP64 run in the stage form, and
the persistent form's equal the stage form's on a noise-free run; new tips
start over.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_ffi import Oracle
from pas_cg_amd import model as M
from pas_cg_amd import params as P


def _prm(window_steps=6, **kw):
    p = P.default_params()
    p.update(qa_on=0, afm_on=1)
    p["afm_stats_window"] = window_steps * p["dt"]
    p.update(kw)
    return p


def _model(n=40, seed=4, force=True):
    m = M.add_pulling(M.synthetic_protein(n, M.job_seed(seed)), vel_A_per_tau=0.5)
    if force:
        mid = n // 2
        m["ftip_res"] = np.array([mid], np.int32)
        m["ftip_force"] = np.array([[0.7, -0.3, 0.2]], np.float64)
    return m


def _run_recording(o, prm, steps):
    """Steps the oracle one step at a time, recording what each force
    evaluation sees: the time, the positions, and the velocities (y1 / dt
    as the last Langevin step left them, zero before the first)."""
    n = o.r().shape[0]
    dt_inv = 1.0 / prm["dt"]
    ts, rs, vs = [], [], []
    for k in range(steps):
        ts.append(o.t)
        rs.append(o.r())
        vs.append(o.y(1) * dt_inv if k > 0 else np.zeros((n, 3)))
        o.step(1)
    return np.array(ts), np.array(rs), np.array(vs)


def _same(a, b, rtol=0.0):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    assert np.allclose(a[ok], b[ok], rtol=rtol, atol=0.0), (a, b)


def test_oracle_stats_match_the_reference():
    import ref_ffi as R
    if not R.available():
        pytest.skip("reference kernels not built (oracle/_ref)")
    L = R.lib()
    L.ref_afm_stats.restype = None
    L.ref_afm_stats.argtypes = [R.I, R.I, R.PD, R.PD, R.PD, R.I, R.PI, R.PD, R.PD, R.I, R.PI,
                                R.PD, R.D, R.D, R.D, R.PD, R.PD, R.PU8]
    prm = _prm(window_steps=6)
    m = _model()
    o = Oracle(prm, m, seed=5)
    steps = 40
    ts, rs, vs = _run_recording(o, prm, steps)
    got = o.afm_stats()
    n = m["n"]
    vres = np.ascontiguousarray(m["tip_res"], np.int32)
    vorig = np.ascontiguousarray(m["tip_orig"], np.float64)
    vvel = np.ascontiguousarray(m["tip_vel"], np.float64)
    fres = np.ascontiguousarray(m["ftip_res"], np.int32)
    ff = np.ascontiguousarray(m["ftip_force"], np.float64)
    k = len(vres) + len(fres)
    avg, perp, has = np.zeros((k, 3)), np.zeros(k), np.zeros(k, np.uint8)
    p = R.p
    L.ref_afm_stats(n, steps, p(np.ascontiguousarray(ts), R.D), p(np.ascontiguousarray(rs), R.D),
                    p(np.ascontiguousarray(vs), R.D), len(vres), p(vres, R.I), p(vorig, R.D),
                    p(vvel, R.D), len(fres), p(fres, R.I), p(ff, R.D), prm["afm_H1"],
                    prm["afm_H2"], prm["afm_stats_window"], p(avg, R.D), p(perp, R.D),
                    p(has, R.U8))
    assert has.astype(bool).tolist() == got["has"].tolist() == [True] * 3
    _same(got["avg"], avg, rtol=1e-15)
    _same(got["perp"][:2], perp[:2], rtol=1e-15)
    assert np.isnan(got["perp"][0])          # the resting tip has no direction
    assert got["avg"][2].any()               # the force tip's residue moves


def test_zero_window_keeps_the_last_sample():
    """window 0: every add closes the window, so the average is the last
    sample: a force tip's is its residue's velocity before the last step."""
    prm = _prm(window_steps=0)
    m = _model()
    o = Oracle(prm, m, seed=6)
    o.step(5)
    v_before = o.y(1)[m["ftip_res"][0]] * (1.0 / prm["dt"])
    o.step(1)
    s = o.afm_stats()
    assert s["has"].all()
    assert np.array_equal(s["avg"][2], v_before)
    F, _ = o.eval_term("vel_afm")
    # the tip's average is minus the force of the last evaluation on its
    # residue (the eval_term above adds one more sample at the same time)
    s2 = o.afm_stats()
    assert np.allclose(s2["avg"][1], -F[m["tip_res"][1]], rtol=1e-14, atol=0)


def test_new_tips_start_over():
    prm = _prm(window_steps=3)
    m = _model()
    o = Oracle(prm, m, seed=6)
    o.step(10)
    assert o.afm_stats()["has"].all()
    o.set_tips({k: m[k] for k in ("tip_res", "tip_orig", "tip_vel")})
    s = o.afm_stats()
    assert len(s["has"]) == 2 and not s["has"].any()


@pytest.mark.gpu
def test_context_stats_follow_oracle():
    from pas_cg_amd import pcg as G
    prm = _prm(window_steps=7)
    m = _model()
    c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=8)
    o = Oracle(prm, m, seed=8)
    for _ in range(3):
        c.step(15)
        o.step(15)
        a, b = c.afm_stats(), o.afm_stats()
        assert a["has"].tolist() == b["has"].tolist()
        _same(a["avg"], b["avg"], rtol=1e-9)
        _same(a["perp"], b["perp"], rtol=1e-9)
    assert a["has"].all()
    # new tips start over on both
    tips = {k: m[k] for k in ("tip_res", "tip_orig", "tip_vel", "ftip_res", "ftip_force")}
    c.set_tips(tips)
    o.set_tips(tips)
    assert not c.afm_stats()["has"].any()
    c.step(10)
    o.step(10)
    a, b = c.afm_stats(), o.afm_stats()
    assert a["has"].tolist() == b["has"].tolist() and a["has"].all()
    _same(a["avg"], b["avg"], rtol=1e-9)
    c.close()


@pytest.mark.gpu
def test_persistent_form_keeps_the_same_stats():
    from pas_cg_amd import pcg as G
    prm = _prm(window_steps=5)
    m = _model(n=90, seed=9)
    ctx = []
    for form in (G.FORM_STAGES, G.FORM_PERSISTENT):
        c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=3, philox_seed=4)
        c.set_noise_scale(0.0)
        assert c.step_form(form) == form
        c.step(37)
        ctx.append(c)
    a, b = ctx[0].afm_stats(), ctx[1].afm_stats()
    assert a["has"].tolist() == b["has"].tolist() and a["has"].all()
    _same(a["avg"], b["avg"], rtol=1e-9)
    _same(a["perp"], b["perp"], rtol=1e-9)
    for c in ctx:
        c.close()
