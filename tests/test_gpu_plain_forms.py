"""The plain context's two step forms (pcg_step_form): the stage kernels and
the persistent single-system kernel take the same FP64 trajectory (noise
off: the forms sum forces in different orders, so they agree to rounding,
not bits); the persistent form equals a one-system batch; switching forms
mid-run rebuilds the lists and stays on the trajectory; ineligible systems
keep the stage form."""
import numpy as np
import pytest

from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G

pytestmark = pytest.mark.gpu


def _prm(**kw):
    p = P.default_params()
    p.update(chir_on=1, qa_on=0, pid_on=0)
    p.update(kw)
    return p


def _ctx(prm, m, form, seed=11):
    c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=seed, philox_seed=5)
    c.set_noise_scale(0.0)
    assert c.step_form(form) == (form if form else G.FORM_PERSISTENT)
    return c


def test_forms_follow_the_same_trajectory():
    m = M.synthetic_protein(150, M.job_seed(21))
    prm = _prm()
    a = _ctx(prm, m, G.FORM_STAGES)
    b = _ctx(prm, m, G.FORM_AUTO)
    for _ in range(4):
        a.step(50)
        b.step(50)
        assert np.abs(a.r() - b.r()).max() < 1e-9
        assert abs(a.energy()[0] - b.energy()[0]) <= 1e-9 * abs(a.energy()[0])
        assert a.time() == b.time()
    assert np.abs(a.F() - b.F()).max() <= 1e-9 * np.abs(a.F()).max()
    # switching forms mid-run: each switch rebuilds the lists (QA and PID
    # off, so a rebuild does not change the forces)
    b.step_form(G.FORM_STAGES)
    b.step(30)
    b.step_form(G.FORM_PERSISTENT)
    b.step(30)
    a.step(60)
    assert np.abs(a.r() - b.r()).max() < 1e-9
    a.close()
    b.close()


def test_persistent_form_equals_a_one_system_batch():
    m = M.synthetic_protein(120, M.job_seed(23))
    prm = _prm(qa_on=1)
    c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=4, philox_seed=9)
    assert c.step_form(G.FORM_AUTO) == G.FORM_PERSISTENT
    b = G.Batch(prm, [m], precision=G.FP64, philox_seed=9, seeds=[4])
    c.step(300)
    b.run(300)
    b.sync()
    assert np.abs(c.r() - b.get(0, 0)).max() < 1e-8
    assert c.time()[1] == 300
    c.close()
    b.close()


def test_ineligible_contexts_keep_the_stages():
    m = M.synthetic_protein(60, M.job_seed(25))
    c = G.Context(_prm(), m, precision=G.FP64, noise=G.RAN2, seed=3)
    assert c.step_form(G.FORM_AUTO) == G.FORM_STAGES
    with pytest.raises(G.PcgError):
        c.step_form(G.FORM_PERSISTENT)
    c.close()
    prm = _prm()
    P.set_solid_walls(prm, m["pos"].min(axis=0) - 1.0, m["pos"].max(axis=0) + 1.0)
    c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=3)
    assert c.step_form(G.FORM_AUTO) == G.FORM_STAGES
    c.close()
