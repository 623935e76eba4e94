"""Step-by-step comparison of the plain context with the oracle on the
oracle's own states (test infrastructure)."""
import numpy as np

from oracle_ffi import Oracle
from pas_cg_amd import pcg as G


def qa_lockstep(prm, m, precision, steps, r0=None, t0=None, seed=31, keep=False):
    """Runs the oracle's noise-free step piece by piece and evaluates the GPU
    context on the oracle's state at each piece: verify (and both rebuild),
    prepare_nh, forces, QA finish. Returns the worst relative force error of
    the whole force and of the two QA terms, and the worst relative energy
    error of the QA terms; asserts that QA contacts, saturations and sync
    budgets stay identical (every QA decision is exact in both precisions).
    keep: also returns the oracle and the context, both on the final state."""
    o = Oracle(prm, m, seed=seed)
    o.set_noise_scale(0.0)
    if r0 is not None:
        o.set_r(r0)
        o.set_t(t0)
    c = G.Context(prm, m, precision=precision, noise=G.RAN2, seed=seed)
    c.set_noise_scale(0.0)
    worst = {"total": 0.0, "qa_cand": 0.0, "qa_cont": 0.0, "V_qa_cand": 0.0, "V_qa_cont": 0.0}
    formed = 0
    for k in range(steps):
        c.set(0, o.r())
        c.set_time(o.t)
        if k == 0 or o.nl_verify():
            o.fix_nl()
            c.fix_nl()
        o.prepare_nh()
        c.prepare_nh()
        o.eval_forces()
        c.eval_forces()
        F1, F2 = o.F(), c.F()
        worst["total"] = max(worst["total"], float(np.abs(F2 - F1).max() / np.abs(F1).max()))
        Vo, Vg = o.term_V(), c.energy()[1]
        for t in ("qa_cand", "qa_cont"):
            if Vo[t] != 0:
                worst["V_" + t] = max(worst["V_" + t], abs(Vg[t] - Vo[t]) / abs(Vo[t]))
        o.qa_finish()
        c.qa_finish()
        so, no = o.sync(), o.num_qa_contacts()
        sg, ng = c.qa_state()
        assert no == ng and np.array_equal(so, sg), (k, no, ng)
        ci, gi = o.list(6), c.list(6)
        assert np.array_equal(ci[0], gi[0]) and np.array_equal(ci[1], gi[1]), k
        # the oracle's flag is 255 for a vacant slot, else the contact type
        assert np.array_equal(ci[3] == 255, gi[3] != 0), k
        assert np.allclose(ci[2], gi[2], rtol=0, atol=1e-12), k
        formed = max(formed, no)
        o.lang_step()
    if keep:
        # both on the oracle's final state
        c.set(0, o.r())
        c.set_time(o.t)
        return worst, formed, o, c
    c.close()
    return worst, formed
