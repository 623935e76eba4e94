"""The component-array entry points (pcg_get3, pcg_set3, pcg_eval_term3)
against the interleaved ones, through the C ABI: unit stride (the
reference's structure-of-arrays vect::vector<vec3r>) and stride 3 (an
interleaved buffer seen through three offset pointers)."""
import ctypes as C

import numpy as np
import pytest

from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G

pytestmark = pytest.mark.gpu
D = C.c_double


def _ctx(precision):
    prm = P.default_params()
    prm.update(qa_on=0, chir_on=1)
    m = M.synthetic_protein(70, M.job_seed(3))
    c = G.Context(prm, m, precision=precision, noise=G.RAN2, seed=3)
    c.step(5)
    return c


def _strided(c, which):
    buf = np.zeros(3 * c.n)
    base = buf.ctypes.data
    px = C.cast(base, C.POINTER(D))
    py = C.cast(base + 8, C.POINTER(D))
    pz = C.cast(base + 16, C.POINTER(D))
    G.check(c.L.pcg_get3(c.h, which, px, py, pz, 3))
    return buf.reshape(c.n, 3)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
def test_component_arrays_match_interleaved(precision):
    c = _ctx(precision)
    for which in (0, 1, 6, 7):
        a = c.get(which)
        np.testing.assert_array_equal(c.get3(which).T, a)
        np.testing.assert_array_equal(_strided(c, which), a)
    r = c.r() + 0.01
    c.set3(0, np.ascontiguousarray(r.T))
    np.testing.assert_array_equal(c.r(), r)
    c.set(0, r - 0.01)
    # two evaluations sum their atomics in their own orders: equal to the
    # rounding of the force accumulators (FP64, or FP32 in mixed)
    tol = 1e-12 if precision == G.FP64 else 1e-6
    for term in ("tether", "chir", "nat_cont", "nat_ang", "cnd"):
        F, V = c.eval_term(term)
        F3, V3 = c.eval_term3(term)
        assert np.abs(F3.T - F).max() <= tol * max(np.abs(F).max(), 1e-300), term
        assert abs(V3 - V) <= 1e-12 * max(abs(V), 1e-300), term
    Vonly = C.c_double()
    G.check(c.L.pcg_eval_term3(c.h, G.TERM_ID["tether"], 0, -1, None, None, None, 1,
                               C.byref(Vonly)))
    assert abs(Vonly.value - c.eval_term("tether")[1]) <= 1e-12 * abs(Vonly.value)
    with pytest.raises(G.PcgError):
        G.check(c.L.pcg_get3(c.h, 0, None, None, None, 1))
    c.close()
