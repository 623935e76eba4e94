"""pcg_batch_report_static, the report entry point that allocates nothing
(its sums live in a device array of the library), against pcg_batch_report
through the C ABI: two calls in a row, the second on a smaller batch, must
give the same scalars; a batch beyond the array's size is refused."""
import numpy as np
import pytest

from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G

pytestmark = pytest.mark.gpu


def _params():
    p = P.default_params()
    p.update(chir_on=1, pid_on=1, qa_on=1)
    return p


def _batch(sizes, seed0):
    models = [M.synthetic_protein(n, M.job_seed(seed0 + k)) for k, n in enumerate(sizes)]
    b = G.Batch(_params(), models, precision=G.MIXED, steps_per_launch=50)
    b.run(100)
    return b


def _same(a, b):
    for ra, rb in zip(a, b):
        for k in G.REPORT_FIELDS:
            assert ra[k] == rb[k] or abs(ra[k] - rb[k]) <= 1e-12 * max(abs(ra[k]), 1.0), (k, ra[k], rb[k])


def test_report_static_matches_report_on_two_calls():
    big = _batch([40, 310, 75, 180, 33, 120, 96], seed0=700)
    ref = big.report()
    got = big.report_static()
    assert len(got) == 7
    _same(ref, got)
    small = _batch([64, 51], seed0=800)
    ref2 = small.report()
    got2 = small.report_static()
    assert len(got2) == 2
    _same(ref2, got2)
    # the larger batch again after the smaller one: nothing left over
    _same(big.report(), big.report_static())
    assert any(r["PID-INTRHC"] > 0 or r["INTRHC"] > 0 for r in ref)
    big.close()
    small.close()


def test_report_static_refuses_more_systems_than_its_array():
    m = M.synthetic_protein(3, M.job_seed(1))
    p = _params()
    p.update(qa_on=0, pid_on=0)
    b = G.Batch(p, [m] * 4097, precision=G.MIXED, steps_per_launch=1)
    with pytest.raises(G.PcgError, match="4096"):
        b.report_static()
    b.close()
