"""The stage rebuild's cell grid (systems of at least 2048 residues in a
cell periodic along every axis): the neighbour list and every list the
partition makes from it are bit-exact with the oracle's O(N^2)
nl::legacy_update (cg/src/cg/nl/legacy_update.cpp:11-55) and the ordered
updates, on the 64 x 100-mer box at 1 mM (6400 residues, 21 cells a side,
chains across the periodic faces), on the same box squeezed to 3 cells a
side (the smallest grid), and after the positions move; a cell too small
for the NL radius falls back to the scan over every column."""
import numpy as np
import pytest

from oracle_ffi import Oracle
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G
from pas_cg_amd import workloads as W

pytestmark = pytest.mark.gpu


def _lists_equal(o, c, tag):
    for kind in range(6):  # NL, native contacts, pauli, dh, pid, qa free pairs
        a, g = o.list(kind), c.list(kind)
        assert len(a[0]) == len(g[0]), (tag, kind, len(a[0]), len(g[0]))
        assert np.array_equal(a[0], g[0]) and np.array_equal(a[1], g[1]), (tag, kind)
        if kind == 0:
            assert np.array_equal(a[2], g[2]), (tag, "NL distances not bit-exact")
    return len(o.list(0)[0])


def _box(scale=1.0, qa=True):
    prm, m = W.c1()
    prm.update(qa_on=1 if qa else 0, pid_on=0 if qa else 1)
    m = dict(m)
    m["pos"] = m["pos"] * scale
    m["orig_pos"] = m["orig_pos"] * scale
    m["cell"] = np.asarray(m["cell"], np.float64) * scale
    return prm, m


@pytest.mark.parametrize("scale,qa", [(1.0, True), (1.0, False), (0.15, True)],
                         ids=["box_qa", "box_pid", "squeezed_3_cells"])
def test_cell_grid_lists_bit_exact(scale, qa):
    prm, m = _box(scale, qa)
    side = m["cell"][0] / ((max(prm["nl_cutoff"], prm["dh_cutoff"]) + prm["nl_pad"]) * (1 + 1e-6))
    assert int(side) >= 3  # the grid is in use
    o = Oracle(prm, m, seed=3)
    c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=3)
    o.fix_nl()
    c.fix_nl()
    assert _lists_equal(o, c, "start") > 0
    # positions moved by up to 6 A, some across the periodic faces
    rng = np.random.default_rng(5)
    r = o.r() + rng.uniform(-6, 6, size=o.r().shape) * P.ANGSTROM
    o.set_r(r)
    c.set(0, r)
    o.fix_nl()
    c.fix_nl()
    _lists_equal(o, c, "moved")
    c.close()


def test_small_cell_falls_back_to_the_column_scan():
    """A cell too small for three grid cells of the NL radius: no grid, the
    lists still equal the oracle's."""
    prm, m = _box(0.12, True)
    o = Oracle(prm, m, seed=4)
    c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=4)
    o.fix_nl()
    c.fix_nl()
    assert _lists_equal(o, c, "small") > 0
    c.close()
