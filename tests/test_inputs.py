"""Input formats (pas_cg_amd.inputs) and the library's SAW generator.

* The package's PDB and sequence-file readers give, from the reference's own
  parity inputs, exactly the models of the golden fixtures (whose V traces
  the oracle and the GPU reproduce; test_golden.py, test_gpu_parity.py).
  These run where the reference tree is mounted.
* Self-contained cases pin the record handling by hand: chains in id order,
  hydrogens and later models skipped, SSBOND contacts, CRYST1, a right
  angle, a trans dihedral, contact maps with shifts and taken angles.
* pcg_morph_saw (host code, no GPU) equals the oracle's SAW bit for bit and
  uses the same number of ran2 draws; the oracle's SAW is pinned by the 9aac
  and 1ubq golden traces, which start from it.
"""
import math
import os

import numpy as np
import pytest

from golden_cases import case, load
from pas_cg_amd import inputs as I
from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G

REF = "/root/reference/parity/tests"
A = P.ANGSTROM
KEYS = ("pos", "orig_pos", "atype", "chain", "seq", "teth", "ang", "ang_has", "ang_theta",
        "dih", "dih_has", "dih_phi", "cont", "cont_len", "cell")

need_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
need_lib = pytest.mark.skipif(not os.path.exists(G.LIB_PATH), reason="libpcg.so not built")


@need_ref
@pytest.mark.parametrize("name,path,kw", [
    ("start2", "start2/cxx_7/start_7_2.pdb", dict(contacts="none")),
    ("TEST", "TEST/cxx/start.pdb", dict(contacts="none", load_structure=False)),
    ("input_test", "input_test/cxx/start_6_2.pdb", dict(contacts="none", load_structure=False)),
    ("sample", "sample/cxx/sample.pdb", dict(contacts="none", load_structure=False)),
    ("1ubq", "1ubq/cxx/1ubq.pdb", dict(contacts="all_atoms")),
])
def test_pdb_reader_gives_the_fixture_models(name, path, kw):
    fx = load(name)
    want = M.from_fixture(fx["model"], pbc=fx["pbc"])
    got = I.model_from_pdb(os.path.join(REF, path), pbc=fx["pbc"], **kw)
    for k in KEYS:
        assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), k
    if name == "1ubq":
        # ubiquitin has no cysteines: every contact is backbone/side typed
        assert set(np.unique(got["cont_type"]).tolist()) == {1, 2, 3, 4}


@need_ref
def test_seq_file_reader_gives_the_fixture_model():
    want = M.from_fixture(load("9aac")["model"], pbc=False)
    got = I.model_from_seq_file(os.path.join(REF, "9aac/cxx/9aac.yml"))
    for k in ("atype", "chain", "seq", "teth", "ang", "ang_has", "dih", "dih_has", "cont"):
        assert np.array_equal(got[k], want[k]), k
    assert not got["pos"].any()


def _atom(serial, name, res, chain, seq, x, y, z):
    return (f"ATOM  {serial:5d} {name:<4s} {res:3s} {chain}{seq:4d}    "
            f"{x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00")


def _pdb_text():
    # chain B first in the file: chains are ordered by id. Chain A: four
    # residues on a right angle and a planar trans zig-zag; hydrogens and a
    # second model are ignored; an SSBOND joins A1 and B2.
    ca_a = [(0.0, 0.0, 0.0), (3.8, 0.0, 0.0), (3.8, 3.8, 0.0), (7.6, 3.8, 0.0)]
    ca_b = [(20.0, 0.0, 0.0), (20.0, 3.8, 0.0), (20.0, 7.6, 0.0)]
    lines = ["CRYST1   50.000   60.000   70.000  90.00  90.00  90.00 P 1           1",
             "SSBOND   1 CYS A    1    CYS B    2                          1555   1555  2.04",
             "MODEL        1"]
    s = 1
    for k, (x, y, z) in enumerate(ca_b):
        lines.append(_atom(s, "CA", "GLY", "B", k + 1, x, y, z)); s += 1
    for k, (x, y, z) in enumerate(ca_a):
        res = "CYS" if k == 0 else "ALA"
        lines.append(_atom(s, "N", res, "A", k + 1, x - 1.0, y, z)); s += 1
        lines.append(_atom(s, "CA", res, "A", k + 1, x, y, z)); s += 1
        lines.append(_atom(s, "H", res, "A", k + 1, x, y + 1.0, z)); s += 1
    lines += ["ENDMDL", "MODEL        2", _atom(s, "CA", "ALA", "C", 1, 1.0, 1.0, 1.0),
              "ENDMDL", "END"]
    return "\n".join(lines) + "\n"


def test_pdb_records_by_hand():
    m = I.model_from_pdb(_pdb_text(), contacts="none", pbc=True)
    assert m["n"] == 7
    assert m["chain"].tolist() == [0, 0, 0, 0, 1, 1, 1]
    assert m["atype"][0] == P.AA_INDEX["CYS"] and m["atype"][4] == P.AA_INDEX["GLY"]
    assert np.allclose(m["pos"][1], [3.8 * A, 0, 0], rtol=0, atol=1e-15)
    assert m["teth"].tolist() == [[0, 1], [1, 2], [2, 3], [4, 5], [5, 6]]
    assert m["ang"].tolist() == [[0, 1, 2], [1, 2, 3], [4, 5, 6]]
    assert abs(m["ang_theta"][0] - math.pi / 2) < 1e-12
    assert abs(m["ang_theta"][2] - math.pi) < 1e-7          # straight chain B
    assert m["dih"].tolist() == [[0, 1, 2, 3]]
    assert abs(abs(m["dih_phi"][0]) - math.pi) < 1e-12      # planar trans
    assert m["cont"].tolist() == [[0, 5]] and m["cont_type"].tolist() == [I.SSBOND]
    d = np.linalg.norm(m["pos"][5] - m["pos"][0])
    assert m["cont_len"][0] == d
    assert np.allclose(m["cell"], [50 * A, 60 * A, 70 * A], rtol=0, atol=1e-14)


def test_ca_contacts_use_residue_radii():
    # CA atoms 1 and 4 about 7.3 A apart, sequence gap 3: ALA + ALA radii
    # 3.35 A each, limit (6.7 A) (26/7)^(1/6) = 8.25 A -> a backbone-backbone
    # contact; moved to 9.2 A, none
    xyz = [(0.0, 0.0, 0.0), (3.0, 1.5, 0.0), (6.0, 0.0, 0.0), (2.0, 7.0, 0.0)]
    lines = [_atom(k + 1, "CA", "ALA", "A", k + 1, *xyz[k]) for k in range(4)]
    m = I.model_from_pdb("\n".join(lines) + "\nEND\n", contacts="ca")
    r03 = np.linalg.norm(m["pos"][3] - m["pos"][0]) / A
    assert r03 < 6.7 * (26 / 7) ** (1 / 6)
    assert m["cont"].tolist() == [[0, 3]] and m["cont_type"].tolist() == [I.BACK_BACK]
    far = I.model_from_pdb("\n".join(lines[:3] + [_atom(4, "CA", "ALA", "A", 4, 2.0, 9.0, 0.0)])
                           + "\nEND\n", contacts="ca")
    assert len(far["cont"]) == 0


def test_seq_file_with_maps(tmp_path):
    (tmp_path / "part.map.yml").write_text(
        "contacts: |-\n  i1,i2,length\n  0,4,1.25\n"
        "angles: |-\n  i1,i2,i3,theta\n  0,1,2,1.9\n"
        "dihedrals: |-\n  i1,i2,i3,i4,phi\n  0,1,2,3,-2.5\n")
    (tmp_path / "seq.yml").write_text(
        "model:\n  chains:\n"
        "    - seq: |\n        ACDE\n        FG\n"
        "      maps:\n        - (at path): part.map.yml\n"
        "        - __source:\n            (at path): part.map.yml\n          shift: 1\n"
        "    - seq: KLM\n")
    m = I.model_from_seq_file(str(tmp_path / "seq.yml"))
    assert m["n"] == 9 and m["chain"].tolist() == [0] * 6 + [1] * 3
    assert m["atype"][0] == P.LETTER_INDEX["A"] and m["atype"][6] == P.LETTER_INDEX["K"]
    assert m["cont"].tolist() == [[0, 4], [1, 5]] and m["cont_len"].tolist() == [1.25, 1.25]
    # map angles first, then every consecutive triple no map gave
    assert m["ang"].tolist()[:2] == [[0, 1, 2], [1, 2, 3]]
    assert m["ang_has"].tolist() == [1, 1, 0, 0, 0]
    assert m["ang"].tolist()[2:] == [[2, 3, 4], [3, 4, 5], [6, 7, 8]]
    assert m["dih"].tolist() == [[0, 1, 2, 3], [1, 2, 3, 4], [2, 3, 4, 5]]
    assert m["dih_phi"].tolist() == [-2.5, -2.5, 0.0]
    assert m["teth"].tolist() == [[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [6, 7], [7, 8]]


@need_lib
@pytest.mark.parametrize("name", ["9aac", "1ubq"])
def test_library_saw_equals_the_oracle_saw(name):
    from oracle_ffi import Oracle
    p, m, seed, saw, fx = case(name)
    o = Oracle(p, m, seed=seed, saw=saw)
    got, draws = I.morph_into_saw(m, seed, bond=None if name == "1ubq" else saw["bond"],
                                  intersection=saw["intersection"])
    assert np.array_equal(got["pos"], o.pos0)
    assert draws == o.saw_draws()
    assert np.array_equal(got["orig_pos"], o.pos0 if saw["update_orig"] else m["orig_pos"])


@need_lib
def test_library_saw_multichain_box():
    from oracle_ffi import Oracle
    ch = np.repeat(np.arange(4, dtype=np.int32), [1, 30, 2, 45])
    m = M.synthetic_protein(78, 5)
    m["chain"] = ch
    m["seq"] = np.concatenate([np.arange(k) for k in (1, 30, 2, 45)]).astype(np.int32)
    side = 40.0 * A
    got, draws = I.morph_into_saw(m, 9, bond=3.8 * A, intersection=4.56 * A, box_size=side,
                                  pbc=True)
    o = Oracle(P.default_params(), m, seed=9,
               saw=dict(bond=3.8 * A, intersection=4.56 * A, box_size=side, pbc=True,
                        update_orig=True))
    assert np.array_equal(got["pos"], o.pos0) and draws == o.saw_draws()
    pos = got["pos"]
    for c in range(4):
        q = pos[ch == c]
        if len(q) > 1:
            assert np.allclose(np.linalg.norm(np.diff(q, axis=0), axis=1), 3.8 * A, rtol=1e-12)
        assert (np.abs(q[0]) <= side / 2).all()
        d = q[:, None] - q[None]
        d -= side * np.round(d / side)
        iu = np.triu_indices(len(q), k=3)
        assert len(iu[0]) == 0 or np.linalg.norm(d, axis=2)[iu].min() >= 4.56 * A
    # density gives the box side cbrt(n / density)
    g2, _ = I.morph_into_saw(m, 9, bond=3.8 * A, density=78 / side ** 3, pbc=True)
    assert np.allclose(g2["cell"], side, rtol=1e-12)


@need_lib
def test_saw_argument_errors():
    with pytest.raises(G.PcgError):
        G.morph_saw([0, 0, 1, 0], 1, 3.8 * A, 4.56 * A)        # chains not contiguous
    with pytest.raises(G.PcgError):
        G.morph_saw([0, 0, 0], 1, 3.8 * A, 4.56 * A, pbc=True)  # pbc without a box
    with pytest.raises(G.PcgError):
        G.morph_saw([0, 0, 0], 1, 0.0, 4.56 * A)                # no bond length
