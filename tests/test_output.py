"""The reference's text outputs (pas_cg_amd.output): the SL4 scalars table
of the `.out` file and the raw-data dump.

CPU: the layout by hand (sl4.cpp:15-123, print_raw_data.cpp:4-9) and the
unit conversion of the report row. GPU: the raw dump of an FP64 ran2 run
of the TEST fixture carries, at its dumped steps, the reference's own V
trace (parity/tests/TEST/comp.csv, the C++ column, which is that program's
raw dump) and the step's time; a `.out` text written from live reports reads
back to the reports.
"""
import io

import pytest

from pas_cg_amd import output as O
from pas_cg_amd import params as P


def test_sl4_table_layout():
    rows = [{"TIME[tau]": 0.5, "EPOT[eps]": -12.25, "ICN": 7},
            {"TIME[tau]": 1000.0, "EPOT[eps]": -3.0e-7, "ICN": 112}]
    # widths 9, 9, 3 (the headers are the widest); rows start with a space
    want = ("#TIME[TAU]   EPOT[EPS]   ICN\n"
            "       0.5      -12.25     7\n"
            "      1000      -3e-07   112")
    assert O.sl4_table(rows) == want
    assert O.out_file([rows[:1], rows[1:]]).splitlines()[0] == "#TRAJECTORY 0"
    assert O.out_file([rows[:1], rows[1:]]).splitlines()[3] == "#TRAJECTORY 1"


def test_raw_data_line():
    assert O.raw_data_line(0, 100, 0.5, -1.0 / 3.0) == "0 100 0.5 -0.33333333333333331\n"
    assert O.raw_data_line(2, 0, 0.0, 1e-20) == "2 0 0 9.9999999999999995e-21\n"


def test_scalars_row_units_and_columns():
    rep = {"TIME": 3.0, "EPOT": -2.0, "ETOT": -1.5, "EKIN": 0.5, "RG": 1.0, "L": 2.0,
           "RMSD": 0.2, "W": 1.0, "ICN": 4.0, "ICNss": 1.0, "INTEHC": 2.0, "INTRHC": 3.0}
    rep.update({k: float(i) for i, k in enumerate(O.PID_COLS)})
    row = O.scalars_row(rep, nat_cont=True, qa=False, pid=True)
    assert list(row) == ["TIME[tau]", "EPOT[eps]", "ETOT[eps]", "RG[A]", "L[A]", "RMSD[A]",
                         "W[A**2]", "ICN", "ICNss"] + O.PID_COLS
    assert row["RG[A]"] == 1.0 / P.ANGSTROM and row["W[A**2]"] == 1.0 / P.ANGSTROM ** 2
    assert isinstance(row["ICN"], int) and row["PID-S1-S2"] == 5


@pytest.mark.gpu
def test_raw_dump_carries_the_reference_trace():
    from golden_cases import case
    from pas_cg_amd import pcg as G
    p, m, seed, saw, fx = case("TEST")
    c = G.Context(p, m, precision=G.FP64, noise=G.RAN2, seed=seed)
    buf = io.StringIO()
    O.run_with_raw_dump(c, 95, 20, buf, traj_idx=3)
    lines = buf.getvalue().splitlines()
    assert [int(ln.split()[1]) for ln in lines] == [0, 20, 40, 60, 80]
    ref = fx["trace"]["V_cxx"]
    dt = p["dt"]
    for ln in lines:
        traj, step, t, V = ln.split()
        assert traj == "3"
        assert abs(float(t) - int(step) * dt) < 1e-12
        assert abs(float(V) - ref[int(step)]) <= 1e-8 * abs(ref[int(step)]), ln
    assert c.time()[1] == 95
    c.close()


@pytest.mark.gpu
def test_out_file_reads_back_to_the_reports():
    from pas_cg_amd import model as M
    from pas_cg_amd import pcg as G
    m = M.synthetic_protein(60, M.job_seed(3))
    prm = P.default_params()
    c = G.Context(prm, m, precision=G.FP64, noise=G.PHILOX, seed=3)
    rows, reps = [], []
    for _ in range(3):
        c.step(100)
        reps.append(c.report())
        rows.append(O.scalars_row(reps[-1], nat_cont=True, qa=bool(prm["qa_on"]), pid=False))
    text = O.out_file([rows])
    lines = text.splitlines()
    head = lines[1][1:].split()
    assert head[:3] == ["TIME[TAU]", "EPOT[EPS]", "ETOT[EPS]"]
    for rep, ln in zip(reps, lines[2:]):
        vals = dict(zip(head, ln.split()))
        assert float(vals["TIME[TAU]"]) == pytest.approx(rep["TIME"], rel=1e-5)
        assert float(vals["EPOT[EPS]"]) == pytest.approx(rep["EPOT"], rel=1e-5)
        assert float(vals["RG[A]"]) == pytest.approx(rep["RG"] / P.ANGSTROM, rel=1e-5)
        assert int(vals["ICN"]) == int(rep["ICN"])
    c.close()


def test_pdb_snapshots_layout_and_round_trip():
    """The `.pdb` snapshots of make_report::emit_pdb: fixed columns of the
    reference's record writers (records.cpp), and a file the PDB reader
    reads back to the same C-alpha positions (to the 3 printed decimals)."""
    import numpy as np
    from pas_cg_amd import inputs as I
    from pas_cg_amd import model as M
    m = M.replicate(M.synthetic_protein(12, 3), 2, 60 * P.ANGSTROM, 3)
    snaps = [(0.0, m["pos"]), (2.5, m["pos"] + 0.01)]
    txt = O.pdb_snapshots(snaps, m["chain"], m["atype"], m["cell"])
    lines = txt.split("\n")
    assert all(len(ln) == 80 for ln in lines) and not txt.endswith("\n")
    assert lines[0].startswith("MODEL        1") and lines[-1].startswith("ENDMDL")
    assert lines[1][:33] == "CRYST1   60.000   60.000   60.000"
    assert lines[1][55:70] == "        P 1   1"
    atom = next(ln for ln in lines if ln.startswith("ATOM"))
    assert atom[6:11] == "    1" and atom[12:16] == "  CA" and atom[21] == "A" and atom[22:26] == "   1"
    x = m["pos"][0] / P.ANGSTROM
    assert atom[30:54] == "%8.3f%8.3f%8.3f" % tuple(x)
    ter = [ln for ln in lines if ln.startswith("TER")]
    assert len(ter) == 4 and ter[0][6:11] == "   13" and ter[0][22:26] == "  12"
    rem = [ln for ln in lines if ln.startswith("REMARK")]
    assert len(rem) == 8 and "T[TAU]" in rem[0] and "RG[A]" in rem[0]
    st = I.read_pdb(txt)   # the first model
    assert list(st.chains) == ["A", "B"]
    got = np.array([a[2] for res in st.chains["A"] + st.chains["B"] for a in res["atoms"]])
    assert np.abs(got - m["pos"]).max() < 6e-4 * P.ANGSTROM
    # the REMARK row carries the chain's radius of gyration in A
    vals = rem[1][11:].split()
    r = m["pos"][m["chain"] == 0]
    rg = np.sqrt(((r - r.mean(axis=0)) ** 2).sum() / 12.0) / P.ANGSTROM
    assert abs(float(vals[1]) - rg) <= 1e-5 * rg
