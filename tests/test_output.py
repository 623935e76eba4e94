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


def test_sl4_cells_and_empty_tables():
    """sl4.cpp:15-44 and table.cpp: a column appears where a row first sets
    it, a row without it leaves a padded blank, a float that is not finite
    does not survive the stream round trip and prints 0, booleans print T/F,
    and a table without rows prints its empty header."""
    import numpy as np
    assert O.sl4_table([]) == "#"
    rows = [{"a": 1, "b": True}, {"a": np.int32(22), "c": float("nan"), "b": np.bool_(False)},
            {"a": 3, "c": -float("inf")}]
    assert O.sl4_table(rows) == ("# A   B   C\n"
                                 "  1   T    \n"
                                 " 22   F   0\n"
                                 "  3       0")
    assert O._comment("T = ", 1e-05) == "#T = 1E-05"
    assert O._comment("n = ", 7) == "#N = 7"


def _map_model():
    import numpy as np
    A = P.ANGSTROM
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1], [3, 0, 0]], float) * A
    m = {"chain": np.array([0, 0, 0, 0, 1]), "seq": np.array([0, 1, 2, 3, 0]),
         "atype": np.array([0, 4, 7, 12, 19], np.uint8),
         "ang": np.array([[0, 1, 2], [1, 2, 3]]), "ang_has": np.array([1, 0]),
         "ang_theta": np.array([2.0, 0.0]), "dih": np.array([[0, 1, 2, 3]]),
         "dih_has": np.array([0]), "dih_phi": np.array([0.0]), "cell": np.zeros(3)}
    return pos, m


def test_map_div_layout():
    """make_report::add_map_data (make_report.cpp:499-656) on a hand-built
    system with closed-form geometry: the bond vectors x, y, z are pairwise
    perpendicular, so every angle between consecutive bonds and the
    dihedral are pi/2 (positive: r12 x r23 points along r34); contact
    distances sqrt(2), 2, sqrt(3) and sqrt(5) A; sequence numbers from 1;
    a same-chain pair labelled "inter" as the reference does; only active
    PID bundles listed; a native value only where the model has one."""
    pos, m = _map_model()
    nat = {"i1": [0, 1], "i2": [2, 4], "type": [1, 5], "active": [True, False],
           "change_t": [-1.0, 3.5]}
    qa = {"i1": [0, 2], "i2": [3, 4], "type": [1, 30]}
    pid = {"i1": [0, 1], "i2": [3, 4], "active": [True, False]}
    want = """#T = 12.5
#NATIVE CONTACTS
#N = 1
#CHAIN1   SEQ1   RES1   CHAIN2   SEQ2   RES2   INTER/INTRA     TYPE   DIST[A]   ACTIVE   TIME[TAU]
      A      1    ALA        A      3    GLY         inter       bb   1.41421        T          -1
      A      2    CYS        B      1    VAL         intra   ssbond         2        F         3.5
#QA
#N = 2
#CHAIN1   SEQ1   RES1   CHAIN2   SEQ2   RES2   INTER/INTRA   TYPE   DIST[A]
      A      1    ALA        A      4    MET         inter     bb   1.73205
      A      3    GLY        B      1    VAL         intra     ss   2.23607
#PID
#N = 1
#CHAIN1   SEQ1   RES1   CHAIN2   SEQ2   RES2   INTER/INTRA   DIST[A]
      A      1    ALA        A      4    MET         inter   1.73205
#ANGLES
#I1   CHAIN1   SEQ1   RES1    THETA   NATIVE
  0        A      1    ALA   1.5708        2
  1        A      2    CYS   1.5708         
#DIHEDRALS
#I1   CHAIN1   SEQ1   RES1      PHI
  0        A      1    ALA   1.5708"""
    assert O.map_div(12.5, pos, m, nat=nat, qa=qa, pid=pid) == want
    # two trajectories, the second with a 2.5 A periodic x axis: the 2 A
    # pair's nearest image is 0.5 A away
    txt = O.map_file([[O.map_div(0.0, pos, m)],
                      [O.map_div(1e-5, pos, m, cell=[2.5 * P.ANGSTROM, 0, 0], nat=nat)]])
    lines = txt.split("\n")
    assert lines[0] == "#TRAJECTORY 0" and lines[1] == "#T = 0"
    k = lines.index("#TRAJECTORY 1")
    assert lines[k + 1] == "#T = 1E-05" and lines[k + 2] == "#NATIVE CONTACTS"
    assert lines[k + 6].split()[8] == "0.5"


def test_map_geometry_closed_forms():
    """angle_value is the angle between consecutive bonds (pi minus the bond
    angle) and dihedral_value takes the sign of (r12 x r23) . r34."""
    import math
    import numpy as np
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, -1], [2, 1, 0]], float)
    assert O.map_angles(pos, [[0, 1, 2]])[0] == pytest.approx(math.pi / 2, abs=1e-15)
    assert O.map_angles(pos, [[0, 1, 0]])[0] == pytest.approx(math.pi, abs=1e-7)
    assert O.map_dihedrals(pos, [[0, 1, 2, 3]])[0] == pytest.approx(-math.pi / 2, abs=1e-15)
    # a planar trans chain: phi = pi; cis: 0
    trans = np.array([[0, 1, 0], [0, 0, 0], [1, 0, 0], [1, -1, 0]], float)
    cis = np.array([[0, 1, 0], [0, 0, 0], [1, 0, 0], [1, 1, 0]], float)
    assert abs(O.map_dihedrals(trans, [[0, 1, 2, 3]])[0]) == pytest.approx(math.pi, abs=1e-7)
    assert O.map_dihedrals(cis, [[0, 1, 2, 3]])[0] == pytest.approx(0.0, abs=1e-7)
    # C round: a separation of exactly half the cell goes to the image away
    # from zero (|d| unchanged), as round() does and numpy's even rounding
    # would not for 2.5 cells
    assert O._c_round(np.array([2.5, -2.5, 0.5, 1.5, 0.49]))[:5].tolist() == [3, -3, 1, 2, 0]


def test_results_div_unfolding_times():
    """make_report::at_simul_end (make_report.cpp:77-128): per contact the
    mean change time over the trajectories that changed it (-1 if none),
    rows sorted by (time, index); the median entry num_traj // 2 of the
    sorted unfolding times, N/A when fewer trajectories unfolded."""
    import numpy as np
    A = P.ANGSTROM
    contacts = {"i1": [0, 2, 5], "i2": [4, 9, 11], "type": [0, 5, 1],
                "nat_dist": [5.0 * A, 6.5 * A, 4.25 * A]}
    times = [np.array([2.0, -1.0, 1.0]), np.array([4.0, -1.0, 0.5]), np.array([3.0, -1.0, -1.0])]
    ct0, u0 = O.unfolding_sample(times[0], 2)
    assert u0 is None and ct0.tolist() == [2.0, -1.0, 1.0]
    _, u1 = O.unfolding_sample(np.array([4.0, 7.0, 0.5]), 3)
    assert u1 == 7.0
    txt = O.results_div(contacts, times, [9.0, 7.0], num_traj=3)
    want = """#RESULTS
#AVERAGE BREAKING/FORMING TIMES
#I1   I2     TYPE   NAT_DIST[A]   AVG_TIME[TAU]
  2    9   ssbond           6.5              -1
  5   11       bb          4.25            0.75
  0    4      any             5               3
#MEDIAN (UN)FOLDING TIME: 9"""
    assert txt == want
    assert O.results_div(contacts, times, [9.0], num_traj=3).endswith("#MEDIAN (UN)FOLDING TIME: N/A")
    assert O.results_div() == "#RESULTS"


def test_afm_tables():
    """make_report::add_afm_stuff (make_report.cpp:264-311): the norms of the
    averages in A/tau and eps/A, -1 before a window closed; a resting
    velocity tip's perpendicular force is NaN and prints 0."""
    import numpy as np
    A = P.ANGSTROM
    stats = {"avg": np.array([[3.0 / A, 4.0 / A, 0.0], [1.0, 0.0, 0.0], [0.0, 0.6 * A, 0.8 * A]]),
             "perp": np.array([float("nan"), 0.5, 0.0]), "has": np.array([True, False, True])}
    tips = {"tip_res": [0, 11], "ftip_res": [7]}
    f, v = O.afm_rows(stats, tips)
    assert f == [{"res_idx": 7, "avg_vel [A/tau]": pytest.approx(1.0, rel=1e-15)}]
    assert v[0]["avg_force [eps/A]"] == pytest.approx(5.0, rel=1e-15)
    assert v[1] == {"res_idx": 11, "avg_force [eps/A]": -1.0, "avg_perp_force [eps/A]": -1.0}
    txt = O.out_file([[{"TIME[tau]": 1.0}]], afm=[(f, v)])
    assert txt.split("\n") == ["#TRAJECTORY 0", "#TIME[TAU]", "         1",
                               "#FORCE AFM TIPS", "#RES_IDX   AVG_VEL [A/TAU]",
                               "       7                 1",
                               "#VELOCITY AFM TIPS",
                               "#RES_IDX   AVG_FORCE [EPS/A]   AVG_PERP_FORCE [EPS/A]",
                               "       0                   5                        0",
                               "      11                  -1                       -1"]
    # no force tips: the force AFM div still comes, with its empty table
    f, v = O.afm_rows(stats, {"tip_res": [0, 11]})
    assert O.out_file([[{"TIME[tau]": 1.0}]], afm=[(f, v)]).split("\n")[3:5] == \
        ["#FORCE AFM TIPS", "#"]


@pytest.mark.gpu
def test_map_of_context_equals_oracle_map():
    """map_snapshot of the plain context, through pcg_nat_cont_state,
    pcg_get_list kind 7 (QA contacts with their types) and kind 4 (PID
    bundles), is the map the oracle's state gives: QA on, after a lockstep
    run in which contacts formed; then PID on, on the oracle's positions."""
    import numpy as np
    from lockstep import qa_lockstep
    from oracle_ffi import Oracle
    from pas_cg_amd import model as M
    from pas_cg_amd import pcg as G
    prm = P.default_params()
    prm.update(qa_on=1, pid_on=0, afm_on=0, temperature=0.3)
    one = M.synthetic_protein(50, M.job_seed(31))
    one["atype"][::5] = P.AA_INDEX["ARG"]
    one["atype"][2::6] = P.AA_INDEX["ASP"]
    m = M.replicate(one, 3, 45.0 * P.ANGSTROM, 31)
    _, formed, o, c = qa_lockstep(prm, m, G.FP64, 60, keep=True)
    assert formed > 0
    k = len(m["cont"])
    act, ct = o.nat_cont_state(k)
    ij, _, ty, _, _, vac = o.qa_contacts()
    live = vac == 0
    gi1, gi2, gty, gvac = c.list(7)
    assert np.array_equal(gvac != 0, ~live)
    assert np.array_equal(gty[live].astype(np.int64), ty[live])
    cont = np.asarray(m["cont"])
    want = O.map_div(o.t, o.r(), m,
                     nat={"i1": cont.min(axis=1), "i2": cont.max(axis=1), "type": m["cont_type"],
                          "active": act, "change_t": ct},
                     qa={"i1": ij[live, 0], "i2": ij[live, 1], "type": ty[live]})
    got = O.map_snapshot(c, m, prm)
    assert got == want
    assert "#QA\n#N = %d\n" % int(live.sum()) in got
    c.close()

    prm = P.default_params()
    prm.update(qa_on=0, pid_on=1)
    m = M.synthetic_protein(80, M.job_seed(12))
    o = Oracle(prm, m, seed=3)
    o.step(40)
    c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=3)
    c.set(0, o.r())
    c.set_time(o.t)
    o.fix_nl()
    c.fix_nl()
    i1, i2, _, _ = o.list(4)
    active = O.pid_active(o.r(), m["cell"], i1, i2, m["atype"], prm)
    assert active.any()
    want = O.map_div(o.t, o.r(), m, pid={"i1": i1, "i2": i2, "active": active})
    got = O.map_snapshot(c, m, dict(prm, nat_cont_on=0))
    assert got == want
    c.close()


class _FakeCtx:
    """A stand-in for pcg.Context with the calls the report driver makes:
    every step adds dt to the time and moves residue 0."""

    def __init__(self, m, dt):
        import numpy as np
        self.m, self.dt, self.t, self.k = m, dt, 0.0, 0
        self.pos = np.asarray(m["pos"], np.float64).copy()
        self.steps_taken = []

    def step(self, k):
        for _ in range(k):
            self.t += self.dt
            self.k += 1
            self.pos[0, 0] += 0.01
        self.steps_taken.append(k)

    def time(self):
        return self.t, self.k

    def r(self):
        return self.pos.copy()

    def report(self):
        rep = {k: 0.0 for k in ["TIME", "EPOT", "ETOT", "EKIN", "RG", "L", "RMSD", "W", "ICN",
                                "ICNss", "INTEHC", "INTRHC"] + O.PID_COLS}
        rep["TIME"] = self.t
        return rep

    def nat_cont_state(self, k):
        import numpy as np
        return np.zeros(k, np.uint8), np.full(k, -1.0), 0

    def list(self, kind):
        import numpy as np
        z = np.zeros(0, np.int32)
        return z, z, np.zeros(0), np.zeros(0, np.uint8)


def test_reporter_timing_and_files(tmp_path):
    """make_report's timing (make_report.cpp:29-46, called once t > 0): the
    first report one step in, then whenever the time passes the last
    report's plus the interval; the .out rows, .map divs and .pdb models
    follow; a trajectory's end adds one of each, the simulation's end the
    results div."""
    import numpy as np
    from pas_cg_amd import model as M
    m = M.synthetic_protein(20, M.job_seed(3))
    prm = P.default_params()
    prm.update(qa_on=0, pid_on=0, nat_cont_on=1)
    dt = prm["dt"]
    ctx = _FakeCtx(m, dt)
    rep = O.Reporter(str(tmp_path / "run"), m, prm, stats_every=10 * dt, struct_every=25 * dt,
                     measure_times=True)
    rep.run(ctx, 60)
    out = (tmp_path / "run.out").read_text().split("\n")
    rows = [ln for ln in out if ln.startswith(" ")]
    times = [float(ln.split()[0]) for ln in rows]
    # steps 1, 11, 21, 31, 41, 51 (the time passes last + 10 dt on the step
    # that reaches it)
    assert len(times) == 6
    assert np.allclose(times, [dt * k for k in (1, 11, 21, 31, 41, 51)], rtol=1e-9)
    mp = (tmp_path / "run.map").read_text()
    assert mp.startswith("#TRAJECTORY 0\n#T = ") and mp.count("#ANGLES") == 6
    pdb = (tmp_path / "run.pdb").read_text()
    assert pdb.count("MODEL ") == 3  # steps 1, 26, 51
    rep.traj_end(ctx)
    rep.simul_end()
    out = (tmp_path / "run.out").read_text()
    assert out.split("#RESULTS")[0].count("\n ") == 7
    assert (tmp_path / "run.pdb").read_text().count("MODEL ") == 4
    assert out.rstrip().split("\n")[-1] == "#MEDIAN (UN)FOLDING TIME: N/A"
    assert "#AVERAGE BREAKING/FORMING TIMES" in out
    assert sum(ctx.steps_taken) == 60


@pytest.mark.gpu
def test_reporter_on_a_context(tmp_path):
    """The report files of a live context with QA and a pulled protein: each
    `.map` div's QA count equals the INTEHC + INTRHC of the `.out` row of
    the same report (both count the live QA contacts of one state), the AFM
    tables get a row a tip a report, and the `.pdb` holds every snapshot."""
    from pas_cg_amd import model as M
    from pas_cg_amd import pcg as G
    m = M.add_pulling(M.synthetic_protein(80, M.job_seed(5)), vel_A_per_tau=0.5)
    prm = P.default_params()
    prm.update(qa_on=1, pid_on=0, afm_on=1, afm_stats_window=20 * prm["dt"])
    c = G.Context(prm, m, precision=G.MIXED, noise=G.PHILOX, seed=5)
    tips = {k: m[k] for k in ("tip_res", "tip_orig", "tip_vel")}
    rep = O.Reporter(str(tmp_path / "p"), m, prm, stats_every=100 * prm["dt"],
                     struct_every=200 * prm["dt"])
    rep.run(c, 400, tips)
    out = (tmp_path / "p.out").read_text().split("\n")
    head = out[1][1:].split()
    rows = []
    for ln in out[2:]:
        if not ln.startswith(" "):
            break
        rows.append(dict(zip(head, ln.split())))
    # the reference's timing: t accumulated step by step, a report once
    # t >= last + interval (the fourth lands a step late by rounding)
    t, last, due = 0.0, -float("inf"), []
    for k in range(1, 401):
        t += prm["dt"]
        if t >= last + 100 * prm["dt"]:
            due.append(k)
            last = t
    assert len(rows) == len(due) and len(due) >= 4
    assert [float(r["TIME[TAU]"]) for r in rows] == pytest.approx([k * prm["dt"] for k in due],
                                                                   rel=1e-5)
    mp = (tmp_path / "p.map").read_text().split("\n")
    qa_counts = [int(mp[k + 1].split("=")[1]) for k, ln in enumerate(mp) if ln == "#QA"]
    assert qa_counts == [int(r["INTEHC"]) + int(r["INTRHC"]) for r in rows]
    k = out.index("#VELOCITY AFM TIPS")
    assert len([ln for ln in out[k + 2:] if ln.startswith(" ")]) == 2 * len(rows)
    assert (tmp_path / "p.pdb").read_text().count("MODEL ") == 2
    c.close()
