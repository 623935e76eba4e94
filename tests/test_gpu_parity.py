"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

* lists (NL pairs with their distances, the partition into native contacts,
  QA free pairs, PID, DH and Pauli lists) bit-exact on identical positions;
* every force term on identical positions and lists: FP64 mode to 1e-9 of
  the largest force component, mixed mode (FP32 radial pair potentials) to
  the bound written below;
* whole trajectories in FP64 with the reference's ran2 noise stream against
  the oracle and against the reference's golden V traces;
* contact-state transitions (active flags, change times) bit-exact.
"""
import numpy as np
import pytest

from golden_cases import case
from lockstep import qa_lockstep
from oracle_ffi import Oracle
from pas_cg_amd import inputs as I
from pas_cg_amd import model as M
from pas_cg_amd import params as P
from pas_cg_amd import pcg as G

pytestmark = pytest.mark.gpu

# mixed precision bounds (north star: force error <= 1e-5, energy <= 1e-6)
MIXED_F_TOL = 1e-5
MIXED_V_TOL = 1e-6
FP64_TOL = 1e-9


def full_params(**kw):
    p = P.default_params()
    p.update(chir_on=1, pid_on=1, afm_on=1)
    p.update(kw)
    return p


def protein(n=90, seed=5, noise_A=0.3):
    m = M.synthetic_protein(n, M.job_seed(seed))
    rng = np.random.default_rng(seed)
    m["pos"] = m["pos"] + rng.normal(scale=noise_A * P.ANGSTROM, size=m["pos"].shape)
    m["atype"][::7] = P.AA_INDEX["LYS"]
    m["atype"][3::11] = P.AA_INDEX["GLU"]
    m["atype"][5::13] = P.AA_INDEX["PRO"]
    return M.add_pulling(m)


def multichain(copies=4, n=40, side_A=70.0, seed=9):
    one = M.synthetic_protein(n, M.job_seed(seed))
    one["atype"][::5] = P.AA_INDEX["ARG"]
    one["atype"][2::6] = P.AA_INDEX["ASP"]
    m = M.replicate(one, copies, side_A * P.ANGSTROM, seed)
    return m


CONFIGS = {
    "protein": lambda: (full_params(qa_on=0), protein()),
    "protein_sink_pid": lambda: (
        _sink(full_params(qa_on=0, nat_cont_on=0, dh_variant=1,
                          dh_perm=4.0 * P.EPS_0 / P.ANGSTROM)), protein(seed=6)),
    "multichain": lambda: (_sink(full_params(qa_on=0, nat_cont_on=0)), multichain()),
    "multichain_qa": lambda: (full_params(qa_on=1, pid_on=0), multichain(seed=3)),
    "protein_force_afm": lambda: (full_params(qa_on=0),
                                  M.add_force_pulling(protein(seed=8), force_eps_per_A=2.5)),
    "protein_walls": lambda: _walled(protein(seed=12)),
}


def _walled(m, gap_A=4.0):
    """Solid walls on the faces of the model's bounding box widened by gap_A."""
    prm = full_params(qa_on=0)
    P.set_solid_walls(prm, m["pos"].min(axis=0) - gap_A * P.ANGSTROM,
                      m["pos"].max(axis=0) + gap_A * P.ANGSTROM)
    return prm, m


def _sink(p):
    P.set_pid_forces(p, "sink lj", bbp_depth=0.3, bbm_depth=0.25,
                     ss_depth_table=[[0.1 + 0.01 * ((a * 7 + b) % 13) for b in range(21)]
                                     for a in range(21)])
    return p


def same_state(cfg, precision, steps=8):
    prm, m = CONFIGS[cfg]()
    o = Oracle(prm, m, seed=17)
    o.set_noise_scale(0.0)
    o.step(steps)
    c = G.Context(prm, m, precision=precision, noise=G.RAN2, seed=17)
    c.set(0, o.r())
    c.set_time(o.t)
    return prm, m, o, c


@pytest.mark.parametrize("cfg", list(CONFIGS))
def test_lists_bit_exact(cfg):
    prm, m, o, c = same_state(cfg, G.FP64)
    o.fix_nl()
    c.fix_nl()
    for kind in range(6):  # NL, native contacts, pauli, dh, pid, qa free pairs
        a = o.list(kind)
        g = c.list(kind)
        assert len(a[0]) == len(g[0]), (cfg, kind, len(a[0]), len(g[0]))
        assert np.array_equal(a[0], g[0]) and np.array_equal(a[1], g[1]), (cfg, kind)
        if kind == 0:
            assert np.array_equal(a[2], g[2]), "NL distances not bit-exact"
    assert len(o.list(0)[0]) > 0


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("cfg", ["protein", "protein_sink_pid", "multichain", "protein_force_afm",
                                 "protein_walls"])
def test_terms_match_oracle(cfg, precision):
    prm, m, o, c = same_state(cfg, precision)
    o.fix_nl()
    c.fix_nl()
    ftol = FP64_TOL if precision == G.FP64 else MIXED_F_TOL
    vtol = FP64_TOL if precision == G.FP64 else MIXED_V_TOL
    checked = 0
    for term in G.TERMS:
        if term.startswith("qa"):
            continue
        F1, V1 = o.eval_term(term)
        F2, V2 = c.eval_term(term)
        scale = np.abs(F1).max()
        if scale == 0 and V1 == 0:
            assert np.abs(F2).max() == 0 and V2 == 0, term
            continue
        checked += 1
        err = np.abs(F2 - F1).max() / scale
        assert err <= ftol, (term, err)
        assert abs(V2 - V1) <= vtol * max(abs(V1), 1e-300) + 1e-12, (term, V1, V2)
    assert checked >= 5


def test_contact_transitions_bit_exact():
    prm = full_params(qa_on=0)
    m = protein(noise_A=0.8)
    o = Oracle(prm, m, seed=21)
    o.set_noise_scale(0.0)
    c = G.Context(prm, m, precision=G.MIXED, noise=G.RAN2, seed=21)
    c.set_noise_scale(0.0)
    k = len(o.static_list(7)[0])
    for _ in range(5):
        o.step(7)
        # GPU evaluates on the oracle's positions and time
        c.set(0, o.r())
        c.set_time(o.t)
        o.fix_nl()
        c.fix_nl()
        F1, _ = o.eval_term("nat_cont")
        c.eval_term("nat_cont")
        a1, t1 = o.nat_cont_state(k)
        a2, t2, _ = c.nat_cont_state(k)
        assert np.array_equal(a1, a2)


@pytest.mark.parametrize("name", ["start2", "TEST", "input_test", "9aac", "1ubq"])
def test_fp64_ran2_trajectory_matches_golden(name):
    """FP64 + the reference's ran2 stream: the GPU trajectory follows the
    oracle and the reference's own V trace step by step."""
    p, m, seed, saw, fx = case(name)
    o = Oracle(p, m, seed=seed, saw=saw)
    skip = 0
    if saw:  # input preparation by the library's SAW (pcg_morph_saw)
        m, skip = I.morph_into_saw(m, seed, bond=saw["bond"], intersection=saw["intersection"],
                                   update_orig=saw["update_orig"])
        assert np.array_equal(m["pos"], o.pos0) and skip == o.saw_draws()
    steps = 100
    ref = np.asarray(fx["trace"]["V_fort_7" if name == "start2" else "V_cxx"][:steps])
    c = G.Context(p, m, precision=G.FP64, noise=G.RAN2, seed=seed, ran2_skip=skip)
    tr_o = o.step(steps)
    tr_g = np.zeros(steps)
    for k in range(steps):
        c.step(1)
        tr_g[k] = c.energy()[0]
    rel_o = np.abs(tr_g - tr_o) / np.abs(tr_o)
    rel_ref = np.abs(tr_g - ref) / np.abs(ref)
    assert rel_o.max() < 1e-9, (int(rel_o.argmax()), rel_o.max())
    assert rel_ref.max() < 1e-8, (int(rel_ref.argmax()), rel_ref.max())


def test_qa_trajectory_fp64_matches_oracle():
    """QA on: free pairs, contacts, saturations and sync budgets follow the
    oracle over a noise-free trajectory in FP64."""
    prm = full_params(qa_on=1, pid_on=0, temperature=0.3)
    m = multichain(copies=3, n=50, side_A=45.0, seed=31)
    o = Oracle(prm, m, seed=31)
    o.set_noise_scale(0.0)
    c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=31)
    c.set_noise_scale(0.0)
    formed = 0
    for _ in range(6):
        o.step(20)
        c.step(20)
        assert np.abs(c.r() - o.r()).max() < 1e-8
        so, no = o.sync(), o.num_qa_contacts()
        sg, ng = c.qa_state()
        assert no == ng and np.array_equal(so, sg)
        ci = o.list(6)
        gi = c.list(6)
        assert np.array_equal(ci[0], gi[0]) and np.array_equal(ci[1], gi[1])
        assert np.allclose(ci[2], gi[2], rtol=0, atol=1e-12)
        # energy of the last step, QA contact term included
        assert abs(c.energy()[0] - o.V) <= 1e-9 * abs(o.V), (c.energy()[0], o.V)
        formed = max(formed, no)
    assert formed > 0


def test_eval_term_slices_sum_to_whole():
    prm, m, o, c = same_state("protein", G.FP64)
    c.fix_nl()
    n = c.term_size("tether")
    F_all, V_all = c.eval_term("tether")
    F_a, V_a = c.eval_term("tether", 0, n // 3)
    F_b, V_b = c.eval_term("tether", n // 3, n)
    assert np.allclose(F_a + F_b, F_all, atol=1e-12)
    assert abs(V_a + V_b - V_all) < 1e-10 * abs(V_all)


def test_c1_box_per_term_errors_mixed():
    """64 copies of a 100-residue chain in a periodic box at 1 mM with Debye-
    Hueckel, QA and PID (the plain path, stage kernels): after a Langevin
    run, every term in mixed precision against the FP64 oracle on the same
    positions, to 1e-5 in force and 1e-6 in energy; then the QA terms with
    the rest, step by step on the oracle's states through QA finishes."""
    from pas_cg_amd import workloads as W
    prm, m = W.c1()
    c = G.Context(prm, m, precision=G.MIXED, noise=G.PHILOX, seed=5)
    c.use_graph(True)
    c.step(300)
    r = c.r()
    c_time = c.time()[0]
    o = Oracle(prm, m, seed=5)
    o.set_r(r)
    o.set_t(c.time()[0])
    o.fix_nl()
    o.prepare_nh()
    c2 = G.Context(prm, m, precision=G.MIXED, noise=G.PHILOX, seed=5)
    c2.set(0, r)
    c2.set_time(c.time()[0])
    c2.fix_nl()
    c2.prepare_nh()
    checked = 0
    for term in G.TERMS:
        if term.startswith("qa"):
            continue
        F1, V1 = o.eval_term(term)
        F2, V2 = c2.eval_term(term)
        scale = np.abs(F1).max()
        if scale == 0:
            assert np.abs(F2).max() == 0, term
            continue
        checked += 1
        assert np.abs(F2 - F1).max() / scale <= MIXED_F_TOL, term
        assert abs(V2 - V1) <= MIXED_V_TOL * abs(V1) + 1e-12, (term, V1, V2)
    assert checked >= 5
    c.close()
    c2.close()
    worst, _ = qa_lockstep(prm, m, G.MIXED, 4, r0=r, t0=c_time, seed=5)
    assert worst["total"] <= MIXED_F_TOL, worst
    assert worst["V_qa_cand"] <= MIXED_V_TOL and worst["V_qa_cont"] <= MIXED_V_TOL, worst


@pytest.mark.parametrize("cfg", ["multichain", "multichain_qa", "protein_force_afm",
                                 "protein_walls"])
def test_batch_periodic_multichain_follows_oracle(cfg):
    """The persistent kernel on a periodic box of several chains (wrapped
    pairs; inter-chain Pauli, DH and PID, or QA with its sync budgets)
    follows the oracle over a noise-free FP64 trajectory."""
    prm, m = CONFIGS[cfg]()
    steps = 60
    b = G.Batch(prm, [m], seeds=[1], precision=G.FP64, steps_per_launch=20)
    b.set_noise_scale(0.0)
    b.run(steps)
    o = Oracle(prm, m, seed=1)
    o.set_noise_scale(0.0)
    tr = o.step(steps)
    err = np.abs(b.get(0, 0) - o.r()).max()
    assert err < 1e-8, err
    V, _ = b.energy(0)
    assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (V, tr[-1])
    b.close()


def test_plain_context_tiny_systems_follow_oracle():
    """The drop-in context on systems of 2 to 5 residues (no pairs, chain
    ends everywhere) follows the oracle with the reference's ran2 noise."""
    prm = full_params(qa_on=1)
    for n in (2, 3, 4, 5):
        m = M.synthetic_protein(n, M.job_seed(800 + n))
        o = Oracle(prm, m, seed=3)
        c = G.Context(prm, m, precision=G.FP64, noise=G.RAN2, seed=3)
        o.step(40)
        c.step(40)
        assert np.abs(c.r() - o.r()).max() < 1e-9, n
        assert abs(c.energy()[0] - o.V) <= 1e-9 * max(abs(o.V), 1e-12), (n, c.energy()[0], o.V)
        c.close()


def _qa_survivors(m, prm):
    """Free pairs (not native, |i - j| >= 3) within their QA first-test
    radius min(max r_high, max(rep_cutoff, req)) at the starting positions."""
    rh = np.asarray(prm["qa_lj_r_high"])
    r, at = m["pos"], m["atype"]
    i, j = np.triu_indices(len(r), k=3)
    d = np.linalg.norm(r[i] - r[j], axis=1)
    nat = set(map(tuple, np.asarray(m["cont"]).tolist()))
    free = np.array([(a, b) not in nat for a, b in zip(i, j)])
    req = np.maximum(rh[1:4].max(), rh[4 + at[i] + 21 * at[j]]) * (1 + prm["qa_formation_tol"])
    thr = np.minimum(rh.max(), np.maximum(prm["rep_cutoff"], req))
    return int((free & (d <= thr)).sum())


@pytest.mark.parametrize("block", [1024, 256])
def test_batch_large_qa_protein_follows_oracle(block):
    """A 600-residue protein with QA on, in the persistent kernel at 16 and 4
    waves. About 1000 pairs pass the first QA test at the start: with 4 waves
    every wave fills and drains its 64-pair survivor stack several times a
    step, with 16 waves about once. Noise-free FP64 positions, energy and the
    QA contact count follow the oracle."""
    prm = full_params(qa_on=1, pid_on=0)
    m = M.synthetic_protein(600, M.job_seed(41))
    assert _qa_survivors(m, prm) > 2 * 64 * 4
    steps = 40
    b = G.Batch(prm, [m], seeds=[5], precision=G.FP64, block=block, steps_per_launch=20)
    b.set_noise_scale(0.0)
    b.run(steps)
    o = Oracle(prm, m, seed=5)
    o.set_noise_scale(0.0)
    tr = o.step(steps)
    assert np.abs(b.get(0, 0) - o.r()).max() < 1e-8
    V, _ = b.energy(0)
    assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (V, tr[-1])
    rep = b.report()[0]
    assert rep["INTRHC"] == o.num_qa_contacts() and o.num_qa_contacts() > 0
    b.close()
