"""GPU parity of the term variants a protein can carry, in every path:

* the simple native dihedral (nat_dih_variant 1, angles/nat_dih/simple.cpp);
* heuristic bond angles and dihedrals beside native ones (a model whose
  angles and dihedrals partly lack native values, angles/heur_ang and
  angles/heur_dih);
* disulfide native contacts (contact type SSBOND) with the harmonic and the
  Lennard-Jones disulfide force (nat_cont/eval_forces.cpp:26-31);
* the reference's own parity inputs (9aac, TEST, input_test) in the batched
  kernel;
* the QA terms in mixed precision on identical states, step by step.

Plain context: per term on identical positions and lists (FP64 to 1e-9,
mixed to 1e-5 in force and 1e-6 in energy). Batched kernel: noise-free FP64
trajectories against the oracle, and the mixed kernel's total force with
energies off (so every step takes the FP32 paths) against the FP64 kernel.
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

MIXED_F_TOL = 1e-5
MIXED_V_TOL = 1e-6
FP64_TOL = 1e-9


def _params(**kw):
    p = P.default_params()
    p.update(chir_on=1, pid_on=1, afm_on=1, qa_on=0)
    p.update(kw)
    return p


def _protein(n=90, seed=5, heuristic_frac=0.0, noise_A=0.3, disulfides=0):
    m = M.synthetic_protein(n, M.job_seed(seed), heuristic_frac=heuristic_frac)
    rng = np.random.default_rng(seed)
    m["pos"] = m["pos"] + rng.normal(scale=noise_A * P.ANGSTROM, size=m["pos"].shape)
    m["atype"][::7] = P.AA_INDEX["LYS"]
    m["atype"][3::11] = P.AA_INDEX["GLU"]
    m["atype"][5::13] = P.AA_INDEX["PRO"]
    m["atype"][9::17] = P.AA_INDEX["GLY"]
    if disulfides:
        m = M.add_disulfides(m, disulfides, seed=seed)
    return M.add_pulling(m)


def _ss_lj(p):
    # disulfide bond force "lj": depth 4 eps, r_min 6 A (default.yml)
    p.update(nc_ss_kind=2, nc_ss_depth=4.0, nc_ss_r_min=6.0 * P.ANGSTROM)
    return p


VARIANTS = {
    "snd": lambda: (_params(nat_dih_variant=1), _protein(seed=21)),
    "heuristic": lambda: (_params(), _protein(seed=22, heuristic_frac=0.5)),
    "all_heuristic": lambda: (_params(), _protein(seed=23, heuristic_frac=1.0)),
    "ss_harmonic": lambda: (_params(nc_ss_kind=1), _protein(seed=24, disulfides=4)),
    "ss_lj": lambda: (_ss_lj(_params()), _protein(seed=25, disulfides=4)),
}


def _same_state(prm, m, precision, steps=8, seed=17):
    o = Oracle(prm, m, seed=seed)
    o.set_noise_scale(0.0)
    o.step(steps)
    c = G.Context(prm, m, precision=precision, noise=G.RAN2, seed=seed)
    c.set(0, o.r())
    c.set_time(o.t)
    return o, c


def test_variant_models_carry_their_items():
    """The variant models really hold what the tests below exercise."""
    _, m = VARIANTS["heuristic"]()
    assert 0 < (m["ang_has"] == 0).sum() < len(m["ang_has"])
    assert 0 < (m["dih_has"] == 0).sum() < len(m["dih_has"])
    _, m = VARIANTS["all_heuristic"]()
    assert (m["ang_has"] == 0).all() and (m["dih_has"] == 0).all()
    _, m = VARIANTS["ss_lj"]()
    assert (m["cont_type"] == 5).sum() == 4


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("cfg", list(VARIANTS))
def test_variant_terms_match_oracle(cfg, precision):
    prm, m = VARIANTS[cfg]()
    o, c = _same_state(prm, m, precision)
    o.fix_nl()
    c.fix_nl()
    ftol = FP64_TOL if precision == G.FP64 else MIXED_F_TOL
    vtol = FP64_TOL if precision == G.FP64 else MIXED_V_TOL
    want = {"snd": ["snd"], "heuristic": ["heur_ang", "heur_dih", "nat_ang", "cnd"],
            "all_heuristic": ["heur_ang", "heur_dih"], "ss_harmonic": ["nat_cont"],
            "ss_lj": ["nat_cont"]}[cfg]
    seen = set()
    for term in G.TERMS:
        if term.startswith("qa"):
            continue
        F1, V1 = o.eval_term(term)
        F2, V2 = c.eval_term(term)
        scale = np.abs(F1).max()
        if scale == 0 and V1 == 0:
            assert np.abs(F2).max() == 0 and V2 == 0, term
            continue
        seen.add(term)
        err = np.abs(F2 - F1).max() / scale
        assert err <= ftol, (term, err)
        assert abs(V2 - V1) <= vtol * max(abs(V1), 1e-300) + 1e-12, (term, V1, V2)
    assert set(want) <= seen, (want, seen)


def _nc_ss_energy_differs(prm, m):
    """The disulfide contacts change the native-contact energy: the same
    state with the disulfide force off gives another value."""
    o, c = _same_state(prm, m, G.FP64)
    o.fix_nl()
    _, V_ss = o.eval_term("nat_cont")
    p2 = dict(prm, nc_ss_kind=0)
    o2, _ = _same_state(p2, m, G.FP64)
    o2.fix_nl()
    _, V_plain = o2.eval_term("nat_cont")
    return abs(V_ss - V_plain) > 1e-6 * abs(V_plain)


@pytest.mark.parametrize("cfg", ["ss_harmonic", "ss_lj"])
def test_disulfide_force_acts(cfg):
    prm, m = VARIANTS[cfg]()
    assert _nc_ss_energy_differs(prm, m)


@pytest.mark.parametrize("cfg", list(VARIANTS))
def test_batch_variant_follows_oracle(cfg):
    """The persistent kernel in FP64 follows the oracle over a noise-free
    trajectory with the variant's items, staged in LDS (one small protein)."""
    prm, m = VARIANTS[cfg]()
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


@pytest.mark.parametrize("cfg", ["snd", "heuristic", "all_heuristic", "ss_lj"])
def test_batch_variant_unstaged_follows_oracle(cfg):
    """The same items read from HBM: a batch whose largest system (1700
    residues) cannot stage its bonded parameters in LDS."""
    prm, m = VARIANTS[cfg]()
    big = M.synthetic_protein(1700, M.job_seed(121), heuristic_frac=0.3 if "heur" in cfg else 0.0)
    steps = 20
    b = G.Batch(prm, [m, big], seeds=[1, 1], precision=G.FP64, steps_per_launch=10)
    b.set_noise_scale(0.0)
    b.run(steps)
    for s, mm in enumerate((m, big)):
        o = Oracle(prm, mm, seed=1)
        o.set_noise_scale(0.0)
        tr = o.step(steps)
        err = np.abs(b.get(s, 0) - o.r()).max()
        assert err < 1e-8, (s, err)
        V, _ = b.energy(s)
        assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (s, V, tr[-1])
    b.close()


@pytest.mark.parametrize("cfg", list(VARIANTS))
def test_batch_variant_mixed_matches_fp64_energies_off(cfg):
    """With energies off every step of the mixed batched kernel takes its
    FP32 paths (heuristic angles from FP32 geometry, dihedrals without acos,
    FP32 disulfide and LJ contacts): the total force stays within 1e-5 of
    the FP64 kernel (relative to the largest component) on evolved
    positions."""
    prm, m = VARIANTS[cfg]()
    if cfg == "ss_lj":
        # The reference hands the Lennard-Jones disulfide force the distance
        # where it expects the inverse distance (nat_cont/eval_forces.cpp:27),
        # which makes it grow as r^12 beyond 4.2 A: a trajectory blows up
        # within a hundred steps, the oracle's as well. The model's own
        # (perturbed) positions are the state here.
        pos = m["pos"]
    else:
        b0 = G.Batch(prm, [m], precision=G.MIXED, steps_per_launch=100)
        b0.run(200)
        pos = b0.get(0, 0)
        b0.close()
    F = {}
    for prec in (G.FP64, G.MIXED):
        b = G.Batch(prm, [m], precision=prec, steps_per_launch=1)
        b.energies(False)
        b.set_noise_scale(0.0)
        b.set(0, 0, pos)
        b.run(1)
        F[prec] = b.get(0, 6)
        b.close()
    err = float(np.abs(F[G.MIXED] - F[G.FP64]).max() / np.abs(F[G.FP64]).max())
    assert err <= MIXED_F_TOL, err


@pytest.mark.parametrize("name", ["TEST", "input_test", "9aac"])
def test_golden_models_in_batch_follow_oracle(name):
    """The reference's own parity inputs (heuristic angles and dihedrals, PID,
    Debye-Hueckel; 9aac from the library's SAW) in the persistent kernel:
    FP64 follows the oracle without noise, and the mixed kernel with energies
    off stays within 1e-5 of the FP64 kernel's force."""
    p, m, seed, saw, _ = case(name)
    if saw:
        m, _skip = I.morph_into_saw(m, seed, bond=saw["bond"], intersection=saw["intersection"],
                                    update_orig=saw["update_orig"])
    steps = 60
    b = G.Batch(p, [m], seeds=[seed], precision=G.FP64, steps_per_launch=20)
    b.set_noise_scale(0.0)
    b.run(steps)
    o = Oracle(p, m, seed=seed)
    o.set_noise_scale(0.0)
    tr = o.step(steps)
    err = np.abs(b.get(0, 0) - o.r()).max()
    assert err < 1e-8, err
    V, _ = b.energy(0)
    assert abs(V - tr[-1]) <= 1e-9 * abs(tr[-1]), (V, tr[-1])
    pos = b.get(0, 0)
    b.close()
    F = {}
    for prec in (G.FP64, G.MIXED):
        bb = G.Batch(p, [m], seeds=[seed], precision=prec, steps_per_launch=1)
        bb.energies(False)
        bb.set_noise_scale(0.0)
        bb.set(0, 0, pos)
        bb.run(1)
        F[prec] = bb.get(0, 6)
        bb.close()
    ferr = float(np.abs(F[G.MIXED] - F[G.FP64]).max() / np.abs(F[G.FP64]).max())
    assert ferr <= MIXED_F_TOL, ferr


def _multichain_qa(seed=31, copies=3, n=50, side_A=45.0):
    one = M.synthetic_protein(n, M.job_seed(seed))
    one["atype"][::5] = P.AA_INDEX["ARG"]
    one["atype"][2::6] = P.AA_INDEX["ASP"]
    return M.replicate(one, copies, side_A * P.ANGSTROM, seed)


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED], ids=["fp64", "mixed"])
def test_qa_terms_lockstep_match_oracle(precision):
    """QA on (candidates, contacts with their saturation, the finish):
    evaluated step by step on the oracle's noise-free states, the whole
    force in mixed precision stays within 1e-5 of the FP64 oracle and the
    QA energies within 1e-6, while contacts and budgets follow exactly."""
    prm = _params(qa_on=1, pid_on=0, afm_on=0, temperature=0.3)
    m = _multichain_qa()
    worst, formed = qa_lockstep(prm, m, precision, 80)
    ftol = FP64_TOL if precision == G.FP64 else MIXED_F_TOL
    vtol = FP64_TOL if precision == G.FP64 else MIXED_V_TOL
    assert formed > 0
    assert worst["total"] <= ftol, worst
    assert worst["V_qa_cand"] <= vtol and worst["V_qa_cont"] <= vtol, worst
