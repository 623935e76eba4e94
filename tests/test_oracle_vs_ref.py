"""The oracle against the reference's own kernel sources (oracle/_ref).

Every force term, the NL rebuild with the ordered list updates, and a full
QA round are run on identical positions and lists by the oracle and by the
reference kernels compiled from their sources (PID through its vectorised
VCL path). Forces and energies must agree to 1e-10 relative (ULP-level
differences from the libm vs VCL transcendentals), lists exactly.
"""
import ctypes as C

import numpy as np
import pytest

import ref_ffi as R
from oracle_ffi import Oracle
from pas_cg_amd import model as M
from pas_cg_amd import params as P

pytestmark = pytest.mark.skipif(not R.available(),
                                reason="oracle/_ref not built (reference tree absent)")

D, I, U8, I8 = C.c_double, C.c_int32, C.c_uint8, C.c_int8


def ptr(a, ct):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ct))


def close(F1, V1, F2, V2, tol=1e-10):
    scale = max(np.abs(F2).max(), 1e-12)
    assert np.abs(F1 - F2).max() <= tol * scale, np.abs(F1 - F2).max() / scale
    assert abs(V1 - V2) <= tol * max(abs(V2), 1e-12), (V1, V2)


def full_params(**kw):
    p = P.default_params()
    p.update(chir_on=1, pid_on=1, afm_on=1)
    p.update(kw)
    return p


def perturbed_protein(n=90, seed=5, noise_A=0.4, heuristic=False):
    m = M.synthetic_protein(n, M.job_seed(seed))
    rng = np.random.default_rng(seed)
    m["pos"] = m["pos"] + rng.normal(scale=noise_A * P.ANGSTROM, size=m["pos"].shape)
    if heuristic:
        m["ang_has"][:] = 0
        m["dih_has"][:] = 0
    # a few ionised residues and one disulfide-typed contact
    m["atype"][::7] = P.AA_INDEX["LYS"]
    m["atype"][3::11] = P.AA_INDEX["GLU"]
    m["atype"][5::13] = P.AA_INDEX["PRO"]
    if len(m["cont_type"]):
        m["cont_type"][0] = 5
    return M.add_pulling(m)


def multichain_box(copies=4, n=40, side_A=70.0, seed=9):
    one = M.synthetic_protein(n, M.job_seed(seed))
    one["atype"][::5] = P.AA_INDEX["ARG"]
    one["atype"][2::6] = P.AA_INDEX["ASP"]
    m = M.replicate(one, copies, side_A * P.ANGSTROM, seed)
    rng = np.random.default_rng(seed)
    m["pos"] = m["pos"] + rng.normal(scale=0.3 * P.ANGSTROM, size=m["pos"].shape)
    return m


def term_ref(term, o, prm, m):
    n = o.n
    r = np.ascontiguousarray(o.r())
    cell = np.ascontiguousarray(m["cell"], np.float64)
    F = np.zeros((n, 3))
    ps = R.params_struct(prm)
    L = R.lib()
    if term in ("chir",):
        idx, v0, v1, _ = o.static_list(2)
        # the oracle stores the quad as the reference evaluates it; hand the
        # harness the setup's (nat_factor, nat_chir) in state.cpp:384 order,
        # i.e. the stored fields in field order (nat_chir field, nat_factor field)
        V = L.ref_chir(n, ptr(r, D), len(idx), ptr(np.ascontiguousarray(idx), I),
                       ptr(np.ascontiguousarray(v1), D), ptr(np.ascontiguousarray(v0), D),
                       prm["e_chi"], ptr(F, D))
    elif term == "tether":
        idx, v0, _, _ = o.static_list(0)
        ij = np.ascontiguousarray(idx[:, :2])
        V = L.ref_tether(n, ptr(r, D), len(ij), ptr(ij, I), ptr(v0, D),
                         prm["tether_H1"], prm["tether_H2"], ptr(F, D))
    elif term == "lrep":
        idx, _, _, _ = o.static_list(1)
        ij = np.ascontiguousarray(idx[:, :2])
        V = L.ref_lrep(n, ptr(r, D), len(ij), ptr(ij, I), prm["lrep_depth"],
                       prm["rep_cutoff"], ptr(F, D))
    elif term == "pauli":
        i1, i2, _, _ = o.list(2)
        ij = np.ascontiguousarray(np.stack([i1, i2], 1))
        V = L.ref_pauli(n, ptr(r, D), ptr(cell, D), len(ij), ptr(ij, I),
                        prm["pauli_depth"], prm["rep_cutoff"], ptr(F, D))
    elif term in ("const_dh", "rel_dh"):
        i1, i2, q, _ = o.list(3)
        ij = np.ascontiguousarray(np.stack([i1, i2], 1))
        V = L.ref_dh(0 if term == "const_dh" else 1, n, ptr(r, D), ptr(cell, D), len(ij),
                     ptr(ij, I), ptr(q, D), C.byref(ps), ptr(F, D))
    elif term == "nat_ang":
        idx, v0, _, _ = o.static_list(3)
        ijk = np.ascontiguousarray(idx[:, :3])
        V = L.ref_nat_ang(n, ptr(r, D), len(ijk), ptr(ijk, I), ptr(v0, D), prm["CBA"],
                          ptr(F, D))
    elif term == "heur_ang":
        idx, _, _, ty = o.static_list(4)
        ijk = np.ascontiguousarray(idx[:, :3])
        V = L.ref_heur_ang(n, ptr(r, D), len(ijk), ptr(ijk, I), ptr(ty, I), C.byref(ps),
                           ptr(F, D))
    elif term in ("cnd", "snd"):
        idx, v0, _, _ = o.static_list(5)
        q4 = np.ascontiguousarray(idx)
        V = L.ref_nat_dih(0 if term == "cnd" else 1, n, ptr(r, D), len(q4), ptr(q4, I),
                          ptr(v0, D), C.byref(ps), ptr(F, D))
    elif term == "heur_dih":
        idx, _, _, ty = o.static_list(6)
        q4 = np.ascontiguousarray(idx)
        V = L.ref_heur_dih(n, ptr(r, D), len(q4), ptr(q4, I), ptr(ty, I), C.byref(ps),
                           ptr(F, D))
    elif term == "pid":
        i1, i2, od, ty = o.list(4)
        ij = np.ascontiguousarray(np.stack([i1, i2], 1))
        tyi = np.ascontiguousarray(ty.astype(np.int32))
        # the oracle's pid list carries type in a uint8 flag: recompute it
        at = np.ascontiguousarray(m["atype"], np.uint8)
        tyi = np.ascontiguousarray((at[i1].astype(np.int32) * 21 + at[i2]).astype(np.int32))
        V = L.ref_pid(n, ptr(r, D), ptr(cell, D), ptr(at, U8), len(ij), ptr(ij, I),
                      ptr(od, D), ptr(tyi, I), C.byref(ps), ptr(F, D))
    elif term == "vel_afm":
        res = np.ascontiguousarray(m["tip_res"], np.int32)
        orig = np.ascontiguousarray(m["tip_orig"], np.float64)
        vel = np.ascontiguousarray(m["tip_vel"], np.float64)
        V = L.ref_vel_afm(n, ptr(r, D), o.t, len(res), ptr(res, I), ptr(orig, D),
                          ptr(vel, D), prm["afm_H1"], prm["afm_H2"], ptr(F, D))
    elif term == "solid_wall":
        V = L.ref_solid_wall(n, ptr(r, D), C.byref(ps), ptr(F, D))
    elif term == "force_afm":
        res = np.ascontiguousarray(m["ftip_res"], np.int32)
        frc = np.ascontiguousarray(m["ftip_force"], np.float64)
        V = L.ref_force_afm(n, o.t, len(res), ptr(res, I), ptr(frc, D), ptr(F, D))
    else:
        raise KeyError(term)
    return F, V


def advance(o, steps):
    o.step(steps)


@pytest.mark.parametrize("heuristic", [False, True], ids=["native", "heuristic"])
def test_bonded_and_afm_terms(heuristic):
    prm = full_params(qa_on=0)
    if heuristic:
        prm["nat_dih_variant"] = 1
    m = perturbed_protein(heuristic=heuristic)
    o = Oracle(prm, m, seed=3)
    advance(o, 30)
    terms = ["chir", "tether", "lrep", "vel_afm"]
    terms += ["heur_ang", "heur_dih"] if heuristic else ["nat_ang", "cnd"]
    if heuristic:
        terms = [t for t in terms]  # snd needs native dihedrals: separate test
    for term in terms:
        F1, V1 = o.eval_term(term)
        F2, V2 = term_ref(term, o, prm, m)
        assert np.abs(F2).max() > 0 or term == "vel_afm"
        close(F1, V1, F2, V2)


def test_simple_native_dihedral():
    prm = full_params(qa_on=0, nat_dih_variant=1)
    m = perturbed_protein()
    o = Oracle(prm, m, seed=4)
    advance(o, 10)
    F1, V1 = o.eval_term("snd")
    F2, V2 = term_ref("snd", o, prm, m)
    close(F1, V1, F2, V2)


@pytest.mark.parametrize("variant", ["lj-algebraic", "sink-cosine"])
def test_pair_terms_multichain_pbc(variant):
    # native contacts off so that close pairs go to the Pauli list
    prm = full_params(qa_on=0, nat_cont_on=0)
    if variant == "sink-cosine":
        prm["pid_lambda_variant"] = 0
        P.set_pid_forces(prm, "sink lj", bbp_depth=0.3, bbm_depth=0.25,
                         ss_depth_table=[[0.1 + 0.01 * ((a * 7 + b) % 13) for b in range(21)]
                                         for a in range(21)])
    m = multichain_box()
    o = Oracle(prm, m, seed=11)
    # the "lj" PID variant is very stiff on compact chains: compare early
    advance(o, 2)
    assert np.isfinite(o.r()).all()
    for term in ["pauli", "const_dh", "pid"]:
        F1, V1 = o.eval_term(term)
        F2, V2 = term_ref(term, o, prm, m)
        assert np.abs(F2).max() > 0, term
        close(F1, V1, F2, V2)


def test_relative_dh():
    prm = full_params(qa_on=0, dh_variant=1, dh_perm=4.0 * P.EPS_0 / P.ANGSTROM)
    m = multichain_box(seed=12)
    o = Oracle(prm, m, seed=12)
    advance(o, 5)
    F1, V1 = o.eval_term("rel_dh")
    F2, V2 = term_ref("rel_dh", o, prm, m)
    assert np.abs(F2).max() > 0
    close(F1, V1, F2, V2)


@pytest.mark.parametrize("ss_kind", [0, 1, 2])
def test_native_contacts_and_transitions(ss_kind):
    prm = full_params(qa_on=0, nc_ss_kind=ss_kind)
    m = perturbed_protein(noise_A=0.8)
    o = Oracle(prm, m, seed=21)
    advance(o, 40)
    # current contacts with their state before this evaluation
    idx, nat, _, ty = o.static_list(7)
    act0, ct0 = o.nat_cont_state(len(idx))
    i1, i2, _, _ = o.list(1)
    cur = {(a, b) for a, b in zip(i1, i2)}
    sel = np.array([(a, b) in cur for a, b in idx[:, :2]])
    ij = np.ascontiguousarray(idx[sel][:, :2])
    natc = np.ascontiguousarray(nat[sel])
    tyc = np.ascontiguousarray(ty[sel])
    act = np.ascontiguousarray(act0[sel])
    ct = np.ascontiguousarray(ct0[sel])
    F2 = np.zeros((o.n, 3))
    nch = C.c_int32(0)
    r = np.ascontiguousarray(o.r())
    cell = np.ascontiguousarray(m["cell"])
    V2 = R.lib().ref_nat_cont(o.n, ptr(r, D), ptr(cell, D), len(ij), ptr(ij, I),
                              ptr(natc, D), ptr(tyc, I), ptr(act, U8), ptr(ct, D), o.t,
                              C.byref(R.params_struct(prm)), C.byref(nch), ptr(F2, D))
    changed0 = o.num_changed()
    F1, V1 = o.eval_term("nat_cont")
    close(F1, V1, F2, V2)
    act1, ct1 = o.nat_cont_state(len(idx))
    assert np.array_equal(act1[sel], act), "active transitions differ"
    assert np.array_equal(ct1[sel], ct), "change times differ"
    assert o.num_changed() - changed0 == nch.value


@pytest.mark.parametrize("cfg", ["single", "multichain_qa", "multichain_pid"])
def test_nl_rebuild_and_list_updates_exact(cfg):
    if cfg == "single":
        prm = full_params(qa_on=0)
        m = perturbed_protein()
    elif cfg == "multichain_qa":
        prm = full_params(qa_on=1, qa_include4=0)
        m = multichain_box(seed=3)
    else:
        prm = full_params(qa_on=0, pid_include4=0)
        m = multichain_box(seed=4)
    o = Oracle(prm, m, seed=1)
    advance(o, 15)
    r = np.ascontiguousarray(o.r())
    o.fix_nl()
    cap = 200000
    sizes = np.zeros(6, np.int32)
    bufs = [np.zeros(2 * cap, np.int32) for _ in range(6)]
    nl_d = np.zeros(cap)
    taken = np.zeros(cap, np.uint8)
    idx, nat, _, _ = o.static_list(7)
    cij = np.ascontiguousarray(idx[:, :2])
    natc = np.ascontiguousarray(nat)
    R.lib().ref_fix_nl(o.n, ptr(r, D), ptr(np.ascontiguousarray(m["cell"]), D),
                       ptr(np.ascontiguousarray(m["atype"], np.uint8), U8),
                       ptr(np.ascontiguousarray(m["chain"], np.int32), I),
                       ptr(np.ascontiguousarray(m["seq"], np.int32), I),
                       C.byref(R.params_struct(prm)), len(cij), ptr(cij, I), ptr(natc, D),
                       cap, ptr(sizes, I), ptr(bufs[0], I), ptr(nl_d, D), ptr(taken, U8),
                       ptr(bufs[1], I), ptr(bufs[2], I), ptr(bufs[3], I), ptr(bufs[4], I),
                       ptr(bufs[5], I))
    kinds = {0: 0, 1: 1, 2: 5, 3: 4, 4: 3, 5: 2}  # harness slot -> oracle list kind
    for slot, kind in kinds.items():
        i1, i2, val, flag = o.list(kind)
        k = sizes[slot]
        assert len(i1) == k, (slot, len(i1), k)
        got = np.stack([i1, i2], 1)
        ref = bufs[slot][:2 * k].reshape(-1, 2)
        assert np.array_equal(got, ref), slot
        if slot == 0:
            assert np.array_equal(val, nl_d[:k])  # bit-exact distances
            assert np.array_equal(flag, taken[:k])
    assert sizes[0] > 0


def test_qa_round_matches_reference():
    prm = full_params(qa_on=1, pid_on=0, temperature=0.3)
    m = multichain_box(copies=3, n=50, side_A=45.0, seed=31)
    o = Oracle(prm, m, seed=31)
    checked = 0
    for _ in range(12):
        advance(o, 25)
        r = np.ascontiguousarray(o.r())
        fi1, fi2, fd, fv = o.list(5)
        cij, cod, cty, csat, csd, cvac = o.qa_contacts()
        sync = np.ascontiguousarray(o.sync())
        # reference round from the same state
        cap = len(fi1) + 4 * len(cij) + 1000
        n_free = C.c_int32(len(fi1))
        n_cont = C.c_int32(len(cij))
        num_cont = C.c_int32(o.num_qa_contacts())
        f_ij = np.zeros(2 * cap, np.int32)
        f_ij[:2 * len(fi1)] = np.stack([fi1, fi2], 1).reshape(-1)
        f_d = np.zeros(cap)
        f_d[:len(fd)] = fd
        f_v = np.zeros(cap, np.uint8)
        f_v[:len(fv)] = fv
        c_ij = np.zeros(2 * cap, np.int32)
        c_ij[:2 * len(cij)] = cij.reshape(-1)
        c_d = np.zeros(cap)
        c_d[:len(cod)] = cod
        c_t = np.zeros(cap, np.int32)
        c_t[:len(cty)] = cty
        c_s = np.zeros(cap)
        c_s[:len(csat)] = csat
        c_sd = np.zeros(8 * cap, np.int8)
        c_sd[:8 * len(csd)] = csd.reshape(-1)
        c_v = np.zeros(cap, np.uint8)
        c_v[:len(cvac)] = cvac
        F2 = np.zeros((o.n, 3))
        sync_ref = sync.copy()
        V2 = R.lib().ref_qa_round(
            o.n, ptr(r, D), ptr(np.ascontiguousarray(m["cell"]), D),
            ptr(np.ascontiguousarray(m["atype"], np.uint8), U8),
            ptr(np.ascontiguousarray(m["chain"], np.int32), I),
            C.byref(R.params_struct(prm)), prm["dt"], ptr(sync_ref, I8), C.byref(n_free),
            ptr(f_ij, I), ptr(f_d, D), ptr(f_v, U8), C.byref(n_cont), ptr(c_ij, I),
            ptr(c_d, D), ptr(c_t, I), ptr(c_s, D), ptr(c_sd, I8), ptr(c_v, U8),
            C.byref(num_cont), cap, ptr(F2, D))
        # the oracle's own round
        o.prepare_nh()
        Fa, Va = o.eval_term("qa_cand")
        Fb, Vb = o.eval_term("qa_cont")
        o.qa_finish()
        close(Fa + Fb, Va + Vb, F2, V2)
        fi1b, fi2b, fdb, fvb = o.list(5)
        assert n_free.value == len(fi1b)
        assert np.array_equal(np.stack([fi1b, fi2b], 1).reshape(-1), f_ij[:2 * n_free.value])
        assert np.array_equal(fvb, f_v[:n_free.value])
        cij2, cod2, cty2, csat2, csd2, cvac2 = o.qa_contacts()
        assert n_cont.value == len(cij2)
        k = n_cont.value
        assert np.array_equal(cij2.reshape(-1), c_ij[:2 * k])
        assert np.array_equal(cty2, c_t[:k])
        assert np.array_equal(cvac2, c_v[:k])
        assert np.allclose(csat2, c_s[:k], rtol=0, atol=1e-14)
        assert np.array_equal(o.sync(), sync_ref)
        assert o.num_qa_contacts() == num_cont.value
        checked += k
        # keep the trajectory going from the oracle's state
    assert checked > 0, "no QA contacts formed"


def _ref_pid_counts(o, prm, m):
    i1, i2, od, _ = o.list(4)
    at = np.ascontiguousarray(m["atype"], np.uint8)
    ij = np.ascontiguousarray(np.stack([i1, i2], 1).astype(np.int32))
    ty = np.ascontiguousarray((at[i1].astype(np.int32) * 21 + at[i2]).astype(np.int32))
    ch = np.ascontiguousarray(m["chain"], np.int32)
    r = np.ascontiguousarray(o.r())
    cell = np.ascontiguousarray(m["cell"], np.float64)
    out = np.zeros(6, np.int32)
    R.lib().ref_pid_count(o.n, ptr(r, D), ptr(cell, D), ptr(at, U8), ptr(ch, I), len(ij),
                          ptr(ij, I), ptr(np.ascontiguousarray(od), D), ptr(ty, I),
                          C.byref(R.params_struct(prm)), ptr(out, I))
    return out


@pytest.mark.parametrize("variant", ["lj-algebraic", "sink-cosine"])
def test_pid_activity_counts(variant):
    """pid::count_active of the report: the oracle's counts over its current
    bundles equal the reference's on the same positions and bundles, with
    intra- and inter-chain, backbone and side-chain counts all exercised."""
    prm = full_params(qa_on=0, nat_cont_on=0)
    if variant == "sink-cosine":
        prm["pid_lambda_variant"] = 0
        P.set_pid_forces(prm, "sink lj", bbp_depth=0.3, bbm_depth=0.25,
                         ss_depth_table=[[0.1 + 0.01 * ((a * 7 + b) % 13) for b in range(21)]
                                         for a in range(21)])
    m = multichain_box()
    o = Oracle(prm, m, seed=11)
    seen = np.zeros(6, np.int64)
    for _ in range(2):
        o.step(3)
        got, want = o.pid_counts(), _ref_pid_counts(o, prm, m)
        assert (got == want).all(), (got, want)
        assert got[0] == got[1] + got[2] and got[3] == got[4] + got[5]
        seen += got
    assert seen[1] > 0 and seen[2] > 0
    if variant == "lj-algebraic":
        assert seen[4] > 0 and seen[5] > 0


def test_force_afm_tips():
    """Constant-force AFM tips: the oracle equals afm::force::eval_forces."""
    prm = full_params(qa_on=0)
    m = M.add_force_pulling(perturbed_protein(), force_eps_per_A=2.5)
    o = Oracle(prm, m, seed=8)
    advance(o, 5)
    F1, V1 = o.eval_term("force_afm")
    F2, V2 = term_ref("force_afm", o, prm, m)
    assert np.abs(F2).max() > 0 and V1 == 0.0 and V2 == 0.0
    close(F1, V1, F2, V2, tol=1e-15)


def walled(m, gap_A=4.0):
    """Solid walls on the six faces of the model's bounding box widened by
    gap_A, so the outermost residues sit inside the wall threshold."""
    prm = full_params(qa_on=0)
    lo = m["pos"].min(axis=0) - gap_A * P.ANGSTROM
    hi = m["pos"].max(axis=0) + gap_A * P.ANGSTROM
    P.set_solid_walls(prm, lo, hi)
    return prm


def test_solid_walls():
    """Solid walls: the oracle equals wall::solid::eval_forces."""
    m = perturbed_protein()
    prm = walled(m)
    o = Oracle(prm, m, seed=9)
    advance(o, 5)
    F1, V1 = o.eval_term("solid_wall")
    F2, V2 = term_ref("solid_wall", o, prm, m)
    assert (np.abs(F2).max(axis=1) > 0).sum() >= 6
    close(F1, V1, F2, V2, tol=1e-14)
