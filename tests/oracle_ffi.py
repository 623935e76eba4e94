"""ctypes binding of the FP64 CPU oracle (oracle/oracle.h).

Test infrastructure: used only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")

TERMS = ["chir", "tether", "lrep", "nat_cont", "pauli", "const_dh", "rel_dh",
         "qa_cand", "qa_cont", "pid", "nat_ang", "heur_ang", "heur_dih", "cnd",
         "snd", "vel_afm", "force_afm", "solid_wall", "hw_free", "hw_conn", "ljw_free",
         "ljw_conn"]
TERM_ID = {t: i for i, t in enumerate(TERMS)}

D = C.c_double
I = C.c_int32


class OraParams(C.Structure):
    _fields_ = [
        ("rep_cutoff", D), ("counting_factor", D), ("nl_cutoff", D), ("nl_pad", D),
        ("tether_on", I), ("tether_H1", D), ("tether_H2", D),
        ("lrep_on", I), ("lrep_depth", D),
        ("chir_on", I), ("e_chi", D),
        ("nat_ang_on", I), ("CBA", D),
        ("heur_ang_on", I), ("heur_ang_poly", (D * 9) * 7),
        ("nat_dih_on", I), ("nat_dih_variant", I), ("CDA", D), ("CDB", D), ("CDH", D),
        ("heur_dih_on", I), ("heur_dih_coeffs", (D * 9) * 6),
        ("pauli_on", I), ("pauli_depth", D),
        ("nat_cont_on", I), ("nat_cont_depth", D), ("nc_ss_kind", I),
        ("nc_ss_H1", D), ("nc_ss_H2", D), ("nc_ss_nat_r", D), ("nc_ss_depth", D),
        ("nc_ss_r_min", D),
        ("dh_on", I), ("dh_variant", I), ("dh_screen_dist", D), ("dh_cutoff", D),
        ("dh_perm", D),
        ("qa_on", I), ("qa_include4", I), ("qa_phase_dur", D), ("qa_formation_tol", D),
        ("qa_min_cos_hr", D), ("qa_min_cos_hh", D), ("qa_max_cos_nr", D),
        ("qa_lj_depth", D * 445), ("qa_lj_r_low", D * 445), ("qa_lj_r_high", D * 445),
        ("pid_on", I), ("pid_include4", I), ("pid_lambda_variant", I),
        ("pid_bbp_alpha", D), ("pid_bbp_psi0", D), ("pid_bbm_alpha", D),
        ("pid_bbm_psi0", D), ("pid_ss_alpha", D), ("pid_ss_psi0", D),
        ("pid_bbp_lj", D * 3), ("pid_bbm_lj", D * 3),
        ("pid_ss_depth", D * 441), ("pid_ss_r_low", D * 441), ("pid_ss_r_high", D * 441),
        ("afm_on", I), ("afm_H1", D), ("afm_H2", D),
        ("lang_on", I), ("gamma", D), ("temperature", D), ("dt", D),
        ("mass", D * 21), ("charge", D * 21), ("ptype", I * 21),
        ("base_sync", (I * 4) * 21),
        ("n_walls", I), ("wall_origin", (D * 3) * 6), ("wall_normal", (D * 3) * 6),
        ("wall_depth", D), ("wall_min_dist", D),
        ("wall_kind", I * 6), ("wall_side", I * 6), ("wall_limit", I * 6),
        ("hw_on", I), ("ljw_on", I),
        ("wall_hh1", D), ("wall_lj_depth", D), ("ljw_breaking_dist", D),
        ("ljw_saturation_diff", D), ("afm_stats_window", D),
    ]


class OraBox(C.Structure):
    """The sbox parameters of the box protocol (ora_box, the fields of pcg_box)."""
    _fields_ = [
        ("box_min", D * 3), ("box_max", D * 3), ("pbc", I * 3), ("total_time", D),
        ("rest_period", D), ("target_vel", D), ("accel_dist", D), ("avg_forces_over", D),
        ("squeezing", I), ("target_density", D), ("vel_above_2V", D),
        ("force_min", I), ("force_for_max_vel", D),
        ("oscillations", I), ("osc_shear", I), ("num_of_cycles", I), ("angular_freq", D),
        ("amp_relative", I), ("amp_abs", D), ("amp_rel", D),
    ]


class OraModel(C.Structure):
    _fields_ = [
        ("n", I), ("pos", C.POINTER(D)), ("orig_pos", C.POINTER(D)),
        ("atype", C.POINTER(C.c_uint8)), ("chain_idx", C.POINTER(I)),
        ("seq_idx", C.POINTER(I)),
        ("n_teth", I), ("teth", C.POINTER(I)),
        ("n_ang", I), ("ang", C.POINTER(I)), ("ang_has", C.POINTER(C.c_uint8)),
        ("ang_theta", C.POINTER(D)),
        ("n_dih", I), ("dih", C.POINTER(I)), ("dih_has", C.POINTER(C.c_uint8)),
        ("dih_phi", C.POINTER(D)),
        ("n_cont", I), ("cont", C.POINTER(I)), ("cont_len", C.POINTER(D)),
        ("cont_type", C.POINTER(I)),
        ("cell", D * 3),
        ("n_tips", I), ("tip_res", C.POINTER(I)), ("tip_orig", C.POINTER(D)),
        ("tip_vel", C.POINTER(D)),
        ("n_ftips", I), ("ftip_res", C.POINTER(I)), ("ftip_force", C.POINTER(D)),
    ]


class OraSaw(C.Structure):
    _fields_ = [("perform", I), ("bond", D), ("intersection", D), ("with_box", I),
                ("box_size", D), ("pbc", I), ("update_orig", I)]


def fill_struct(st, p: dict):
    for name, ctype in st._fields_:
        if name not in p:
            continue
        v = p[name]
        if hasattr(ctype, "_length_"):
            arr = getattr(st, name)
            flat = np.asarray(v).reshape(-1)
            inner = getattr(ctype, "_type_")
            if hasattr(inner, "_length_"):
                k = inner._length_
                for a in range(ctype._length_):
                    for b in range(k):
                        arr[a][b] = flat[a * k + b]
            else:
                for a in range(ctype._length_):
                    arr[a] = flat[a]
        else:
            setattr(st, name, v)
    return st


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.ora_create.restype = P
        L.ora_create.argtypes = [C.POINTER(OraParams), C.POINTER(OraModel), I,
                                 C.POINTER(OraSaw), C.POINTER(D)]
        L.ora_destroy.argtypes = [P]
        L.ora_step.argtypes = [P, I, C.POINTER(D)]
        L.ora_last_V.restype = D
        L.ora_last_V.argtypes = [P]
        L.ora_last_term_V.argtypes = [P, C.POINTER(D)]
        L.ora_nl_verify.restype = I
        L.ora_nl_verify.argtypes = [P]
        for f in ("ora_fix_nl", "ora_eval_forces", "ora_qa_finish", "ora_lang_step",
                  "ora_prepare_nh"):
            getattr(L, f).argtypes = [P]
        L.ora_eval_term.argtypes = [P, I, C.POINTER(D), C.POINTER(D)]
        for f in ("ora_get_r", "ora_set_r", "ora_get_F"):
            getattr(L, f).argtypes = [P, C.POINTER(D)]
        L.ora_get_y.argtypes = [P, I, C.POINTER(D)]
        L.ora_set_y.argtypes = [P, I, C.POINTER(D)]
        L.ora_get_t.restype = D
        L.ora_get_t.argtypes = [P]
        L.ora_set_t.argtypes = [P, D]
        L.ora_set_noise_scale.argtypes = [P, D]
        L.ora_list.restype = I
        L.ora_list.argtypes = [P, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(D),
                               C.POINTER(C.c_uint8)]
        L.ora_nat_cont_state.argtypes = [P, C.POINTER(C.c_uint8), C.POINTER(D)]
        for f in ("ora_num_changed", "ora_num_qa_contacts", "ora_nl_rebuilds"):
            getattr(L, f).restype = I
            getattr(L, f).argtypes = [P]
        L.ora_sync.argtypes = [P, C.POINTER(C.c_int8)]
        L.ora_pid_counts.argtypes = [P, C.POINTER(I)]
        L.ora_set_tips.argtypes = [P, I, C.POINTER(I), C.POINTER(D), C.POINTER(D), I,
                                   C.POINTER(I), C.POINTER(D)]
        L.ora_saw_draws.restype = C.c_int64
        L.ora_saw_draws.argtypes = [P]
        L.ora_static_list.restype = I
        L.ora_static_list.argtypes = [P, I, I, C.POINTER(I), C.POINTER(D), C.POINTER(D),
                                      C.POINTER(I)]
        L.ora_qa_contacts.restype = I
        L.ora_qa_contacts.argtypes = [P, I, C.POINTER(I), C.POINTER(D), C.POINTER(I),
                                      C.POINTER(D), C.POINTER(C.c_int8),
                                      C.POINTER(C.c_uint8)]
        L.ora_afm_stats.restype = I
        L.ora_afm_stats.argtypes = [P, C.POINTER(D), C.POINTER(D), C.POINTER(C.c_uint8)]
        L.ora_box_setup.argtypes = [P, C.POINTER(OraBox)]
        L.ora_box_start.argtypes = [P]
        L.ora_box_state.argtypes = [P, C.POINTER(D)]
        L.ora_ljw_process.argtypes = [P]
        L.ora_wall_conns.restype = I
        L.ora_wall_conns.argtypes = [P, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(D),
                                     C.POINTER(D)]
        L.ora_ljw_slots.restype = I
        L.ora_ljw_slots.argtypes = [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(D),
                                    C.POINTER(D), C.POINTER(C.c_uint8)]
        L.ora_ljw_pending.argtypes = [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I),
                                      C.POINTER(D), C.POINTER(I), C.POINTER(I),
                                      C.POINTER(C.c_uint8), C.POINTER(I)]
        L.ora_rng_new.restype = P
        L.ora_rng_new.argtypes = [I]
        L.ora_rng_free.argtypes = [P]
        L.ora_rng_uniform.restype = D
        L.ora_rng_uniform.argtypes = [P]
        L.ora_rng_normal.restype = D
        L.ora_rng_normal.argtypes = [P]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Oracle:
    """One trajectory of the oracle."""

    def __init__(self, params: dict, model: dict, seed: int = 448, saw: dict | None = None):
        L = lib()
        self.n = int(model["n"])
        self._keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(np.asarray(x, dt))
            self._keep.append(a)
            return a

        m = OraModel()
        m.n = self.n
        pos = arr(model["pos"], np.float64)
        m.pos = _ptr(pos, D)
        m.orig_pos = _ptr(arr(model["orig_pos"], np.float64), D)
        m.atype = _ptr(arr(model["atype"], np.uint8), C.c_uint8)
        m.chain_idx = _ptr(arr(model["chain"], np.int32), I)
        m.seq_idx = _ptr(arr(model["seq"], np.int32), I)
        m.n_teth = len(model["teth"])
        m.teth = _ptr(arr(model["teth"], np.int32), I)
        m.n_ang = len(model["ang"])
        m.ang = _ptr(arr(model["ang"], np.int32), I)
        m.ang_has = _ptr(arr(model["ang_has"], np.uint8), C.c_uint8)
        m.ang_theta = _ptr(arr(model["ang_theta"], np.float64), D)
        m.n_dih = len(model["dih"])
        m.dih = _ptr(arr(model["dih"], np.int32), I)
        m.dih_has = _ptr(arr(model["dih_has"], np.uint8), C.c_uint8)
        m.dih_phi = _ptr(arr(model["dih_phi"], np.float64), D)
        m.n_cont = len(model["cont"])
        m.cont = _ptr(arr(model["cont"], np.int32), I)
        m.cont_len = _ptr(arr(model["cont_len"], np.float64), D)
        m.cont_type = _ptr(arr(model["cont_type"], np.int32), I)
        for k in range(3):
            m.cell[k] = float(model["cell"][k])
        m.n_tips = len(model["tip_res"])
        m.tip_res = _ptr(arr(model["tip_res"], np.int32), I)
        m.tip_orig = _ptr(arr(model["tip_orig"], np.float64), D)
        m.tip_vel = _ptr(arr(model["tip_vel"], np.float64), D)
        ftip_res = model.get("ftip_res", np.zeros(0, np.int32))
        m.n_ftips = len(ftip_res)
        m.ftip_res = _ptr(arr(ftip_res, np.int32), I)
        m.ftip_force = _ptr(arr(model.get("ftip_force", np.zeros((0, 3))), np.float64), D)
        self._model = m
        p = fill_struct(OraParams(), params)
        s = OraSaw()
        if saw:
            s.perform = 1
            s.bond = saw["bond"]
            s.intersection = saw["intersection"]
            s.with_box = 1 if saw.get("box_size") else 0
            s.box_size = saw.get("box_size", 0.0)
            s.pbc = 1 if saw.get("pbc") else 0
            s.update_orig = 1 if saw.get("update_orig") else 0
        self.pos0 = np.zeros((self.n, 3))
        self.h = L.ora_create(C.byref(p), C.byref(m), seed, C.byref(s),
                              _ptr(self.pos0, D))
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_destroy(self.h)
            self.h = None

    def step(self, k: int = 1) -> np.ndarray:
        tr = np.zeros(k)
        self.L.ora_step(self.h, k, _ptr(tr, D))
        return tr

    @property
    def V(self) -> float:
        return self.L.ora_last_V(self.h)

    def term_V(self) -> dict:
        out = np.zeros(len(TERMS))
        self.L.ora_last_term_V(self.h, _ptr(out, D))
        return dict(zip(TERMS, out))

    def nl_verify(self) -> bool:
        return bool(self.L.ora_nl_verify(self.h))

    def fix_nl(self):
        self.L.ora_fix_nl(self.h)

    def prepare_nh(self):
        self.L.ora_prepare_nh(self.h)

    def eval_forces(self):
        self.L.ora_eval_forces(self.h)

    def eval_term(self, term: str):
        F = np.zeros((self.n, 3))
        V = C.c_double(0.0)
        self.L.ora_eval_term(self.h, TERM_ID[term], _ptr(F, D), C.byref(V))
        return F, V.value

    def qa_finish(self):
        self.L.ora_qa_finish(self.h)

    def lang_step(self):
        self.L.ora_lang_step(self.h)

    def r(self) -> np.ndarray:
        out = np.zeros((self.n, 3))
        self.L.ora_get_r(self.h, _ptr(out, D))
        return out

    def set_r(self, r):
        a = np.ascontiguousarray(np.asarray(r, np.float64))
        self.L.ora_set_r(self.h, _ptr(a, D))

    def F(self) -> np.ndarray:
        out = np.zeros((self.n, 3))
        self.L.ora_get_F(self.h, _ptr(out, D))
        return out

    def y(self, k) -> np.ndarray:
        out = np.zeros((self.n, 3))
        self.L.ora_get_y(self.h, k, _ptr(out, D))
        return out

    def set_y(self, k, y):
        a = np.ascontiguousarray(np.asarray(y, np.float64))
        self.L.ora_set_y(self.h, k, _ptr(a, D))

    @property
    def t(self) -> float:
        return self.L.ora_get_t(self.h)

    def set_t(self, t: float):
        self.L.ora_set_t(self.h, t)

    def set_noise_scale(self, s: float):
        self.L.ora_set_noise_scale(self.h, s)

    def list(self, kind: int):
        size = self.L.ora_list(self.h, kind, 0, None, None, None, None)
        i1 = np.zeros(size, np.int32)
        i2 = np.zeros(size, np.int32)
        val = np.zeros(size)
        flag = np.zeros(size, np.uint8)
        self.L.ora_list(self.h, kind, size, _ptr(i1, I), _ptr(i2, I), _ptr(val, D),
                        _ptr(flag, C.c_uint8))
        return i1, i2, val, flag

    def static_list(self, kind: int):
        size = self.L.ora_static_list(self.h, kind, 0, None, None, None, None)
        idx = np.zeros((size, 4), np.int32)
        v0 = np.zeros(size)
        v1 = np.zeros(size)
        ity = np.zeros(size, np.int32)
        self.L.ora_static_list(self.h, kind, size, _ptr(idx, I), _ptr(v0, D), _ptr(v1, D),
                               _ptr(ity, I))
        return idx, v0, v1, ity

    def qa_contacts(self):
        size = self.L.ora_qa_contacts(self.h, 0, None, None, None, None, None, None)
        ij = np.zeros((size, 2), np.int32)
        orig = np.zeros(size)
        ty = np.zeros(size, np.int32)
        sat = np.zeros(size)
        sd = np.zeros((size, 8), np.int8)
        vac = np.zeros(size, np.uint8)
        self.L.ora_qa_contacts(self.h, size, _ptr(ij, I), _ptr(orig, D), _ptr(ty, I),
                               _ptr(sat, D), _ptr(sd, C.c_int8), _ptr(vac, C.c_uint8))
        return ij, orig, ty, sat, sd, vac

    def nat_cont_state(self, k: int):
        a = np.zeros(k, np.uint8)
        ct = np.zeros(k)
        self.L.ora_nat_cont_state(self.h, _ptr(a, C.c_uint8), _ptr(ct, D))
        return a, ct

    def num_changed(self) -> int:
        return self.L.ora_num_changed(self.h)

    def num_qa_contacts(self) -> int:
        return self.L.ora_num_qa_contacts(self.h)

    def saw_draws(self) -> int:
        return self.L.ora_saw_draws(self.h)

    def nl_rebuilds(self) -> int:
        return self.L.ora_nl_rebuilds(self.h)

    def sync(self) -> np.ndarray:
        out = np.zeros((self.n, 4), np.int8)
        self.L.ora_sync(self.h, _ptr(out, C.c_int8))
        return out

    def set_tips(self, tips: dict):
        """Replaces the AFM tips (reference time = now); {} removes them."""
        vr = np.ascontiguousarray(tips.get("tip_res", []), np.int32)
        vo = np.ascontiguousarray(np.reshape(tips.get("tip_orig", np.zeros((0, 3))), (-1, 3)), np.float64)
        vv = np.ascontiguousarray(np.reshape(tips.get("tip_vel", np.zeros((0, 3))), (-1, 3)), np.float64)
        fr = np.ascontiguousarray(tips.get("ftip_res", []), np.int32)
        ff = np.ascontiguousarray(np.reshape(tips.get("ftip_force", np.zeros((0, 3))), (-1, 3)), np.float64)
        self.L.ora_set_tips(self.h, len(vr), _ptr(vr, I), _ptr(vo, D), _ptr(vv, D), len(fr),
                            _ptr(fr, I), _ptr(ff, D))

    def pid_counts(self) -> np.ndarray:
        """{intra, intra_bb, intra_ss, inter, inter_bb, inter_ss} of the
        current PID bundles at the current positions."""
        out = np.zeros(6, np.int32)
        self.L.ora_pid_counts(self.h, _ptr(out, I))
        return out

    def afm_stats(self) -> dict:
        """The tips' moving averages: avg [tips, 3], perp [tips], has [tips]
        (velocity tips first, then force tips)."""
        k = self.L.ora_afm_stats(self.h, None, None, None)
        avg, perp, has = np.zeros((k, 3)), np.zeros(k), np.zeros(k, np.uint8)
        if k:
            self.L.ora_afm_stats(self.h, _ptr(avg, D), _ptr(perp, D), _ptr(has, C.c_uint8))
        return {"avg": avg, "perp": perp, "has": has.astype(bool)}

    def box_setup(self, box: dict):
        self._box = fill_struct(OraBox(), box)
        self.L.ora_box_setup(self.h, C.byref(self._box))

    def box_start(self):
        self.L.ora_box_start(self.h)

    def box_state(self) -> np.ndarray:
        out = np.zeros(120, np.float64)
        self.L.ora_box_state(self.h, _ptr(out, D))
        return out

    def ljw_process(self):
        self.L.ora_ljw_process(self.h)

    def wall_conns(self, kind: int) -> dict:
        """kind 0 harmonic, 1 LJ (live slots): residue, plane, offset, saturation."""
        k = self.L.ora_wall_conns(self.h, kind, 0, None, None, None, None)
        res, wall = np.zeros(k, np.int32), np.zeros(k, np.int32)
        off, sat = np.zeros((k, 3), np.float64), np.zeros(k, np.float64)
        if k:
            self.L.ora_wall_conns(self.h, kind, k, _ptr(res, I), _ptr(wall, I), _ptr(off, D),
                                  _ptr(sat, D))
        return {"res": res, "wall": wall, "off": off, "sat": sat}

    def ljw_slots(self) -> dict:
        k = self.L.ora_ljw_slots(self.h, 0, None, None, None, None, None)
        res, wall = np.zeros(k, np.int32), np.zeros(k, np.int32)
        off, sat, vac = np.zeros((k, 3)), np.zeros(k), np.zeros(k, np.uint8)
        if k:
            self.L.ora_ljw_slots(self.h, k, _ptr(res, I), _ptr(wall, I), _ptr(off, D),
                                 _ptr(sat, D), _ptr(vac, C.c_uint8))
        return {"res": res, "wall": wall, "off": off, "sat": sat, "vac": vac}

    def ljw_pending(self, cap: int = 100000) -> dict:
        nc, nr = I(0), I(0)
        cr, cw, cd = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap)
        rem = np.zeros(cap, np.int32)
        conn = np.zeros(self.n, np.uint8)
        num = np.zeros(6, np.int32)
        self.L.ora_ljw_pending(self.h, cap, C.byref(nc), _ptr(cr, I), _ptr(cw, I), _ptr(cd, D),
                               C.byref(nr), _ptr(rem, I), _ptr(conn, C.c_uint8), _ptr(num, I))
        k, m = nc.value, nr.value
        return {"cand_res": cr[:k], "cand_wall": cw[:k], "cand_d": cd[:k], "removed": rem[:m],
                "is_conn": conn, "num_conns": num}


class Ran2:
    def __init__(self, seed: int):
        self.L = lib()
        self.h = self.L.ora_rng_new(seed)

    def uniform(self) -> float:
        return self.L.ora_rng_uniform(self.h)

    def normal(self) -> float:
        return self.L.ora_rng_normal(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ora_rng_free(self.h)
            self.h = None
