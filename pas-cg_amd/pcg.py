"""ctypes binding of the C ABI (include/pcg.h) of libpcg.so.

`Context` mirrors the reference's per-step loop and its per-term tasks
(thread::advance_by_step, sliceable_task::for_slice); `Batch` runs many
systems in one persistent launch; `Server` admits jobs continuously.
There is no CPU fallback: without the HIP library these classes raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpcg.so")

TERMS = ["chir", "tether", "lrep", "nat_cont", "pauli", "const_dh", "rel_dh",
         "qa_cand", "qa_cont", "pid", "nat_ang", "heur_ang", "heur_dih", "cnd",
         "snd", "vel_afm", "force_afm", "solid_wall", "hw_free", "hw_conn", "ljw_free",
         "ljw_conn"]
TERM_ID = {t: i for i, t in enumerate(TERMS)}
# pcg_report_field order: the scalars of make_report::add_cur_scalars
REPORT_FIELDS = ["TIME", "EPOT", "ETOT", "EKIN", "RG", "L", "RMSD", "W", "ICN", "ICNss",
                 "INTEHC", "INTRHC", "PID-INTRHC", "PID-B1-B1", "PID-B1-S1", "PID-INTEHC",
                 "PID-B1-B2", "PID-S1-S2"]
MIXED, FP64 = 0, 1
FORM_AUTO, FORM_STAGES, FORM_PERSISTENT = 0, 1, 2
PHILOX, RAN2 = 0, 1

D, I, I64, U8 = C.c_double, C.c_int32, C.c_int64, C.c_uint8


class PcgParams(C.Structure):
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


class PcgBox(C.Structure):
    """pcg_box: the sbox parameters of the box protocol (params.box_params)."""
    _fields_ = [
        ("box_min", D * 3), ("box_max", D * 3), ("pbc", I * 3), ("total_time", D),
        ("rest_period", D), ("target_vel", D), ("accel_dist", D), ("avg_forces_over", D),
        ("squeezing", I), ("target_density", D), ("vel_above_2V", D),
        ("force_min", I), ("force_for_max_vel", D),
        ("oscillations", I), ("osc_shear", I), ("num_of_cycles", I), ("angular_freq", D),
        ("amp_relative", I), ("amp_abs", D), ("amp_rel", D),
    ]


# pcg_box_state layout and phase ids
BOX_STATE = {"phase": 0, "since": 1, "until": 2, "min": 3, "max": 6, "amplitude": 9,
             "omega": 10, "displacement": 11, "shear": 12, "avg_z_has": 13, "avg_z": 14,
             "origin": 15, "normal": 33, "wall_F": 51, "shift": 69, "work": 87,
             "avg_F_has": 93, "avg_F": 99, "cell": 117}
BOX_STATE_NUM = 120
PHASES = ["NONE", "SQUEEZING", "REST_AFTER_SQUEEZING", "FIND_FORCE_MIN",
          "REST_AFTER_FORCE_MIN", "MAX_AMPLITUDE", "REST_AFTER_MAX_AMP", "OSCILLATIONS",
          "REST_AFTER_OSCILLATIONS", "FREEFORM"]


def box_state_dict(v: np.ndarray) -> dict:
    """The pcg_box_state array as named fields."""
    S = BOX_STATE
    return {
        "phase": PHASES[int(v[S["phase"]])], "since": v[S["since"]], "until": v[S["until"]],
        "min": v[S["min"]:S["min"] + 3], "max": v[S["max"]:S["max"] + 3],
        "amplitude": v[S["amplitude"]], "omega": v[S["omega"]],
        "displacement": v[S["displacement"]], "shear": bool(v[S["shear"]]),
        "avg_z": v[S["avg_z"]] if v[S["avg_z_has"]] else None,
        "origin": v[S["origin"]:S["origin"] + 18].reshape(6, 3),
        "normal": v[S["normal"]:S["normal"] + 18].reshape(6, 3),
        "wall_F": v[S["wall_F"]:S["wall_F"] + 18].reshape(6, 3),
        "shift": v[S["shift"]:S["shift"] + 18].reshape(6, 3),
        "work": v[S["work"]:S["work"] + 6],
        "avg_F_has": v[S["avg_F_has"]:S["avg_F_has"] + 6].astype(bool),
        "avg_F": v[S["avg_F"]:S["avg_F"] + 18].reshape(6, 3),
        "cell": v[S["cell"]:S["cell"] + 3],
    }


class PcgModel(C.Structure):
    _fields_ = [
        ("n", I), ("pos", C.POINTER(D)), ("orig_pos", C.POINTER(D)),
        ("atype", C.POINTER(U8)), ("chain_idx", C.POINTER(I)), ("seq_idx", C.POINTER(I)),
        ("n_teth", I), ("teth", C.POINTER(I)),
        ("n_ang", I), ("ang", C.POINTER(I)), ("ang_has", C.POINTER(U8)),
        ("ang_theta", C.POINTER(D)),
        ("n_dih", I), ("dih", C.POINTER(I)), ("dih_has", C.POINTER(U8)),
        ("dih_phi", C.POINTER(D)),
        ("n_cont", I), ("cont", C.POINTER(I)), ("cont_len", C.POINTER(D)),
        ("cont_type", C.POINTER(I)),
        ("cell", D * 3),
        ("n_tips", I), ("tip_res", C.POINTER(I)), ("tip_orig", C.POINTER(D)),
        ("tip_vel", C.POINTER(D)),
        ("n_ftips", I), ("ftip_res", C.POINTER(I)), ("ftip_force", C.POINTER(D)),
    ]


class PcgOptions(C.Structure):
    _fields_ = [("precision", I), ("noise", I), ("seed", I), ("philox_seed", C.c_uint64),
                ("temperature", D), ("device", I), ("ran2_skip", I64)]


class PcgBatchOptions(C.Structure):
    _fields_ = [("precision", I), ("philox_seed", C.c_uint64), ("block", I),
                ("steps_per_launch", I), ("use_graph", I), ("device", I)]


class PcgSawOptions(C.Structure):
    _fields_ = [("bond", D), ("intersection", D), ("box_size", D), ("density", D),
                ("pbc", I), ("max_tries", I)]


class PcgServerOptions(C.Structure):
    _fields_ = [("precision", I), ("hbm_budget_bytes", I64), ("max_resident", I),
                ("chunk_steps", I), ("philox_seed", C.c_uint64), ("device", I)]


class PcgError(RuntimeError):
    pass


_V = C.c_void_p
# argument types of every entry point of include/pcg.h (all return int
# status codes except pcg_last_error, which returns the message)
SIGNATURES = {
    "pcg_last_error": [],
    "pcg_device_count": [C.POINTER(C.c_int)],
    "pcg_create": [C.POINTER(PcgParams), C.POINTER(PcgModel), C.POINTER(PcgOptions),
                   C.POINTER(_V)],
    "pcg_destroy": [_V], "pcg_step": [_V, I],
    "pcg_nl_verify": [_V, C.POINTER(I)], "pcg_fix_nl": [_V], "pcg_prepare_nh": [_V],
    "pcg_eval_forces": [_V],
    "pcg_eval_term": [_V, I, I, I, C.POINTER(D), C.POINTER(D)],
    "pcg_term_size": [_V, I, C.POINTER(I)],
    "pcg_qa_finish": [_V], "pcg_lang_step": [_V],
    "pcg_get": [_V, I, C.POINTER(D)], "pcg_set": [_V, I, C.POINTER(D)],
    "pcg_get_energy": [_V, C.POINTER(D), C.POINTER(D)],
    "pcg_report": [_V, C.POINTER(D)],
    "pcg_get_time": [_V, C.POINTER(D), C.POINTER(I64)],
    "pcg_set_time": [_V, D], "pcg_set_noise_scale": [_V, D],
    "pcg_get_list": [_V, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I),
                     C.POINTER(D), C.POINTER(U8)],
    "pcg_nat_cont_state": [_V, C.POINTER(U8), C.POINTER(D), C.POINTER(I)],
    "pcg_qa_state": [_V, C.POINTER(C.c_int8), C.POINTER(I)],
    "pcg_nl_rebuilds": [_V, C.POINTER(I)],
    "pcg_use_graph": [_V, I],
    "pcg_stream": [_V, C.POINTER(_V)],
    "pcg_batch_create": [C.POINTER(PcgParams), I, C.POINTER(PcgModel), C.POINTER(D),
                         C.POINTER(I), C.POINTER(PcgBatchOptions), C.POINTER(_V)],
    "pcg_batch_destroy": [_V], "pcg_batch_run": [_V, I], "pcg_batch_sync": [_V],
    "pcg_batch_get": [_V, I, I, C.POINTER(D)],
    "pcg_batch_set": [_V, I, I, C.POINTER(D)],
    "pcg_batch_energy": [_V, I, C.POINTER(D), C.POINTER(D)],
    "pcg_batch_info": [_V, C.POINTER(I), C.POINTER(I), C.POINTER(I64)],
    "pcg_batch_last_ms": [_V, C.POINTER(D), C.POINTER(I)],
    "pcg_server_create": [C.POINTER(PcgParams), C.POINTER(PcgServerOptions),
                          C.POINTER(_V)],
    "pcg_server_destroy": [_V],
    "pcg_server_submit": [_V, C.POINTER(PcgModel), D, I, I64, C.POINTER(I64)],
    "pcg_server_round": [_V, C.POINTER(I)],
    "pcg_server_status": [_V, I64, C.POINTER(I), C.POINTER(I64)],
    "pcg_server_result": [_V, I64, C.POINTER(D), C.POINTER(D)],
    "pcg_server_cancel": [_V, I64],
    "pcg_server_get": [_V, I64, I, C.POINTER(D)],
    "pcg_server_set_noise_scale": [_V, D],
    "pcg_checkpoint_size": [_V, C.POINTER(I64)],
    "pcg_checkpoint": [_V, _V, I64],
    "pcg_restore": [_V, _V, I64],
    "pcg_server_stats": [_V, C.POINTER(I64), C.POINTER(I), C.POINTER(I)],
    "pcg_server_job_times": [_V, I64, C.POINTER(D), C.POINTER(D), C.POINTER(D)],
    "pcg_batch_set_noise_scale": [_V, D],
    "pcg_batch_energies": [_V, I],
    "pcg_batch_report": [_V, C.POINTER(D)],
    "pcg_batch_report_static": [_V, C.POINTER(D)],
    "pcg_batch_checkpoint_size": [_V, C.POINTER(I64)],
    "pcg_batch_checkpoint": [_V, _V, I64],
    "pcg_batch_restore": [_V, _V, I64],
    "pcg_batch_sizes": [_V, I, C.POINTER(I)],
    "pcg_batch_autotune": [_V, I, C.POINTER(I), C.POINTER(D)],
    "pcg_stream_copy": [I64, I, I, C.POINTER(D), C.POINTER(D)],
    "pcg_set_tips": [C.c_void_p, I, C.POINTER(I), C.POINTER(D), C.POINTER(D), I, C.POINTER(I),
                     C.POINTER(D)],
    "pcg_afm_stats": [C.c_void_p, I, C.POINTER(I), C.POINTER(D), C.POINTER(D), C.POINTER(U8)],
    "pcg_batch_set_tips": [C.c_void_p, I, I, C.POINTER(I), C.POINTER(D), C.POINTER(D), I,
                           C.POINTER(I), C.POINTER(D)],
    "pcg_morph_saw": [C.POINTER(PcgSawOptions), I, C.POINTER(I), I, C.POINTER(D),
                      C.POINTER(I64)],
    "pcg_step_form": [_V, I, C.POINTER(I)],
    "pcg_get3": [_V, I, C.POINTER(D), C.POINTER(D), C.POINTER(D), I64],
    "pcg_set3": [_V, I, C.POINTER(D), C.POINTER(D), C.POINTER(D), I64],
    "pcg_eval_term3": [_V, I, I, I, C.POINTER(D), C.POINTER(D), C.POINTER(D), I64,
                       C.POINTER(D)],
    "pcg_box_setup": [_V, C.POINTER(PcgBox)],
    "pcg_box_start": [_V],
    "pcg_box_state": [_V, C.POINTER(D)],
    "pcg_ljw_process": [_V],
    "pcg_wall_conns": [_V, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(D),
                       C.POINTER(D)],
}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PcgError(f"HIP library not built: {LIB_PATH} (run pas-cg_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        L.pcg_last_error.restype = C.c_char_p
        L.pcg_last_error.argtypes = []
        for name, args in SIGNATURES.items():
            if name == "pcg_last_error":
                continue
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        raise PcgError(f"pcg error {rc}: {lib().pcg_last_error().decode()}")


def fill_struct(st, p: dict):
    for name, ctype in st._fields_:
        if name not in p:
            continue
        v = p[name]
        if hasattr(ctype, "_length_"):
            arr = getattr(st, name)
            flat = np.asarray(v).reshape(-1)
            inner = ctype._type_
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


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _tip_args(tips: dict):
    """pcg_set_tips arguments from a dict with tip_res/tip_orig/tip_vel and
    ftip_res/ftip_force (missing keys: no tips of that kind); the arrays are
    returned too, to keep them alive for the call."""
    vr = np.ascontiguousarray(tips.get("tip_res", []), np.int32)
    vo = np.ascontiguousarray(np.reshape(tips.get("tip_orig", np.zeros((0, 3))), (-1, 3)), np.float64)
    vv = np.ascontiguousarray(np.reshape(tips.get("tip_vel", np.zeros((0, 3))), (-1, 3)), np.float64)
    fr = np.ascontiguousarray(tips.get("ftip_res", []), np.int32)
    ff = np.ascontiguousarray(np.reshape(tips.get("ftip_force", np.zeros((0, 3))), (-1, 3)), np.float64)
    keep = (vr, vo, vv, fr, ff)
    return keep, (len(vr), _p(vr, I), _p(vo, D), _p(vv, D), len(fr), _p(fr, I), _p(ff, D))


class _ModelRef:
    """Keeps the numpy arrays behind a PcgModel alive."""

    def __init__(self, m: dict):
        self.keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(np.asarray(x, dt))
            self.keep.append(a)
            return a

        s = PcgModel()
        s.n = int(m["n"])
        s.pos = _p(arr(m["pos"], np.float64), D)
        s.orig_pos = _p(arr(m["orig_pos"], np.float64), D)
        s.atype = _p(arr(m["atype"], np.uint8), U8)
        s.chain_idx = _p(arr(m["chain"], np.int32), I)
        s.seq_idx = _p(arr(m["seq"], np.int32), I)
        s.n_teth = len(m["teth"])
        s.teth = _p(arr(m["teth"], np.int32), I)
        s.n_ang = len(m["ang"])
        s.ang = _p(arr(m["ang"], np.int32), I)
        s.ang_has = _p(arr(m["ang_has"], np.uint8), U8)
        s.ang_theta = _p(arr(m["ang_theta"], np.float64), D)
        s.n_dih = len(m["dih"])
        s.dih = _p(arr(m["dih"], np.int32), I)
        s.dih_has = _p(arr(m["dih_has"], np.uint8), U8)
        s.dih_phi = _p(arr(m["dih_phi"], np.float64), D)
        s.n_cont = len(m["cont"])
        s.cont = _p(arr(m["cont"], np.int32), I)
        s.cont_len = _p(arr(m["cont_len"], np.float64), D)
        s.cont_type = _p(arr(m["cont_type"], np.int32), I)
        for k in range(3):
            s.cell[k] = float(m["cell"][k])
        s.n_tips = len(m["tip_res"])
        s.tip_res = _p(arr(m["tip_res"], np.int32), I)
        s.tip_orig = _p(arr(m["tip_orig"], np.float64), D)
        s.tip_vel = _p(arr(m["tip_vel"], np.float64), D)
        ftip_res = m.get("ftip_res", np.zeros(0, np.int32))
        s.n_ftips = len(ftip_res)
        s.ftip_res = _p(arr(ftip_res, np.int32), I)
        s.ftip_force = _p(arr(m.get("ftip_force", np.zeros((0, 3))), np.float64), D)
        self.s = s


class Context:
    """Plain single-system context: the drop-in per-step loop."""

    def __init__(self, params: dict, model: dict, precision=MIXED, noise=PHILOX,
                 seed: int = 448, philox_seed: int = 1, temperature: float = -1.0,
                 device: int = 0, ran2_skip: int = 0):
        L = lib()
        self.n = int(model["n"])
        self._m = _ModelRef(model)
        self._p = fill_struct(PcgParams(), params)
        o = PcgOptions(precision, noise, seed, philox_seed, temperature, device, ran2_skip)
        h = C.c_void_p()
        check(L.pcg_create(C.byref(self._p), C.byref(self._m.s), C.byref(o), C.byref(h)))
        self.h = h
        self.L = L

    def close(self):
        if getattr(self, "h", None):
            self.L.pcg_destroy(self.h)
            self.h = None

    __del__ = close

    def step(self, k: int = 1):
        check(self.L.pcg_step(self.h, k))

    def use_graph(self, on=True):
        check(self.L.pcg_use_graph(self.h, int(on)))

    def step_form(self, form: int = FORM_AUTO) -> int:
        """pcg_step_form: FORM_AUTO, FORM_STAGES or FORM_PERSISTENT; returns
        the form the next step takes."""
        v = C.c_int32()
        check(self.L.pcg_step_form(self.h, form, C.byref(v)))
        return v.value

    def nl_verify(self) -> bool:
        v = C.c_int32()
        check(self.L.pcg_nl_verify(self.h, C.byref(v)))
        return bool(v.value)

    def fix_nl(self):
        check(self.L.pcg_fix_nl(self.h))

    def prepare_nh(self):
        check(self.L.pcg_prepare_nh(self.h))

    def eval_forces(self):
        check(self.L.pcg_eval_forces(self.h))

    def eval_term(self, term, frm=0, to=-1):
        F = np.zeros((self.n, 3))
        V = C.c_double()
        tid = TERM_ID[term] if isinstance(term, str) else int(term)
        check(self.L.pcg_eval_term(self.h, tid, frm, to, _p(F, D), C.byref(V)))
        return F, V.value

    def term_size(self, term) -> int:
        v = C.c_int32()
        check(self.L.pcg_term_size(self.h, TERM_ID[term], C.byref(v)))
        return v.value

    def qa_finish(self):
        check(self.L.pcg_qa_finish(self.h))

    def lang_step(self):
        check(self.L.pcg_lang_step(self.h))

    def get(self, which: int) -> np.ndarray:
        out = np.zeros((self.n, 3))
        check(self.L.pcg_get(self.h, which, _p(out, D)))
        return out

    def get3(self, which: int) -> np.ndarray:
        """pcg_get3 into a structure of arrays: [3, n] (x, y, z rows)."""
        out = np.zeros((3, self.n))
        check(self.L.pcg_get3(self.h, which, _p(out[0], D), _p(out[1], D), _p(out[2], D), 1))
        return out

    def set3(self, which: int, soa):
        a = np.ascontiguousarray(np.asarray(soa, np.float64))
        check(self.L.pcg_set3(self.h, which, _p(a[0], D), _p(a[1], D), _p(a[2], D), 1))

    def eval_term3(self, term, frm=0, to=-1):
        """pcg_eval_term3: the term's forces as [3, n] and its energy."""
        F = np.zeros((3, self.n))
        V = C.c_double()
        tid = TERM_ID[term] if isinstance(term, str) else int(term)
        check(self.L.pcg_eval_term3(self.h, tid, frm, to, _p(F[0], D), _p(F[1], D), _p(F[2], D),
                                    1, C.byref(V)))
        return F, V.value

    def r(self):
        return self.get(0)

    def F(self):
        return self.get(6)

    def set(self, which: int, a):
        a = np.ascontiguousarray(np.asarray(a, np.float64))
        check(self.L.pcg_set(self.h, which, _p(a, D)))

    def energy(self):
        V = C.c_double()
        tv = np.zeros(len(TERMS))
        check(self.L.pcg_get_energy(self.h, C.byref(V), _p(tv, D)))
        return V.value, dict(zip(TERMS, tv))

    def report(self) -> dict:
        """The trajectory's report scalars (REPORT_FIELDS) of the current state."""
        out = np.zeros(len(REPORT_FIELDS))
        check(self.L.pcg_report(self.h, _p(out, D)))
        return dict(zip(REPORT_FIELDS, out.tolist()))

    def time(self):
        t = C.c_double()
        s = C.c_int64()
        check(self.L.pcg_get_time(self.h, C.byref(t), C.byref(s)))
        return t.value, s.value

    def set_time(self, t):
        check(self.L.pcg_set_time(self.h, t))

    def set_noise_scale(self, s):
        check(self.L.pcg_set_noise_scale(self.h, s))

    def set_tips(self, tips: dict):
        """Replaces the AFM tips (pcg_set_tips); an empty dict turns AFM off."""
        keep, args = _tip_args(tips)
        check(self.L.pcg_set_tips(self.h, *args))

    def afm_stats(self) -> dict:
        """The tips' moving averages (pcg_afm_stats): avg [tips, 3], perp
        [tips], has [tips]; velocity tips first, then force tips."""
        size = C.c_int32()
        check(self.L.pcg_afm_stats(self.h, 0, C.byref(size), None, None, None))
        k = size.value
        avg, perp, has = np.zeros((k, 3)), np.zeros(k), np.zeros(k, np.uint8)
        if k:
            check(self.L.pcg_afm_stats(self.h, k, C.byref(size), _p(avg, D), _p(perp, D),
                                       _p(has, U8)))
        return {"avg": avg, "perp": perp, "has": has.astype(bool)}

    def list(self, kind: int):
        size = C.c_int32()
        check(self.L.pcg_get_list(self.h, kind, 0, C.byref(size), None, None, None, None))
        n = size.value
        i1 = np.zeros(n, np.int32)
        i2 = np.zeros(n, np.int32)
        val = np.zeros(n)
        flag = np.zeros(n, np.uint8)
        check(self.L.pcg_get_list(self.h, kind, n, C.byref(size), _p(i1, I), _p(i2, I),
                                  _p(val, D), _p(flag, U8)))
        return i1, i2, val, flag

    def nat_cont_state(self, k: int):
        a = np.zeros(max(k, 1), np.uint8)
        ct = np.zeros(max(k, 1))
        nc = C.c_int32()
        check(self.L.pcg_nat_cont_state(self.h, _p(a, U8), _p(ct, D), C.byref(nc)))
        return a[:k], ct[:k], nc.value

    def qa_state(self):
        sync = np.zeros((self.n, 4), np.int8)
        num = C.c_int32()
        check(self.L.pcg_qa_state(self.h, _p(sync, C.c_int8), C.byref(num)))
        return sync, num.value

    def checkpoint(self) -> bytes:
        """The context's whole state, with its ran2 stream (restore())."""
        n = I64()
        check(self.L.pcg_checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(self.L.pcg_checkpoint(self.h, buf, n.value))
        return buf.raw

    def restore(self, ck: bytes):
        """Resumes from a checkpoint of a context of the same model,
        parameters, precision and noise source."""
        buf = C.create_string_buffer(ck, len(ck))
        check(self.L.pcg_restore(self.h, buf, len(ck)))

    def box_setup(self, box: dict):
        """The simulation box and its walls (pcg_box_setup); box from
        params.box_params."""
        self._box = fill_struct(PcgBox(), box)
        check(self.L.pcg_box_setup(self.h, C.byref(self._box)))

    def box_start(self):
        """Hands the trajectory over to the box protocol (pcg_box_start)."""
        check(self.L.pcg_box_start(self.h))

    def box_state(self, raw: bool = False):
        out = np.zeros(BOX_STATE_NUM)
        check(self.L.pcg_box_state(self.h, _p(out, D)))
        return out if raw else box_state_dict(out)

    def ljw_process(self):
        check(self.L.pcg_ljw_process(self.h))

    def wall_conns(self, kind: int) -> dict:
        """kind 0 harmonic, 1 LJ: residue, plane, bead offset, saturation."""
        size = C.c_int32()
        check(self.L.pcg_wall_conns(self.h, kind, 0, C.byref(size), None, None, None, None))
        k = size.value
        res, wall = np.zeros(k, np.int32), np.zeros(k, np.int32)
        off, sat = np.zeros((k, 3)), np.zeros(k)
        if k:
            check(self.L.pcg_wall_conns(self.h, kind, k, C.byref(size), _p(res, I), _p(wall, I),
                                        _p(off, D), _p(sat, D)))
        return {"res": res, "wall": wall, "off": off, "sat": sat}

    def stream(self) -> int:
        """The HIP stream every launch of this context is enqueued on."""
        st = C.c_void_p()
        check(self.L.pcg_stream(self.h, C.byref(st)))
        return st.value or 0

    def nl_rebuilds(self) -> int:
        v = C.c_int32()
        check(self.L.pcg_nl_rebuilds(self.h, C.byref(v)))
        return v.value


class Batch:
    """Many systems advanced together by the persistent kernel."""

    def __init__(self, params: dict, models: list, temps=None, seeds=None, precision=MIXED,
                 philox_seed: int = 1, block: int = 0, steps_per_launch: int = 0,
                 use_graph: bool = False, device: int = 0):
        L = lib()
        self.ns = [int(m["n"]) for m in models]
        self._refs = [_ModelRef(m) for m in models]
        arr = (PcgModel * len(models))(*[r.s for r in self._refs])
        self._arr = arr
        self._p = fill_struct(PcgParams(), params)
        t = np.ascontiguousarray(np.asarray(temps if temps is not None else [-1.0] * len(models),
                                            np.float64))
        s = np.ascontiguousarray(np.asarray(seeds if seeds is not None else
                                            [448 + k for k in range(len(models))], np.int32))
        self._t, self._s = t, s
        o = PcgBatchOptions(precision, philox_seed, block, steps_per_launch, int(use_graph), device)
        h = C.c_void_p()
        check(L.pcg_batch_create(C.byref(self._p), len(models), arr, _p(t, D), _p(s, I),
                                 C.byref(o), C.byref(h)))
        self.h = h
        self.L = L

    def close(self):
        if getattr(self, "h", None):
            self.L.pcg_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, nsteps: int):
        check(self.L.pcg_batch_run(self.h, nsteps))

    def sync(self):
        check(self.L.pcg_batch_sync(self.h))

    def autotune(self, trial_steps: int = 100):
        """Fastest workgroup size for this batch (state left unchanged)."""
        blk = C.c_int32()
        ms = C.c_double()
        check(self.L.pcg_batch_autotune(self.h, trial_steps, C.byref(blk), C.byref(ms)))
        return blk.value, ms.value

    def get(self, sys: int, which: int) -> np.ndarray:
        out = np.zeros((self.ns[sys], 3))
        check(self.L.pcg_batch_get(self.h, sys, which, _p(out, D)))
        return out

    def set(self, sys: int, which: int, a):
        a = np.ascontiguousarray(np.asarray(a, np.float64))
        check(self.L.pcg_batch_set(self.h, sys, which, _p(a, D)))

    def energy(self, sys: int):
        V = C.c_double()
        tv = np.zeros(len(TERMS))
        check(self.L.pcg_batch_energy(self.h, sys, C.byref(V), _p(tv, D)))
        return V.value, dict(zip(TERMS, tv))

    def checkpoint(self) -> bytes:
        """The batch's whole device state (restore with restore())."""
        n = I64()
        check(self.L.pcg_batch_checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(self.L.pcg_batch_checkpoint(self.h, buf, n.value))
        return buf.raw

    def restore(self, ck: bytes):
        """Resumes from a checkpoint of a batch of the same models,
        parameters and precision."""
        buf = C.create_string_buffer(ck, len(ck))
        check(self.L.pcg_batch_restore(self.h, buf, len(ck)))

    def report(self) -> list:
        """The report scalars (REPORT_FIELDS) of every system, one dict each."""
        out = np.zeros((len(self.ns), len(REPORT_FIELDS)))
        check(self.L.pcg_batch_report(self.h, _p(out, D)))
        return [dict(zip(REPORT_FIELDS, row.tolist())) for row in out]

    def report_static(self) -> list:
        """report() through pcg_batch_report_static (no allocation)."""
        out = np.zeros((len(self.ns), len(REPORT_FIELDS)))
        check(self.L.pcg_batch_report_static(self.h, _p(out, D)))
        return [dict(zip(REPORT_FIELDS, row.tolist())) for row in out]

    def info(self):
        b = C.c_int32()
        s = C.c_int32()
        by = C.c_int64()
        check(self.L.pcg_batch_info(self.h, C.byref(b), C.byref(s), C.byref(by)))
        return {"block": b.value, "steps_per_launch": s.value, "bytes": by.value}

    def last_ms(self):
        ms = C.c_double()
        n = C.c_int32()
        check(self.L.pcg_batch_last_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_noise_scale(self, s: float):
        check(self.L.pcg_batch_set_noise_scale(self.h, s))

    def set_tips(self, sys: int, tips: dict):
        """Replaces the AFM tips of one system (pcg_batch_set_tips)."""
        keep, args = _tip_args(tips)
        check(self.L.pcg_batch_set_tips(self.h, sys, *args))

    def energies(self, on: bool = True):
        """Whether the last step of a launch evaluates energies (default on);
        off, every step takes the mixed precision's FP32 force paths."""
        check(self.L.pcg_batch_energies(self.h, int(on)))

    def sizes(self, sys: int) -> dict:
        out = np.zeros(len(TERMS) + 3, np.int32)
        check(self.L.pcg_batch_sizes(self.h, sys, _p(out, I)))
        d = dict(zip(TERMS, (int(x) for x in out[:len(TERMS)])))
        d["nl"] = int(out[len(TERMS)])
        d["n"] = int(out[len(TERMS) + 1])
        d["rebuilds"] = int(out[len(TERMS) + 2])
        return d


JOB_STATUS = {0: "queued", 1: "running", 2: "done", 3: "cancelled", 4: "failed"}


class Server:
    """Continuous batching of simulation jobs (admission by HBM budget)."""

    def __init__(self, params: dict, precision=MIXED, hbm_budget_bytes: int = 8 << 30,
                 max_resident: int = 1024, chunk_steps: int = 1000, philox_seed: int = 1,
                 device: int = 0):
        L = lib()
        self._p = fill_struct(PcgParams(), params)
        o = PcgServerOptions(precision, hbm_budget_bytes, max_resident, chunk_steps,
                             philox_seed, device)
        h = C.c_void_p()
        check(L.pcg_server_create(C.byref(self._p), C.byref(o), C.byref(h)))
        self.h = h
        self.L = L
        self.ns = {}

    def close(self):
        if getattr(self, "h", None):
            self.L.pcg_server_destroy(self.h)
            self.h = None

    __del__ = close

    def submit(self, model: dict, nsteps: int, temperature: float = -1.0, seed: int = 448) -> int:
        ref = _ModelRef(model)
        jid = C.c_int64()
        check(self.L.pcg_server_submit(self.h, C.byref(ref.s), temperature, seed, nsteps,
                                       C.byref(jid)))
        self.ns[jid.value] = int(model["n"])
        return jid.value

    def round(self) -> int:
        a = C.c_int32()
        check(self.L.pcg_server_round(self.h, C.byref(a)))
        return a.value

    def status(self, job: int):
        s = C.c_int32()
        d = C.c_int64()
        check(self.L.pcg_server_status(self.h, job, C.byref(s), C.byref(d)))
        return JOB_STATUS[s.value], d.value

    def result(self, job: int):
        pos = np.zeros((self.ns[job], 3))
        V = C.c_double()
        check(self.L.pcg_server_result(self.h, job, _p(pos, D), C.byref(V)))
        return pos, V.value

    def get(self, job: int, which: int) -> np.ndarray:
        """Positions (0), last forces (6) or velocities (7) of a finished job."""
        out = np.zeros((self.ns[job], 3))
        check(self.L.pcg_server_get(self.h, job, which, _p(out, D)))
        return out

    def set_noise_scale(self, s: float):
        check(self.L.pcg_server_set_noise_scale(self.h, s))

    def cancel(self, job: int):
        check(self.L.pcg_server_cancel(self.h, job))

    def stats(self):
        b = C.c_int64()
        r = C.c_int32()
        q = C.c_int32()
        check(self.L.pcg_server_stats(self.h, C.byref(b), C.byref(r), C.byref(q)))
        return {"resident_bytes": b.value, "resident": r.value, "queued": q.value}

    def times(self, job: int):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self.L.pcg_server_job_times(self.h, job, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


def stream_copy(nbytes: int = 1 << 30, iters: int = 20, device: int = 0):
    """Achieved bandwidth (GB/s, read + write) and time per copy (ms) of a
    streaming copy of `nbytes`."""
    L = lib()
    g = C.c_double()
    ms = C.c_double()
    check(L.pcg_stream_copy(nbytes, iters, device, C.byref(g), C.byref(ms)))
    return g.value, ms.value


def morph_saw(chain_idx, seed: int, bond: float, intersection: float, box_size: float = 0.0,
              density: float = 0.0, pbc: bool = False, max_tries: int = 0):
    """Self-avoiding-walk positions [n, 3] of the chains given by chain_idx,
    drawn from ran2(seed), and the number of draws used (pcg_morph_saw)."""
    L = lib()
    ch = np.ascontiguousarray(chain_idx, np.int32)
    pos = np.zeros((len(ch), 3), np.float64)
    o = PcgSawOptions(bond, intersection, box_size, density, 1 if pbc else 0, max_tries)
    draws = I64()
    check(L.pcg_morph_saw(C.byref(o), len(ch), _p(ch, I), seed, _p(pos, D), C.byref(draws)))
    return pos, draws.value
