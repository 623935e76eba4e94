"""ctypes binding of oracle/_ref/libpcgref.so: the reference's own kernel
sources compiled by oracle/Makefile and driven by oracle/ref_harness.cpp.
Test infrastructure; absent when the reference tree was not available at
build time (the tests using it are then skipped)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from oracle_ffi import ROOT, OraParams, fill_struct

REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libpcgref.so")
D, I, U8, I8 = C.c_double, C.c_int32, C.c_uint8, C.c_int8
PD, PI, PU8, PI8 = C.POINTER(D), C.POINTER(I), C.POINTER(U8), C.POINTER(I8)
PP = C.POINTER(OraParams)

_lib = None


def available() -> bool:
    return os.path.exists(REF_PATH)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(REF_PATH)
        sig = {
            "ref_chir": [I, PD, I, PI, PD, PD, D, PD],
            "ref_tether": [I, PD, I, PI, PD, D, D, PD],
            "ref_lrep": [I, PD, I, PI, D, D, PD],
            "ref_pauli": [I, PD, PD, I, PI, D, D, PD],
            "ref_nat_cont": [I, PD, PD, I, PI, PD, PI, PU8, PD, D, PP, PI, PD],
            "ref_dh": [I, I, PD, PD, I, PI, PD, PP, PD],
            "ref_nat_ang": [I, PD, I, PI, PD, D, PD],
            "ref_heur_ang": [I, PD, I, PI, PI, PP, PD],
            "ref_nat_dih": [I, I, PD, I, PI, PD, PP, PD],
            "ref_heur_dih": [I, PD, I, PI, PI, PP, PD],
            "ref_pid": [I, PD, PD, PU8, I, PI, PD, PI, PP, PD],
            "ref_vel_afm": [I, PD, D, I, PI, PD, PD, D, D, PD],
            "ref_force_afm": [I, D, I, PI, PD, PD],
            "ref_solid_wall": [I, PD, PP, PD],
            "ref_hw_free": [I, PD, PP, PU8, PD, PD],
            "ref_hw_conn": [I, PD, PP, I, PI, PI, PD, PD, PD, PD],
            "ref_ljw_sift": [I, PD, PP, PU8, PI, PD, PD, I, PI, PI, PD, PI],
            "ref_ljw_conn": [I, PD, PP, I, PI, PI, PD, PD, PU8, PD, PD, I, PI, PI],
            "ref_qa_round": [I, PD, PD, PU8, PI, PP, D, PI8, PI, PI, PD, PU8, PI, PI, PD,
                             PI, PD, PI8, PU8, PI, I, PD],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = D
        L.ref_fix_nl.argtypes = [I, PD, PD, PU8, PI, PI, PP, I, PI, PD, I, PI, PI, PD,
                                 PU8, PI, PI, PI, PI, PI]
        L.ref_fix_nl.restype = None
        L.ref_pid_count.argtypes = [I, PD, PD, PU8, PI, I, PI, PD, PI, PP, PI]
        L.ref_pid_count.restype = None
        L.ref_ljw_process.argtypes = [I, PD, PP, I, PI, PI, PD, PD, PU8, I, PI, I, PI, PI, PD,
                                      PU8, PI, I, PI, PI, PI, PD, PD, PU8]
        L.ref_ljw_process.restype = None
        L.ref_log_forces.argtypes = [I, PD, PD, PD, PD, PI, I, PI, D, PD, PD, PD, PD, PD]
        L.ref_log_forces.restype = None
        _lib = L
    return _lib


def p(a, ct):
    return np.ascontiguousarray(a).ctypes.data_as(C.POINTER(ct))


def params_struct(prm: dict) -> OraParams:
    return fill_struct(OraParams(), prm)
