"""The C-ABI library loads and exports every entry point the public header
declares (no compute: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from pas_cg_amd import pcg as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "pcg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w ]+\*?\s*\b(pcg_\w+)\s*\(", src, re.M)))


def test_header_declares_the_interface():
    names = declared()
    for must in ("pcg_create", "pcg_step", "pcg_eval_term", "pcg_batch_run",
                 "pcg_server_submit", "pcg_last_error"):
        assert must in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    if not os.path.exists(G.LIB_PATH):
        pytest.skip("libpcg.so not built")
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    if not os.path.exists(G.LIB_PATH):
        pytest.skip("libpcg.so not built")
    L = G.lib()
    bound = set(G.SIGNATURES)
    assert set(declared()) <= bound, sorted(set(declared()) - bound)
    assert L is G.lib()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(G, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(G, "_lib", None)
    with pytest.raises(G.PcgError):
        G.lib()


def _enum_names(first: str, last: str):
    src = open(os.path.join(ROOT, "include", "pcg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    body = src[src.index(first):src.index(last)]
    return [n.split("=")[0].strip() for n in body.split(",") if n.strip()]


def test_term_and_report_enums_match_the_python_mirror():
    """The term ids and the report fields of pcg.h, in order, are the ones the
    Python mirror indexes with (TERMS, REPORT_FIELDS)."""
    terms = _enum_names("PCG_CHIR", "PCG_NUM_TERMS")
    assert [t[4:].lower() for t in terms] == G.TERMS
    fields = _enum_names("PCG_REP_TIME", "PCG_REPORT_NUM")
    assert len(fields) == len(G.REPORT_FIELDS)
    assert fields[0] == "PCG_REP_TIME" and fields[-1] == "PCG_REP_PID_S1S2"


def test_oracle_terms_follow_the_library_order():
    """The oracle's term ids (oracle/oracle.h) are the library's, so per-term
    parity compares like with like."""
    import oracle_ffi
    assert oracle_ffi.TERMS == G.TERMS
    src = open(os.path.join(ROOT, "oracle", "oracle.h")).read()
    body = src[src.index("ORA_CHIR"):src.index("ORA_NUM_TERMS")]
    names = [n.split("=")[0].strip()[4:].lower() for n in body.split(",") if n.strip()]
    assert names == G.TERMS
