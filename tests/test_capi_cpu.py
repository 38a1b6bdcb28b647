"""CPU-side checks of the C ABI: the library loads and exports every symbol
declared in include/maxlq.h (no compute calls without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "maxlq.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mlq_[a-zA-Z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from maxlq_amd import capi
    assert set(_header_symbols()) == set(capi.EXPORTS)


def test_library_exports_every_symbol():
    from maxlq_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmaxlq.so not built (run __graft_entry__.build())")
    lib = capi.load()
    for s in _header_symbols():
        assert hasattr(lib, s), s


def test_default_params_match_reference_globals():
    from maxlq_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmaxlq.so not built")
    p = capi.default_params(6.0)
    assert (p.q, p.B, p.visc, p.lambda1) == (6.0, 1.0, 1.0, 0.1)
    assert p.toDealias == 1 and p.use_e_u == -1 and p.dealiase_if_mult_by_e_u == 1


def test_missing_library_fails_loudly(tmp_path):
    from maxlq_amd import capi
    saved = capi._lib
    capi._lib = None
    try:
        with pytest.raises(RuntimeError):
            capi.load(str(tmp_path / "nope.so"))
    finally:
        capi._lib = saved
