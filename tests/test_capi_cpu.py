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


def _header_constants():
    txt = open(os.path.join(ROOT, "include", "maxlq.h")).read()
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define\s+(MLQ_[A-Z0-9_]+)\s+(-?\d+)\b", txt)}


def test_constants_agree_across_bindings():
    """Every integer constant of the header that the Python or the Fortran
    binding restates has the header's value there (flags, kinds, codes)."""
    from maxlq_amd import capi
    hdr = _header_constants()
    assert "MLQ_RCG_LEAN" in hdr and "MLQ_RCG_REFERENCE" in hdr
    py = {k: getattr(capi, k) for k in hdr if hasattr(capi, k)}
    assert py, "no constant of the header in capi.py"
    for k, v in py.items():
        assert v == hdr[k], (k, v, hdr[k])
    f90 = open(os.path.join(ROOT, "maxlq_amd", "fortran", "maxlq_c.f90")).read()
    ft = {m.group(1): int(m.group(2)) for m in re.finditer(r"\b(MLQ_[A-Z0-9_]+)\s*=\s*(-?\d+)", f90)}
    assert ft, "no constant in the Fortran interface"
    for k, v in ft.items():
        if k in hdr:
            assert v == hdr[k], (k, v, hdr[k])
    for k in ("MLQ_RCG_REFERENCE", "MLQ_RCG_KAPPA", "MLQ_RCG_LEAN"):
        assert py.get(k) == hdr[k] and ft.get(k) == hdr[k], k


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


def test_fortran_interface_compiles(tmp_path):
    """maxlq_c.f90 (the iso_c_binding block for include/maxlq.h) compiles."""
    import shutil
    import subprocess
    fc = "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    src = os.path.join(ROOT, "maxlq_amd", "fortran", "maxlq_c.f90")
    r = subprocess.run([fc, "-c", src, "-J", str(tmp_path), "-o", str(tmp_path / "m.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_fortran_dropin_compiles_against_reference_globals(tmp_path):
    """The drop-in fftwfunction / mlq_field_ops modules compile against the
    reference's own global_variables module (only where the reference is
    mounted; objects go to a temporary directory)."""
    import subprocess
    fc = "/opt/rocm/bin/amdflang"
    gv = "/root/reference/global_variables.f90"
    if not (os.path.exists(fc) and os.path.exists(gv)):
        pytest.skip("amdflang or reference sources not available")
    fdir = os.path.join(ROOT, "maxlq_amd", "fortran")
    for src in (gv, os.path.join(fdir, "maxlq_c.f90"), os.path.join(fdir, "fftwfunction_gpu.f90"),
                os.path.join(fdir, "mlq_field_ops.f90")):
        r = subprocess.run([fc, "-c", src, "-J", str(tmp_path), "-o",
                            str(tmp_path / (os.path.basename(src) + ".o"))],
                           capture_output=True, text=True, cwd=str(tmp_path))
        assert r.returncode == 0, (src, r.stderr)


def test_init_rejects_unsupported_grids():
    """Non-cubic grids and sizes that are not a power of two in [8, 1024]
    return MLQ_EUNSUP before any device call."""
    import ctypes as C
    from maxlq_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmaxlq.so not built (run __graft_entry__.build())")
    lib = capi.load()
    p = capi.Params()
    lib.mlq_default_params(C.byref(p), 4.0)
    for dims in ((48, 48, 48), (64, 64, 32), (4, 4, 4), (2048, 2048, 2048), (96, 96, 96)):
        ctx = C.c_void_p()
        assert lib.mlq_init(C.byref(ctx), *dims, C.byref(p)) == capi.MLQ_EUNSUP, dims
        assert not ctx.value
