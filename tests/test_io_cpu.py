"""Field files (NetCDF classic Ux/Uy/Uz, data_ops.f90:390-447, 1198-1397),
result-file names and their parsing (data_ops.f90:31, function_ops.f90:
4590-4674, initialize.f90:49-70), and the fftw3-mpi.f03 stand-in
(initialize.f90:136-153): library host code, no GPU.  Files written by the
library are read back with scipy.io.netcdf_file and files scipy writes are read
by the library."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lq():
    from maxlq_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmaxlq.so not built (run __graft_entry__.build())")
    from maxlq_amd import lq as L
    return L


def _field(shape, seed=0):
    return np.random.default_rng(seed).standard_normal((3,) + shape)


@pytest.mark.parametrize("layout,cdf,version", [(None, None, 1), ("fixed", "2", 2),
                                                 ("record", None, 2)])
def test_field_file_read_by_scipy(lq, tmp_path, monkeypatch, layout, cdf, version):
    if layout:
        monkeypatch.setenv("MLQ_NC_LAYOUT", layout)
    if cdf:
        monkeypatch.setenv("MLQ_NC_CDF", cdf)
    u = _field((8, 12, 16))
    path = tmp_path / "u_result.nc"
    lq.save_field_nc(path, u)
    with sio.netcdf_file(str(path), "r", mmap=False) as f:
        assert f.version_byte == version
        assert f.dimensions["x"] == 16 and f.dimensions["y"] == 12
        assert f.dimensions["z"] == (None if layout == "record" else 8)
        for c, name in enumerate(("Ux", "Uy", "Uz")):
            v = f.variables[name]
            assert v.dimensions == ("z", "y", "x") and v.typecode() == "d"
            assert np.array_equal(v[:], u[c])
    # and back through the library's reader
    n = 8
    w = _field((n, n, n), 1)
    lq.save_field_nc(path, w)
    assert np.array_equal(lq.read_field_nc(path, n), w)


@pytest.mark.parametrize("version,unlimited", [(1, False), (2, False), (1, True), (2, True)])
def test_library_reads_scipy_files(lq, tmp_path, version, unlimited):
    n = 8
    u = _field((n, n, n), 2)
    path = str(tmp_path / "from_scipy.nc")
    with sio.netcdf_file(path, "w", version=version) as f:
        f.history = "written by scipy"                  # a global attribute to skip
        if unlimited:                                   # scipy defines the record dimension first
            f.createDimension("z", None)
        f.createDimension("x", n)
        f.createDimension("y", n)
        if not unlimited:
            f.createDimension("z", n)
        f.createVariable("extra", "i", ("x",))[:] = np.arange(n)
        for c, name in enumerate(("Ux", "Uy", "Uz")):
            v = f.createVariable(name, "d", ("z", "y", "x"))
            v.units = "1"                               # a variable attribute to skip
            v[:] = u[c]
    assert np.array_equal(lq.read_field_nc(path, n), u)


def test_reader_rejects_other_grids_and_files(lq, tmp_path):
    from maxlq_amd.capi import MlqError
    path = tmp_path / "u.nc"
    lq.save_field_nc(path, _field((8, 8, 8)))
    with pytest.raises(MlqError):
        lq.read_field_nc(path, 16)
    (tmp_path / "junk.nc").write_bytes(b"not a netcdf file")
    with pytest.raises(MlqError):
        lq.read_field_nc(tmp_path / "junk.nc", 8)
    with pytest.raises(MlqError):
        lq.read_field_nc(tmp_path / "missing.nc", 8)


def test_result_filenames(lq):
    # data_ops.f90:31 with lebesgueQTxt (initialize.f90:49-70)
    assert lq.result_filename(4.0, 64, 3, 12, dir_="output/ncFiles/") == \
        "output/ncFiles/u_result_q4_n64_B003_iter00012.nc"
    assert lq.result_filename(9.0, 1024, 22, -1) == "u_result_q9_n1024_B022_iterend.nc"
    assert lq.result_filename(12.0, 128, 1, 0, field_name="d_used") == \
        "d_used_q12_n128_B001_iter00000.nc"
    assert [lq.q_text(q) for q in (3.0, 4.0, 9.0, 12.0, 2.5, 3.75, 150.0)] == \
        ["3", "4", "9", "12", "2.50E+00", "3.75E+00", "1.50E+02"]


def test_parse_result_filenames(lq):
    # the reference's own example (maxdLpdt.f90:63): resume inside B solve 22
    assert lq.parse_result_filename("u_result_0501_q9_n1024_B022_iter00150.nc") == (9.0, 1024, 21, 150)
    # a finished B solve: resume with the next B at iteration 0
    assert lq.parse_result_filename("out/ncFiles/u_result_q4_n64_B003_iterend.nc") == (4.0, 64, 3, 0)
    assert lq.parse_result_filename("u_result_q2.50E+00_n128_B010_iter00007.nc") == (2.5, 128, 9, 7)
    for name in (lq.result_filename(6.0, 256, 5, 40), lq.result_filename(6.0, 256, 5, -1)):
        q, n, b, i = lq.parse_result_filename(name)
        assert (q, n) == (6.0, 256) and (b, i) in ((4, 40), (5, 0))


def test_fftw_mpi_standin(lq):
    """fftw_mpi_local_size_3d through maxlq_amd/fortran/fftw3-mpi.f03, used as
    initialize.f90:131-153 does (one rank: local_N = n3, offset 0)."""
    import ctypes as C
    from maxlq_amd import capi
    lib = capi.load()
    f = lib.mlq_fftw_mpi_local_size_3d
    f.restype = C.c_ssize_t
    f.argtypes = [C.c_ssize_t, C.c_ssize_t, C.c_ssize_t, C.c_int32, C.POINTER(C.c_ssize_t),
                  C.POINTER(C.c_ssize_t)]
    ln, off = C.c_ssize_t(), C.c_ssize_t()
    assert f(32, 16, 8, 91, C.byref(ln), C.byref(off)) == 32 * 16 * 8
    assert (ln.value, off.value) == (32, 0)
    exe = os.path.join(ROOT, "maxlq_amd", "fortran", "mlq_fftwmpi_check")
    if not os.path.exists(exe):
        pytest.skip("Fortran checks not built (amdflang)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["64", "0", "262144"]
