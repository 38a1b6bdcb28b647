"""The Fortran callers on the GPU: the B-continuation driver
(maxlq_amd/fortran/mlq_continuation.f90, maxdLpdt.f90:458-512) with the
reference's output files, resumed from one of its own field files
(function_ops.f90:4635-4674, optimization.f90:81-86), against the same
solves driven from Python; and the drop-in modules fftwfunction /
mlq_field_ops run against the reference's global_variables module."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "maxlq_amd",
                    "fortran")
N, Q, MAXIT, SAVE = 32, 4.0, 6, 3


def _exe(name):
    path = os.path.join(FDIR, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (amdflang / reference sources)")
    return path


def _run(args, cwd):
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, cwd=cwd)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _B(j):        # maxdLpdt.f90:477-481 for q != 3
    b = 1.0
    for _ in range(j - 1):
        b = b * 10.0 ** (1.0 / 8.0)
    return b


def _python_continuation(u0, js, start_iter=0):
    """The same B solves driven from Python: fields, results and iteration records."""
    from maxlq_amd.lq import MaxLq
    g = MaxLq(N, q=Q, B=_B(js[0]))
    h = g.field(3)
    if u0 is None:
        g.seeded_field(h, 1, "gauss", 16.0)
    else:
        g.upload(h, u0)
    out = []
    for k, j in enumerate(js):
        g.set_params(B=_B(j))
        res, tr = g.maxdLqdt(h, max_iter=MAXIT, optim_tol=0.0, start_iter=start_iter if k == 0 else 0,
                             save_every=SAVE)
        out.append((j, res, tr, g.download(h)))
    g.close()
    return out


def _results(path):
    rows = open(path).read().splitlines()
    assert rows[0].split() == ["B#", "B", "J1", "iter", "enstrophy", "enstrophy_lhs", "q"]
    return [r.split() for r in rows[1:]]


def _iteration_info(path):
    rows = open(path).read().splitlines()
    return [tuple(float(x) for x in r.split()) for r in rows if not r.strip().startswith("Iter")]


def _check_against_python(home, ref, lq):
    res_rows = {int(r[0]): r for r in _results(os.path.join(home, "results_q4.dat"))}
    for j, res, tr, u in ref:
        row = res_rows[j]
        assert abs(float(row[1]) - _B(j)) <= 1e-11 * _B(j)
        assert abs(float(row[2]) - res.J1) <= 1e-11 * abs(res.J1)
        assert int(row[3]) == (res.iters if res.successful else -res.iters)
        assert abs(float(row[4]) - res.enstrophy) <= 1e-11 * abs(res.enstrophy)
        assert abs(float(row[5]) - res.dEnstrophydt) <= 1e-11 * abs(res.dEnstrophydt)
        bdir = [d for d in os.listdir(os.path.join(home, "constraintDirs"))
                if d.startswith(f"B{j:03d}-")]
        assert bdir == [f"B{j:03d}-{_B(j):.1E}"]
        info = _iteration_info(os.path.join(home, "constraintDirs", bdir[0],
                                            f"iteration-info-B{j:03d}.dat"))
        want = [r for r in tr.iteration_info]
        assert [int(r[0]) for r in info[-len(want):]] == [w[0] for w in want]
        for a, w in zip(info[-len(want):], want):
            assert abs(a[3] - w[3]) <= 1e-11 * abs(w[3])
        nc = os.path.join(home, "ncFiles")
        np.testing.assert_allclose(lq.read_field_nc(os.path.join(nc, lq.result_filename(
            Q, N, j, -1)), N), u, rtol=0, atol=1e-13 * np.abs(u).max())
        for it, f in tr.fields:
            got = lq.read_field_nc(os.path.join(nc, lq.result_filename(Q, N, j, it)), N)
            np.testing.assert_allclose(got, f, rtol=0, atol=1e-13 * np.abs(f).max())


def test_continuation_driver_and_resume(tmp_path):
    from maxlq_amd import lq
    exe = _exe("mlq_continuation")
    home = str(tmp_path / "output") + "/"
    _run([exe, home, str(N), "4", "3", str(MAXIT), "0.0", "fp64", str(SAVE)], str(tmp_path))
    ref = _python_continuation(None, [1, 2, 3])
    _check_against_python(home, ref, lq)
    assert [it for it, _ in ref[1][2].fields] == [3, 6]

    # resume inside B solve 2 from its iteration-3 file: B numbers 2 and 3 again
    start = os.path.join(home, "ncFiles", lq.result_filename(Q, N, 2, 3))
    home2 = str(tmp_path / "resumed") + "/"
    out = _run([exe, home2, str(N), "4", "2", str(MAXIT), "0.0", "fp64", str(SAVE), start],
               str(tmp_path))
    assert "bIterOffset" in out
    ref2 = _python_continuation(lq.read_field_nc(start, N), [2, 3], start_iter=3)
    _check_against_python(home2, ref2, lq)
    assert ref2[0][2].iteration_info[0][0] == 3          # the solve starts at iteration 3


def test_dropin_modules_with_reference_globals():
    out = _run([_exe("mlq_dropin_check")], None)
    assert "PASS mlq_dropin_check" in out


def test_continuation_driver_kappa_files(tmp_path):
    """kappa=2: save_kappa_test's files (data_ops.f90:1448-1491) for one B
    solve, against the kappa records of the same solve driven from Python."""
    from maxlq_amd.lq import MaxLq
    exe = _exe("mlq_continuation")
    home = str(tmp_path / "output") + "/"
    _run([exe, home, str(N), "4", "1", "4", "0.0", "fp64", "0", "kappa=2"], str(tmp_path))
    g = MaxLq(N, q=Q, B=_B(1))
    h = g.field(3)
    g.seeded_field(h, 1, "gauss", 16.0)
    res, tr = g.maxdLqdt(h, max_iter=4, optim_tol=0.0, kappa=True, kappa_every=2, kappa_seed=1)
    g.close()
    kdir = os.path.join(home, "constraintDirs", f"B001-{_B(1):.1E}", "kappaTest")
    assert sorted(os.listdir(kdir)) == sorted(
        ["kappa_B001_initial_grad_L2.dat", "kappa_B001_gradL2_iter00002.dat",
         "kappa_B001_gradL2_iter00004.dat", "kappa_B001_end_d1.dat", "kappa_B001_end_gradJ.dat"])
    for it, name, adjusted, tab, frac, best, row in tr.kappa:
        ident = f"gradL2_iter{it:05d}" if name == "gradL2_iter" else name
        rows = open(os.path.join(kdir, f"kappa_B001_{ident}.dat")).read().splitlines()
        vals = np.array([[float(x) for x in r.split()] for r in rows[1:]])
        if adjusted:
            assert rows[0].split()[:5] == ["eps", "deltaJ", "deltaJ/eps", "inner_prod", "kappa_adj"]
            assert vals.shape == (16, 9)
            np.testing.assert_allclose(vals[:, 4], tab[:, 4], rtol=1e-11)
            np.testing.assert_allclose(vals[:, 8], frac, rtol=1e-11)
        else:
            assert rows[0].split() == ["eps", "deltaJ", "deltaJ/eps", "inner_prod", "kappa",
                                       "LOG10|kap-1|"]
            vals = vals[1:]          # the first row is written twice
            assert vals.shape == (16, 6)
            np.testing.assert_allclose(vals[:, 4], tab[:, 3], rtol=1e-11)
        np.testing.assert_allclose(vals[:, 0], tab[:, 0], rtol=1e-11)
        np.testing.assert_allclose(vals[:, 3], tab[:, 1], rtol=1e-11)
