"""1024^3, q = 9, mixed precision: one full RCG iteration of maxdLqdt
(GradL2ForLq, Hilbert gradient, Leray projection, tangential projection, CG
direction, mnbrak + brent line search, vector transport, retraction) on the
seeded field exp(-|m|^2/1024), B = 2.2, within the device's 288 GB, and R(u)
of the result from the mixed path against FP64 on the same field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# MI355X: 288 GiB of HBM3E (the device reports 309.2e9 bytes)
DEVICE_BYTES = 288.0 * 2**30


@pytest.mark.timeout(900)
def test_config5_1024_q9_one_iteration(monkeypatch):
    from maxlq_amd.lq import MaxLq
    monkeypatch.setenv("MLQ_AUTOTUNE", "0")   # default tiles: keeps the test short
    n, q, B = 1024, 9.0, 2.2
    g = MaxLq(n, q=q, B=B, precision="mixed")
    h = g.field(3)
    g.seeded_field(h, 5, "gauss", 1024.0)
    # the resident state does not fit: the reference-structured gradient with
    # the line search from coefficient fields
    path = g.rcg_path()
    print("1024^3 q=9 mixed: iteration path", path)
    assert path == 2
    res, tr = g.maxdLqdt(h, max_iter=1, optim_tol=0.0, batch_linesearch=3)
    used, peak, free, total = g.mem_info()
    _, live_peak = g.mem_live()
    (_, _, _, J0), (it, tau1, beta, J1) = tr.iteration_info[0], tr.iteration_info[-1]
    print(f"1024^3 q=9 mixed: {res.seconds:.1f} s for the solve, tau1 {tau1:.6e}, J0 {J0:.10e}, "
          f"J1 {J1:.10e}, {res.n_J} J requests; device memory in use at its peak "
          f"{live_peak / 1e9:.1f} GB, allocator high-water mark (cached temporaries included) "
          f"{peak / 1e9:.1f} GB, device {total / 1e9:.1f} GB")
    assert res.iters == 1 and it == 1
    assert J1 >= J0                                   # R increases along the step
    assert live_peak < DEVICE_BYTES and peak <= total
    # temporaries carved out of slabs: the allocator's high-water mark leaves
    # room on the device (285 GiB with a pool of blocks keyed by size)
    assert peak < 0.9 * DEVICE_BYTES
    assert tau1 > 0.0
    u = g.download(h)
    Pm = g.calc_local_dLqdt_inclParts(u)
    g.close()
    g64 = MaxLq(n, q=q, B=B, precision="fp64")
    P64 = g64.calc_local_dLqdt_inclParts(u)
    L64 = g64.calc_global_Lq_norm(u)
    g64.close()
    dR = abs(Pm[0] - P64[0]) / abs(P64[0])
    print(f"R of the iterate fp64 {P64} mixed {Pm}: same-field |dR|/R {dR:.2e}; "
          f"||u||_q = {L64:.15f} (B = {B})")
    # the retracted field is stored in FP32: its norm is B to FP32 storage accuracy
    assert abs(L64 - B) / B < 1.0e-7
    assert abs(J1 - Pm[0]) <= 1.0e-12 * abs(Pm[0])    # J1 = R(u_new) as the solve reports it
    assert dR < 1.0e-6
