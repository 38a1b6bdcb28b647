"""Write the golden fixtures of the full-size parity cases from the FP64 oracle.

    python tests/golden/make_golden.py rcg64      # RCG solve at 64^3 (q = 4)
    python tests/golden/make_golden.py fd256      # R, gradient and FD at 256^3

rcg64: 64^3, q = 4, nu = 1, seed 1, spectrum exp(-|m|^2/16), B = 1; one RCG
solve (optimization.f90:7-514) with MAX_ITER = 50, OPTIM_TOL = 1e-8.  The
fixture holds the per-iteration record (tau bracket, tau, beta, J0, J1, the
number of line-search evaluations, message codes) and probes of the final
field (L2 and Lq norms, R parts, a few point values).

fd256: 256^3, q = 4, seed 2, spectrum |m|^2 exp(-|m|^2/64), B = 1.  R parts
(calc_local_dLqdt_inclParts), G = GradL2ForLq(u), H = HilbertGradient(G,
3/2 - 1/q), and for 8 directions phi_i (seeds 1001..1008, same spectrum,
divergence-free, scaled to ||phi_i||_2 = ||u||_2): <G, phi_i>_L2, <H,
phi_i>_L2 and the central difference (R(u + h phi_i) - R(u - h phi_i)) / 2h
with h = 1e-6 (optimization.f90:1044-1069, central variant).

The fixtures are data only (inputs are regenerated from the seeds by the
seeded generator on either side).
"""
from __future__ import annotations

import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import lq_oracle as L          # noqa: E402
from oracle.seeded import seeded_field    # noqa: E402


def probes(orc, u):
    n = u.shape[-1]
    R = orc.calc_local_dLqdt_inclParts(u)
    pts = [(0, 0, 0, 0), (1, n // 3, n // 5, n // 7), (2, n - 1, n // 2, 3)]
    return {
        "L2sq": float(orc.global_summed_field_inner_product(u, u, L.INNER_L2)),
        "Lq": float(orc.calc_global_Lq_norm(u)),
        "R_parts": [float(x) for x in R],
        "points": [[c, z, y, x, float(u[c, z, y, x])] for c, z, y, x in pts],
    }


def rcg64():
    n, q, B = 64, 4.0, 1.0
    u0 = seeded_field(n, 1, "gauss", 16.0, q=q, B=B)
    orc = L.Oracle(L.Params(n=n, q=q, B=B, MAX_ITER=50, OPTIM_TOL=1e-8))
    t0 = time.time()

    def on_iter(rec, u):
        print(f"iter {rec.iter} J1={rec.J1:.15e} tau={rec.tau:.6e} nJ={rec.nJ} "
              f"({time.time() - t0:.0f}s)", flush=True)

    res = L.maxdLqdt(orc, u0, on_iter=on_iter)
    recs = [{"iter": r.iter, "tau_brack": [float(x) for x in r.tau_brack], "tau": r.tau,
             "beta": r.beta, "J0": r.J0, "J1": r.J1, "deltaJ": r.deltaJ, "nJ": r.nJ,
             "msgs": list(r.msgs)} for r in res.records]
    out = {"case": "rcg64", "n": n, "q": q, "B": B, "seed": 1, "envelope": ["gauss", 16.0],
           "max_iter": 50, "optim_tol": 1e-8, "J1": res.J1, "iters": res.iters,
           "successful": bool(res.successful), "while_flag": res.while_flag,
           "stop_reason": res.stop_reason, "records": recs, "final": probes(orc, res.u)}
    with open(os.path.join(HERE, "rcg64_q4.json"), "w") as f:
        json.dump(out, f, indent=1)


def fd256():
    n, q, B, h = 256, 4.0, 1.0, 1e-6
    orc = L.Oracle(L.Params(n=n, q=q, B=B))
    u = seeded_field(n, 2, "k2gauss", 64.0, q=q, B=B)
    t0 = time.time()
    R = orc.calc_local_dLqdt_inclParts(u)
    G = orc.GradL2ForLq(u)
    H = orc.HilbertGradient(G.copy(), 1.5 - 1.0 / q)
    print(f"R, G, H done ({time.time() - t0:.0f}s)", flush=True)
    ip = orc.global_summed_field_inner_product
    uu = ip(u, u, L.INNER_L2)
    dirs = []
    for i in range(8):
        phi = seeded_field(n, 1001 + i, "k2gauss", 64.0, q=q, B=1.0)
        phi *= math.sqrt(uu / ip(phi, phi, L.INNER_L2))
        Gp = ip(G, phi, L.INNER_L2)
        Hp = ip(H, phi, L.INNER_L2)
        Rp = orc.eval_J(u + h * phi, "maxdLqdt")
        Rm = orc.eval_J(u - h * phi, "maxdLqdt")
        dirs.append({"seed": 1001 + i, "G_phi": Gp, "H_phi": Hp, "R_plus": Rp, "R_minus": Rm,
                     "fd": (Rp - Rm) / (2.0 * h)})
        print(f"dir {i}: <G,phi>={Gp:.15e} fd={(Rp - Rm) / (2 * h):.15e} "
              f"({time.time() - t0:.0f}s)", flush=True)
    out = {"case": "fd256", "n": n, "q": q, "B": B, "seed": 2, "envelope": ["k2gauss", 64.0],
           "h": h, "R_parts": [float(x) for x in R], "L2sq_u": uu,
           "G_L2sq": ip(G, G, L.INNER_L2), "H_L2sq": ip(H, H, L.INNER_L2), "directions": dirs}
    with open(os.path.join(HERE, "fd256_q4.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    L.set_threads(int(os.environ.get("ORACLE_THREADS", "4")))
    {"rcg64": rcg64, "fd256": fd256}[sys.argv[1]]()
