"""Profiling driver for the apply stage of the batched matvec (FP32 by default).

Operator: d=5 regular sparse grid of level 7 (N = 61,183), exp-sine coefficient with
16 xi drawn as in bench.py (PCG64 seed 20261203), --batch instances; or, with --grid, an
adaptive grid from an .npz (sparse path) with a = 1 + |x|^2.  The first
matvec autotunes the apply block sizes; then --reps matvecs follow, each timed with
HIP events around the apply launches (lsg_last_timing).  Finally one elementwise
kernel y = 2 x over --copy-mib MiB of float64 runs, whose HBM traffic is known
(reads and writes --copy-mib MiB each): it calibrates FETCH_SIZE / WRITE_SIZE.

Meant to run under rocprofv3, e.g.
    rocprofv3 --kernel-trace --stats -d OUT -o mv -- python3 loadsg_amd/profile_matvec.py
    rocprofv3 --pmc FETCH_SIZE -d OUT -o mv -- python3 loadsg_amd/profile_matvec.py
and summarised with loadsg_amd/rocpd_summary.py.  Prints one JSON line.
"""
import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--precision", type=int, default=32, choices=(32, 64))
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--grid", default=None,
                    help="an .npz of idx / active (an adaptive grid, sparse path) instead of d=5 level 7; "
                         "coefficient 1 + |x|^2 there")
    args = ap.parse_args()

    import torch
    import loadsg_amd as L

    B = args.batch
    rng = np.random.default_rng(20261203)
    xi = rng.normal(0.0, 0.5, (B, 16))
    if args.grid:
        z = np.load(args.grid)
        grid = L.AdaptiveSparseGrid(z["idx"], z["active"])
        op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_CFG3), precision=args.precision, batch=B)
    else:
        grid = L.AdaptiveSparseGrid.regular(5, 7)
        op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[0]),
                                   precision=args.precision, batch=B)
        for b in range(B):
            op.set_coefficient(b, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[b]))
    dt = torch.float64 if args.precision == 64 else torch.float32
    dev = torch.device("cuda:0")
    x = torch.randn(B, grid.size, dtype=dt, device=dev)
    y = torch.zeros_like(x)
    op.multiplication(x, y)                 # autotunes the apply stage
    torch.cuda.synchronize()
    apply_ms, launches = [], 0
    for _ in range(args.reps):
        op.multiplication(x, y)
        t = op.last_timing()
        apply_ms.append(t["apply_ms"])
        launches = t["apply_launches"]
    torch.cuda.synchronize()
    n = args.copy_mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    c = torch.empty_like(a)
    torch.mul(a, 2.0, out=c)
    torch.cuda.synchronize()
    st = op.stats()
    print(json.dumps({"batch": B, "precision": args.precision, "reps": args.reps,
                      "apply_launches_per_matvec": launches,
                      "apply_ms_per_matvec_events": float(np.mean(apply_ms)),
                      "apply_values_per_instance": st["apply_values"],
                      "copy_bytes_read": n * 8, "copy_bytes_written": n * 8}))


if __name__ == "__main__":
    main()
