"""Profiling driver for the timed solve of bench.py: d=5 regular level 7 (N = 61,183),
exp-sine coefficients (16 xi per instance, PCG64 seed 20261203 as in bench.py), 8-term sine
right-hand sides, --batch instances, FP32 PCG with the depth-block (or diagonal)
preconditioner under FP64 iterative refinement to 1e-8 from a zero start.

One untimed solve (autotuning, graph capture), then one elementwise marker kernel
(y = 2 x over 1 MiB), then --solves timed solves of fresh instances, each with the batch
set-up.  Under rocprofv3 --kernel-trace, `rocpd_summary trace DB --after elementwise` splits
the timed solves by kernel.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=500)
    ap.add_argument("--solves", type=int, default=1)
    ap.add_argument("--precond", choices=("block", "diag"), default="block")
    ap.add_argument("--inner-tol", type=float, default=1e-4, help="inner_rel_tol of the refinement")
    args = ap.parse_args()

    import torch
    import loadsg_amd as L

    B = args.batch
    rng = np.random.default_rng(20261203)
    grid = L.AdaptiveSparseGrid.regular(5, 7)
    m = rng.integers(1, 4, 8).astype(np.int32)
    st0 = L.VariableCoefficient(L.COEF_EXPSIN, xi=np.zeros(16))
    op64 = L.MatrixVectorHomogen(grid, st0, precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, st0, precision=32, batch=B)
    op32.set_precond(L.PRECOND_DEPTH_BLOCK if args.precond == "block" else L.PRECOND_DIAG)
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(f)

    def solve():
        cs = [L.VariableCoefficient(L.COEF_EXPSIN, xi=rng.normal(0.0, 0.5, 16)) for _ in range(B)]
        op64.set_coefficients(cs)
        op32.set_coefficients(cs)
        op64.rhs_sines(rng.standard_normal((B, 8)), m, out=f)
        x.zero_()
        return L.CG.solveMixed(1e-8, x, f, op64, op32, inner_rel_tol=args.inner_tol)

    solve()
    torch.cuda.synchronize()
    a = torch.ones(1 << 17, dtype=torch.float64, device="cuda")
    torch.mul(a, 2.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = []
    for _ in range(args.solves):
        infos += solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.solves
    print(json.dumps({"batch": B, "precond": args.precond, "inner_tol": args.inner_tol, "solves": args.solves,
                      "s_per_solve_batch": dt,
                      "solves_per_s": B / dt,
                      "inner_iterations_mean": float(np.mean([i["iterations"] for i in infos])),
                      "outer_sweeps_mean": float(np.mean([i["outer"] for i in infos]))}))


if __name__ == "__main__":
    main()
