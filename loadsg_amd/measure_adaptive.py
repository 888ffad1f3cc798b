"""Throughput of the sparse path on a locally adaptive grid with hanging nodes.

Workload: a d=3 grid refined towards the origin (a regular start grid whose active
points with support box closer than 0.15 to the origin are refined in every dimension
up to a sum level, then completeDirichletGrid), written by the reference's own grid
code into an .npz of `idx` (n x d int32, slot order) and `active` (n uint8).
Coefficient a(x) = 1 + |x|^2, c = 0.  Right-hand sides are the load vectors of
f = sum_j eta_j prod_i sin(m_j pi x_i), eta ~ N(0, 1), m = (1, 2), one per instance.

Measured, with the inputs resident on the device:
  * FP64 and FP32 matvecs at batch 1 and at --batch (HIP events around the launches);
  * the manufactured solution u = |x|^(2/3) prod x_i (1 - x_i) (loadsg_amd.manufactured.cfg3):
    the reference driver's right-hand side f_h = M prew(f_I) (calcPrewByNodal on the grid's
    hanging-node structure, then the mass operator), the plain FP64 PCG and the
    mixed-precision solve of it to 1e-10, the error of the mixed solution against the FP64
    one and the discretisation errors of both (largest nodal error, L2 and H1 norms of the
    nodal error, applications/H1_Norm.cpp);
  * the plain FP64 PCG to 1e-10 relative residual at batch 1;
  * the mixed-precision solve (FP32 inner PCG, FP64 refinement) to 1e-10 at batch 1
    and at --batch, with its error against the FP64 solve of the same systems.
Prints one JSON line.  Run as
    python3 loadsg_amd/measure_adaptive.py --grid tests/golden/grids/cfg3_start4_level8.npz
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _matvec_ms(op, x, y, reps):
    import torch
    op.multiplication(x, y)            # first call autotunes / warms up
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        op.multiplication(x, y)
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / reps


def run(grid_path, batch=64, reps=20, tol=1e-10, matvec_only=False, precond="block"):
    """Measure the sparse path on the grid in `grid_path`; returns a dict.  precond: the
    preconditioner of the FP32 inner solves of the mixed-precision method ("block": depth
    blocks over the active points, "diag": the diagonal); the batched mixed solve is also timed
    with the diagonal for comparison."""
    import torch
    import loadsg_amd as L

    z = np.load(grid_path)
    grid = L.AdaptiveSparseGrid(z["idx"], z["active"])
    st = grid.stats()
    N = grid.size
    dev = torch.device("cuda:0")
    coef = L.VariableCoefficient(L.COEF_CFG3)
    m = np.array([1, 2], dtype=np.int32)
    rng = np.random.default_rng(20261203)
    out = {"workload": "adaptive d=3 cfg3 " + os.path.basename(grid_path), "points": int(N),
           "hanging": int((z["active"] == 0).sum()), "depths": st["depths"],
           "multilevel_size": st["multilevel_size"], "apply_items": st["apply_items"],
           "wavefront_levels": st["wavefront_levels"], "batch": batch}

    # matvec timings
    for B in (1, batch):
        for prec, dt in ((64, torch.float64), (32, torch.float32)):
            op = L.MatrixVectorHomogen(grid, coef, precision=prec, batch=B)
            x = torch.randn(B, N, dtype=dt, device=dev)
            y = torch.zeros_like(x)
            ms = _matvec_ms(op, x, y, reps)
            out["matvec_fp%d_b%d_ms" % (prec, B)] = ms
            out["matvec_fp%d_b%d_per_s" % (prec, B)] = B * 1e3 / ms
            del op
    print("matvecs done", flush=True, file=sys.stderr)
    if matvec_only:
        return out

    pk = L.PRECOND_DEPTH_BLOCK if precond == "block" else L.PRECOND_DIAG
    out["inner_preconditioner"] = "depth-block" if precond == "block" else "diagonal"

    def ops(B):
        op32 = L.MatrixVectorHomogen(grid, coef, precision=32, batch=B)
        t0 = time.perf_counter()
        op32.set_precond(pk)
        out["precond_setup_s"] = time.perf_counter() - t0
        return L.MatrixVectorHomogen(grid, coef, precision=64, batch=B), op32

    out["manufactured"] = manufactured(grid, tol, pk)
    print("manufactured solution done", flush=True, file=sys.stderr)

    # plain FP64 PCG, batch 1
    op64, op32 = ops(1)
    f1 = torch.zeros(1, N, dtype=torch.float64, device=dev)
    op64.rhs_sines(rng.standard_normal((1, 2)), m, out=f1)
    xr = torch.zeros_like(f1)
    L.CG.solveHomogen(tol, xr.clone(), f1, op64, stop_rule=L.STOP_REL, maxit=5000)   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = L.CG.solveHomogen(tol, xr, f1, op64, stop_rule=L.STOP_REL, maxit=5000)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    out.update(fp64_cg_b1_s=t64, fp64_cg_b1_solves_per_s=1.0 / t64, fp64_cg_iterations=info["iterations"],
               fp64_cg_rel_residual=info["rel_residual"])
    # mixed precision, batch 1
    xm = torch.zeros_like(f1)
    L.CG.solveMixed(tol, xm.clone(), f1, op64, op32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    im = L.CG.solveMixed(tol, xm, f1, op64, op32)
    torch.cuda.synchronize()
    tm = time.perf_counter() - t0
    out.update(mixed_b1_s=tm, mixed_b1_solves_per_s=1.0 / tm, mixed_b1_inner_iterations=im["iterations"],
               mixed_b1_rel_residual=im["rel_residual"],
               mixed_b1_err_vs_fp64=float(torch.linalg.norm(xm - xr) / torch.linalg.norm(xr)))
    del op64, op32
    print("batch-1 solves done", flush=True, file=sys.stderr)

    # mixed precision, batched
    B = batch
    op64, op32 = ops(B)
    f = torch.zeros(B, N, dtype=torch.float64, device=dev)
    op64.rhs_sines(rng.standard_normal((B, 2)), m, out=f)
    xb = torch.zeros_like(f)
    L.CG.solveMixed(tol, xb.clone(), f, op64, op32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ib = L.CG.solveMixed(tol, xb, f, op64, op32)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    xf = torch.zeros_like(f)
    L.CG.solveHomogen(tol * 1e-3, xf, f, op64, stop_rule=L.STOP_REL, maxit=5000)
    err = torch.linalg.norm(xb - xf, dim=1) / torch.linalg.norm(xf, dim=1)
    out.update(mixed_batch_s=tb, mixed_batch_solves_per_s=B / tb,
               mixed_batch_mean_inner_iterations=float(np.mean([i["iterations"] for i in ib])),
               mixed_batch_max_rel_residual=float(max(i["rel_residual"] for i in ib)),
               mixed_batch_all_converged=bool(all(i["converged"] == 1 for i in ib)),
               mixed_batch_max_err_vs_fp64=float(err.max()))
    out["speedup_mixed_batch_vs_fp64_b1"] = out["mixed_batch_solves_per_s"] / out["fp64_cg_b1_solves_per_s"]
    if pk != L.PRECOND_DIAG:
        # the same batch with the diagonal in the inner solves
        op32.set_precond(L.PRECOND_DIAG)
        xd = torch.zeros_like(f)
        L.CG.solveMixed(tol, xd.clone(), f, op64, op32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        idg = L.CG.solveMixed(tol, xd, f, op64, op32)
        torch.cuda.synchronize()
        td = time.perf_counter() - t0
        out.update(mixed_batch_diag_solves_per_s=B / td,
                   mixed_batch_diag_mean_inner_iterations=float(np.mean([i["iterations"] for i in idg])))
    return out


def manufactured(grid, tol, precond):
    """Config 3's manufactured solution on `grid`: right-hand side, FP64 and mixed solves
    (inner preconditioner `precond`), errors (see the module docstring)."""
    import torch
    import loadsg_amd as L
    from loadsg_amd.manufactured import cfg3, coordinates
    P = cfg3()
    r = {"u": "|x|^(2/3) prod x_i (1 - x_i)", "a": "1 + |x|^2", "rhs": "f_h = M prew(f_I)"}
    op64 = L.MatrixVectorHomogen(grid, P.stencil, precision=64)
    op32 = L.MatrixVectorHomogen(grid, P.stencil, precision=32).set_precond(precond)
    t0 = time.perf_counter()
    f = P.reference_rhs(grid, op64)          # first call: transform tables and the mass operator
    r["rhs_first_call_s"] = time.perf_counter() - t0
    # device-resident right-hand side construction: prewavelets of f_I, then M
    _, act = grid.slot_order()
    dev = grid.device_order()
    fI = np.zeros(grid.size)
    fI[dev] = P.f(coordinates(grid)) * act
    mass = L.MatrixVectorHomogen(grid, L.HelmHoltz(), precision=64)
    tin = torch.tensor(fI[None, :], device="cuda")
    tp, tf = torch.zeros_like(tin), torch.zeros_like(tin)
    op64.prew_by_nodal(tin, tp)
    mass.multiplication(tp, tf)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    reps = 5
    ev[0].record()
    for _ in range(reps):
        op64.prew_by_nodal(tin, tp)
    ev[1].record()
    for _ in range(reps):
        mass.multiplication(tp, tf)
    ev[2].record()
    torch.cuda.synchronize()
    r["prew_by_nodal_ms"] = ev[0].elapsed_time(ev[1]) / reps
    r["mass_matvec_ms"] = ev[1].elapsed_time(ev[2]) / reps
    assert np.allclose(tf.cpu().numpy()[0][dev], f, rtol=0, atol=1e-12 * np.abs(f).max())
    # FP64 PCG and the mixed solve of f_h
    x = np.zeros(grid.size)
    L.CG.solveHomogen(tol, x.copy(), f, op64, stop_rule=L.STOP_REL, maxit=5000)   # warm-up
    t0 = time.perf_counter()
    info = L.CG.solveHomogen(tol, x, f, op64, stop_rule=L.STOP_REL, maxit=5000)
    r["fp64_cg_s"] = time.perf_counter() - t0
    r["fp64_cg_iterations"] = info["iterations"]
    xm = np.zeros(grid.size)
    L.CG.solveMixed(tol, xm.copy(), f, op64, op32)
    t0 = time.perf_counter()
    im = L.CG.solveMixed(tol, xm, f, op64, op32)
    r["mixed_s"] = time.perf_counter() - t0
    r["mixed_inner_iterations"] = im["iterations"]
    r["mixed_rel_residual"] = im["rel_residual"]
    r["mixed_err_vs_fp64"] = float(np.linalg.norm(xm - x) / np.linalg.norm(x))
    r["errors_fp64"] = P.errors(grid, x, op64)
    r["errors_mixed"] = P.errors(grid, xm, op64)
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", default=os.path.join(ROOT, "tests", "golden", "grids", "cfg3_start4_level8.npz"))
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--matvec-only", action="store_true", help="time the matvecs only (for profiling)")
    ap.add_argument("--precond", choices=("block", "diag"), default="block",
                    help="preconditioner of the FP32 inner solves")
    args = ap.parse_args()
    print(json.dumps(run(args.grid, args.batch, args.reps, args.tol, args.matvec_only, args.precond)), flush=True)


if __name__ == "__main__":
    main()
