"""Benchmark: solves/s to a 1e-8 relative residual on the d=5 level-7 sparse grid
(N = 61,183) with the variable coefficient a(x) = exp(sum_k xi_k sin(k pi x_{1+(k mod d)})/k)
(16 xi ~ N(0, 0.25)) and right-hand sides f = sum_j eta_j prod_i sin(m_j pi x_i)
(8 eta ~ N(0,1), m_j in {1,2,3}); synthetic seeded inputs (PCG64 seed 20261203).

One step = one batched solve of --batch (1,000) fresh instances out of 2,000 held-out
instances (seeds disjoint from the training set; the default 2 steps solve each once):
the instances' set-up (coefficient tables of the FP64 and FP32 operators, load vectors,
and the preconditioner diagonal that the new coefficients invalidate), the learned warm
start x0 = MLP(xi, eta) (training on 20,000 GPU-generated solves happens before, untimed,
and is reported), then FP32 inner PCG under FP64 iterative refinement, every instance to
||f - A x||_2 <= 1e-8 ||f||_2 (checked in FP64).  The zero-start solve of the same
instances is timed the same way and reported beside it, with the set-up cost alone.
The grid and operators are resident in HBM before the timed region.  Also reported: the
plain FP64 GPU CG (batch 1) and the batched FP64 PCG, the FP64 error of the mixed
solutions, the apply-kernel roofline (FP32 batched matvec, the inner-solver kernel), and
the reference CPU path timed on this host.

Prints one JSON line.

--workload selects one of the other configurations instead (one JSON line each, with
its own metric; the default line above is the headline):
  cfg2  d=4 regular level 8, a = 1 + 0.5 prod cos(pi x_i), c = 1 + x_1: FP64 matvec and
        FP64 PCG to 1e-10 (one instance, and matvecs at batch 64), beside the reference
        CPU matvec sampled on this host;
  cfg3  the locally adaptive d=3 grid (start level 4, refined towards the origin to sum
        level 11, tests/golden/grids), a = 1 + |x|^2: the manufactured solution
        u = |x|^(2/3) prod x_i (1 - x_i) with the reference driver's right-hand side
        f_h = M prew(f_I), solved in FP64 and mixed precision to 1e-10 with its
        discretisation errors, and batched FP32 PCG under FP64 refinement to 1e-10 on
        synthetic right-hand sides against the FP64 solve (loadsg_amd/measure_adaptive.py);
  cfg4  64 instances with d uniform in {3..6} and level in {5..8} (seeded), exp-sine
        coefficients (16 xi), 8-term sine right-hand sides, solved together to 1e-8
        (lsg_ir_solve_many: one host thread advancing every grid shape's solve, the
        shapes side by side on the GPU);
  transform  nodal values -> prewavelets (calcPrewByNodal), the inverse and the H1 norm on
        the d=5 level-7 grid for --batch instances, beside the reference's calcPrewByNodal.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12       # MI355X HBM3E bytes/s
SEED = 20261203
D, LEVEL = 5, 7
NXI, NETA = 16, 8


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def inputs(batch, rng):
    xi = rng.normal(0.0, 0.5, (batch, NXI))          # variance 0.25
    eta = rng.standard_normal((batch, NETA))
    m = rng.integers(1, 4, NETA).astype(np.int32)
    return xi, eta, m


def physical_cores():
    """Physical cores available to this process (lscpu: cores per socket x sockets, capped
    by the affinity mask)."""
    avail = len(os.sched_getaffinity(0))
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=30).stdout
        cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        if cores:
            return max(1, min(avail, len(cores)))
    except (OSError, subprocess.SubprocessError):
        pass
    return avail


def baseline_cores():
    """Threads for the reference CPU path: the physical cores this process may use, at most
    16 (the CPU share of one GPU on the measurement host)."""
    return min(physical_cores(), 16)


def cpu_model():
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return None


def apply_traffic_from_profile(batch):
    """HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the apply stage of one
    FP32 matvec per batch instance, as recorded from rocprofv3 --pmc runs at this batch."""
    p = os.path.join(ROOT, "profiles", "apply_sep_kernel_fp32_d5l7_b%d_pmc.md" % batch)
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_matvec_per_instance:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(iters_fp64, budget_s):
    """Reference CPU path (oracle/_ref/ref_harness_d5, built from the unmodified
    LoAdSG sources) timed on this host: calcNodal plus a sample of the 32 cases of one
    d=5 level-7 variable-coefficient matvec, ratio-scaled to the full matvec by the
    per-case cell counts; solves/s = 1 / ((iters+1) * t_matvec), preconditioner set-up
    excluded (favours the CPU)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_d5")
    if not os.path.exists(exe):
        return None
    cores = baseline_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close")
    xi = ",".join("%.6f" % v for v in np.random.default_rng(SEED).normal(0.0, 0.5, NXI))
    try:
        r = subprocess.run([exe, "samplemv", str(LEVEL), "expsin:" + xi, str(budget_s)], env=env,
                           capture_output=True, text=True, timeout=budget_s * 6 + 120)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        return None
    res = json.loads(r.stdout.strip().splitlines()[-1])
    t_mv = res["matvec_s_estimate"]
    return {"value": 1.0 / ((iters_fp64 + 1) * t_mv), "unit": "solves/s", "cores": cores, "kind": "reference",
            "cpu_model": cpu_model(),
            "sample": "LoAdSG MatrixVectorHomogen (unmodified sources, %d OpenMP threads): calcNodal + cases %s of the "
                      "32 of one d=5 L7 var-coeff matvec (%.1f s timed), scaled by per-case cell counts to %.1f s per "
                      "matvec; x (FP64 GPU CG iterations + 1 = %d) matvecs per solve, preconditioner set-up "
                      "excluded" % (cores, res["cases"], res["timed_s"], t_mv, iters_fp64 + 1),
            "matvec_s": t_mv}


def ref_sample_matvec(d, level, coef, budget_s):
    """Reference CPU matvec (oracle/_ref/ref_harness_d<d> samplemv, the unmodified
    LoAdSG sources) timed on this host: seconds per matvec, or None."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_d%d" % d)
    if not os.path.exists(exe):
        return None, 0
    cores = baseline_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close")
    try:
        r = subprocess.run([exe, "samplemv", str(level), coef, str(budget_s)], env=env, capture_output=True,
                           text=True, timeout=budget_s * 6 + 120)
    except subprocess.TimeoutExpired:
        return None, cores
    if r.returncode != 0:
        return None, cores
    return json.loads(r.stdout.strip().splitlines()[-1]), cores


def workload_cfg1(args):
    """Config 1: 2D Poisson on the regular level-6 grid (N = 769), u = prod sin(pi x_i),
    unpreconditioned CG from x0 = 0 to ||f - A x|| <= 1e-10 ||f|| on b_p = int f psi_p;
    matvecs/s, solves/s and the discretisation errors, beside the reference's whole PCG
    solve (oracle/_ref/ref_harness_d2 timesolve) timed on this host."""
    import torch
    import loadsg_amd as L
    from loadsg_amd import manufactured as MS
    d, level = 2, 6
    grid = L.AdaptiveSparseGrid.regular(d, level)
    n = grid.size
    prob = MS.cfg1(d)
    op = L.MatrixVectorHomogen(grid, L.Poisson())
    b = prob.load(op)
    perm = grid.device_order()
    dev = torch.device("cuda:0")
    xs = torch.tensor((1.0 / (1.0 + (np.arange(n) % 7)))[np.argsort(perm)][None, :], device=dev)
    ys = torch.zeros_like(xs)
    op.multiplication(xs, ys)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(100):
        op.multiplication(xs, ys)
    torch.cuda.synchronize()
    t_mv = (time.perf_counter() - t0) / 100
    fd = np.zeros(n)
    fd[perm] = b
    f = torch.tensor(fd[None, :], device=dev)
    x = torch.zeros_like(f)
    rule = L.STOP_REL | L.CG_NO_PRECOND
    L.CG.solveHomogen(1e-10, x, f, op, stop_rule=rule, maxit=5000)
    reps = 20
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        x.zero_()
        info = L.CG.solveHomogen(1e-10, x, f, op, stop_rule=rule, maxit=5000)
    torch.cuda.synchronize()
    t_cg = (time.perf_counter() - t0) / reps
    errs = prob.errors(grid, x[0].cpu().numpy()[perm], b, op)
    log("cfg1: matvec %.1f us; plain CG %d iterations %.2f ms; L2 error %.3e, energy error %.3e"
        % (t_mv * 1e6, info["iterations"], t_cg * 1e3, errs["l2"], errs["energy"]))
    cpu = None
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_d2")
    if not args.no_cpu_baseline and os.path.exists(exe):
        cores = baseline_cores()
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close")
        r = subprocess.run([exe, "timesolve", str(level), "const:1:0", "1e-10"], env=env, capture_output=True,
                           text=True, timeout=600)
        if r.returncode == 0:
            rr = json.loads(r.stdout.strip().splitlines()[-1])
            cpu = {"value": 1.0 / rr["cg_s"], "unit": "solves/s", "cores": cores, "kind": "reference",
                   "cpu_model": cpu_model(),
                   "sample": "LoAdSG CG::solveHomogen (diagonal PCG built inside, reference stop rule eps 1e-10, "
                             "%d iterations) on the same grid, whole solve timed; matvec %.2f ms"
                             % (rr["iterations"], 1e3 * rr["matvec_s"]),
                   "matvec_s": rr["matvec_s"], "cg_s": rr["cg_s"], "setup_s": rr["setup_s"]}
    out = {"metric": "unpreconditioned CG solves/s to 1e-10 rel. residual (d=2 L6 Poisson)", "value": 1.0 / t_cg,
           "unit": "solves/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "d=2 regular sparse grid level 6 (N=%d), Poisson, u = prod sin(pi x_i)" % n},
           "cpu_baseline": cpu,
           "extra": {"matvec_us": t_mv * 1e6, "cg_iterations": info["iterations"],
                     "cg_rel_residual": info["rel_residual"], "cg_ms": t_cg * 1e3, "l2_error": errs["l2"],
                     "energy_error": errs["energy"], "nodal_max_error": errs["nodal_max"]}}
    print(json.dumps(out))


def workload_cfg2(args):
    import torch
    import loadsg_amd as L
    d, level = 4, 8
    grid = L.AdaptiveSparseGrid.regular(d, level)
    n = grid.size
    coef = L.VariableCoefficient(L.COEF_CFG2)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED)
    res = {}
    for B in (1, 64):
        op = L.MatrixVectorHomogen(grid, coef, precision=64, batch=B)
        xv = torch.randn(B, n, dtype=torch.float64, device=dev)
        yv = torch.zeros_like(xv)
        op.multiplication(xv, yv)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(10):
            op.multiplication(xv, yv)
        ev1.record()
        torch.cuda.synchronize()
        res[B] = ev0.elapsed_time(ev1) / 10
    from loadsg_amd import manufactured as MS
    prob = MS.cfg2(d)
    op = L.MatrixVectorHomogen(grid, coef, precision=64, batch=1)
    # load vector of the manufactured u = prod x_i (1 - x_i) exp(x_1): b_p = int f psi_p
    b_host = prob.load(op)
    perm = grid.device_order()
    fd = np.zeros(n)
    fd[perm] = b_host
    f = torch.tensor(fd[None, :], device=dev)
    x = torch.zeros_like(f)
    L.CG.solveHomogen(1e-10, x, f, op, stop_rule=L.STOP_REL, maxit=5000)
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info = L.CG.solveHomogen(1e-10, x, f, op, stop_rule=L.STOP_REL, maxit=5000)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    errs = prob.errors(grid, x[0].cpu().numpy()[perm], b_host, op)
    log("cfg2: matvec %.3f ms (b1), %.3f ms (b64); CG %d iterations %.3f s; L2 error %.3e, nodal max %.3e"
        % (res[1], res[64], info["iterations"], t_cg, errs["l2"], errs["nodal_max"]))
    cpu = None
    if not args.no_cpu_baseline:
        r, cores = ref_sample_matvec(d, level, "cfg2", args.cpu_budget)
        if r is not None:
            t_mv = r["matvec_s_estimate"]
            cpu = {"value": 1.0 / t_mv, "unit": "matvecs/s", "cores": cores, "kind": "reference",
                   "cpu_model": cpu_model(),
                   "sample": "LoAdSG MatrixVectorHomogen, d=4 level 8, CFG2 coefficient (%d OpenMP threads): "
                             "calcNodal + cases %s (%.0f%% of the cell work), scaled to the full matvec"
                             % (cores, r["cases"], 100 * r["case_weight_fraction"]),
                   "matvec_s": t_mv, "cg_s_estimate": (info["iterations"] + 1) * t_mv}
    out = {"metric": "FP64 matvecs/s and PCG to 1e-10 (d=4 L8, variable coefficient)", "value": 1e3 / res[1],
           "unit": "matvecs/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "d=4 regular sparse grid level 8 (N=%d), a=1+0.5 prod cos(pi x_i), c=1+x_1" % n},
           "cpu_baseline": cpu,
           "extra": {"matvec_b1_ms": res[1], "matvec_b64_ms": res[64], "matvecs_per_s_b64": 64e3 / res[64],
                     "rhs": "b_p = int f psi_p, f = -div(a grad u) + c u for u = prod x_i (1 - x_i) exp(x_1)",
                     "l2_error": errs["l2"], "nodal_max_error": errs["nodal_max"],
                     "cg_iterations": info["iterations"], "cg_rel_residual": info["rel_residual"],
                     "cg_s": t_cg, "cg_solves_per_s": 1.0 / t_cg,
                     "speedup_matvec_vs_cpu": None if cpu is None else cpu["matvec_s"] * 1e3 / res[1],
                     "speedup_cg_vs_cpu": None if cpu is None else cpu["cg_s_estimate"] / t_cg}}
    print(json.dumps(out))


def workload_cfg3(args):
    from loadsg_amd.measure_adaptive import run
    r = run(os.path.join(ROOT, "tests", "golden", "grids", "cfg3_start4_level11.npz"), batch=128, reps=10, tol=1e-10,
            precond=args.precond)
    out = {"metric": "solves/s to 1e-10 rel. residual, locally adaptive d=3 grid (FP32 PCG under FP64 IR)",
           "value": r["mixed_batch_solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
           "dtype": "f32 inner / f64 refinement",
           "data": "reference-generated grid; manufactured and synthetic right-hand sides",
           "config": {"workload": r["workload"], "points": r["points"], "hanging": r["hanging"], "batch": 128,
                      "inner_preconditioner": r["inner_preconditioner"]},
           "extra": r}
    print(json.dumps(out))


def workload_cfg4(args):
    import torch
    import loadsg_amd as L
    rng = np.random.default_rng(SEED + 4)
    B = 64
    ds = rng.integers(3, 7, B)
    ls = rng.integers(5, 9, B)
    xi = rng.normal(0.0, 0.5, (B, NXI))
    eta = rng.standard_normal((B, NETA))
    m = rng.integers(1, 4, NETA).astype(np.int32)
    groups = {}
    for b in range(B):
        groups.setdefault((int(ds[b]), int(ls[b])), []).append(b)
    probs = []
    for (d, level), idx in sorted(groups.items()):
        grid = L.AdaptiveSparseGrid.regular(d, level)
        nb = len(idx)
        o64 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[idx[0]]), precision=64, batch=nb)
        o32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[idx[0]]), precision=32, batch=nb)
        o32.set_precond(L.PRECOND_DEPTH_BLOCK if args.precond == "block" else L.PRECOND_DIAG)
        cs = [L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[b]) for b in idx]
        o64.set_coefficients(cs)
        o32.set_coefficients(cs)
        f = torch.zeros(nb, grid.size, dtype=torch.float64, device="cuda")
        o64.rhs_sines(eta[idx], m, out=f)
        probs.append((torch.zeros_like(f), f, o64, o32))
    log("cfg4: %d groups set up" % len(probs))
    L.solve_many(probs, 1e-8)                     # warm-up (autotuning)
    for x, _, _, _ in probs:
        x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = L.solve_many(probs, 1e-8)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    flat = [i for g in infos for i in g]
    assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-8 for i in flat)
    out = {"metric": "solves/s to 1e-8 rel. residual, 64 uneven instances (d 3..6, level 5..8)", "value": B / dt,
           "unit": "solves/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32 inner / f64 refinement",
           "data": "synthetic (seeded PCG64 %d)" % (SEED + 4),
           "config": {"workload": "64 instances, %d (d, level) groups, %s preconditioner in the inner FP32 solves"
                                  % (len(probs), "depth-block" if args.precond == "block" else "diagonal"),
                      "groups": {"%d,%d" % k: len(v) for k, v in sorted(groups.items())}},
           "extra": {"s": dt, "inner_iterations_mean": float(np.mean([i["iterations"] for i in flat])),
                     "inner_iterations_max": int(max(i["iterations"] for i in flat)),
                     "points_total": int(sum(p[1].shape[0] * p[1].shape[1] for p in probs))}}
    print(json.dumps(out))


def workload_transform(args):
    """Basis transformations on the d=5 level-7 grid: calcPrewByNodal (nodal values ->
    prewavelets), calcNodalByPrew and the H1 energy norm (H1::getValue) of --batch
    instances, inputs resident in HBM; beside the reference's own calcPrewByNodal on this
    host (the reference's cost grows about 6.7x per level, so it is timed at level 6 and
    the level-7 figure is an extrapolation by the measured level-5 -> 6 growth)."""
    import torch
    import loadsg_amd as L
    grid = L.AdaptiveSparseGrid.regular(D, LEVEL)
    n = grid.size
    nml = grid.stats()["multilevel_size"]
    B = args.batch
    dev = torch.device("cuda:0")
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64, batch=B)
    g = torch.Generator(device=dev)
    g.manual_seed(SEED)
    u = torch.randn(B, n, dtype=torch.float64, device=dev, generator=g)
    p = torch.zeros_like(u)
    w = torch.zeros_like(u)
    op.prew_by_nodal(u, p)
    op.nodal_by_prew(p, w)
    op.energy_norm(u)          # first call autotunes the matvec launch shapes (untimed)
    torch.cuda.synchronize()
    rt_err = float((w - u).abs().max() / u.abs().max())

    def timed(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    reps = max(args.steps, 3)
    t_pn = timed(lambda: op.prew_by_nodal(u, p), reps)
    t_np = timed(lambda: op.nodal_by_prew(p, w), reps)
    t_h = timed(lambda: op.hierarchize(u, w), reps)
    t_norm = timed(lambda: op.energy_norm(u), reps)
    # streaming model of calcPrewByNodal per instance: box gathers of the hierarchization
    # and of every level (2 N_ML), the scratch boxes written and read back (2 N_ML), the
    # row updates of the running representation (N_ML), the surpluses and prewavelets
    # written (2 N)
    alg = 8.0 * (5 * nml + 2 * n) * B
    log("transform: prew_by_nodal %.3f ms, nodal_by_prew %.3f ms, hierarchize %.3f ms, H1 norm %.3f ms (batch %d)"
        % (t_pn * 1e3, t_np * 1e3, t_h * 1e3, t_norm * 1e3, B))
    cpu = None
    if not args.no_cpu_baseline:
        exe = os.path.join(ROOT, "oracle", "_ref", "ref_harness_d5")
        cores = baseline_cores()
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close")
        res = {}
        for lev in (5, 6):
            if not os.path.exists(exe):
                break
            log("transform: reference calcPrewByNodal at level %d" % lev)
            r = subprocess.run([exe, "timetr", str(lev)], env=env, capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                break
            res[lev] = json.loads(r.stdout.strip().splitlines()[-1])
        if 6 in res and 5 in res:
            growth = res[6]["prew_by_nodal_s"] / res[5]["prew_by_nodal_s"]
            t7 = res[6]["prew_by_nodal_s"] * growth
            cpu = {"value": 1.0 / t7, "unit": "transforms/s", "cores": cores, "kind": "reference",
                   "cpu_model": cpu_model(),
                   "sample": "LoAdSG calcPrewByNodal (unmodified sources, %d OpenMP threads) timed at d=5 level 6 "
                             "(%.1f s, N=%d) and level 5 (%.2f s); level 7 extrapolated by the %.2fx growth: %.0f s"
                             % (cores, res[6]["prew_by_nodal_s"], res[6]["n"], res[5]["prew_by_nodal_s"], growth, t7),
                   "level6_s": res[6]["prew_by_nodal_s"], "level7_s_extrapolated": t7}
    out = {"metric": "prewavelet transforms/s (calcPrewByNodal, d=5 L7)", "value": B / t_pn, "unit": "transforms/s",
           "n_gpus": 1, "steps": reps, "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded normal)",
           "config": {"workload": "d=5 regular sparse grid level 7 (N=%d), nodal values -> prewavelet coefficients, "
                                  "batch %d" % (n, B), "batch": B, "N": n},
           "roofline": {"bound": "hbm", "achieved": alg / t_pn / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": alg / t_pn / HBM_PEAK, "traffic": None,
                        "kernel": "tr_box_kernel + tr_gather_kernel: the whole calcPrewByNodal launch sequence "
                                  "(timed as a sequence; per-kernel split in profiles/)",
                        "algorithmic_bytes": alg, "time_ms": t_pn * 1e3},
           "cpu_baseline": cpu,
           "extra": {"prew_by_nodal_ms": t_pn * 1e3, "nodal_by_prew_ms": t_np * 1e3, "hierarchize_ms": t_h * 1e3,
                     "h1_norm_ms": t_norm * 1e3, "round_trip_rel_err": rt_err, "n_ml": nml,
                     "speedup_vs_reference": None if cpu is None else cpu["level7_s_extrapolated"] * B / t_pn}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2,
                    help="timed steps; step k solves held-out batch k mod (eval / batch)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1000)
    ap.add_argument("--eval", type=int, default=2000, help="held-out instances (a multiple of --batch)")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--warm-xi", type=int, default=2500,
                    help="coefficient samples for the warm-start training set (x 8 basis solves = 20,000 "
                         "GPU-generated solves); 0 = zero start")
    ap.add_argument("--warm-steps", type=int, default=96000)
    ap.add_argument("--workload", choices=("d5l7", "cfg1", "cfg2", "cfg3", "cfg4", "transform"), default="d5l7")
    ap.add_argument("--precond", choices=("block", "diag"), default="block",
                    help="preconditioner of the inner FP32 solves: depth-block Jacobi or the reference's diagonal")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (batched solves shard by instances across GPUs)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload != "d5l7":
        {"cfg1": workload_cfg1, "cfg2": workload_cfg2, "cfg3": workload_cfg3, "cfg4": workload_cfg4,
         "transform": workload_transform}[args.workload](args)
        return

    import torch
    import loadsg_amd as L

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED)
    B = args.batch
    if args.eval % B:
        sys.exit("bench.py: --eval must be a multiple of --batch")
    # held-out evaluation instances (seeds disjoint from the training draw below), plus one
    # batch for the untimed warm-up; the RHS modes m are shared by all instances
    xi_all, eta_all, m = inputs(args.eval + B, rng)
    nb = args.eval // B

    def batch_of(k):       # k = -1: the warm-up batch
        sl = slice(args.eval, args.eval + B) if k < 0 else slice(k * B, (k + 1) * B)
        return xi_all[sl], eta_all[sl]

    xi, eta = batch_of(0)
    grid = L.AdaptiveSparseGrid.regular(D, LEVEL)
    n = grid.size
    st0 = L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[0])
    op64 = L.MatrixVectorHomogen(grid, st0, precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, st0, precision=32, batch=B)
    inner_pre = L.PRECOND_DEPTH_BLOCK if args.precond == "block" else L.PRECOND_DIAG
    op32.set_precond(inner_pre)
    f = torch.zeros(B, n, dtype=torch.float64, device=dev)
    x = torch.zeros(B, n, dtype=torch.float64, device=dev)

    def set_up(k):
        """Per-batch set-up: coefficient tables of both operators for every instance
        (lsg_op_set_coef) and the load vectors; the preconditioner diagonal follows in
        the solve (it is invalidated by the new coefficients)."""
        xk, ek = batch_of(k)
        cs = [L.VariableCoefficient(L.COEF_EXPSIN, xi=xk[b]) for b in range(B)]
        op64.set_coefficients(cs)
        op32.set_coefficients(cs)
        op64.rhs_sines(ek, m, out=f)
        return xk, ek

    set_up(0)
    stats = op32.stats()        # the plan the operators run on

    log("grid N=%d, %d instances per batch, %d held-out instances" % (n, B, args.eval))
    # ---- plain FP64 GPU CG, one instance (the baseline the mixed/batched path is compared to)
    op1 = L.MatrixVectorHomogen(grid, st0, precision=64, batch=1)
    f1 = f[0:1].clone()
    x1 = torch.zeros_like(f1)
    L.CG.solveHomogen(1e-8, x1, f1, op1, stop_rule=L.STOP_REL)     # warm-up / autotune
    x1.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    info1 = L.CG.solveHomogen(1e-8, x1, f1, op1, stop_rule=L.STOP_REL)
    torch.cuda.synchronize()
    t_fp64 = time.perf_counter() - t0
    fp64_solves = 1.0 / t_fp64

    log("plain FP64 CG: %d iterations, %.3f s" % (info1["iterations"], t_fp64))
    # ---- reference FP64 solutions (tight) for the error statement: first instances of batch 0
    nchk = min(B, 4)
    opc = L.MatrixVectorHomogen(grid, st0, precision=64, batch=nchk)
    for b in range(nchk):
        opc.set_coefficient(b, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[b]))
    xref = torch.zeros(nchk, n, dtype=torch.float64, device=dev)
    L.CG.solveHomogen(1e-13, xref, f[:nchk].clone(), opc, stop_rule=L.STOP_REL, maxit=2000)

    log("tight FP64 reference solves done")

    # ---- learned warm start (untimed training on GPU-generated solves)
    ws = None
    if args.warm_xi > 0:
        from loadsg_amd.warmstart import WarmStart
        ws = WarmStart(grid, m, n_xi=NXI, progress=True)
        xi_train = np.random.default_rng(SEED + 1).normal(0.0, 0.5, (args.warm_xi, NXI))
        U = ws.basis_solutions(xi_train, batch=512)
        ws.fit(xi_train, U, steps=args.warm_steps)
        del U
        log("warm start trained: %s" % json.dumps(ws.stats))

    def solve_step(k, warm):
        """One step: set-up of batch k, start vector, mixed-precision solve."""
        xk, ek = set_up(k)
        if warm:
            ws.predict(xk, ek, out=x)         # inference is part of the timed solve
        else:
            x.zero_()
        return L.CG.solveMixed(1e-8, x, f, op64, op32, inner_rel_tol=args.inner_tol)

    def timed_pass(warm, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = []
        for s_ in range(steps):
            r = solve_step(s_ % nb, warm)
            infos += r if B > 1 else [r]
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, infos

    # ---- per-batch set-up cost alone (inside every timed step below)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    set_up(0)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0

    # ---- batched FP64 PCG from zero (batching alone, without mixed precision or warm start);
    # one matvec first so that the FP64 operator's first-use autotuning is not timed
    set_up(0)
    op64.multiplication(x, torch.zeros_like(x))
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos64 = L.CG.solveHomogen(1e-8, x, f, op64, stop_rule=L.STOP_REL)
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t0
    if B == 1:
        infos64 = [infos64]
    log("batched FP64 PCG: %.3f s" % dt64)

    # ---- zero start over the held-out instances (same steps, reported beside the value), and
    # one zero-start step with the reference's diagonal in the inner solves
    solve_step(-1, False)
    dt0, infos0 = timed_pass(False, args.steps)
    log("zero-start mixed-precision solves: %.3f s per step" % dt0)
    dtd, infosd = None, []
    if inner_pre != L.PRECOND_DIAG:
        op32.set_precond(L.PRECOND_DIAG)
        solve_step(-1, False)
        dtd, infosd = timed_pass(False, 1)
        op32.set_precond(inner_pre)
        log("zero-start mixed-precision solves, diagonal inner preconditioner: %.3f s per step" % dtd)

    # ---- timed: batched mixed-precision solves of fresh instances (warm-started when trained)
    for _ in range(args.warmup):
        solve_step(-1, ws is not None)
    dt, infos = timed_pass(ws is not None, args.steps)
    log("timed mixed-precision solves: %.3f s per step" % dt)
    assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-8 for i in infos), infos
    # error of the mixed solutions against the tight FP64 solves (batch 0)
    solve_step(0, ws is not None)
    err = float(torch.max(torch.linalg.norm(x[:nchk] - xref, dim=1) / torch.linalg.norm(xref, dim=1)))

    # ---- roofline of the dominant kernel: the apply stage of the FP32 batched matvec
    # (the inner PCG iterations are FP32 matvecs); timed with HIP events on the
    # operator's stream around the apply launches.  FP64 figures are reported beside it.
    def apply_timing(op, xv, reps=5):
        yv = torch.zeros_like(xv)
        op.multiplication(xv, yv)
        a_ms = t_ms = 0.0
        nl = 0
        for _ in range(reps):
            op.multiplication(xv, yv)
            tm = op.last_timing()
            a_ms += tm["apply_ms"]
            t_ms += tm["total_ms"]
            nl = tm["apply_launches"]
        return a_ms / reps, t_ms / reps, nl

    apply_ms64, total_ms64, _ = apply_timing(op64, x)
    apply_ms, total_ms, launches = apply_timing(op32, x.float())
    # algorithmic bytes: the streaming-model values (each value each stage of the
    # reference data flow reads or writes, counted once) of the stages the apply
    # kernel performs, x 4 bytes x batch
    apply_bytes = stats["stream_apply_values"] * 4.0 * B
    achieved = apply_bytes / (apply_ms * 1e-3)
    own_bw = stats["apply_values"] * 4.0 * B / (apply_ms * 1e-3)
    stream_bw = stats["stream_values"] * 4.0 * B / (total_ms * 1e-3)
    stream_bw64 = stats["stream_values"] * 8.0 * B / (total_ms64 * 1e-3)
    # the plain FP64 CG's own matvec (batch 1), streaming-model bytes over its time
    _, total_ms1, _ = apply_timing(op1, x1)
    stream_bw1 = stats["stream_values"] * 8.0 / (total_ms1 * 1e-3)
    achieved64 = stats["stream_apply_values"] * 8.0 * B / (apply_ms64 * 1e-3)
    tr = apply_traffic_from_profile(B)

    log("apply stage: FP32 %.3f ms, FP64 %.3f ms per batched matvec" % (apply_ms, apply_ms64))
    cpu = None if args.no_cpu_baseline else cpu_baseline(info1["iterations"], args.cpu_budget)

    out = {
        "metric": "solves/s to 1e-8 rel. residual (d=5 L7) + matvec HBM GB/s vs roofline",
        "value": B / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 refinement",
        "data": "synthetic (seeded PCG64 %d)" % SEED,
        "config": {"workload": "d=5 regular sparse grid level 7 (N=%d), var. coefficient exp-sine (16 xi), "
                               "8-term sine RHS, batch %d, FP32 PCG (%s preconditioner) under FP64 IR to 1e-8%s" %
                               (n, B, "depth-block" if inner_pre != L.PRECOND_DIAG else "diagonal",
                                ", MLP warm start" if ws is not None else ""),
                   "batch": B, "N": n, "inner_rel_tol": args.inner_tol},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": None if tr is None else tr * B,
                     "kernel": "apply_sep_kernel<float> (separable-coefficient apply stage): the %d apply "
                               "launches of one FP32 matvec of batch %d (one 'launch' = one matvec's apply stage)"
                               % (launches, B),
                     "algorithmic_bytes": apply_bytes, "time_ms": apply_ms,
                     "kernel_io_GBps": own_bw / 1e9,
                     "traffic_GBps": None if tr is None else tr * B / (apply_ms * 1e-3) / 1e9},
        "cpu_baseline": cpu,
        "extra": {
            "fp64_plain_cg_solves_per_s": fp64_solves,
            "fp64_plain_cg_iterations": info1["iterations"],
            "speedup_vs_fp64_plain": (B / dt) / fp64_solves,
            "fp64_batched_cg_solves_per_s": B / dt64,
            "speedup_vs_fp64_batched": (B / dt) / (B / dt64),
            "setup_s_per_batch": setup_s,
            "setup_share_of_step": setup_s / dt,
            "held_out_instances": args.eval,
            "instances_timed": args.steps * B,
            "fp64_batched_cg_iterations_mean": float(np.mean([i["iterations"] for i in infos64])),
            "mixed_inner_iterations_mean": float(np.mean([i["iterations"] for i in infos])),
            "mixed_outer_sweeps_mean": float(np.mean([i["outer"] for i in infos])),
            "zero_start_solves_per_s": B / dt0,
            "warm_start_gain": dt0 / dt,
            "zero_start_inner_iterations_mean": float(np.mean([i["iterations"] for i in infos0])),
            "zero_start_outer_sweeps_mean": float(np.mean([i["outer"] for i in infos0])),
            "inner_preconditioner": "depth-block Jacobi (constant-coefficient Kronecker blocks scaled by the "
                                    "diagonal)" if inner_pre != L.PRECOND_DIAG else "diagonal (precond3)",
            "zero_start_diag_inner_solves_per_s": None if dtd is None else B / dtd,
            "zero_start_diag_inner_iterations_mean": (float(np.mean([i["iterations"] for i in infosd]))
                                                      if infosd else None),
            "warm_start": None if ws is None else dict(ws.stats, xi_samples=args.warm_xi,
                                                      mlp="24-512-512-512-%d GELU, BF16 MFMA" % n),
            "max_rel_err_vs_fp64_solve": err,
            "matvec_fp32_batched_ms": total_ms,
            "matvec_fp64_batched_ms": total_ms64,
            "matvec_fp64_batched_stream_model_frac": stream_bw64 / HBM_PEAK,
            "matvec_fp64_b1_ms": total_ms1,
            "matvec_fp64_b1_stream_model_frac": stream_bw1 / HBM_PEAK,
            "apply_fp64_ms": apply_ms64,
            "apply_fp64_GBps": achieved64 / 1e9,
            "apply_fp64_frac": achieved64 / HBM_PEAK,
            "matvec_stream_model_GBps": stream_bw / 1e9,
            "matvec_stream_model_frac": stream_bw / HBM_PEAK,
            "stream_values_per_matvec": stats["stream_values"],
            "moved_values_per_matvec": stats["moved_values"],
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
