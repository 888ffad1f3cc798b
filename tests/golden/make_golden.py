"""Writes the golden fixtures under tests/golden/<case>/ by running the compiled
LoAdSG reference (oracle/_ref/ref_harness_d<D>, built by `make -C oracle ref`
from the unmodified sources in /root/reference).

Each case directory holds .npy arrays written by the harness:
  idx (n x d int32, reference slot order), active (n uint8), x (input vector),
  Ax_poisson / Ax_helm / Ax_var (MatrixVectorHomogen::multiplication),
  diag_poisson / diag_var (Preconditioning<S>::precond3),
  f_prew / f_rhs (calcPrewByNodal and the driver's f_h = M prew(f_I)),
  cg_poisson_x / cg_poisson_iters / cg_var_x / cg_var_iters (CG::solveHomogen, eps 1e-10).
  tr_u (nodal values prod sin(pi x_i) + x_1), tr_hier (CalcHierarchicalBasis), tr_prew
  (calcPrewByNodal), tr_nodal_of_x (calcNodalByPrew of x), tr_norms (H1::getValue,
  L2::getValue of tr_u).
The reference preconditioner is run with OMP_NUM_THREADS=1 so that its stale
hanging-node work-vector values (cg_method.h:356-433 resets v with an
active-only `v = 0.0`) are deterministic.

It also writes tests/golden/grids/cfg3_start4_level<L>.npz (idx, active only): the
harness's config-3 style refinement (start level 4, support boxes within 0.15 of the
origin, sum level L) followed by completeDirichletGrid, for the throughput measurement
and the larger sparse-path tests.  The generator's time grows about 3.5x per level
(level 10: about 7 minutes on one core).

It also writes tests/golden/gridgen_sets.npz (--gridgen): seeded random sets of interior
points in d = 2, 3, each added with AddPoint and completed with completeDirichletGrid by
the harness's `points` mode -- the inputs (pts<k>) and the reference's grids (idx<k>,
active<k>).  Most of these sets make addPoints2 add hanging nodes, which the refined
benchmark grids never do.

Run:  python tests/golden/make_golden.py [--grids 7,8,9,10] [--gridgen]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

# name: (d, argv after the binary)
CASES = {
    "d2_l6_cfg2": (2, ["regular", "6", "cfg2"]),
    "d3_l4_cfg3": (3, ["regular", "4", "cfg3"]),
    "d4_l3_cfg2": (4, ["regular", "3", "cfg2"]),
    "d5_l3_expsin": (5, ["regular", "3", "expsin:0.3,-0.5,0.2,0.4,-0.1,0.25,-0.35,0.15"]),
    "d3_adaptive_cfg3": (3, ["adaptive", "3", "0.15", "6", "cfg3"]),
    "d2_adaptive_cfg2": (2, ["adaptive", "3", "0.2", "8", "cfg2"]),
    "d3_adaptive_expsin": (3, ["adaptive", "2", "0.3", "5", "expsin:0.4,-0.3,0.25,-0.2,0.35,0.1"]),
    # non-separable a and c (the harness's "gen" plug-in), for the per-cell coefficient path
    "d3_l4_gen": (3, ["regular", "4", "gen"]),
    "d5_l3_gen": (5, ["regular", "3", "gen"]),
    "d3_adaptive_gen": (3, ["adaptive", "3", "0.15", "6", "gen"]),
    # refined near the origin and near (1, ..., 1): the finest depths have two separate
    # components (prewavelet line solves per component)
    "d2_adaptive2_cfg2": (2, ["adaptive2", "3", "0.2", "7", "cfg2"]),
    "d3_adaptive2_cfg3": (3, ["adaptive2", "3", "0.15", "6", "cfg3"]),
}


def make_grids(levels):
    exe = os.path.join(REF, "ref_harness_d3")
    if not os.path.exists(exe):
        sys.exit("missing %s: run `make -C oracle ref`" % exe)
    os.makedirs(os.path.join(HERE, "grids"), exist_ok=True)
    for lev in levels:
        with tempfile.TemporaryDirectory() as tmp:
            cmd = [exe, "adaptive", "4", "0.15", str(lev), "cfg3", tmp, "none"]
            print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
            idx = np.load(os.path.join(tmp, "idx.npy")).astype(np.int32)
            act = np.load(os.path.join(tmp, "active.npy")).astype(np.uint8)
        np.savez_compressed(os.path.join(HERE, "grids", "cfg3_start4_level%d.npz" % lev), idx=idx, active=act)


def make_gridgen(nsets=16, seed=20261204):
    rng = np.random.default_rng(seed)
    out = {}
    for k in range(nsets):
        d = int(rng.integers(2, 4))
        pts = []
        for _ in range(int(rng.integers(2, 12))):
            q = []
            for i in range(d):
                t = int(rng.integers(1, 7))
                q.append((1 << t) + int(rng.integers(0, 1 << (t - 1))))
            pts.append(q)
        exe = os.path.join(REF, "ref_harness_d%d" % d)
        if not os.path.exists(exe):
            sys.exit("missing %s: run `make -C oracle ref`" % exe)
        with tempfile.TemporaryDirectory() as tmp:
            f = os.path.join(tmp, "pts.txt")
            with open(f, "w") as fh:
                fh.write("\n".join(" ".join(map(str, q)) for q in pts) + "\n")
            subprocess.run([exe, "points", f, tmp, "none"], check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
            out["pts%d" % k] = np.array(pts, dtype=np.int32)
            out["idx%d" % k] = np.load(os.path.join(tmp, "idx.npy")).astype(np.int32)
            out["active%d" % k] = np.load(os.path.join(tmp, "active.npy")).astype(np.uint8)
        print("set %d: d=%d, %d points -> %d" % (k, d, len(pts), len(out["idx%d" % k])), flush=True)
    np.savez_compressed(os.path.join(HERE, "gridgen_sets.npz"), **out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grids", default="", help="comma-separated sum levels of the refined grids to write")
    ap.add_argument("--only", default="", help="comma-separated case names to (re)write; default all")
    ap.add_argument("--gridgen", action="store_true", help="write gridgen_sets.npz only")
    args = ap.parse_args()
    if args.grids:
        make_grids([int(v) for v in args.grids.split(",")])
        return
    if args.gridgen:
        make_gridgen()
        return
    only = set(filter(None, args.only.split(",")))
    meta = {}
    for name, (d, args) in CASES.items():
        meta[name] = {"d": d, "argv": args}
        if only and name not in only:
            continue
        out = os.path.join(HERE, name)
        os.makedirs(out, exist_ok=True)
        exe = os.path.join(REF, "ref_harness_d%d" % d)
        if not os.path.exists(exe):
            sys.exit("missing %s: run `make -C oracle ref`" % exe)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        cmd = [exe] + args[:-1 if args[0] == "regular" else None]
        if args[0] == "regular":
            cmd = [exe, "regular", args[1], args[2], out]
        else:
            cmd = [exe, args[0], args[1], args[2], args[3], args[4], out]
        print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, env=env)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
