"""Profiling driver for the warm-start network's training step and inference.

Network: 24 -> 512 x 3 (GELU) -> 61,183 (the d=5 level-7 grid size), BF16 MFMA GEMMs with
FP32 accumulation.  Runs --steps Adam steps on a batch of --batch random samples (the
training batch of bench.py) and --infer forward passes of 512 samples (the batch of the
timed solves), timed with a device synchronisation around each loop.  Prints one JSON line.

Meant to run under rocprofv3, e.g.
    rocprofv3 --kernel-trace -d OUT -o mlp -- python3 loadsg_amd/profile_mlp.py
    python3 -m loadsg_amd.rocpd_summary sequence OUT/mlp_results.db --kernel to_bf16 --tail 150
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--infer", type=int, default=20)
    a = ap.parse_args()
    import torch
    import loadsg_amd as L
    n_out = 61183
    m = L.MLP(24, 512, 3, n_out, seed=1)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    x = torch.randn(a.batch, 24, device="cuda", generator=g)
    t = torch.randn(a.batch, n_out, device="cuda", generator=g)
    for _ in range(5):
        m.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        m.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    xb = torch.randn(512, 24, device="cuda", generator=g)
    for _ in range(3):
        m.forward(xb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.infer):
        m.forward(xb)
    torch.cuda.synchronize()
    di = (time.perf_counter() - t0) / a.infer
    flop_step = 3 * 2.0 * a.batch * (24 * 512 + 2 * 512 * 512 + 512 * n_out)
    print(json.dumps({"train_step_ms": dt * 1e3, "train_tflops": flop_step / dt / 1e12,
                      "forward_512_ms": di * 1e3,
                      "forward_tflops": 2.0 * 512 * (24 * 512 + 2 * 512 * 512 + 512 * n_out) / di / 1e12}))


if __name__ == "__main__":
    main()
