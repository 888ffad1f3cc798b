"""Learned warm start for batched solves of the exp-sine coefficient family.

Problem family: -div(a grad u) = f on [0,1]^d, zero Dirichlet boundary, with
    a(x) = exp(sum_k xi_k sin(k pi x_{k mod d}) / k)          (LSG_COEF_EXPSIN)
    f(x) = sum_j eta_j prod_i sin(m_j pi x_i)                    (lsg_rhs_sines)
for fixed modes m_j.  Because u is linear in eta, u(xi, eta) = sum_j eta_j U_j(xi),
where U_j(xi) solves the problem with right-hand side mode j alone.  Training data are
therefore generated as the n_eta "basis" solves per sampled xi (batched on the GPU with
the mixed-precision solver), and every training sample (xi, eta) is synthesised from
them with fresh eta.  The network (loadsg_amd.MLP: n_xi + n_eta -> width x hidden ->
N) maps (xi / xi_scale, eta) to u / u_scale in the grid's device order.

The network predicts the solution scaled by the diagonal D of the constant-coefficient
operator (y = D u / u_scale): since the operator is close to D in the prewavelet
basis, the squared error in y approximates ||A (x0 - u)||^2, the residual the solver
starts from, rather than the plain coefficient error (which weights the fine levels,
where A is large, far too little).

The prediction is used as the start vector of CG.solveMixed, which replaces any start
worse than zero (||f - A x0|| > ||f||) by zero, so a poor prediction costs at most one
FP64 residual evaluation.
"""
import sys
import time

import numpy as np

from . import (COEF_EXPSIN, CG, MLP, PRECOND_DEPTH_BLOCK, MatrixVectorHomogen, Poisson, VariableCoefficient)


class WarmStart:
    def __init__(self, grid, modes, n_xi=16, xi_scale=0.5, width=512, hidden=3, seed=0, progress=False):
        self.progress = progress      # a line on stderr about every 20 s of data generation or training
        self.grid = grid
        self.modes = np.asarray(modes, dtype=np.int32)
        self.n_xi = n_xi
        self.n_eta = len(self.modes)
        self.xi_scale = float(xi_scale)
        self.net = MLP(n_xi + self.n_eta, width, hidden, grid.size, seed=seed)
        self.u_scale = 1.0
        self.stats = {}
        import torch
        op = MatrixVectorHomogen(grid, Poisson(), precision=64, batch=1)
        d = torch.zeros(1, grid.size, dtype=torch.float64, device="cuda")
        op.diagonal(out=d)
        self.D = d[0].float()

    # -------------------------------------------------------------- data
    def basis_solutions(self, xis, batch=64, rel_tol=1e-5, inner_rel_tol=1e-4, precond=PRECOND_DEPTH_BLOCK):
        """U[s, j] (float32, device order) for every xi sample s and mode j, from batched
        mixed-precision solves to rel_tol (inner solves with the preconditioner `precond`)."""
        import torch
        xis = np.asarray(xis, dtype=np.float64)
        S, J, n = len(xis), self.n_eta, self.grid.size
        per = max(1, batch // J)              # xi samples per batched solve
        B = per * J
        st0 = VariableCoefficient(COEF_EXPSIN, xi=xis[0])
        op64 = MatrixVectorHomogen(self.grid, st0, precision=64, batch=B)
        op32 = MatrixVectorHomogen(self.grid, st0, precision=32, batch=B).set_precond(precond)
        eta = np.tile(np.eye(J), (per, 1))   # instance b solves mode b % J
        f = torch.zeros(B, n, dtype=torch.float64, device="cuda")
        op64.rhs_sines(eta, self.modes, out=f)
        x = torch.zeros_like(f)
        U = torch.zeros(S, J, n, dtype=torch.float32, device="cuda")
        solves = 0
        t0 = time.perf_counter()
        for s0 in range(0, S, per):
            cnt = min(per, S - s0)
            cs = [VariableCoefficient(COEF_EXPSIN, xi=xis[min(s0 + b // J, S - 1)]) for b in range(B)]
            op64.set_coefficients(cs)
            op32.set_coefficients(cs)
            x.zero_()
            CG.solveMixed(rel_tol, x, f, op64, op32, inner_rel_tol=inner_rel_tol)
            U[s0:s0 + cnt] = x[:cnt * J].view(cnt, J, n).float()
            solves += cnt * J
            self._report(t0, "basis solves %d / %d" % (solves, S * J))
        torch.cuda.synchronize()
        self.stats.update({"basis_solves": solves, "basis_solve_s": time.perf_counter() - t0})
        return U

    def _report(self, t0, msg):
        now = time.perf_counter()
        if self.progress and now - getattr(self, "_last_report", 0.0) > 20.0:
            self._last_report = now
            print("warm start: %s (%.0f s)" % (msg, now - t0), file=sys.stderr, flush=True)

    def features(self, xi, eta):
        import torch
        f = np.concatenate([np.asarray(xi, dtype=np.float64) / self.xi_scale, np.asarray(eta, dtype=np.float64)],
                           axis=1)
        return torch.tensor(f, dtype=torch.float32, device="cuda")

    # ------------------------------------------------------------- train
    def fit(self, xis, U, steps=2000, batch=256, lr=1e-3, lr_final=1e-5, seed=0):
        """Adam on synthetic samples (xi_s, eta) -> sum_j eta_j U[s, j]; cosine-decayed lr."""
        import torch
        rng = np.random.default_rng(seed)
        xis = np.asarray(xis, dtype=np.float64)
        S, J = U.shape[0], U.shape[1]
        # output scale: RMS of D u over random eta
        Us = (U * self.D).contiguous()
        probe = torch.randn(min(S, 256), 1, J, device="cuda") @ Us[:min(S, 256)]
        self.u_scale = float(torch.sqrt(torch.mean(probe.double() ** 2)))
        Us /= self.u_scale
        t0 = time.perf_counter()
        losses = []
        tgt = torch.empty(batch, Us.shape[2], dtype=torch.float32, device="cuda")
        for it in range(steps):
            idx = rng.integers(0, S, batch)
            eta = rng.standard_normal((batch, J))
            self.net.mix_rows(Us, torch.tensor(idx, dtype=torch.int64, device="cuda"),
                              torch.tensor(eta, dtype=torch.float32, device="cuda"), out=tgt)
            c = 0.5 * (1 + np.cos(np.pi * it / steps))
            losses.append(self.net.train_step(self.features(xis[idx], eta), tgt, lr_final + (lr - lr_final) * c))
            if it % 1000 == 0:
                self._report(t0, "training step %d / %d, loss %.3e" % (it, steps, losses[-1]))
        torch.cuda.synchronize()
        self.stats.update({"train_steps": steps, "train_batch": batch, "train_s": time.perf_counter() - t0,
                           "train_loss_first": losses[0], "train_loss_last": float(np.mean(losses[-50:]))})
        return losses

    # ----------------------------------------------------------- predict
    def predict(self, xi, eta, out=None):
        """x0 (float64, device order) for each row of (xi, eta)."""
        import torch
        y = self.net.forward(self.features(xi, eta))
        y.mul_(self.u_scale).div_(self.D)
        return y.double() if out is None else out.copy_(y)
