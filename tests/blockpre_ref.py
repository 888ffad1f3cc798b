"""CPU restatement of the depth-block preconditioner (lsg_op_set_precond with
LSG_PRECOND_DEPTH_BLOCK) for the tests: no counterpart in the reference, so this numpy
version is its checker.

For every depth block T the preconditioner holds the Galerkin block of the
constant-coefficient Laplacian on the prewavelets of depth T,
    B_T = sum_i K_(T_i) (x) prod_(j != i) M_(T_j),
with the 1D prewavelet stiffness and mass matrices of level t (prewavelets of the odd level-t
positions: 1 / -0.6 / 0.1, 0.9 at the centre next to the boundary, as in
indices/index.cc:180-327), and applies h = s .* (B_T^-1 (s .* r)) with s = sqrt(D0 / D).  The
library keeps dense inverses up to max_dense points (128 on FP32 operators, 64 on FP64 ones) and
applies larger blocks of up to max_fd points whose levels are at most fd_maxlev by fast
diagonalisation, which is the same B_T^-1 up to rounding; other blocks fall back to h = r / D.
tests/test_blockpre_ref.py checks the Kronecker blocks against the oracle's constant-coefficient
matvec and the fast diagonalisation against the dense inverse.
"""
import numpy as np


def prewavelet_1d(t):
    """(K, M): 1D Galerkin matrices of the level-t prewavelets, m = 2^(t-1)."""
    nn, m = (1 << t) - 1, 1 << (t - 1)
    P = np.zeros((nn, m))
    if t == 1:
        P[0, 0] = 1.0
    else:
        for j in range(m):
            k = 2 * j + 1
            c = 0.9 if (k == 1 or k == nn) else 1.0
            for off, v in ((0, c), (-1, -0.6), (1, -0.6), (-2, 0.1), (2, 0.1)):
                q = k + off
                if 1 <= q <= nn:
                    P[q - 1, j] = v
    h = 2.0 ** -t
    K = (np.diag(2 * np.ones(nn)) - np.diag(np.ones(nn - 1), 1) - np.diag(np.ones(nn - 1), -1)) / h
    M = (np.diag(4 * np.ones(nn)) + np.diag(np.ones(nn - 1), 1) + np.diag(np.ones(nn - 1), -1)) * h / 6
    return P.T @ K @ P, P.T @ M @ P


def depth_blocks(idx):
    """{depth tuple: slot indices} and the local 1D indices of every slot."""
    idx = np.asarray(idx)
    T = np.floor(np.log2(idx)).astype(int)
    loc = idx - (1 << T)
    groups = {}
    for i in range(len(idx)):
        groups.setdefault(tuple(T[i]), []).append(i)
    return groups, loc


def constant_blocks(idx):
    """{depth: (slots, B_T)} of the constant-coefficient Galerkin blocks."""
    groups, loc = depth_blocks(idx)
    d = np.asarray(idx).shape[1]
    mats = {}
    out = {}
    for T, ids in groups.items():
        ids = np.array(ids)
        L = loc[ids]
        B = np.zeros((len(ids), len(ids)))
        for i in range(d):
            term = None
            for j in range(d):
                t = T[j]
                if t not in mats:
                    mats[t] = prewavelet_1d(t)
                A = mats[t][0] if j == i else mats[t][1]
                f = A[np.ix_(L[:, j], L[:, j])]
                term = f if term is None else term * f
            B += term
        out[T] = (ids, B)
    return out


def fast_diagonal_inverse(T, idx_block):
    """B_T^-1 by fast diagonalisation: (x)_i V_(T_i) diag(1 / sum_i lam) (x)_i V_(T_i)^T with
    the generalised eigenvectors K V = M V diag(lam), V^T M V = I of every 1D factor."""
    import scipy.linalg as sl
    groups, loc = depth_blocks(idx_block)
    L = loc
    d = len(T)
    fac = []
    for t in T:
        K, M = prewavelet_1d(t)
        lam, V = sl.eigh(K, M)
        fac.append((lam, V))
    nb = len(idx_block)
    Vfull = np.ones((nb, nb))
    lamsum = np.zeros(nb)
    for i in range(d):
        lam, V = fac[i]
        Vfull = Vfull * V[np.ix_(L[:, i], L[:, i])]
        lamsum += lam[L[:, i]]
    return Vfull @ np.diag(1.0 / lamsum) @ Vfull.T


def make_preconditioner(idx, D, max_dense=128, max_fd=1024, fd_maxlev=9):
    """The preconditioner as a function r -> h for the diagonal D of the operator."""
    blocks = constant_blocks(idx)
    n = len(idx)
    D0 = np.zeros(n)
    parts = []
    for T, (ids, B) in blocks.items():
        D0[ids] = np.diag(B)
        exact = len(ids) <= max_dense or (len(ids) <= max_fd and max(T) <= fd_maxlev)
        parts.append((ids, np.linalg.inv(B) if exact else None))
    s = np.sqrt(D0 / D)

    def apply(r):
        h = np.empty_like(r)
        for ids, Binv in parts:
            if Binv is None:
                h[ids] = r[ids] / D[ids]
            else:
                h[ids] = s[ids] * (Binv @ (s[ids] * r[ids]))
        return h
    return apply


def pcg(A_mul, b, M, tol, maxit=5000):
    """PCG with the relative-residual stop ||r|| <= tol ||b||; returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    nb = np.linalg.norm(b)
    for it in range(1, maxit + 1):
        q = A_mul(p)
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= tol * nb:
            return x, it
        z = M(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxit
