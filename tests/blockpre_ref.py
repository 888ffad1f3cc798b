"""CPU restatement of the depth-block preconditioner (lsg_op_set_precond with
LSG_PRECOND_DEPTH_BLOCK) for the tests: no counterpart in the reference, so this numpy
version is its checker.

For every depth block T the preconditioner holds the Galerkin block of the
constant-coefficient Laplacian on the prewavelets of depth T,
    B_T = sum_i K_(T_i) (x) prod_(j != i) M_(T_j),
with the 1D prewavelet stiffness and mass matrices of level t (prewavelets of the odd level-t
positions: 1 / -0.6 / 0.1, 0.9 at the centre next to the boundary, as in
indices/index.cc:180-327), and applies h = s .* (B_T^-1 (s .* r)) with s = sqrt(D0 / D).  On adaptive grids the blocks
are taken over the active points (hanging points keep h = r / D).  The library keeps dense
inverses up to max_dense points (128 on FP32 operators, 64 on FP64 ones) and applies larger
blocks whose active points form a product set of 1D positions (up to fd_maxm per dimension and
max_fd points) by fast diagonalisation, which is the same B_T^-1 up to rounding; other blocks
fall back to h = r / D.
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
    """B_T^-1 by fast diagonalisation, for a block whose points form a product set
    S_1 x ... x S_d of 1D positions: (x)_i V_i diag(1 / sum_i lam_i) (x)_i V_i^T with the
    generalised eigenvectors K_i V_i = M_i V_i diag(lam_i), V_i^T M_i V_i = I of the restrictions
    K_i = K_(T_i)[S_i, S_i], M_i = M_(T_i)[S_i, S_i]."""
    import scipy.linalg as sl
    _, loc = depth_blocks(idx_block)
    d = len(T)
    nb = len(idx_block)
    Vfull = np.ones((nb, nb))
    lamsum = np.zeros(nb)
    for i in range(d):
        S, rank = np.unique(loc[:, i], return_inverse=True)
        K, M = prewavelet_1d(T[i])
        lam, V = sl.eigh(K[np.ix_(S, S)], M[np.ix_(S, S)])
        Vfull = Vfull * V[np.ix_(rank, rank)]
        lamsum += lam[rank]
    return Vfull @ np.diag(1.0 / lamsum) @ Vfull.T


def make_preconditioner(idx, D, max_dense=128, max_fd=2048, fd_maxm=256, active=None):
    """The preconditioner as a function r -> h for the diagonal D of the operator.

    Blocks are taken over the active points of a depth (hanging points: h = r / D, which is r
    as the operator's diagonal is 1 there).  A block keeps its exact inverse when it has at most
    max_dense points (hanging ones included), or when its active points form a product set
    S_1 x ... x S_d of 1D positions with |S_i| <= fd_maxm and at most max_fd points (the library
    applies those by fast diagonalisation); other blocks take h = r / D."""
    blocks = constant_blocks(idx)
    _, loc = depth_blocks(idx)
    n = len(idx)
    act = np.ones(n, bool) if active is None else np.asarray(active) == 1
    D0 = np.ones(n)
    parts = []
    for T, (ids, B) in blocks.items():
        keep = act[ids]
        ida = ids[keep]
        if len(ida) == 0:
            continue
        Bk = B[np.ix_(keep, keep)]
        D0[ida] = np.diag(Bk)
        sets = [np.unique(loc[ida, i]) for i in range(len(T))]
        product = (int(np.prod([len(q) for q in sets])) == len(ida) and len(ida) <= max_fd
                   and max(len(q) for q in sets) <= fd_maxm)
        exact = len(ids) <= max_dense or product
        parts.append((ida, np.linalg.inv(Bk) if exact else None))
    s = np.sqrt(D0 / D)

    def apply(r):
        h = r / D
        for ids, Binv in parts:
            if Binv is not None:
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
