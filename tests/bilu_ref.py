"""Test-side restatement of block ILU(0) (Saad, Iterative Methods for Sparse
Linear Systems, 2nd ed., sec. 10.3, over b x b blocks): the checker of the
library's MRILU-interface preconditioner.  Dense numpy blocks, plain loops:
for small matrices only."""
import numpy as np
import scipy.sparse as sp


def block_pattern(A, b):
    A = sp.csr_matrix(A)
    nb = A.shape[0] // b
    pat = [set([I]) for I in range(nb)]
    coo = A.tocoo()
    for r, c in zip(coo.row, coo.col):
        pat[r // b].add(c // b)
    return [sorted(s) for s in pat]


def bilu0(A, b, lutol=1e-14):
    """Returns (pattern, blocks dict (I, J) -> b x b factor block, Dinv list)."""
    A = sp.csr_matrix(A).toarray()
    nb = A.shape[0] // b
    pat = block_pattern(A, b)
    blk = {(I, J): A[I * b:(I + 1) * b, J * b:(J + 1) * b].copy() for I in range(nb) for J in pat[I]}
    dinv = [None] * nb
    for I in range(nb):
        row = pat[I]
        for K in [K for K in row if K < I]:
            blk[(I, K)] = blk[(I, K)] @ dinv[K]
            for J in [J for J in row if J > K]:
                if (K, J) in blk:
                    blk[(I, J)] = blk[(I, J)] - blk[(I, K)] @ blk[(K, J)]
        D = blk[(I, I)].copy()
        scale = np.abs(D).max() if np.abs(D).max() > 0 else 1.0
        # partial pivoting with the floor, as LAPACK-style elimination
        M = np.hstack([D, np.eye(b)])
        for c in range(b):
            p = c + int(np.argmax(np.abs(M[c:, c])))
            M[[c, p]] = M[[p, c]]
            if abs(M[c, c]) < lutol * scale:
                M[c, c] = -lutol * scale if M[c, c] < 0 else lutol * scale
            M[c] = M[c] / M[c, c]
            for r in range(b):
                if r != c and M[r, c] != 0.0:
                    M[r] = M[r] - M[r, c] * M[c]
        dinv[I] = M[:, b:]
    return pat, blk, dinv


def bilu0_apply(fac, v, b):
    pat, blk, dinv = fac
    nb = len(pat)
    y = np.zeros_like(v)
    for I in range(nb):
        acc = v[I * b:(I + 1) * b].copy()
        for K in pat[I]:
            if K < I:
                acc -= blk[(I, K)] @ y[K * b:(K + 1) * b]
        y[I * b:(I + 1) * b] = acc
    x = np.zeros_like(v)
    for I in range(nb - 1, -1, -1):
        acc = y[I * b:(I + 1) * b].copy()
        for J in pat[I]:
            if J > I:
                acc -= blk[(I, J)] @ x[J * b:(J + 1) * b]
        x[I * b:(I + 1) * b] = dinv[I] @ acc
    return x


def lu_product(fac, b):
    """L U of the factors (unit block-lower L, block-upper U with D_I = inv(Dinv_I))."""
    pat, blk, dinv = fac
    nb = len(pat)
    n = nb * b
    L = np.eye(n)
    U = np.zeros((n, n))
    for I in range(nb):
        for J in pat[I]:
            if J < I:
                L[I * b:(I + 1) * b, J * b:(J + 1) * b] = blk[(I, J)]
            elif J > I:
                U[I * b:(I + 1) * b, J * b:(J + 1) * b] = blk[(I, J)]
        U[I * b:(I + 1) * b, I * b:(I + 1) * b] = np.linalg.inv(dinv[I])
    return L @ U


def grid_matrix(nx, ny, b, seed):
    """A nonsymmetric 5-point-stencil matrix over b unknowns per grid point,
    diagonally dominant by blocks."""
    rng = np.random.default_rng(seed)
    nb = nx * ny
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            I = i + nx * j
            for di, dj in ((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny:
                    J = ii + nx * jj
                    B = rng.normal(size=(b, b)) * (0.3 if I != J else 1.0)
                    if I == J:
                        B += 4.0 * b * np.eye(b)
                    for r in range(b):
                        for c in range(b):
                            rows.append(I * b + r); cols.append(J * b + c); vals.append(B[r, c])
    return sp.csr_matrix((vals, (rows, cols)), shape=(nb * b, nb * b))
