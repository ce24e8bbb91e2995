"""The checker of the MRILU-interface preconditioner pinned by what block
ILU(0) must satisfy: (L U)_IJ = A_IJ on the block pattern of A; exact LU
(L U = A) when the pattern is closed under elimination (block tridiagonal)
or full; the apply is then a solve with A."""
import numpy as np
import pytest
import scipy.sparse as sp

from bilu_ref import bilu0, bilu0_apply, block_pattern, grid_matrix, lu_product


@pytest.mark.parametrize("b", [1, 2, 3])
def test_ilu0_matches_a_on_its_pattern(b):
    A = grid_matrix(5, 4, b, 1)
    fac = bilu0(A, b)
    LU = lu_product(fac, b)
    Ad = A.toarray()
    pat = block_pattern(A, b)
    for I in range(len(pat)):
        for J in pat[I]:
            np.testing.assert_allclose(LU[I * b:(I + 1) * b, J * b:(J + 1) * b],
                                       Ad[I * b:(I + 1) * b, J * b:(J + 1) * b], rtol=0, atol=1e-12 * np.abs(Ad).max())
    assert np.abs(LU - Ad).max() > 1e-6   # fill outside the pattern is dropped


@pytest.mark.parametrize("b", [1, 4])
def test_exact_on_closed_patterns(b):
    rng = np.random.default_rng(2)
    nb = 12
    A = sp.lil_matrix((nb * b, nb * b))
    for I in range(nb):
        for J in (I - 1, I, I + 1):
            if 0 <= J < nb:
                B = rng.normal(size=(b, b)) + (6.0 * b * np.eye(b) if I == J else 0)
                A[I * b:(I + 1) * b, J * b:(J + 1) * b] = B
    A = A.tocsr()
    v = rng.normal(size=nb * b)
    x = bilu0_apply(bilu0(A, b), v, b)
    assert np.abs(A @ x - v).max() <= 1e-12 * np.abs(v).max()
    Af = sp.csr_matrix(rng.normal(size=(3 * b, 3 * b)) + 10 * b * np.eye(3 * b))
    w = rng.normal(size=3 * b)
    np.testing.assert_allclose(bilu0_apply(bilu0(Af, b), w, b), np.linalg.solve(Af.toarray(), w), rtol=1e-12, atol=1e-13)
