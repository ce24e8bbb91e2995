"""The preconditioner plug-in behind the reference's MRILU interface
(mrilucpp.F90:120-554, Ifpack_MRILU.cpp:25-37) through the C ABI: handle
life cycle as the reference defines it, the GPU block ILU(0) against the
checker's restatement (tests/bilu_ref.py, pinned by the ILU(0) identities),
exactness on closed patterns, the reference's own symbol names, and its use
in the slot TRIOS gives it: the temperature/salinity block of a THCM
Jacobian (TRIOS_BlockPreconditioner.C:1919-1969 solves ATS with it)."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import iemic_amd
from bilu_ref import bilu0, bilu0_apply, grid_matrix
from gpu_helpers import gpu_ocean
from thcm_cases import box16_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("b", [1, 2, 3, 6])
def test_gpu_bilu_equals_the_restatement(b):
    A = grid_matrix(7, 5, b, 11)
    v = np.random.default_rng(3).normal(size=A.shape[0])
    P = iemic_amd.MriluPrec(A, blocksize=b)
    P.compute()
    got = P.apply(v)
    want = bilu0_apply(bilu0(A, b), v, b)
    err = np.abs(got - want).max() / np.abs(want).max()
    blocks, levels, floored = P.info()
    print(f"b={b}: n {A.shape[0]}, {blocks} blocks, {levels} levels, max difference {err:.2e}")
    assert err <= 1e-12 and floored == 0
    P.close()


def test_gpu_bilu_is_exact_on_a_block_tridiagonal_matrix():
    rng = np.random.default_rng(4)
    b, nb = 4, 300
    A = sp.lil_matrix((nb * b, nb * b))
    for I in range(nb):
        for J in (I - 1, I, I + 1):
            if 0 <= J < nb:
                A[I * b:(I + 1) * b, J * b:(J + 1) * b] = rng.normal(size=(b, b)) + (6.0 * b * np.eye(b) if I == J else 0)
    A = A.tocsr()
    P = iemic_amd.MriluPrec(A, blocksize=b)
    P.compute()
    v = rng.normal(size=nb * b)
    x = P.apply(v)
    print(f"A M^-1 v - v: {np.abs(A @ x - v).max():.2e}")
    assert np.abs(A @ x - v).max() <= 1e-12 * np.abs(v).max()
    P.close()


def test_life_cycle_follows_the_reference():
    """create with id 0 returns a new id; create with that id overwrites it;
    apply before compute copies the right-hand side; compute twice is an
    error; an unknown id is an error; destroy sets the id to 0."""
    lib = iemic_amd.load()
    A = sp.csr_matrix(np.diag(np.arange(1.0, 9.0)))
    beg, jco, co = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    args = (8, len(co), beg.ctypes.data_as(ip), jco.ctypes.data_as(ip), co.ctypes.data_as(dp))
    pid = C.c_int(0)
    assert lib.iemic_mrilu_create(C.byref(pid), *args) == 0 and pid.value > 0
    first = pid.value
    assert lib.iemic_mrilu_create(C.byref(pid), *args) == 0 and pid.value == first
    v = np.arange(8.0)
    out = np.zeros(8)
    assert lib.iemic_mrilu_apply(first, 8, v.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    assert np.array_equal(out, v)
    assert lib.iemic_mrilu_compute(first) == 0
    assert lib.iemic_mrilu_compute(first) == -5          # IEMIC_ESTATE: factored once per matrix
    assert lib.iemic_mrilu_apply(first, 8, v.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    np.testing.assert_allclose(out, v / np.arange(1.0, 9.0), rtol=1e-15)
    assert lib.iemic_mrilu_apply(first, 7, v.ctypes.data_as(dp), out.ctypes.data_as(dp)) == -1
    assert lib.iemic_mrilu_compute(first + 37) == -1      # unknown id
    assert lib.iemic_mrilu_destroy(C.byref(pid)) == 0 and pid.value == 0
    assert lib.iemic_mrilu_compute(first) == -1


def test_the_reference_symbol_names():
    """mrilucpp_create / set_params / compute / apply / destroy with every
    scalar by address, as Ifpack_MRILU.cpp calls them."""
    lib = C.CDLL(iemic_amd.LIB_PATH)
    A = grid_matrix(4, 4, 2, 5)
    beg, jco, co = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
    n, nnz = C.c_int(A.shape[0]), C.c_int(len(co))
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    pid = C.c_int(0)
    lib.mrilucpp_create(C.byref(pid), C.byref(n), C.byref(nnz), beg.ctypes.data_as(ip), jco.ctypes.data_as(ip),
                        co.ctypes.data_as(dp))
    I = lambda v: C.byref(C.c_int(v))
    D = lambda v: C.byref(C.c_double(v))
    lib.mrilucpp_set_params(C.byref(pid), I(2), I(0), I(0), I(0), I(0), D(1.0), D(1e-2), D(0.1), I(0), D(1.0),
                            D(0.8), D(1e-6), D(1e-3), D(1.0), D(1.0), D(0.675), I(9), D(1e-4), D(1.0), D(0.1),
                            D(1e-14), I(1), I(2))
    lib.mrilucpp_compute(C.byref(pid))
    v = np.random.default_rng(6).normal(size=A.shape[0])
    sol = np.zeros_like(v)
    lib.mrilucpp_apply(C.byref(pid), C.byref(n), v.ctypes.data_as(dp), sol.ctypes.data_as(dp))
    np.testing.assert_allclose(sol, bilu0_apply(bilu0(A, 2), v, 2), rtol=1e-12, atol=1e-14)
    lib.mrilucpp_destroy(C.byref(pid))
    assert pid.value == 0


def _fgmres(A, M, b, tol, maxit):
    n = b.size
    nb = np.linalg.norm(b)
    V = np.zeros((maxit + 1, n)); Z = np.zeros((maxit, n)); H = np.zeros((maxit + 1, maxit))
    V[0] = b / nb
    for j in range(maxit):
        Z[j] = M(V[j]); w = A @ Z[j]
        for _ in range(2):
            h = V[:j + 1] @ w; H[:j + 1, j] += h; w -= V[:j + 1].T @ h
        H[j + 1, j] = np.linalg.norm(w); V[j + 1] = w / H[j + 1, j]
        e = np.zeros(j + 2); e[0] = nb
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e, rcond=None)[0]
        x = Z[:j + 1].T @ y
        rel = np.linalg.norm(b - A @ x) / nb
        if rel <= tol:
            return x, j + 1, rel
    return x, maxit, rel


def test_ats_block_of_a_thcm_jacobian():
    """The T,S block of the box16 Jacobian (the block TRIOS hands to its
    ATS solver), T and S of a cell as one 2 x 2 block: FGMRES with the block
    ILU(0) reaches 1e-8 in a fraction of the iterations it takes without."""
    case = box16_case()
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    rp, ci, val, _ = oc.getJacobian()
    J = sp.csr_matrix((val, ci, rp), shape=(oc.ndim, oc.ndim))
    ts = np.nonzero(np.isin(np.arange(oc.ndim) % 6, (4, 5)))[0]
    ric = oc.rowintcon()
    A = J[ts][:, ts].tocsr()
    if ric >= 0:   # the dense integral row is not part of ATS in TRIOS: its S equation's diagonal stands in
        k = int(np.nonzero(ts == ric)[0][0])
        A = A.tolil(); d = A[k, k]; A[k, :] = 0; A[k, k] = d if d != 0 else -1.0; A = A.tocsr()
    b = A @ np.random.default_rng(7).normal(size=A.shape[0])
    P = iemic_amd.MriluPrec(A, blocksize=2)
    P.compute()
    x, its, rel = _fgmres(A, P.apply, b, 1e-8, 200)
    _, its0, rel0 = _fgmres(A, lambda v: v, b, 1e-8, 200)
    print(f"ATS of box16 (n {A.shape[0]}): block ILU(0) {its} iterations (residual {rel:.1e}), "
          f"none {its0} (residual {rel0:.1e}); levels {P.info()[1]}")
    assert rel <= 1e-8 and its < its0 / 3
    P.close()
    oc.close()
