"""Device products and solver: J x, J^T x against scipy on the exported
reference-order CSR, the preconditioner, and FGMRES to the FP64 residual bar."""
import numpy as np
import pytest
import scipy.sparse as sp

from gpu_helpers import gpu_ocean
from thcm_cases import box16_case, load_golden

pytestmark = pytest.mark.gpu


def _csr(oc):
    rp, ci, val, dB = oc.getJacobian()
    return sp.csr_matrix((val, ci, rp), shape=(oc.ndim, oc.ndim))


@pytest.mark.parametrize("name", ["natl8", "gateway16", "box16"])
def test_matvec_and_transpose_match_csr(name):
    case, _ = load_golden(name)
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    J = _csr(oc)
    rng = np.random.default_rng(5)
    v = rng.normal(size=oc.ndim)
    y = oc.applyMatrix(v)
    ref = J @ v
    scale = np.abs(J) @ np.abs(v)
    assert np.all(np.abs(y - ref) <= 1e-13 * scale + 1e-300)
    yt = oc.applyMatrixTranspose(v)
    reft = J.T @ v
    scalet = np.abs(J.T) @ np.abs(v)
    assert np.all(np.abs(yt - reft) <= 1e-13 * scalet + 1e-300)


def test_mass_matrix_matches_diag_B():
    case, exp = load_golden("box16")
    oc = gpu_ocean(case)
    v = np.random.default_rng(1).normal(size=oc.ndim)
    np.testing.assert_array_equal(oc.applyMassMat(v), exp["diagB"] * v)


def test_solve_reaches_fp64_residual_bar():
    """Ocean::solve at the 16x16x12 box state: ||b - J x|| / ||b|| <= 1e-8,
    checked on the host against the exported Jacobian."""
    case = box16_case()
    oc = gpu_ocean(case)
    F = oc.computeRHS(case.x)
    oc.computeJacobian(case.x)
    b = -F
    x, info = oc.solve(b, oc.solver_params(tol=1e-8, restart=250, maxit=2000), raise_on_noconv=False)
    J = _csr(oc)
    rel = np.linalg.norm(b - J @ x) / np.linalg.norm(b)
    print(f"iterations {info.iterations} reported {info.rel_residual:.3e} host {rel:.3e}")
    assert abs(rel - info.rel_residual) <= 1e-3 * max(rel, 1e-12) + 1e-14
    assert rel <= 1e-8
