"""Device products and solver: J x, J^T x against scipy on the exported
reference-order CSR, the preconditioners, and FGMRES to the FP64 residual bar.

The bar for a solve is the one of Ocean::solve: a true FP64 relative
residual ||b - J x|| / ||b|| <= 1e-8, recomputed here on the host from the
exported Jacobian (the reference's own iterates are not comparable: its
Belos/TRIOS solver is a different algorithm)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import iemic_amd
from gpu_helpers import gpu_ocean
from thcm_cases import (OracleModel, box16_case, checkerboard_modes, load_golden, pressure_projection,
                        seeded_state)

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


@pytest.mark.parametrize("name", ["natl8", "box16", "gateway16"])
def test_pressure_projection(name):
    """Ocean::pressureProjection against its restatement; the projected
    vector is orthogonal to both modes, and the modes are null vectors of J
    (so the projection changes no product with J)."""
    case, _ = load_golden(name)
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    lm = OracleModel(case).landm()
    v = np.random.default_rng(4).normal(size=oc.ndim) + 3.0
    got = oc.pressureProjection(v)
    ref = pressure_projection(case.cfg, lm, v)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(v))
    s = checkerboard_modes(case.cfg, lm)
    assert np.all(np.abs(s @ got) <= 1e-12 * np.linalg.norm(v))
    J = _csr(oc)
    assert np.linalg.norm(J @ s[0]) <= 1e-12 * abs(J).max()
    assert np.linalg.norm(J @ s[1]) <= 1e-12 * abs(J).max()


def test_solve_ignores_the_inconsistent_part_of_b():
    """J is singular (one null vector per pressure basin), so J x = b has a
    solution only for b orthogonal to the left null vectors m of J.  For
    b = J y + e m the solve returns x with J x = J y to the FP64 bar, e.g.
    a Newton residual whose continuity rows carry rounding along m."""
    case, _ = load_golden("natl8")
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    J = _csr(oc)
    U, S, _ = np.linalg.svd(J.toarray())
    left = U[:, S <= 1e-13 * S[0]]
    assert left.shape[1] == 2
    Jy = J @ np.random.default_rng(6).normal(size=oc.ndim)
    for eps in (1e-3, 1.0):
        b = Jy + eps * np.linalg.norm(Jy) * left[:, 0]
        x, info = oc.solve(b, oc.solver_params(tol=1e-10, restart=50, maxit=200), raise_on_noconv=False)
        rel = np.linalg.norm(Jy - J @ x) / np.linalg.norm(Jy)
        print(f"eps {eps}: {info.iterations} iterations, consistent-part residual {rel:.2e}, "
              f"reported {info.rel_residual:.2e}")
        assert rel <= 1e-8
        assert info.iterations <= 10


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
    x, info = oc.solve(b, oc.solver_params(tol=1e-8, restart=50, maxit=200), raise_on_noconv=False)
    J = _csr(oc)
    rel = np.linalg.norm(b - J @ x) / np.linalg.norm(b)
    print(f"iterations {info.iterations} reported {info.rel_residual:.3e} host {rel:.3e}")
    assert abs(rel - info.rel_residual) <= 1e-3 * max(rel, 1e-12) + 1e-14
    assert rel <= 1e-8


def _case(name):
    case, _ = load_golden(name)
    if name in ("gateway16", "global96"):
        case.x = seeded_state(case.cfg, case.landm, 11)   # away from the stored equilibrium
    return case


@pytest.mark.parametrize("name", ["natl8", "natl8_eos", "box16", "gateway16", "global96"])
def test_row_lu_solve_fp64(name):
    """Newton system J y = F with the row-block LU preconditioner."""
    case = _case(name)
    oc = gpu_ocean(case)
    F = oc.computeRHS(case.x)
    oc.computeJacobian(case.x)
    y, info = oc.solve(F, oc.solver_params(tol=1e-8, restart=50, maxit=200), raise_on_noconv=False)
    rel = np.linalg.norm(F - _csr(oc) @ y) / np.linalg.norm(F)
    print(f"{name}: {info.iterations} iterations, residual {rel:.3e}")
    assert rel <= 1e-8
    assert info.iterations <= 30


@pytest.mark.parametrize("lu", ["rocsolver", "tournament", "tournament_ipiv"])
@pytest.mark.parametrize("name", ["natl8", "box16", "gateway16", "global96"])
def test_row_lu_is_exact_for_the_factored_matrix(name, lu):
    """M^{-1} is an exact inverse of J except in the rows the factored matrix
    replaces (one continuity row per pressure basin, the integral row):
    J M^{-1} v = v on every other row.  Both dense LUs of the row blocks:
    rocSOLVER getrf (partial pivoting) and the library's tournament-pivoted
    LU (dense_lu.hip), the latter with its row moves worked out from the
    chosen rows (the default) and from LAPACK's interchange sequence
    (IEMIC_LU_IPIV=1)."""
    case = _case(name)
    os.environ["IEMIC_DENSE_LU"] = "0" if lu == "rocsolver" else "1"
    os.environ["IEMIC_LU_IPIV"] = "1" if lu == "tournament_ipiv" else "0"
    try:
        oc = gpu_ocean(case)
        oc.computeJacobian(case.x)
        oc.buildPreconditioner()
    finally:
        del os.environ["IEMIC_DENSE_LU"]
        del os.environ["IEMIC_LU_IPIV"]
    pins = oc.preconditionerInfo().pins
    J = _csr(oc)
    v = J @ np.random.default_rng(3).normal(size=oc.ndim)
    z = oc.applyPrecon(v)
    # componentwise backward error of each row: rounding-level on the rows
    # the factored matrix keeps (up to ~1e-8 where the explicit block
    # inverses are ill-conditioned), O(1) on the replaced ones
    r = np.abs(J @ z - v) / (abs(J) @ np.abs(z) + np.abs(v))
    bad = np.nonzero(r > 1e-6)[0]
    print(f"{name}: pins {pins}, rows off {len(bad)} {np.sort(r)[::-1][:pins + 3]}, "
          f"max backward error elsewhere {np.delete(r, bad).max():.2e}")
    assert len(bad) <= pins + 1


@pytest.mark.parametrize("name", ["gateway16", "global96"])
def test_row_lu_with_look_ahead(name):
    """The tournament-pivoted LU with its panels factored one outer block
    ahead on a second stream (the row interchanges outside a panel replayed
    later from the recorded move lists): the exactness of the plain build
    (J M^{-1} v = v off the replaced rows, to the same backward error), and
    the same M^{-1} v up to the rounding of the differently split trailing
    products, which the explicit inverses of these ill-conditioned blocks
    amplify (gateway16: 5e-6 of max |z|; a misplaced interchange would break
    the exactness instead)."""
    case = _case(name)
    outs = {}
    for mode in ("0", "1"):
        os.environ["IEMIC_LU_LOOKAHEAD"] = mode
        try:
            oc = gpu_ocean(case)
            oc.computeJacobian(case.x)
            oc.buildPreconditioner()
        finally:
            del os.environ["IEMIC_LU_LOOKAHEAD"]
        J = _csr(oc)
        v = J @ np.random.default_rng(3).normal(size=oc.ndim)
        outs[mode] = (oc.applyPrecon(v), oc.preconditionerInfo().pins)
        oc.close()
    z0, pins = outs["0"]
    z1, _ = outs["1"]
    d = np.abs(z1 - z0).max() / np.abs(z0).max()
    r0 = np.abs(J @ z0 - v) / (abs(J) @ np.abs(z0) + np.abs(v))
    r1 = np.abs(J @ z1 - v) / (abs(J) @ np.abs(z1) + np.abs(v))
    e0, e1 = np.sort(r0)[-pins - 2], np.sort(r1)[-pins - 2]
    print(f"{name}: look-ahead against one stream {d:.2e}, rows off {(r1 > 1e-6).sum()} (pins {pins}), "
          f"backward error off the replaced rows {e1:.1e} (one stream {e0:.1e})")
    assert (r1 > 1e-6).sum() <= pins + 1
    assert e1 <= 10 * e0 + 1e-12
    assert d <= 1e-4


@pytest.mark.parametrize("name", ["gateway16", "global96"])
def test_row_lu_with_rocsolver_lu(name):
    """The row blocks factored by rocSOLVER getrf (partial pivoting) and
    inverted by getrs instead of the library's tournament-pivoted LU: the
    same FP64 solve bar."""
    case = _case(name)
    os.environ["IEMIC_DENSE_LU"] = "0"
    try:
        oc = gpu_ocean(case)
        F = oc.computeRHS(case.x)
        oc.computeJacobian(case.x)
        y, info = oc.solve(F, oc.solver_params(tol=1e-8, restart=50, maxit=200), raise_on_noconv=False)
    finally:
        del os.environ["IEMIC_DENSE_LU"]
    rel = np.linalg.norm(F - _csr(oc) @ y) / np.linalg.norm(F)
    print(f"{name}: {info.iterations} iterations, residual {rel:.3e}")
    assert rel <= 1e-8
    assert info.iterations <= 30


@pytest.mark.parametrize("name", ["natl8", "box16"])
def test_column_ilu_solve_fp64(name):
    """The water-column block ILU(0) alternative converges on the small cases."""
    case = _case(name)
    oc = gpu_ocean(case)
    oc.setPreconditioner(iemic_amd.PREC_COLUMN_ILU)
    F = oc.computeRHS(case.x)
    oc.computeJacobian(case.x)
    y, info = oc.solve(F, oc.solver_params(tol=1e-8, restart=250, maxit=1000), raise_on_noconv=False)
    rel = np.linalg.norm(F - _csr(oc) @ y) / np.linalg.norm(F)
    print(f"{name}: {info.iterations} iterations, residual {rel:.3e}")
    assert rel <= 1e-8


def test_column_ilu_dataflow_sweep_matches_level_launches():
    """The one-launch dataflow sweep computes bit-for-bit what one launch per
    level computes (same per-column arithmetic, only the scheduling differs)."""
    case = _case("gateway16")
    v = np.random.default_rng(5).normal(size=case.ndim)
    out = {}
    for mode in ("1", "0"):
        os.environ["IEMIC_LEVEL_LAUNCH"] = mode
        try:
            oc = gpu_ocean(case)
            oc.setPreconditioner(iemic_amd.PREC_COLUMN_ILU)
            oc.computeJacobian(case.x)
            out[mode] = oc.applyPrecon(v)
            oc.close()
        finally:
            del os.environ["IEMIC_LEVEL_LAUNCH"]
    assert np.array_equal(out["0"], out["1"])


def test_solve_with_an_all_land_row_and_two_basins():
    """A 16x16x8 box cut by a zonal land wall (row j = 8 all land): a row
    block with no ocean unknowns, and two closed basins whose pressure modes
    each need a pin.  Salinity restoring (sres = 1): with a salt flux each
    closed basin would carry its own mean-salinity null mode, which the one
    integral condition cannot fix (as in the reference)."""
    from iemic_amd import basins
    from thcm_cases import Case, Config, seeded_state as _seed
    cfg_d, lm, frc, x0, pars = basins.make_box(16, 16, 8)
    cfg_d = dict(cfg_d)
    cfg_d["sres"] = 1
    lm = lm.reshape(10, 18, 18).copy()
    lm[:, 8, :] = 1
    cfg = Config(**{f: cfg_d[f] for f, _ in Config._fields_})
    case = Case("wall", cfg, lm.reshape(-1), frc, pars, _seed(cfg, lm.reshape(-1), 11))
    oc = gpu_ocean(case)
    F = oc.computeRHS(case.x)
    oc.computeJacobian(case.x)
    y, info = oc.solve(F, oc.solver_params(tol=1e-8, restart=50, maxit=200), raise_on_noconv=False)
    rel = np.linalg.norm(F - _csr(oc) @ y) / np.linalg.norm(F)
    pinfo = oc.preconditionerInfo()
    print(f"wall box: pins {pinfo.pins}, rows {pinfo.blocks}, {info.iterations} iterations, residual {rel:.2e}")
    assert pinfo.pins >= 4
    assert rel <= 1e-8
    assert info.iterations <= 30
