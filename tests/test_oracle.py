"""The CPU oracle (oracle/thcm_oracle.c) against the reference's golden
vectors and against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from thcm_cases import (OCEAN, LAND, PAR, Case, OracleModel, box_mask, box16_case,
                        expected_rows, golden_names, load_golden, make_config, restrict_rows)


@pytest.mark.parametrize("name", golden_names())
def test_oracle_matches_reference_fixture(name):
    """Bit-exact F, J (pattern, order and values), diag(B), integral row."""
    case, exp = load_golden(name)
    orc = OracleModel(case)
    assert orc.rowintcon() == int(exp["rowintcon"])
    np.testing.assert_array_equal(orc.coef(), exp["coef"])
    np.testing.assert_array_equal(orc.rhs(), exp["F"])
    rp, ci, val, dB, dropped = orc.jacobian()
    assert dropped == 0
    np.testing.assert_array_equal(dB, exp["diagB"])
    rows = expected_rows(exp, case.ndim)
    srp, sci, sval = restrict_rows(rp, ci, val, rows)
    np.testing.assert_array_equal(srp, exp["J_rowptr"])
    np.testing.assert_array_equal(sci, exp["J_col"])
    np.testing.assert_array_equal(sval, exp["J_val"])


def test_graph_row_lengths_interior():
    """CreateMaximalGraph: 24/22/7/11/20/20 entries per interior cell
    (THCM.C:2320-2325), sorted and unique; the dense integral row holds every
    S column."""
    case = box16_case()
    orc = OracleModel(case)
    rp, ci = orc.graph()
    n, m, l = 16, 16, 12
    lens = np.diff(rp).reshape(l, m, n, 6)
    np.testing.assert_array_equal(lens[5, 5, 5], [24, 22, 7, 11, 20, 20])
    for r in range(0, len(rp) - 1, 997):
        seg = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(seg) > 0)
    ric = orc.rowintcon()
    assert ric == case.ndim - 1
    np.testing.assert_array_equal(ci[rp[ric]:rp[ric + 1]], np.arange(5, case.ndim, 6))


def test_rhs_zero_at_rest_without_forcing():
    """test_ocean.C:42-50: ||F(0)|| = 0 at COMB = 0."""
    case = box16_case()
    case.pars = dict(COMB=0.0, WIND=1.0, TEMP=1.0, SALT=1.0)
    orc = OracleModel(case)
    assert np.abs(orc.rhs(np.zeros(case.ndim), 0.0)).max() < 1e-12


def test_mass_matrix_values():
    """test_ocean.C:61-125: diag(B) = -Ro (u, v), 0 (w, p), -1 (T, S), 0 at
    the integral row."""
    case = box16_case()
    orc = OracleModel(case)
    _, _, _, dB, _ = orc.jacobian()
    Ro = orc.par("ROSB")
    d = dB.reshape(-1, 6)
    n, m, l = 16, 16, 12
    dd = d.reshape(l, m, n, 6)
    # u: -Ro unless the east neighbour is land; v: unless the north one is
    assert np.all(dd[:, :, :-1, 0] == -Ro) and np.all(dd[:, :, -1, 0] == 0)
    assert np.all(dd[:, :-1, :, 1] == -Ro) and np.all(dd[:, -1, :, 1] == 0)
    assert np.all(d[:, 2:4] == 0)
    assert np.all(d[:-1, 4:6] == -1) and d[-1, 4] == -1 and d[-1, 5] == 0


def test_jacobian_at_rest_is_linear_part():
    """At x = 0 every nonlinear atom vanishes, so J(0) is the linear
    operator: (F(t d) - F(0)) / t - J(0) d = O(t) along any direction d
    (Mixing = 0 removes the convective term, whose taper is not smooth)."""
    case = box16_case()
    case.cfg = make_config(n=16, m=16, l=12, sres=0, iza=3, ite=2, its=2, vmix=0)
    orc = OracleModel(case)
    rp, ci, val, _, _ = orc.jacobian(np.zeros(case.ndim))
    rng = np.random.default_rng(1)
    d = rng.normal(size=case.ndim)
    t = 1e-4
    F0 = orc.rhs(np.zeros(case.ndim), 0.0)
    Ft = orc.rhs(t * d, 0.0)
    Jd = np.zeros(case.ndim)
    for r in range(case.ndim):
        Jd[r] = val[rp[r]:rp[r + 1]] @ d[ci[rp[r]:rp[r + 1]]]
    # quadratic terms are O(t^2): (F(td) - F(0))/t - J d = O(t)
    err = np.abs((Ft - F0) / t - Jd).max()
    assert err < 1e-3 * np.abs(Jd).max()


def _central_fd_columns(orc, x, cols, h):
    out = []
    for c in cols:
        xp = x.copy(); xm = x.copy()
        xp[c] += h; xm[c] -= h
        out.append((orc.rhs(xp) - orc.rhs(xm)) / (2 * h))
    return np.array(out).T


def test_jacobian_columns_against_central_differences():
    """check_jac (legacy/check_jac.f:1-68): J columns against central
    differences of F, land rows excluded.  The convective-mixing block of J
    is itself a forward difference with eps = 1e-8 (mix_imp.f:750-771), so the
    bound is relative to the column scale."""
    case = box16_case()
    orc = OracleModel(case)
    rp, ci, val, _, _ = orc.jacobian()
    rng = np.random.default_rng(7)
    cols = np.sort(rng.choice(case.ndim, 60, replace=False))
    fd = _central_fd_columns(orc, case.x, cols, 1e-6)
    import scipy.sparse as sp
    J = sp.csr_matrix((val, ci, rp), shape=(case.ndim, case.ndim)).tocsc()
    land = np.repeat(case.landm.reshape(14, 18, 18)[1:-1, 1:-1, 1:-1].reshape(-1) == LAND, 6)
    for q, c in enumerate(cols):
        jc = J[:, c].toarray().ravel()
        diff = np.abs(jc - fd[:, q])[~land]
        scale = max(np.abs(jc).max(), 1.0)
        assert diff.max() < 2e-5 * scale, (c, diff.max(), scale)


@pytest.mark.parametrize("name", ["natl8", "box16"])
def test_checkerboard_modes_span_the_null_space(name):
    """The pressure modes that Ocean::pressureProjection removes are null
    vectors of J (J s = 0 to rounding); on natl8 (dense SVD) they span the
    whole null space, and the projection is idempotent."""
    import scipy.sparse as sp
    from thcm_cases import checkerboard_modes, pressure_projection
    case, _ = load_golden(name)
    orc = OracleModel(case)
    rp, ci, val, _, _ = orc.jacobian(case.x)
    J = sp.csr_matrix((val, ci, rp), shape=(case.ndim, case.ndim))
    s = checkerboard_modes(case.cfg, orc.landm())
    assert np.linalg.norm(J @ s.T, axis=0).max() <= 1e-13 * abs(J).max()
    v = np.random.default_rng(2).normal(size=case.ndim)
    pv = pressure_projection(case.cfg, orc.landm(), v)
    assert np.abs(pressure_projection(case.cfg, orc.landm(), pv) - pv).max() <= 1e-14
    assert np.abs(J @ pv - J @ v).max() <= 1e-12 * (abs(J) @ np.abs(v)).max()
    if case.ndim <= 2000:
        _, S, Vt = np.linalg.svd(J.toarray())
        null = Vt[S <= 1e-14 * S[0]]
        assert null.shape[0] == 2
        assert np.linalg.norm(null - (null @ s.T) @ s) <= 1e-8


def test_stored_state_of_reft_ocean_is_a_steady_state():
    """The steady state the reference's own end-to-end test stores
    (test/ocean/ocean_reference.h5, kept as data in
    tests/golden/reft/ocean_reference_state.npz) is one of the restated
    reft_ocean configuration: F is 7.7e-6 in the 2-norm there with the
    convective mixing on, against the 1e-2 Newton tolerance of
    test/ocean/reft_continuation_params.xml, while with the mixing off the
    same state leaves 0.157.  This pins the configuration (grid, mask,
    forcing types, the 30 stored parameters) and shows that the reference's
    Mixing = 2 run had its mixing terms on."""
    import os
    from thcm_cases import GOLDEN, Config, config_dict
    case, exp = load_golden("reft16")
    z = np.load(os.path.join(GOLDEN, "reft", "ocean_reference_state.npz"), allow_pickle=False)
    np.testing.assert_array_equal(case.x, z["x"])
    # the u, v, T, S norms reft_ocean.C:69-88 compares
    np.testing.assert_allclose([np.linalg.norm(case.x[v::6]) for v in (0, 1, 4, 5)],
                               [0.097907, 0.022403, 3.516206, 0.035162], atol=1e-6)
    F = OracleModel(case).rhs()
    np.testing.assert_array_equal(F, exp["F"])
    assert np.linalg.norm(F) < 1e-5
    cfg0 = Config(**config_dict(case.cfg))
    cfg0.vmix = 0
    F0 = OracleModel(Case("reft16_nomix", cfg0, case.landm, case.fields, case.pars, case.x)).rhs()
    assert np.linalg.norm(F0) > 0.1


@pytest.mark.parametrize("name", ["natl8", "gateway16"])
def test_fortran_layout_fixture_is_the_jacobian_fixture(name):
    """The reference's own Fortran buffers (tests/golden/fortran) hold the
    fixture's Jacobian: scattered into the fixed graph they equal J in every
    row but the integral row (THCM.C:1082-1169 copies them there), whose
    Fortran row is the S equation; B = -F in every row but that one; coB
    equals diag(B) but at the integral row (-1, an ocean S row, which THCM
    zeroes)."""
    import os
    from thcm_cases import GOLDEN
    case, exp = load_golden(name)
    z = np.load(os.path.join(GOLDEN, "fortran", name + ".npz"), allow_pickle=False)
    nd = case.ndim
    ric = int(exp["rowintcon"])
    beg, jco, co = z["begA"] - 1, z["jcoA"] - 1, z["coA"]
    rows = np.repeat(np.arange(nd), np.diff(beg))
    J = sp.csr_matrix((co, (rows, jco)), shape=(nd, nd))
    G = sp.csr_matrix((exp["J_val"], exp["J_col"], exp["J_rowptr"]), shape=(nd, nd))
    keep = np.arange(nd) != ric
    D = (J - G).tocsr()[keep]
    assert D.nnz == 0 or np.abs(D.data).max() == 0.0
    assert np.array_equal(-z["B"][keep], exp["F"][keep])
    assert np.array_equal(z["coB"][keep], exp["diagB"][keep]) and z["coB"][ric] == -1.0
    # each row's entries in stencil-position order, so not sorted by column
    assert np.all(np.abs(co) > 1e-10)
