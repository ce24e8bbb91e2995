"""Live comparison of the oracle with the reference Fortran build on fresh
seeded states (skipped where oracle/_ref is not built)."""
import numpy as np
import pytest

from thcm_cases import OracleModel, RefModel, box16_case, load_golden, rand_state

pytestmark = pytest.mark.ref


@pytest.mark.parametrize("name,seed", [("natl8", 101), ("natl8_eos", 102), ("gateway16", 103)])
def test_oracle_bit_exact_on_new_states(name, seed):
    case, _ = load_golden(name)
    case.x = rand_state(case.cfg, seed, 1.5)
    ref, orc = RefModel(case), OracleModel(case)
    np.testing.assert_array_equal(orc.rhs(), ref.rhs())
    r1, r2 = ref.jacobian(), orc.jacobian()
    for a, b in zip(r1[:4], r2[:4]):
        np.testing.assert_array_equal(a, b)


def test_oracle_bit_exact_after_parameter_changes():
    case = box16_case()
    case.pars.update(COMB=0.35, HMTP=0.0, SPL1=1500.0, BIOT=5.0, NLES=1.0, CMPR=0.2)
    ref, orc = RefModel(case), OracleModel(case)
    np.testing.assert_array_equal(orc.rhs(), ref.rhs())
    np.testing.assert_array_equal(orc.jacobian()[2], ref.jacobian()[2])


def _psim_closed_form_box(case, ref_grid_z):
    """psiM of the flat-bottom box (no land, walls on all sides) for v = 1 at
    every (i, j, k) and u = w = p = T = S = 0.  usol (usrc.F90:1014-1121)
    zeroes v on the east and north walls (i = n, j = m), so vs(j, k) =
    (n - 1) dx for 1 <= j < m and 0 for j = 0, m; compute_psim
    (thcm_utils.F90:95-118) then gives psiM(j, k) = -cos(yv_j) (n - 1) dx dz
    sum over k' <= k of dfzT(k') on the levels with z(k) hdim < -500, and 0
    on the others."""
    c = case.cfg
    d2r = np.pi / 180.0
    dx = (c.xmax_deg - c.xmin_deg) * d2r / c.n
    dy = (c.ymax_deg - c.ymin_deg) * d2r / c.m
    yv = c.ymin_deg * d2r + np.arange(c.m + 1) * dy
    z, dfzT = ref_grid_z
    dz = 1.0 / c.l
    deep = np.concatenate([[False], z * c.hdim < -500])
    col = np.concatenate([[0.0], np.cumsum(np.where(deep[1:], dfzT, 0.0))])
    psim = -np.cos(yv)[:, None] * ((c.n - 1) * dx * dz) * col[None, :]
    psim[:, ~deep] = 0.0
    psim[0, :] = 0.0
    psim[c.m, :] = 0.0
    return psim.T.reshape(-1)   # j fastest


def _z_dfzT(cfg):
    """The vertical grid of grid.F90 at the cell centres s = -1 + (k - 1/2) / l:
    z = -1 + tanh(qz (s + 1)) / tanh(qz) for qz > 1, else z = s + (1 - qz) s (1 - s),
    dfzT = dz/ds."""
    l, qz = cfg.l, cfg.qz
    s = -1.0 + (np.arange(l) + 0.5) / l
    if qz > 1.0:
        z = -1.0 + np.tanh(qz * (s + 1.0)) / np.tanh(qz)
        dfz = qz * (1.0 - np.tanh(qz * (s + 1.0)) ** 2) / np.tanh(qz)
    else:
        z = s + (1.0 - qz) * s * (1.0 - s)
        dfz = 1.0 + (1.0 - qz) * (1.0 - 2.0 * s)
    return z, dfz


@pytest.mark.parametrize("name", ["natl8", "gateway16", "global96"])
def test_reference_psim_fixture(name):
    """The fixtures' psiM equals the live reference pipeline on the fixture
    state; the pipeline is linear in the state and sees only the v
    unknowns."""
    case, exp = load_golden(name)
    ref = RefModel(case)
    mx, ps = ref.psim()
    assert mx == float(exp["psim_max"]) and np.array_equal(ps, exp["psim"])
    x = case.x.copy()
    x[np.arange(x.size) % 6 != 1] = 0.0
    assert np.array_equal(ref.psim(x)[1], ps)
    assert np.array_equal(ref.psim(4.0 * case.x)[1], 4.0 * ps)


def test_reference_psim_closed_form_on_the_box():
    """v = 1 everywhere on the 16x16x12 box: the closed form of
    _psim_closed_form_box to rounding, which pins the wall zeroing of usol
    and the depth condition of compute_psim."""
    case = box16_case()
    ref = RefModel(case)
    ones = np.zeros(case.ndim)
    ones[1::6] = 1.0
    got = ref.psim(ones)[1]
    want = _psim_closed_form_box(case, _z_dfzT(case.cfg))
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
