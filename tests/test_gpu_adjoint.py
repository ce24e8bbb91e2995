"""Adjoint path: J^T products, the transposed preconditioner, the adjoint
FGMRES, and sensitivities of steady-state diagnostics (mean SST, maximum
overturning) to forcing parameters, checked against central finite
differences of steady states and against the oracle's Jacobian.  The
reference has no transposed solve (TRIOS_BlockPreconditioner.H:122), so
these checks rest on the mathematics: dot-product identities, duality of the
forward and adjoint solves, and finite differences."""
import numpy as np
import pytest
import scipy.sparse as sp

import iemic_amd
from iemic_amd import adjoint
from gpu_helpers import gpu_ocean
from thcm_cases import Case, Config, OracleModel, config_dict, load_golden, seeded_state

pytestmark = pytest.mark.gpu


def _case(name):
    case, _ = load_golden(name)
    if name in ("gateway16", "global96"):
        case.x = seeded_state(case.cfg, case.landm, 11)
    return case


def _csr(oc):
    rp, ci, val, dB = oc.getJacobian()
    return sp.csr_matrix((val, ci, rp), shape=(oc.ndim, oc.ndim))


@pytest.mark.parametrize("name", ["natl8", "box16", "gateway16"])
def test_transposed_preconditioner_is_the_transpose(name):
    """u . M^{-1} v = (M^{-T} u) . v for random u, v."""
    case = _case(name)
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    rng = np.random.default_rng(7)
    u, v = rng.normal(size=oc.ndim), rng.normal(size=oc.ndim)
    a = u @ oc.applyPrecon(v)
    b = oc.applyPreconTranspose(u) @ v
    print(f"{name}: {a:.15e} {b:.15e}")
    assert abs(a - b) <= 1e-9 * max(abs(a), abs(b))


@pytest.mark.parametrize("name", ["natl8", "gateway16", "global96"])
def test_adjoint_solve_and_duality(name):
    """J^T lam = g to the FP64 bar, and g . x = lam . b for J x = b
    (g is the gradient of the mean SST, orthogonal to the pressure null
    vectors of J; b = J y is consistent).  The two sides differ by the
    solves' forward errors along J's smallest singular directions, which a
    residual of 1e-10 leaves determined only to residual / sigma_min: on
    global96 the gap is 1.2e-7 of g . x, the same with both solves run to
    1e-12 (where FGMRES stagnates near 1e-11), so its bound is 1e-6; natl8
    agrees to 13 digits."""
    case = _case(name)
    oc = gpu_ocean(case)
    oc.computeJacobian(case.x)
    J = _csr(oc)
    _, g = oc.diagnostic(iemic_amd.DIAG_MEAN_SST, case.x)
    p = oc.solver_params(tol=1e-10, restart=50, maxit=200)
    lam, info = oc.solveTranspose(g, p)
    rel = np.linalg.norm(g - J.T @ lam) / np.linalg.norm(g)
    b = J @ np.random.default_rng(9).normal(size=oc.ndim)
    x, _ = oc.solve(b, p)
    gx, lb = g @ x, lam @ b
    print(f"{name}: adjoint its {info.iterations} residual {rel:.2e}; g.x {gx:.12e} lam.b {lb:.12e}")
    assert rel <= 1e-8
    bound = 1e-6 if name == "global96" else 1e-7
    assert abs(gx - lb) <= bound * max(abs(gx), abs(lb))


def _smooth_natl8():
    case, _ = load_golden("natl8")
    cfg = Config(**config_dict(case.cfg))
    cfg.vmix = 0
    return Case(case.name, cfg, case.landm, case.fields, case.pars, case.x, case.intcorr)


def test_newton_reaches_the_rounding_level():
    """Newton from rest (x = 0) at COMB = 0.01 converges quadratically to
    ||F|| at the rounding level of its evaluation (a dense least-squares
    Newton on the host reaches ~4e-15 from the same start)."""
    oc = gpu_ocean(_smooth_natl8())
    oc.setPar("COMB", 0.01)
    steps = []
    x, hist = adjoint.newton(oc, np.zeros(oc.ndim), tol=1e-14, steps=steps)
    print(f"|F| {np.array(hist)}\nsteps {np.array(steps)}")
    assert hist[-1] <= 1e-13
    assert len(hist) <= 12


# wind amplitude, SST restoring amplitude, salt-flux amplitude, horizontal diffusivity
PARS = ["WIND", "TEMP", "SALT", "PE_H"]


COMB = 0.01


@pytest.fixture(scope="module")
def natl8_steady():
    """A steady state of the 8x8x4 North Atlantic (mask_natl8) at COMB = 0.01,
    reached by natural continuation from rest (x = 0 at COMB = 0).  The
    convective mixing is off (vmix = 0): the reference forms its Jacobian by
    forward differences with step 1e-8 (mix_imp.f:729-815), so with mixing
    on J is not the derivative of F to the accuracy a 1e-6 comparison with
    finite differences of steady states needs.  Without mixing this branch
    ends in a saddle-node fold near COMB = 0.022 (the overturning maximum
    grows like a square root of the distance to it), so the state is taken
    well before it, where J is well conditioned."""
    case = _smooth_natl8()
    oc = gpu_ocean(case)
    oc.setPar("COMB", 0.0)
    states = adjoint.continuation(oc, np.zeros(oc.ndim), "COMB", [0.005, COMB], start=0.0)
    xs = states[-1]
    print(f"steady state |F| {np.linalg.norm(oc.computeRHS(xs)):.2e} |x| {np.linalg.norm(xs):.3e}")
    return case, oc, xs


DIAGS = [iemic_amd.DIAG_MEAN_SST, iemic_amd.DIAG_MAX_PSIM]


@pytest.fixture(scope="module")
def fd_table(natl8_steady):
    """dg/dp by finite differences of steady states: g at x*(p0 + d h) for
    d = -1, -1/2, 1/2, 1 (Newton from the steady state at p0, h = 1% of p0),
    central differences D(h), D(h/2) and the Richardson value
    (4 D(h/2) - D(h)) / 3, whose truncation error is O(h^4).  The step is
    large because each steady state carries a rounding-level error of about
    cond(J) eps, which the difference divides by h."""
    case, oc, xs = natl8_steady
    out = {}
    for par in PARS:
        p0 = oc.getPar(par)
        h = adjoint.par_step(p0, 1e-2)
        g = {}
        for d in (-1.0, -0.5, 0.5, 1.0):
            oc.setPar(par, p0 + d * h)
            steps = []
            x, hist = adjoint.newton(oc, xs, tol=1e-14, steps=steps)
            print(f"  {par} {d * h:+.2e}: Newton |F| {np.array(hist)} steps {np.array(steps)}")
            for which in DIAGS:
                g[which, d] = oc.diagnostic(which, x)[0]
        oc.setPar(par, p0)
        for which in DIAGS:
            d1 = (g[which, 1.0] - g[which, -1.0]) / (2 * h)
            d2 = (g[which, 0.5] - g[which, -0.5]) / h
            out[which, par] = (d1, d2, (4 * d2 - d1) / 3, h)
    return out


@pytest.mark.parametrize("which", DIAGS)
def test_sensitivities_match_finite_differences(natl8_steady, fd_table, which):
    """dg/dp from one adjoint solve against the extrapolated finite
    differences of steady states, to 1e-6 relative.  Sensitivities that
    vanish (the mean SST is fixed by the zero net restoring heat flux, so
    only TEMP moves it) are held to 1e-10 of the scale of their terms, and
    the differences to the rounding of g over the step, 1e-12 |g| / h."""
    case, oc, xs = natl8_steady
    g0, sens, info, scale = adjoint.sensitivity(oc, xs, which, PARS)
    for par in PARS:
        d1, d2, fd, h = fd_table[which, par]
        print(f"diag {which} g {g0:.6e} {par}: adjoint {sens[par]:.12e} fd {fd:.12e} "
              f"(D(h) {d1:.8e}, D(h/2) {d2:.8e}; term scale {scale[par]:.3e})")
        assert abs(sens[par] - fd) <= 1e-6 * abs(fd) + 1e-10 * scale[par] + 1e-12 * abs(g0) / h


def test_sensitivity_matches_oracle_jacobian(natl8_steady):
    """The same dg/dp with the adjoint solved on the oracle's Jacobian (dense
    least squares on the host) and dF/dp from the oracle's F."""
    case, oc, xs = natl8_steady
    g0, sens, _, scale = adjoint.sensitivity(oc, xs, iemic_amd.DIAG_MEAN_SST, PARS)
    case.x = xs
    case.pars = {k: oc.getPar(k) for k in ("COMB", "WIND", "SALT", "TEMP", "PE_H")}
    orc = OracleModel(case)
    rp, ci, val, dB, _ = orc.jacobian(xs)
    J = sp.csr_matrix((val, ci, rp), shape=(oc.ndim, oc.ndim)).toarray()
    _, gx = oc.diagnostic(iemic_amd.DIAG_MEAN_SST, xs)
    lam = np.linalg.lstsq(J.T, gx, rcond=None)[0]
    for par in PARS:
        p0 = orc.par(par)
        h = adjoint.par_step(p0, 1e-5)
        orc.set_par(par, p0 + h)
        Fp = orc.rhs(xs)
        orc.set_par(par, p0 - h)
        Fm = orc.rhs(xs)
        orc.set_par(par, p0)
        ref = -lam @ ((Fp - Fm) / (2 * h))
        print(f"{par}: gpu {sens[par]:.12e} oracle {ref:.12e} (term scale {scale[par]:.3e})")
        # vanishing sensitivities: both sides are rounding, below 1e-12 |g| / |p|
        assert abs(sens[par] - ref) <= 1e-8 * abs(ref) + 1e-12 * (scale[par] + abs(g0) / abs(p0))


def test_torch_autograd_wrapper_gives_the_adjoint_gradient(natl8_steady, fd_table):
    """The steady-state diagnostic as a torch.autograd.Function of the four
    parameters (test harness only): forward = Newton to the steady state on
    the device model, backward = the one-solve adjoint.  p.grad must match
    the extrapolated finite differences of steady states."""
    torch = pytest.importorskip("torch")
    case, oc, xs = natl8_steady
    which = iemic_amd.DIAG_MAX_PSIM

    class SteadyDiagnostic(torch.autograd.Function):
        @staticmethod
        def forward(ctx, p):
            old = [oc.getPar(k) for k in PARS]
            for k, v in zip(PARS, p.tolist()):
                oc.setPar(k, v)
            x, _ = adjoint.newton(oc, xs, tol=1e-14)
            g, sens, _, _ = adjoint.sensitivity(oc, x, which, PARS)
            for k, v in zip(PARS, old):
                oc.setPar(k, v)
            ctx.save_for_backward(torch.tensor([sens[k] for k in PARS], dtype=torch.float64))
            return torch.tensor(g, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            (dg,) = ctx.saved_tensors
            return grad_out * dg

    p = torch.tensor([oc.getPar(k) for k in PARS], dtype=torch.float64, requires_grad=True)
    g = SteadyDiagnostic.apply(p)
    g.backward()
    for i, par in enumerate(PARS):
        _, _, fd, h = fd_table[which, par]
        print(f"{par}: autograd {p.grad[i].item():.12e} fd {fd:.12e}")
        assert abs(p.grad[i].item() - fd) <= 1e-6 * abs(fd) + 1e-12 * abs(g.item()) / h


@pytest.mark.parametrize("name", ["box16", "natl8", "natl8_eos", "gateway16", "global96", "reft16"])
def test_max_psim_matches_the_reference(name):
    """g1 = max psiM equals, bit for bit, what the reference's own pipeline
    (usol1D, the sum over i and compute_psim; OceanGrid::recomputePsiM,
    OceanGrid.C:264-345) gives on each fixture's state, stored in the fixture
    by tests/golden/make_golden.py.  Its gradient is exact: psiM is linear in
    v, so a step along the gradient's support that keeps the maximiser moves
    g1 by grad . d."""
    case, exp = load_golden(name)
    oc = gpu_ocean(case)
    g, grad = oc.diagnostic(iemic_amd.DIAG_MAX_PSIM, case.x)
    print(f"{name}: max psiM {g!r} (reference {float(exp['psim_max'])!r})")
    assert g == float(exp["psim_max"])
    sup = np.nonzero(grad)[0]
    assert len(sup) > 0 and np.all(sup % 6 == 1)
    d = np.zeros(oc.ndim)
    d[sup] = np.random.default_rng(3).normal(size=len(sup))
    h = 1e-6 * max(abs(g), 1e-12) / np.abs(grad @ d)
    g2, _ = oc.diagnostic(iemic_amd.DIAG_MAX_PSIM, case.x + h * d)
    assert abs((g2 - g) - h * (grad @ d)) <= 1e-9 * abs(h * (grad @ d))
    oc.close()
