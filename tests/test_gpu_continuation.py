"""Callers of the path on device state: the damped Newton iteration and the
pseudo-arclength continuation of continuation.hip (after Continuation.H).
Checks: Newton from rest to the rounding level of F, agreeing with the
host-side Newton of iemic_amd.adjoint; continuation that passes the
saddle-node fold of the 8x8x4 North Atlantic without mixing (natural
continuation in COMB stops there); continuation to a target value landing
on the steady state natural continuation finds."""
import numpy as np
import pytest

from iemic_amd import adjoint
from gpu_helpers import gpu_ocean
from thcm_cases import Case, Config, config_dict, load_golden

pytestmark = pytest.mark.gpu

COMB = 19


def _smooth(name):
    """The fixture's basin and forcing with the convective mixing off (its
    Jacobian is a forward difference in the reference, mix_imp.f:729-815)."""
    case, _ = load_golden(name)
    cfg = Config(**config_dict(case.cfg))
    cfg.vmix = 0
    return Case(case.name, cfg, case.landm, case.fields, case.pars, case.x, case.intcorr)


def test_device_newton_matches_host_newton():
    oc = gpu_ocean(_smooth("natl8"))
    oc.setPar("COMB", 0.01)
    x, its, fn = oc.newton(np.zeros(oc.ndim), tol=1e-13, step_tol=1e-14)
    print(f"device Newton: {its} iterations, |F| {fn:.2e}")
    assert fn <= 1e-13
    xh, hist = adjoint.newton(oc, np.zeros(oc.ndim), tol=1e-14)
    # the states agree up to the pressure modes the projection leaves out
    d = oc.pressureProjection(x - xh)
    print(f"host Newton: |F| {hist[-1]:.2e}; |x_dev - x_host| {np.linalg.norm(d):.2e} of |x| {np.linalg.norm(x):.3e}")
    assert np.linalg.norm(d) <= 1e-9 * np.linalg.norm(x)


def test_pseudo_arclength_passes_the_fold():
    """natl8 without mixing: natural continuation in COMB fails beyond
    about 0.0223 (a saddle-node).  Pseudo-arclength continuation follows
    the branch around it: COMB rises to a maximum there and then falls,
    and the last point is a steady state on another branch than the one
    natural continuation from rest reaches at the same COMB (the branch
    folds back again near 0.0211: two states coexist in between)."""
    oc = gpu_ocean(_smooth("natl8"))
    oc.setPar("COMB", 0.0)
    # |dx/dCOMB| ~ 200: with the state tangent scaled by 1e-2 (zeta =
    # 1e-2 / n) an arclength step of 5e-3 moves COMB by about 4e-3 away
    # from the fold
    cp = oc.continuation_params(par=COMB, target=0.05, ds=5e-3, ds_min=1e-7, ds_max=1e-2, max_steps=60,
                                newton_tol=1e-9, tan_scaling=1e-2)
    x, ph, nh, info = oc.continuation(np.zeros(oc.ndim), cp)
    print(f"steps {info.steps}, rejected {info.rejected}, Newton {info.newton_iterations}, "
          f"linear {info.linear_iterations}, {info.seconds:.1f} s")
    print("COMB", np.array2string(ph, precision=5), "\n|x|", np.array2string(nh, precision=3))
    k = int(np.argmax(ph))
    assert 0 < k < len(ph) - 1, "no turning point in the parameter history"
    assert 0.020 <= ph[k] <= 0.025
    assert ph[-1] < ph[k] - 1e-4, "the branch did not turn back"
    oc.setPar("COMB", float(ph[-1]))
    F = oc.computeRHS(x)
    print(f"final COMB {ph[-1]:.6f}, |F| {np.linalg.norm(F):.2e}")
    assert np.linalg.norm(F) <= 1e-7
    # below the fold the first branch is still there: natural continuation
    # from rest to the same COMB finds a different steady state
    oc.setPar("COMB", 0.0)
    xl = adjoint.continuation(oc, np.zeros(oc.ndim), "COMB", [0.01, 0.015, 0.02, float(ph[-1])], start=0.0)[-1]
    d = np.linalg.norm(oc.pressureProjection(x - xl))
    print(f"|x_arclength - x_lower branch| {d:.3e} of |x| {np.linalg.norm(x):.3e}")
    assert d >= 1e-2 * np.linalg.norm(x)


def test_continuation_reaches_the_target():
    """box16 without mixing, COMB 0 -> 0.05 (below its fold near 0.096):
    the run ends on the steady state at COMB = 0.05 that natural
    continuation reaches."""
    oc = gpu_ocean(_smooth("box16"))
    oc.setPar("COMB", 0.0)
    cp = oc.continuation_params(par=COMB, target=0.05, ds=1e-2, ds_max=2e-2, max_steps=30, newton_tol=1e-9)
    x, ph, nh, info = oc.continuation(np.zeros(oc.ndim), cp)
    print(f"steps {info.steps}, rejected {info.rejected}, Newton {info.newton_iterations}, reached {info.reached}, "
          f"COMB {np.array2string(ph, precision=4)}")
    assert info.reached and abs(info.par - 0.05) <= 1e-14
    oc.setPar("COMB", 0.05)
    assert np.linalg.norm(oc.computeRHS(x)) <= 1e-8
    oc.setPar("COMB", 0.0)
    xn = adjoint.continuation(oc, np.zeros(oc.ndim), "COMB", [0.01, 0.02, 0.03, 0.04, 0.05], start=0.0)[-1]
    d = oc.pressureProjection(x - xn)
    print(f"|x_arclength - x_natural| {np.linalg.norm(d):.2e} of |x| {np.linalg.norm(x):.3e}")
    assert np.linalg.norm(d) <= 1e-8 * np.linalg.norm(x)


def test_factor_reuse_gives_the_same_branch():
    """Row-LU factors kept over several Jacobians (iemic_set_precond_reuse):
    the corrector solves still reach their tolerance (stale factors that
    miss it are rebuilt), so the branch and the end state are the same as
    with a factorisation per Jacobian."""
    runs = {}
    for reuse in (0, 8):
        oc = gpu_ocean(_smooth("box16"))
        oc.setPreconditionerReuse(reuse)
        oc.setPar("COMB", 0.0)
        cp = oc.continuation_params(par=COMB, target=0.05, ds=1e-2, ds_max=2e-2, max_steps=30, newton_tol=1e-9)
        runs[reuse] = oc.continuation(np.zeros(oc.ndim), cp)
        x, ph, nh, info = runs[reuse]
        print(f"reuse {reuse}: steps {info.steps}, Newton {info.newton_iterations}, "
              f"linear {info.linear_iterations}, {info.seconds:.2f} s")
        oc.close()
    (x0, p0, _, i0), (x8, p8, _, i8) = runs[0], runs[8]
    assert i0.reached and i8.reached
    assert np.allclose(p0, p8, rtol=1e-6, atol=1e-9)
    assert np.linalg.norm(x0 - x8) <= 1e-7 * np.linalg.norm(x0)
