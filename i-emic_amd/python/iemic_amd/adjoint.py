"""Steady states and adjoint sensitivities on the GPU model.

* newton(oc, x0): Newton's method on F(x) = 0 with the device solver
  (the corrector of Continuation::newtonCorrector, Continuation.H:411-443,
  without the arclength equation).
* continuation(oc, x0, par, values): natural continuation in one parameter,
  a Newton solve per value (enough to reach a steady state from rest).
* sensitivity(oc, x, which, par): dg/dp of a steady-state diagnostic g
  (iemic_amd.DIAG_MEAN_SST, DIAG_MAX_PSIM) with respect to a continuation
  parameter.  At a steady state F(x*(p), p) = 0, so
      dg/dp = g_x . dx*/dp = -lambda . dF/dp,   J^T lambda = g_x,
  one adjoint solve for all parameters; dF/dp is a central difference of F
  in p at fixed x (the way Continuation::computeDFDPar forms it,
  Continuation.H:609-640).
"""
import numpy as np


def newton(oc, x0, tol=1e-10, step_tol=1e-13, floor_tol=1e-6, maxit=40, params=None, steps=None):
    """Damped Newton on F(x) = 0.  Converged when ||F|| <= tol * max(1,
    ||F(x0)||), when a step is below step_tol * max(1, ||x||), or at the
    rounding floor: F carries an evaluation error of about eps |J| |x|, so
    near the root ||F|| stops falling and the steps J^{-1} F stop shrinking;
    a step below floor_tol * max(1, ||x||) that is not at least half the
    previous one ends the iteration.  A step that does not reduce ||F|| is
    halved (Armijo backtracking, down to 1/64).  Returns (x, list of ||F||)
    and appends the step norms to `steps` if given; raises RuntimeError if
    it stalls above the floor or runs out of iterations."""
    x = np.array(x0, dtype=np.float64, copy=True)
    F = oc.computeRHS(x)
    f0 = np.linalg.norm(F)
    hist = [f0]
    dxs = [] if steps is None else steps
    p = params if params is not None else oc.solver_params(tol=1e-10, restart=50, maxit=200)
    for _ in range(maxit):
        if hist[-1] <= tol * max(1.0, f0):
            return x, hist
        oc.computeJacobian(x)
        dx, _ = oc.solve(F, p, raise_on_noconv=False)   # J dx = F, x -= t dx
        # J is singular in the checkerboard pressure modes; rounding noise
        # in F that J cannot reach puts large components along them into dx
        dx = oc.pressureProjection(dx)
        ndx, nx = np.linalg.norm(dx), max(1.0, np.linalg.norm(x))
        floor = ndx <= floor_tol * nx and len(dxs) > 0 and ndx > 0.5 * dxs[-1]
        dxs.append(ndx)
        if ndx <= step_tol * nx or floor:
            return x, hist
        t = 1.0
        while True:
            xt = x - t * dx
            Ft = oc.computeRHS(xt)
            ft = np.linalg.norm(Ft)
            if ft <= (1.0 - 1e-4 * t) * hist[-1] or t < 1.0 / 64:
                break
            t *= 0.5
        if ft > hist[-1]:
            if ndx <= floor_tol * nx:
                return x, hist
            raise RuntimeError(f"Newton stalled: |F| history {hist}, steps {dxs}")
        x, F = xt, Ft
        hist.append(ft)
    raise RuntimeError(f"Newton did not converge: |F| history {hist}, steps {dxs}")


def continuation(oc, x0, par, values, start=None, **newton_args):
    """Steady states along par = values, by natural continuation from the
    steady state x0 at par = start (default: the current value).  A failed
    Newton solve is retried from the last state with the step halved."""
    x = np.array(x0, dtype=np.float64, copy=True)
    p_prev = oc.getPar(par) if start is None else start
    states = []
    for v in values:
        step = v - p_prev
        while True:
            oc.setPar(par, p_prev + step)
            try:
                x, _ = newton(oc, x, **newton_args)
            except RuntimeError:
                step *= 0.5
                if abs(step) < 1e-4 * max(abs(v), 1.0):
                    raise
                continue
            p_prev += step
            if abs(p_prev - v) <= 1e-14 * max(abs(v), 1.0):
                break
            step = v - p_prev
        states.append(x.copy())
    return states


def par_step(p0, rel):
    """Difference step for a parameter of value p0: rel * |p0| (rel if 0)."""
    return rel * abs(p0) if p0 != 0.0 else rel


def dF_dp(oc, x, par, rel=1e-5):
    """Central difference of F(x; p) in the parameter par at fixed x (the
    forcing and mixing parameters enter F linearly, so the difference is
    exact up to rounding)."""
    p0 = oc.getPar(par)
    h = par_step(p0, rel)
    oc.setPar(par, p0 + h)
    Fp = oc.computeRHS(x)
    oc.setPar(par, p0 - h)
    Fm = oc.computeRHS(x)
    oc.setPar(par, p0)
    return (Fp - Fm) / (2.0 * h)


def sensitivity(oc, x, which, pars, params=None):
    """(g(x), {par: dg/dpar}, solve info, {par: sum_i |lambda_i dF_i/dpar|})
    at the steady state x, one adjoint solve.  The last item is the scale of
    the terms the sensitivity sums, the yardstick for sensitivities that
    vanish (mean SST against parameters that leave the net heat flux
    unchanged)."""
    g, gx = oc.diagnostic(which, x)
    oc.computeJacobian(x)
    p = params if params is not None else oc.solver_params(tol=1e-10, restart=50, maxit=200)
    lam, info = oc.solveTranspose(gx, p)
    sens, scale = {}, {}
    for par in pars:
        Fp = dF_dp(oc, x, par)
        sens[par] = -float(lam @ Fp)
        scale[par] = float(np.abs(lam) @ np.abs(Fp))
    return g, sens, info, scale
