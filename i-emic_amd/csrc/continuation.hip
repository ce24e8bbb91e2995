// Callers of the solve path on device state: a damped Newton iteration on
// F(x) = 0 and pseudo-arclength continuation in one parameter, after
// Continuation::run / step / eulerPredictor / newtonCorrector / createTangent
// / adjustStep / reset (continuation/Continuation.H:200-1075).  Every vector
// stays in HBM; the host keeps scalars only.  The corrector's two solves per
// iteration with the same J (J y = dF/dpar, J z = -F) are one two-column
// solve (krylov_solve_multi), so the preconditioner is applied to both in
// one batched sweep.
#include <hip/hip_runtime.h>

#include <cstdio>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "ctx.hpp"

namespace iemic {

int krylov_solve(iemic_ctx *c, const iemic_solver_params *p, const double *d_b, double *d_x, iemic_solve_info *info,
                 bool transpose);
int krylov_solve_multi(iemic_ctx *c, const iemic_solver_params *p, int q, const double *d_B, double *d_X,
                       iemic_solve_info *infos, bool transpose);

namespace {

__global__ void k_axpy(double *__restrict__ y, const double *__restrict__ x, double a, int n)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) y[r] += a * x[r];
}
// y = a x + b y
__global__ void k_axpby2(double *__restrict__ y, const double *__restrict__ x, double a, double b, int n)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) y[r] = a * x[r] + b * y[r];
}
// y = (x1 - x0) s
__global__ void k_diff_scale(double *__restrict__ y, const double *__restrict__ x1, const double *__restrict__ x0,
                             double s, int n)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) y[r] = (x1[r] - x0[r]) * s;
}
// block maxima of |x| (one per block); reduced on the host
__global__ void __launch_bounds__(256) k_absmax(const double *__restrict__ x, int n, double *__restrict__ part)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) m = fmax(m, fabs(x[r]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

struct Work {
    iemic_ctx *c;
    int n;
    dim3 g, b{256};
    explicit Work(iemic_ctx *cc) : c(cc), n(cc->sc.ndim), g((cc->sc.ndim + 255) / 256) {}
    double dot(const double *a, const double *bb) { return device_dot(c, a, bb, n); }
    double norm(const double *a) { return std::sqrt(dot(a, a)); }
    int axpy(double *y, const double *x, double a)
    {
        hipLaunchKernelGGL(k_axpy, g, b, 0, c->stream, y, x, a, n);
        HIPCHK(hipGetLastError());
        return IEMIC_OK;
    }
    int axpby(double *y, const double *x, double a, double bb)
    {
        hipLaunchKernelGGL(k_axpby2, g, b, 0, c->stream, y, x, a, bb, n);
        HIPCHK(hipGetLastError());
        return IEMIC_OK;
    }
    int copy(double *y, const double *x)
    {
        HIPCHK(hipMemcpyAsync(y, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        return IEMIC_OK;
    }
    double absmax(const double *x, double *part)
    {
        const int nb = std::min(1024, (n + 255) / 256);
        hipLaunchKernelGGL(k_absmax, dim3(nb), b, 0, c->stream, x, n, part);
        std::vector<double> h(nb);
        if (hipMemcpyAsync(h.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            return NAN;
        double m = 0.0;
        for (double v : h) m = std::max(m, v);
        return m;
    }
};

// device buffers of one run, freed on every exit
struct Buffers {
    std::vector<double *> all;
    ~Buffers()
    {
        for (double *p : all) (void)hipFree(p);
    }
    double *get(iemic_ctx *c, size_t n)
    {
        double *p = nullptr;
        if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) return nullptr;
        all.push_back(p);
        (void)c;
        return p;
    }
};

int set_par(iemic_ctx *c, int idx, double v)
{
    c->h.par[idx] = v;
    host_parameters_changed(c);
    return upload_tables(c);
}

// dF/dpar = (F(x, p + eps) - F(x, p)) / eps into d_dF; F(x, p) into d_F
// (Continuation::computeDFDPar, Continuation.H:413-450)
int dfdpar(iemic_ctx *c, Work &w, int idx, double p, double eps, const double *d_x, double *d_F, double *d_dF)
{
    int rc;
    if ((rc = launch_rhs(c, d_x, d_F))) return rc;
    if ((rc = set_par(c, idx, p + eps))) return rc;
    if ((rc = launch_rhs(c, d_x, d_dF))) return rc;
    if ((rc = set_par(c, idx, p))) return rc;
    return w.axpby(d_dF, d_F, -1.0 / eps, 1.0 / eps);
}

}  // namespace

// Damped Newton: x -= t dx with J dx = F(x), the step's pressure-mode part
// removed (Ocean::pressureProjection), t halved from 1 while ||F|| does not
// fall (down to 1/64).  Stops when ||F|| <= tol or a step's infinity norm
// <= step_tol.  On return x holds the last iterate.
int newton_dev(iemic_ctx *c, const iemic_solver_params *sp, double tol, double step_tol, int maxit, double *d_x,
               int *iterations, double *fnorm, int *linear)
{
    Work w(c);
    Buffers B;
    double *F = B.get(c, w.n), *dx = B.get(c, w.n), *xt = B.get(c, w.n), *part = B.get(c, 1024);
    if (!F || !dx || !xt || !part) return set_error(IEMIC_EHIP, "newton: device allocation failed");
    int rc, its = 0, lin = 0;
    bool converged = false;
    if ((rc = launch_rhs(c, d_x, F))) return rc;
    double f = w.norm(F);
    converged = f <= tol;
    while (its < maxit && !converged) {
        if ((rc = launch_jacobian(c, d_x))) return rc;
        iemic_solve_info info{};
        rc = krylov_solve(c, sp, F, dx, &info, false);
        if (rc && rc != IEMIC_ENOCONV) return rc;
        lin += info.iterations;
        if ((rc = launch_pressure_projection(c, dx))) return rc;
        const double step = w.absmax(dx, part);
        double t = 1.0, ft = 0.0;
        while (true) {
            if ((rc = w.copy(xt, d_x)) || (rc = w.axpy(xt, dx, -t))) return rc;
            if ((rc = launch_rhs(c, xt, F))) return rc;
            ft = w.norm(F);
            if (ft <= (1.0 - 1e-4 * t) * f || t < 1.0 / 64) break;
            t *= 0.5;
        }
        its++;
        if (ft > f) {
            // no decrease even for small steps: keep x; converged only if
            // the step was already at the rounding level
            converged = step <= step_tol;
            break;
        }
        if ((rc = w.copy(d_x, xt))) return rc;
        f = ft;
        converged = f <= tol || t * step <= step_tol;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (iterations) *iterations = its;
    if (fnorm) *fnorm = f;
    if (linear) *linear = lin;
    return converged ? IEMIC_OK : set_error(IEMIC_ENOCONV, "newton: no convergence");
}

// Pseudo-arclength continuation (Keller), the reference's defaults: the
// quadratic normalisation zeta ||x - x0||^2 + (p - p0)^2 = ds^2 with
// zeta = tan_scaling / n ('N'), an Euler initial tangent and secant tangents
// after it ('S'), the corrector residual max(|dp|, ||dx||_inf) ('D'), step
// control ds *= clamp(opt_newton / iterations, 1/2, 2) (adjustStep), and a
// halved step after a failed one (reset).  The run stops when the parameter
// passes the target; a Newton iteration at p = target then puts the state on
// it.
int continuation_dev(iemic_ctx *c, const iemic_solver_params *sp, const iemic_continuation_params *cp, double *d_x,
                     double *par_hist, double *norm_hist, iemic_continuation_info *out)
{
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    if (cp->par < 1 || cp->par > 30 || cp->ds == 0.0 || cp->max_steps < 1 || cp->newton_maxit < 1 ||
        cp->eps == 0.0 || cp->ds_min <= 0.0 || cp->ds_max < cp->ds_min)
        return set_error(IEMIC_EINVAL, "continuation parameters out of range");
    Work w(c);
    const int n = w.n, idx = cp->par;
    Buffers B;
    double *x0 = B.get(c, n), *xdot = B.get(c, n), *FB = B.get(c, 2 * (size_t)n), *YZ = B.get(c, 2 * (size_t)n),
           *F = B.get(c, n), *diff = B.get(c, n), *part = B.get(c, 1024);
    if (!x0 || !xdot || !FB || !YZ || !F || !diff || !part)
        return set_error(IEMIC_EHIP, "continuation: device allocation failed");
    double *dFdp = FB, *R = FB + n, *y = YZ, *z = YZ + n;
    const double zeta = cp->tan_scaling / n;
    double p = c->h.par[idx], ds = cp->ds, pdot;
    int rc, steps = 0, rejected = 0, newton_total = 0, linear_total = 0, reached = 0;
    auto record = [&](int k) {
        if (par_hist) par_hist[k] = p;
        if (norm_hist) norm_hist[k] = w.norm(d_x);
    };
    record(0);
    // initial tangent (Euler): J xdot = -dF/dp, normalised with pdot
    if ((rc = dfdpar(c, w, idx, p, cp->eps, d_x, F, dFdp))) return rc;
    if ((rc = launch_jacobian(c, d_x))) return rc;
    {
        iemic_solve_info info{};
        rc = krylov_solve(c, sp, dFdp, xdot, &info, false);
        if (rc && rc != IEMIC_ENOCONV) return rc;
        linear_total += info.iterations;
        if ((rc = launch_pressure_projection(c, xdot))) return rc;
        if ((rc = w.axpby(xdot, xdot, -1.0, 0.0))) return rc;
        const double nx = w.norm(xdot);
        pdot = 1.0 / std::sqrt(zeta * nx * nx + 1.0);   // normalize 'N'
        if ((rc = w.axpby(xdot, xdot, pdot, 0.0))) return rc;
    }
    const double sgn0 = (cp->target >= p) ? 1.0 : -1.0;
    while (steps < cp->max_steps) {
        // store (Continuation::store)
        if ((rc = w.copy(x0, d_x))) return rc;
        const double p0 = p;
        // Euler predictor
        if ((rc = w.axpy(d_x, xdot, ds))) return rc;
        p = p0 + ds * pdot;
        if ((rc = set_par(c, idx, p))) return rc;
        if ((rc = launch_rhs(c, d_x, F))) return rc;
        bool ok = w.norm(F) <= cp->predictor_bound;
        int it = 0;
        double res = INFINITY;
        while (ok && it < cp->newton_maxit) {
            // corrector: R = -F, dF/dp, rbp, then [J y, J z] = [dF/dp, R]
            if ((rc = dfdpar(c, w, idx, p, cp->eps, d_x, F, dFdp))) return rc;
            if ((rc = w.axpby(R, F, -1.0, 0.0))) return rc;
            hipLaunchKernelGGL(k_diff_scale, w.g, w.b, 0, c->stream, diff, d_x, x0, 1.0, n);
            const double pd = p - p0;
            const double rbp = ds * ds - zeta * w.dot(diff, diff) - pd * pd;
            if ((rc = launch_jacobian(c, d_x))) return rc;
            iemic_solve_info infos[2];
            rc = krylov_solve_multi(c, sp, 2, FB, YZ, infos, false);
            if (rc && rc != IEMIC_ENOCONV) return rc;
            linear_total += infos[0].iterations + infos[1].iterations;
            if ((rc = launch_pressure_projection(c, y)) || (rc = launch_pressure_projection(c, z))) return rc;
            const double dp = (rbp - 2.0 * zeta * w.dot(diff, z)) / (2.0 * pd - 2.0 * zeta * w.dot(diff, y));
            if (!std::isfinite(dp)) {
                ok = false;
                break;
            }
            if ((rc = w.axpy(z, y, -dp))) return rc;   // dx = z - dp y
            if ((rc = w.axpy(d_x, z, 1.0))) return rc;
            p += dp;
            if ((rc = set_par(c, idx, p))) return rc;
            it++;
            if ((rc = launch_rhs(c, d_x, F))) return rc;
            if (!(w.norm(F) <= cp->predictor_bound)) {
                ok = false;
                break;
            }
            res = std::max(std::fabs(dp), w.absmax(z, part));
            if (res < cp->newton_tol) break;
        }
        newton_total += it;
        ok = ok && res < cp->newton_tol;
        if (!ok) {
            // reset: back to the stored point, half the step
            rejected++;
            if ((rc = w.copy(d_x, x0))) return rc;
            p = p0;
            if ((rc = set_par(c, idx, p))) return rc;
            if (std::fabs(ds) <= cp->ds_min) {
                HIPCHK(hipStreamSynchronize(c->stream));
                return set_error(IEMIC_ENOCONV, "continuation: a step at the minimum step size failed");
            }
            ds = (ds > 0 ? 1.0 : -1.0) * std::max(std::fabs(ds) / 2.0, cp->ds_min);
            continue;
        }
        steps++;
        record(steps);
        if (sp->verbose)
            std::fprintf(stderr, "continuation: step %d, parameter %.8g, |x| %.6g, ds %.4g, %d Newton iterations\n",
                         steps, p, w.norm(d_x), ds, it);
        // secant tangent: (x - x0, p - p0) / ds
        hipLaunchKernelGGL(k_diff_scale, w.g, w.b, 0, c->stream, xdot, d_x, x0, 1.0 / ds, n);
        pdot = (p - p0) / ds;
        // passed the target: Newton at p = target
        if ((p - cp->target) * sgn0 >= 0.0) {
            if ((rc = set_par(c, idx, cp->target))) return rc;
            int nit = 0, lin = 0;
            double fn = 0.0;
            rc = newton_dev(c, sp, cp->newton_tol * 1e-4, cp->newton_tol * 1e-2, 2 * cp->newton_maxit, d_x, &nit, &fn,
                            &lin);
            if (rc && rc != IEMIC_ENOCONV) return rc;
            newton_total += nit;
            linear_total += lin;
            p = cp->target;
            reached = rc == IEMIC_OK;
            if (par_hist) par_hist[steps] = p;
            if (norm_hist) norm_hist[steps] = w.norm(d_x);
            break;
        }
        // step control (adjustStep)
        double fac = cp->opt_newton / std::max(it, 1);
        fac = std::min(2.0, std::max(0.5, fac));
        ds *= fac;
        ds = (ds > 0 ? 1.0 : -1.0) * std::min(std::max(std::fabs(ds), cp->ds_min), cp->ds_max);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (out) {
        out->steps = steps;
        out->rejected = rejected;
        out->newton_iterations = newton_total;
        out->linear_iterations = linear_total;
        out->par = p;
        out->reached = reached;
        out->seconds = std::chrono::duration<double>(clk::now() - t0).count();
    }
    return IEMIC_OK;
}

}  // namespace iemic
