// Right-preconditioned flexible GMRES on the device (the solve of
// Ocean::solve, Ocean.C:1070-1147, which runs Belos FGMRES with DGKS
// orthogonalisation; the in-tree template GMRESSolver.H:80-255 has the same
// structure).  Krylov vectors live in HBM; each iteration is
//   z_j = M^{-1} v_j,  w = J z_j,
//   two passes of classical Gram-Schmidt (h = V^T w, w -= V h),
//   h_{j+1,j} = ||w||, v_{j+1} = w / h_{j+1,j},
// and the small Hessenberg least-squares problem is solved on the host with
// Givens rotations (GMRESSolver.H:257-289).  The FP64 true residual
// ||b - J x|| / ||b|| is checked at every restart and at the end.
//
// Mixed precision (precision = 1): the preconditioner is factored and
// applied in FP32 while the basis, the products with J and every residual
// stay FP64, so the accuracy of the result is that of the FP64 iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "ctx.hpp"

namespace iemic {

int precond_ready(iemic_ctx *c, int precision);
int precond_apply(iemic_ctx *c, int precision, const double *in, double *out, float *tmp32);
int precond_check(iemic_ctx *c);
int precond_apply_t(iemic_ctx *c, int precision, const double *in, double *out);
int precond_null_space(iemic_ctx *c, int precision, bool left, const double **vecs, const double **nrm2, int *count);
int precond_stale(iemic_ctx *c, int precision);
int precond_build(iemic_ctx *c, int precision);
int precond_apply_multi(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                        size_t ldout, float *tmp32);
int precond_apply_multi_t(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                          size_t ldout);

template <typename T>
struct Basis {
    T *V = nullptr, *Z = nullptr, *w = nullptr;
    int cap = 0;
};

struct Krylov {
    Basis<double> b64;
    double *d_h = nullptr;     // Gram-Schmidt coefficients, two passes
    double *d_part = nullptr;  // partial sums (blocks x (restart+2))
    int nblocks = 0;
    double *d_r = nullptr;     // FP64 residual
    double *d_bp = nullptr;    // right-hand side with its inconsistent part removed
    float *tmp32 = nullptr;    // 2 ndim: FP32 preconditioner input/output
    int hcap = 0;
    int precision = 0;         // preconditioner precision of the current solve
    bool transpose = false;    // solving J^T x = b (adjoint)
    // several right-hand sides: Krylov blocks j hold the j-th basis vector
    // of every column, so column q's basis is a strided sequence
    double *mV = nullptr, *mZ = nullptr, *mW = nullptr, *mB = nullptr, *mh = nullptr;
    int mm = 0, mq = 0;
    // block FGMRES: tall-skinny Gram partials, a Gram result and the small
    // coefficient matrices of the block updates
    double *bwork = nullptr, *bG = nullptr, *bH = nullptr;
    size_t bwork_elems = 0, bsmall_elems = 0;
};

// ------------------------------------------------------------------ kernels
// part[q * gridDim.x + block] = sum over this block's rows of V[q] . w for
// the NV vectors q = blockIdx.y * NV + (0..NV-1) below nv: one pass over the
// rows reads each basis vector once and w once per group of NV.
constexpr int NV = 8;

template <typename T>
__global__ void __launch_bounds__(256) k_multidot(const T *__restrict__ V, size_t ld, int nv,
                                                   const T *__restrict__ w, int n,
                                                   double *__restrict__ part)
{
    __shared__ double sh[4][NV];
    const int q0 = blockIdx.y * NV;
    const int cnt = min(NV, nv - q0);
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const double wr = (double)w[r];
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < cnt) acc[k] += (double)V[(size_t)(q0 + k) * ld + r] * wr;
    }
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[wave][k] = acc[k];
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
        const int k = threadIdx.x;
        part[(size_t)(q0 + k) * gridDim.x + blockIdx.x] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}

__global__ void k_sum_parts(const double *__restrict__ part, int nblocks, int nv,
                            double *__restrict__ out, int accumulate)
{
    // one block per vector: out[q] (+)= sum_b part[q*nblocks + b]
    __shared__ double sh[256];
    int q = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += part[(size_t)q * nblocks + b];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = accumulate ? out[q] + sh[0] : sh[0];
}

template <typename T>
__global__ void __launch_bounds__(256) k_update(T *__restrict__ w, const T *__restrict__ V, size_t ld,
                                                 int nv, const double *__restrict__ h, double sign, int n)
{
    // w += sign * V h
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc = 0.0;
    for (int q = 0; q < nv; q++) acc += (double)V[(size_t)q * ld + r] * h[q];
    w[r] = (T)((double)w[r] + sign * acc);
}

template <typename T>
__global__ void k_scale_copy(T *__restrict__ dst, const T *__restrict__ src, double a, int n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) dst[r] = (T)(a * (double)src[r]);
}

__global__ void k_zero(double *__restrict__ y, int n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) y[r] = 0.0;
}

__global__ void k_axpby(double *__restrict__ y, const double *__restrict__ x, double a, double b, int n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) y[r] = a * x[r] + b * y[r];
}

// ------------------------------------------------------------------ host
static inline int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

void krylov_free(iemic_ctx *c)
{
    Krylov *k = c->kry;
    if (!k) return;
    void *bufs[] = {k->b64.V, k->b64.Z, k->b64.w, k->d_h, k->d_part, k->d_r, k->d_bp, k->tmp32,
                    k->mV, k->mZ, k->mW, k->mB, k->mh, k->bwork, k->bG, k->bH};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete k;
    c->kry = nullptr;
}

template <typename T>
static int ensure_basis(Basis<T> &B, int m, int n)
{
    if (B.cap >= m) return IEMIC_OK;
    if (B.V) { (void)hipFree(B.V); (void)hipFree(B.Z); (void)hipFree(B.w); }
    HIPCHK(hipMalloc(&B.V, (size_t)(m + 1) * n * sizeof(T)));
    HIPCHK(hipMalloc(&B.Z, (size_t)m * n * sizeof(T)));
    HIPCHK(hipMalloc(&B.w, (size_t)n * sizeof(T)));
    B.cap = m;
    return IEMIC_OK;
}

static int ensure_krylov(iemic_ctx *c, int m)
{
    if (!c->kry) c->kry = new Krylov();
    Krylov *k = c->kry;
    if (k->hcap < m) {
        if (k->d_h) { (void)hipFree(k->d_h); (void)hipFree(k->d_part); }
        k->nblocks = 480;
        HIPCHK(hipMalloc(&k->d_h, (size_t)2 * (m + 2) * sizeof(double) + 64));
        HIPCHK(hipMalloc(&k->d_part, (size_t)(m + 2) * k->nblocks * sizeof(double)));
        k->hcap = m;
    }
    if (!k->d_r) HIPCHK(hipMalloc(&k->d_r, (size_t)c->sc.ndim * sizeof(double)));
    if (!k->tmp32) HIPCHK(hipMalloc(&k->tmp32, 2 * (size_t)c->sc.ndim * sizeof(float)));
    return IEMIC_OK;
}

// matrix and preconditioner of the iteration
template <typename T> struct Ops;
template <> struct Ops<double> {
    static int matvec(iemic_ctx *c, const double *x, double *y)
    {
        return c->kry->transpose ? launch_spmv_t(c, x, y) : launch_spmv(c, x, y);
    }
    static int prec(iemic_ctx *c, const double *x, double *y)
    {
        if (c->kry->transpose) return precond_apply_t(c, c->kry->precision, x, y);
        return precond_apply(c, c->kry->precision, x, y, c->kry->tmp32);
    }
};

// ||v||_2 via the multidot machinery (one vector against itself)
template <typename T>
static int norm2(iemic_ctx *c, const T *v, int n, double &out)
{
    Krylov *k = c->kry;
    hipLaunchKernelGGL(k_multidot<T>, dim3(k->nblocks, 1), dim3(256), 0, c->stream, v, (size_t)n, 1, v, n, k->d_part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, k->d_part, k->nblocks, 1, k->d_h + 2 * (k->hcap + 2), 0);
    double r;
    HIPCHK(hipMemcpyAsync(&r, k->d_h + 2 * (k->hcap + 2), sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    out = std::sqrt(r);
    return IEMIC_OK;
}

__global__ void k_add_into(double *__restrict__ a, const double *__restrict__ b, int n)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] += b[q];
}

// One FGMRES(m) solve of A x = b in precision T from x0 = 0.  The recurrence
// residual is trusted only to decide when to stop a cycle: convergence is
// declared on the true residual ||b - A x|| <= tol ||b||, computed after every
// cycle (explicit residual check of GMRESSolver.H:229-248).
template <typename T>
static int fgmres(iemic_ctx *c, const T *b, T *x, int m, double tol, int maxit, int verbose, int *its,
                  int *cycles, double *est)
{
    Krylov *k = c->kry;
    Basis<T> &B = k->b64;
    const int n = c->sc.ndim;
    int rc = ensure_basis(B, m, n);
    if (rc) return rc;
    const size_t ld = n;
    double bnorm;
    rc = norm2(c, b, n, bnorm);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(x, 0, (size_t)n * sizeof(T), c->stream));
    *its = 0;
    *cycles = 0;
    *est = 1.0;
    if (bnorm == 0.0) { *est = 0.0; return IEMIC_OK; }
    std::vector<double> H((size_t)(m + 1) * m), cs(m + 1), sn(m + 1), s(m + 1), h(m + 2), y(m + 1);
    double *d_h = k->d_h, *d_h2 = k->d_h + (m + 2);
    T *r = B.w;
    HIPCHK(hipMemcpyAsync(r, b, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    double beta = bnorm;
    int total = 0;
    const int nb = nblk(n, 256);
    while (true) {
        hipLaunchKernelGGL(k_scale_copy<T>, dim3(nb), dim3(256), 0, c->stream, B.V, r, 1.0 / beta, n);
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        int j = 0;
        for (; j < m && total < maxit; j++, total++) {
            T *vj = B.V + (size_t)j * ld, *zj = B.Z + (size_t)j * ld;
            if ((rc = Ops<T>::prec(c, vj, zj))) return rc;
            if ((rc = Ops<T>::matvec(c, zj, B.w))) return rc;
            // classical Gram-Schmidt, twice; coefficients summed on the device
            const dim3 gd(k->nblocks, (j + 1 + NV - 1) / NV);
            hipLaunchKernelGGL(k_multidot<T>, gd, dim3(256), 0, c->stream, B.V, ld, j + 1, B.w, n, k->d_part);
            hipLaunchKernelGGL(k_sum_parts, dim3(j + 1), dim3(256), 0, c->stream, k->d_part, k->nblocks, j + 1, d_h, 0);
            hipLaunchKernelGGL(k_update<T>, dim3(nb), dim3(256), 0, c->stream, B.w, B.V, ld, j + 1, d_h, -1.0, n);
            hipLaunchKernelGGL(k_multidot<T>, gd, dim3(256), 0, c->stream, B.V, ld, j + 1, B.w, n, k->d_part);
            hipLaunchKernelGGL(k_sum_parts, dim3(j + 1), dim3(256), 0, c->stream, k->d_part, k->nblocks, j + 1, d_h2, 0);
            hipLaunchKernelGGL(k_update<T>, dim3(nb), dim3(256), 0, c->stream, B.w, B.V, ld, j + 1, d_h2, -1.0, n);
            hipLaunchKernelGGL(k_add_into, dim3(1), dim3(256), 0, c->stream, d_h, d_h2, j + 1);
            hipLaunchKernelGGL(k_multidot<T>, dim3(k->nblocks, 1), dim3(256), 0, c->stream, B.w, (size_t)n, 1, B.w, n, k->d_part);
            hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, k->d_part, k->nblocks, 1, d_h + j + 1, 0);
            HIPCHK(hipMemcpyAsync(h.data(), d_h, (j + 2) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            double hn = std::sqrt(h[j + 1]);
            h[j + 1] = hn;
            T *vn = B.V + (size_t)(j + 1) * ld;
            hipLaunchKernelGGL(k_scale_copy<T>, dim3(nb), dim3(256), 0, c->stream, vn, B.w, hn > 0 ? 1.0 / hn : 0.0, n);
            // Givens rotations on column j (GMRESSolver.H:190-200)
            for (int q = 0; q < j; q++) {
                double t = cs[q] * h[q] + sn[q] * h[q + 1];
                h[q + 1] = -sn[q] * h[q] + cs[q] * h[q + 1];
                h[q] = t;
            }
            double den = std::hypot(h[j], h[j + 1]);
            cs[j] = den == 0.0 ? 1.0 : h[j] / den;
            sn[j] = den == 0.0 ? 0.0 : h[j + 1] / den;
            h[j] = cs[j] * h[j] + sn[j] * h[j + 1];
            h[j + 1] = 0.0;
            s[j + 1] = -sn[j] * s[j];
            s[j] = cs[j] * s[j];
            for (int q = 0; q <= j; q++) H[(size_t)q * m + j] = h[q];
            *est = std::fabs(s[j + 1]) / bnorm;
            if (verbose > 1 && (total % 10 == 0)) std::printf("  fgmres it %d est %.3e\n", total, *est);
            if (*est < tol || hn == 0.0) {
                j++;
                total++;
                break;
            }
        }
        // x += Z y with y = H^{-1} s (back substitution)
        for (int q = j - 1; q >= 0; q--) {
            double acc = s[q];
            for (int p = q + 1; p < j; p++) acc -= H[(size_t)q * m + p] * y[p];
            y[q] = acc / H[(size_t)q * m + q];
        }
        HIPCHK(hipMemcpyAsync(d_h, y.data(), j * sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_update<T>, dim3(nb), dim3(256), 0, c->stream, x, B.Z, ld, j, d_h, 1.0, n);
        *its = total;
        (*cycles)++;
        // explicit residual r = b - A x
        if ((rc = Ops<T>::matvec(c, x, B.w))) return rc;
        std::vector<double> one(1, 1.0);
        HIPCHK(hipMemcpyAsync(d_h2, one.data(), sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_scale_copy<T>, dim3(nb), dim3(256), 0, c->stream, B.w, B.w, -1.0, n);
        hipLaunchKernelGGL(k_update<T>, dim3(nb), dim3(256), 0, c->stream, B.w, b, ld, 1, d_h2, 1.0, n);
        r = B.w;
        if ((rc = norm2(c, r, n, beta))) return rc;
        *est = beta / bnorm;
        if (verbose) std::printf("  fgmres cycle end at %d: true residual %.3e\n", total, *est);
        if (*est < tol || total >= maxit) break;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return IEMIC_OK;
}

// FP64 true residual r = b - J x and its norm ratio
static int true_residual(iemic_ctx *c, const double *b, const double *x, double *r, double bnorm, double *rel)
{
    int rc = Ops<double>::matvec(c, x, r);
    if (rc) return rc;
    hipLaunchKernelGGL(k_axpby, dim3(nblk(c->sc.ndim, 256)), dim3(256), 0, c->stream, r, b, 1.0, -1.0, c->sc.ndim);
    double rn;
    if ((rc = norm2(c, r, c->sc.ndim, rn))) return rc;
    *rel = bnorm > 0 ? rn / bnorm : 0.0;
    return IEMIC_OK;
}

// b' = b minus its components along the null vectors of J^T (J^T x = b:
// along those of J), one per pressure basin.  J x = b is solvable only for
// b orthogonal to them; a residual F(x) carries rounding of order eps times
// the divergence terms in that direction, which near a root of F is no
// longer small relative to ||F||.  FGMRES on the unprojected b cannot pass
// that level and its iterates gather large null-space components.  The
// vectors of different basins have disjoint supports, so the projections
// are applied one after the other.
static int consistent_rhs(iemic_ctx *c, int precision, bool transpose, const double *b, const double **out)
{
    *out = b;
    const double *vecs, *nrm2;
    int cnt;
    int rc = precond_null_space(c, precision, !transpose, &vecs, &nrm2, &cnt);
    if (rc || cnt == 0) return rc;
    Krylov *k = c->kry;
    const int n = c->sc.ndim;
    if (!k->d_bp) HIPCHK(hipMalloc(&k->d_bp, (size_t)n * sizeof(double)));
    HIPCHK(hipMemcpyAsync(k->d_bp, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    for (int q = 0; q < cnt; q++) {
        if (!(nrm2[q] > 0.0)) continue;
        const double *v = vecs + (size_t)q * n;
        const double a = device_dot(c, v, k->d_bp, n) / nrm2[q];
        if (std::isnan(a)) return set_error(IEMIC_EHIP, "null-space projection: reduction failed");
        hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 256)), dim3(256), 0, c->stream, k->d_bp, v, -a, 1.0, n);
    }
    HIPCHK(hipGetLastError());
    *out = k->d_bp;
    return IEMIC_OK;
}

// J x = b, or J^T x = b when transpose is set (the adjoint solve: transposed
// products and the transposed preconditioner, same FGMRES).  FGMRES runs on
// the consistent part of b; the reported residual is that of b itself.
int krylov_solve(iemic_ctx *c, const iemic_solver_params *p, const double *d_b, double *d_x,
                 iemic_solve_info *info, bool transpose)
{
    using clk = std::chrono::steady_clock;
    if (p->restart < 1 || p->maxit < 1 || p->tol <= 0.0 || (p->precision != 0 && p->precision != 1))
        return set_error(IEMIC_EINVAL, "solver parameters out of range");
    int rc = ensure_krylov(c, p->restart);
    if (rc) return rc;
    Krylov *k = c->kry;
    k->precision = p->precision;
    k->transpose = transpose;
    iemic_solve_info inf{};
    auto t0 = clk::now();
    if ((rc = precond_ready(c, p->precision))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    auto t1 = clk::now();
    inf.t_precon_setup = std::chrono::duration<double>(t1 - t0).count();
    double bnorm;
    if ((rc = norm2(c, d_b, c->sc.ndim, bnorm))) return rc;
    const double *d_bc;
    if ((rc = consistent_rhs(c, p->precision, transpose, d_b, &d_bc))) return rc;
    int its, cycles;
    double est;
    if ((rc = fgmres<double>(c, d_bc, d_x, p->restart, p->tol, p->maxit, p->verbose, &its, &cycles, &est))) return rc;
    inf.iterations = its;
    inf.restarts = cycles;
    if ((rc = true_residual(c, d_b, d_x, k->d_r, bnorm, &inf.rel_residual))) return rc;
    // factors kept from an older Jacobian (iemic_set_precond_reuse) that did
    // not reach the tolerance: refactor once and solve again
    if (inf.rel_residual > p->tol && precond_stale(c, p->precision)) {
        if ((rc = precond_build(c, p->precision))) return rc;
        if ((rc = consistent_rhs(c, p->precision, transpose, d_b, &d_bc))) return rc;
        if ((rc = fgmres<double>(c, d_bc, d_x, p->restart, p->tol, p->maxit, p->verbose, &its, &cycles, &est)))
            return rc;
        inf.iterations += its;
        inf.restarts += cycles;
        if ((rc = true_residual(c, d_b, d_x, k->d_r, bnorm, &inf.rel_residual))) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    inf.t_solve = std::chrono::duration<double>(clk::now() - t1).count();
    if ((rc = precond_check(c))) return rc;
    inf.converged = inf.rel_residual <= p->tol;
    if (info) *info = inf;
    return inf.converged ? IEMIC_OK : set_error(IEMIC_ENOCONV, "solver did not reach the tolerance");
}

// ------------------------------------------------------------------ several right-hand sides
static int ensure_multi(iemic_ctx *c, int m, int q)
{
    Krylov *k = c->kry;
    if (k->mm >= m && k->mq >= q) return IEMIC_OK;
    void *bufs[] = {k->mV, k->mZ, k->mW, k->mB, k->mh};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    k->mV = k->mZ = k->mW = k->mB = k->mh = nullptr;
    k->mm = k->mq = 0;
    const size_t n = c->sc.ndim;
    HIPCHK(hipMalloc(&k->mV, (size_t)(m + 1) * q * n * sizeof(double)));
    HIPCHK(hipMalloc(&k->mZ, (size_t)m * q * n * sizeof(double)));
    HIPCHK(hipMalloc(&k->mW, (size_t)q * n * sizeof(double)));
    HIPCHK(hipMalloc(&k->mB, (size_t)q * n * sizeof(double)));
    HIPCHK(hipMalloc(&k->mh, (size_t)q * 2 * (m + 2) * sizeof(double)));
    k->mm = m;
    k->mq = q;
    return IEMIC_OK;
}

// per-column 2-norms of a multi-vector (column r at Y + r n)
static int norms_multi(iemic_ctx *c, int q, const double *Y, std::vector<double> &out)
{
    out.resize(q);
    for (int r = 0; r < q; r++) {
        int rc = norm2(c, Y + (size_t)r * c->sc.ndim, c->sc.ndim, out[r]);
        if (rc) return rc;
    }
    return IEMIC_OK;
}

// J X = B for q right-hand sides (column r at B + r ndim), FGMRES per column
// run in lock step: every iteration applies the preconditioner to the q new
// basis vectors in one batched sweep and J to them in one SpMM, then
// orthogonalises each column against its own basis (CGS2) and updates its
// Givens least-squares problem.  A column that reaches the tolerance is
// frozen (its later basis vectors are zero); after each cycle the true
// residuals decide which columns restart.  Like krylov_solve, each column
// runs on the consistent part of its right-hand side.
int krylov_solve_multi(iemic_ctx *c, const iemic_solver_params *p, int q, const double *d_B, double *d_X,
                       iemic_solve_info *infos, bool transpose)
{
    using clk = std::chrono::steady_clock;
    if (q < 1 || p->restart < 1 || p->maxit < 1 || p->tol <= 0.0 || (p->precision != 0 && p->precision != 1))
        return set_error(IEMIC_EINVAL, "solver parameters out of range");
    const int m = p->restart;
    int rc = ensure_krylov(c, m);
    if (rc) return rc;
    if ((rc = ensure_multi(c, m, q))) return rc;
    Krylov *k = c->kry;
    k->precision = p->precision;
    k->transpose = transpose;
    const int n = c->sc.ndim;
    const size_t N = (size_t)n, ldk = (size_t)q * n;
    auto prec = [&](const double *in, double *out) {
        return transpose ? precond_apply_multi_t(c, p->precision, q, in, N, out, N)
                         : precond_apply_multi(c, p->precision, q, in, N, out, N, k->tmp32);
    };
    auto mat = [&](const double *in, double *out) {
        return transpose ? launch_spmm_t(c, q, in, N, out, N) : launch_spmm(c, q, in, N, out, N);
    };   // stride between a column's basis vectors
    const int nb = nblk(n, 256);
    auto t0 = clk::now();
    if ((rc = precond_ready(c, p->precision))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    auto t1 = clk::now();
    // consistent right-hand sides into mB; original norms
    std::vector<double> bnorm;
    if ((rc = norms_multi(c, q, d_B, bnorm))) return rc;
    for (int r = 0; r < q; r++) {
        const double *bc;
        if ((rc = consistent_rhs(c, p->precision, transpose, d_B + r * N, &bc))) return rc;
        HIPCHK(hipMemcpyAsync(k->mB + r * N, bc, N * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    std::vector<double> bcn;
    if ((rc = norms_multi(c, q, k->mB, bcn))) return rc;
    for (int r = 0; r < q; r++) hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, c->stream, d_X + r * N, n);
    // residual block: starts as the consistent right-hand sides
    double *R = k->mW;
    HIPCHK(hipMemcpyAsync(R, k->mB, ldk * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    std::vector<double> beta = bcn;
    std::vector<int> its(q, 0), cycles(q, 0);
    std::vector<char> done(q, 0);
    for (int r = 0; r < q; r++) done[r] = !(bcn[r] > 0.0);
    std::vector<double> H((size_t)q * (m + 1) * m), cs((size_t)q * (m + 1)), sn((size_t)q * (m + 1)),
        sv((size_t)q * (m + 1)), hh((size_t)q * 2 * (m + 2)), y(m + 1);
    std::vector<int> jend(q, 0);
    std::vector<char> live(q, 0);
    while (true) {
        int nlive = 0;
        for (int r = 0; r < q; r++) {
            live[r] = !done[r] && its[r] < p->maxit;
            nlive += live[r];
            jend[r] = 0;
            double *v0 = k->mV + r * N;
            if (live[r]) {
                hipLaunchKernelGGL(k_scale_copy<double>, dim3(nb), dim3(256), 0, c->stream, v0, R + r * N, 1.0 / beta[r], n);
                std::fill(sv.begin() + (size_t)r * (m + 1), sv.begin() + (size_t)(r + 1) * (m + 1), 0.0);
                sv[(size_t)r * (m + 1)] = beta[r];
            } else {
                hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, c->stream, v0, n);
            }
        }
        if (nlive == 0) break;
        int j = 0;
        for (; j < m; j++) {
            bool any = false;
            for (int r = 0; r < q; r++) any |= live[r] && jend[r] == 0;
            if (!any) break;
            // Z_j = M^{-1} V_j and W = J Z_j on all q columns
            double *Vj = k->mV + (size_t)j * ldk, *Zj = k->mZ + (size_t)j * ldk, *W = k->mW;
            if ((rc = prec(Vj, Zj))) return rc;
            if ((rc = mat(Zj, W))) return rc;
            for (int r = 0; r < q; r++) {
                if (!live[r] || jend[r]) continue;
                const double *Vr = k->mV + r * N;
                double *w = W + r * N, *dh = k->mh + (size_t)r * 2 * (m + 2), *dh2 = dh + (m + 2);
                const dim3 gd(k->nblocks, (j + 1 + NV - 1) / NV);
                hipLaunchKernelGGL(k_multidot<double>, gd, dim3(256), 0, c->stream, Vr, ldk, j + 1, w, n, k->d_part);
                hipLaunchKernelGGL(k_sum_parts, dim3(j + 1), dim3(256), 0, c->stream, k->d_part, k->nblocks, j + 1, dh, 0);
                hipLaunchKernelGGL(k_update<double>, dim3(nb), dim3(256), 0, c->stream, w, Vr, ldk, j + 1, dh, -1.0, n);
                hipLaunchKernelGGL(k_multidot<double>, gd, dim3(256), 0, c->stream, Vr, ldk, j + 1, w, n, k->d_part);
                hipLaunchKernelGGL(k_sum_parts, dim3(j + 1), dim3(256), 0, c->stream, k->d_part, k->nblocks, j + 1, dh2, 0);
                hipLaunchKernelGGL(k_update<double>, dim3(nb), dim3(256), 0, c->stream, w, Vr, ldk, j + 1, dh2, -1.0, n);
                hipLaunchKernelGGL(k_add_into, dim3(1), dim3(256), 0, c->stream, dh, dh2, j + 1);
                hipLaunchKernelGGL(k_multidot<double>, dim3(k->nblocks, 1), dim3(256), 0, c->stream, w, N, 1, w, n,
                                   k->d_part);
                hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, k->d_part, k->nblocks, 1, dh + j + 1, 0);
            }
            HIPCHK(hipMemcpyAsync(hh.data(), k->mh, (size_t)q * 2 * (m + 2) * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            for (int r = 0; r < q; r++) {
                double *vn = k->mV + (size_t)(j + 1) * ldk + r * N;
                if (!live[r] || jend[r]) {
                    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, c->stream, vn, n);
                    continue;
                }
                double *h = hh.data() + (size_t)r * 2 * (m + 2);
                double *Hr = H.data() + (size_t)r * (m + 1) * m, *csr = cs.data() + (size_t)r * (m + 1),
                       *snr = sn.data() + (size_t)r * (m + 1), *s = sv.data() + (size_t)r * (m + 1);
                const double hn = std::sqrt(h[j + 1]);
                h[j + 1] = hn;
                hipLaunchKernelGGL(k_scale_copy<double>, dim3(nb), dim3(256), 0, c->stream, vn, W + r * N,
                                   hn > 0 ? 1.0 / hn : 0.0, n);
                for (int t = 0; t < j; t++) {
                    const double a = csr[t] * h[t] + snr[t] * h[t + 1];
                    h[t + 1] = -snr[t] * h[t] + csr[t] * h[t + 1];
                    h[t] = a;
                }
                const double den = std::hypot(h[j], h[j + 1]);
                csr[j] = den == 0.0 ? 1.0 : h[j] / den;
                snr[j] = den == 0.0 ? 0.0 : h[j + 1] / den;
                h[j] = csr[j] * h[j] + snr[j] * h[j + 1];
                h[j + 1] = 0.0;
                s[j + 1] = -snr[j] * s[j];
                s[j] = csr[j] * s[j];
                for (int t = 0; t <= j; t++) Hr[(size_t)t * m + j] = h[t];
                its[r]++;
                const double est = std::fabs(s[j + 1]) / bcn[r];
                if (est < p->tol || hn == 0.0 || its[r] >= p->maxit || j + 1 == m) jend[r] = j + 1;
            }
        }
        // x_r += Z_r y_r, y_r = H_r^{-1} s_r
        for (int r = 0; r < q; r++) {
            if (!live[r]) continue;
            const int je = jend[r] ? jend[r] : j;
            const double *Hr = H.data() + (size_t)r * (m + 1) * m, *s = sv.data() + (size_t)r * (m + 1);
            for (int t = je - 1; t >= 0; t--) {
                double acc = s[t];
                for (int u = t + 1; u < je; u++) acc -= Hr[(size_t)t * m + u] * y[u];
                y[t] = acc / Hr[(size_t)t * m + t];
            }
            double *dh = k->mh + (size_t)r * 2 * (m + 2);
            HIPCHK(hipMemcpyAsync(dh, y.data(), je * sizeof(double), hipMemcpyHostToDevice, c->stream));
            hipLaunchKernelGGL(k_update<double>, dim3(nb), dim3(256), 0, c->stream, d_X + r * N, k->mZ + r * N, ldk, je,
                               dh, 1.0, n);
            cycles[r]++;
            // the host copy of y is reused for the next column: wait for this one
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        // true residuals R = B' - J X of the consistent systems
        if ((rc = mat(d_X, R))) return rc;
        std::vector<double> one(1, 1.0);
        double *d_one = k->mh;
        HIPCHK(hipMemcpyAsync(d_one, one.data(), sizeof(double), hipMemcpyHostToDevice, c->stream));
        for (int r = 0; r < q; r++) {
            hipLaunchKernelGGL(k_scale_copy<double>, dim3(nb), dim3(256), 0, c->stream, R + r * N, R + r * N, -1.0, n);
            hipLaunchKernelGGL(k_update<double>, dim3(nb), dim3(256), 0, c->stream, R + r * N, k->mB + r * N, N, 1,
                               d_one, 1.0, n);
        }
        if ((rc = norms_multi(c, q, R, beta))) return rc;
        for (int r = 0; r < q; r++)
            if (live[r] && beta[r] / bcn[r] < p->tol) done[r] = 1;
    }
    // reported residuals against the original right-hand sides
    if ((rc = mat(d_X, R))) return rc;
    for (int r = 0; r < q; r++)
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, c->stream, R + r * N, d_B + r * N, 1.0, -1.0, n);
    std::vector<double> rn;
    if ((rc = norms_multi(c, q, R, rn))) return rc;
    const double t_solve = std::chrono::duration<double>(clk::now() - t1).count();
    if ((rc = precond_check(c))) return rc;
    bool all = true;
    for (int r = 0; r < q; r++)
        if (bnorm[r] > 0 && !(rn[r] / bnorm[r] <= p->tol)) all = false;
    if (!all && precond_stale(c, p->precision)) {
        // factors from an older Jacobian missed the tolerance: refactor once
        if ((rc = precond_build(c, p->precision))) return rc;
        return krylov_solve_multi(c, p, q, d_B, d_X, infos, transpose);
    }
    all = true;
    for (int r = 0; r < q; r++) {
        iemic_solve_info inf{};
        inf.iterations = its[r];
        inf.restarts = cycles[r];
        inf.rel_residual = bnorm[r] > 0 ? rn[r] / bnorm[r] : 0.0;
        inf.t_precon_setup = std::chrono::duration<double>(t1 - t0).count();
        inf.t_solve = t_solve;
        inf.converged = inf.rel_residual <= p->tol;
        all = all && inf.converged;
        if (infos) infos[r] = inf;
    }
    return all ? IEMIC_OK : set_error(IEMIC_ENOCONV, "solver did not reach the tolerance for every right-hand side");
}

// ------------------------------------------------------------------ block FGMRES
size_t ts_gram_work(int n, int a, int b);
hipError_t ts_gram(hipStream_t st, int n, int a, int b, const double *A, size_t lda, const double *B, size_t ldb,
                   double *G, int ldg, double *work);
hipError_t ts_update(hipStream_t st, int n, int a, int b, double alpha, const double *A, size_t lda, const double *H,
                     int ldh, double beta, double *C, size_t ldc);

namespace {

// S = Q diag(w) Q^T for the symmetric p x p matrix S (column-major), by
// cyclic Jacobi rotations (Golub and Van Loan, Matrix Computations, 8.5);
// p is at most the block size, so this costs nothing next to one pass
// over the basis.
void sym_eig(int p, std::vector<double> S, std::vector<double> &w, std::vector<double> &Q)
{
    Q.assign((size_t)p * p, 0.0);
    for (int i = 0; i < p; i++) Q[i + (size_t)i * p] = 1.0;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < p; j++)
            for (int i = 0; i < p; i++) {
                const double v = S[i + (size_t)j * p] * S[i + (size_t)j * p];
                tot += v;
                if (i != j) off += v;
            }
        if (!(off > 1e-32 * tot)) break;
        for (int i = 0; i < p - 1; i++)
            for (int j = i + 1; j < p; j++) {
                const double aij = S[i + (size_t)j * p];
                if (aij == 0.0) continue;
                const double theta = (S[j + (size_t)j * p] - S[i + (size_t)i * p]) / (2.0 * aij);
                const double t = std::copysign(1.0, theta) / (std::fabs(theta) + std::hypot(theta, 1.0));
                const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < p; k++) {   // columns i and j
                    const double ski = S[k + (size_t)i * p], skj = S[k + (size_t)j * p];
                    S[k + (size_t)i * p] = cs * ski - sn * skj;
                    S[k + (size_t)j * p] = sn * ski + cs * skj;
                }
                for (int k = 0; k < p; k++) {   // rows i and j
                    const double sik = S[i + (size_t)k * p], sjk = S[j + (size_t)k * p];
                    S[i + (size_t)k * p] = cs * sik - sn * sjk;
                    S[j + (size_t)k * p] = sn * sik + cs * sjk;
                }
                for (int k = 0; k < p; k++) {
                    const double qki = Q[k + (size_t)i * p], qkj = Q[k + (size_t)j * p];
                    Q[k + (size_t)i * p] = cs * qki - sn * qkj;
                    Q[k + (size_t)j * p] = sn * qki + cs * qkj;
                }
            }
    }
    w.resize(p);
    for (int i = 0; i < p; i++) w[i] = S[i + (size_t)i * p];
}

// G = R^T R, R upper triangular (p x p, column-major); false if G is not
// numerically positive definite
bool cholesky_upper(int p, const std::vector<double> &G, std::vector<double> &R)
{
    R.assign((size_t)p * p, 0.0);
    for (int j = 0; j < p; j++) {
        for (int i = 0; i <= j; i++) {
            double s = G[i + (size_t)j * p];
            for (int k = 0; k < i; k++) s -= R[k + (size_t)i * p] * R[k + (size_t)j * p];
            if (i == j) {
                if (!(s > 0.0)) return false;
                R[j + (size_t)j * p] = std::sqrt(s);
            } else {
                R[i + (size_t)j * p] = s / R[i + (size_t)i * p];
            }
        }
    }
    return true;
}

// inverse of the upper triangular R (p x p, column-major)
std::vector<double> upper_inverse(int p, const std::vector<double> &R)
{
    std::vector<double> X((size_t)p * p, 0.0);
    for (int j = 0; j < p; j++) {
        X[j + (size_t)j * p] = 1.0 / R[j + (size_t)j * p];
        for (int i = j - 1; i >= 0; i--) {
            double s = 0.0;
            for (int k = i + 1; k <= j; k++) s += R[i + (size_t)k * p] * X[k + (size_t)j * p];
            X[i + (size_t)j * p] = -s / R[i + (size_t)i * p];
        }
    }
    return X;
}

// Rank split of a block W (p columns) from its Gram matrix G = W^T W: the
// eigen-decomposition G = Q diag(w) Q^T, the directions with w above
// max(rel * largest, floor) kept.  Returns their number k and the p x p
// rotation H whose first k columns are Q_k diag(w_k)^{-1/2} (W H then has
// orthonormal leading columns, up to the rounding of G) and whose other
// columns are the dropped eigenvectors (W H there holds the dropped part);
// *dropped is an upper bound of the dropped singular values, including what
// the rounding of G hides (eps of the largest eigenvalue).
int rank_split(int p, const std::vector<double> &G, double rel, double floor, std::vector<double> &H, double *dropped)
{
    std::vector<double> w, Q;
    sym_eig(p, G, w, Q);
    std::vector<int> order(p);
    for (int i = 0; i < p; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return w[x] > w[y]; });
    const double wmax = std::max(0.0, w[order[0]]);
    int k = 0;
    double wd = 0.0;
    for (int i = 0; i < p; i++) {
        const double v = w[order[i]];
        if (v > rel * wmax && v > floor) k++;
        else wd = std::max(wd, v);
    }
    H.assign((size_t)p * p, 0.0);
    for (int c = 0; c < p; c++) {
        const int e = order[c];
        const double sc = c < k ? 1.0 / std::sqrt(w[e]) : 1.0;
        for (int i = 0; i < p; i++) H[i + (size_t)c * p] = Q[i + (size_t)e * p] * sc;
    }
    *dropped = k == p ? 0.0 : std::sqrt(std::max(wd, 0.0) + 1e-16 * wmax * p);
    return k;
}

}  // namespace

static int ensure_block(iemic_ctx *c, int m, int q)
{
    Krylov *k = c->kry;
    const int capV = (m + 1) * q;
    const size_t wk = ts_gram_work(c->sc.ndim, capV, q), sm = (size_t)capV * q;
    if (k->bwork_elems < wk) {
        if (k->bwork) (void)hipFree(k->bwork);
        k->bwork = nullptr;
        k->bwork_elems = 0;
        HIPCHK(hipMalloc(&k->bwork, wk * sizeof(double)));
        k->bwork_elems = wk;
    }
    if (k->bsmall_elems < sm) {
        if (k->bG) (void)hipFree(k->bG);
        if (k->bH) (void)hipFree(k->bH);
        k->bG = k->bH = nullptr;
        k->bsmall_elems = 0;
        HIPCHK(hipMalloc(&k->bG, sm * sizeof(double)));
        HIPCHK(hipMalloc(&k->bH, sm * sizeof(double)));
        k->bsmall_elems = sm;
    }
    return IEMIC_OK;
}

// J X = B (or J^T X = B) for q <= 32 right-hand sides by block FGMRES: the
// columns share one Krylov space, so a direction found for one column serves
// all of them.  A cycle starts from the block of residuals of the columns
// not yet converged; each iteration
//   Z_j = M^{-1} V_j (one batched sweep), W = J Z_j (one SpMM),
//   block classical Gram-Schmidt twice against [V_0 .. V_j] (W -= V (V^T W)),
//   then W orthonormalised by the eigen-decomposition of its Gram matrix
//   (SVQB, Stathopoulos and Wu 2002) and once more by Cholesky QR,
// every tall-skinny product on the matrix cores (block_gs.hip).  Directions
// whose singular value falls below 1e-7 of the largest (or below 1e-12 of
// the block's scale) are dropped: they already lie in the space (the
// preconditioner of this library differs from J by a few pinned rows, so
// after one iteration the new block has only a few directions left, whatever
// the number of right-hand sides).  The block upper-Hessenberg matrix is
// reduced by Householder reflections as it grows; the least-squares
// residual of each column decides when the cycle ends, the true FP64
// residuals decide which columns restart.  Each column runs on the
// consistent part of its right-hand side, as in krylov_solve.
int krylov_solve_block(iemic_ctx *c, const iemic_solver_params *p, int q, const double *d_B, double *d_X,
                       iemic_solve_info *infos, bool transpose)
{
    using clk = std::chrono::steady_clock;
    if (q < 1 || p->restart < 1 || p->maxit < 1 || p->tol <= 0.0 || (p->precision != 0 && p->precision != 1))
        return set_error(IEMIC_EINVAL, "solver parameters out of range");
    if (q > 32) return set_error(IEMIC_EINVAL, "block solve: at most 32 right-hand sides per call");
    const int m = p->restart;
    int rc = ensure_krylov(c, m);
    if (rc) return rc;
    if ((rc = ensure_multi(c, m, q))) return rc;
    if ((rc = ensure_block(c, m, q))) return rc;
    Krylov *k = c->kry;
    k->precision = p->precision;
    k->transpose = transpose;
    const int n = c->sc.ndim;
    const size_t N = (size_t)n;
    hipStream_t st = c->stream;
    auto prec = [&](int nr, const double *in, double *out) {
        return transpose ? precond_apply_multi_t(c, p->precision, nr, in, N, out, N)
                         : precond_apply_multi(c, p->precision, nr, in, N, out, N, k->tmp32);
    };
    auto mat = [&](int nr, const double *in, double *out) {
        return transpose ? launch_spmm_t(c, nr, in, N, out, N) : launch_spmm(c, nr, in, N, out, N);
    };
    // G = A^T Bm (a x b) on the host
    auto gram = [&](int a, const double *A, int b, const double *Bm, std::vector<double> &out) -> int {
        HIPCHK(ts_gram(st, n, a, b, A, N, Bm, N, k->bG, a, k->bwork));
        out.resize((size_t)a * b);
        HIPCHK(hipMemcpyAsync(out.data(), k->bG, out.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return IEMIC_OK;
    };
    // Cm = beta Cm + alpha A H (H: a x b on the host, leading dimension ldh)
    auto update = [&](int a, const double *A, int b, const double *H, int ldh, double alpha, double beta,
                      double *Cm) -> int {
        if (b <= 0) return IEMIC_OK;
        if (a > 0) HIPCHK(hipMemcpyAsync(k->bH, H, (size_t)ldh * b * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(ts_update(st, n, a, b, alpha, A, N, k->bH, ldh, beta, Cm, N));
        // the host matrix may change as soon as this returns
        HIPCHK(hipStreamSynchronize(st));
        return IEMIC_OK;
    };
    const int nb = nblk(n, 256);
    auto t0 = clk::now();
    if ((rc = precond_ready(c, p->precision))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    auto t1 = clk::now();
    // column norms of a block of at most 32 columns: the diagonal of its
    // Gram matrix (one pass)
    auto col_norms = [&](int nq, const double *P, std::vector<double> &out) -> int {
        std::vector<double> gg;
        if ((rc = gram(nq, P, nq, P, gg))) return rc;
        out.resize(nq);
        for (int i = 0; i < nq; i++) out[i] = std::sqrt(std::max(gg[i + (size_t)i * nq], 0.0));
        return IEMIC_OK;
    };
    std::vector<double> bnorm, bcn;
    if ((rc = col_norms(q, d_B, bnorm))) return rc;
    // consistent right-hand sides: the components along the null vectors
    // (mutually orthogonal: disjoint pressure basins) removed from all
    // columns at once, B' = B - N diag(1 / |n_i|^2) N^T B
    HIPCHK(hipMemcpyAsync(k->mB, d_B, (size_t)q * N * sizeof(double), hipMemcpyDeviceToDevice, st));
    {
        const double *vecs, *nrm2;
        int cnt;
        if ((rc = precond_null_space(c, p->precision, !transpose, &vecs, &nrm2, &cnt))) return rc;
        const int capA = (m + 1) * q;
        for (int i0 = 0; i0 < cnt; i0 += capA) {
            const int ci = std::min(capA, cnt - i0);
            const double *Vn = vecs + (size_t)i0 * N;
            std::vector<double> gv;
            if ((rc = gram(ci, Vn, q, k->mB, gv))) return rc;
            for (int j2 = 0; j2 < q; j2++)
                for (int i = 0; i < ci; i++) {
                    const double nn = nrm2[i0 + i];
                    gv[i + (size_t)j2 * ci] = nn > 0.0 ? gv[i + (size_t)j2 * ci] / nn : 0.0;
                }
            if ((rc = update(ci, Vn, q, gv.data(), ci, -1.0, 1.0, k->mB))) return rc;
        }
    }
    if ((rc = col_norms(q, k->mB, bcn))) return rc;
    for (int r = 0; r < q; r++) hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, st, d_X + r * N, n);
    double *R = k->mW;
    HIPCHK(hipMemcpyAsync(R, k->mB, (size_t)q * N * sizeof(double), hipMemcpyDeviceToDevice, st));
    std::vector<int> its(q, 0), cycles(q, 0);
    std::vector<char> done(q, 0);
    for (int r = 0; r < q; r++) done[r] = !(bcn[r] > 0.0);
    const int capV = (m + 1) * q, capZ = m * q;
    std::vector<double> Hq((size_t)capV * capZ), Gr, Hcol, tau, beta = bcn;
    std::vector<std::vector<double>> refl;
    double *V = k->mV, *Z = k->mZ;
    // Orthonormalise the pc columns at V + nbas N against V[0, nbas) and
    // among themselves: block CGS twice, then the rank split of what is
    // left, the kept directions first and orthonormalised once more by
    // Cholesky QR, the dropped remainder after them.  sref (< 0: measured
    // here, from the block before projection) scales the absolute floor
    // tau_fac * sref below which directions are dropped.
    auto ortho_block = [&](int nbas, int pc, double tau_fac, double &sref, int &kept, double &dropped) -> int {
        double *W = V + (size_t)nbas * N;
        const int lg = nbas + pc;
        std::vector<double> g1, g2, Gw((size_t)pc * pc), Hr, M;
        if ((rc = gram(lg, V, pc, W, g1))) return rc;
        if (sref < 0.0) {
            sref = 0.0;
            for (int i = 0; i < pc; i++) sref = std::max(sref, g1[nbas + i + (size_t)i * lg]);
            sref = std::sqrt(sref);
        }
        if (nbas > 0) {
            // W <- W - V C1 = [V W] [-C1; I] (W is the tail of [V W])
            M.assign((size_t)lg * pc, 0.0);
            for (int j2 = 0; j2 < pc; j2++) {
                for (int t = 0; t < nbas; t++) M[t + (size_t)j2 * lg] = -g1[t + (size_t)j2 * lg];
                M[nbas + j2 + (size_t)j2 * lg] = 1.0;
            }
            if ((rc = update(lg, V, pc, M.data(), lg, 1.0, 0.0, W))) return rc;
            if ((rc = gram(lg, V, pc, W, g2))) return rc;
        } else {
            g2 = g1;
        }
        for (int b2 = 0; b2 < pc; b2++)
            for (int a2 = 0; a2 < pc; a2++) {
                double sum = g2[nbas + a2 + (size_t)b2 * lg];
                for (int t = 0; t < nbas; t++) sum -= g2[t + (size_t)a2 * lg] * g2[t + (size_t)b2 * lg];
                Gw[a2 + (size_t)b2 * pc] = sum;
            }
        const double fl = tau_fac * sref;
        int kk = rank_split(pc, Gw, 1e-14, fl * fl, Hr, &dropped);
        // second projection and the rank split's rotation in one pass:
        // W <- (W - V C2) Hr = [V W] [-C2 Hr; Hr]
        M.assign((size_t)lg * pc, 0.0);
        for (int j2 = 0; j2 < pc; j2++) {
            for (int t = 0; t < nbas; t++) {
                double sum = 0.0;
                for (int u = 0; u < pc; u++) sum += g2[t + (size_t)u * lg] * Hr[u + (size_t)j2 * pc];
                M[t + (size_t)j2 * lg] = -sum;
            }
            for (int i = 0; i < pc; i++) M[nbas + i + (size_t)j2 * lg] = Hr[i + (size_t)j2 * pc];
        }
        if ((rc = update(lg, V, pc, M.data(), lg, 1.0, 0.0, W))) return rc;
        if (kk > 0) {
            // Cholesky QR of the kept columns, with their last projection:
            // W_k <- (W_k - V C3) R3^{-1} = [V W_k] [-C3 R3^{-1}; R3^{-1}]
            std::vector<double> g3, G3((size_t)kk * kk), R3;
            const int l3 = nbas + kk;
            if ((rc = gram(l3, V, kk, W, g3))) return rc;
            for (int b2 = 0; b2 < kk; b2++)
                for (int a2 = 0; a2 < kk; a2++) {
                    double sum = g3[nbas + a2 + (size_t)b2 * l3];
                    for (int t = 0; t < nbas; t++) sum -= g3[t + (size_t)a2 * l3] * g3[t + (size_t)b2 * l3];
                    G3[a2 + (size_t)b2 * kk] = sum;
                }
            if (!cholesky_upper(kk, G3, R3)) {
                kk = 0;   // lost to rounding: drop the whole block
                dropped = sref;
            } else {
                std::vector<double> Ri = upper_inverse(kk, R3);
                M.assign((size_t)l3 * kk, 0.0);
                for (int j2 = 0; j2 < kk; j2++) {
                    for (int t = 0; t < nbas; t++) {
                        double sum = 0.0;
                        for (int u = 0; u <= j2; u++) sum += g3[t + (size_t)u * l3] * Ri[u + (size_t)j2 * kk];
                        M[t + (size_t)j2 * l3] = -sum;
                    }
                    for (int i = 0; i <= j2; i++) M[nbas + i + (size_t)j2 * l3] = Ri[i + (size_t)j2 * kk];
                }
                if ((rc = update(l3, V, kk, M.data(), l3, 1.0, 0.0, W))) return rc;
            }
        }
        kept = kk;
        return IEMIC_OK;
    };
    // up to three rank splits, each resolving the previous one's dropped
    // remainder relative to its own scale, until what is dropped lies below
    // the absolute floor; returns the number of orthonormal columns added
    auto extend_basis = [&](int nbas, int pc, double tau_fac, int &added) -> int {
        double sref = -1.0;
        added = 0;
        for (int lev = 0; lev < 3 && added < pc; lev++) {
            int kk;
            double dropped;
            if ((rc = ortho_block(nbas + added, pc - added, tau_fac, sref, kk, dropped))) return rc;
            added += kk;
            if (kk == 0 || dropped <= tau_fac * sref) break;
        }
        return IEMIC_OK;
    };
    const double tau_fac = 0.1 * p->tol;
    auto ms = [](clk::time_point a, clk::time_point b) { return 1e3 * std::chrono::duration<double>(b - a).count(); };
    while (true) {
        std::vector<int> live;
        for (int r = 0; r < q; r++)
            if (!done[r] && its[r] < p->maxit) live.push_back(r);
        const int ql = (int)live.size();
        if (ql == 0) break;
        // V_0 S_0 = the residuals of the live columns: moved to the front of
        // R (live[i] >= i, so no column is overwritten before it is read),
        // scaled to unit norm into V, orthonormalised, S_0 = V_0^T R
        for (int i = 0; i < ql; i++) {
            if (live[i] != i)
                HIPCHK(hipMemcpyAsync(R + i * N, R + live[i] * N, N * sizeof(double), hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_scale_copy<double>, dim3(nb), dim3(256), 0, st, V + i * N, R + i * N,
                               1.0 / beta[live[i]], n);
        }
        int p0 = 0;
        if ((rc = extend_basis(0, ql, tau_fac, p0))) return rc;
        if (p0 == 0) break;
        std::vector<double> s0;
        if ((rc = gram(p0, V, ql, R, s0))) return rc;
        Gr.assign((size_t)capV * ql, 0.0);
        for (int cq = 0; cq < ql; cq++)
            for (int i = 0; i < p0; i++) Gr[i + (size_t)cq * capV] = s0[i + (size_t)cq * p0];
        int nV = p0, nZ = 0, off = 0, pj = p0;
        refl.clear();
        tau.clear();
        std::fill(Hq.begin(), Hq.end(), 0.0);
        for (int j = 0; j < m; j++) {
            const int pin = pj;
            double *Vj = V + (size_t)off * N, *Zj = Z + (size_t)off * N, *W = V + (size_t)nV * N;
            // Z_j = M^{-1} V_j; J Z_j kept in R (free until the restart) and
            // copied to the next block's place in V
            auto ta = clk::now();
            if ((rc = prec(pj, Vj, Zj))) return rc;
            if (p->verbose) HIPCHK(hipStreamSynchronize(st));
            auto tb = clk::now();
            if ((rc = mat(pj, Zj, R))) return rc;
            HIPCHK(hipMemcpyAsync(W, R, (size_t)pj * N * sizeof(double), hipMemcpyDeviceToDevice, st));
            if (p->verbose) HIPCHK(hipStreamSynchronize(st));
            auto tc = clk::now();
            int p1 = 0;
            if ((rc = extend_basis(nV, pj, tau_fac, p1))) return rc;
            if (nV + p1 > capV) p1 = std::max(0, capV - nV);
            // H block column: [V_0 .. V_j+1]^T J Z_j
            const int L = nV + p1;
            if ((rc = gram(L, V, pj, R, Hcol))) return rc;
            auto td = clk::now();
            // Householder QR of the grown Hessenberg matrix, applied to Gr
            for (int i = 0; i < pj; i++) {
                const int col = nZ + i;
                double *h = Hq.data() + (size_t)col * capV;
                for (int t = 0; t < L; t++) h[t] = Hcol[t + (size_t)i * L];
                for (int kq = 0; kq < col; kq++) {
                    const std::vector<double> &v = refl[kq];
                    double w = 0.0;
                    for (size_t t = 0; t < v.size(); t++) w += v[t] * h[kq + t];
                    w *= tau[kq];
                    for (size_t t = 0; t < v.size(); t++) h[kq + t] -= w * v[t];
                }
                std::vector<double> v(L - col, 0.0);
                const double alpha = h[col];
                double xn = 0.0;
                for (int t = col + 1; t < L; t++) xn = std::hypot(xn, h[t]);
                double tk = 0.0, bet = alpha;
                v[0] = 1.0;
                if (xn > 0.0) {
                    bet = -std::copysign(std::hypot(alpha, xn), alpha);
                    tk = (bet - alpha) / bet;
                    for (int t = col + 1; t < L; t++) v[t - col] = h[t] / (alpha - bet);
                }
                h[col] = bet;
                for (int t = col + 1; t < L; t++) h[t] = 0.0;
                for (int cq = 0; cq < ql; cq++) {
                    double *gcol = Gr.data() + (size_t)cq * capV;
                    double w = 0.0;
                    for (int t = col; t < L; t++) w += v[t - col] * gcol[t];
                    w *= tk;
                    for (int t = col; t < L; t++) gcol[t] -= w * v[t - col];
                }
                refl.push_back(std::move(v));
                tau.push_back(tk);
            }
            nZ += pj;
            off = nV;
            nV = L;
            pj = p1;
            bool conv = true;
            for (int cq = 0; cq < ql; cq++) {
                its[live[cq]]++;
                double e = 0.0;
                for (int t = nZ; t < nV; t++) e = std::hypot(e, Gr[t + (size_t)cq * capV]);
                if (!(e < p->tol * bcn[live[cq]])) conv = false;
            }
            if (p->verbose)
                std::fprintf(stderr,
                             "block FGMRES: iteration %d, block %d -> %d directions, basis %d; ms: preconditioner "
                             "%.2f, J %.2f, orthonormalisation %.2f\n",
                             j + 1, pin, p1, nV, ms(ta, tb), ms(tb, tc), ms(tc, td));
            if (conv || p1 == 0 || its[live[0]] >= p->maxit || nZ + p1 > capZ) break;
        }
        // Y = R^{-1} Gr (least squares), X += Z Y
        std::vector<double> Y((size_t)nZ * q, 0.0);
        double dmax = 0.0;
        for (int i = 0; i < nZ; i++) dmax = std::max(dmax, std::fabs(Hq[i + (size_t)i * capV]));
        for (int cq = 0; cq < ql; cq++) {
            double *y = Y.data() + (size_t)live[cq] * nZ;
            for (int i = nZ - 1; i >= 0; i--) {
                double s = Gr[i + (size_t)cq * capV];
                for (int u = i + 1; u < nZ; u++) s -= Hq[i + (size_t)u * capV] * y[u];
                const double d = Hq[i + (size_t)i * capV];
                y[i] = std::fabs(d) > 1e-14 * dmax ? s / d : 0.0;
            }
        }
        if ((rc = update(nZ, Z, q, Y.data(), nZ, 1.0, 1.0, d_X))) return rc;
        for (int r : live) cycles[r]++;
        // true residuals R = B' - J X of the consistent systems
        if ((rc = mat(q, d_X, R))) return rc;
        for (int r = 0; r < q; r++)
            hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, st, R + r * N, k->mB + r * N, 1.0, -1.0, n);
        if ((rc = col_norms(q, R, beta))) return rc;
        for (int r : live)
            if (beta[r] / bcn[r] < p->tol) done[r] = 1;
    }
    // reported residuals against the original right-hand sides
    if ((rc = mat(q, d_X, R))) return rc;
    for (int r = 0; r < q; r++)
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, st, R + r * N, d_B + r * N, 1.0, -1.0, n);
    std::vector<double> rn;
    if ((rc = col_norms(q, R, rn))) return rc;
    const double t_solve = std::chrono::duration<double>(clk::now() - t1).count();
    if ((rc = precond_check(c))) return rc;
    bool all = true;
    for (int r = 0; r < q; r++)
        if (bnorm[r] > 0 && !(rn[r] / bnorm[r] <= p->tol)) all = false;
    if (!all && precond_stale(c, p->precision)) {
        if ((rc = precond_build(c, p->precision))) return rc;
        return krylov_solve_block(c, p, q, d_B, d_X, infos, transpose);
    }
    all = true;
    for (int r = 0; r < q; r++) {
        iemic_solve_info inf{};
        inf.iterations = its[r];
        inf.restarts = cycles[r];
        inf.rel_residual = bnorm[r] > 0 ? rn[r] / bnorm[r] : 0.0;
        inf.t_precon_setup = std::chrono::duration<double>(t1 - t0).count();
        inf.t_solve = t_solve;
        inf.converged = inf.rel_residual <= p->tol;
        all = all && inf.converged;
        if (infos) infos[r] = inf;
    }
    return all ? IEMIC_OK : set_error(IEMIC_ENOCONV, "solver did not reach the tolerance for every right-hand side");
}

}  // namespace iemic
