// Preconditioner slot of Ocean::buildPreconditioner / applyPrecon
// (Ocean.C:1377-1416), where the reference plugs in the TRIOS block
// preconditioner with ML/MRILU inner solvers.
//
// Cell block-Jacobi: the 6x6 diagonal block of every cell (all unknowns of a
// cell) is factored with partial pivoting, tiny pivots replaced by a signed
// 1e-10 x (largest pivot) so that blocks made singular by boundary identity
// rows stay usable; the dense integral-condition row is represented by its
// diagonal entry.  One thread per cell, factors kept in HBM as 36 values + 6
// pivot indices per cell.
#include <hip/hip_runtime.h>

#include <cmath>

#include "ctx.hpp"

namespace iemic {

struct Precond {
    double *lu = nullptr;   // 36 per cell, slot-major [36][ncell]
    float *lu32 = nullptr;
    int8_t *piv = nullptr;  // 6 per cell
    bool built = false, built32 = false;
};

// diagonal-block entries (row e, column v) -> slot index (or -1)
__device__ int diag_slot(int e, int v)
{
#pragma unroll
    for (int s = 0; s < NSLOT; s++)
        if (SLOTS[s].eq == e && SLOTS[s].var == v && SLOTS[s].kk == 5) return s;
    return -1;
}

template <typename T>
__global__ void __launch_bounds__(256) k_bj_factor(Scalars s, const double *__restrict__ J,
                                                   const double *__restrict__ coef, double sign,
                                                   T *__restrict__ lu, int8_t *__restrict__ piv)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const size_t C = (size_t)s.ncell;
    double a[6][6];
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
        for (int v = 0; v < 6; v++) {
            int sl = diag_slot(e, v);
            a[e][v] = sl >= 0 ? J[(size_t)sl * C + cell] : 0.0;
        }
    if (6 * cell + S == s.rowintcon) {
        for (int v = 0; v < 6; v++) a[S][v] = 0.0;
        a[S][S] = sign * coef[6 * (size_t)cell + S];
        if (a[S][S] == 0.0) a[S][S] = 1.0;
    }
    double big = 0.0;
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
        for (int v = 0; v < 6; v++) big = fmax(big, fabs(a[e][v]));
    const double tiny = 1.0e-10 * (big > 0 ? big : 1.0);
    int p[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        int r = q;
        double best = fabs(a[q][q]);
#pragma unroll
        for (int t = q + 1; t < 6; t++)
            if (fabs(a[t][q]) > best) { best = fabs(a[t][q]); r = t; }
        p[q] = r;
        if (r != q)
#pragma unroll
            for (int v = 0; v < 6; v++) { double t = a[q][v]; a[q][v] = a[r][v]; a[r][v] = t; }
        if (fabs(a[q][q]) < tiny) a[q][q] = (a[q][q] < 0 ? -tiny : tiny);
        double inv = 1.0 / a[q][q];
#pragma unroll
        for (int t = q + 1; t < 6; t++) {
            a[t][q] *= inv;
#pragma unroll
            for (int v = q + 1; v < 6; v++) a[t][v] -= a[t][q] * a[q][v];
        }
    }
#pragma unroll
    for (int e = 0; e < 6; e++) {
        piv[6 * (size_t)cell + e] = (int8_t)p[e];
#pragma unroll
        for (int v = 0; v < 6; v++) lu[(size_t)(6 * e + v) * C + cell] = (T)a[e][v];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bj_apply(Scalars s, const T *__restrict__ lu,
                                                  const int8_t *__restrict__ piv,
                                                  const T *__restrict__ in, T *__restrict__ out)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const size_t C = (size_t)s.ncell;
    T y[6];
#pragma unroll
    for (int e = 0; e < 6; e++) y[e] = in[6 * (size_t)cell + e];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        int r = piv[6 * (size_t)cell + q];
        if (r != q) { T t = y[q]; y[q] = y[r]; y[r] = t; }
    }
#pragma unroll
    for (int e = 1; e < 6; e++)
#pragma unroll
        for (int v = 0; v < e; v++) y[e] -= lu[(size_t)(6 * e + v) * C + cell] * y[v];
#pragma unroll
    for (int e = 5; e >= 0; e--) {
#pragma unroll
        for (int v = e + 1; v < 6; v++) y[e] -= lu[(size_t)(6 * e + v) * C + cell] * y[v];
        y[e] /= lu[(size_t)(6 * e + e) * C + cell];
    }
#pragma unroll
    for (int e = 0; e < 6; e++) out[6 * (size_t)cell + e] = y[e];
}

static inline int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

void precond_free(iemic_ctx *c)
{
    Precond *p = c->prec;
    if (!p) return;
    if (p->lu) (void)hipFree(p->lu);
    if (p->lu32) (void)hipFree(p->lu32);
    if (p->piv) (void)hipFree(p->piv);
    delete p;
    c->prec = nullptr;
}

static int ensure(iemic_ctx *c)
{
    if (!c->prec) c->prec = new Precond();
    Precond *p = c->prec;
    if (!p->piv) HIPCHK(hipMalloc(&p->piv, (size_t)6 * c->sc.ncell));
    return IEMIC_OK;
}

int precond_build(iemic_ctx *c)
{
    int rc = ensure(c);
    if (rc) return rc;
    Precond *p = c->prec;
    if (!p->lu) HIPCHK(hipMalloc(&p->lu, (size_t)36 * c->sc.ncell * sizeof(double)));
    hipLaunchKernelGGL(k_bj_factor<double>, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc,
                       c->d_J, c->d_coef, (double)c->cfg.int_sign, p->lu, p->piv);
    HIPCHK(hipGetLastError());
    p->built = true;
    return IEMIC_OK;
}

int precond_build_f32(iemic_ctx *c)
{
    int rc = ensure(c);
    if (rc) return rc;
    Precond *p = c->prec;
    if (!p->lu32) HIPCHK(hipMalloc(&p->lu32, (size_t)36 * c->sc.ncell * sizeof(float)));
    hipLaunchKernelGGL(k_bj_factor<float>, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc,
                       c->d_J, c->d_coef, (double)c->cfg.int_sign, p->lu32, p->piv);
    HIPCHK(hipGetLastError());
    p->built32 = true;
    return IEMIC_OK;
}

int precond_apply(iemic_ctx *c, const double *d_in, double *d_out)
{
    if (!c->prec || !c->prec->built) {
        int rc = precond_build(c);
        if (rc) return rc;
    }
    timed_begin_ext(c, 3);
    hipLaunchKernelGGL(k_bj_apply<double>, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc,
                       c->prec->lu, c->prec->piv, d_in, d_out);
    timed_end_ext(c, 3);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

int precond_apply_f32(iemic_ctx *c, const float *d_in, float *d_out)
{
    if (!c->prec || !c->prec->built32) {
        int rc = precond_build_f32(c);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_bj_apply<float>, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc,
                       c->prec->lu32, c->prec->piv, d_in, d_out);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

}  // namespace iemic
