// Dense LU with partial pivoting for the row-block preconditioner
// (rowlu.hip), same contract as LAPACK getrf: A = P L U in place, L unit
// lower, ipiv 1-based (row i was interchanged with row ipiv[i]), info = first
// zero pivot (1-based) or 0.
//
// Right-looking, blocked twice.  Outer blocks of NB columns update the
// trailing matrix with rocBLAS TRSM/GEMM (MFMA).  Inside an outer block,
// sub-panels of IB columns are factored by one workgroup that keeps the whole
// sub-panel in registers (1024 threads x RPT rows x IB columns): the pivot
// search is a wavefront/LDS reduction, the pivot row goes through LDS, and
// no kernel launch separates two columns.  The rest of the outer block is
// brought up to date after each sub-panel by two kernels (row interchanges,
// then the IB x w triangular solve and the rank-IB update).  This replaces
// the column-by-column panel kernels whose launch latency dominated the
// factorisation (three launches per column).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <climits>
#include <string>

#include "ctx.hpp"

namespace iemic {

constexpr int LU_THREADS = 1024;
constexpr int LU_IB = 16;       // widest sub-panel
constexpr int LU_NB = 128;      // outer block width

// Factor the m x nc sub-panel at P (leading dimension lda), nc <= IB,
// with partial pivoting.  Writes the factored rows back in pivoted order and
// ipiv[0..nc) (1-based, relative to the sub-panel's first row, plus rowoff).
// RPT x IB values per thread stay in registers: a 1024-thread workgroup has
// 128 registers per lane, so IB shrinks as the panel gets taller.
template <typename T, int RPT, int IB>
__global__ void __launch_bounds__(LU_THREADS) k_lu_panel(T *__restrict__ Pm, int lda, int m, int nc, int rowoff,
                                                         int *__restrict__ ipiv, int *__restrict__ info, int coloff)
{
    __shared__ T prow[IB];
    __shared__ T wv[LU_THREADS / 64];
    __shared__ int wi[LU_THREADS / 64];
    __shared__ int s_piv;
    __shared__ int prows[IB];
    __shared__ int trk_row[2 * IB], trk_pos[2 * IB];
    __shared__ int ntrk;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    T a[RPT][IB];
    bool used[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = t + LU_THREADS * q;
        used[q] = false;
#pragma unroll
        for (int c = 0; c < IB; c++) a[q][c] = (r < m && c < nc) ? Pm[r + (size_t)c * lda] : (T)0;
    }
    for (int c = 0; c < nc; c++) {
        // pivot search: largest magnitude, smallest row on ties (LAPACK iamax)
        T best = (T)-1;
        int bi = INT_MAX;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = t + LU_THREADS * q;
            T v = (T)-1;
#pragma unroll
            for (int cc = 0; cc < IB; cc++)
                if (cc == c) v = fabs(a[q][cc]);
            if (r < m && !used[q] && (v > best || (v == best && r < bi))) {
                best = v;
                bi = r;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            wv[wave] = best;
            wi[wave] = bi;
        }
        __syncthreads();
        if (t == 0) {
            T b = wv[0];
            int i = wi[0];
            for (int w = 1; w < LU_THREADS / 64; w++)
                if (wv[w] > b || (wv[w] == b && wi[w] < i)) {
                    b = wv[w];
                    i = wi[w];
                }
            s_piv = i;
            prows[c] = i;
        }
        __syncthreads();
        const int pr = s_piv;
        if (pr % LU_THREADS == t) {
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == pr / LU_THREADS) {
#pragma unroll
                    for (int cc = 0; cc < IB; cc++) prow[cc] = a[q][cc];
                    used[q] = true;
                }
        }
        __syncthreads();
        const T piv = prow[c];
        if (piv == (T)0) {
            if (t == 0 && *info == 0) *info = coloff + c + 1;
            continue;
        }
        const T rinv = (T)1 / piv;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = t + LU_THREADS * q;
            if (r < m && !used[q]) {
                T l = 0;
#pragma unroll
                for (int cc = 0; cc < IB; cc++)
                    if (cc == c) {
                        l = a[q][cc] * rinv;
                        a[q][cc] = l;
                    }
#pragma unroll
                for (int cc = 0; cc < IB; cc++)
                    if (cc > c) a[q][cc] -= l * prow[cc];
            }
        }
    }
    // LAPACK interchange sequence: at step c the rows at positions c and
    // p_c swap, p_c = current position of pivot row prows[c]
    if (t == 0) {
        int n = 0;
        for (int c = 0; c < nc; c++) {
            const int row = prows[c];
            int p = row;
            for (int e = 0; e < n; e++)
                if (trk_row[e] == row) p = trk_pos[e];
            ipiv[c] = rowoff + p + 1;
            int rc = c;   // row currently at position c
            for (int e = 0; e < n; e++)
                if (trk_pos[e] == c) rc = trk_row[e];
            // row -> position c, rc -> position p
            bool fr = false, fc = false;
            for (int e = 0; e < n; e++) {
                if (trk_row[e] == row) { trk_pos[e] = c; fr = true; }
                else if (trk_row[e] == rc) { trk_pos[e] = p; fc = true; }
            }
            if (!fr) { trk_row[n] = row; trk_pos[n] = c; n++; }
            if (!fc && rc != row) { trk_row[n] = rc; trk_pos[n] = p; n++; }
        }
        ntrk = n;
    }
    __syncthreads();
    const int nt = ntrk;
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = t + LU_THREADS * q;
        if (r >= m) continue;
        int fp = r;
        for (int e = 0; e < nt; e++)
            if (trk_row[e] == r) fp = trk_pos[e];
#pragma unroll
        for (int c = 0; c < IB; c++)
            if (c < nc) Pm[fp + (size_t)c * lda] = a[q][c];
    }
}

// Row interchanges ipiv[k1..k2) (1-based global rows) on columns [c0, c1),
// one thread per column.
template <typename T>
__global__ void k_lu_laswp(T *__restrict__ A, int lda, int c0, int c1, int k1, int k2, const int *__restrict__ ipiv)
{
    const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= c1) return;
    T *a = A + (size_t)col * lda;
    for (int i = k1; i < k2; i++) {
        const int p = ipiv[i] - 1;
        if (p != i) {
            const T tmp = a[i];
            a[i] = a[p];
            a[p] = tmp;
        }
    }
}

// After a sub-panel at rows/columns [k, k+ib): for the columns [c0, c1) of
// the outer block, U12 = L11^{-1} A12 on rows [k, k+ib), then
// A22 -= L21 U12 on rows [k+ib, n).  Grid: (column strips of 16, row chunks);
// every workgroup solves its strip's small triangular system itself, the
// row-chunk-0 workgroups write U12.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_update(T *__restrict__ A, int lda, int n, int k, int ib, int c0, int c1,
                                                   int rows_per_chunk)
{
    __shared__ T L11[LU_IB][LU_IB];
    __shared__ T U[LU_IB][16];
    const int cs = c0 + blockIdx.x * 16;
    const int w = min(16, c1 - cs);
    const int t = threadIdx.x;
    for (int e = t; e < ib * ib; e += blockDim.x) {
        const int i = e % ib, p = e / ib;
        L11[i][p] = A[(k + i) + (size_t)(k + p) * lda];
    }
    for (int e = t; e < ib * 16; e += blockDim.x) {
        const int i = e % ib, cc = e / ib;
        U[i][cc] = cc < w ? A[(k + i) + (size_t)(cs + cc) * lda] : (T)0;
    }
    __syncthreads();
    if (t < 16) {   // forward substitution with the unit lower L11, one column per thread
        for (int i = 1; i < ib; i++) {
            T s = U[i][t];
            for (int p = 0; p < i; p++) s -= L11[i][p] * U[p][t];
            U[i][t] = s;
        }
    }
    __syncthreads();
    if (blockIdx.y == 0)
        for (int e = t; e < ib * 16; e += blockDim.x) {
            const int i = e % ib, cc = e / ib;
            if (cc < w) A[(k + i) + (size_t)(cs + cc) * lda] = U[i][cc];
        }
    const int r0 = k + ib + blockIdx.y * rows_per_chunk, r1 = min(n, r0 + rows_per_chunk);
    for (int r = r0 + t; r < r1; r += blockDim.x) {
        T l[LU_IB];
#pragma unroll
        for (int p = 0; p < LU_IB; p++) l[p] = p < ib ? A[r + (size_t)(k + p) * lda] : (T)0;
        for (int cc = 0; cc < w; cc++) {
            T s = A[r + (size_t)(cs + cc) * lda];
#pragma unroll
            for (int p = 0; p < LU_IB; p++) s -= l[p] * U[p][cc];
            A[r + (size_t)(cs + cc) * lda] = s;
        }
    }
}

static rocblas_status trsm_llnu(rocblas_handle h, int m, int n, const double *L, int lda, double *B, int ldb)
{
    const double one = 1.0;
    return rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit, m, n,
                         &one, L, lda, B, ldb);
}
static rocblas_status trsm_llnu(rocblas_handle h, int m, int n, const float *L, int lda, float *B, int ldb)
{
    const float one = 1.0f;
    return rocblas_strsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit, m, n,
                         &one, L, lda, B, ldb);
}
static rocblas_status gemm_nn_sub(rocblas_handle h, int m, int n, int k, const double *A, int lda, const double *B,
                                  int ldb, double *C, int ldc)
{
    const double mone = -1.0, one = 1.0;
    return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &mone, A, lda, B, ldb, &one, C, ldc);
}
static rocblas_status gemm_nn_sub(rocblas_handle h, int m, int n, int k, const float *A, int lda, const float *B,
                                  int ldb, float *C, int ldc)
{
    const float mone = -1.0f, one = 1.0f;
    return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &mone, A, lda, B, ldb, &one, C, ldc);
}

template <typename T, int RPT, int IB>
static void launch_panel(hipStream_t s, T *P, int lda, int m, int nc, int rowoff, int *ipiv, int *info, int coloff)
{
    hipLaunchKernelGGL((k_lu_panel<T, RPT, IB>), dim3(1), dim3(LU_THREADS), 0, s, P, lda, m, nc, rowoff, ipiv, info,
                       coloff);
}

// Sub-panel width for an m-row panel, so that the RPT x IB values of a lane
// stay in its 128 registers (no scratch): FP64 24 values, FP32 48.
template <typename T>
static int panel_width(int m)
{
    const int rpt = (m + LU_THREADS - 1) / LU_THREADS;
    if (sizeof(T) == 8) return rpt <= 1 ? 16 : rpt <= 3 ? 8 : rpt <= 6 ? 4 : rpt <= 12 ? 2 : 1;
    return rpt <= 2 ? 16 : rpt <= 6 ? 8 : rpt <= 12 ? 4 : rpt <= 24 ? 2 : 1;
}

template <typename T>
static int panel(hipStream_t s, T *P, int lda, int m, int nc, int rowoff, int *ipiv, int *info, int coloff)
{
    const int rpt = (m + LU_THREADS - 1) / LU_THREADS;
    const int ib = panel_width<T>(m);
    if (nc > ib) return set_error(IEMIC_EINVAL, "dense LU: sub-panel wider than its register budget");
#define LU_CASE(R, W) \
    if (rpt <= R && ib == W) { launch_panel<T, R, W>(s, P, lda, m, nc, rowoff, ipiv, info, coloff); return IEMIC_OK; }
    if constexpr (sizeof(T) == 8) {
        LU_CASE(1, 16) LU_CASE(3, 8) LU_CASE(6, 4) LU_CASE(12, 2) LU_CASE(24, 1)
    } else {
        LU_CASE(2, 16) LU_CASE(6, 8) LU_CASE(12, 4) LU_CASE(24, 2) LU_CASE(24, 1)
    }
#undef LU_CASE
    return set_error(IEMIC_EUNSUPPORTED, "dense LU: panel taller than 24576 rows");
}

#define LU_BLASCHK(expr)                                                                            \
    do {                                                                                            \
        rocblas_status st_ = (expr);                                                                \
        if (st_ != rocblas_status_success)                                                          \
            return set_error(IEMIC_EHIP, std::string(#expr) + ": status " + std::to_string((int)st_)); \
    } while (0)

// A = P L U for the n x n column-major A (lda = n); ipiv and info on the device.
template <typename T>
int dense_getrf(rocblas_handle h, hipStream_t s, int n, T *A, int *ipiv, int *info)
{
    HIPCHK(hipMemsetAsync(info, 0, sizeof(int), s));
    for (int K = 0; K < n; K += LU_NB) {
        const int nb = std::min(LU_NB, n - K);
        for (int k = K, ib = 0; k < K + nb; k += ib) {
            ib = std::min(panel_width<T>(n - k), K + nb - k);
            int rc = panel<T>(s, A + k + (size_t)k * n, n, n - k, ib, k, ipiv + k, info, k);
            if (rc) return rc;
            // interchanges on the rest of the outer block, left and right of the sub-panel
            if (k > K)
                hipLaunchKernelGGL(k_lu_laswp<T>, dim3((k - K + 63) / 64), dim3(64), 0, s, A, n, K, k, k, k + ib, ipiv);
            const int r0 = k + ib, r1 = K + nb;
            if (r1 > r0) {
                hipLaunchKernelGGL(k_lu_laswp<T>, dim3((r1 - r0 + 63) / 64), dim3(64), 0, s, A, n, r0, r1, k, k + ib,
                                   ipiv);
                const int rows = n - (k + ib);
                const int chunk = 1024;
                dim3 grid((r1 - r0 + 15) / 16, std::max(1, (rows + chunk - 1) / chunk));
                hipLaunchKernelGGL(k_lu_update<T>, grid, dim3(256), 0, s, A, n, n, k, ib, r0, r1, chunk);
            }
        }
        // outer block: interchanges on the columns outside it, then U12 and the trailing update
        if (K > 0)
            hipLaunchKernelGGL(k_lu_laswp<T>, dim3((K + 63) / 64), dim3(64), 0, s, A, n, 0, K, K, K + nb, ipiv);
        const int c0 = K + nb;
        if (c0 < n) {
            hipLaunchKernelGGL(k_lu_laswp<T>, dim3((n - c0 + 63) / 64), dim3(64), 0, s, A, n, c0, n, K, K + nb, ipiv);
            LU_BLASCHK(trsm_llnu(h, nb, n - c0, A + K + (size_t)K * n, n, A + K + (size_t)c0 * n, n));
            LU_BLASCHK(gemm_nn_sub(h, n - c0, n - c0, nb, A + c0 + (size_t)K * n, n, A + K + (size_t)c0 * n, n,
                                   A + c0 + (size_t)c0 * n, n));
        }
    }
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

template int dense_getrf<double>(rocblas_handle, hipStream_t, int, double *, int *, int *);
template int dense_getrf<float>(rocblas_handle, hipStream_t, int, float *, int *, int *);

}  // namespace iemic
