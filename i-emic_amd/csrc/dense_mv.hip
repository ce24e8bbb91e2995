// Dense products of the row-LU apply: Y = alpha A X + beta Y with A column-
// major (m x k, leading dimension lda) and X, Y holding q <= 32 columns.
// The sweep applies one explicit block inverse or coupling block per zonal
// row, so the whole apply streams the factors (81 GB at 180x80x24 in FP64)
// once; these products are HBM-bound for q = 1 and stay close to it for a
// few tens of right-hand sides.
//
// Layout of the work: blocks of 256 rows (one row per thread) times K
// slices.  Every thread walks the columns of its slice reading A[c lda + r]
// (consecutive rows of a wavefront read one contiguous 512-byte line per
// column), with the slice of X staged in LDS 64 columns at a time and read
// as a broadcast.  Each K slice writes its partial sums to a workspace; a
// second kernel adds the slices in a fixed order and applies alpha and
// beta, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace iemic {

constexpr int MM_ROWS = 256;   // rows per block, one per thread
constexpr int MM_KC = 64;      // columns of X staged in LDS per step
constexpr int MM_BLOCKS = 2048;  // default blocks aimed at per product (8 per CU); tuned by the caller

template <typename T, int NC>
__global__ void __launch_bounds__(MM_ROWS) k_mm_partial(int m, int k, int q, int kslice, const T *__restrict__ A,
                                                        size_t lda, const T *__restrict__ X, size_t ldx,
                                                        T *__restrict__ P)
{
    __shared__ T xs[MM_KC][NC];
    const int r = blockIdx.x * MM_ROWS + threadIdx.x;
    const int s = blockIdx.y;
    const int c0 = s * kslice, c1 = min(k, c0 + kslice);
    T acc[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) acc[j] = T(0);
    for (int cb = c0; cb < c1; cb += MM_KC) {
        const int nc = min(MM_KC, c1 - cb);
        // X[cb .. cb + nc) of every column: consecutive threads read
        // consecutive rows of one column of X
        for (int e = threadIdx.x; e < MM_KC * NC; e += MM_ROWS) {
            const int j = e / MM_KC, cc = e % MM_KC;
            xs[cc][j] = (cc < nc && j < q) ? X[(size_t)j * ldx + cb + cc] : T(0);
        }
        __syncthreads();
        if (r < m) {
            const T *a = A + (size_t)cb * lda + r;
            int cc = 0;
            // unrolled so that the loads of several steps are in flight at
            // once (the same sums in the same order)
#pragma unroll 4
            for (; cc + 4 <= nc; cc += 4) {
                const T a0 = a[(size_t)cc * lda], a1 = a[(size_t)(cc + 1) * lda];
                const T a2 = a[(size_t)(cc + 2) * lda], a3 = a[(size_t)(cc + 3) * lda];
#pragma unroll
                for (int j = 0; j < NC; j++)
                    acc[j] += a0 * xs[cc][j] + a1 * xs[cc + 1][j] + a2 * xs[cc + 2][j] + a3 * xs[cc + 3][j];
            }
            for (; cc < nc; cc++) {
                const T a0 = a[(size_t)cc * lda];
#pragma unroll
                for (int j = 0; j < NC; j++) acc[j] += a0 * xs[cc][j];
            }
        }
        __syncthreads();
    }
    if (r < m)
#pragma unroll
        for (int j = 0; j < NC; j++)
            if (j < q) P[((size_t)s * NC + j) * m + r] = acc[j];
}

template <typename T, int NC>
__global__ void __launch_bounds__(256) k_mm_reduce(int m, int q, int ns, const T *__restrict__ P, T alpha, T beta,
                                                   T *__restrict__ Y, size_t ldy)
{
    const int r = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (r >= m || j >= q) return;
    T acc = T(0);
    for (int s = 0; s < ns; s++) acc += P[((size_t)s * NC + j) * m + r];
    T *y = Y + (size_t)j * ldy + r;
    *y = beta == T(0) ? alpha * acc : alpha * acc + beta * *y;
}

// FP64 matrix-core variant for 9..32 right-hand sides (CT column tiles of
// 16), v_mfma_f64_16x16x4f64: D(16x16) += A(16x4) B(4x16) with lane l
// holding A[l % 16][l / 16], B[l / 16][l % 16] and the four accumulator
// values v = 0..3 of D[l / 16 + 4 v][l % 16].  A wavefront owns 64 rows (four
// 16-row tiles) of its block's 256; each step of 4 columns loads one factor
// element per tile and lane (a column's 64 rows are one contiguous 512-byte
// line across the four loads) and takes the B tiles from the LDS slice of X.
typedef double mm_d4 __attribute__((ext_vector_type(4)));

template <int CT>
__global__ void __launch_bounds__(MM_ROWS) k_mm_partial_mfma(int m, int k, int q, int kslice,
                                                             const double *__restrict__ A, size_t lda,
                                                             const double *__restrict__ X, size_t ldx,
                                                             double *__restrict__ P)
{
    constexpr int NC = 16 * CT;
    __shared__ double xs[MM_KC][NC + 1];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int row0 = blockIdx.x * MM_ROWS + wave * 64;
    const int s = blockIdx.y;
    const int c0 = s * kslice, c1 = min(k, c0 + kslice);
    const int lr = lane % 16, lk = lane / 16;
    mm_d4 acc[4][CT];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < CT; u++) acc[t][u] = mm_d4{0.0, 0.0, 0.0, 0.0};
    for (int cb = c0; cb < c1; cb += MM_KC) {
        const int nc = min(MM_KC, c1 - cb);
        for (int e = threadIdx.x; e < MM_KC * NC; e += MM_ROWS) {
            const int j = e / MM_KC, cc = e % MM_KC;
            xs[cc][j] = (cc < nc && j < q) ? X[(size_t)j * ldx + cb + cc] : 0.0;
        }
        __syncthreads();
        // a fixed trip count (steps at or beyond nc multiply zeros), so that
        // the loop around the convergent matrix-core builtin unrolls and
        // the factor loads of later steps are issued early
#pragma unroll
        for (int kk = 0; kk < MM_KC; kk += 4) {
            const int kc = kk + lk;   // < MM_KC; rows of xs at or beyond nc are zero
            double a[4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                // an unconditional load (of A[0] where out of range) and a
                // select: loads under a branch were each waited for before
                // the step's products, one round trip per step
                const int r = row0 + 16 * t + lr;
                const bool in = kc < nc && r < m;
                const double v = A[in ? (size_t)(cb + kc) * lda + r : 0];
                a[t] = in ? v : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CT; u++) {
                const double b = xs[kc][16 * u + lr];
#pragma unroll
                for (int t = 0; t < 4; t++) acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b, acc[t][u], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < CT; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int r = row0 + 16 * t + lk + 4 * v, c = 16 * u + lr;
                if (r < m && c < q) P[((size_t)s * NC + c) * m + r] = acc[t][u][v];
            }
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

template <int CT>
static hipError_t mm_launch_mfma(hipStream_t st, int m, int k, int q, double alpha, const double *A, size_t lda,
                                 const double *X, size_t ldx, double beta, double *Y, size_t ldy, double *work,
                                 size_t work_elems, int blocks)
{
    constexpr int NC = 16 * CT;
    const int rb = cdiv(m, MM_ROWS);
    int ns = 0, kslice = 0;
    if (k > 0) {
        ns = std::max(1, std::min(cdiv(blocks > 0 ? blocks : MM_BLOCKS, rb), cdiv(k, MM_KC)));
        while (ns > 1 && (size_t)ns * NC * m > work_elems) ns--;
        if ((size_t)ns * NC * m > work_elems) return hipErrorInvalidValue;
        kslice = cdiv(cdiv(k, ns), MM_KC) * MM_KC;
        ns = cdiv(k, kslice);
        hipLaunchKernelGGL((k_mm_partial_mfma<CT>), dim3(rb, ns), dim3(MM_ROWS), 0, st, m, k, q, kslice, A, lda, X,
                           ldx, work);
    }
    hipLaunchKernelGGL((k_mm_reduce<double, NC>), dim3(cdiv(m, 256), q), dim3(256), 0, st, m, q, ns, work, alpha,
                       beta, Y, ldy);
    return hipGetLastError();
}

template <typename T, int NC>
static hipError_t mm_launch(hipStream_t st, int m, int k, int q, T alpha, const T *A, size_t lda, const T *X,
                            size_t ldx, T beta, T *Y, size_t ldy, T *work, size_t work_elems, int blocks)
{
    const int rb = cdiv(m, MM_ROWS);
    int ns = 0, kslice = 0;
    if (k > 0) {
        ns = std::max(1, std::min(cdiv(blocks > 0 ? blocks : MM_BLOCKS, rb), cdiv(k, MM_KC)));
        while (ns > 1 && (size_t)ns * NC * m > work_elems) ns--;
        if ((size_t)ns * NC * m > work_elems) return hipErrorInvalidValue;
        kslice = cdiv(cdiv(k, ns), MM_KC) * MM_KC;
        ns = cdiv(k, kslice);
        hipLaunchKernelGGL((k_mm_partial<T, NC>), dim3(rb, ns), dim3(MM_ROWS), 0, st, m, k, q, kslice, A, lda, X, ldx,
                           work);
    }
    hipLaunchKernelGGL((k_mm_reduce<T, NC>), dim3(cdiv(m, 256), q), dim3(256), 0, st, m, q, ns, work, alpha, beta, Y,
                       ldy);
    return hipGetLastError();
}

// FP64 products with 9..32 right-hand sides run on the matrix cores unless
// IEMIC_NO_MFMA is set (comparison runs)
static bool mm_use_mfma()
{
    static const bool use = std::getenv("IEMIC_NO_MFMA") == nullptr;
    return use;
}

// Y = alpha A X + beta Y; q columns (column j of X at X + j ldx); work holds
// at least dense_mm_work(m) elements; blocks: the number of blocks (row
// blocks times K slices) to aim at, 0 for the default.  Returns a HIP error
// code.
template <typename T>
hipError_t dense_mm(hipStream_t st, int m, int k, int q, T alpha, const T *A, size_t lda, const T *X, size_t ldx,
                    T beta, T *Y, size_t ldy, T *work, size_t work_elems, int blocks)
{
    if (m <= 0 || q <= 0) return hipSuccess;
    for (int j0 = 0; j0 < q; j0 += 32) {
        const int qq = std::min(32, q - j0);
        const T *Xj = X + (size_t)j0 * ldx;
        T *Yj = Y + (size_t)j0 * ldy;
        hipError_t e;
        if constexpr (std::is_same<T, double>::value) {
            if (qq > 8 && mm_use_mfma()) {
                e = qq > 16 ? mm_launch_mfma<2>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks)
                            : mm_launch_mfma<1>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
                if (e != hipSuccess) return e;
                continue;
            }
        }
        if (qq == 1) e = mm_launch<T, 1>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
        else if (qq <= 2) e = mm_launch<T, 2>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
        else if (qq <= 4) e = mm_launch<T, 4>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
        else if (qq <= 8) e = mm_launch<T, 8>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
        else if (qq <= 16) e = mm_launch<T, 16>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
        else e = mm_launch<T, 32>(st, m, k, qq, alpha, A, lda, Xj, ldx, beta, Yj, ldy, work, work_elems, blocks);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Y = alpha A^T X + beta Y with A column-major (m x k): output row c is the
// dot product of column c of A (contiguous) with each column of X.  Each
// wavefront owns CW output rows and walks their columns 64 rows at a time;
// the four wavefronts of a block share the chunk of X staged in LDS, and a
// wavefront reads each staged X value once for its CW columns (CW = 4 cuts
// the LDS reads and the X staging per output row fourfold when NC >= 4).
constexpr int MT_CHUNK = 1024;   // rows of X staged per step

template <typename T, int NC, int CW>
__global__ void __launch_bounds__(256) k_mmt(int m, int k, int q, const T *__restrict__ A, size_t lda,
                                             const T *__restrict__ X, size_t ldx, T alpha, T beta,
                                             T *__restrict__ Y, size_t ldy)
{
    __shared__ T xs[NC][MT_CHUNK];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int c0 = (blockIdx.x * 4 + wave) * CW;
    T acc[CW][NC];
#pragma unroll
    for (int u = 0; u < CW; u++)
#pragma unroll
        for (int j = 0; j < NC; j++) acc[u][j] = T(0);
    for (int r0 = 0; r0 < m; r0 += MT_CHUNK) {
        const int nr = min(MT_CHUNK, m - r0);
        for (int e = threadIdx.x; e < NC * MT_CHUNK; e += 256) {
            const int j = e / MT_CHUNK, rr = e % MT_CHUNK;
            xs[j][rr] = (rr < nr && j < q) ? X[(size_t)j * ldx + r0 + rr] : T(0);
        }
        __syncthreads();
        for (int rr = lane; rr < nr; rr += 64) {
            T av[CW];
#pragma unroll
            for (int u = 0; u < CW; u++) av[u] = c0 + u < k ? A[(size_t)(c0 + u) * lda + r0 + rr] : T(0);
#pragma unroll
            for (int j = 0; j < NC; j++) {
                const T xv = xs[j][rr];
#pragma unroll
                for (int u = 0; u < CW; u++) acc[u][j] += av[u] * xv;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < CW; u++)
#pragma unroll
        for (int j = 0; j < NC; j++) {
            T v = acc[u][j];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            acc[u][j] = v;
        }
#pragma unroll
    for (int u = 0; u < CW; u++) {
        const int c = c0 + u;
        if (c < k && lane < NC && lane < q) {
            T v = acc[u][0];
#pragma unroll
            for (int j = 1; j < NC; j++)
                if (lane == j) v = acc[u][j];
            T *y = Y + (size_t)lane * ldy + c;
            *y = beta == T(0) ? alpha * v : alpha * v + beta * *y;
        }
    }
}

// Y = alpha A^T X + beta Y, q <= 8 columns; A column-major m x k (lda)
template <typename T>
hipError_t dense_mm_t(hipStream_t st, int m, int k, int q, T alpha, const T *A, size_t lda, const T *X, size_t ldx,
                      T beta, T *Y, size_t ldy)
{
    if (k <= 0 || q <= 0) return hipSuccess;
    if (q > 8) return hipErrorInvalidValue;
    const dim3 g1((k + 3) / 4), g4((k + 15) / 16), b(256);
    if (q == 1) hipLaunchKernelGGL((k_mmt<T, 1, 1>), g1, b, 0, st, m, k, q, A, lda, X, ldx, alpha, beta, Y, ldy);
    else if (q <= 2) hipLaunchKernelGGL((k_mmt<T, 2, 1>), g1, b, 0, st, m, k, q, A, lda, X, ldx, alpha, beta, Y, ldy);
    else if (q <= 4) hipLaunchKernelGGL((k_mmt<T, 4, 4>), g4, b, 0, st, m, k, q, A, lda, X, ldx, alpha, beta, Y, ldy);
    else hipLaunchKernelGGL((k_mmt<T, 8, 4>), g4, b, 0, st, m, k, q, A, lda, X, ldx, alpha, beta, Y, ldy);
    return hipGetLastError();
}

template hipError_t dense_mm_t<double>(hipStream_t, int, int, int, double, const double *, size_t, const double *,
                                       size_t, double, double *, size_t);
template hipError_t dense_mm_t<float>(hipStream_t, int, int, int, float, const float *, size_t, const float *,
                                      size_t, float, float *, size_t);

// elements of workspace dense_mm needs for products with m rows
size_t dense_mm_work(int m) { return (size_t)32 * 32 * std::max(m, 1); }

template hipError_t dense_mm<double>(hipStream_t, int, int, int, double, const double *, size_t, const double *,
                                     size_t, double, double *, size_t, double *, size_t, int);
template hipError_t dense_mm<float>(hipStream_t, int, int, int, float, const float *, size_t, const float *, size_t,
                                    float, float *, size_t, float *, size_t, int);

}  // namespace iemic
