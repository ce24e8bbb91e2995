// Tall-skinny products of the block Krylov solver on the matrix cores
// (v_mfma_f64_16x16x4f64):
//   ts_gram:   G = A^T B      (A: n x a basis columns, B: n x b, b <= 32)
//   ts_update: C = beta C + alpha A H   (H: a x b, small)
// with n in the millions and a, b in the tens.  Both stream A once from HBM;
// the matrix cores form the a*b products of each group of 4 rows with one
// instruction per 16 x 16 output tile instead of 16 vector FMAs per lane,
// which keeps a block Gram-Schmidt pass at the cost of reading the basis.
//
// Lane layout of v_mfma_f64_16x16x4f64 (D = A B + C, A 16x4, B 4x16),
// checked with exact small-integer products before use: lane l holds
// A[l % 16][l / 16], B[l / 16][l % 16], and accumulator value v holds
// D[l / 16 + 4 v][l % 16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <mutex>
#include <string>

#include "ctx.hpp"

namespace iemic {

typedef double ts_d4 __attribute__((ext_vector_type(4)));

constexpr int TS_MAXB = 32;      // columns of B (ts_gram) or of C (ts_update) per launch
constexpr int TS_WAVES = 4;      // waves per workgroup
constexpr int TS_SLICE_MAX = 512;    // workgroups of one Gram product (two per CU)

// G = A^T B over one slice of rows, one partial per workgroup.  Workgroup
// (x, y): A columns [16 MT x, 16 MT (x + 1)), rows of slice y in blocks of
// 16 taken by its four waves in turn.  MFMA operands: M = A column, N = B
// column, K = row.  In a block starting at row rb, lane l supplies rows
// rb + 4 (l / 16) + s at step s = 0..3, so each lane reads 4 consecutive
// rows of its column (two 16-byte loads) and every 4 lanes read one whole
// 128-byte line.  The four waves' sums are added in LDS in a fixed order
// and written to P[y][c][j] (c < a, j < b).
template <int MT, int NT>
__global__ void __launch_bounds__(256) k_tsgram(int n, int a, int b, const double *__restrict__ A, size_t lda,
                                                const double *__restrict__ B, size_t ldb, int slice_rows, int vec,
                                                double *__restrict__ P)
{
    constexpr int NACC = MT * NT * 4;
    __shared__ double red[2][NACC][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kk = lane >> 4, ii = lane & 15;
    const int cbase = blockIdx.x * 16 * MT;
    const int r0 = blockIdx.y * slice_rows, r1 = min(n, r0 + slice_rows);
    const double *ap[MT];
    bool av[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) {
        const int c = cbase + 16 * t + ii;
        av[t] = c < a;
        ap[t] = A + (size_t)(av[t] ? c : 0) * lda;
    }
    const double *bp[NT];
    bool bv[NT];
#pragma unroll
    for (int u = 0; u < NT; u++) {
        const int j = 16 * u + ii;
        bv[u] = j < b;
        bp[u] = B + (size_t)(bv[u] ? j : 0) * ldb;
    }
    ts_d4 acc[MT][NT];
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++) acc[t][u] = ts_d4{0.0, 0.0, 0.0, 0.0};
    // the loop count and the load path are uniform across the wave (MFMA
    // needs every lane)
    for (int rb = r0 + 16 * wave; rb < r1; rb += 16 * TS_WAVES) {
        const int rl = rb + 4 * kk;
        double x[MT][4], y[NT][4];
        if (vec && rb + 16 <= r1) {
#pragma unroll
            for (int t = 0; t < MT; t++) {
                const double2 *p2 = reinterpret_cast<const double2 *>(ap[t] + rl);
                const double2 v0 = p2[0], v1 = p2[1];
                x[t][0] = av[t] ? v0.x : 0.0;
                x[t][1] = av[t] ? v0.y : 0.0;
                x[t][2] = av[t] ? v1.x : 0.0;
                x[t][3] = av[t] ? v1.y : 0.0;
            }
#pragma unroll
            for (int u = 0; u < NT; u++) {
                const double2 *p2 = reinterpret_cast<const double2 *>(bp[u] + rl);
                const double2 v0 = p2[0], v1 = p2[1];
                y[u][0] = bv[u] ? v0.x : 0.0;
                y[u][1] = bv[u] ? v0.y : 0.0;
                y[u][2] = bv[u] ? v1.x : 0.0;
                y[u][3] = bv[u] ? v1.y : 0.0;
            }
        } else {
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int r = rl + s;
                const bool rv = r < r1;
                const size_t rr = rv ? (size_t)r : (size_t)r0;
#pragma unroll
                for (int t = 0; t < MT; t++) {
                    const double v = ap[t][rr];
                    x[t][s] = (rv && av[t]) ? v : 0.0;
                }
#pragma unroll
                for (int u = 0; u < NT; u++) {
                    const double v = bp[u][rr];
                    y[u][s] = (rv && bv[u]) ? v : 0.0;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int u = 0; u < NT; u++)
                    acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[t][s], y[u][s], acc[t][u], 0, 0, 0);
    }
    // waves 2, 3 into LDS, added to waves 0, 1; wave 1 into LDS, added to 0
    if (wave >= 2) {
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int u = 0; u < NT; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) red[wave - 2][(t * NT + u) * 4 + v][lane] = acc[t][u][v];
    }
    __syncthreads();
    if (wave < 2) {
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int u = 0; u < NT; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[t][u][v] += red[wave][(t * NT + u) * 4 + v][lane];
    }
    __syncthreads();
    if (wave == 1) {
#pragma unroll
        for (int t = 0; t < MT; t++)
#pragma unroll
            for (int u = 0; u < NT; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) red[0][(t * NT + u) * 4 + v][lane] = acc[t][u][v];
    }
    __syncthreads();
    if (wave != 0) return;
    double *out = P + (size_t)blockIdx.y * a * b;
#pragma unroll
    for (int t = 0; t < MT; t++)
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int c = cbase + 16 * t + kk + 4 * v, j = 16 * u + ii;
                if (c < a && j < b) out[(size_t)c * b + j] = acc[t][u][v] + red[0][(t * NT + u) * 4 + v][lane];
            }
}

// G[c + j ldg] = sum over the np partials of P[p][c][j], in a fixed order:
// thread (e, g) of a workgroup sums partials g, g + 8, ... of entry e, and
// the eight sums are added in order of g.
__global__ void __launch_bounds__(256) k_tsgram_reduce(int a, int b, int np, const double *__restrict__ P,
                                                       double *__restrict__ G, int ldg)
{
    __shared__ double s[8][32];
    const int e = blockIdx.x * 32 + (threadIdx.x & 31), g = threadIdx.x >> 5;
    const int ab = a * b;
    double acc = 0.0;
    if (e < ab) {
#pragma unroll 8
        for (int p = g; p < np; p += 8) acc += P[(size_t)p * ab + e];
    }
    s[g][threadIdx.x & 31] = acc;
    __syncthreads();
    if (g == 0 && e < ab) {
        double v = s[0][threadIdx.x];
#pragma unroll
        for (int q = 1; q < 8; q++) v += s[q][threadIdx.x];
        const int c = e / b, j = e % b;
        G[c + (size_t)j * ldg] = v;
    }
}

// C = beta C + alpha A H, C with b <= 16 NT columns.  Each wave owns RT
// tiles of 16 rows and every output column of them, so it reads all of its
// rows of A before it writes its rows of C: C may be A itself (b <= a).
// MFMA operands (the transposed product C^T = H^T A^T): M = output column,
// N = row, K = column of A: lane l loads H[c + l / 16][j + l % 16] and
// A[r + l % 16][c + l / 16] (16 lanes read 128 contiguous bytes of one
// column), and value v of the accumulator is C[r + l % 16][j + l / 16 + 4 v].
template <int NT, int RT>
__global__ void __launch_bounds__(256) k_tsupdate(int n, int a, int b, double alpha, const double *A, size_t lda,
                                                  const double *__restrict__ H, int ldh, double beta, double *C,
                                                  size_t ldc)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kk = lane >> 4, ii = lane & 15;
    const long long rbase = ((long long)blockIdx.x * TS_WAVES + wave) * 16 * RT;
    if (rbase >= n) return;   // uniform per wave
    ts_d4 acc[NT][RT];
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
        for (int t = 0; t < RT; t++) acc[u][t] = ts_d4{0.0, 0.0, 0.0, 0.0};
    size_t rowoff[RT];
    bool rv[RT];
#pragma unroll
    for (int t = 0; t < RT; t++) {
        const long long r = rbase + 16 * t + ii;
        rv[t] = r < n;
        rowoff[t] = rv[t] ? (size_t)r : (size_t)rbase;
    }
    for (int c0 = 0; c0 < a; c0 += 4) {
        const int c = c0 + kk;
        const bool cv = c < a;
        const int cc = cv ? c : 0;
        double h[NT], x[RT];
#pragma unroll
        for (int u = 0; u < NT; u++) {
            const int j = 16 * u + ii;
            const double v = H[cc + (size_t)(j < b ? j : 0) * ldh];
            h[u] = (cv && j < b) ? v : 0.0;
        }
#pragma unroll
        for (int t = 0; t < RT; t++) {
            const double v = A[rowoff[t] + (size_t)cc * lda];
            x[t] = (cv && rv[t]) ? v : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NT; u++)
#pragma unroll
            for (int t = 0; t < RT; t++) acc[u][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(h[u], x[t], acc[u][t], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < NT; u++)
#pragma unroll
        for (int t = 0; t < RT; t++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int j = 16 * u + kk + 4 * v;
                if (j < b && rv[t]) {
                    double *cp = C + rowoff[t] + (size_t)j * ldc;
                    *cp = beta == 0.0 ? alpha * acc[u][t][v] : alpha * acc[u][t][v] + beta * *cp;
                }
            }
}

// ---------------------------------------------------------------- host
static int ts_col_tiles(int a) { return a <= 16 ? 1 : a <= 32 ? 2 : 4; }

static int ts_slices(int n, int a)
{
    const int mt = ts_col_tiles(a);
    const int gx = (a + 16 * mt - 1) / (16 * mt);
    const int want = std::max(1, TS_SLICE_MAX / gx);
    const int by_rows = std::max(1, (n + 2047) / 2048);   // at least ~512 rows per wave
    return std::min(want, by_rows);
}

// doubles of workspace ts_gram needs for n rows, up to a columns of A and
// b <= 32 columns of B
size_t ts_gram_work(int n, int a, int b)
{
    size_t best = 0;
    for (int aa = 1; aa <= a; aa++)
        best = std::max(best, (size_t)ts_slices(n, aa) * aa * std::min(b, TS_MAXB));
    return best;
}

template <int MT>
static void launch_gram_mt(hipStream_t st, dim3 g, int n, int a, int b, const double *A, size_t lda, const double *B,
                           size_t ldb, int rows, int vec, double *P)
{
    if (b <= 16) hipLaunchKernelGGL((k_tsgram<MT, 1>), g, dim3(256), 0, st, n, a, b, A, lda, B, ldb, rows, vec, P);
    else hipLaunchKernelGGL((k_tsgram<MT, 2>), g, dim3(256), 0, st, n, a, b, A, lda, B, ldb, rows, vec, P);
}

// G (a x b, column-major, leading dimension ldg) = A^T B; work holds
// ts_gram_work(n, a, b) doubles.  The result does not depend on scheduling.
hipError_t ts_gram(hipStream_t st, int n, int a, int b, const double *A, size_t lda, const double *B, size_t ldb,
                   double *G, int ldg, double *work)
{
    if (a <= 0 || b <= 0) return hipSuccess;
    if (b > TS_MAXB || n <= 0) return hipErrorInvalidValue;
    const int mt = ts_col_tiles(a);
    const int gx = (a + 16 * mt - 1) / (16 * mt);
    const int S = ts_slices(n, a);
    int rows = (n + S - 1) / S;
    rows = (rows + 15) / 16 * 16;
    const dim3 g(gx, S);
    // 16-byte loads of 4 rows need 16-byte aligned columns
    const int vec = (lda % 2 == 0 && ldb % 2 == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(B) & 15) == 0);
    switch (mt) {
    case 1: launch_gram_mt<1>(st, g, n, a, b, A, lda, B, ldb, rows, vec, work); break;
    case 2: launch_gram_mt<2>(st, g, n, a, b, A, lda, B, ldb, rows, vec, work); break;
    default: launch_gram_mt<4>(st, g, n, a, b, A, lda, B, ldb, rows, vec, work);
    }
    hipLaunchKernelGGL(k_tsgram_reduce, dim3((a * b + 31) / 32), dim3(256), 0, st, a, b, S, work, G, ldg);
    return hipGetLastError();
}

// C (n x b, b <= 32) = beta C + alpha A H, A n x a, H a x b (leading
// dimension ldh); C may be A itself when b <= a.
hipError_t ts_update(hipStream_t st, int n, int a, int b, double alpha, const double *A, size_t lda, const double *H,
                     int ldh, double beta, double *C, size_t ldc)
{
    if (b <= 0 || n <= 0) return hipSuccess;
    if (b > TS_MAXB || a < 0) return hipErrorInvalidValue;
    constexpr int RT = 4;
    const long long waves = ((long long)n + 16 * RT - 1) / (16 * RT);
    const dim3 g((unsigned)((waves + TS_WAVES - 1) / TS_WAVES));
    if (b <= 16)
        hipLaunchKernelGGL((k_tsupdate<1, RT>), g, dim3(256), 0, st, n, a, b, alpha, A, lda, H, ldh, beta, C, ldc);
    else
        hipLaunchKernelGGL((k_tsupdate<2, RT>), g, dim3(256), 0, st, n, a, b, alpha, A, lda, H, ldh, beta, C, ldc);
    return hipGetLastError();
}

}  // namespace iemic

// The two products on their own, for checking them against a host product:
// column-major operands with leading dimension n, the null stream, and the
// call returns when the result is in place.
extern "C" int iemic_block_gram_dev(int n, int a, int b, const double *d_A, const double *d_B, double *d_G)
{
    using namespace iemic;
    if (n < 1 || a < 1 || b < 1 || b > TS_MAXB) return set_error(IEMIC_EINVAL, "block Gram: sizes out of range");
    if (!d_A || !d_B || !d_G) return set_error(IEMIC_EINVAL, "block Gram: null operand");
    double *work = nullptr;
    HIPCHK(hipMalloc(&work, ts_gram_work(n, a, b) * sizeof(double)));
    hipError_t e = ts_gram(nullptr, n, a, b, d_A, (size_t)n, d_B, (size_t)n, d_G, a, work);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    (void)hipFree(work);
    if (e != hipSuccess) return set_error(IEMIC_EHIP, std::string("block Gram: ") + hipGetErrorString(e));
    return IEMIC_OK;
}

extern "C" int iemic_block_update_dev(int n, int a, int b, double alpha, const double *d_A, const double *d_H,
                                      double beta, double *d_C)
{
    using namespace iemic;
    if (n < 1 || a < 0 || b < 1 || b > TS_MAXB) return set_error(IEMIC_EINVAL, "block update: sizes out of range");
    if (!d_A || !d_C || (a > 0 && !d_H)) return set_error(IEMIC_EINVAL, "block update: null operand");
    if (d_C == d_A && b > a) return set_error(IEMIC_EINVAL, "block update: in place needs b <= a");
    hipError_t e = ts_update(nullptr, n, a, b, alpha, d_A, (size_t)n, d_H, a, beta, d_C, (size_t)n);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return set_error(IEMIC_EHIP, std::string("block update: ") + hipGetErrorString(e));
    return IEMIC_OK;
}

namespace iemic {

// y = a x + b y; with a = 0, x is not read, with b = 0, y is not read (so
// a = b = 0 fills y with zeros whatever it held)
__global__ void k_vec_axpby(long long n, double a, const double *x, double b, double *y)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = (a == 0.0 ? 0.0 : a * x[i]) + (b == 0.0 ? 0.0 : b * y[i]);
}

}  // namespace iemic

// Level-1 operations on device vectors for host-driven iterations (the
// Vector of the GMRES template contract, include/iemic_gmres.hpp): the null
// stream, and the call returns when the result is in place.
extern "C" int iemic_vec_axpby_dev(long long n, double a, const double *d_x, double b, double *d_y)
{
    using namespace iemic;
    if (n < 0) return set_error(IEMIC_EINVAL, "vector update: negative length");
    if (n == 0) return IEMIC_OK;
    if ((a != 0.0 && !d_x) || !d_y) return set_error(IEMIC_EINVAL, "vector update: null operand");
    const long long blocks = std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_vec_axpby, dim3((unsigned)blocks), dim3(256), 0, nullptr, n, a, d_x, b, d_y);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return set_error(IEMIC_EHIP, std::string("vector update: ") + hipGetErrorString(e));
    return IEMIC_OK;
}

namespace iemic {

// The partials of iemic_vec_dot_dev: one buffer per device, grown on demand
// and kept until the process ends; the mutex serialises the dots of one
// process (each returns only when its result is on the host).
static std::mutex g_dot_mu;
static double *g_dot_work[64] = {};
static size_t g_dot_cap[64] = {};

}  // namespace iemic

extern "C" int iemic_vec_dot_dev(long long n, const double *d_x, const double *d_y, double *out)
{
    using namespace iemic;
    if (!out) return set_error(IEMIC_EINVAL, "vector dot: null result");
    *out = 0.0;
    if (n <= 0) return n < 0 ? set_error(IEMIC_EINVAL, "vector dot: negative length") : IEMIC_OK;
    if (n > INT_MAX) return set_error(IEMIC_EINVAL, "vector dot: length beyond 2^31 - 1");
    if (!d_x || !d_y) return set_error(IEMIC_EINVAL, "vector dot: null operand");
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_error(IEMIC_EUNSUPPORTED, "vector dot: device index out of range");
    std::lock_guard<std::mutex> lock(g_dot_mu);
    const size_t need = ts_gram_work((int)n, 1, 1) + 1;
    if (g_dot_cap[dev] < need) {
        if (g_dot_work[dev]) (void)hipFree(g_dot_work[dev]);
        g_dot_work[dev] = nullptr;
        g_dot_cap[dev] = 0;
        HIPCHK(hipMalloc(&g_dot_work[dev], need * sizeof(double)));
        g_dot_cap[dev] = need;
    }
    double *work = g_dot_work[dev];
    hipError_t e = ts_gram(nullptr, (int)n, 1, 1, d_x, (size_t)n, d_y, (size_t)n, work, 1, work + 1);
    if (e == hipSuccess) e = hipMemcpy(out, work, sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(IEMIC_EHIP, std::string("vector dot: ") + hipGetErrorString(e));
    return IEMIC_OK;
}
