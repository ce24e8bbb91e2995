// Preconditioner slot of Ocean::buildPreconditioner / applyPrecon
// (Ocean.C:1377-1416), where the reference plugs in the TRIOS block
// preconditioner with ML/MRILU inner solvers (TRIOS_BlockPreconditioner.C).
//
// Water-column block ILU(0).  Land rows of J are identity rows and no ocean
// row references a land unknown, so land unknowns pass through unchanged and
// only ocean cells enter the factors.  The ocean cells of a water column
// (i,j) are the top d levels k0..l-1 (k0 = l - d); their unknowns form one
// dense block of size b = 6 d (local index 6 (k - k0) + var), so the
// vertical hydrostatic / continuity / mixing couplings are factored exactly.
// Between columns the block pattern is the horizontal 9-point neighbourhood
// and the factorisation is block ILU(0) in the natural column order:
//     for r in order:  for lower neighbours c of r (ascending):
//         L_rc = A_rc D_c^{-1};  A_rq -= L_rc U_cq  for q > c adjacent to r
//     D_r^{-1} = inverse of the updated A_rr (partial pivoting; a pivot below
//                1e-10 x the largest entry is replaced by a signed 1e-10 x largest).
// The periodic wrap-around couplings (i = 1 <-> i = n) are left out of the
// factors so that the column dependency graph has about n + 2 m levels
// instead of n m; the outer FGMRES sees the exact J.  The dense
// integral-condition row is represented by its diagonal entry (a rank-one
// change the Krylov solver absorbs).
//
// Blocks have variable shape b_r x b_q and are stored column-major (element
// (row, col) at col * b_r + row) at per-(column, neighbour) offsets, so that a
// wavefront reading one block column is unit-stride.  Columns with no level
// dependency between them are processed by one launch per level, one
// workgroup per column.  The FP32 factors used by the mixed-precision solver
// are the FP64 factors rounded to FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "ctx.hpp"

namespace iemic {

constexpr int NB9 = 9;   // neighbour slot (dj+1)*3 + (di+1); 4 is the column itself
constexpr int BMAX = 256;   // largest column block (6 l <= 256, i.e. l <= 42)

template <typename T>
struct Factors {
    T *blk = nullptr;
    unsigned version = 0;   // Jacobian version the factors belong to
    bool built = false;
};

struct Precond {
    int ncol = 0, nactive = 0, nlevels = 0, bmax = 0, maxw = 0;
    size_t total = 0;              // elements of all blocks
    int *d_nbr = nullptr;          // ncol*9 neighbour column or -1 (no wrap, no dry column)
    int *d_k0 = nullptr;           // ncol first ocean level (l for a dry column)
    int *d_rank = nullptr;         // ncol position in the elimination order
    uint8_t *d_pin = nullptr;      // ncell: 1 where the continuity row is replaced by p = 0
    int npins = 0;
    size_t *d_off = nullptr;       // ncol*9 block offsets (elements)
    int *d_order = nullptr;        // active columns sorted by level
    std::vector<int> lvl_off;      // nlevels+1 offsets into order
    Factors<double> f64;
    Factors<float> f32;
    int *d_flag = nullptr;         // ncol completion flags of the dataflow sweeps
    int *d_ticket = nullptr;       // [0] ticket counter, [1] error flag
    int epoch = 0;
    bool dataflow = true;          // one launch per sweep instead of one per level
    int *d_piv = nullptr;          // pivot scratch: widest level * bmax
    double *d_scratch = nullptr;   // block scratch: widest level * bmax^2
};

static inline int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

struct ColGeom {   // what a kernel needs to map a column block entry to the vector
    const int *nbr, *k0, *rank;
    const size_t *off;
    int ncol, l;
    __device__ int b(int c) const { return 6 * (l - k0[c]); }
    __device__ size_t g(int c, int e) const { return 6 * ((size_t)(k0[c] + e / 6) * ncol + c) + (e % 6); }
};

// ------------------------------------------------------------------ block assembly
template <typename T>
__global__ void k_fill(T *__restrict__ a, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        a[q] = (T)0;
}

// Scatter the slot values of every ocean cell into the column blocks.
__global__ void __launch_bounds__(256) k_scatter_blocks(Scalars s, ColGeom G, const double *__restrict__ J,
                                                        const double *__restrict__ coef, double sign,
                                                        const uint8_t *__restrict__ pin,
                                                        double *__restrict__ blk)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const int col = cell % G.ncol, k = cell / G.ncol;
    const int i = col % s.n, j = col / s.n;
    const int k0r = G.k0[col];
    if (k < k0r) return;                       // land: identity, not in the factors
    const int br = G.b(col);
    const size_t C = (size_t)s.ncell;
    const bool introw = (6 * cell + S == s.rowintcon);
    const bool pinned = pin[cell] != 0;
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        const int kk = SLOTS[sl].kk, e = SLOTS[sl].eq, v = SLOTS[sl].var;
        const int di = kk_di(kk), dj = kk_dj(kk), dk = kk_dk(kk);
        const int nk = k + dk;
        if ((introw && e == S) || (pinned && e == P)) continue;
        if (nk >= s.l || j + dj < 0 || j + dj >= s.m) continue;
        const int sn = (dj + 1) * 3 + (di + 1);
        const int q = G.nbr[(size_t)col * NB9 + sn];
        if (q < 0 || nk < G.k0[q]) continue;   // wrap, dry column or land cell: zero entry
        const int row = 6 * (k - k0r) + e, cc = 6 * (nk - G.k0[q]) + v;
        blk[G.off[(size_t)col * NB9 + sn] + (size_t)cc * br + row] = J[(size_t)sl * C + cell];
    }
    if (pinned) {
        const int row = 6 * (k - k0r) + P;
        blk[G.off[(size_t)col * NB9 + 4] + (size_t)row * br + row] = 1.0;
    }
    if (introw) {
        const double d = sign * coef[6 * (size_t)cell + S];
        const int row = 6 * (k - k0r) + S;
        blk[G.off[(size_t)col * NB9 + 4] + (size_t)row * br + row] = (d != 0.0 ? d : 1.0);
    }
}

// ------------------------------------------------------------------ dense kernels (one workgroup)
// C (+)= alpha * A * B, column-major, A: M x K (lda), B: K x N (ldb),
// C: M x N (ldc).  64 x 64 output tiles through LDS; 16 x 16 threads, each
// owning a 4 x 4 sub-tile.
template <typename T>
__device__ void wg_gemm(const T *__restrict__ A, int lda, const T *__restrict__ B, int ldb, T *__restrict__ Cm,
                        int ldc, int M, int N, int K, T alpha, bool accumulate)
{
    __shared__ T As[64][17];
    __shared__ T Bs[16][65];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    for (int m0 = 0; m0 < M; m0 += 64)
        for (int n0 = 0; n0 < N; n0 += 64) {
            T acc[4][4];
#pragma unroll
            for (int p = 0; p < 4; p++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc[p][q] = 0;
            for (int k0 = 0; k0 < K; k0 += 16) {
                for (int e = threadIdx.x; e < 64 * 16; e += blockDim.x) {
                    int r = e % 64, c = e / 64;
                    As[r][c] = (m0 + r < M && k0 + c < K) ? A[(size_t)(k0 + c) * lda + m0 + r] : (T)0;
                }
                for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
                    int r = e % 16, c = e / 16;
                    Bs[r][c] = (k0 + r < K && n0 + c < N) ? B[(size_t)(n0 + c) * ldb + k0 + r] : (T)0;
                }
                __syncthreads();
#pragma unroll 4
                for (int kk = 0; kk < 16; kk++) {
                    T a[4], bv[4];
#pragma unroll
                    for (int p = 0; p < 4; p++) a[p] = As[tx + 16 * p][kk];
#pragma unroll
                    for (int q = 0; q < 4; q++) bv[q] = Bs[kk][ty + 16 * q];
#pragma unroll
                    for (int p = 0; p < 4; p++)
#pragma unroll
                        for (int q = 0; q < 4; q++) acc[p][q] += a[p] * bv[q];
                }
                __syncthreads();
            }
#pragma unroll
            for (int p = 0; p < 4; p++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    int r = m0 + tx + 16 * p, c = n0 + ty + 16 * q;
                    if (r < M && c < N) {
                        T v = alpha * acc[p][q];
                        T *dst = Cm + (size_t)c * ldc + r;
                        *dst = accumulate ? *dst + v : v;
                    }
                }
            __syncthreads();
        }
}

// In-place inverse of a b x b column-major block by Gauss-Jordan elimination
// with partial pivoting (tiny pivots replaced, see file comment).
template <typename T>
__device__ void wg_invert(T *__restrict__ A, int b, int *__restrict__ piv, double thr, double rep)
{
    __shared__ T sbig;
    __shared__ int sp;
    __shared__ T red[256];
    __shared__ int redi[256];
    T mx = 0;
    for (size_t e = threadIdx.x; e < (size_t)b * b; e += blockDim.x) mx = fmax(mx, (T)fabs((double)A[e]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) sbig = red[0] > 0 ? red[0] : (T)1;
    __syncthreads();
    const T tiny = (T)thr * sbig, repl = (T)rep * sbig;
    for (int q = 0; q < b; q++) {
        T best = -1;
        int bi = q;
        for (int r = q + threadIdx.x; r < b; r += blockDim.x) {
            T v = (T)fabs((double)A[(size_t)q * b + r]);
            if (v > best) { best = v; bi = r; }
        }
        red[threadIdx.x] = best;
        redi[threadIdx.x] = bi;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                T o = red[threadIdx.x + w];
                int oi = redi[threadIdx.x + w];
                if (o > red[threadIdx.x] || (o == red[threadIdx.x] && oi < redi[threadIdx.x])) {
                    red[threadIdx.x] = o;
                    redi[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            sp = redi[0];
            piv[q] = sp;
        }
        __syncthreads();
        const int p = sp;
        if (p != q)
            for (int c = threadIdx.x; c < b; c += blockDim.x) {
                T t = A[(size_t)c * b + q];
                A[(size_t)c * b + q] = A[(size_t)c * b + p];
                A[(size_t)c * b + p] = t;
            }
        __syncthreads();
        if (threadIdx.x == 0) {
            T d = A[(size_t)q * b + q];
            if (fabs((double)d) < (double)tiny) d = (d < 0) ? -repl : repl;
            A[(size_t)q * b + q] = d;
        }
        __syncthreads();
        const T inv = (T)1 / A[(size_t)q * b + q];
        for (int c = threadIdx.x; c < b; c += blockDim.x)
            if (c != q) A[(size_t)c * b + q] *= inv;
        __syncthreads();
        if (threadIdx.x == 0) A[(size_t)q * b + q] = inv;
        __syncthreads();
        for (size_t e = threadIdx.x; e < (size_t)b * b; e += blockDim.x) {
            int r = (int)(e % b), c = (int)(e / b);
            if (r == q || c == q) continue;
            A[(size_t)c * b + r] -= A[(size_t)q * b + r] * A[(size_t)c * b + q];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < b; r += blockDim.x)
            if (r != q) A[(size_t)q * b + r] = -A[(size_t)q * b + r] * inv;
        __syncthreads();
    }
    for (int q = b - 1; q >= 0; q--) {
        int p = piv[q];
        if (p != q)
            for (int r = threadIdx.x; r < b; r += blockDim.x) {
                T t = A[(size_t)q * b + r];
                A[(size_t)q * b + r] = A[(size_t)p * b + r];
                A[(size_t)p * b + r] = t;
            }
        __syncthreads();
    }
}

// One ILU(0) row step per workgroup for the columns of one level (FP64).
__global__ void __launch_bounds__(256) k_cbilu_level(const int *__restrict__ cols, ColGeom G,
                                                     double *__restrict__ blk, double *__restrict__ scratch,
                                                     int bmax, int *__restrict__ pivs, double thr, double rep)
{
    const int r = cols[blockIdx.x];
    const int br = G.b(r);
    double *W = scratch + (size_t)blockIdx.x * bmax * bmax;
    int nr[NB9];
#pragma unroll
    for (int s = 0; s < NB9; s++) nr[s] = G.nbr[(size_t)r * NB9 + s];
    int ord[NB9], no = 0;
    for (int s = 0; s < NB9; s++)
        if (nr[s] >= 0 && G.rank[nr[s]] < G.rank[r]) ord[no++] = s;
    for (int a = 1; a < no; a++)
        for (int z = a; z > 0 && G.rank[nr[ord[z]]] < G.rank[nr[ord[z - 1]]]; z--) {
            int t = ord[z]; ord[z] = ord[z - 1]; ord[z - 1] = t;
        }
    for (int a = 0; a < no; a++) {
        const int sc = ord[a], c = nr[sc];
        const int bc = G.b(c);
        double *Arc = blk + G.off[(size_t)r * NB9 + sc];
        for (size_t e = threadIdx.x; e < (size_t)br * bc; e += blockDim.x) W[e] = Arc[e];
        __syncthreads();
        // L_rc = A_rc D_c^{-1}
        wg_gemm<double>(W, br, blk + G.off[(size_t)c * NB9 + 4], bc, Arc, br, br, bc, bc, 1.0, false);
        for (int sq = 0; sq < NB9; sq++) {
            const int q = G.nbr[(size_t)c * NB9 + sq];
            if (q < 0 || G.rank[q] <= G.rank[c]) continue;
            int sr = -1;
            for (int s = 0; s < NB9; s++)
                if (nr[s] == q) sr = s;
            if (sr < 0) continue;
            const int bq = G.b(q);
            // A_rq -= L_rc U_cq
            wg_gemm<double>(Arc, br, blk + G.off[(size_t)c * NB9 + sq], bc, blk + G.off[(size_t)r * NB9 + sr], br,
                            br, bq, bc, -1.0, true);
        }
    }
    wg_invert<double>(blk + G.off[(size_t)r * NB9 + 4], br, pivs + (size_t)blockIdx.x * bmax, thr, rep);
}

template <typename T> struct Pair;
template <> struct Pair<double> { using type = double2; };
template <> struct Pair<float> { using type = float2; };

constexpr int SWEEP_THREADS = 1024;
constexpr int MAXNB = 9;   // neighbours of a column on either side, plus one

// acc[e] -= sum_s sum_q B_s(e, q) xs[xo_s + q] for e < br over nb blocks of
// br rows (br even, column-major, 16-byte aligned), all threads of the
// workgroup: thread t takes the row pair 2 (t % h), h = br / 2, and every
// S-th column of the concatenated blocks from t / h; four columns in flight
// per thread.  Partial sums meet in LDS (part: 2 * blockDim.x entries).
template <typename T>
__device__ void wg_multi_mv_sub(const T *const *Bs, const int *bcs, const int *xo, int nb, int br,
                                const T *__restrict__ xs, T *__restrict__ part, T *__restrict__ acc)
{
    using V2 = typename Pair<T>::type;
    const int h = br / 2;
    const int S = (int)blockDim.x / h;
    const int e2 = threadIdx.x % h, p = threadIdx.x / h;
    T a0 = 0, a1 = 0;
    if (p < S) {
        for (int s = 0; s < nb; s++) {
            const V2 *B = reinterpret_cast<const V2 *>(Bs[s]) + e2;
            const T *x = xs + xo[s];
            const int bc = bcs[s];
            int q = p;
            for (; q + 3 * S < bc; q += 4 * S) {
                const V2 b0 = B[(size_t)q * h], b1 = B[(size_t)(q + S) * h];
                const V2 b2 = B[(size_t)(q + 2 * S) * h], b3 = B[(size_t)(q + 3 * S) * h];
                const T x0 = x[q], x1 = x[q + S], x2 = x[q + 2 * S], x3 = x[q + 3 * S];
                a0 += b0.x * x0;
                a1 += b0.y * x0;
                a0 += b1.x * x1;
                a1 += b1.y * x1;
                a0 += b2.x * x2;
                a1 += b2.y * x2;
                a0 += b3.x * x3;
                a1 += b3.y * x3;
            }
            for (; q < bc; q += S) {
                const V2 b0 = B[(size_t)q * h];
                a0 += b0.x * x[q];
                a1 += b0.y * x[q];
            }
        }
        part[p * br + 2 * e2] = a0;
        part[p * br + 2 * e2 + 1] = a1;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < br; e += blockDim.x) {
        T sum = 0;
        for (int z = 0; z < S; z++) sum += part[z * br + e];
        acc[e] -= sum;
    }
    __syncthreads();
}

// Forward step of column r: y_r = v_r - sum_{c<r} L_rc y_c
// Backward step of column r: x_r = D_r^{-1} (y_r - sum_{q>r} U_rq x_q)
// WAIT(cs, nb) is called by every thread before the values of the nb
// neighbours cs[] are read.
template <typename T, bool BACKWARD, typename Wait>
__device__ void sweep_column(int r, const ColGeom &G, const T *__restrict__ blk, const T *__restrict__ in,
                             T *__restrict__ io, T *xs, T *acc, T *part, Wait wait)
{
    const int br = G.b(r);
    const T *Bs[MAXNB];
    int bcs[MAXNB], xo[MAXNB], cs[MAXNB], nb = 0, tot = 0;
    for (int s = 0; s < NB9; s++) {
        const int c = G.nbr[(size_t)r * NB9 + s];
        if (c < 0 || (BACKWARD ? G.rank[c] <= G.rank[r] : G.rank[c] >= G.rank[r])) continue;
        Bs[nb] = blk + G.off[(size_t)r * NB9 + s];
        bcs[nb] = G.b(c);
        xo[nb] = tot;
        cs[nb] = c;
        tot += bcs[nb];
        nb++;
    }
    wait(cs, nb);
    for (int a = 0; a < nb; a++)
        for (int e = threadIdx.x; e < bcs[a]; e += blockDim.x) xs[xo[a] + e] = io[G.g(cs[a], e)];
    for (int e = threadIdx.x; e < br; e += blockDim.x) acc[e] = BACKWARD ? io[G.g(r, e)] : in[G.g(r, e)];
    __syncthreads();
    if (nb) wg_multi_mv_sub<T>(Bs, bcs, xo, nb, br, xs, part, acc);
    if (BACKWARD) {
        for (int e = threadIdx.x; e < br; e += blockDim.x) {
            xs[e] = acc[e];
            acc[e] = 0;
        }
        __syncthreads();
        const T *D = blk + G.off[(size_t)r * NB9 + 4];
        const int zero = 0;
        wg_multi_mv_sub<T>(&D, &br, &zero, 1, br, xs, part, acc);
        for (int e = threadIdx.x; e < br; e += blockDim.x) acc[e] = -acc[e];
        __syncthreads();
    }
    for (int e = threadIdx.x; e < br; e += blockDim.x) io[G.g(r, e)] = acc[e];
}

// One level per launch: the launch boundary orders the levels.
template <typename T, bool BACKWARD>
__global__ void __launch_bounds__(SWEEP_THREADS) k_cbilu_sweep(const int *__restrict__ cols, ColGeom G,
                                                               const T *__restrict__ blk, const T *__restrict__ in,
                                                               T *__restrict__ io)
{
    __shared__ T xs[(MAXNB - 1) * BMAX], acc[BMAX], part[2 * SWEEP_THREADS];
    sweep_column<T, BACKWARD>(cols[blockIdx.x], G, blk, in, io, xs, acc, part, [](const int *, int) {});
}

// All levels in one launch (dataflow).  Each workgroup draws a ticket and
// takes the column at that position of the level order (reversed for the
// backward sweep), so every column it waits for was drawn earlier by a
// running workgroup.  Completion is published per column: after a workgroup
// barrier thread 0 releases at agent scope and sets flag[r] = epoch; a
// consumer's thread 0 spins on the flags (bounded: past SPIN_LIMIT polls it
// records an error and goes on, so the grid always drains), acquires, and a
// barrier passes the ordering on to the rest of the workgroup.
constexpr int SPIN_LIMIT = 1 << 20;

template <typename T, bool BACKWARD>
__global__ void __launch_bounds__(SWEEP_THREADS) k_cbilu_sweep_df(const int *__restrict__ order, int nact, ColGeom G,
                                                        const T *__restrict__ blk, const T *__restrict__ in,
                                                        T *__restrict__ io, int *__restrict__ flag,
                                                        int *__restrict__ ticket, int epoch, int *__restrict__ err)
{
    __shared__ T xs[(MAXNB - 1) * BMAX], acc[BMAX], part[2 * SWEEP_THREADS];
    __shared__ int s_r;
    if (threadIdx.x == 0) {
        const int t = atomicAdd(ticket, 1);
        s_r = order[BACKWARD ? nact - 1 - t : t];
    }
    __syncthreads();
    const int r = s_r;
    // thread 0 polls and then acquires; the workgroup barrier extends the
    // ordering to the other waves, which share its caches
    auto wait = [&](const int *cs, int nb) {
        if (threadIdx.x == 0) {
            for (int a = 0; a < nb; a++) {
                int spins = 0;
                while (__hip_atomic_load(flag + cs[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                    if (++spins > SPIN_LIMIT) {
                        atomicOr(err, 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
    };
    sweep_column<T, BACKWARD>(r, G, blk, in, io, xs, acc, part, wait);
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(flag + r, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_zero_int(int *a, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = 0;
}

// land unknowns: identity rows
template <typename T>
__global__ void k_copy_land(ColGeom G, int ncell, const T *__restrict__ in, T *__restrict__ out)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const int col = cell % G.ncol, k = cell / G.ncol;
    if (k >= G.k0[col]) return;
#pragma unroll
    for (int v = 0; v < 6; v++) out[6 * (size_t)cell + v] = in[6 * (size_t)cell + v];
}

template <typename A, typename B>
__global__ void k_cast(A *__restrict__ dst, const B *__restrict__ src, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = (A)src[q];
}

__global__ void k_round32(float *__restrict__ dst, const double *__restrict__ src, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = (float)src[q];
}

// ------------------------------------------------------------------ host
static void cilu_free(iemic_ctx *c)
{
    Precond *p = c->prec;
    if (!p) return;
    void *bufs[] = {p->d_nbr, p->d_k0, p->d_rank, p->d_pin, p->d_flag, p->d_ticket, p->d_off, p->d_order, p->f64.blk, p->f32.blk, p->d_piv, p->d_scratch};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
    c->prec = nullptr;
}

static ColGeom geom(iemic_ctx *c)
{
    Precond *p = c->prec;
    return ColGeom{p->d_nbr, p->d_k0, p->d_rank, p->d_off, p->ncol, c->cfg.l};
}

// First ocean level (0-based) of every water column i + n j; l for a dry
// column.  Ocean cells are contiguous from the surface down.
std::vector<int> column_tops(const iemic_ctx *c)
{
    const int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    std::vector<int> k0((size_t)n * m, l);
    const auto &lm = c->h.landm;
    for (int j = 0; j < m; j++)
        for (int i = 0; i < n; i++) {
            int kk = l;
            while (kk > 0 && lm[(size_t)(i + 1) + (size_t)(n + 2) * ((j + 1) + (size_t)(m + 2) * kk)] == OCEAN) kk--;
            k0[i + n * j] = kk;
        }
    return k0;
}

// Pressure pins.  On the B grid the pressure gradient at a velocity point
// whose four surrounding cells are ocean vanishes exactly when
// p(i,j) = p(i+1,j+1) and p(i+1,j) = p(i,j+1), and the hydrostatic rows tie
// the levels of a column together.  The connected components of these
// couplings are the pressure basins: J has one null vector per basin (p
// constant on it, zero elsewhere) and one dependent combination of continuity
// rows.  A preconditioner that factors J replaces, per basin, the continuity
// row of the cell that comes last in its elimination order (highest column
// rank, then highest level) by the row p = 0, which makes the factors regular;
// J itself is unchanged.  Returns one flag per cell.
std::vector<uint8_t> pressure_pins(const iemic_ctx *c, const std::vector<int> &k0, const std::vector<int> &rank)
{
    const int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l, ncol = n * m;
    const size_t C = (size_t)n * m * l;
    std::vector<uint8_t> pinv(C, 0);
    std::vector<int> par(C);
    for (size_t q = 0; q < C; q++) par[q] = (int)q;
    auto find = [&](int a) {
        while (par[a] != a) a = par[a] = par[par[a]];
        return a;
    };
    auto unite = [&](int a, int b) { par[find(a)] = find(b); };
    auto cell_of = [&](int i, int j, int k) { return i + n * (j + m * k); };
    auto wet = [&](int i, int j, int k) { return k >= k0[i + n * j]; };
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n; i++) {
                if (!wet(i, j, k)) continue;
                if (k + 1 < l) unite(cell_of(i, j, k), cell_of(i, j, k + 1));
                int i1 = i + 1;
                if (i1 == n) {
                    if (!c->cfg.periodic) continue;
                    i1 = 0;
                }
                if (j + 1 >= m) continue;
                if (wet(i1, j, k) && wet(i, j + 1, k) && wet(i1, j + 1, k)) {
                    unite(cell_of(i, j, k), cell_of(i1, j + 1, k));
                    unite(cell_of(i1, j, k), cell_of(i, j + 1, k));
                }
            }
    // per basin: the cell of the highest-ranked column, then the highest level
    std::vector<long long> best(C, -1);
    std::vector<int> bestcell(C, -1);
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n; i++) {
                if (!wet(i, j, k)) continue;
                const int cl = cell_of(i, j, k), root = find(cl);
                const long long key = (long long)rank[i + n * j] * l + k;
                if (best[root] < 0 || key > best[root]) {
                    best[root] = key;
                    bestcell[root] = cl;
                }
            }
    for (size_t q = 0; q < C; q++)
        if (bestcell[q] >= 0) pinv[(size_t)bestcell[q]] = 1;
    (void)ncol;
    return pinv;
}

static int ensure(iemic_ctx *c)
{
    if (c->prec) return IEMIC_OK;
    const int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    if (6 * l > BMAX) return set_error(IEMIC_EUNSUPPORTED, "column preconditioner supports at most 42 levels");
    Precond *p = new Precond();
    c->prec = p;
    p->ncol = n * m;
    const std::vector<int> k0 = column_tops(c);
    // elimination order: natural (i fastest), without the periodic wrap
    std::vector<int> rank(p->ncol), byrank(p->ncol);
    for (int j = 0; j < m; j++)
        for (int i = 0; i < n; i++) rank[i + n * j] = i + n * j;
    for (int r = 0; r < p->ncol; r++) byrank[rank[r]] = r;
    std::vector<int> nbr((size_t)p->ncol * NB9, -1), level(p->ncol, -1);
    std::vector<size_t> off((size_t)p->ncol * NB9, 0);
    size_t total = 0;
    int nlev = 0;
    for (int t = 0; t < p->ncol; t++) {
        const int r = byrank[t], i = r % n, j = r / n;
        if (k0[r] >= l) continue;
        const size_t br = 6 * (size_t)(l - k0[r]);
        int lv = 0;
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const int ii = i + di, jj = j + dj;
                if (ii < 0 || ii >= n || jj < 0 || jj >= m) continue;
                const int q = ii + n * jj;
                if (k0[q] >= l) continue;
                const int s = (dj + 1) * 3 + (di + 1);
                nbr[(size_t)r * NB9 + s] = q;
                off[(size_t)r * NB9 + s] = total;
                total += br * 6 * (size_t)(l - k0[q]);
                if (rank[q] < rank[r]) lv = std::max(lv, level[q] + 1);
            }
        level[r] = lv;
        nlev = std::max(nlev, lv + 1);
        p->nactive++;
        p->bmax = std::max(p->bmax, (int)br);
    }
    std::vector<uint8_t> pinv = pressure_pins(c, k0, rank);
    for (uint8_t v : pinv) p->npins += v;
    p->total = total;
    p->nlevels = nlev;
    std::vector<int> order(p->nactive);
    p->lvl_off.assign(p->nlevels + 1, 0);
    for (int r = 0; r < p->ncol; r++)
        if (level[r] >= 0) p->lvl_off[level[r] + 1]++;
    for (int q = 0; q < p->nlevels; q++) p->lvl_off[q + 1] += p->lvl_off[q];
    std::vector<int> fill(p->lvl_off.begin(), p->lvl_off.end() - 1);
    for (int r = 0; r < p->ncol; r++)
        if (level[r] >= 0) order[fill[level[r]]++] = r;
    p->maxw = 0;
    for (int q = 0; q < p->nlevels; q++) p->maxw = std::max(p->maxw, p->lvl_off[q + 1] - p->lvl_off[q]);
    HIPCHK(hipMalloc(&p->d_nbr, nbr.size() * sizeof(int)));
    HIPCHK(hipMemcpy(p->d_nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&p->d_k0, k0.size() * sizeof(int)));
    HIPCHK(hipMemcpy(p->d_k0, k0.data(), k0.size() * sizeof(int), hipMemcpyHostToDevice));
    if (const char *v = std::getenv("IEMIC_LEVEL_LAUNCH")) p->dataflow = std::atoi(v) == 0;
    HIPCHK(hipMalloc(&p->d_flag, (size_t)p->ncol * sizeof(int)));
    HIPCHK(hipMemset(p->d_flag, 0, (size_t)p->ncol * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_ticket, 2 * sizeof(int)));
    HIPCHK(hipMemset(p->d_ticket, 0, 2 * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_pin, pinv.size()));
    HIPCHK(hipMemcpy(p->d_pin, pinv.data(), pinv.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&p->d_rank, rank.size() * sizeof(int)));
    HIPCHK(hipMemcpy(p->d_rank, rank.data(), rank.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&p->d_off, off.size() * sizeof(size_t)));
    HIPCHK(hipMemcpy(p->d_off, off.data(), off.size() * sizeof(size_t), hipMemcpyHostToDevice));
    if (p->nactive) {
        HIPCHK(hipMalloc(&p->d_order, order.size() * sizeof(int)));
        HIPCHK(hipMemcpy(p->d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&p->d_piv, (size_t)p->maxw * p->bmax * sizeof(int)));
        HIPCHK(hipMalloc(&p->d_scratch, (size_t)p->maxw * p->bmax * p->bmax * sizeof(double)));
    }
    return IEMIC_OK;
}

static int factor64(iemic_ctx *c)
{
    Precond *p = c->prec;
    Factors<double> &F = p->f64;
    if (!F.blk) HIPCHK(hipMalloc(&F.blk, std::max<size_t>(p->total, 1) * sizeof(double)));
    const ColGeom G = geom(c);
    hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, c->stream, F.blk, p->total);
    hipLaunchKernelGGL(k_scatter_blocks, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc, G, c->d_J,
                       c->d_coef, (double)c->cfg.int_sign, p->d_pin, F.blk);
    for (int lv = 0; lv < p->nlevels; lv++) {
        const int w = p->lvl_off[lv + 1] - p->lvl_off[lv];
        hipLaunchKernelGGL(k_cbilu_level, dim3(w), dim3(256), 0, c->stream, p->d_order + p->lvl_off[lv], G, F.blk,
                           p->d_scratch, p->bmax, p->d_piv, 1e-10, 1e-10);
    }
    HIPCHK(hipGetLastError());
    F.built = true;
    F.version = c->jac_version;
    return IEMIC_OK;
}

template <typename T>
static int sweep(iemic_ctx *c, const T *blk, const T *in, T *out)
{
    Precond *p = c->prec;
    const ColGeom G = geom(c);
    hipLaunchKernelGGL(k_copy_land<T>, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, G, c->sc.ncell, in,
                       out);
    if (p->nactive == 0) return IEMIC_OK;
    if (p->dataflow) {
        for (int dir = 0; dir < 2; dir++) {
            const int epoch = ++p->epoch;
            hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, p->d_ticket, 1);
            if (dir == 0)
                hipLaunchKernelGGL((k_cbilu_sweep_df<T, false>), dim3(p->nactive), dim3(SWEEP_THREADS), 0, c->stream,
                                   p->d_order, p->nactive, G, blk, in, out, p->d_flag, p->d_ticket, epoch,
                                   p->d_ticket + 1);
            else
                hipLaunchKernelGGL((k_cbilu_sweep_df<T, true>), dim3(p->nactive), dim3(SWEEP_THREADS), 0, c->stream,
                                   p->d_order, p->nactive, G, blk, in, out, p->d_flag, p->d_ticket, epoch,
                                   p->d_ticket + 1);
        }
        HIPCHK(hipGetLastError());
        return IEMIC_OK;
    }
    for (int lv = 0; lv < p->nlevels; lv++) {
        const int w = p->lvl_off[lv + 1] - p->lvl_off[lv];
        hipLaunchKernelGGL((k_cbilu_sweep<T, false>), dim3(w), dim3(SWEEP_THREADS), 0, c->stream, p->d_order + p->lvl_off[lv],
                           G, blk, in, out);
    }
    for (int lv = p->nlevels - 1; lv >= 0; lv--) {
        const int w = p->lvl_off[lv + 1] - p->lvl_off[lv];
        hipLaunchKernelGGL((k_cbilu_sweep<T, true>), dim3(w), dim3(SWEEP_THREADS), 0, c->stream, p->d_order + p->lvl_off[lv],
                           G, blk, in, out);
    }
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

// Error flag of the dataflow sweeps (a wait that ran past its spin limit).
static int cilu_check(iemic_ctx *c)
{
    Precond *p = c->prec;
    if (!p || !p->d_ticket) return IEMIC_OK;
    int e = 0;
    HIPCHK(hipMemcpyAsync(&e, p->d_ticket + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!e) return IEMIC_OK;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, p->d_ticket + 1, 1);
    return set_error(IEMIC_EHIP, "preconditioner sweep: dependency wait exceeded its spin limit");
}

// the column ILU's sweeps read the couplings from J itself, so its factors
// are always rebuilt for a new Jacobian (reuse applies to the row LU only)
static bool current(const iemic_ctx *c, bool built, unsigned version)
{
    return built && version == c->jac_version;
}

static int cilu_build(iemic_ctx *c)
{
    int rc = ensure(c);
    if (rc) return rc;
    timed_begin_ext(c, 4);
    rc = factor64(c);
    timed_end_ext(c, 4);
    return rc;
}

static int cilu_build_f32(iemic_ctx *c)
{
    int rc = ensure(c);
    if (rc) return rc;
    Precond *p = c->prec;
    if (!current(c, p->f64.built, p->f64.version) && (rc = cilu_build(c))) return rc;
    timed_begin_ext(c, 5);
    if (!p->f32.blk) HIPCHK(hipMalloc(&p->f32.blk, std::max<size_t>(p->total, 1) * sizeof(float)));
    hipLaunchKernelGGL(k_round32, dim3(2048), dim3(256), 0, c->stream, p->f32.blk, p->f64.blk, p->total);
    timed_end_ext(c, 5);
    HIPCHK(hipGetLastError());
    p->f32.built = true;
    p->f32.version = p->f64.version;
    return IEMIC_OK;
}

// Factors for the current Jacobian in the given precision (0 FP64, 1 FP32),
// built only if the Jacobian changed since the last build.
static int cilu_ready(iemic_ctx *c, int precision)
{
    int rc = ensure(c);
    if (rc) return rc;
    Precond *p = c->prec;
    if (precision == 0) return current(c, p->f64.built, p->f64.version) ? IEMIC_OK : cilu_build(c);
    return current(c, p->f32.built, p->f32.version) ? IEMIC_OK : cilu_build_f32(c);
}

static int cilu_apply(iemic_ctx *c, const double *d_in, double *d_out)
{
    if (!c->prec || !current(c, c->prec->f64.built, c->prec->f64.version)) {
        int rc = cilu_build(c);
        if (rc) return rc;
    }
    timed_begin_ext(c, 3);
    int rc = sweep<double>(c, c->prec->f64.blk, d_in, d_out);
    timed_end_ext(c, 3);
    return rc;
}

static int cilu_apply_f32(iemic_ctx *c, const float *d_in, float *d_out)
{
    if (!c->prec || !current(c, c->prec->f32.built, c->prec->f32.version)) {
        int rc = cilu_build_f32(c);
        if (rc) return rc;
    }
    timed_begin_ext(c, 6);
    int rc = sweep<float>(c, c->prec->f32.blk, d_in, d_out);
    timed_end_ext(c, 6);
    return rc;
}

// ------------------------------------------------------------------ preconditioner slot
int rowlu_build(iemic_ctx *c, int precision);
int rowlu_ready(iemic_ctx *c, int precision);
int rowlu_apply(iemic_ctx *c, int precision, const double *in, double *out);
int rowlu_apply_t(iemic_ctx *c, int precision, const double *in, double *out);
void rowlu_free(iemic_ctx *c);

void precond_free(iemic_ctx *c)
{
    cilu_free(c);
    rowlu_free(c);
}

int precond_build(iemic_ctx *c, int precision)
{
    if (c->prec_kind == IEMIC_PREC_ROW_LU) return rowlu_build(c, precision);
    return precision ? cilu_build_f32(c) : cilu_build(c);
}

int precond_ready(iemic_ctx *c, int precision)
{
    if (c->prec_kind == IEMIC_PREC_ROW_LU) return rowlu_ready(c, precision);
    return cilu_ready(c, precision);
}

// z = M^{-1} v on the model's unknowns, factors and sweeps in FP64
// (precision 0) or FP32 (precision 1); tmp32 holds 2 ndim floats for the
// FP32 column sweeps.
int precond_apply(iemic_ctx *c, int precision, const double *in, double *out, float *tmp32)
{
    if (c->prec_kind == IEMIC_PREC_ROW_LU) return rowlu_apply(c, precision, in, out);
    if (!precision) return cilu_apply(c, in, out);
    const int n = c->sc.ndim;
    hipLaunchKernelGGL((k_cast<float, double>), dim3(nblk(n, 256)), dim3(256), 0, c->stream, tmp32, in, n);
    int rc = cilu_apply_f32(c, tmp32, tmp32 + n);
    if (rc) return rc;
    hipLaunchKernelGGL((k_cast<double, float>), dim3(nblk(n, 256)), dim3(256), 0, c->stream, out, tmp32 + n, n);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

// z = M^{-T} v (adjoint solves); the row LU only
int precond_apply_t(iemic_ctx *c, int precision, const double *in, double *out)
{
    if (c->prec_kind != IEMIC_PREC_ROW_LU)
        return set_error(IEMIC_EUNSUPPORTED, "transposed preconditioner: row LU only");
    return rowlu_apply_t(c, precision, in, out);
}

int rowlu_null_space(iemic_ctx *c, int precision, bool left, const double **vecs, const double **nrm2, int *count);
int rowlu_apply_multi(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out, size_t ldout);
int rowlu_apply_multi_t(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                        size_t ldout);

// M^{-T} on nrhs vectors (row LU only, as the single transposed apply)
int precond_apply_multi_t(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                          size_t ldout)
{
    if (c->prec_kind != IEMIC_PREC_ROW_LU)
        return set_error(IEMIC_EUNSUPPORTED, "transposed preconditioner: row LU only");
    return rowlu_apply_multi_t(c, precision, nrhs, in, ldin, out, ldout);
}

// M^{-1} on nrhs vectors (column q at in + q ldin): one batched sweep for
// the row LU, a sweep per column for the column ILU
int precond_apply_multi(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                        size_t ldout, float *tmp32)
{
    if (c->prec_kind == IEMIC_PREC_ROW_LU) return rowlu_apply_multi(c, precision, nrhs, in, ldin, out, ldout);
    for (int q = 0; q < nrhs; q++) {
        int rc = precond_apply(c, precision, in + (size_t)q * ldin, out + (size_t)q * ldout, tmp32);
        if (rc) return rc;
    }
    return IEMIC_OK;
}

// null vectors of J per pressure basin (row LU only; none for the column ILU)
int precond_null_space(iemic_ctx *c, int precision, bool left, const double **vecs, const double **nrm2, int *count)
{
    *count = 0;
    if (c->prec_kind != IEMIC_PREC_ROW_LU) return IEMIC_OK;
    return rowlu_null_space(c, precision, left, vecs, nrm2, count);
}

int rowlu_stale(iemic_ctx *c, int precision);

// 1 if the current factors were built from an older Jacobian (reuse)
int precond_stale(iemic_ctx *c, int precision)
{
    return c->prec_kind == IEMIC_PREC_ROW_LU ? rowlu_stale(c, precision) : 0;
}

int precond_check(iemic_ctx *c) { return c->prec_kind == IEMIC_PREC_ROW_LU ? IEMIC_OK : cilu_check(c); }

int rowlu_info(iemic_ctx *c, iemic_precond_info *info);

int precond_info(iemic_ctx *c, iemic_precond_info *info)
{
    *info = iemic_precond_info{};
    info->kind = c->prec_kind;
    if (c->prec_kind == IEMIC_PREC_ROW_LU) return rowlu_info(c, info);
    int rc = ensure(c);
    if (rc) return rc;
    Precond *p = c->prec;
    info->blocks = p->nactive;
    info->max_block = p->bmax;
    info->pins = p->npins;
    info->levels = p->nlevels;
    info->factor_bytes = (double)p->total * sizeof(double);
    return IEMIC_OK;
}

}  // namespace iemic
