// Preconditioner slot of Ocean::buildPreconditioner / applyPrecon
// (Ocean.C:1377-1416), where the reference plugs in the TRIOS block
// preconditioner with ML/MRILU inner solvers (TRIOS_BlockPreconditioner.C).
//
// Water-column block ILU(0).  All unknowns of one water column (i,j) form a
// dense block of size b = 6 l (local index 6 (k-1) + var), so the vertical
// hydrostatic / continuity / mixing couplings are factored exactly; between
// columns the block pattern is the horizontal 9-point neighbourhood and the
// factorisation is block ILU(0) in the natural column order:
//     for r in order:  for lower neighbours c of r (ascending):
//         L_rc = A_rc D_c^{-1};  A_rq -= L_rc U_cq  for q > c adjacent to r
//     D_r^{-1} = inverse of the updated A_rr (partial pivoting; a pivot below
//                1e-10 x the largest is replaced by a signed 1e-10 x largest).
// The periodic wrap-around couplings (i = 1 <-> i = n) are left out of the
// factors so that the column dependency graph has n + 2 m levels instead of
// n m; the outer FGMRES sees the exact J.  The dense integral-condition row is
// represented by its diagonal entry (a rank-one change the Krylov solver
// absorbs).
//
// Columns with no level dependency between them are processed by one kernel
// launch per level, one workgroup per column; blocks are stored column-major
// (element (row, col) at col * b + row) so that a wavefront reading one block
// column is unit-stride.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.hpp"

namespace iemic {

constexpr int NB9 = 9;   // neighbour slot (dj+1)*3 + (di+1)

template <typename T>
struct Factors {
    T *blk = nullptr;    // ncol * 9 * b * b
    bool built = false;
};

struct Precond {
    int b = 0, ncol = 0, nlevels = 0;
    int *d_nbr = nullptr;          // ncol*9 neighbour column or -1 (no wrap)
    int *d_order = nullptr;        // columns sorted by level
    std::vector<int> lvl_off;      // nlevels+1 offsets into order
    Factors<double> f64;
    Factors<float> f32;
    double *d_tmp64 = nullptr;     // ndim scratch for the sweeps
    float *d_tmp32 = nullptr;
    int *d_piv = nullptr;          // pivot scratch: widest level * b
    void *d_scratch = nullptr;     // block scratch: widest level * b * b doubles
    int maxw = 0;                  // widest level
};

static inline int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

// ------------------------------------------------------------------ block assembly
// Scatter the slot values of every cell into the dense column blocks.
template <typename T>
__global__ void __launch_bounds__(256) k_scatter_blocks(Scalars s, const double *__restrict__ J,
                                                        const double *__restrict__ coef, double sign,
                                                        const int *__restrict__ nbr, int b,
                                                        T *__restrict__ blk)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const int i = cell % s.n, j = (cell / s.n) % s.m, k = cell / (s.n * s.m);
    const int col = i + s.n * j;
    const size_t C = (size_t)s.ncell, bb = (size_t)b * b;
    const bool introw = (6 * cell + S == s.rowintcon);
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        const int kk = SLOTS[sl].kk, e = SLOTS[sl].eq, v = SLOTS[sl].var;
        int di = kk_di(kk), dj = kk_dj(kk), dk = kk_dk(kk);
        int nk = k + dk;
        if (nk < 0 || nk >= s.l) continue;
        int q = nbr[(size_t)col * NB9 + (dj + 1) * 3 + (di + 1)];
        if (q < 0) continue;
        double val = J[(size_t)sl * C + cell];
        if (introw && e == S) continue;
        int row = 6 * k + e, cc = 6 * nk + v;
        blk[((size_t)col * NB9 + (dj + 1) * 3 + (di + 1)) * bb + (size_t)cc * b + row] = (T)val;
    }
    if (introw) {
        double d = sign * coef[6 * (size_t)cell + S];
        int row = 6 * k + S;
        blk[((size_t)col * NB9 + 4) * bb + (size_t)row * b + row] = (T)(d != 0.0 ? d : 1.0);
    }
}

// ------------------------------------------------------------------ dense kernels (one workgroup)
// C (+)= alpha * A * B for b x b column-major blocks in global memory, tiled
// through LDS: 16 x 16 threads, each owning rows tx + 16 p and columns
// ty + 16 q (p, q < R = ceil(b/16)).
template <typename T, int R>
__device__ void wg_gemm(const T *__restrict__ A, const T *__restrict__ B, T *__restrict__ Cm, int b,
                        T alpha, bool accumulate)
{
    __shared__ T As[16 * R][16 + 1];
    __shared__ T Bs[16][16 * R + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    T acc[R][R];
#pragma unroll
    for (int p = 0; p < R; p++)
#pragma unroll
        for (int q = 0; q < R; q++) acc[p][q] = 0;
    for (int k0 = 0; k0 < b; k0 += 16) {
        // A tile: rows 0..b-1, cols k0..k0+15 ; B tile: rows k0..k0+15, cols 0..b-1
        for (int e = threadIdx.x; e < 16 * R * 16; e += blockDim.x) {
            int r = e % (16 * R), c = e / (16 * R);
            As[r][c] = (r < b && k0 + c < b) ? A[(size_t)(k0 + c) * b + r] : (T)0;
        }
        for (int e = threadIdx.x; e < 16 * 16 * R; e += blockDim.x) {
            int r = e % 16, c = e / 16;
            Bs[r][c] = (k0 + r < b && c < b) ? B[(size_t)c * b + k0 + r] : (T)0;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < 16; kk++) {
            T a[R], bv[R];
#pragma unroll
            for (int p = 0; p < R; p++) a[p] = As[tx + 16 * p][kk];
#pragma unroll
            for (int q = 0; q < R; q++) bv[q] = Bs[kk][ty + 16 * q];
#pragma unroll
            for (int p = 0; p < R; p++)
#pragma unroll
                for (int q = 0; q < R; q++) acc[p][q] += a[p] * bv[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < R; p++)
#pragma unroll
        for (int q = 0; q < R; q++) {
            int r = tx + 16 * p, c = ty + 16 * q;
            if (r < b && c < b) {
                T v = alpha * acc[p][q];
                Cm[(size_t)c * b + r] = accumulate ? Cm[(size_t)c * b + r] + v : v;
            }
        }
    __syncthreads();
}

// In-place inverse of a b x b column-major block by Gauss-Jordan elimination
// with partial pivoting (tiny pivots replaced, see file comment).
template <typename T>
__device__ void wg_invert(T *__restrict__ A, int b, int *__restrict__ piv)
{
    __shared__ T sbig;
    __shared__ int sp;
    __shared__ T red[256];
    __shared__ int redi[256];
    // largest magnitude for the pivot floor
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
    const T tiny = (T)1.0e-10 * sbig;
    for (int q = 0; q < b; q++) {
        // pivot search in column q, rows q..b-1
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
            if (fabs((double)d) < (double)tiny) d = (d < 0) ? -tiny : tiny;
            A[(size_t)q * b + q] = d;
        }
        __syncthreads();
        const T inv = (T)1 / A[(size_t)q * b + q];
        // scale pivot row (all columns), pivot entry becomes 1/d
        for (int c = threadIdx.x; c < b; c += blockDim.x)
            if (c != q) A[(size_t)c * b + q] *= inv;
        __syncthreads();
        if (threadIdx.x == 0) A[(size_t)q * b + q] = inv;
        __syncthreads();
        // eliminate column q from the other rows
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
    // undo the row interchanges as column interchanges, last first
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

// One ILU(0) row step per workgroup for the columns of one level.
template <typename T, int R>
__global__ void __launch_bounds__(256) k_cbilu_level(const int *__restrict__ cols, const int *__restrict__ nbr,
                                                     int b, T *__restrict__ blk, T *__restrict__ scratch,
                                                     int *__restrict__ pivs)
{
    const int r = cols[blockIdx.x];
    const size_t bb = (size_t)b * b;
    T *row = blk + (size_t)r * NB9 * bb;
    T *W = scratch + (size_t)blockIdx.x * bb;
    int nr[NB9];
#pragma unroll
    for (int s = 0; s < NB9; s++) nr[s] = nbr[(size_t)r * NB9 + s];
    int ord[NB9], no = 0;
    for (int s = 0; s < NB9; s++)
        if (nr[s] >= 0 && nr[s] < r) ord[no++] = s;
    for (int a = 1; a < no; a++)
        for (int z = a; z > 0 && nr[ord[z]] < nr[ord[z - 1]]; z--) {
            int t = ord[z]; ord[z] = ord[z - 1]; ord[z - 1] = t;
        }
    for (int a = 0; a < no; a++) {
        const int sc = ord[a], c = nr[sc];
        T *Arc = row + (size_t)sc * bb;
        const T *crow = blk + (size_t)c * NB9 * bb;
        for (size_t e = threadIdx.x; e < bb; e += blockDim.x) W[e] = Arc[e];
        __syncthreads();
        wg_gemm<T, R>(W, crow + 4 * bb, Arc, b, (T)1, false);   // L_rc = A_rc D_c^{-1}
        for (int sq = 0; sq < NB9; sq++) {
            int q = nbr[(size_t)c * NB9 + sq];
            if (q <= c) continue;
            int sr = -1;
            for (int s = 0; s < NB9; s++)
                if (nr[s] == q) sr = s;
            if (sr < 0) continue;
            wg_gemm<T, R>(Arc, crow + (size_t)sq * bb, row + (size_t)sr * bb, b, (T)-1, true);
        }
    }
    wg_invert<T>(row + 4 * bb, b, pivs + (size_t)blockIdx.x * b);
}

// Forward sweep of one level: y_r = v_r - sum_{c<r} L_rc y_c
// Backward sweep of one level: x_r = D_r^{-1} (y_r - sum_{q>r} U_rq x_q)
// Vectors are in the model's unknown order; column r's local unknown
// 6 k + var lives at 6 (k n m + col) + var.
template <typename T, bool BACKWARD>
__global__ void __launch_bounds__(256) k_cbilu_sweep(const int *__restrict__ cols, const int *__restrict__ nbr,
                                                     int b, int ncolumns, const T *__restrict__ blk,
                                                     const T *__restrict__ in, T *__restrict__ io)
{
    extern __shared__ unsigned char smem[];
    T *xs = reinterpret_cast<T *>(smem);        // neighbour values, b
    T *acc = xs + b;                            // accumulator, b
    const int r = cols[blockIdx.x];
    const size_t bb = (size_t)b * b;
    auto gidx = [&](int colid, int e) { return 6 * ((size_t)(e / 6) * ncolumns + colid) + (e % 6); };
    for (int e = threadIdx.x; e < b; e += blockDim.x) acc[e] = BACKWARD ? io[gidx(r, e)] : in[gidx(r, e)];
    __syncthreads();
    for (int s = 0; s < NB9; s++) {
        int c = nbr[(size_t)r * NB9 + s];
        if (c < 0 || (BACKWARD ? c <= r : c >= r)) continue;
        for (int e = threadIdx.x; e < b; e += blockDim.x) xs[e] = io[gidx(c, e)];
        __syncthreads();
        const T *B = blk + ((size_t)r * NB9 + s) * bb;
        for (int e = threadIdx.x; e < b; e += blockDim.x) {
            T v = 0;
            for (int q = 0; q < b; q++) v += B[(size_t)q * b + e] * xs[q];
            acc[e] -= v;
        }
        __syncthreads();
    }
    if (BACKWARD) {
        const T *D = blk + ((size_t)r * NB9 + 4) * bb;
        for (int e = threadIdx.x; e < b; e += blockDim.x) xs[e] = acc[e];
        __syncthreads();
        for (int e = threadIdx.x; e < b; e += blockDim.x) {
            T v = 0;
            for (int q = 0; q < b; q++) v += D[(size_t)q * b + e] * xs[q];
            acc[e] = v;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < b; e += blockDim.x) io[gidx(r, e)] = acc[e];
}

// ------------------------------------------------------------------ host
void precond_free(iemic_ctx *c)
{
    Precond *p = c->prec;
    if (!p) return;
    void *bufs[] = {p->d_nbr, p->d_order, p->f64.blk, p->f32.blk, p->d_tmp64, p->d_tmp32, p->d_piv,
                    p->d_scratch};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
    c->prec = nullptr;
}

static int ensure(iemic_ctx *c)
{
    if (c->prec) return IEMIC_OK;
    Precond *p = new Precond();
    c->prec = p;
    const int n = c->cfg.n, m = c->cfg.m;
    p->b = 6 * c->cfg.l;
    p->ncol = n * m;
    // neighbour table without periodic wrap; levels of the natural order
    std::vector<int> nbr((size_t)p->ncol * NB9, -1), level(p->ncol, 0);
    for (int j = 0; j < m; j++)
        for (int i = 0; i < n; i++) {
            int r = i + n * j, lv = 0;
            for (int dj = -1; dj <= 1; dj++)
                for (int di = -1; di <= 1; di++) {
                    int ii = i + di, jj = j + dj;
                    if (ii < 0 || ii >= n || jj < 0 || jj >= m) continue;
                    int q = ii + n * jj;
                    nbr[(size_t)r * NB9 + (dj + 1) * 3 + (di + 1)] = q;
                    if (q < r) lv = std::max(lv, level[q] + 1);
                }
            level[r] = lv;
        }
    p->nlevels = *std::max_element(level.begin(), level.end()) + 1;
    std::vector<int> order(p->ncol);
    p->lvl_off.assign(p->nlevels + 1, 0);
    for (int r = 0; r < p->ncol; r++) p->lvl_off[level[r] + 1]++;
    for (int q = 0; q < p->nlevels; q++) p->lvl_off[q + 1] += p->lvl_off[q];
    std::vector<int> fill(p->lvl_off.begin(), p->lvl_off.end() - 1);
    for (int r = 0; r < p->ncol; r++) order[fill[level[r]]++] = r;
    p->maxw = 0;
    for (int q = 0; q < p->nlevels; q++) p->maxw = std::max(p->maxw, p->lvl_off[q + 1] - p->lvl_off[q]);
    HIPCHK(hipMalloc(&p->d_nbr, nbr.size() * sizeof(int)));
    HIPCHK(hipMemcpy(p->d_nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&p->d_order, order.size() * sizeof(int)));
    HIPCHK(hipMemcpy(p->d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&p->d_piv, (size_t)p->maxw * p->b * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_scratch, (size_t)p->maxw * p->b * p->b * sizeof(double)));
    return IEMIC_OK;
}

template <typename T, int R>
static void launch_level(iemic_ctx *c, Precond *p, T *blk, int lv)
{
    int w = p->lvl_off[lv + 1] - p->lvl_off[lv];
    hipLaunchKernelGGL((k_cbilu_level<T, R>), dim3(w), dim3(256), 0, c->stream, p->d_order + p->lvl_off[lv],
                       p->d_nbr, p->b, blk, reinterpret_cast<T *>(p->d_scratch), p->d_piv);
}

template <typename T>
static int factor(iemic_ctx *c, Factors<T> &F)
{
    int rc = ensure(c);
    if (rc) return rc;
    Precond *p = c->prec;
    const size_t bb = (size_t)p->b * p->b, total = (size_t)p->ncol * NB9 * bb;
    if (!F.blk) HIPCHK(hipMalloc(&F.blk, total * sizeof(T)));
    HIPCHK(hipMemsetAsync(F.blk, 0, total * sizeof(T), c->stream));
    hipLaunchKernelGGL(k_scatter_blocks<T>, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc, c->d_J,
                       c->d_coef, (double)c->cfg.int_sign, p->d_nbr, p->b, F.blk);
    HIPCHK(hipGetLastError());
    const int R = (p->b + 15) / 16;
    for (int lv = 0; lv < p->nlevels; lv++) {
        switch (R) {
        case 1: launch_level<T, 1>(c, p, F.blk, lv); break;
        case 2: launch_level<T, 2>(c, p, F.blk, lv); break;
        case 3: launch_level<T, 3>(c, p, F.blk, lv); break;
        case 4: launch_level<T, 4>(c, p, F.blk, lv); break;
        case 5: launch_level<T, 5>(c, p, F.blk, lv); break;
        case 6: launch_level<T, 6>(c, p, F.blk, lv); break;
        case 7: launch_level<T, 7>(c, p, F.blk, lv); break;
        case 8: launch_level<T, 8>(c, p, F.blk, lv); break;
        case 9: launch_level<T, 9>(c, p, F.blk, lv); break;
        case 10: launch_level<T, 10>(c, p, F.blk, lv); break;
        default: return set_error(IEMIC_EUNSUPPORTED, "column preconditioner supports at most 26 levels");
        }
    }
    HIPCHK(hipGetLastError());
    F.built = true;
    return IEMIC_OK;
}

template <typename T>
static int sweep(iemic_ctx *c, const Factors<T> &F, const T *in, T *out)
{
    Precond *p = c->prec;
    const size_t sh = 2 * (size_t)p->b * sizeof(T);
    for (int lv = 0; lv < p->nlevels; lv++) {
        int w = p->lvl_off[lv + 1] - p->lvl_off[lv];
        hipLaunchKernelGGL((k_cbilu_sweep<T, false>), dim3(w), dim3(256), sh, c->stream,
                           p->d_order + p->lvl_off[lv], p->d_nbr, p->b, p->ncol, F.blk, in, out);
    }
    for (int lv = p->nlevels - 1; lv >= 0; lv--) {
        int w = p->lvl_off[lv + 1] - p->lvl_off[lv];
        hipLaunchKernelGGL((k_cbilu_sweep<T, true>), dim3(w), dim3(256), sh, c->stream,
                           p->d_order + p->lvl_off[lv], p->d_nbr, p->b, p->ncol, F.blk, in, out);
    }
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

int precond_build(iemic_ctx *c)
{
    int rc = ensure(c);
    if (rc) return rc;
    timed_begin_ext(c, 4);
    rc = factor<double>(c, c->prec->f64);
    timed_end_ext(c, 4);
    return rc;
}

int precond_build_f32(iemic_ctx *c)
{
    int rc = ensure(c);
    if (rc) return rc;
    timed_begin_ext(c, 5);
    rc = factor<float>(c, c->prec->f32);
    timed_end_ext(c, 5);
    return rc;
}

int precond_apply(iemic_ctx *c, const double *d_in, double *d_out)
{
    if (!c->prec || !c->prec->f64.built) {
        int rc = precond_build(c);
        if (rc) return rc;
    }
    timed_begin_ext(c, 3);
    int rc = sweep<double>(c, c->prec->f64, d_in, d_out);
    timed_end_ext(c, 3);
    return rc;
}

int precond_apply_f32(iemic_ctx *c, const float *d_in, float *d_out)
{
    if (!c->prec || !c->prec->f32.built) {
        int rc = precond_build_f32(c);
        if (rc) return rc;
    }
    return sweep<float>(c, c->prec->f32, d_in, d_out);
}

}  // namespace iemic
