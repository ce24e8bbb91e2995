// Row-block LU preconditioner: the default behind Ocean::buildPreconditioner
// / applyPrecon (Ocean.C:1377-1416), in the slot where the reference uses the
// TRIOS block preconditioner (TRIOS_BlockPreconditioner.C).
//
// The ocean unknowns of one zonal row j (all water columns i at that
// latitude, all their wet levels, all six variables) form one dense block of
// size n_j; local index = base(i) + 6 (k - k0(i)) + var.  The 27-point
// stencil couples a row only to rows j - 1 and j + 1, so J restricted to the
// ocean unknowns is block tridiagonal in j (the periodic wrap in i stays
// inside the row blocks) and factors exactly by block elimination:
//     D_0 = A_00,   X_j = D_{j-1}^{-1} A_{j-1,j},   D_j = A_jj - A_{j,j-1} X_j,
//     Minv_j = D_j^{-1}  (LU with partial pivoting, then D_j Minv_j = I solved
//                         with the LU factors).
// The couplings A_{j,j-1}, A_{j-1,j} stay sparse (at most the stencil's
// entries per unknown, read from the Jacobian slots through precomputed
// index lists); Minv_j and X_j are dense, column-major.  Applying M^{-1} is a
// forward sweep z_j = Minv_j (b_j - A_{j,j-1} z_{j-1}) and a backward sweep
// z_{j-1} -= X_j z_j, i.e. 2 m - 1 dense matrix-vector products.
//
// The factored matrix differs from J in three ways, all absorbed by the
// outer FGMRES: land rows are identity rows of J and pass through unchanged;
// the dense integral-condition row is represented by its diagonal entry; and
// one continuity row per pressure basin is replaced by p = 0 (pressure_pins),
// which removes the pressure null vectors of J from the factors.
//
// The factored matrix is equilibrated first, R M C with power-of-two row and
// column scales (exact in floating point), and M^{-1} = C (R M C)^{-1} R.
//
// Dense LU is the tournament-pivoted dense_getrf of dense_lu.hip (rocSOLVER
// getrf + getrs with IEMIC_DENSE_LU=0); the inverse comes from two rocBLAS
// triangular solves; the block assembly, the sparse Schur-complement updates
// and the row gathers are kernels of this file.  The FP32 factors are computed
// in FP32 (from the FP64 slot values) for the mixed-precision solve.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "ctx.hpp"
#include "dense_lu.hpp"

namespace iemic {

std::vector<int> column_tops(const iemic_ctx *c);
template <typename T>
hipError_t dense_mm(hipStream_t st, int m, int k, int q, T alpha, const T *A, size_t lda, const T *X, size_t ldx,
                    T beta, T *Y, size_t ldy, T *work, size_t work_elems, int blocks);
size_t dense_mm_work(int m);
template <typename T>
hipError_t dense_mm_t(hipStream_t st, int m, int k, int q, T alpha, const T *A, size_t lda, const T *X, size_t ldx,
                      T beta, T *Y, size_t ldy);
std::vector<uint8_t> pressure_pins(const iemic_ctx *c, const std::vector<int> &k0, const std::vector<int> &rank);

template <typename T>
struct RowFactors {
    T *blk = nullptr;        // Minv_j and X_j
    unsigned version = 0;    // Jacobian version the factors belong to
    bool built = false;
};

struct RowLU {
    int m = 0, nmax = 0, nocean = 0, npins = 0;
    std::vector<int> nj, ubase;          // unknowns per row, first ocean unknown of the row
    std::vector<size_t> doff, xoff;      // element offsets of Minv_j and X_j (j >= 1)
    size_t total = 0;
    int *d_loc = nullptr;                // per cell: local index of its first unknown in its row, -1 on land
    int *d_perm = nullptr;               // per ocean unknown: model index 6 cell + var
    int *d_lptr = nullptr;               // A_{j,j-1} by rows of row j (CSR over ocean unknowns)
    int2 *d_lent = nullptr;              //   {local column in row j-1, slot-value index}
    int *d_cptr = nullptr;               // A_{j+1,j} by columns of row j (transposed sweep)
    int2 *d_cent = nullptr;              //   {local row in row j+1, index of the entry in lent/lval}
    int *d_uptr = nullptr;               // A_{j-1,j} by columns in row j
    int2 *d_uent = nullptr;              //   {local row in row j-1, slot-value index}
    uint8_t *d_pin = nullptr;            // per cell: continuity row replaced by p = 0
    double *d_rs = nullptr, *d_cs = nullptr;   // ndim row / column scales (powers of two)
    double *d_lval = nullptr, *d_uval = nullptr;  // scaled coupling values of lent / uent
    int nl = 0, nu = 0;
    RowFactors<double> f64;
    RowFactors<float> f32;
    int *d_ipiv = nullptr, *d_info = nullptr;
    void *d_luw = nullptr;               // dense_getrf workspace
    int *d_rowperm = nullptr;            // its row permutation (nmax)
    void *d_v = nullptr, *d_t = nullptr;  // row-layout work vectors (nocean, nmax) in the widest type
    void *d_vm = nullptr, *d_tm = nullptr; // the same for mcap right-hand sides (multi-vector sweeps)
    int mcap = 0;
    void *d_w = nullptr;                  // nmax^2 scratch: right-hand sides of the inversion
    rocblas_handle blas = nullptr;
    LuAhead ahead;                        // look-ahead stream, handle and events of dense_getrf
    bool lookahead = false;               // true (IEMIC_LU_LOOKAHEAD=1): panels one block ahead on a second stream
    // null vectors of J, one per pressure basin b (pinned row r_b):
    // left m_b = M^{-T} e_{r_b}, right n_b = M^{-1} e_{r_b}
    std::vector<int> pin_rows;
    int *d_pin_rows = nullptr;
    double *d_null[2] = {nullptr, nullptr};   // [0] left, [1] right: npins x ndim
    std::vector<double> null_nrm2[2];
    unsigned null_version[2] = {0, 0};
    int null_prec[2] = {-1, -1};
    double *d_unit = nullptr;
    bool dense_lu = true;                 // tournament-pivoted dense_getrf; false: rocSOLVER getrf + getrs
    bool blas_apply = false;              // rocBLAS gemv/gemm in the sweeps instead of dense_mm (comparison runs)
    void *d_mmw = nullptr;                // dense_mm workspace, dense_mm_work(nmax) doubles
    std::map<int, int> mm_blocks;         // tuned dense_mm block target per number of right-hand sides
};

static inline int nblk(size_t n, int b) { return (int)((n + b - 1) / b); }

// at most this many pressure basins get explicit null vectors (two vectors
// of ndim doubles each)
constexpr int ROWLU_NULL_MAX = 64;

struct RowGeom {
    const int *loc;
    int n, m, l, periodic;
};

// ------------------------------------------------------------------ assembly
template <typename T>
__global__ void k_fill_t(T *__restrict__ a, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        a[q] = (T)0;
}

template <typename T>
__global__ void k_identity(T *__restrict__ W, int n)
{
    const size_t nn = (size_t)n * n;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (size_t)gridDim.x * blockDim.x)
        W[q] = (q % ((size_t)n + 1) == 0) ? (T)1 : (T)0;
}

// ------------------------------------------------------------------ scaling
// The factored matrix is R M C with diagonal R, C of powers of two (exact in
// floating point): R equilibrates the rows, C the columns of R M.  This keeps
// the FP32 factors in range and well scaled; M^{-1} = C (R M C)^{-1} R.
__device__ __forceinline__ double pow2_inv(double mx) { return mx > 0.0 ? ldexp(1.0, -ilogb(mx)) : 1.0; }

// target cell of slot sl of cell (i,j,k) in the row numbering, -1 if none
__device__ __forceinline__ int slot_target(const RowGeom &G, int i, int j, int k, int kk)
{
    int ni = i + kk_di(kk);
    const int nj = j + kk_dj(kk), nk = k + kk_dk(kk);
    if (G.periodic) ni = (ni + G.n) % G.n;
    if (ni < 0 || ni >= G.n || nj < 0 || nj >= G.m || nk < 0 || nk >= G.l) return -1;
    const int t = ni + G.n * (nj + G.m * nk);
    return G.loc[t] >= 0 ? t : -1;
}

__global__ void __launch_bounds__(256) k_row_scale(Scalars s, RowGeom G, const double *__restrict__ J,
                                                   const double *__restrict__ coef, double sign,
                                                   const uint8_t *__restrict__ pin, double *__restrict__ rs)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    if (G.loc[cell] < 0) {
        for (int e = 0; e < 6; e++) rs[6 * (size_t)cell + e] = 1.0;
        return;
    }
    const int i = cell % G.n, j = (cell / G.n) % G.m, k = cell / (G.n * G.m);
    const size_t C = (size_t)s.ncell;
    double mx[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++)
        if (slot_target(G, i, j, k, SLOTS[sl].kk) >= 0)
            mx[SLOTS[sl].eq] = fmax(mx[SLOTS[sl].eq], fabs(J[(size_t)sl * C + cell]));
    if (pin[cell]) mx[P] = 1.0;
    if (6 * cell + S == s.rowintcon) {
        const double d = sign * coef[6 * (size_t)cell + S];
        mx[S] = d != 0.0 ? fabs(d) : 1.0;
    }
    for (int e = 0; e < 6; e++) rs[6 * (size_t)cell + e] = pow2_inv(mx[e]);
}

// column scales of R M: each cell gathers the entries of its six columns
// from the cells whose slots reference it
__global__ void __launch_bounds__(256) k_col_scale(Scalars s, RowGeom G, const double *__restrict__ J,
                                                   const double *__restrict__ coef, double sign,
                                                   const uint8_t *__restrict__ pin, const double *__restrict__ rs,
                                                   double *__restrict__ cs)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    if (G.loc[cell] < 0) {
        for (int e = 0; e < 6; e++) cs[6 * (size_t)cell + e] = 1.0;
        return;
    }
    const int i = cell % G.n, j = (cell / G.n) % G.m, k = cell / (G.n * G.m);
    const size_t C = (size_t)s.ncell;
    double mx[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        const int kk = SLOTS[sl].kk;
        // the owner of this entry sits at -offset(kk) from this column cell
        int oi = i - kk_di(kk);
        const int oj = j - kk_dj(kk), ok = k - kk_dk(kk);
        if (G.periodic) oi = (oi + G.n) % G.n;
        if (oi < 0 || oi >= G.n || oj < 0 || oj >= G.m || ok < 0 || ok >= G.l) continue;
        const int owner = oi + G.n * (oj + G.m * ok);
        if (G.loc[owner] < 0) continue;
        const int e = SLOTS[sl].eq;
        if ((pin[owner] && e == P) || (6 * owner + S == s.rowintcon && e == S)) continue;
        mx[SLOTS[sl].var] = fmax(mx[SLOTS[sl].var], fabs(J[(size_t)sl * C + owner]) * rs[6 * (size_t)owner + e]);
    }
    if (pin[cell]) mx[P] = fmax(mx[P], rs[6 * (size_t)cell + P]);
    if (6 * cell + S == s.rowintcon) {
        const double d = sign * coef[6 * (size_t)cell + S];
        mx[S] = fmax(mx[S], (d != 0.0 ? fabs(d) : 1.0) * rs[6 * (size_t)cell + S]);
    }
    for (int e = 0; e < 6; e++) cs[6 * (size_t)cell + e] = pow2_inv(mx[e]);
}

// scaled values of coupling entries {other-row local index, slot index}:
// val = J * rs(row) * cs(col); row = the entry's own unknown, col its target
__global__ void k_scale_entries(int nent, const int2 *__restrict__ ent, Scalars s, RowGeom G,
                                const double *__restrict__ J, const double *__restrict__ rs,
                                const double *__restrict__ cs, double *__restrict__ val)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nent) return;
    const int vidx = ent[q].y;
    const int sl = vidx / s.ncell, cell = vidx % s.ncell;
    const int i = cell % G.n, j = (cell / G.n) % G.m, k = cell / (G.n * G.m);
    const int t = slot_target(G, i, j, k, SLOTS[sl].kk);
    val[q] = t < 0 ? 0.0 : J[vidx] * rs[6 * (size_t)cell + SLOTS[sl].eq] * cs[6 * (size_t)t + SLOTS[sl].var];
}

// D_j = A_jj: the slots of the cells of row j whose target lies in row j.
template <typename T>
__global__ void __launch_bounds__(256) k_row_block(Scalars s, RowGeom G, int j, int nrow, const double *__restrict__ J,
                                                   const double *__restrict__ coef, double sign,
                                                   const uint8_t *__restrict__ pin, const double *__restrict__ rs,
                                                   const double *__restrict__ cs, T *__restrict__ D)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G.n * G.l) return;
    const int i = t % G.n, k = t / G.n;
    const int cell = i + G.n * (j + G.m * k);
    const int lr = G.loc[cell];
    if (lr < 0) return;
    const size_t C = (size_t)s.ncell;
    const bool introw = (6 * cell + S == s.rowintcon), pinned = pin[cell] != 0;
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        const int kk = SLOTS[sl].kk, e = SLOTS[sl].eq, v = SLOTS[sl].var;
        if (kk_dj(kk) != 0) continue;
        if ((introw && e == S) || (pinned && e == P)) continue;
        const int nk = k + kk_dk(kk);
        int ni = i + kk_di(kk);
        if (G.periodic) ni = (ni + G.n) % G.n;
        if (nk < 0 || nk >= G.l || ni < 0 || ni >= G.n) continue;
        const int tc = ni + G.n * (j + G.m * nk);
        const int lc = G.loc[tc];
        if (lc < 0) continue;
        D[(size_t)(lc + v) * nrow + lr + e] =
            (T)(J[(size_t)sl * C + cell] * rs[6 * (size_t)cell + e] * cs[6 * (size_t)tc + v]);
    }
    if (introw) {
        const double d = sign * coef[6 * (size_t)cell + S];
        D[(size_t)(lr + S) * nrow + lr + S] =
            (T)((d != 0.0 ? d : 1.0) * rs[6 * (size_t)cell + S] * cs[6 * (size_t)cell + S]);
    }
    if (pinned) D[(size_t)(lr + P) * nrow + lr + P] = (T)(rs[6 * (size_t)cell + P] * cs[6 * (size_t)cell + P]);
}

// X_j(:, c) = sum over the entries (r, a) of column c of A_{j-1,j}: a Minv_{j-1}(:, r)
template <typename T>
__global__ void __launch_bounds__(256) k_row_x(int nprev, const int *__restrict__ uptr, const int2 *__restrict__ uent,
                                               const double *__restrict__ uval, const T *__restrict__ Minv,
                                               T *__restrict__ X)
{
    const int c = blockIdx.x;
    const int q0 = uptr[c], q1 = uptr[c + 1];
    T *out = X + (size_t)c * nprev;
    for (int e = threadIdx.x; e < nprev; e += blockDim.x) {
        T acc = 0;
        int q = q0;
        // four entries' loads in flight at once, summed in the same order
        for (; q + 4 <= q1; q += 4) {
            const int2 u0 = uent[q], u1 = uent[q + 1], u2 = uent[q + 2], u3 = uent[q + 3];
            const T m0 = Minv[(size_t)u0.x * nprev + e], m1 = Minv[(size_t)u1.x * nprev + e];
            const T m2 = Minv[(size_t)u2.x * nprev + e], m3 = Minv[(size_t)u3.x * nprev + e];
            acc += (T)uval[q] * m0;
            acc += (T)uval[q + 1] * m1;
            acc += (T)uval[q + 2] * m2;
            acc += (T)uval[q + 3] * m3;
        }
        for (; q < q1; q++) acc += (T)uval[q] * Minv[(size_t)uent[q].x * nprev + e];
        out[e] = acc;
    }
}

// D_j(:, c) -= A_{j,j-1} X_j(:, c)
template <typename T>
__global__ void __launch_bounds__(256) k_row_schur(int nrow, int nprev, const int *__restrict__ lptr,
                                                   const int2 *__restrict__ lent, const double *__restrict__ lval,
                                                   const T *__restrict__ X, T *__restrict__ D)
{
    const int c = blockIdx.x;
    const T *xc = X + (size_t)c * nprev;
    T *dc = D + (size_t)c * nrow;
    for (int e = threadIdx.x; e < nrow; e += blockDim.x) {
        T acc = 0;
        int q = lptr[e];
        const int q1 = lptr[e + 1];
        // four entries' index loads, then their four gathers, in flight at
        // once (one chain of dependent loads per entry before); summed in
        // the same order
        for (; q + 4 <= q1; q += 4) {
            const int2 l0 = lent[q], l1 = lent[q + 1], l2 = lent[q + 2], l3 = lent[q + 3];
            const T x0 = xc[l0.x], x1 = xc[l1.x], x2 = xc[l2.x], x3 = xc[l3.x];
            acc += (T)lval[q] * x0;
            acc += (T)lval[q + 1] * x1;
            acc += (T)lval[q + 2] * x2;
            acc += (T)lval[q + 3] * x3;
        }
        for (; q < q1; q++) acc += (T)lval[q] * xc[lent[q].x];
        dc[e] -= acc;
    }
}

// ------------------------------------------------------------------ apply
template <typename T>
__global__ void k_row_gather(int nocean, const int *__restrict__ perm, const double *__restrict__ rs,
                             const double *__restrict__ in, T *__restrict__ v)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nocean) v[u] = (T)(rs[perm[u]] * in[perm[u]]);
}

template <typename T>
__global__ void k_row_scatter(int nocean, const int *__restrict__ perm, const double *__restrict__ cs,
                              const T *__restrict__ v, double *__restrict__ out)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < nocean) out[perm[u]] = cs[perm[u]] * (double)v[u];
}

__global__ void k_row_land(int ncell, const int *__restrict__ loc, const double *__restrict__ in, double *__restrict__ out)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell || loc[cell] >= 0) return;
#pragma unroll
    for (int v = 0; v < 6; v++) out[6 * (size_t)cell + v] = in[6 * (size_t)cell + v];
}

// t = v_j - A_{j,j-1} z_{j-1}
template <typename T>
__global__ void k_row_lmv(int nrow, const int *__restrict__ lptr, const int2 *__restrict__ lent,
                          const double *__restrict__ lval, const T *__restrict__ vj, const T *__restrict__ zprev,
                          T *__restrict__ t)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrow) return;
    T acc = vj[e];
    for (int q = lptr[e]; q < lptr[e + 1]; q++) acc -= (T)lval[q] * zprev[lent[q].x];
    t[e] = acc;
}

// t = y_j - A_{j+1,j}^T x_{j+1}  (entries of A_{j+1,j} listed by column of row j)
template <typename T>
__global__ void k_row_ltmv(int nrow, const int *__restrict__ cptr, const int2 *__restrict__ cent,
                           const double *__restrict__ lval, const T *__restrict__ yj, const T *__restrict__ xnext,
                           T *__restrict__ t)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrow) return;
    T acc = yj[e];
    for (int q = cptr[e]; q < cptr[e + 1]; q++) acc -= (T)lval[cent[q].y] * xnext[cent[q].x];
    t[e] = acc;
}

// ------------------------------------------------------------------ host
#define BLASCHK(expr)                                                                  \
    do {                                                                               \
        rocblas_status st_ = (expr);                                                   \
        if (st_ != rocblas_status_success)                                             \
            return set_error(IEMIC_EHIP, std::string(#expr) + ": rocBLAS/rocSOLVER status " + \
                                             std::to_string((int)st_));                \
    } while (0)

void rowlu_free(iemic_ctx *c)
{
    RowLU *p = c->rowlu;
    if (!p) return;
    void *bufs[] = {p->d_loc, p->d_perm, p->d_lptr, p->d_lent, p->d_cptr, p->d_cent, p->d_uptr, p->d_uent, p->d_pin,
                    p->d_rs, p->d_cs, p->d_lval, p->d_uval,
                    p->f64.blk, p->f32.blk, p->d_ipiv, p->d_info, p->d_v, p->d_t, p->d_w,
                    p->d_null[0], p->d_null[1], p->d_unit, p->d_vm, p->d_tm, p->d_mmw, p->d_pin_rows, p->d_luw, p->d_rowperm};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (p->blas) rocblas_destroy_handle(p->blas);
    if (p->ahead.h2) rocblas_destroy_handle(p->ahead.h2);
    if (p->ahead.ev_panel) (void)hipEventDestroy(p->ahead.ev_panel);
    if (p->ahead.ev_rest) (void)hipEventDestroy(p->ahead.ev_rest);
    if (p->ahead.s2) (void)hipStreamDestroy(p->ahead.s2);
    delete p;
    c->rowlu = nullptr;
}

static int rowlu_ensure(iemic_ctx *c)
{
    if (c->rowlu) return IEMIC_OK;
    const int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    const size_t C = (size_t)n * m * l;
    if ((size_t)NSLOT * C >= (size_t)1 << 31)
        return set_error(IEMIC_EUNSUPPORTED, "row preconditioner: grid too large for 32-bit slot indices");
    RowLU *p = new RowLU();
    c->rowlu = p;
    p->m = m;
    const std::vector<int> k0 = column_tops(c);
    std::vector<int> rank((size_t)n * m);
    for (int q = 0; q < n * m; q++) rank[q] = q;   // rows eliminated in j order
    const std::vector<uint8_t> pin = pressure_pins(c, k0, rank);
    for (size_t q = 0; q < C; q++)
        if (pin[q]) p->pin_rows.push_back(6 * (int)q + P);
    p->npins = (int)p->pin_rows.size();
    // local numbering within rows and the ocean-unknown permutation
    std::vector<int> loc(C, -1), perm;
    p->nj.assign(m, 0);
    p->ubase.assign(m + 1, 0);
    for (int j = 0; j < m; j++) {
        int base = 0;
        for (int i = 0; i < n; i++)
            for (int k = k0[i + n * j]; k < l; k++) {
                const int cell = i + n * (j + m * k);
                loc[cell] = base;
                for (int v = 0; v < 6; v++) perm.push_back(6 * cell + v);
                base += 6;
            }
        p->nj[j] = base;
        p->ubase[j + 1] = p->ubase[j] + base;
        p->nmax = std::max(p->nmax, base);
    }
    p->nocean = p->ubase[m];
    p->doff.assign(m, 0);
    p->xoff.assign(m, 0);
    size_t tot = 0;
    for (int j = 0; j < m; j++) {
        p->doff[j] = tot;
        tot += (size_t)p->nj[j] * p->nj[j];
        if (j > 0) {
            p->xoff[j] = tot;
            tot += (size_t)p->nj[j - 1] * p->nj[j];
        }
    }
    p->total = tot;
    // couplings between neighbouring rows, by rows of A_{j,j-1} and by columns of A_{j-1,j}
    const int ric = c->sc.rowintcon;
    std::vector<std::vector<int2>> lrow(p->nocean), ucol(p->nocean);
    for (int k = 0; k < l; k++)
        for (int j = 0; j < m; j++)
            for (int i = 0; i < n; i++) {
                const int cell = i + n * (j + m * k);
                if (loc[cell] < 0) continue;
                const bool introw = (6 * cell + S == ric), pinned = pin[cell] != 0;
                for (int sl = 0; sl < NSLOT; sl++) {
                    const SlotDef &sd = SLOTS[sl];
                    const int dj = kk_dj(sd.kk);
                    if (dj == 0) continue;
                    if ((introw && sd.eq == S) || (pinned && sd.eq == P)) continue;
                    const int nj2 = j + dj, nk = k + kk_dk(sd.kk);
                    int ni = i + kk_di(sd.kk);
                    if (c->cfg.periodic) ni = (ni + n) % n;
                    if (nj2 < 0 || nj2 >= m || nk < 0 || nk >= l || ni < 0 || ni >= n) continue;
                    const int tcell = ni + n * (nj2 + m * nk);
                    if (loc[tcell] < 0) continue;
                    const int vidx = (int)((size_t)sl * C + cell);
                    const int r_local = loc[cell] + sd.eq, c_local = loc[tcell] + sd.var;
                    if (dj < 0)   // A_{j,j-1}: row in row j, column in row j-1
                        lrow[p->ubase[j] + r_local].push_back(make_int2(c_local, vidx));
                    else          // A_{j,j+1} = A_{(j+1)-1, j+1}: listed by its column in row j+1
                        ucol[p->ubase[nj2] + c_local].push_back(make_int2(r_local, vidx));
                }
            }
    // per-row CSR pointers are relative to the row (lptr_j[e] indexes the global entry array)
    std::vector<int> lptr(p->nocean + 1, 0), uptr(p->nocean + 1, 0);
    std::vector<int2> lent, uent;
    for (int u = 0; u < p->nocean; u++) {
        lptr[u] = (int)lent.size();
        lent.insert(lent.end(), lrow[u].begin(), lrow[u].end());
        uptr[u] = (int)uent.size();
        uent.insert(uent.end(), ucol[u].begin(), ucol[u].end());
    }
    lptr[p->nocean] = (int)lent.size();
    uptr[p->nocean] = (int)uent.size();
    // A_{j+1,j} by columns of row j, pointing at the lent entries
    std::vector<std::vector<int2>> lcol(p->nocean);
    for (int j = 1; j < m; j++)
        for (int e = 0; e < p->nj[j]; e++) {
            const int u = p->ubase[j] + e;
            for (int q = lptr[u]; q < lptr[u + 1]; q++) lcol[p->ubase[j - 1] + lent[q].x].push_back(make_int2(e, q));
        }
    std::vector<int> cptr(p->nocean + 1, 0);
    std::vector<int2> cent;
    for (int u = 0; u < p->nocean; u++) {
        cptr[u] = (int)cent.size();
        cent.insert(cent.end(), lcol[u].begin(), lcol[u].end());
    }
    cptr[p->nocean] = (int)cent.size();
    if (cent.empty()) cent.push_back(make_int2(0, 0));
    if (lent.empty()) lent.push_back(make_int2(0, 0));
    if (uent.empty()) uent.push_back(make_int2(0, 0));
    auto up = [&](auto **dst, const auto &v) -> int {
        HIPCHK(hipMalloc((void **)dst, v.size() * sizeof(v[0])));
        HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
        return IEMIC_OK;
    };
    int rc;
    if ((rc = up(&p->d_loc, loc)) || (rc = up(&p->d_perm, perm)) || (rc = up(&p->d_lptr, lptr)) ||
        (rc = up(&p->d_lent, lent)) || (rc = up(&p->d_uptr, uptr)) || (rc = up(&p->d_uent, uent)) ||
        (rc = up(&p->d_cptr, cptr)) || (rc = up(&p->d_cent, cent)) ||
        (rc = up(&p->d_pin, pin)) || (!p->pin_rows.empty() && (rc = up(&p->d_pin_rows, p->pin_rows))))
        return rc;
    p->nl = (int)lent.size();
    p->nu = (int)uent.size();
    HIPCHK(hipMalloc(&p->d_rs, (size_t)c->sc.ndim * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_cs, (size_t)c->sc.ndim * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_lval, (size_t)p->nl * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_uval, (size_t)p->nu * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_ipiv, (size_t)std::max(p->nmax, 1) * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_info, (size_t)m * sizeof(int)));
    HIPCHK(hipMalloc(&p->d_v, (size_t)std::max(p->nocean, 1) * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_t, (size_t)std::max(p->nmax, 1) * sizeof(double)));
    if (const char *v = std::getenv("IEMIC_DENSE_LU")) p->dense_lu = std::atoi(v) != 0;
    if (const char *v = std::getenv("IEMIC_BLAS_APPLY")) p->blas_apply = std::atoi(v) != 0;
    HIPCHK(hipMalloc(&p->d_mmw, dense_mm_work(p->nmax) * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_luw, dense_getrf_work_bytes(p->nmax)));
    HIPCHK(hipMalloc(&p->d_rowperm, (size_t)std::max(p->nmax, 1) * sizeof(int)));
    BLASCHK(rocblas_create_handle(&p->blas));
    BLASCHK(rocblas_set_stream(p->blas, c->stream));
    // size the handle's workspace once for the largest row, so that the
    // factorisation makes no allocation of its own per call
    if (p->nmax > 0) {
        size_t ws = 0;
        double *dd = reinterpret_cast<double *>(p->d_v);
        float *df = reinterpret_cast<float *>(p->d_v);
        BLASCHK(rocblas_start_device_memory_size_query(p->blas));
        rocsolver_dgetrf(p->blas, p->nmax, p->nmax, dd, p->nmax, p->d_ipiv, p->d_info);
        rocsolver_dgetrs(p->blas, rocblas_operation_none, p->nmax, p->nmax, dd, p->nmax, p->d_ipiv, dd, p->nmax);
        rocsolver_sgetrf(p->blas, p->nmax, p->nmax, df, p->nmax, p->d_ipiv, p->d_info);
        rocsolver_sgetrs(p->blas, rocblas_operation_none, p->nmax, p->nmax, df, p->nmax, p->d_ipiv, df, p->nmax);
        dense_getrf_size_query(p->blas, p->nmax);
        BLASCHK(rocblas_stop_device_memory_size_query(p->blas, &ws));
        BLASCHK(rocblas_set_device_memory_size(p->blas, ws));
        // the look-ahead stream of the dense LU: its own handle, sized the
        // same way (the two streams' TRSM and GEMM may run at once)
        if (const char *v = std::getenv("IEMIC_LU_LOOKAHEAD")) p->lookahead = std::atoi(v) != 0;
        if (p->lookahead) {
            size_t ws2 = 0;
            HIPCHK(hipStreamCreateWithFlags(&p->ahead.s2, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&p->ahead.ev_panel, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&p->ahead.ev_rest, hipEventDisableTiming));
            BLASCHK(rocblas_create_handle(&p->ahead.h2));
            BLASCHK(rocblas_set_stream(p->ahead.h2, p->ahead.s2));
            BLASCHK(rocblas_start_device_memory_size_query(p->ahead.h2));
            dense_getrf_size_query(p->ahead.h2, p->nmax);
            BLASCHK(rocblas_stop_device_memory_size_query(p->ahead.h2, &ws2));
            BLASCHK(rocblas_set_device_memory_size(p->ahead.h2, ws2));
        }
    }
    return IEMIC_OK;
}

// typed rocSOLVER / rocBLAS entry points
static rocblas_status getrf(rocblas_handle h, int n, double *A, int *ipiv, int *info)
{
    return rocsolver_dgetrf(h, n, n, A, n, ipiv, info);
}
static rocblas_status getrf(rocblas_handle h, int n, float *A, int *ipiv, int *info)
{
    return rocsolver_sgetrf(h, n, n, A, n, ipiv, info);
}
// B = A^{-1} B from the LU factors of A (two triangular solves with n
// right-hand sides, BLAS-3 shaped)
static rocblas_status getrs(rocblas_handle h, int n, const double *A, const int *ipiv, double *B)
{
    return rocsolver_dgetrs(h, rocblas_operation_none, n, n, const_cast<double *>(A), n, ipiv, B, n);
}
static rocblas_status getrs(rocblas_handle h, int n, const float *A, const int *ipiv, float *B)
{
    return rocsolver_sgetrs(h, rocblas_operation_none, n, n, const_cast<float *>(A), n, ipiv, B, n);
}
static rocblas_status gemv(rocblas_handle h, int rows, int cols, double alpha, const double *A, const double *x,
                           double beta, double *y)
{
    return rocblas_dgemv(h, rocblas_operation_none, rows, cols, &alpha, A, rows, x, 1, &beta, y, 1);
}
static rocblas_status gemv(rocblas_handle h, int rows, int cols, float alpha, const float *A, const float *x,
                           float beta, float *y)
{
    float a = alpha, b = beta;
    return rocblas_sgemv(h, rocblas_operation_none, rows, cols, &a, A, rows, x, 1, &b, y, 1);
}

static rocblas_status gemv_t(rocblas_handle h, int rows, int cols, double alpha, const double *A, const double *x,
                             double beta, double *y)
{
    return rocblas_dgemv(h, rocblas_operation_transpose, rows, cols, &alpha, A, rows, x, 1, &beta, y, 1);
}
static rocblas_status gemv_t(rocblas_handle h, int rows, int cols, float alpha, const float *A, const float *x,
                             float beta, float *y)
{
    return rocblas_sgemv(h, rocblas_operation_transpose, rows, cols, &alpha, A, rows, x, 1, &beta, y, 1);
}

template <typename T>
static int rowlu_factor(iemic_ctx *c, RowFactors<T> &F)
{
    RowLU *p = c->rowlu;
    if (!F.blk) HIPCHK(hipMalloc(&F.blk, std::max<size_t>(p->total, 1) * sizeof(T)));
    if (!p->d_w) HIPCHK(hipMalloc(&p->d_w, std::max<size_t>((size_t)p->nmax * p->nmax, 1) * sizeof(double)));
    const RowGeom G{p->d_loc, c->cfg.n, c->cfg.m, c->cfg.l, c->cfg.periodic};
    BLASCHK(rocblas_set_pointer_mode(p->blas, rocblas_pointer_mode_host));
    if (p->ahead.h2) BLASCHK(rocblas_set_pointer_mode(p->ahead.h2, rocblas_pointer_mode_host));
    const double sign = (double)c->cfg.int_sign;
    const int nbc = nblk(c->sc.ncell, 256);
    hipLaunchKernelGGL(k_row_scale, dim3(nbc), dim3(256), 0, c->stream, c->sc, G, c->d_J, c->d_coef, sign, p->d_pin,
                       p->d_rs);
    hipLaunchKernelGGL(k_col_scale, dim3(nbc), dim3(256), 0, c->stream, c->sc, G, c->d_J, c->d_coef, sign, p->d_pin,
                       p->d_rs, p->d_cs);
    hipLaunchKernelGGL(k_scale_entries, dim3(nblk(p->nl, 256)), dim3(256), 0, c->stream, p->nl, p->d_lent, c->sc, G,
                       c->d_J, p->d_rs, p->d_cs, p->d_lval);
    hipLaunchKernelGGL(k_scale_entries, dim3(nblk(p->nu, 256)), dim3(256), 0, c->stream, p->nu, p->d_uent, c->sc, G,
                       c->d_J, p->d_rs, p->d_cs, p->d_uval);
    for (int j = 0; j < p->m; j++) {
        const int nr = p->nj[j];
        if (nr == 0) continue;
        T *D = F.blk + p->doff[j];
        hipLaunchKernelGGL(k_fill_t<T>, dim3(std::min(nblk((size_t)nr * nr, 256), 4096)), dim3(256), 0, c->stream, D,
                           (size_t)nr * nr);
        hipLaunchKernelGGL(k_row_block<T>, dim3(nblk((size_t)c->cfg.n * c->cfg.l, 256)), dim3(256), 0, c->stream,
                           c->sc, G, j, nr, c->d_J, c->d_coef, sign, p->d_pin, p->d_rs, p->d_cs, D);
        const int np = j > 0 ? p->nj[j - 1] : 0;
        if (np > 0) {
            T *X = F.blk + p->xoff[j];
            const T *Mprev = F.blk + p->doff[j - 1];
            hipLaunchKernelGGL(k_row_x<T>, dim3(nr), dim3(256), 0, c->stream, np, p->d_uptr + p->ubase[j], p->d_uent,
                               p->d_uval, Mprev, X);
            hipLaunchKernelGGL(k_row_schur<T>, dim3(nr), dim3(256), 0, c->stream, nr, np, p->d_lptr + p->ubase[j],
                               p->d_lent, p->d_lval, X, D);
        }
        HIPCHK(hipGetLastError());
        // Minv_j: solve D W = I, then W replaces the factors
        T *W = reinterpret_cast<T *>(p->d_w);
        if (p->dense_lu) {   // the inverse replaces the factors in D
            int rc = dense_getrf<T>(p->blas, c->stream, nr, D, p->d_ipiv, p->d_info + j, p->d_luw, p->d_rowperm,
                                    p->lookahead ? &p->ahead : nullptr);
            if (!rc) rc = dense_inverse<T>(p->blas, c->stream, nr, D, p->d_rowperm, W);
            if (rc) return rc;
            continue;
        }
        BLASCHK(getrf(p->blas, nr, D, p->d_ipiv, p->d_info + j));
        hipLaunchKernelGGL(k_identity<T>, dim3(std::min(nblk((size_t)nr * nr, 256), 8192)), dim3(256), 0, c->stream, W,
                           nr);
        BLASCHK(getrs(p->blas, nr, D, p->d_ipiv, W));
        HIPCHK(hipMemcpyAsync(D, W, (size_t)nr * nr * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    }
    std::vector<int> info(p->m, 0);
    HIPCHK(hipMemcpyAsync(info.data(), p->d_info, p->m * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < p->m; j++)
        if (p->nj[j] > 0 && info[j] != 0)
            return set_error(IEMIC_ENOCONV, "row preconditioner: row block " + std::to_string(j) +
                                                " is singular (pivot " + std::to_string(info[j]) + ")");
    F.built = true;
    F.version = c->jac_version;
    return IEMIC_OK;
}

template <typename T>
static int rowlu_sweep(iemic_ctx *c, const RowFactors<T> &F, const double *in, double *out, int blocks = 0)
{
    RowLU *p = c->rowlu;
    T *v = reinterpret_cast<T *>(p->d_v), *t = reinterpret_cast<T *>(p->d_t);
    T *W = reinterpret_cast<T *>(p->d_mmw);
    const size_t nw = dense_mm_work(p->nmax) * (sizeof(double) / sizeof(T));
    hipLaunchKernelGGL(k_row_land, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc.ncell, p->d_loc, in, out);
    if (p->nocean == 0) return IEMIC_OK;
    hipLaunchKernelGGL(k_row_gather<T>, dim3(nblk(p->nocean, 256)), dim3(256), 0, c->stream, p->nocean, p->d_perm,
                       p->d_rs, in, v);
    // forward: z_j = Minv_j (v_j - A_{j,j-1} z_{j-1}), z overwrites v
    for (int j = 0; j < p->m; j++) {
        const int nr = p->nj[j];
        if (nr == 0) continue;
        T *vj = v + p->ubase[j];
        const int np = j > 0 ? p->nj[j - 1] : 0;
        if (np > 0)
            hipLaunchKernelGGL(k_row_lmv<T>, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, nr,
                               p->d_lptr + p->ubase[j], p->d_lent, p->d_lval, vj, v + p->ubase[j - 1], t);
        else
            HIPCHK(hipMemcpyAsync(t, vj, (size_t)nr * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        if (p->blas_apply) BLASCHK(gemv(p->blas, nr, nr, (T)1, F.blk + p->doff[j], t, (T)0, vj));
        else HIPCHK(dense_mm<T>(c->stream, nr, nr, 1, (T)1, F.blk + p->doff[j], nr, t, nr, (T)0, vj, nr, W, nw, blocks));
    }
    // backward: z_{j-1} -= X_j z_j
    for (int j = p->m - 1; j >= 1; j--) {
        const int nr = p->nj[j], np = p->nj[j - 1];
        if (nr == 0 || np == 0) continue;
        if (p->blas_apply)
            BLASCHK(gemv(p->blas, np, nr, (T)-1, F.blk + p->xoff[j], v + p->ubase[j], (T)1, v + p->ubase[j - 1]));
        else
            HIPCHK(dense_mm<T>(c->stream, np, nr, 1, (T)-1, F.blk + p->xoff[j], np, v + p->ubase[j], nr, (T)1,
                               v + p->ubase[j - 1], np, W, nw, blocks));
    }
    hipLaunchKernelGGL(k_row_scatter<T>, dim3(nblk(p->nocean, 256)), dim3(256), 0, c->stream, p->nocean, p->d_perm,
                       p->d_cs, v, out);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

// M^{-T}: with M = L U, L block lower (D_j on the diagonal, A_{j,j-1} below)
// and U unit block upper (X_{j+1} above), M^{-T} = L^{-T} U^{-T}:
//   forward  y_0 = v_0,  y_{j+1} = v_{j+1} - X_{j+1}^T y_j
//   backward x_j = D_j^{-T} (y_j - A_{j+1,j}^T x_{j+1})
// and the scales swap roles: M^{-T} = R (R M C)^{-T} C.
template <typename T>
static int rowlu_sweep_t(iemic_ctx *c, const RowFactors<T> &F, const double *in, double *out)
{
    RowLU *p = c->rowlu;
    T *v = reinterpret_cast<T *>(p->d_v), *t = reinterpret_cast<T *>(p->d_t);
    hipLaunchKernelGGL(k_row_land, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc.ncell, p->d_loc, in, out);
    if (p->nocean == 0) return IEMIC_OK;
    hipLaunchKernelGGL(k_row_gather<T>, dim3(nblk(p->nocean, 256)), dim3(256), 0, c->stream, p->nocean, p->d_perm,
                       p->d_cs, in, v);
    for (int j = 0; j + 1 < p->m; j++) {
        const int nr = p->nj[j], nn = p->nj[j + 1];
        if (nr == 0 || nn == 0) continue;
        if (p->blas_apply)
            BLASCHK(gemv_t(p->blas, nr, nn, (T)-1, F.blk + p->xoff[j + 1], v + p->ubase[j], (T)1, v + p->ubase[j + 1]));
        else
            HIPCHK(dense_mm_t<T>(c->stream, nr, nn, 1, (T)-1, F.blk + p->xoff[j + 1], nr, v + p->ubase[j], nr, (T)1,
                                 v + p->ubase[j + 1], nn));
    }
    for (int j = p->m - 1; j >= 0; j--) {
        const int nr = p->nj[j];
        if (nr == 0) continue;
        T *vj = v + p->ubase[j];
        if (j + 1 < p->m && p->nj[j + 1] > 0)
            hipLaunchKernelGGL(k_row_ltmv<T>, dim3(nblk(nr, 256)), dim3(256), 0, c->stream, nr,
                               p->d_cptr + p->ubase[j], p->d_cent, p->d_lval, vj, v + p->ubase[j + 1], t);
        else
            HIPCHK(hipMemcpyAsync(t, vj, (size_t)nr * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        if (p->blas_apply) BLASCHK(gemv_t(p->blas, nr, nr, (T)1, F.blk + p->doff[j], t, (T)0, vj));
        else HIPCHK(dense_mm_t<T>(c->stream, nr, nr, 1, (T)1, F.blk + p->doff[j], nr, t, nr, (T)0, vj, nr));
    }
    hipLaunchKernelGGL(k_row_scatter<T>, dim3(nblk(p->nocean, 256)), dim3(256), 0, c->stream, p->nocean, p->d_perm,
                       p->d_rs, v, out);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

static bool row_current(const iemic_ctx *c, bool built, unsigned version)
{
    return built && c->jac_version - version <= c->prec_max_age;
}

int rowlu_build(iemic_ctx *c, int precision)
{
    int rc = rowlu_ensure(c);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    timed_begin_ext(c, precision ? 5 : 4);
    rc = precision ? rowlu_factor<float>(c, p->f32) : rowlu_factor<double>(c, p->f64);
    timed_end_ext(c, precision ? 5 : 4);
    return rc;
}

int rowlu_ready(iemic_ctx *c, int precision)
{
    int rc = rowlu_ensure(c);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    const bool cur = precision ? row_current(c, p->f32.built, p->f32.version) : row_current(c, p->f64.built, p->f64.version);
    return cur ? IEMIC_OK : rowlu_build(c, precision);
}

// M^{-1} in on the model's unknowns (FP64 in and out); the factors and the
// sweeps run in FP64 (precision 0) or FP32 (precision 1).
// 1 if the factors of this precision were built from an older Jacobian
int rowlu_stale(iemic_ctx *c, int precision)
{
    RowLU *p = c->rowlu;
    if (!p) return 0;
    const RowFactors<double> &F64 = p->f64;
    return precision ? (p->f32.built && p->f32.version != c->jac_version)
                     : (F64.built && F64.version != c->jac_version);
}

// Autotuning of the dense products' launch shape: the number of blocks
// (row blocks times K slices) each product aims at decides how far a long
// factor row is split; fewer slices read the right-hand sides less often,
// more keep more CUs streaming.  The first apply with a given number of
// right-hand sides times every candidate on the real operands (the sweep
// is a pure function of its input, so each run simply rewrites the output)
// and keeps the fastest for this context.
static const int MM_CANDIDATES[] = {512, 1024, 2048, 4096, 8192};

template <typename Run>
static int tune_blocks(iemic_ctx *c, int q, Run run)
{
    RowLU *p = c->rowlu;
    auto it = p->mm_blocks.find(q);
    if (it != p->mm_blocks.end()) return it->second;
    if (std::getenv("IEMIC_NO_TUNE")) return p->mm_blocks[q] = 2048;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 2048;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 2048;
    }
    int best = 2048;
    float tbest = 1e30f;
    for (int cand : MM_CANDIDATES) {
        float tmin = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            (void)hipEventRecord(e0, c->stream);
            if (run(cand)) break;
            (void)hipEventRecord(e1, c->stream);
            float ms = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
            tmin = std::min(tmin, ms);
        }
        if (tmin < tbest) {
            tbest = tmin;
            best = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return p->mm_blocks[q] = best;
}

int rowlu_tuned_blocks(iemic_ctx *c, int nrhs)
{
    if (!c->rowlu) return 0;
    auto it = c->rowlu->mm_blocks.find(nrhs);
    return it == c->rowlu->mm_blocks.end() ? 0 : it->second;
}

int rowlu_apply(iemic_ctx *c, int precision, const double *in, double *out)
{
    int rc = rowlu_ready(c, precision);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    if (!p->blas_apply && in != out)   // tuning rewrites out, so not in place
        (void)tune_blocks(c, precision ? -1 : 1, [&](int b) {
            return precision ? rowlu_sweep<float>(c, p->f32, in, out, b) : rowlu_sweep<double>(c, p->f64, in, out, b);
        });
    auto tb = p->mm_blocks.find(precision ? -1 : 1);
    const int blocks = p->blas_apply || tb == p->mm_blocks.end() ? 0 : tb->second;
    timed_begin_ext(c, precision ? 6 : 3);
    rc = precision ? rowlu_sweep<float>(c, p->f32, in, out, blocks) : rowlu_sweep<double>(c, p->f64, in, out, blocks);
    timed_end_ext(c, precision ? 6 : 3);
    return rc;
}

// ------------------------------------------------------------------ several right-hand sides
// Column q of a multi-vector lives at base + q * stride; blockIdx.y is q.
template <typename T>
__global__ void k_rowm_gather(int nocean, const int *__restrict__ perm, const double *__restrict__ rs,
                              const double *__restrict__ in, size_t ldin, T *__restrict__ v)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nocean) return;
    const size_t q = blockIdx.y;
    v[q * nocean + u] = (T)(rs[perm[u]] * in[q * ldin + perm[u]]);
}

template <typename T>
__global__ void k_rowm_scatter(int nocean, const int *__restrict__ perm, const double *__restrict__ cs,
                               const T *__restrict__ v, double *__restrict__ out, size_t ldout)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nocean) return;
    const size_t q = blockIdx.y;
    out[q * ldout + perm[u]] = cs[perm[u]] * (double)v[q * nocean + u];
}

__global__ void k_rowm_land(int ncell, const int *__restrict__ loc, const double *__restrict__ in, size_t ldin,
                            double *__restrict__ out, size_t ldout)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell || loc[cell] >= 0) return;
    const size_t q = blockIdx.y;
#pragma unroll
    for (int v = 0; v < 6; v++) out[q * ldout + 6 * (size_t)cell + v] = in[q * ldin + 6 * (size_t)cell + v];
}

// t(:, q) = v_j(:, q) - A_{j,j-1} z_{j-1}(:, q); v columns at stride nocean, t packed (nrow)
template <typename T>
__global__ void k_rowm_lmv(int nrow, const int *__restrict__ lptr, const int2 *__restrict__ lent,
                           const double *__restrict__ lval, const T *__restrict__ vj, const T *__restrict__ zprev,
                           size_t ldv, T *__restrict__ t)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrow) return;
    const size_t q = blockIdx.y;
    const T *zq = zprev + q * ldv;
    T acc = vj[q * ldv + e];
    for (int k = lptr[e]; k < lptr[e + 1]; k++) acc -= (T)lval[k] * zq[lent[k].x];
    t[q * (size_t)nrow + e] = acc;
}

static rocblas_status gemm(rocblas_handle h, int m, int n, int k, double alpha, const double *A, int lda,
                           const double *B, int ldb, double beta, double *C, int ldc)
{
    return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                         ldc);
}
static rocblas_status gemm(rocblas_handle h, int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                           int ldb, float beta, float *C, int ldc)
{
    return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                         ldc);
}

// M^{-1} on nrhs vectors at once: the same sweep as rowlu_sweep with the
// dense products as GEMMs (nr x nr times nr x nrhs), so the factors stream
// from HBM once for all right-hand sides.
template <typename T>
static int rowlu_sweep_multi(iemic_ctx *c, const RowFactors<T> &F, int nrhs, const double *in, size_t ldin,
                             double *out, size_t ldout, int blocks = 0)
{
    RowLU *p = c->rowlu;
    if (p->mcap < nrhs) {
        if (p->d_vm) { (void)hipFree(p->d_vm); (void)hipFree(p->d_tm); p->d_vm = p->d_tm = nullptr; }
        HIPCHK(hipMalloc(&p->d_vm, (size_t)std::max(p->nocean, 1) * nrhs * sizeof(double)));
        HIPCHK(hipMalloc(&p->d_tm, (size_t)std::max(p->nmax, 1) * nrhs * sizeof(double)));
        p->mcap = nrhs;
    }
    T *v = reinterpret_cast<T *>(p->d_vm), *t = reinterpret_cast<T *>(p->d_tm);
    const size_t ldv = (size_t)p->nocean;
    T *W = reinterpret_cast<T *>(p->d_mmw);
    const size_t nw = dense_mm_work(p->nmax) * (sizeof(double) / sizeof(T));
    hipLaunchKernelGGL(k_rowm_land, dim3(nblk(c->sc.ncell, 256), nrhs), dim3(256), 0, c->stream, c->sc.ncell, p->d_loc,
                       in, ldin, out, ldout);
    if (p->nocean == 0) return IEMIC_OK;
    hipLaunchKernelGGL(k_rowm_gather<T>, dim3(nblk(p->nocean, 256), nrhs), dim3(256), 0, c->stream, p->nocean, p->d_perm,
                       p->d_rs, in, ldin, v);
    for (int j = 0; j < p->m; j++) {
        const int nr = p->nj[j];
        if (nr == 0) continue;
        T *vj = v + p->ubase[j];
        const int np = j > 0 ? p->nj[j - 1] : 0;
        if (np > 0) {
            hipLaunchKernelGGL(k_rowm_lmv<T>, dim3(nblk(nr, 256), nrhs), dim3(256), 0, c->stream, nr,
                               p->d_lptr + p->ubase[j], p->d_lent, p->d_lval, vj, v + p->ubase[j - 1], ldv, t);
        } else {
            HIPCHK(hipMemcpy2DAsync(t, (size_t)nr * sizeof(T), vj, ldv * sizeof(T), (size_t)nr * sizeof(T), nrhs,
                                    hipMemcpyDeviceToDevice, c->stream));
        }
        if (p->blas_apply)
            BLASCHK(gemm(p->blas, nr, nrhs, nr, (T)1, F.blk + p->doff[j], nr, t, nr, (T)0, vj, (int)ldv));
        else
            HIPCHK(dense_mm<T>(c->stream, nr, nr, nrhs, (T)1, F.blk + p->doff[j], nr, t, nr, (T)0, vj, ldv, W, nw,
                               blocks));
    }
    for (int j = p->m - 1; j >= 1; j--) {
        const int nr = p->nj[j], np = p->nj[j - 1];
        if (nr == 0 || np == 0) continue;
        if (p->blas_apply)
            BLASCHK(gemm(p->blas, np, nrhs, nr, (T)-1, F.blk + p->xoff[j], np, v + p->ubase[j], (int)ldv, (T)1,
                         v + p->ubase[j - 1], (int)ldv));
        else
            HIPCHK(dense_mm<T>(c->stream, np, nr, nrhs, (T)-1, F.blk + p->xoff[j], np, v + p->ubase[j], ldv, (T)1,
                               v + p->ubase[j - 1], ldv, W, nw, blocks));
    }
    hipLaunchKernelGGL(k_rowm_scatter<T>, dim3(nblk(p->nocean, 256), nrhs), dim3(256), 0, c->stream, p->nocean,
                       p->d_perm, p->d_cs, v, out, ldout);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

// t(:, q) = y_j(:, q) - A_{j+1,j}^T x_{j+1}(:, q); y, x at stride ldv, t packed (nrow)
template <typename T>
__global__ void k_rowm_ltmv(int nrow, const int *__restrict__ cptr, const int2 *__restrict__ cent,
                            const double *__restrict__ lval, const T *__restrict__ yj, const T *__restrict__ xnext,
                            size_t ldv, T *__restrict__ t)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrow) return;
    const size_t q = blockIdx.y;
    const T *xq = xnext + q * ldv;
    T acc = yj[q * ldv + e];
    for (int k = cptr[e]; k < cptr[e + 1]; k++) acc -= (T)lval[cent[k].y] * xq[cent[k].x];
    t[q * (size_t)nrow + e] = acc;
}

static rocblas_status gemm_t(rocblas_handle h, int m, int n, int k, double alpha, const double *A, int lda,
                             const double *B, int ldb, double beta, double *C, int ldc)
{
    return rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta,
                         C, ldc);
}
static rocblas_status gemm_t(rocblas_handle h, int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                             int ldb, float beta, float *C, int ldc)
{
    return rocblas_sgemm(h, rocblas_operation_transpose, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta,
                         C, ldc);
}

// M^{-T} on nrhs vectors: rowlu_sweep_t with GEMMs over all columns
template <typename T>
static int rowlu_sweep_t_multi(iemic_ctx *c, const RowFactors<T> &F, int nrhs, const double *in, size_t ldin,
                               double *out, size_t ldout)
{
    RowLU *p = c->rowlu;
    if (p->mcap < nrhs) {
        if (p->d_vm) { (void)hipFree(p->d_vm); (void)hipFree(p->d_tm); p->d_vm = p->d_tm = nullptr; }
        HIPCHK(hipMalloc(&p->d_vm, (size_t)std::max(p->nocean, 1) * nrhs * sizeof(double)));
        HIPCHK(hipMalloc(&p->d_tm, (size_t)std::max(p->nmax, 1) * nrhs * sizeof(double)));
        p->mcap = nrhs;
    }
    T *v = reinterpret_cast<T *>(p->d_vm), *t = reinterpret_cast<T *>(p->d_tm);
    const size_t ldv = (size_t)p->nocean;
    hipLaunchKernelGGL(k_rowm_land, dim3(nblk(c->sc.ncell, 256), nrhs), dim3(256), 0, c->stream, c->sc.ncell, p->d_loc,
                       in, ldin, out, ldout);
    if (p->nocean == 0) return IEMIC_OK;
    hipLaunchKernelGGL(k_rowm_gather<T>, dim3(nblk(p->nocean, 256), nrhs), dim3(256), 0, c->stream, p->nocean, p->d_perm,
                       p->d_cs, in, ldin, v);
    // Y = alpha A^T X + beta Y (A column-major m x k): the library's
    // transposed product, up to 8 columns per pass over A, or rocBLAS
    auto mmt = [&](int m, int k, T alpha, const T *A, const T *X, size_t ldx, T beta, T *Y, size_t ldy) -> int {
        if (p->blas_apply) {
            BLASCHK(gemm_t(p->blas, k, nrhs, m, alpha, A, m, X, (int)ldx, beta, Y, (int)ldy));
            return IEMIC_OK;
        }
        for (int c0 = 0; c0 < nrhs; c0 += 8)
            HIPCHK(dense_mm_t<T>(c->stream, m, k, std::min(8, nrhs - c0), alpha, A, m, X + c0 * ldx, ldx, beta,
                                 Y + c0 * ldy, ldy));
        return IEMIC_OK;
    };
    for (int j = 0; j + 1 < p->m; j++) {
        const int nr = p->nj[j], nn = p->nj[j + 1];
        if (nr == 0 || nn == 0) continue;
        if (int rc = mmt(nr, nn, (T)-1, F.blk + p->xoff[j + 1], v + p->ubase[j], ldv, (T)1, v + p->ubase[j + 1], ldv))
            return rc;
    }
    for (int j = p->m - 1; j >= 0; j--) {
        const int nr = p->nj[j];
        if (nr == 0) continue;
        T *vj = v + p->ubase[j];
        if (j + 1 < p->m && p->nj[j + 1] > 0)
            hipLaunchKernelGGL(k_rowm_ltmv<T>, dim3(nblk(nr, 256), nrhs), dim3(256), 0, c->stream, nr,
                               p->d_cptr + p->ubase[j], p->d_cent, p->d_lval, vj, v + p->ubase[j + 1], ldv, t);
        else
            HIPCHK(hipMemcpy2DAsync(t, (size_t)nr * sizeof(T), vj, ldv * sizeof(T), (size_t)nr * sizeof(T), nrhs,
                                    hipMemcpyDeviceToDevice, c->stream));
        if (int rc = mmt(nr, nr, (T)1, F.blk + p->doff[j], t, nr, (T)0, vj, ldv)) return rc;
    }
    hipLaunchKernelGGL(k_rowm_scatter<T>, dim3(nblk(p->nocean, 256), nrhs), dim3(256), 0, c->stream, p->nocean,
                       p->d_perm, p->d_rs, v, out, ldout);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

int rowlu_apply_multi_t(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                        size_t ldout)
{
    int rc = rowlu_ready(c, precision);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    return precision ? rowlu_sweep_t_multi<float>(c, p->f32, nrhs, in, ldin, out, ldout)
                     : rowlu_sweep_t_multi<double>(c, p->f64, nrhs, in, ldin, out, ldout);
}

int rowlu_apply_multi(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out, size_t ldout)
{
    int rc = rowlu_ready(c, precision);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    const int key = precision ? -nrhs : nrhs;
    if (!p->blas_apply && in != out)
        (void)tune_blocks(c, key, [&](int b) {
            return precision ? rowlu_sweep_multi<float>(c, p->f32, nrhs, in, ldin, out, ldout, b)
                             : rowlu_sweep_multi<double>(c, p->f64, nrhs, in, ldin, out, ldout, b);
        });
    auto tb = p->mm_blocks.find(key);
    const int blocks = p->blas_apply || tb == p->mm_blocks.end() ? 0 : tb->second;
    timed_begin_ext(c, 7);
    rc = precision ? rowlu_sweep_multi<float>(c, p->f32, nrhs, in, ldin, out, ldout, blocks)
                   : rowlu_sweep_multi<double>(c, p->f64, nrhs, in, ldin, out, ldout, blocks);
    timed_end_ext(c, 7);
    return rc;
}

int rowlu_apply_t(iemic_ctx *c, int precision, const double *in, double *out)
{
    int rc = rowlu_ready(c, precision);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    return precision ? rowlu_sweep_t<float>(c, p->f32, in, out) : rowlu_sweep_t<double>(c, p->f64, in, out);
}

__global__ void k_unit(double *__restrict__ v, int n, int r)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) v[q] = q == r ? 1.0 : 0.0;
}
__global__ void k_sub_from(double *__restrict__ y, const double *__restrict__ x, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] -= x[q];
}
__global__ void k_add_to(double *__restrict__ y, const double *__restrict__ x, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] += x[q];
}
__global__ void k_rsub(double *__restrict__ y, const double *__restrict__ t, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) y[q] = t[q] - y[q];
}
// the rows M replaces (pins, then the integral row if any): y[r] = mode 0:
// v[r] (pins) / d v[r] (integral row); mode 1: 0; mode 2: y[r] + (same as 0)
__global__ void k_replaced_rows(double *__restrict__ y, const double *__restrict__ v, const int *__restrict__ rows,
                                int nrows, int ric, double d, int mode)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > nrows) return;
    int r;
    double w;
    if (q < nrows) {
        r = rows[q];
        w = v[r];
    } else {
        if (ric < 0) return;
        r = ric;
        w = d * v[r];
    }
    y[r] = mode == 1 ? 0.0 : (mode == 2 ? y[r] + w : w);
}

// y = M v (or M^T v) through products with J: M = J + sum_c e_{r_c} (e_{r_c}
// - J[r_c,:]) + e_ric (d e_ric - J[ric,:]), so M v is J v with the replaced
// rows set to v[r_c] and d v[ric], and M^T v = J^T v~ plus those diagonal
// terms, v~ = v with the replaced entries zeroed.  tmp: ndim scratch.
static int apply_pinned(iemic_ctx *c, RowLU *p, bool transpose, const double *v, double *y, double *tmp)
{
    const int nd = c->sc.ndim, ric = c->sc.rowintcon, nr = p->npins;
    const double d = ric >= 0 && c->h.coef[ric] != 0.0 ? c->cfg.int_sign * c->h.coef[ric] : 1.0;
    int rc;
    if (!transpose) {
        if ((rc = launch_spmv(c, v, y))) return rc;
        hipLaunchKernelGGL(k_replaced_rows, dim3(nblk(nr + 1, 64)), dim3(64), 0, c->stream, y, v, p->d_pin_rows, nr,
                           ric, d, 0);
    } else {
        HIPCHK(hipMemcpyAsync(tmp, v, (size_t)nd * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        hipLaunchKernelGGL(k_replaced_rows, dim3(nblk(nr + 1, 64)), dim3(64), 0, c->stream, tmp, v, p->d_pin_rows, nr,
                           ric, d, 1);
        if ((rc = launch_spmv_t(c, tmp, y))) return rc;
        hipLaunchKernelGGL(k_replaced_rows, dim3(nblk(nr + 1, 64)), dim3(64), 0, c->stream, y, v, p->d_pin_rows, nr,
                           ric, d, 2);
    }
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

template <typename T>
static int pinned_solve(iemic_ctx *c, RowLU *p, const RowFactors<T> &F, bool transpose, const double *in, double *out)
{
    return transpose ? rowlu_sweep_t<T>(c, F, in, out) : rowlu_sweep<T>(c, F, in, out);
}

// ||J v|| (J^T v when left) <= 1e-10 ||J e_r|| ||v||: v is a null vector
static int null_check(iemic_ctx *c, bool left, const double *v, double colnorm, double *w, bool *ok, double *vv)
{
    int rc = left ? launch_spmv_t(c, v, w) : launch_spmv(c, v, w);
    if (rc) return rc;
    const int nd = c->sc.ndim;
    *vv = device_dot(c, v, v, nd);
    const double res = std::sqrt(device_dot(c, w, w, nd));
    HIPCHK(hipGetLastError());
    *ok = res <= 1e-10 * colnorm * std::sqrt(*vv);   // false for NaN (factors that broke down)
    return IEMIC_OK;
}

// Null vectors of J, one per pressure basin b with pinned row r_b.  The
// factored matrix M equals J except in the replaced rows (see apply_pinned).
// A right null vector n of J with n[r_c] = delta_bc satisfies M n = e_{r_b};
// a left null vector m of J (m^T J = 0, supported on the continuity rows of
// basin b, m[r_c] = delta_bc, m[ric] = 0) satisfies M^T m = e_{r_b} - J^T
// e_{r_b}.  Each is solved with the factors of the given precision, refined
// twice against FP64 products with M, and kept if null_check passes; a
// basin whose vector fails gets none (its norm is stored as 0).  The
// vectors involve only the continuity rows and the pressure columns of J,
// which are linear in the state with coefficients fixed by the grid and
// mask, so they are computed once; after a new Jacobian they are checked
// again (one product each) and recomputed only if a check fails.
int rowlu_null_space(iemic_ctx *c, int precision, bool left, const double **vecs, const double **nrm2, int *count)
{
    *count = 0;
    int rc = rowlu_ready(c, precision);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    const int side = left ? 0 : 1, nb = p->npins;
    const size_t nd = (size_t)c->sc.ndim;
    if (nb == 0 || nb > ROWLU_NULL_MAX) return IEMIC_OK;
    if (!p->d_unit) HIPCHK(hipMalloc(&p->d_unit, 3 * nd * sizeof(double)));
    double *t = p->d_unit, *w1 = p->d_unit + nd, *w2 = p->d_unit + 2 * nd;
    const dim3 g(nblk(nd, 256)), blk(256);
    // scale of J near each pinned unknown: ||J e_r|| (||J^T e_r|| when left)
    auto colnorm = [&](int b, double *out) -> int {
        hipLaunchKernelGGL(k_unit, g, blk, 0, c->stream, t, (int)nd, p->pin_rows[b]);
        int r = left ? launch_spmv_t(c, t, w2) : launch_spmv(c, t, w2);
        if (r) return r;
        *out = std::sqrt(device_dot(c, w2, w2, (int)nd));
        return IEMIC_OK;
    };
    if (p->d_null[side] && p->null_prec[side] == 1 && precision == 0)
        for (double v2 : p->null_nrm2[side])
            if (v2 == 0.0) p->null_prec[side] = -1;   // retry with the FP64 factors
    if (p->d_null[side] && p->null_prec[side] >= 0 && p->null_version[side] != c->jac_version) {
        bool all = true;
        for (int b = 0; b < nb && all; b++) {
            if (p->null_nrm2[side][b] == 0.0) continue;
            double cn, vv;
            bool ok;
            if ((rc = colnorm(b, &cn))) return rc;
            if ((rc = null_check(c, left, p->d_null[side] + (size_t)b * nd, cn, w1, &ok, &vv))) return rc;
            all = ok;
        }
        if (all) p->null_version[side] = c->jac_version;
        else p->null_prec[side] = -1;   // recompute below
    }
    if (!p->d_null[side] || p->null_prec[side] < 0) {
        if (!p->d_null[side]) HIPCHK(hipMalloc(&p->d_null[side], (size_t)nb * nd * sizeof(double)));
        p->null_nrm2[side].assign(nb, 0.0);
        for (int b = 0; b < nb; b++) {
            double *v = p->d_null[side] + (size_t)b * nd;
            double cn;
            if ((rc = colnorm(b, &cn))) return rc;      // leaves t = e_r, w2 = J e_r / J^T e_r
            if (left) hipLaunchKernelGGL(k_sub_from, g, blk, 0, c->stream, t, w2, (int)nd);
            rc = precision ? pinned_solve<float>(c, p, p->f32, left, t, v) : pinned_solve<double>(c, p, p->f64, left, t, v);
            if (rc) return rc;
            for (int it = 0; it < 2; it++) {
                if ((rc = apply_pinned(c, p, left, v, w1, w2))) return rc;
                hipLaunchKernelGGL(k_rsub, g, blk, 0, c->stream, w1, t, (int)nd);
                rc = precision ? pinned_solve<float>(c, p, p->f32, left, w1, w2)
                               : pinned_solve<double>(c, p, p->f64, left, w1, w2);
                if (rc) return rc;
                hipLaunchKernelGGL(k_add_to, g, blk, 0, c->stream, v, w2, (int)nd);
            }
            double vv;
            bool ok;
            if ((rc = null_check(c, left, v, cn, w1, &ok, &vv))) return rc;
            p->null_nrm2[side][b] = ok ? vv : 0.0;
        }
        p->null_prec[side] = precision;
        p->null_version[side] = c->jac_version;
    }
    *vecs = p->d_null[side];
    *nrm2 = p->null_nrm2[side].data();
    *count = nb;
    return IEMIC_OK;
}

int rowlu_info(iemic_ctx *c, iemic_precond_info *info)
{
    int rc = rowlu_ensure(c);
    if (rc) return rc;
    RowLU *p = c->rowlu;
    info->blocks = p->m;
    info->max_block = p->nmax;
    info->pins = p->npins;
    info->levels = 2 * p->m - 1;
    info->factor_bytes = (double)p->total * sizeof(double);
    return IEMIC_OK;
}

}  // namespace iemic
