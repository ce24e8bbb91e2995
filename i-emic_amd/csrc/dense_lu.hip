// Dense LU with tournament pivoting for the row-block preconditioner
// (rowlu.hip), in LAPACK getrf's layout: A = P L U in place, L unit lower,
// info = first zero pivot (1-based) or 0; the row permutation is returned as
// perm (and, with IEMIC_LU_IPIV=1, also as LAPACK's 1-based interchanges
// ipiv, so that LAPACK-style solvers such as rocSOLVER getrs apply the
// factors as is); dense_inverse forms A^{-1} from the factors and perm in
// place.
//
// Why not rocSOLVER getrf: its panel factorisation launches three kernels
// per column (pivot search, scale/update, singularity check) of 2-4 us each,
// and at the benchmark sizes those latencies are more than half of the
// preconditioner build.  Here the pivots of a whole sub-panel of CA_W
// columns are chosen at once by tournament pivoting (communication-avoiding
// LU, Grigori, Demmel and Xiang, SIAM J. Matrix Anal. Appl. 32 (2011)):
//   1. every group of (up to) CA_G rows runs Gaussian elimination with partial
//      pivoting on its rows of the sub-panel, in registers, and nominates the
//      CA_W rows it pivots on;
//   2. the nominees are merged CA_FAN sets at a time by the same elimination
//      on their (original) rows until one set is left;
//   3. those rows are moved to the top, in the order chosen, and the
//      sub-panel is factored without further pivoting: U11 and L11 are what
//      the last elimination computed, and L21 = A21 U11^{-1}.
// The pivot rows of a panel then span its column space whenever the panel
// has full column rank, and the element growth stays close to partial
// pivoting in practice; only the pivot order differs from LAPACK's.
//
// Sub-panels are grouped into outer blocks of CA_NB columns: inside a block
// the sub-panel's rank-CA_W update reaches the block's columns only; after
// the block, rocBLAS TRSM and GEMM (matrix cores) update the trailing
// matrix with depth CA_NB.  Per sub-panel there are 5-6 launches instead of
// three per column.
//
// Look-ahead (with a second stream): the panel chain above is latency-bound
// and occupies few CUs, the trailing update is GEMM-bound, so block b + 1's
// panel is factored while block b's update of the columns beyond it runs on
// the second stream.  For that, a panel's row interchanges are applied at
// once only to its own columns; the columns left of it and beyond the next
// panel receive them from the second stream (replayed from the recorded
// move lists, in order) just before that stream's TRSM and GEMM, and the
// next panel's columns receive them on the panel stream.  The two streams
// then never write the same columns: the panel stream owns the current and
// next panel, the second stream everything else, and events order the hand-
// overs (the next panel waits for the previous trailing update, the trailing
// update for the panel whose factors it applies).  Measured at 180x80x24
// (one MI355X): the two streams overlap for 0.63 s of the 5.3 s build, but
// the trailing updates total only ~0.7 s (the inverse's TRSM and TRTRI are
// the larger part of the matrix-core work) and the GEMMs that run beside the
// panel slow its kernels by about as much, so the build time is unchanged
// (5.33 s against 5.32 s); the row LU therefore runs without it unless asked
// (IEMIC_LU_LOOKAHEAD=1).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <string>
#include <type_traits>

#include "ctx.hpp"
#include "dense_lu.hpp"

namespace iemic {

constexpr int CA_W = 32;               // sub-panel width
constexpr int CA_G = 512;              // rows per selection group (one row per thread), at most
constexpr int CA_FAN = CA_G / CA_W;    // candidate sets merged per group above the first level
constexpr int CA_NB = 512;             // outer block width (TRSM/GEMM depth), default (128: +4 %, 256: +1.5 %, 1024: same, at 180x80x24)
constexpr int CA_NB_MAX = 1024;

// Outer block width: CA_NB, or IEMIC_LU_NB (128, 256, 512 or 1024) for
// comparison runs.
static int lu_block()
{
    static const int nb = [] {
        const char *v = std::getenv("IEMIC_LU_NB");
        const int b = v ? std::atoi(v) : CA_NB;
        return (b == 128 || b == 256 || b == 512 || b == 1024) ? b : CA_NB;
    }();
    return nb;
}
constexpr int CA_MOVES = 2 * CA_W;     // rows a sub-panel's interchanges can touch
constexpr int CA_MSTRIDE = 1 + 2 * CA_MOVES;   // one move list: count, destinations, sources
constexpr int CA_SUBS = CA_NB_MAX / CA_W;      // sub-panels of an outer block, at most
constexpr int CA_STRIP = 16;           // columns per workgroup in the U12 solve and block update

// Workspace: two candidate lists, the CA_W x CA_W factors of the chosen
// rows, the move lists of the sub-panels of two outer blocks.
static size_t cand_elems(int n) { return (size_t)((n + CA_G - 1) / CA_G) * CA_W; }

size_t dense_getrf_work_bytes(int n)
{
    return 2 * cand_elems(n) * sizeof(int) + (size_t)CA_W * CA_W * sizeof(double) +
           (size_t)2 * CA_SUBS * CA_MSTRIDE * sizeof(int) + 256;
}

// wave64 maximum of unsigned 32-bit values, returned to every lane: DPP row
// shifts by 1, 2, 4, 8 (lane 15 of each row of 16 then holds the row's
// maximum), row broadcasts of lanes 15 and 31, and lane 63 read out.  Lanes
// without a source read 0 (bound_ctrl), the identity of the maximum, so the
// compiler folds each step into one v_max_u32 with a DPP operand.
__device__ __forceinline__ unsigned wave_max_u32(unsigned x)
{
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, true));
    x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, true));
    return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}

// The value v of lane `src` (uniform) in every lane, through scalar registers.
__device__ __forceinline__ double lane_value(double v, int src)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Order-preserving key of |v| over the lanes of a wave: the bits of |v| (a
// non-negative double orders like its bit pattern) with the top bit set on
// live lanes, 0 on the others.  Returns the first lane whose key's high word
// (sign-free exponent and 20 leading mantissa bits of |v|) is the largest
// (-1 when no lane is live) and that lane's full key.  Taking the first of
// the lanes that tie in the high word rather than the largest full key saves
// a second wave reduction per column step; the pivot chosen is within
// 2^-20 of the largest magnitude, which partial pivoting's bound on the
// multipliers does not notice.
template <typename T>
__device__ __forceinline__ int wave_argmax_abs(T v, bool live, unsigned long long *key)
{
    unsigned hi, lo;
    if constexpr (sizeof(T) == 8) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs((double)v));
        hi = (unsigned)(b >> 32);
        lo = (unsigned)b;
    } else {
        hi = (unsigned)__float_as_uint(fabsf((float)v));
        lo = 0u;
    }
    hi = live ? (hi | 0x80000000u) : 0u;
    const unsigned mh = wave_max_u32(hi);
    const unsigned long long bal = __ballot(hi == mh);
    const int fl = mh ? (int)__ffsll(bal) - 1 : -1;
    const unsigned ml = fl >= 0 ? (unsigned)__builtin_amdgcn_readlane((int)lo, fl) : 0u;
    *key = ((unsigned long long)mh << 32) | ml;
    return fl;
}

// The LAPACK interchange sequence of one sub-panel (ipiv[k .. k + w),
// 1-based) and its net row moves (moves: count, destinations, sources), from
// the chosen rows in pivot order, lane c holding the c-th (wave 0 only).
// Lane e keeps entry e of the map "position key[e] now holds original row
// val[e]" over the positions touched so far (at most 2 w <= 64).
__device__ __forceinline__ void calu_interchanges(int lane, int k, int w, int sel, int *__restrict__ ipiv,
                                                  int *__restrict__ moves)
{
    int key = -1, val = -1, size = 0;
    for (int c = 0; c < w; c++) {
        const int s = __builtin_amdgcn_readlane(sel, c);
        const unsigned long long bv = __ballot(lane < size && val == s);
        const int p = bv ? __builtin_amdgcn_readlane(key, __ffsll((unsigned long long)bv) - 1) : s;
        const unsigned long long bk = __ballot(lane < size && key == k + c);
        const int cur = bk ? __builtin_amdgcn_readlane(val, __ffsll((unsigned long long)bk) - 1) : k + c;
        if (lane == 0) ipiv[k + c] = p + 1;
        // position k + c receives s, position p receives cur
        if (bk) {
            if (lane == __ffsll((unsigned long long)bk) - 1) val = s;
        } else {
            if (lane == size) {
                key = k + c;
                val = s;
            }
            size++;
        }
        if (p != k + c) {
            const unsigned long long bp = __ballot(lane < size && key == p);
            if (bp) {
                if (lane == __ffsll((unsigned long long)bp) - 1) val = cur;
            } else {
                if (lane == size) {
                    key = p;
                    val = cur;
                }
                size++;
            }
        }
    }
    const bool mv = lane < size && key != val;
    const unsigned long long bm = __ballot(mv);
    if (mv) {
        const int slot = __popcll(bm & ((1ull << lane) - 1));
        moves[1 + slot] = key;              // destination
        moves[1 + CA_MOVES + slot] = val;   // source
    }
    if (lane == 0) moves[0] = __popcll(bm);
}

// lu11 (column-major, leading dimension CA_W) from the rows of L11 \ U11
// staged in LDS by the threads that hold them: coalesced stores by all nt
// threads instead of CA_W scattered stores per chosen row (4.7 us less per
// 512-row final launch).
template <typename T>
__device__ __forceinline__ void calu_store_lu11(int t, int nt, int w, const T (*s_lu)[CA_W + 1],
                                                T *__restrict__ lu11)
{
    for (int e = t; e < w * CA_W; e += nt) {
        const int r = e % CA_W, p = e / CA_W;
        if (r < w) lu11[e] = s_lu[r][p];
    }
}

// Gaussian elimination with partial pivoting on the rows of one group,
// restricted to the sub-panel columns [k, k + w): nominates up to w rows, in
// pivot order (largest magnitude; on ties the group's first row).  The group's
// rows are rows_in[g G + t] (absolute row indices, -1 = none), or
// k + g G + t when rows_in is null; G rows per group (64 to CA_G, one per
// thread; a level that fits one group runs in the smallest G that holds it,
// since a column step of a 64-row group takes about half the time of a
// 512-row group's).  At the final level (one group) the
// workgroup also writes the factors of the chosen rows (lu11, column-major,
// leading dimension CA_W), info at the first zero pivot and, when ipiv is
// given, the LAPACK interchanges ipiv[k .. k + w) and the move list that
// performs them.
template <typename T, int G, bool FINAL>
__global__ void __launch_bounds__(G)
    k_calu_select(const T *__restrict__ A, int lda, int k, int w, const int *__restrict__ rows_in, int cnt,
                  int *__restrict__ rows_out, T *__restrict__ lu11, int *__restrict__ ipiv, int *__restrict__ moves,
                  int *__restrict__ info)
{
    // per wave and column parity: the wave's best row (its key, thread, and
    // values), so that one barrier per column suffices
    __shared__ T wrow[2][G / 64][CA_W];
    __shared__ unsigned long long wk[2][G / 64];
    __shared__ int wt[2][G / 64];
    __shared__ int s_sel[CA_W];
    __shared__ T s_lu[FINAL ? CA_W : 1][CA_W + 1];   // final level: the rows of L11 \ U11, staged
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int idx = blockIdx.x * G + t;
    const int myrow = idx < cnt ? (rows_in ? rows_in[idx] : k + idx) : -1;
    bool live = myrow >= 0;
    int chosen = -1;   // final level: the column step that took this row
    // The row stays in registers.  The column loop is unrolled, so column c
    // is a fixed register and each step touches only the columns after c:
    // the pivot row's live part is broadcast through scalar registers and
    // the rows updated branch-free (about 4k instructions in all; a step
    // that updated every column, unrolled, was ~20k, beyond the instruction
    // cache).  At the final level each multiplier replaces the entry it
    // eliminates, so the chosen rows end as the rows of L11 \ U11.
    T a[CA_W];
#pragma unroll
    for (int c = 0; c < CA_W; c++) a[c] = (live && c < w) ? A[myrow + (size_t)(k + c) * lda] : (T)0;
    if (t < CA_W) s_sel[t] = -1;
    __syncthreads();
    bool done = false;   // uniform: no rows left in this group
#pragma unroll
    for (int cb = 0; cb < CA_W; cb += 8) {
        if (cb >= w || done) break;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = cb + j;   // a compile-time constant in each unrolled step
            if (c >= w || done) break;
            const T ac = a[c];
            const int par = c & 1;
            unsigned long long key;
            const int fl = wave_argmax_abs(ac, live, &key);
            if (lane == fl) {
#pragma unroll
                for (int cc = 0; cc < CA_W; cc++)
                    if (cc >= c) wrow[par][wave][cc] = a[cc];
            }
            if (lane == 0) {
                wk[par][wave] = key;
                wt[par][wave] = wave * 64 + (fl < 0 ? 0 : fl);
            }
            __syncthreads();
            unsigned long long bk = wk[par][0];
            int bw = 0;
#pragma unroll
            for (int q = 1; q < G / 64; q++)
                if (wk[par][q] > bk) {
                    bk = wk[par][q];
                    bw = q;
                }
            if (bk == 0) {   // uniform
                done = true;
                break;
            }
            const int pt = wt[par][bw];
            // the pivot row to every lane: lane i reads element i once, the
            // values are then broadcast from lanes into scalar registers (a
            // broadcast LDS read of the whole row by every lane would move
            // 64 copies of it through the LDS return path)
            const T myv = wrow[par][bw][lane & (CA_W - 1)];
            const T piv = lane_value(myv, c);
            if (t == pt) {
                s_sel[c] = myrow;
                live = false;
                chosen = c;
            }
            if (FINAL && t == 0 && piv == (T)0 && *info == 0) *info = k + c + 1;
            // branch-free: rows that are no longer live (and all rows at a
            // zero pivot) take a zero multiplier and keep their values
            const T l = (live && piv != (T)0) ? ac / piv : (T)0;
            if (FINAL && live) a[c] = l;   // column c is done: keep the multiplier in its place
#pragma unroll
            for (int cc = 0; cc < CA_W; cc++)
                if (cc > c) a[cc] = fma(-l, lane_value(myv, cc), a[cc]);
        }
    }
    if (FINAL && chosen >= 0) {
        // the row chosen at step `chosen` holds its multipliers against the
        // pivots before it and, from there on, its row of U11 (a chosen row
        // takes no further updates): row `chosen` of L11 \ U11
#pragma unroll
        for (int p = 0; p < CA_W; p++)
            if (p < w) s_lu[chosen][p] = a[p];
    }
    __syncthreads();
    if (FINAL) calu_store_lu11(t, G, w, s_lu, lu11);
    if (t < CA_W) rows_out[blockIdx.x * CA_W + t] = s_sel[t];
    if (!FINAL || wave != 0 || !ipiv) return;
    calu_interchanges(lane, k, w, lane < CA_W ? s_sel[lane] : -1, ipiv, moves);
}

// The same selection with two rows per thread, rows t and t + G/2 of the
// group on G/2 threads: a 512-row group runs one wave per SIMD instead of
// two, and each broadcast of the pivot row (the readlanes) serves both rows.
// On ties between a thread's two rows the first is taken.
template <typename T, int G, bool FINAL>
__global__ void __launch_bounds__(G / 2)
    k_calu_select2(const T *__restrict__ A, int lda, int k, int w, const int *__restrict__ rows_in, int cnt,
                   int *__restrict__ rows_out, T *__restrict__ lu11, int *__restrict__ ipiv, int *__restrict__ moves,
                   int *__restrict__ info)
{
    constexpr int NT = G / 2, NW = NT / 64;
    __shared__ T wrow[2][NW][CA_W];
    __shared__ unsigned long long wk[2][NW];
    __shared__ int wt[2][NW];
    __shared__ int s_sel[CA_W];
    __shared__ T s_lu[FINAL ? CA_W : 1][CA_W + 1];   // final level: the rows of L11 \ U11, staged
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int idx0 = blockIdx.x * G + t, idx1 = idx0 + NT;
    int chosen0 = -1, chosen1 = -1;   // final level: the column steps that took the rows
    const int row0 = idx0 < cnt ? (rows_in ? rows_in[idx0] : k + idx0) : -1;
    const int row1 = idx1 < cnt ? (rows_in ? rows_in[idx1] : k + idx1) : -1;
    bool live0 = row0 >= 0, live1 = row1 >= 0;
    T a0[CA_W], a1[CA_W];
#pragma unroll
    for (int c = 0; c < CA_W; c++) {
        a0[c] = (live0 && c < w) ? A[row0 + (size_t)(k + c) * lda] : (T)0;
        a1[c] = (live1 && c < w) ? A[row1 + (size_t)(k + c) * lda] : (T)0;
    }
    if (t < CA_W) s_sel[t] = -1;
    __syncthreads();
    bool done = false;   // uniform: no rows left in this group
#pragma unroll
    for (int cb = 0; cb < CA_W; cb += 8) {
        if (cb >= w || done) break;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = cb + j;   // a compile-time constant in each unrolled step
            if (c >= w || done) break;
            const T ac0 = a0[c], ac1 = a1[c];
            const bool take1 = live1 && (!live0 || fabs(ac1) > fabs(ac0));
            const int par = c & 1;
            unsigned long long key;
            const int fl = wave_argmax_abs(take1 ? ac1 : ac0, live0 || live1, &key);
            if (lane == fl) {
#pragma unroll
                for (int cc = 0; cc < CA_W; cc++)
                    if (cc >= c) wrow[par][wave][cc] = take1 ? a1[cc] : a0[cc];
                wt[par][wave] = wave * 64 + lane + (take1 ? NT : 0);
            }
            if (lane == 0) wk[par][wave] = key;
            __syncthreads();
            unsigned long long bk = wk[par][0];
            int bw = 0;
#pragma unroll
            for (int q = 1; q < NW; q++)
                if (wk[par][q] > bk) {
                    bk = wk[par][q];
                    bw = q;
                }
            if (bk == 0) {   // uniform
                done = true;
                break;
            }
            const int pt = wt[par][bw];   // row of the group: thread pt % NT, half pt / NT
            const T myv = wrow[par][bw][lane & (CA_W - 1)];
            const T piv = lane_value(myv, c);
            if (t == pt) {
                s_sel[c] = row0;
                live0 = false;
                chosen0 = c;
            } else if (t + NT == pt) {
                s_sel[c] = row1;
                live1 = false;
                chosen1 = c;
            }
            if (FINAL && t == 0 && piv == (T)0 && *info == 0) *info = k + c + 1;
            // one reciprocal for both rows (LAPACK's getf2 scales by the
            // pivot's reciprocal too)
            const T rinv = piv != (T)0 ? (T)1 / piv : (T)0;
            const T l0 = live0 ? ac0 * rinv : (T)0;
            const T l1 = live1 ? ac1 * rinv : (T)0;
            if (FINAL) {   // column c is done: keep the multipliers in its place
                if (live0) a0[c] = l0;
                if (live1) a1[c] = l1;
            }
#pragma unroll
            for (int cc = 0; cc < CA_W; cc++)
                if (cc > c) {
                    const T pv = lane_value(myv, cc);
                    a0[cc] = fma(-l0, pv, a0[cc]);
                    a1[cc] = fma(-l1, pv, a1[cc]);
                }
        }
    }
    if (FINAL) {   // rows of L11 \ U11, as in k_calu_select
        if (chosen0 >= 0) {
#pragma unroll
            for (int p = 0; p < CA_W; p++)
                if (p < w) s_lu[chosen0][p] = a0[p];
        }
        if (chosen1 >= 0) {
#pragma unroll
            for (int p = 0; p < CA_W; p++)
                if (p < w) s_lu[chosen1][p] = a1[p];
        }
    }
    __syncthreads();
    if (FINAL) calu_store_lu11(t, NT, w, s_lu, lu11);
    if (t < CA_W) rows_out[blockIdx.x * CA_W + t] = s_sel[t];
    if (!FINAL || wave != 0 || !ipiv) return;
    calu_interchanges(lane, k, w, lane < CA_W ? s_sel[lane] : -1, ipiv, moves);
}


// The row moves of one sub-panel on columns [c0, c1) and, as one more
// column, on the row permutation perm (when given): all sources of a column
// are read before any destination is written (one column per 64 threads,
// the barrier orders them).
template <typename T>
__global__ void __launch_bounds__(256)
    k_calu_swap(T *__restrict__ A, int lda, int c0, int c1, const int *__restrict__ moves, int *__restrict__ perm)
{
    const int e = threadIdx.x & 63;
    const int col = c0 + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int cnt = moves[0];
    const bool act = col < c1 && e < cnt;
    const bool actp = perm && col == c1 && e < cnt;
    T v = (T)0;
    int pv = 0;
    if (act) v = A[moves[1 + CA_MOVES + e] + (size_t)col * lda];
    if (actp) pv = perm[moves[1 + CA_MOVES + e]];
    __syncthreads();
    if (act) A[moves[1 + e] + (size_t)col * lda] = v;
    if (actp) perm[moves[1 + e]] = pv;
}

// The row moves of one sub-panel worked out from its chosen rows sel[0 .. w)
// (pivot order, all >= k) instead of the LAPACK interchange sequence: row
// sel[c] moves to position k + c, and each row of [k, k + w) that was not
// chosen moves to one of the positions left by the chosen rows from below
// k + w (the j-th such row, in order, to the j-th such position).  Every
// wave works the moves out for itself (a few ballots, LDS per wave) and
// applies them to one column of [c0, c1), or to perm as the column c1, all
// sources read before any destination is written; workgroup 0 also records
// the move list (count, destinations, sources) for the look-ahead replay.
template <typename T>
__global__ void __launch_bounds__(256)
    k_calu_swap_sel(T *__restrict__ A, int lda, int c0, int c1, int k, int w, const int *__restrict__ sel,
                    int *__restrict__ moves, int *__restrict__ perm)
{
    __shared__ int s_flag[4][CA_W];   // per wave: row k + i was chosen
    __shared__ int s_vac[4][CA_W];    // per wave: the positions left below k + w, in order
    const int e = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = c0 + blockIdx.x * 4 + wv;
    const int s = e < w ? sel[e] : -1;
    const bool valid = e < w && s >= k;
    const bool inq = valid && s < k + w;
    if (e < CA_W) s_flag[wv][e] = 0;
    __syncthreads();
    if (inq) s_flag[wv][s - k] = 1;
    __syncthreads();
    const bool outq = valid && !inq;                 // position s is left free
    const unsigned long long below = (1ull << e) - 1;
    const unsigned long long vm = __ballot(outq);
    // row k + e leaves [k, k + w) (as many as positions are left, which a
    // complete selection guarantees; the bound keeps every write in range)
    const bool leaves = e < w && s_flag[wv][e] == 0;
    const unsigned long long dm0 = __ballot(leaves);
    const bool disp = leaves && __popcll(dm0 & below) < __popcll(vm);
    const unsigned long long dm = __ballot(disp);
    if (outq) s_vac[wv][__popcll(vm & below)] = s;
    __syncthreads();
    const bool mc = valid && s != k + e;
    const int vdst = disp ? s_vac[wv][__popcll(dm & below)] : -1;
    if (blockIdx.x == 0 && wv == 0) {
        const unsigned long long cm = __ballot(mc);
        const int nc = __popcll(cm);
        if (mc) {
            const int slot = __popcll(cm & below);
            moves[1 + slot] = k + e;
            moves[1 + CA_MOVES + slot] = s;
        }
        if (disp) {
            const int slot = nc + __popcll(dm & below);
            moves[1 + slot] = vdst;
            moves[1 + CA_MOVES + slot] = k + e;
        }
        if (e == 0) moves[0] = nc + __popcll(dm);
    }
    const bool act = col < c1, actp = perm && col == c1;
    T v1 = (T)0, v2 = (T)0;
    int p1 = 0, p2 = 0;
    if (act) {
        if (mc) v1 = A[s + (size_t)col * lda];
        if (disp) v2 = A[(k + e) + (size_t)col * lda];
    }
    if (actp) {
        if (mc) p1 = perm[s];
        if (disp) p2 = perm[k + e];
    }
    __syncthreads();
    if (act) {
        if (mc) A[(k + e) + (size_t)col * lda] = v1;
        if (disp) A[vdst + (size_t)col * lda] = v2;
    }
    if (actp) {
        if (mc) perm[k + e] = p1;
        if (disp) perm[vdst] = p2;
    }
}

// The moves of nsub consecutive sub-panels (lists CA_MSTRIDE apart), in
// order, on the columns [a0, a1) and [b0, b1): one column per 64 threads;
// the barriers (with their workgroup fence) make each sub-panel's writes
// visible to the next one's reads.
template <typename T>
__global__ void __launch_bounds__(256) k_calu_swap_replay(T *__restrict__ A, int lda, int a0, int a1, int b0, int b1,
                                                          const int *__restrict__ moves, int nsub)
{
    const int e = threadIdx.x & 63;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int na = a1 - a0;
    const int col = g < na ? a0 + g : b0 + (g - na);
    const bool colok = g < na || col < b1;
    for (int sp = 0; sp < nsub; sp++) {
        const int *mv = moves + (size_t)sp * CA_MSTRIDE;
        const bool act = colok && e < mv[0];
        T v = (T)0;
        if (act) v = A[mv[1 + CA_MOVES + e] + (size_t)col * lda];
        __syncthreads();
        if (act) A[mv[1 + e] + (size_t)col * lda] = v;
        __syncthreads();
    }
}

__global__ void k_lu_iota(int *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// After trtri of the unit lower factor: the strictly upper part of W
// cleared and its diagonal set to 1 (trtri leaves both unspecified).
template <typename T>
__global__ void k_lu_unit_lower(T *__restrict__ W, int n)
{
    const size_t tot = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        if (i <= j) W[e] = i == j ? (T)1 : (T)0;
    }
}

// A^{-1} = X P^T with X = U^{-1} L^{-1}: column k of X is column perm[k]
// of the inverse.  Grid-stride over the n x n elements, column by column.
template <typename T>
__global__ void k_lu_scatter_cols(const T *__restrict__ X, T *__restrict__ out, int n, const int *__restrict__ perm)
{
    const size_t tot = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), k = (int)(e / n);
        out[i + (size_t)perm[k] * n] = X[e];
    }
}

// After the interchanges: L21 = A21 U11^{-1} on rows [k + w, n) of the
// sub-panel (workgroups [0, nrowblk)), the factors of the chosen rows into
// the top w x w block (workgroup 0), and U12 = L11^{-1} A12 on the rows
// [k, k + w) of the block columns [k + w, c1) (workgroups nrowblk + strip).
template <typename T>
__global__ void __launch_bounds__(256) k_calu_l21(T *__restrict__ A, int lda, int n, int k, int w, int c1,
                                                  const T *__restrict__ lu11, int nrowblk)
{
    __shared__ T s_rd[CA_W];
    const int t = threadIdx.x;
    if (t < w) s_rd[t] = (T)1 / lu11[t + t * CA_W];
    const int b = blockIdx.x;
    if (b < nrowblk) {
        if (b == 0)
            for (int e = t; e < w * w; e += blockDim.x)
                A[(k + e % w) + (size_t)(k + e / w) * lda] = lu11[(e % w) + (e / w) * CA_W];
        __syncthreads();
        const int r = k + w + b * blockDim.x + t;
        if (r >= n) return;
        // x U11 = a, column by column: x_c = a_c / u_cc, then a_q -= x_c u_cq.
        // U11 is read from global memory at uniform addresses (scalar loads
        // into scalar registers); from LDS every one of the ~500 products
        // would be a broadcast LDS read.
        T a[CA_W];
#pragma unroll
        for (int c = 0; c < CA_W; c++) a[c] = c < w ? A[r + (size_t)(k + c) * lda] : (T)0;
#pragma unroll
        for (int c = 0; c < CA_W; c++) {
            if (c < w) {
                const T x = a[c] * s_rd[c];
                a[c] = x;
#pragma unroll
                for (int q = c + 1; q < CA_W; q++) a[q] = fma(-x, lu11[c + q * CA_W], a[q]);
            }
        }
#pragma unroll
        for (int c = 0; c < CA_W; c++)
            if (c < w) A[r + (size_t)(k + c) * lda] = a[c];
        return;
    }
    // U12 = L11^{-1} A12: one column per thread, held in registers;
    // column-oriented substitution with the unit lower L11 read at uniform
    // addresses (a serial per-column loop over an LDS copy was this
    // kernel's critical path)
    const int cs = k + w + (b - nrowblk) * CA_STRIP;
    const int nc = min(CA_STRIP, c1 - cs);
    if (t >= nc) return;
    const size_t col = (size_t)(cs + t) * lda;
    T x[CA_W];
#pragma unroll
    for (int i = 0; i < CA_W; i++) x[i] = i < w ? A[(k + i) + col] : (T)0;
#pragma unroll
    for (int p = 0; p < CA_W; p++) {
        if (p < w) {
#pragma unroll
            for (int i = p + 1; i < CA_W; i++) x[i] = fma(-lu11[i + p * CA_W], x[p], x[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < CA_W; i++)
        if (i < w) A[(k + i) + col] = x[i];
}

// Rank-w update of the block columns [c0, c1) on rows [k + w, n):
// A22 -= L21 U12, with L21 and U12 already in A.  Grid: (column strips,
// row chunks of rows_per_chunk).
template <typename T>
__global__ void __launch_bounds__(256) k_calu_update(T *__restrict__ A, int lda, int n, int k, int w, int c0, int c1,
                                                     int rows_per_chunk)
{
    __shared__ T U[CA_W][CA_STRIP];
    const int cs = c0 + blockIdx.x * CA_STRIP;
    const int nc = min(CA_STRIP, c1 - cs);
    const int t = threadIdx.x;
    for (int e = t; e < CA_W * CA_STRIP; e += blockDim.x) {
        const int i = e % CA_W, cc = e / CA_W;
        U[i][cc] = (i < w && cc < nc) ? A[(k + i) + (size_t)(cs + cc) * lda] : (T)0;
    }
    __syncthreads();
    const int r0 = k + w + blockIdx.y * rows_per_chunk, r1 = min(n, r0 + rows_per_chunk);
    for (int r = r0 + t; r < r1; r += blockDim.x) {
        T acc[CA_STRIP];
#pragma unroll
        for (int cc = 0; cc < CA_STRIP; cc++) acc[cc] = cc < nc ? A[r + (size_t)(cs + cc) * lda] : (T)0;
#pragma unroll 4
        for (int p = 0; p < w; p++) {
            const T l = A[r + (size_t)(k + p) * lda];
#pragma unroll
            for (int cc = 0; cc < CA_STRIP; cc++) acc[cc] = fma(-l, U[p][cc], acc[cc]);
        }
#pragma unroll
        for (int cc = 0; cc < CA_STRIP; cc++)
            if (cc < nc) A[r + (size_t)(cs + cc) * lda] = acc[cc];
    }
}

typedef double lu_d4 __attribute__((ext_vector_type(4)));

// Double precision: the rank-w update of k_calu_update on the matrix cores
// (v_mfma_f64_16x16x4f64, lane l: A[l % 16][l / 16], B[l / 16][l % 16],
// value v: D[l / 16 + 4 v][l % 16]).  Workgroup (x, y): rows [r0, r0 + 64)
// below the sub-panel, columns [jc, jc + 64) of [k + w, c1), one 16-column
// chunk per wave.  The product is taken transposed, (L21 U12)^T = U12^T
// L21^T: M = column, N = row, K = sub-panel column c; at step m lane l holds
// U12[c][j + l % 16] and L21[r + l % 16][c] (c = 4 m + l / 16: 16 lanes read
// 128 contiguous bytes of one column of L21), and value v of row tile t is
// the update of row r0 + 16 t + l % 16, column j + l / 16 + 4 v.
__global__ void __launch_bounds__(256) k_calu_update_mfma(double *__restrict__ A, int lda, int n, int k, int w,
                                                          int c1)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kk = lane >> 4, ii = lane & 15;
    const int r0 = k + w + blockIdx.x * 64;
    const int jc = k + w + blockIdx.y * 64 + 16 * wave;
    if (jc >= c1) return;   // uniform per wave
    const int j = jc + ii;
    lu_d4 acc[4];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) acc[rt] = lu_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < CA_W / 4; m++) {
        const int c = 4 * m + kk;
        const bool cv = c < w;
        const double h = (cv && j < c1) ? A[(k + c) + (size_t)j * lda] : 0.0;
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
            const int r = r0 + 16 * rt + ii;
            const double x = (cv && r < n) ? A[r + (size_t)(k + c) * lda] : 0.0;
            acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(h, x, acc[rt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        const int r = r0 + 16 * rt + ii;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int jj = jc + kk + 4 * v;
            if (r < n && jj < c1) A[r + (size_t)jj * lda] -= acc[rt][v];
        }
    }
}

// the block update on the matrix cores (double only; IEMIC_LU_MFMA=0 keeps
// k_calu_update for comparison runs)
static bool lu_mfma()
{
    static const bool on = [] {
        const char *v = std::getenv("IEMIC_LU_MFMA");
        return !(v && std::atoi(v) == 0);
    }();
    return on;
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

// Rows per selection group for a level of cnt rows: the smallest of 64,
// 128, 256 that holds them all, else CA_G (several groups).  (One group of
// 2 CA_G rows for merge levels of up to 1024 candidates, which saves the
// third tree level of row blocks above 8192 rows, measured slower: 69.7 us
// per launch against 33.6 + 32.0 us for the two levels it replaces.)
static int select_group(int cnt)
{
    for (int g = 64; g < CA_G; g *= 2)
        if (cnt <= g) return g;
    return CA_G;
}

// LAPACK interchanges (ipiv) computed by the final selection level and the
// row moves taken from them (IEMIC_LU_IPIV=1), instead of the moves worked
// out from the chosen rows by the swap kernel with ipiv left unwritten; read
// at every factorisation, so that one process can run both forms
static bool lapack_ipiv()
{
    const char *v = std::getenv("IEMIC_LU_IPIV");
    return v && std::atoi(v) != 0;
}

// two rows per thread for the 512-row groups (IEMIC_LU_SELECT2=0: one)
static bool select_two_rows()
{
    static const bool on = [] {
        const char *v = std::getenv("IEMIC_LU_SELECT2");
        return !(v && std::atoi(v) == 0);
    }();
    return on;
}

template <typename T, int G>
static void launch_select_g(int groups, hipStream_t s, const T *A, int lda, int k, int w, const int *in, int cnt,
                            int *out, int fin, T *lu11, int *ipiv, int *moves, int *info)
{
    if (G == CA_G && select_two_rows()) {
        if (fin)
            hipLaunchKernelGGL((k_calu_select2<T, G, true>), dim3(groups), dim3(G / 2), 0, s, A, lda, k, w, in, cnt,
                               out, lu11, ipiv, moves, info);
        else
            hipLaunchKernelGGL((k_calu_select2<T, G, false>), dim3(groups), dim3(G / 2), 0, s, A, lda, k, w, in, cnt,
                               out, lu11, ipiv, moves, info);
        return;
    }
    if (fin)
        hipLaunchKernelGGL((k_calu_select<T, G, true>), dim3(groups), dim3(G), 0, s, A, lda, k, w, in, cnt, out, lu11,
                           ipiv, moves, info);
    else
        hipLaunchKernelGGL((k_calu_select<T, G, false>), dim3(groups), dim3(G), 0, s, A, lda, k, w, in, cnt, out,
                           lu11, ipiv, moves, info);
}

template <typename T>
static void launch_select(int g, int groups, hipStream_t s, const T *A, int lda, int k, int w, const int *in, int cnt,
                          int *out, int fin, T *lu11, int *ipiv, int *moves, int *info)
{
    switch (g) {
    case 64: launch_select_g<T, 64>(groups, s, A, lda, k, w, in, cnt, out, fin, lu11, ipiv, moves, info); break;
    case 128: launch_select_g<T, 128>(groups, s, A, lda, k, w, in, cnt, out, fin, lu11, ipiv, moves, info); break;
    case 256: launch_select_g<T, 256>(groups, s, A, lda, k, w, in, cnt, out, fin, lu11, ipiv, moves, info); break;
    default: launch_select_g<T, CA_G>(groups, s, A, lda, k, w, in, cnt, out, fin, lu11, ipiv, moves, info);
    }
}

constexpr int INV_NB = 2048;  // block rows of dense_inverse's GEMM-based solve with U (default; 256: +5 %, 512: +1.9 %, 1024: +0.4 % build time at 180x80x24)

// block rows of the solve with U: INV_NB, or IEMIC_INV_NB (256, 512, 1024
// or 2048) for comparison runs
static int inv_block()
{
    static const int nb = [] {
        const char *v = std::getenv("IEMIC_INV_NB");
        const int b = v ? std::atoi(v) : INV_NB;
        return (b == 256 || b == 512 || b == 1024 || b == 2048) ? b : INV_NB;
    }();
    return nb;
}

static rocblas_status gemm_nn(rocblas_handle h, int m, int n, int k, double alpha, const double *A, int lda,
                              const double *B, int ldb, double beta, double *C, int ldc)
{
    return rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                         ldc);
}
static rocblas_status gemm_nn(rocblas_handle h, int m, int n, int k, float alpha, const float *A, int lda,
                              const float *B, int ldb, float beta, float *C, int ldc)
{
    return rocblas_sgemm(h, rocblas_operation_none, rocblas_operation_none, m, n, k, &alpha, A, lda, B, ldb, &beta, C,
                         ldc);
}

// Inside rocblas_start/stop_device_memory_size_query: the rocBLAS calls of
// dense_getrf at their largest shapes for an n x n block, so that the
// handle's workspace covers them.
void dense_getrf_size_query(rocblas_handle h, int n)
{
    double *dd = nullptr;
    float *df = nullptr;
    const double one = 1.0;
    const float onef = 1.0f;
    for (int nb = 128; nb <= CA_NB_MAX; nb *= 2)
        if (n > nb) {
            (void)trsm_llnu(h, nb, n - nb, dd, n, dd, n);
            (void)gemm_nn_sub(h, n - nb, n - nb, nb, dd, n, dd, n, dd, n);
            (void)trsm_llnu(h, nb, n - nb, df, n, df, n);
            (void)gemm_nn_sub(h, n - nb, n - nb, nb, df, n, df, n, df, n);
        }
    // the blocked solve of dense_inverse: its GEMMs and the diagonal-block solves
    for (int nb = 256; nb <= 2048; nb *= 2)
        if (n > nb) {
            (void)gemm_nn(h, nb, n, n - nb, -1.0, dd, n, dd, n, 1.0, dd, n);
            (void)rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                rocblas_diagonal_non_unit, nb, n, &one, dd, n, dd, n);
            (void)gemm_nn(h, nb, n, n - nb, -1.0f, df, n, df, n, 1.0f, df, n);
            (void)rocblas_strsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                rocblas_diagonal_non_unit, nb, n, &onef, df, n, df, n);
        }
    // the trtri and the n x n solve of dense_inverse
    (void)rocblas_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_unit, n, dd, n, dd, n);
    (void)rocblas_strtri(h, rocblas_fill_lower, rocblas_diagonal_unit, n, df, n, df, n);
    (void)rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit, n,
                        n, &one, dd, n, dd, n);
    (void)rocblas_strsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit, n,
                        n, &onef, df, n, df, n);
}

#define LU_BLASCHK(expr)                                                                        \
    do {                                                                                            \
        rocblas_status st_ = (expr);                                                                \
        if (st_ != rocblas_status_success)                                                          \
            return set_error(IEMIC_EHIP, std::string(#expr) + ": status " + std::to_string((int)st_)); \
    } while (0)

// A = P L U for the n x n column-major A (lda = n); perm, ipiv and info on
// the device; work holds dense_getrf_work_bytes(n) bytes of device memory.
// perm receives the row permutation (row i of P^T A is row perm[i] of A).
// ipiv (LAPACK's interchanges) is written only with IEMIC_LU_IPIV=1; by
// default the row moves of each sub-panel come from its chosen rows alone
// (k_calu_swap_sel), which takes the serial interchange sequence off the
// panel's critical path, and perm is then required.  With la (a second stream, its own rocBLAS handle and two
// events) the panels are factored one block ahead of the trailing update;
// every launch has finished with respect to stream s when the last one
// enqueued on s has.
template <typename T>
int dense_getrf(rocblas_handle h, hipStream_t s, int n, T *A, int *ipiv, int *info, void *work, int *perm,
                const LuAhead *la)
{
    if (n <= 0) return IEMIC_OK;
    if (!work) return set_error(IEMIC_EINVAL, "dense LU: no workspace");
    const bool lap = lapack_ipiv();
    if (!perm && !lap) return set_error(IEMIC_EINVAL, "dense LU: no row permutation buffer");
    const size_t ce = cand_elems(n);
    int *cand[2] = {reinterpret_cast<int *>(work), reinterpret_cast<int *>(work) + ce};
    T *lu11 = reinterpret_cast<T *>(reinterpret_cast<char *>(work) + ((2 * ce * sizeof(int) + 255) / 256) * 256);
    int *movebuf = reinterpret_cast<int *>(lu11 + CA_W * CA_W);
    HIPCHK(hipMemsetAsync(info, 0, sizeof(int), s));
    if (perm) hipLaunchKernelGGL(k_lu_iota, dim3((n + 255) / 256), dim3(256), 0, s, perm, n);
    const int NB = lu_block();
    const bool ahead = la && la->s2 && n > NB;
    // columns [K, K + nb): tournament pivoting and the in-block updates per
    // sub-panel; interchanges on the block's columns only (ahead) or on all
    // columns at once
    auto panel = [&](int K, int nb, int *mv) {
        for (int k = K, sp = 0; k < K + nb; k += CA_W, sp++) {
            const int w = std::min(CA_W, K + nb - k);
            int *moves = mv + (size_t)sp * CA_MSTRIDE;
            // tournament over rows [k, n)
            int cnt = n - k;
            const int *in = nullptr;
            int lev = 0;
            const int *chosen = nullptr;   // the final level's rows, in pivot order
            for (;;) {
                const int g = select_group(cnt), groups = (cnt + g - 1) / g;
                const int fin = groups == 1;
                int *out = cand[lev & 1];
                launch_select<T>(g, groups, s, A, n, k, w, in, cnt, out, fin, lu11, lap ? ipiv : nullptr, moves,
                                 info);
                if (fin) {
                    chosen = out;
                    break;
                }
                in = out;
                cnt = groups * CA_W;
                lev++;
            }
            const int s0 = ahead ? K : 0, s1 = ahead ? K + nb : n;
            if (lap)
                hipLaunchKernelGGL(k_calu_swap<T>, dim3((s1 - s0 + 1 + 3) / 4), dim3(256), 0, s, A, n, s0, s1, moves,
                                   perm);
            else
                hipLaunchKernelGGL(k_calu_swap_sel<T>, dim3((s1 - s0 + 1 + 3) / 4), dim3(256), 0, s, A, n, s0, s1, k,
                                   w, chosen, moves, perm);
            const int below = n - (k + w);
            const int nrowblk = std::max(1, (below + 255) / 256);
            const int c1 = K + nb;
            const int strips = (c1 - (k + w) + CA_STRIP - 1) / CA_STRIP;
            hipLaunchKernelGGL(k_calu_l21<T>, dim3(nrowblk + strips), dim3(256), 0, s, A, n, n, k, w, c1, lu11,
                               nrowblk);
            if constexpr (std::is_same<T, double>::value) {
                if (lu_mfma()) {
                    if (strips > 0 && below > 0)
                        hipLaunchKernelGGL(k_calu_update_mfma, dim3((below + 63) / 64, (c1 - (k + w) + 63) / 64),
                                           dim3(256), 0, s, A, n, n, k, w, c1);
                    continue;
                }
            }
            if (strips > 0 && below > 0) {
                const int chunk = 256;
                dim3 grid(strips, (below + chunk - 1) / chunk);
                hipLaunchKernelGGL(k_calu_update<T>, grid, dim3(256), 0, s, A, n, n, k, w, k + w, c1, chunk);
            }
        }
    };
    if (!ahead) {
        for (int K = 0; K < n; K += NB) {
            const int nb = std::min(NB, n - K);
            panel(K, nb, movebuf);
            const int c0 = K + nb;
            if (c0 < n) {
                LU_BLASCHK(trsm_llnu(h, nb, n - c0, A + K + (size_t)K * n, n, A + K + (size_t)c0 * n, n));
                LU_BLASCHK(gemm_nn_sub(h, n - c0, n - c0, nb, A + c0 + (size_t)K * n, n, A + K + (size_t)c0 * n, n,
                                       A + c0 + (size_t)c0 * n, n));
            }
        }
        HIPCHK(hipGetLastError());
        return IEMIC_OK;
    }
    // look-ahead: panel b + 1 on s while block b's update of the columns
    // beyond it runs on s2
    hipStream_t s2 = la->s2;
    rocblas_handle h2 = la->h2;
    const int nblocks = (n + NB - 1) / NB;
    panel(0, std::min(NB, n), movebuf);
    for (int b = 0; b < nblocks; b++) {
        const int K = b * NB, nb = std::min(NB, n - K), c0 = K + nb;
        const int nb1 = c0 < n ? std::min(NB, n - c0) : 0;   // the next panel's width
        const int c2 = c0 + nb1;                              // first column beyond the next panel
        const int nsub = (nb + CA_W - 1) / CA_W;
        int *mv = movebuf + (size_t)(b & 1) * CA_SUBS * CA_MSTRIDE;
        const T *L11 = A + K + (size_t)K * n, *L21 = A + c0 + (size_t)K * n;
        HIPCHK(hipEventRecord(la->ev_panel, s));   // panel b factored
        if (nb1 > 0) {
            // the next panel's columns: block b - 1's update of them (on s2)
            // first, then block b's interchanges, U12 and update, then its
            // factorisation
            if (b > 0) HIPCHK(hipStreamWaitEvent(s, la->ev_rest, 0));
            hipLaunchKernelGGL(k_calu_swap_replay<T>, dim3((nb1 + 3) / 4), dim3(256), 0, s, A, n, c0, c2, 0, 0, mv,
                               nsub);
            LU_BLASCHK(trsm_llnu(h, nb, nb1, L11, n, A + K + (size_t)c0 * n, n));
            LU_BLASCHK(gemm_nn_sub(h, n - c0, nb1, nb, L21, n, A + K + (size_t)c0 * n, n, A + c0 + (size_t)c0 * n, n));
            panel(c0, nb1, movebuf + (size_t)((b + 1) & 1) * CA_SUBS * CA_MSTRIDE);
        }
        // s2: block b's interchanges on the columns left of it and beyond the
        // next panel, then U12 and the trailing update beyond the next panel
        HIPCHK(hipStreamWaitEvent(s2, la->ev_panel, 0));
        const int nrest = n - c2;
        if (K + nrest > 0)
            hipLaunchKernelGGL(k_calu_swap_replay<T>, dim3((K + nrest + 3) / 4), dim3(256), 0, s2, A, n, 0, K, c2, n,
                               mv, nsub);
        if (nrest > 0) {
            LU_BLASCHK(trsm_llnu(h2, nb, nrest, L11, n, A + K + (size_t)c2 * n, n));
            LU_BLASCHK(gemm_nn_sub(h2, n - c0, nrest, nb, L21, n, A + K + (size_t)c2 * n, n, A + c0 + (size_t)c2 * n,
                                   n));
        }
        HIPCHK(hipEventRecord(la->ev_rest, s2));
    }
    // everything after this call on s waits for the last trailing work
    HIPCHK(hipStreamWaitEvent(s, la->ev_rest, 0));
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

static rocblas_status trsm_lunn(rocblas_handle h, int n, const double *U, double *B)
{
    const double one = 1.0;
    return rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                         n, n, &one, U, n, B, n);
}
static rocblas_status trsm_lunn(rocblas_handle h, int n, const float *U, float *B)
{
    const float one = 1.0f;
    return rocblas_strsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                         n, n, &one, U, n, B, n);
}

static rocblas_status trtri_lower_unit(rocblas_handle h, int n, const double *L, double *W)
{
    return rocblas_dtrtri(h, rocblas_fill_lower, rocblas_diagonal_unit, n, L, n, W, n);
}
static rocblas_status trtri_lower_unit(rocblas_handle h, int n, const float *L, float *W)
{
    return rocblas_strtri(h, rocblas_fill_lower, rocblas_diagonal_unit, n, L, n, W, n);
}

static rocblas_status trsm_lunn_block(rocblas_handle h, int m, int n, const double *U, int ldu, double *B, int ldb)
{
    const double one = 1.0;
    return rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                         m, n, &one, U, ldu, B, ldb);
}
static rocblas_status trsm_lunn_block(rocblas_handle h, int m, int n, const float *U, int ldu, float *B, int ldb)
{
    const float one = 1.0f;
    return rocblas_strsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none, rocblas_diagonal_non_unit,
                         m, n, &one, U, ldu, B, ldb);
}

// the blocked solve with U (IEMIC_INV_BLOCKED=0: one rocBLAS TRSM)
static bool inv_blocked()
{
    static const bool on = [] {
        const char *v = std::getenv("IEMIC_INV_BLOCKED");
        return !(v && std::atoi(v) == 0);
    }();
    return on;
}

// A^{-1} from the factors of dense_getrf, written over them: W = L^{-1}
// (trtri, n^3/3 flops), X = U^{-1} W (n^3), then the columns of X scattered
// by the row permutation (A^{-1} = U^{-1} L^{-1} P^T).  Against two
// triangular solves with the permuted identity (2 n^3) this saves a third of
// the inverse's flops.  W is n x n workspace.
//
// The solve with U runs by block rows of INV_NB from the bottom, in place
// on W:  W_i <- W_i - U_i,>i X_>i  (one GEMM of INV_NB x n x the rows
// below), then  X_i = U_ii^{-1} W_i  by a triangular solve with the
// diagonal block alone.  The GEMMs carry all but INV_NB / n of the flops:
// rocBLAS's TRSM of a whole 8192 x 8192 block ran at 18 TFLOP/s on one
// MI355X against 71 for its GEMM.  (Multiplying by explicit inverses of the
// diagonal blocks instead was as fast but lost two digits on the
// ill-conditioned global basins.)
template <typename T>
int dense_inverse(rocblas_handle h, hipStream_t s, int n, T *LU, const int *perm, T *W)
{
    if (n <= 0) return IEMIC_OK;
    const size_t tot = (size_t)n * n;
    const unsigned grid = (unsigned)std::min<size_t>((tot + 255) / 256, 8192);
    LU_BLASCHK(trtri_lower_unit(h, n, LU, W));
    hipLaunchKernelGGL(k_lu_unit_lower<T>, dim3(grid), dim3(256), 0, s, W, n);
    HIPCHK(hipGetLastError());
    const int NB = inv_block();
    if (!inv_blocked() || n <= NB) {
        LU_BLASCHK(trsm_lunn(h, n, LU, W));
    } else {
        for (int i0 = ((n - 1) / NB) * NB; i0 >= 0; i0 -= NB) {
            const int b = std::min(NB, n - i0), r = i0 + b;
            if (r < n)
                LU_BLASCHK(gemm_nn(h, b, n, n - r, (T)-1, LU + i0 + (size_t)r * n, n, W + r, n, (T)1, W + i0, n));
            LU_BLASCHK(trsm_lunn_block(h, b, n, LU + i0 + (size_t)i0 * n, n, W + i0, n));
        }
    }
    hipLaunchKernelGGL(k_lu_scatter_cols<T>, dim3(grid), dim3(256), 0, s, W, LU, n, perm);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

template int dense_getrf<double>(rocblas_handle, hipStream_t, int, double *, int *, int *, void *, int *,
                                 const LuAhead *);
template int dense_getrf<float>(rocblas_handle, hipStream_t, int, float *, int *, int *, void *, int *,
                                const LuAhead *);
template int dense_inverse<double>(rocblas_handle, hipStream_t, int, double *, const int *, double *);
template int dense_inverse<float>(rocblas_handle, hipStream_t, int, float *, const int *, float *);

}  // namespace iemic
