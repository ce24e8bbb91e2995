// Block ILU(0) of a general sparse matrix on the GPU, behind the C interface
// of the reference's MRILU plug-in (mrilucpp.F90:120-554, declared for C++ in
// Ifpack_MRILU.cpp:25-37): create from a 0-based CSR with an integer handle,
// set_params, compute, apply, destroy.
//
// The factorisation is not MRILU's multilevel renumbering: it is the
// incomplete LU with no fill (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., sec. 10.3) over b x b blocks (b = the MRILU parameter
// `blocksize`), on the block pattern of the matrix:
//   for block row I in order, for K < I with A_IK in the pattern:
//       A_IK <- A_IK D_K^{-1};   A_IJ <- A_IJ - A_IK A_KJ  for J > K with
//       both A_IJ and A_KJ in the pattern;
//   D_I = A_II, inverted with partial pivoting.
// A pivot below lutol times the largest entry of its block is replaced by
// that threshold (keeping its sign), so singular blocks still give a usable
// operator (MRILU's `lutol`).  Rows of one dependency level are independent:
// the factorisation and both sweeps of the apply run one launch per level,
// one thread per block row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace iemic {

constexpr int BILU_BMAX = 8;

struct BiluParams {
    int blocksize = 1, cutmck = 0, scarow = 0, xactelm = 0, clsonce = 0, gusmod = 0, ilutype = 9, singlu = 1;
    int outlev = 2;
    double nlsfctr = 1.0, epsw = 1.0e-2, elmfctr = 0.1, gusfctr = 1.0, redfctr = 0.8, schtol = 1.0e-6;
    double denslim = 1.0e-3, globfrac = 1.0, locfrac = 1.0, sparslim = 0.675, droptol = 1.0e-4;
    double compfct = 1.0, cpivtol = 0.1, lutol = 1.0e-14;
};

struct Bilu {
    BiluParams par;
    bool created = false, computed = false;
    int n = 0, b = 1, nbr = 0, nnzb = 0;
    // the matrix as handed over (0-based CSR), kept until compute
    std::vector<int> rp, ci;
    std::vector<double> co;
    // device factors
    hipStream_t stream = nullptr;
    int *d_rp = nullptr, *d_ci = nullptr, *d_diag = nullptr, *d_ordf = nullptr, *d_ordb = nullptr;
    double *d_val = nullptr, *d_dinv = nullptr, *d_x = nullptr, *d_y = nullptr;
    std::vector<int> offf, offb;   // level offsets into the forward / backward orders
    int pivots_floored = 0;

    void release()
    {
        for (void *p : {(void *)d_rp, (void *)d_ci, (void *)d_diag, (void *)d_ordf, (void *)d_ordb, (void *)d_val,
                        (void *)d_dinv, (void *)d_x, (void *)d_y})
            if (p) (void)hipFree(p);
        d_rp = d_ci = d_diag = d_ordf = d_ordb = nullptr;
        d_val = d_dinv = d_x = d_y = nullptr;
        rp.clear(); ci.clear(); co.clear();
        offf.clear(); offb.clear();
        created = computed = false;
    }
};

constexpr int BILU_MAX_INST = 64;   // instances at once (the reference's max_inst)
static Bilu *g_bilu[BILU_MAX_INST] = {};
static std::mutex g_bilu_mu;

// ---------------------------------------------------------------- kernels
// block entries are row-major b x b: element (r, c) of block q at q b^2 + r b + c

__device__ int bilu_find(const int *ci, int lo, int hi, int col)   // position of col in ci[lo, hi), or -1
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int v = ci[mid];
        if (v == col) return mid;
        if (v < col) lo = mid + 1; else hi = mid;
    }
    return -1;
}

// Gauss-Jordan inverse of a b x b block with partial pivoting; pivots below
// tol * max|D| are replaced by that threshold.  Returns the floored count.
__device__ int bilu_invert(const double *D, double *Dinv, int b, double tol)
{
    double a[BILU_BMAX][2 * BILU_BMAX];
    double mx = 0.0;
    for (int r = 0; r < b; r++)
        for (int c = 0; c < b; c++) {
            a[r][c] = D[r * b + c];
            a[r][b + c] = (r == c) ? 1.0 : 0.0;
            mx = fmax(mx, fabs(a[r][c]));
        }
    const double floor_ = (mx > 0.0 ? mx : 1.0) * tol;
    int floored = 0;
    for (int c = 0; c < b; c++) {
        int p = c;
        for (int r = c + 1; r < b; r++)
            if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
        if (p != c)
            for (int q = 0; q < 2 * b; q++) {
                const double t = a[c][q];
                a[c][q] = a[p][q];
                a[p][q] = t;
            }
        double piv = a[c][c];
        if (fabs(piv) < floor_) {
            piv = piv < 0.0 ? -floor_ : floor_;
            a[c][c] = piv;
            floored++;
        }
        const double ip = 1.0 / piv;
        for (int q = 0; q < 2 * b; q++) a[c][q] *= ip;
        for (int r = 0; r < b; r++) {
            if (r == c) continue;
            const double f = a[r][c];
            if (f == 0.0) continue;
            for (int q = 0; q < 2 * b; q++) a[r][q] -= f * a[c][q];
        }
    }
    for (int r = 0; r < b; r++)
        for (int c = 0; c < b; c++) Dinv[r * b + c] = a[r][b + c];
    return floored;
}

__global__ void k_bilu_factor(const int *__restrict__ order, int cnt, const int *__restrict__ rp,
                              const int *__restrict__ ci, const int *__restrict__ diag, double *__restrict__ val,
                              double *__restrict__ dinv, int b, double tol, int *__restrict__ floored)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int I = order[t];
    const int bb = b * b;
    double tmp[BILU_BMAX * BILU_BMAX];
    for (int p = rp[I]; p < diag[I]; p++) {   // K < I, ascending
        const int K = ci[p];
        double *AIK = val + (size_t)p * bb;
        const double *DK = dinv + (size_t)K * bb;
        for (int r = 0; r < b; r++)
            for (int c = 0; c < b; c++) {
                double s = 0.0;
                for (int q = 0; q < b; q++) s += AIK[r * b + q] * DK[q * b + c];
                tmp[r * b + c] = s;
            }
        for (int e = 0; e < bb; e++) AIK[e] = tmp[e];
        for (int p2 = p + 1; p2 < rp[I + 1]; p2++) {   // J > K in row I
            const int J = ci[p2];
            const int pk = bilu_find(ci, diag[K] + 1, rp[K + 1], J);
            if (pk < 0) continue;
            double *AIJ = val + (size_t)p2 * bb;
            const double *AKJ = val + (size_t)pk * bb;
            for (int r = 0; r < b; r++)
                for (int c = 0; c < b; c++) {
                    double s = 0.0;
                    for (int q = 0; q < b; q++) s += AIK[r * b + q] * AKJ[q * b + c];
                    AIJ[r * b + c] -= s;
                }
        }
    }
    const int f = bilu_invert(val + (size_t)diag[I] * bb, dinv + (size_t)I * bb, b, tol);
    if (f) atomicAdd(floored, f);
}

// y_I = r_I - sum_{K<I} L_IK y_K
__global__ void k_bilu_lower(const int *__restrict__ order, int cnt, const int *__restrict__ rp,
                             const int *__restrict__ ci, const int *__restrict__ diag,
                             const double *__restrict__ val, const double *__restrict__ r, double *__restrict__ y,
                             int b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int I = order[t];
    double acc[BILU_BMAX];
    for (int e = 0; e < b; e++) acc[e] = r[(size_t)I * b + e];
    for (int p = rp[I]; p < diag[I]; p++) {
        const int K = ci[p];
        const double *L = val + (size_t)p * b * b;
        for (int e = 0; e < b; e++) {
            double s = 0.0;
            for (int q = 0; q < b; q++) s += L[e * b + q] * y[(size_t)K * b + q];
            acc[e] -= s;
        }
    }
    for (int e = 0; e < b; e++) y[(size_t)I * b + e] = acc[e];
}

// x_I = D_I^{-1} (y_I - sum_{J>I} U_IJ x_J)
__global__ void k_bilu_upper(const int *__restrict__ order, int cnt, const int *__restrict__ rp,
                             const int *__restrict__ ci, const int *__restrict__ diag,
                             const double *__restrict__ val, const double *__restrict__ dinv,
                             const double *__restrict__ y, double *__restrict__ x, int b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int I = order[t];
    double acc[BILU_BMAX];
    for (int e = 0; e < b; e++) acc[e] = y[(size_t)I * b + e];
    for (int p = diag[I] + 1; p < rp[I + 1]; p++) {
        const int J = ci[p];
        const double *U = val + (size_t)p * b * b;
        for (int e = 0; e < b; e++) {
            double s = 0.0;
            for (int q = 0; q < b; q++) s += U[e * b + q] * x[(size_t)J * b + q];
            acc[e] -= s;
        }
    }
    const double *D = dinv + (size_t)I * b * b;
    for (int e = 0; e < b; e++) {
        double s = 0.0;
        for (int q = 0; q < b; q++) s += D[e * b + q] * acc[q];
        x[(size_t)I * b + e] = s;
    }
}

// ---------------------------------------------------------------- host side

static int bilu_lookup(int id, Bilu **out)
{
    if (id <= 0 || id > BILU_MAX_INST || !g_bilu[id - 1]) return set_error(IEMIC_EINVAL, "invalid preconditioner id");
    *out = g_bilu[id - 1];
    return IEMIC_OK;
}

static int bilu_create(int *id, int n, int nnz, const int *beg, const int *jco, const double *co)
{
    if (!id || !beg || (nnz > 0 && (!jco || !co))) return set_error(IEMIC_EINVAL, "create: null pointer");
    if (n <= 0 || nnz < 0) return set_error(IEMIC_EINVAL, "create: bad dimensions");
    if (beg[0] != 0 || beg[n] != nnz) return set_error(IEMIC_EINVAL, "create: row pointer is not 0-based or does not end at nnz");
    for (int r = 0; r < n; r++)
        if (beg[r + 1] < beg[r]) return set_error(IEMIC_EINVAL, "create: row pointer decreases");
    for (int q = 0; q < nnz; q++)
        if (jco[q] < 0 || jco[q] >= n) return set_error(IEMIC_EINVAL, "create: column index out of range");
    std::lock_guard<std::mutex> lock(g_bilu_mu);
    Bilu *p = nullptr;
    if (*id == 0) {   // a new instance: the first unused id
        int slot = -1;
        for (int q = 0; q < BILU_MAX_INST; q++)
            if (!g_bilu[q]) { slot = q; break; }
        if (slot < 0) return set_error(IEMIC_EUNSUPPORTED, "create: maximum number of instances exceeded");
        p = new Bilu();
        g_bilu[slot] = p;
        *id = slot + 1;
        HIPCHK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    } else {   // overwrite an existing instance (the reference resets it)
        int rc = bilu_lookup(*id, &p);
        if (rc) return rc;
        hipStream_t s = p->stream;
        p->release();
        p->stream = s;
        p->par = BiluParams();
    }
    p->n = n;
    p->rp.assign(beg, beg + n + 1);
    p->ci.assign(jco, jco + nnz);
    p->co.assign(co, co + nnz);
    p->created = true;
    p->computed = false;
    return IEMIC_OK;
}

static int bilu_compute(Bilu *p)
{
    if (!p->created) return set_error(IEMIC_ESTATE, "compute: no matrix (create must be called first)");
    if (p->computed) return set_error(IEMIC_ESTATE, "compute: a matrix can be factored only once; create it again");
    const int n = p->n, b = p->par.blocksize;
    if (b < 1 || b > BILU_BMAX) return set_error(IEMIC_EUNSUPPORTED, "compute: blocksize must be 1.." + std::to_string(BILU_BMAX));
    if (n % b) return set_error(IEMIC_EINVAL, "compute: the dimension is not a multiple of blocksize");
    const int nbr = n / b;
    // block pattern: sorted unique block columns per block row, diagonal included
    std::vector<int> brp(nbr + 1, 0), bci;
    std::vector<int> mark(nbr, -1), cols;
    std::vector<std::vector<int>> rows(nbr);
    for (int I = 0; I < nbr; I++) {
        cols.clear();
        cols.push_back(I);
        mark[I] = I;
        for (int r = I * b; r < (I + 1) * b; r++)
            for (int q = p->rp[r]; q < p->rp[r + 1]; q++) {
                const int J = p->ci[q] / b;
                if (mark[J] != I) { mark[J] = I; cols.push_back(J); }
            }
        std::sort(cols.begin(), cols.end());
        rows[I] = cols;
        brp[I + 1] = brp[I] + (int)cols.size();
    }
    const int nnzb = brp[nbr];
    bci.reserve(nnzb);
    std::vector<int> diag(nbr);
    for (int I = 0; I < nbr; I++) {
        for (size_t q = 0; q < rows[I].size(); q++)
            if (rows[I][q] == I) diag[I] = brp[I] + (int)q;
        bci.insert(bci.end(), rows[I].begin(), rows[I].end());
    }
    const size_t bb = (size_t)b * b;
    std::vector<double> val((size_t)nnzb * bb, 0.0);
    for (int r = 0; r < n; r++) {
        const int I = r / b, ri = r % b;
        for (int q = p->rp[r]; q < p->rp[r + 1]; q++) {
            const int J = p->ci[q] / b, cj = p->ci[q] % b;
            const auto it = std::lower_bound(rows[I].begin(), rows[I].end(), J);
            val[(size_t)(brp[I] + (it - rows[I].begin())) * bb + ri * b + cj] += p->co[q];
        }
    }
    // dependency levels: forward over K < I, backward over J > I
    std::vector<int> lf(nbr, 0), lb(nbr, 0);
    int nlf = 0, nlb = 0;
    for (int I = 0; I < nbr; I++) {
        int lv = 0;
        for (int q = brp[I]; q < diag[I]; q++) lv = std::max(lv, lf[bci[q]] + 1);
        lf[I] = lv;
        nlf = std::max(nlf, lv + 1);
    }
    for (int I = nbr - 1; I >= 0; I--) {
        int lv = 0;
        for (int q = diag[I] + 1; q < brp[I + 1]; q++) lv = std::max(lv, lb[bci[q]] + 1);
        lb[I] = lv;
        nlb = std::max(nlb, lv + 1);
    }
    auto schedule = [nbr](const std::vector<int> &lev, int nl, std::vector<int> &off, std::vector<int> &ord) {
        off.assign(nl + 1, 0);
        for (int I = 0; I < nbr; I++) off[lev[I] + 1]++;
        for (int q = 0; q < nl; q++) off[q + 1] += off[q];
        std::vector<int> fill(off.begin(), off.end() - 1);
        ord.assign(nbr, 0);
        for (int I = 0; I < nbr; I++) ord[fill[lev[I]]++] = I;
    };
    std::vector<int> ordf, ordb;
    schedule(lf, nlf, p->offf, ordf);
    schedule(lb, nlb, p->offb, ordb);
    auto up = [](void **dst, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) return e;
        return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    HIPCHK(up((void **)&p->d_rp, brp.data(), brp.size() * sizeof(int)));
    HIPCHK(up((void **)&p->d_ci, bci.data(), bci.size() * sizeof(int)));
    HIPCHK(up((void **)&p->d_diag, diag.data(), diag.size() * sizeof(int)));
    HIPCHK(up((void **)&p->d_ordf, ordf.data(), ordf.size() * sizeof(int)));
    HIPCHK(up((void **)&p->d_ordb, ordb.data(), ordb.size() * sizeof(int)));
    HIPCHK(up((void **)&p->d_val, val.data(), val.size() * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_dinv, (size_t)nbr * bb * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_x, (size_t)n * sizeof(double)));
    HIPCHK(hipMalloc(&p->d_y, (size_t)n * sizeof(double)));
    int *d_fl = nullptr;
    HIPCHK(hipMalloc(&d_fl, sizeof(int)));
    HIPCHK(hipMemsetAsync(d_fl, 0, sizeof(int), p->stream));
    for (int L = 0; L < nlf; L++) {
        const int cnt = p->offf[L + 1] - p->offf[L];
        hipLaunchKernelGGL(k_bilu_factor, dim3((cnt + 127) / 128), dim3(128), 0, p->stream, p->d_ordf + p->offf[L], cnt,
                           p->d_rp, p->d_ci, p->d_diag, p->d_val, p->d_dinv, b, p->par.lutol, d_fl);
    }
    HIPCHK(hipGetLastError());
    int fl = 0;
    HIPCHK(hipMemcpyAsync(&fl, d_fl, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    (void)hipFree(d_fl);
    p->pivots_floored = fl;
    p->b = b;
    p->nbr = nbr;
    p->nnzb = nnzb;
    p->computed = true;
    p->created = false;   // the reference releases the matrix once it is factored
    p->rp.clear(); p->ci.clear(); p->co.clear();
    if (p->par.outlev >= 3)
        std::fprintf(stderr, "bilu: n %d, blocksize %d, %d block entries, %d / %d levels, %d pivots floored\n", n, b,
                     nnzb, nlf, nlb, fl);
    return IEMIC_OK;
}

static int bilu_apply(Bilu *p, int n, const double *rhs, double *sol)
{
    if (!rhs || !sol) return set_error(IEMIC_EINVAL, "apply: null pointer");
    if (!p->computed) {   // the reference warns and copies the right-hand side
        if (n != p->n) return set_error(IEMIC_EINVAL, "apply: dimension mismatch");
        std::fprintf(stderr, "WARNING: apply called before the preconditioner was computed\n");
        std::copy(rhs, rhs + n, sol);
        return IEMIC_OK;
    }
    if (n != p->n) return set_error(IEMIC_EINVAL, "apply: dimension mismatch");
    HIPCHK(hipMemcpyAsync(p->d_x, rhs, (size_t)n * sizeof(double), hipMemcpyHostToDevice, p->stream));
    for (size_t L = 0; L + 1 < p->offf.size(); L++) {
        const int cnt = p->offf[L + 1] - p->offf[L];
        hipLaunchKernelGGL(k_bilu_lower, dim3((cnt + 127) / 128), dim3(128), 0, p->stream, p->d_ordf + p->offf[L], cnt,
                           p->d_rp, p->d_ci, p->d_diag, p->d_val, p->d_x, p->d_y, p->b);
    }
    for (size_t L = 0; L + 1 < p->offb.size(); L++) {
        const int cnt = p->offb[L + 1] - p->offb[L];
        hipLaunchKernelGGL(k_bilu_upper, dim3((cnt + 127) / 128), dim3(128), 0, p->stream, p->d_ordb + p->offb[L], cnt,
                           p->d_rp, p->d_ci, p->d_diag, p->d_val, p->d_dinv, p->d_y, p->d_x, p->b);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(sol, p->d_x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return IEMIC_OK;
}

}  // namespace iemic

using namespace iemic;

extern "C" {

int iemic_mrilu_create(int *id, int n, int nnz, const int *beg, const int *jco, const double *co)
{
    return bilu_create(id, n, nnz, beg, jco, co);
}

int iemic_mrilu_set_params(int id, int blocksize, int cutmck, int scarow, int xactelm, int clsonce, double nlsfctr,
                           double epsw, double elmfctr, int gusmod, double gusfctr, double redfctr, double schtol,
                           double denslim, double globfrac, double locfrac, double sparslim, int ilutype,
                           double droptol, double compfct, double cpivtol, double lutol, int singlu, int outlev)
{
    std::lock_guard<std::mutex> lock(g_bilu_mu);
    Bilu *p = nullptr;
    if (int rc = bilu_lookup(id, &p)) return rc;
    if (blocksize < 1 || blocksize > BILU_BMAX)
        return set_error(IEMIC_EUNSUPPORTED, "set_params: blocksize must be 1.." + std::to_string(BILU_BMAX));
    if (!(lutol >= 0.0)) return set_error(IEMIC_EINVAL, "set_params: lutol must be >= 0");
    BiluParams &q = p->par;
    q.blocksize = blocksize; q.cutmck = cutmck; q.scarow = scarow; q.xactelm = xactelm; q.clsonce = clsonce;
    q.nlsfctr = nlsfctr; q.epsw = epsw; q.elmfctr = elmfctr; q.gusmod = gusmod; q.gusfctr = gusfctr;
    q.redfctr = redfctr; q.schtol = schtol; q.denslim = denslim; q.globfrac = globfrac; q.locfrac = locfrac;
    q.sparslim = sparslim; q.ilutype = ilutype; q.droptol = droptol; q.compfct = compfct; q.cpivtol = cpivtol;
    q.lutol = lutol; q.singlu = singlu; q.outlev = outlev;
    return IEMIC_OK;
}

int iemic_mrilu_compute(int id)
{
    std::lock_guard<std::mutex> lock(g_bilu_mu);
    Bilu *p = nullptr;
    if (int rc = bilu_lookup(id, &p)) return rc;
    return bilu_compute(p);
}

int iemic_mrilu_apply(int id, int n, const double *rhs, double *sol)
{
    Bilu *p = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_bilu_mu);
        if (int rc = bilu_lookup(id, &p)) return rc;
    }
    return bilu_apply(p, n, rhs, sol);
}

int iemic_mrilu_info(int id, int *blocks, int *levels, int *pivots_floored)
{
    std::lock_guard<std::mutex> lock(g_bilu_mu);
    Bilu *p = nullptr;
    if (int rc = bilu_lookup(id, &p)) return rc;
    if (!p->computed) return set_error(IEMIC_ESTATE, "info: not computed");
    if (blocks) *blocks = p->nnzb;
    if (levels) *levels = (int)p->offf.size() - 1;
    if (pivots_floored) *pivots_floored = p->pivots_floored;
    return IEMIC_OK;
}

int iemic_mrilu_destroy(int *id)
{
    if (!id) return set_error(IEMIC_EINVAL, "destroy: null pointer");
    std::lock_guard<std::mutex> lock(g_bilu_mu);
    if (*id > 0 && *id <= BILU_MAX_INST && g_bilu[*id - 1]) {
        Bilu *p = g_bilu[*id - 1];
        p->release();
        if (p->stream) (void)hipStreamDestroy(p->stream);
        delete p;
        g_bilu[*id - 1] = nullptr;
    }
    *id = 0;   // so the caller can create again
    return IEMIC_OK;
}

// The reference's own symbols (mrilucpp.F90, bind(C)), with its argument
// passing (every scalar by address) and its error behaviour (Fortran `stop`:
// the message, then the process ends).
static void stop_on(int rc)
{
    if (rc) {
        std::fprintf(stderr, "mrilucpp: %s\n", iemic_last_error());
        std::exit(1);
    }
}
void mrilucpp_create(int *id, int *n, int *nnz, int *beg, int *jco, double *co)
{
    stop_on(iemic_mrilu_create(id, *n, *nnz, beg, jco, co));
}
void mrilucpp_destroy(int *id) { stop_on(iemic_mrilu_destroy(id)); }
void mrilucpp_set_params(const int *id, int *blocksize, int *cutmck, int *scarow, int *xactelm, int *clsonce,
                         double *nlsfctr, double *epsw, double *elmfctr, int *gusmod, double *gusfctr, double *redfctr,
                         double *schtol, double *denslim, double *globfrac, double *locfrac, double *sparslim,
                         int *ilutype, double *droptol, double *compfct, double *cpivtol, double *lutol, int *singlu,
                         int *outlev)
{
    stop_on(iemic_mrilu_set_params(*id, *blocksize, *cutmck, *scarow, *xactelm, *clsonce, *nlsfctr, *epsw, *elmfctr,
                                   *gusmod, *gusfctr, *redfctr, *schtol, *denslim, *globfrac, *locfrac, *sparslim,
                                   *ilutype, *droptol, *compfct, *cpivtol, *lutol, *singlu, *outlev));
}
void mrilucpp_compute(const int *id) { stop_on(iemic_mrilu_compute(*id)); }
void mrilucpp_apply(const int *id, int *dim, const double *rhs, double *sol)
{
    stop_on(iemic_mrilu_apply(*id, *dim, rhs, sol));
}

}  // extern "C"
