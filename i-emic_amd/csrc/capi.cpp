// extern "C" layer of include/iemic_amd.h: argument checks, host/device
// staging, and the Ocean-facade entry points.  The numerical work is in
// host_setup.cpp (tables), assembly.hip (F, J, products), precond.hip and
// solver.hip.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"

using namespace iemic;

namespace iemic {
const char *last_error();
int precond_build(iemic_ctx *c, int precision);
int precond_apply(iemic_ctx *c, int precision, const double *in, double *out, float *tmp32);
void precond_free(iemic_ctx *c);
int precond_check(iemic_ctx *c);
int precond_info(iemic_ctx *c, iemic_precond_info *info);
int krylov_solve(iemic_ctx *c, const iemic_solver_params *p, const double *d_b, double *d_x,
                 iemic_solve_info *info, bool transpose);
int precond_apply_t(iemic_ctx *c, int precision, const double *in, double *out);
int krylov_solve_block(iemic_ctx *c, const iemic_solver_params *p, int q, const double *d_B, double *d_X,
                       iemic_solve_info *infos, bool transpose);
int krylov_solve_multi(iemic_ctx *c, const iemic_solver_params *p, int q, const double *d_B, double *d_X,
                       iemic_solve_info *infos, bool transpose);
int precond_apply_multi(iemic_ctx *c, int precision, int nrhs, const double *in, size_t ldin, double *out,
                        size_t ldout, float *tmp32);
void krylov_free(iemic_ctx *c);
int rowlu_tuned_blocks(iemic_ctx *c, int nrhs);
int newton_dev(iemic_ctx *c, const iemic_solver_params *sp, double tol, double step_tol, int maxit, double *d_x,
               int *iterations, double *fnorm, int *linear);
int continuation_dev(iemic_ctx *c, const iemic_solver_params *sp, const iemic_continuation_params *cp, double *d_x,
                     double *par_hist, double *norm_hist, iemic_continuation_info *out);
}  // namespace iemic

#define CHECK_CTX(c)                                                            \
    do {                                                                        \
        if (!(c)) return set_error(IEMIC_EINVAL, "null context");               \
    } while (0)

static int ensure_staging(iemic_ctx *c)
{
    if (!c->d_x) {
        HIPCHK(hipMalloc(&c->d_x, (size_t)c->sc.ndim * sizeof(double)));
        HIPCHK(hipMalloc(&c->d_y, (size_t)c->sc.ndim * sizeof(double)));
    }
    return IEMIC_OK;
}

extern "C" {

const char *iemic_last_error(void) { return iemic::last_error(); }

int iemic_device_count(int *count)
{
    if (!count) return set_error(IEMIC_EINVAL, "null pointer");
    HIPCHK(hipGetDeviceCount(count));
    return IEMIC_OK;
}

int iemic_create(const iemic_config *cfg, const int *landm, const double *taux, const double *tauy,
                 const double *tatm, const double *emip, const double *spert, iemic_ctx **out)
{
    if (!cfg || !landm || !out) return set_error(IEMIC_EINVAL, "null argument");
    *out = nullptr;
    iemic_ctx *c = new iemic_ctx();
    c->cfg = *cfg;
    int rc = host_init(c, landm, taux, tauy, tatm, emip, spert);
    if (rc != IEMIC_OK) {
        delete c;
        return rc;
    }
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        delete c;
        return set_error(IEMIC_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    rc = upload_tables(c);
    if (rc != IEMIC_OK) {
        iemic_destroy(c);
        return rc;
    }
    *out = c;
    return IEMIC_OK;
}

int iemic_destroy(iemic_ctx *c)
{
    if (!c) return IEMIC_OK;
    precond_free(c);
    krylov_free(c);
    void *bufs[] = {c->d_landm, c->d_al, c->d_jt, c->d_kt, c->d_frc, c->d_coef, c->d_diagB,
                    c->d_J, c->d_x, c->d_y, c->d_red, c->d_src, c->d_col_export, c->d_icrow};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &ev : c->ev)
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return IEMIC_OK;
}

int iemic_set_par(iemic_ctx *c, int idx, double value)
{
    CHECK_CTX(c);
    if (idx < 1 || idx > 30) return set_error(IEMIC_EINVAL, "parameter index must be 1..30");
    c->h.par[idx] = value;
    if (idx == 6 || idx == 29) {   // MIXP, MKAP: neutral physics / GM
        if (value != 0.0) return set_error(IEMIC_EUNSUPPORTED, "neutral physics (MIXP/MKAP) is not supported");
    }
    host_parameters_changed(c);
    return upload_tables(c);
}

int iemic_get_par(const iemic_ctx *c, int idx, double *value)
{
    CHECK_CTX(c);
    if (idx < 1 || idx > 30 || !value) return set_error(IEMIC_EINVAL, "parameter index must be 1..30");
    *value = c->h.par[idx];
    return IEMIC_OK;
}

int iemic_ndim(const iemic_ctx *c, int *ndim)
{
    CHECK_CTX(c);
    *ndim = c->sc.ndim;
    return IEMIC_OK;
}

int iemic_graph_nnz(const iemic_ctx *c, int *nnz)
{
    CHECK_CTX(c);
    if (!c->graph_built) build_graph(const_cast<iemic_ctx *>(c)), const_cast<iemic_ctx *>(c)->graph_built = true;
    *nnz = c->graph.nnz;
    return IEMIC_OK;
}

int iemic_rowintcon(const iemic_ctx *c, int *row)
{
    CHECK_CTX(c);
    *row = c->sc.rowintcon;
    return IEMIC_OK;
}

int iemic_intcond_coef(const iemic_ctx *c, double *coef)
{
    CHECK_CTX(c);
    std::memcpy(coef, c->h.coef.data(), sizeof(double) * c->sc.ndim);
    return IEMIC_OK;
}

int iemic_forcing(const iemic_ctx *c, double *frc)
{
    CHECK_CTX(c);
    std::memcpy(frc, c->h.frc.data(), sizeof(double) * c->sc.ndim);
    return IEMIC_OK;
}

int iemic_set_intcond_correction(iemic_ctx *c, const double *x)
{
    CHECK_CTX(c);
    double d = 0.0;
    for (int r = 0; r < c->sc.ndim; r++) d += c->h.coef[r] * x[r];
    c->intcorr = d;
    return IEMIC_OK;
}

int iemic_set_intcond_correction_value(iemic_ctx *c, double value)
{
    CHECK_CTX(c);
    c->intcorr = value;
    return IEMIC_OK;
}

int iemic_rhs_dev(iemic_ctx *c, const double *d_x, double *d_F)
{
    CHECK_CTX(c);
    return launch_rhs(c, d_x, d_F);
}

int iemic_rhs(iemic_ctx *c, const double *x, double *F)
{
    CHECK_CTX(c);
    int rc = ensure_staging(c);
    if (rc) return rc;
    size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_x, x, bytes, hipMemcpyHostToDevice, c->stream));
    rc = launch_rhs(c, c->d_x, c->d_y);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(F, c->d_y, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IEMIC_OK;
}

int iemic_jacobian_dev(iemic_ctx *c, const double *d_x)
{
    CHECK_CTX(c);
    return launch_jacobian(c, d_x);
}

int iemic_jacobian(iemic_ctx *c, const double *x)
{
    CHECK_CTX(c);
    int rc = ensure_staging(c);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->d_x, x, (size_t)c->sc.ndim * sizeof(double), hipMemcpyHostToDevice, c->stream));
    rc = launch_jacobian(c, c->d_x);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return IEMIC_OK;
}

// ---- the Fortran interface of THCM.C:49-176 (usrc.F90, mat.F90) --------
int iemic_thcm_get_array_sizes(const iemic_ctx *c, int *nrows, int *nnz)
{
    CHECK_CTX(c);
    if (!nrows || !nnz) return set_error(IEMIC_EINVAL, "null pointer");
    *nrows = c->sc.ndim;
    const long long cap = (long long)c->sc.ndim * (6 * 27 + 1);   // ndim (nun np + 1), mat.F90:58-70
    if (cap > INT32_MAX) return set_error(IEMIC_EUNSUPPORTED, "Fortran CSR capacity exceeds 32-bit indices");
    *nnz = (int)cap;
    return IEMIC_OK;
}

int iemic_thcm_set_pointers(iemic_ctx *c, int nrows, int nnz, int *begA, int *jcoA, double *coA, double *coB,
                            int *begF, int *jcoF, double *coF)
{
    CHECK_CTX(c);
    (void)begF; (void)jcoF; (void)coF;   // the stochastic-forcing matrix (get_stochastic_forcing) is not filled
    if (nrows != c->sc.ndim) return set_error(IEMIC_EINVAL, "set_pointers: nrows differs from the model's ndim");
    if (nnz < 1 || !begA || !jcoA || !coA || !coB) return set_error(IEMIC_EINVAL, "set_pointers: null or empty buffer");
    c->f_nrows = nrows;
    c->f_nnz = nnz;
    c->f_begA = begA;
    c->f_jcoA = jcoA;
    c->f_coA = coA;
    c->f_coB = coB;
    return IEMIC_OK;
}

int iemic_thcm_rhs(iemic_ctx *c, const double *un, double *B)
{
    CHECK_CTX(c);
    if (!un || !B) return set_error(IEMIC_EINVAL, "null pointer");
    int rc = ensure_staging(c);
    if (rc) return rc;
    const size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_x, un, bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = launch_rhs_fortran(c, c->d_x, c->d_y))) return rc;
    HIPCHK(hipMemcpyAsync(B, c->d_y, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int r = 0; r < c->sc.ndim; r++) B[r] = -B[r];   // F = -B (THCM.C:1011)
    return IEMIC_OK;
}

int iemic_thcm_matrix(iemic_ctx *c, const double *un)
{
    CHECK_CTX(c);
    if (!un) return set_error(IEMIC_EINVAL, "null pointer");
    if (!c->f_begA) return set_error(IEMIC_ESTATE, "matrix: call iemic_thcm_set_pointers first");
    int rc = ensure_staging(c);
    if (rc) return rc;
    const int nd = c->sc.ndim;
    HIPCHK(hipMemcpyAsync(c->d_x, un, (size_t)nd * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = launch_jacobian(c, c->d_x))) return rc;
    int *d_cnt = nullptr, *d_col = nullptr, *d_bad = nullptr;
    double *d_val = nullptr;
    std::vector<int> beg0(nd + 1);
    auto done = [&](int code) {
        (void)hipStreamSynchronize(c->stream);
        for (void *b : {(void *)d_cnt, (void *)d_col, (void *)d_bad, (void *)d_val})
            if (b) (void)hipFree(b);
        return code;
    };
    if (hipMalloc(&d_cnt, (size_t)(nd + 1) * sizeof(int)) != hipSuccess || hipMalloc(&d_bad, sizeof(int)) != hipSuccess)
        return done(set_error(IEMIC_ENOMEM, "matrix: staging"));
    if ((rc = launch_fortran_count(c, d_cnt))) return done(rc);
    if (hipMemcpyAsync(beg0.data(), d_cnt, (size_t)nd * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return done(set_error(IEMIC_EHIP, "matrix: row counts"));
    long long tot = 0;
    for (int r = 0; r < nd; r++) {
        const int cnt = beg0[r];
        beg0[r] = (int)tot;
        tot += cnt;
    }
    beg0[nd] = (int)tot;
    if (tot > c->f_nnz) return done(set_error(IEMIC_EINVAL, "matrix: the caller's CSR buffers are too small"));
    for (int r = 0; r <= nd; r++) c->f_begA[r] = beg0[r] + 1;   // 1-based, as fillcolA writes begA
    if (hipMalloc(&d_col, (size_t)std::max<long long>(tot, 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_val, (size_t)std::max<long long>(tot, 1) * sizeof(double)) != hipSuccess)
        return done(set_error(IEMIC_ENOMEM, "matrix: staging"));
    int bad = 0;
    if (hipMemcpyAsync(d_cnt, beg0.data(), (size_t)(nd + 1) * sizeof(int), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipMemsetAsync(d_bad, 0, sizeof(int), c->stream) != hipSuccess)
        return done(set_error(IEMIC_EHIP, "matrix: offsets"));
    if ((rc = launch_fortran_fill(c, d_cnt, d_col, d_val, d_bad))) return done(rc);
    if (hipMemcpyAsync(c->f_jcoA, d_col, (size_t)tot * sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->f_coA, d_val, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return done(set_error(IEMIC_EHIP, "matrix: copy back"));
    if (bad) return done(set_error(IEMIC_EINVAL, "matrix: entries outside the grid"));
    // fillcolB (assemble.F90:18-54): THCM's diag(B) except the integral
    // row, which THCM zeroes in its own copy (THCM.C:2197-2198)
    std::memcpy(c->f_coB, c->h.diagB.data(), (size_t)nd * sizeof(double));
    if (c->sc.rowintcon >= 0) c->f_coB[c->sc.rowintcon] = -1.0;
    return done(IEMIC_OK);
}

int iemic_jacobian_csr(iemic_ctx *c, int *rowptr, int *col, double *val, double *diagB)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    if (!c->graph_built) {
        build_graph(c);
        c->graph_built = true;
    }
    Graph &g = c->graph;
    if (!c->d_src) {
        HIPCHK(hipMalloc(&c->d_src, (size_t)g.nnz * sizeof(int)));
        HIPCHK(hipMalloc(&c->d_col_export, (size_t)g.nnz * sizeof(int)));
        HIPCHK(hipMemcpy(c->d_src, g.src.data(), (size_t)g.nnz * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_col_export, g.col.data(), (size_t)g.nnz * sizeof(int), hipMemcpyHostToDevice));
    }
    double *d_val = nullptr;
    HIPCHK(hipMalloc(&d_val, (size_t)g.nnz * sizeof(double)));
    int rc = launch_export(c, d_val);
    if (rc == IEMIC_OK) {
        hipError_t e = hipMemcpyAsync(val, d_val, (size_t)g.nnz * sizeof(double), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = set_error(IEMIC_EHIP, hipGetErrorString(e));
    }
    (void)hipFree(d_val);
    if (rc) return rc;
    if (rowptr) std::memcpy(rowptr, g.rowptr.data(), sizeof(int) * (c->sc.ndim + 1));
    if (col) std::memcpy(col, g.col.data(), sizeof(int) * g.nnz);
    if (diagB) std::memcpy(diagB, c->h.diagB.data(), sizeof(double) * c->sc.ndim);
    return IEMIC_OK;
}

int iemic_apply_matrix_dev(iemic_ctx *c, const double *d_v, double *d_out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    return launch_spmv(c, d_v, d_out);
}

int iemic_apply_matrix_transpose_dev(iemic_ctx *c, const double *d_v, double *d_out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    return launch_spmv_t(c, d_v, d_out);
}

int iemic_apply_matrix(iemic_ctx *c, const double *v, double *out)
{
    CHECK_CTX(c);
    int rc = ensure_staging(c);
    if (rc) return rc;
    size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_x, v, bytes, hipMemcpyHostToDevice, c->stream));
    rc = iemic_apply_matrix_dev(c, c->d_x, c->d_y);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, c->d_y, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IEMIC_OK;
}

int iemic_pressure_projection_dev(iemic_ctx *c, double *d_v)
{
    CHECK_CTX(c);
    if (!d_v) return set_error(IEMIC_EINVAL, "null pointer");
    return launch_pressure_projection(c, d_v);
}

int iemic_pressure_projection(iemic_ctx *c, double *v)
{
    CHECK_CTX(c);
    if (!v) return set_error(IEMIC_EINVAL, "null pointer");
    int rc = ensure_staging(c);
    if (rc) return rc;
    size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_x, v, bytes, hipMemcpyHostToDevice, c->stream));
    rc = launch_pressure_projection(c, c->d_x);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(v, c->d_x, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return IEMIC_OK;
}

int iemic_apply_mass(iemic_ctx *c, const double *v, double *out)
{
    CHECK_CTX(c);
    for (int r = 0; r < c->sc.ndim; r++) out[r] = c->h.diagB[r] * v[r];
    return IEMIC_OK;
}

int iemic_set_preconditioner(iemic_ctx *c, int kind)
{
    CHECK_CTX(c);
    if (kind != IEMIC_PREC_ROW_LU && kind != IEMIC_PREC_COLUMN_ILU)
        return set_error(IEMIC_EINVAL, "unknown preconditioner kind");
    HIPCHK(hipStreamSynchronize(c->stream));
    precond_free(c);
    c->prec_kind = kind;
    return IEMIC_OK;
}

int iemic_apply_tuning(iemic_ctx *c, int nrhs, int *blocks)
{
    CHECK_CTX(c);
    if (!blocks) return set_error(IEMIC_EINVAL, "null pointer");
    *blocks = c->prec_kind == IEMIC_PREC_ROW_LU ? rowlu_tuned_blocks(c, nrhs) : 0;
    return IEMIC_OK;
}

int iemic_spmm_tuning(iemic_ctx *c, int nrhs, int *columns)
{
    CHECK_CTX(c);
    if (!columns) return set_error(IEMIC_EINVAL, "null pointer");
    auto it = c->spmm_nc.find(nrhs);
    *columns = it == c->spmm_nc.end() ? 0 : it->second;
    return IEMIC_OK;
}

int iemic_set_precond_reuse(iemic_ctx *c, int max_age)
{
    CHECK_CTX(c);
    if (max_age < 0) return set_error(IEMIC_EINVAL, "max_age must be >= 0");
    c->prec_max_age = (unsigned)max_age;
    return IEMIC_OK;
}

int iemic_preconditioner_info(iemic_ctx *c, iemic_precond_info *info)
{
    CHECK_CTX(c);
    if (!info) return set_error(IEMIC_EINVAL, "null pointer");
    return precond_info(c, info);
}

int iemic_build_preconditioner(iemic_ctx *c)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    return precond_build(c, 0);
}

int iemic_apply_precon_dev(iemic_ctx *c, const double *d_v, double *d_out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    return precond_apply(c, 0, d_v, d_out, nullptr);
}

int iemic_apply_precon(iemic_ctx *c, const double *v, double *out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    int rc = ensure_staging(c);
    if (rc) return rc;
    size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_x, v, bytes, hipMemcpyHostToDevice, c->stream));
    rc = precond_apply(c, 0, c->d_x, c->d_y, nullptr);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, c->d_y, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return precond_check(c);
}

int iemic_apply_precon_fp32(iemic_ctx *c, const double *v, double *out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    int rc = ensure_staging(c);
    if (rc) return rc;
    float *tmp = nullptr;
    if (c->prec_kind != IEMIC_PREC_ROW_LU)
        HIPCHK(hipMalloc(&tmp, 2 * (size_t)c->sc.ndim * sizeof(float)));
    size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    HIPCHK(hipMemcpyAsync(c->d_x, v, bytes, hipMemcpyHostToDevice, c->stream));
    rc = precond_apply(c, 1, c->d_x, c->d_y, tmp);
    if (rc == IEMIC_OK) {
        HIPCHK(hipMemcpyAsync(out, c->d_y, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        rc = precond_check(c);
    }
    if (tmp) (void)hipFree(tmp);
    return rc;
}

void iemic_solver_defaults(iemic_solver_params *p)
{
    if (!p) return;
    p->tol = 1e-8;
    p->restart = 250;
    p->maxit = 1000;
    p->precision = 0;
    p->verbose = 0;
}

static int solve_dev(iemic_ctx *c, const iemic_solver_params *p, const double *d_b, double *d_x,
                     iemic_solve_info *info, bool transpose)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return krylov_solve(c, p ? p : &def, d_b, d_x, info, transpose);
}

static int solve_host(iemic_ctx *c, const iemic_solver_params *p, const double *b, double *x,
                      iemic_solve_info *info, bool transpose)
{
    CHECK_CTX(c);
    int rc = ensure_staging(c);
    if (rc) return rc;
    size_t bytes = (size_t)c->sc.ndim * sizeof(double);
    double *db = nullptr, *dx = nullptr;
    HIPCHK(hipMalloc(&db, bytes));
    HIPCHK(hipMalloc(&dx, bytes));
    rc = IEMIC_OK;
    if (hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = set_error(IEMIC_EHIP, "copy b");
    if (!rc) rc = solve_dev(c, p, db, dx, info, transpose);
    if (rc == IEMIC_OK || rc == IEMIC_ENOCONV) {
        if (hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = set_error(IEMIC_EHIP, "copy x");
    }
    (void)hipFree(db);
    (void)hipFree(dx);
    return rc;
}

int iemic_solve_dev(iemic_ctx *c, const iemic_solver_params *p, const double *d_b, double *d_x,
                    iemic_solve_info *info)
{
    return solve_dev(c, p, d_b, d_x, info, false);
}

int iemic_solve(iemic_ctx *c, const iemic_solver_params *p, const double *b, double *x, iemic_solve_info *info)
{
    return solve_host(c, p, b, x, info, false);
}

int iemic_solve_transpose_dev(iemic_ctx *c, const iemic_solver_params *p, const double *d_b, double *d_x,
                              iemic_solve_info *info)
{
    return solve_dev(c, p, d_b, d_x, info, true);
}

int iemic_solve_transpose(iemic_ctx *c, const iemic_solver_params *p, const double *b, double *x,
                          iemic_solve_info *info)
{
    return solve_host(c, p, b, x, info, true);
}

int iemic_apply_precon_transpose_dev(iemic_ctx *c, const double *d_v, double *d_out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    return precond_apply_t(c, 0, d_v, d_out);
}

int iemic_solve_multi_dev(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *d_B, double *d_X,
                          iemic_solve_info *infos)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    if (nrhs < 1 || !d_B || !d_X) return set_error(IEMIC_EINVAL, "bad right-hand side block");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return krylov_solve_multi(c, p ? p : &def, nrhs, d_B, d_X, infos, false);
}

int iemic_solve_multi_transpose_dev(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *d_B,
                                    double *d_X, iemic_solve_info *infos)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    if (nrhs < 1 || !d_B || !d_X) return set_error(IEMIC_EINVAL, "bad right-hand side block");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return krylov_solve_multi(c, p ? p : &def, nrhs, d_B, d_X, infos, true);
}

int iemic_solve_block_dev(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *d_B, double *d_X,
                          iemic_solve_info *infos)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    if (nrhs < 1 || nrhs > 32 || !d_B || !d_X) return set_error(IEMIC_EINVAL, "bad right-hand side block (1 to 32 columns)");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return krylov_solve_block(c, p ? p : &def, nrhs, d_B, d_X, infos, false);
}

int iemic_solve_block_transpose_dev(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *d_B,
                                    double *d_X, iemic_solve_info *infos)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    if (nrhs < 1 || nrhs > 32 || !d_B || !d_X) return set_error(IEMIC_EINVAL, "bad right-hand side block (1 to 32 columns)");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return krylov_solve_block(c, p ? p : &def, nrhs, d_B, d_X, infos, true);
}

// host copies around the device solves; block selects block FGMRES
static int solve_multi_host(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                            iemic_solve_info *infos, bool transpose, bool block = false)
{
    CHECK_CTX(c);
    if (nrhs < 1 || !B || !X) return set_error(IEMIC_EINVAL, "bad right-hand side block");
    const size_t bytes = (size_t)nrhs * c->sc.ndim * sizeof(double);
    double *dB = nullptr, *dX = nullptr;
    HIPCHK(hipMalloc(&dB, bytes));
    if (hipMalloc(&dX, bytes) != hipSuccess) {
        (void)hipFree(dB);
        return set_error(IEMIC_EHIP, "allocation of the solution block");
    }
    int rc = IEMIC_OK;
    if (hipMemcpy(dB, B, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = set_error(IEMIC_EHIP, "copy B");
    if (!rc) {
        if (block)
            rc = transpose ? iemic_solve_block_transpose_dev(c, p, nrhs, dB, dX, infos)
                           : iemic_solve_block_dev(c, p, nrhs, dB, dX, infos);
        else
            rc = transpose ? iemic_solve_multi_transpose_dev(c, p, nrhs, dB, dX, infos)
                           : iemic_solve_multi_dev(c, p, nrhs, dB, dX, infos);
    }
    if (rc == IEMIC_OK || rc == IEMIC_ENOCONV) {
        const int keep = rc;
        if (hipMemcpy(X, dX, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = set_error(IEMIC_EHIP, "copy X");
        else rc = keep;
    }
    (void)hipFree(dB);
    (void)hipFree(dX);
    return rc;
}

int iemic_solve_multi(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                      iemic_solve_info *infos)
{
    return solve_multi_host(c, p, nrhs, B, X, infos, false);
}

int iemic_solve_multi_transpose(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                                iemic_solve_info *infos)
{
    return solve_multi_host(c, p, nrhs, B, X, infos, true);
}

int iemic_solve_block(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                      iemic_solve_info *infos)
{
    return solve_multi_host(c, p, nrhs, B, X, infos, false, true);
}

int iemic_solve_block_transpose(iemic_ctx *c, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                                iemic_solve_info *infos)
{
    return solve_multi_host(c, p, nrhs, B, X, infos, true, true);
}

int iemic_apply_precon_multi_dev(iemic_ctx *c, int nrhs, const double *d_in, double *d_out)
{
    CHECK_CTX(c);
    if (!c->have_jac) return set_error(IEMIC_ESTATE, "no Jacobian computed yet");
    if (nrhs < 1 || !d_in || !d_out) return set_error(IEMIC_EINVAL, "bad vector block");
    const size_t n = c->sc.ndim;
    return precond_apply_multi(c, 0, nrhs, d_in, n, d_out, n, nullptr);
}

int iemic_newton_dev(iemic_ctx *c, const iemic_solver_params *p, double tol, double step_tol, int maxit, double *d_x,
                     int *iterations, double *fnorm)
{
    CHECK_CTX(c);
    if (!d_x || maxit < 1 || !(tol >= 0.0) || !(step_tol >= 0.0)) return set_error(IEMIC_EINVAL, "bad Newton arguments");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return newton_dev(c, p ? p : &def, tol, step_tol, maxit, d_x, iterations, fnorm, nullptr);
}

void iemic_continuation_defaults(iemic_continuation_params *p)
{
    if (!p) return;
    p->par = 19;
    p->target = 1.0;
    p->ds = 1e-2;
    p->ds_min = 1e-8;
    p->ds_max = 1e3;
    p->max_steps = 100;
    p->newton_maxit = 7;
    p->newton_tol = 1e-4;
    p->eps = 1e-5;
    p->tan_scaling = 1.0;
    p->opt_newton = 3.5;
    p->predictor_bound = 1e3;
}

int iemic_continuation_dev(iemic_ctx *c, const iemic_solver_params *p, const iemic_continuation_params *cp,
                           double *d_x, double *par_hist, double *norm_hist, iemic_continuation_info *info)
{
    CHECK_CTX(c);
    if (!d_x || !cp) return set_error(IEMIC_EINVAL, "null argument");
    if (cp->par == 6 || cp->par == 29)   // MIXP, MKAP: as iemic_set_par
        return set_error(IEMIC_EUNSUPPORTED, "neutral physics (MIXP/MKAP) is not supported");
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return continuation_dev(c, p ? p : &def, cp, d_x, par_hist, norm_hist, info);
}

int iemic_diagnostic(const iemic_ctx *c, int which, const double *x, double *value, double *grad)
{
    CHECK_CTX(c);
    if (!x || !value) return set_error(IEMIC_EINVAL, "null pointer");
    return diagnostic(c, which, x, value, grad);
}

// One Newton step of one model: F(x), J(x), the preconditioner for J, and
// J dx = F (x is not changed; x_new = x - dx).
static int newton_step_one(iemic_ctx *c, const iemic_solver_params *p, const double *d_x, double *d_dx,
                           iemic_solve_info *info)
{
    int rc = ensure_staging(c);
    if (rc) return rc;
    if ((rc = launch_rhs(c, d_x, c->d_y))) return rc;
    if ((rc = launch_jacobian(c, d_x))) return rc;
    iemic_solver_params def;
    iemic_solver_defaults(&def);
    return krylov_solve(c, p ? p : &def, c->d_y, d_dx, info, false);
}

int iemic_newton_step_many(iemic_ctx *const *ctx, int count, const double *const *d_x, double *const *d_dx,
                           const iemic_solver_params *p, iemic_solve_info *infos, int concurrent)
{
    if (!ctx || !d_x || !d_dx || count < 1) return set_error(IEMIC_EINVAL, "bad model list");
    for (int q = 0; q < count; q++) {
        if (!ctx[q] || !d_x[q] || !d_dx[q]) return set_error(IEMIC_EINVAL, "null model or vector");
        for (int r = 0; r < q; r++)
            if (ctx[r] == ctx[q]) return set_error(IEMIC_EINVAL, "a context appears twice in the model list");
    }
    std::vector<int> rcs(count, IEMIC_OK);
    std::vector<std::string> errs(count);
    std::vector<iemic_solve_info> inf(count);
    auto run = [&](int q) {
        rcs[q] = newton_step_one(ctx[q], p, d_x[q], d_dx[q], &inf[q]);
        if (rcs[q]) errs[q] = iemic_last_error();   // the message is per thread
    };
    if (concurrent) {
        std::vector<std::thread> th;
        th.reserve(count);
        for (int q = 0; q < count; q++) th.emplace_back(run, q);
        for (auto &t : th) t.join();
    } else {
        for (int q = 0; q < count; q++) run(q);
    }
    if (infos)
        for (int q = 0; q < count; q++) infos[q] = inf[q];
    int first = IEMIC_OK;
    for (int q = 0; q < count && first == IEMIC_OK; q++)
        if (rcs[q] && rcs[q] != IEMIC_ENOCONV) first = set_error(rcs[q], "model " + std::to_string(q) + ": " + errs[q]);
    if (first) return first;
    for (int q = 0; q < count; q++)
        if (rcs[q] == IEMIC_ENOCONV) return set_error(IEMIC_ENOCONV, "model " + std::to_string(q) + ": " + errs[q]);
    return IEMIC_OK;
}

int iemic_read_mask(const char *path, int n, int m, int l, int *landm)
{
    return read_mask_text(path, n, m, l, landm);
}

int iemic_dev_alloc(void **ptr, long long bytes)
{
    if (!ptr || bytes < 0) return set_error(IEMIC_EINVAL, "bad allocation request");
    hipError_t e = hipMalloc(ptr, (size_t)bytes);
    if (e != hipSuccess) return set_error(IEMIC_ENOMEM, hipGetErrorString(e));
    return IEMIC_OK;
}
int iemic_dev_free(void *ptr)
{
    if (ptr) HIPCHK(hipFree(ptr));
    return IEMIC_OK;
}
int iemic_memcpy_h2d(void *dst, const void *src, long long bytes)
{
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    return IEMIC_OK;
}
int iemic_memcpy_d2h(void *dst, const void *src, long long bytes)
{
    HIPCHK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    return IEMIC_OK;
}
int iemic_synchronize(iemic_ctx *c)
{
    CHECK_CTX(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    return IEMIC_OK;
}

int iemic_last_kernel_ms(const iemic_ctx *c, int kernel, double *ms)
{
    CHECK_CTX(c);
    if (kernel < 0 || kernel >= 8 || !ms) return set_error(IEMIC_EINVAL, "bad kernel id");
    if (!c->ev[kernel][0]) return set_error(IEMIC_ESTATE, "kernel not launched yet");
    float t = 0.f;
    HIPCHK(hipEventSynchronize(c->ev[kernel][1]));
    HIPCHK(hipEventElapsedTime(&t, c->ev[kernel][0], c->ev[kernel][1]));
    *ms = t;
    return IEMIC_OK;
}

}  // extern "C"
