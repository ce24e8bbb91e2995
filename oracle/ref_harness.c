/*
 * Test infrastructure only: a one-rank driver for the reference THCM Fortran
 * (the .F90 files of src/ocean and mix_imp.f), compiled from the reference sources by
 * oracle/Makefile into oracle/_ref/libthcm_ref.so.  It is used by the tests to
 * pin the C restatement in thcm_oracle.c; nothing in the product links it.
 *
 * The reference drives its Fortran from C++ (THCM.C) through Trilinos, which
 * is not available here.  This file restates the parts of that glue that a
 * single MPI rank executes, without any Trilinos type:
 *   - THCM::THCM (THCM.C:180-795): m_global::initialize, the global grid
 *     setters, init_, the CSR buffers (get_array_sizes/set_pointers), the
 *     integral-condition row and coefficients, the maximal graph, evaluateB;
 *   - THCM::evaluate (THCM.C:957-1199): rhs_ + negation + integral row, and
 *     matrix_ + per-row replacement into the fixed graph + intcond_S;
 *   - THCM::CreateMaximalGraph / insert_graph_entry (THCM.C:2300-2598);
 *   - the Fortran callbacks thcm_forcing_integral_ (THCM.C:2653-2686),
 *     thcm_throw_error_ (THCM.C:2690) and the TIMER_START/STOP hooks
 *     (GlobalDefinitions.C), which are no-ops here.
 * Global row index: 6*((k*M + j)*N + i) + XX-1, XX=1..6 for u,v,w,p,T,S
 * (THCMdefs.H FIND_ROW2).  Graph rows are sorted by column and duplicate-free,
 * as Epetra_CrsGraph::FillComplete leaves them on one rank.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- reference Fortran entry points (flang mangling for module procs) ---- */
extern void _QMm_globalPinitialize(int *n, int *m, int *l, double *xmin, double *xmax,
                                   double *ymin, double *ymax, double *hdim, double *qz,
                                   int *periodic, int *itopo, int *flat, int *rd_mask,
                                   int *tres, int *sres, int *iza, int *ite, int *its,
                                   int *rd_spertm, int *coupled_t, int *coupled_s,
                                   int *forcing_type, const char *maskfile,
                                   const char *spertmaskfile, const char *windfile,
                                   const char *sstfile, const char *sssfile);
extern void _QMm_globalPget_landm(int *cland);
extern void set_global_x(int *n, double *p);
extern void set_global_y(int *n, double *p);
extern void set_global_z(int *n, double *p);
extern void set_global_xu(int *n, double *p);
extern void set_global_yv(int *n, double *p);
extern void set_global_zw(int *n, double *p);
extern void init_(int *n, int *m, int *l, int *nmlglob, double *xmin, double *xmax,
                  double *ymin, double *ymax, double *alphaT, double *alphaS, int *ih,
                  int *vmix, int *tap, int *rho_mixing, int *coriolis_on, int *periodic,
                  int *landm, double *taux, double *tauy, double *tatm, double *emip,
                  double *spert);
extern void finalize_(void);
extern void setparcs_(int *idx, double *val);
extern void getparcs_(int *idx, double *val);
extern void rhs_(double *un, double *b);
extern void matrix_(double *un);
extern void fillcolb_(void);
extern void _QMm_matPget_array_sizes(int *nrows, int *nnz);
extern void _QMm_matPset_pointers(int *nrows, int *nnz, int *beg, int *jco, double *co,
                                  double *cob, int *begf, int *jcof, double *cof);
extern void _QMm_thcm_utilsPintcond_scaling(double *val, int *ind, int *len);
extern void _QMm_mixPset_vmix_fix(int *vmix_fix);
extern int _QMm_mixEvmix_dim;   /* m_mix::vmix_dim, mix.F90 */

/* ---- callbacks the Fortran needs ---------------------------------------- */
static int g_n, g_m, g_l, g_periodic;

void timer_start_(const char *s) { (void)s; }
void timer_stop_(const char *s) { (void)s; }
void thcm_throw_error_(const char *msg)
{
    fprintf(stderr, "thcm_ref: fortran error: %.80s\n", msg);
    abort();
}

/* THCM.C:2653-2686 on one rank: area-weighted surface mean of a 2-D field. */
void thcm_forcing_integral_(double *qfun2, double *y, int *landm, double *fsint)
{
    int n = g_n, m = g_m, l = g_l;
    double lsint = 0.0, lfsint = 0.0;
    for (int j = 0; j < m; j++)
        for (int i = 0; i < n; i++) {
            int pl = (i + 1) + (n + 2) * ((j + 1) + (m + 2) * l);
            int pq = i + n * j;
            lfsint = qfun2[pq] * cos(y[j]) * (1 - landm[pl]) + lfsint;
            lsint = cos(y[j]) * (1 - landm[pl]) + lsint;
        }
    *fsint = lfsint / lsint;
}

/* ---- configuration ------------------------------------------------------ */
typedef struct {
    int n, m, l;
    double xmin_deg, xmax_deg, ymin_deg, ymax_deg;
    int periodic;
    double hdim, qz;
    double alphaT, alphaS;
    int ih, vmix, tap, rho_mixing, coriolis_on;
    int tres, sres, iza, ite, its, coupled_t, coupled_s, forcing_type;
    int int_sign, nic, mic;
} thcmref_config;

static thcmref_config g_cfg;
static int g_ndim, g_nnzmax, g_inited;
static int *g_begA, *g_jcoA, *g_begF, *g_jcoF;
static double *g_coA, *g_coB, *g_coF;
static double *g_coef;      /* intcondCoeff_, length ndim */
static int g_rowintcon = -1;
static double *g_gx, *g_gy, *g_gz, *g_gxu, *g_gyv, *g_gzw;
/* maximal graph (CreateMaximalGraph), 0-based CSR */
static int *g_grow, *g_gcol, g_gnnz;

#define ROW(i, j, k, xx) (6 * (((k) * g_m + (j)) * g_n + (i)) + (xx) - 1)

static double fz(double z, double qz)
{
    double th = tanh(qz * (z + 1)), tth = tanh(qz);
    if (qz > 1.0) return -1 + th / tth;
    return z + (1. - qz) * z * (1 - z);
}

/* TRIOS_Domain.C:34-75: the global grid handed to m_global. */
static void make_global_grid(void)
{
    const double PI_ = 3.14159265358979323846;
    int n = g_cfg.n, m = g_cfg.m, l = g_cfg.l;
    double xmin = g_cfg.xmin_deg * PI_ / 180.0, xmax = g_cfg.xmax_deg * PI_ / 180.0;
    double ymin = g_cfg.ymin_deg * PI_ / 180.0, ymax = g_cfg.ymax_deg * PI_ / 180.0;
    double zmin = -1.0, zmax = 0.0;
    double dx = (xmax - xmin) / n, dy = (ymax - ymin) / m, dz = (zmax - zmin) / l;
    free(g_gx); free(g_gy); free(g_gz); free(g_gxu); free(g_gyv); free(g_gzw);
    g_gx = malloc(sizeof(double) * n); g_gy = malloc(sizeof(double) * m);
    g_gz = malloc(sizeof(double) * l); g_gxu = malloc(sizeof(double) * (n + 1));
    g_gyv = malloc(sizeof(double) * (m + 1)); g_gzw = malloc(sizeof(double) * (l + 1));
    g_gxu[0] = xmin; g_gyv[0] = ymin; g_gzw[0] = zmin;
    for (int i = 0; i < n; i++) { g_gx[i] = xmin + (i + 0.5) * dx; g_gxu[i + 1] = xmin + (i + 1.0) * dx; }
    for (int j = 0; j < m; j++) { g_gy[j] = ymin + (j + 0.5) * dy; g_gyv[j + 1] = ymin + (j + 1.0) * dy; }
    for (int k = 0; k < l; k++) {
        g_gz[k] = fz(zmin + (k + 0.5) * dz, g_cfg.qz);
        g_gzw[k + 1] = fz(zmin + (k + 1.0) * dz, g_cfg.qz);
    }
}

static void global_initialize(int itopo, int rd_mask, const char *maskfile)
{
    thcmref_config *c = &g_cfg;
    double xmin = c->xmin_deg * 3.14159265358979323846 / 180.0;
    double xmax = c->xmax_deg * 3.14159265358979323846 / 180.0;
    double ymin = c->ymin_deg * 3.14159265358979323846 / 180.0;
    double ymax = c->ymax_deg * 3.14159265358979323846 / 180.0;
    int flat = 0, rd_spertm = 0;
    char mf[512];
    snprintf(mf, sizeof mf, "%s", maskfile ? maskfile : "no_mask_specified");
    _QMm_globalPinitialize(&c->n, &c->m, &c->l, &xmin, &xmax, &ymin, &ymax, &c->hdim, &c->qz,
                           &c->periodic, &itopo, &flat, &rd_mask, &c->tres, &c->sres, &c->iza,
                           &c->ite, &c->its, &rd_spertm, &c->coupled_t, &c->coupled_s,
                           &c->forcing_type, mf, "no_mask_specified", "wind/trtau.dat",
                           "levitus/new/t00an1", "levitus/new/s00an1");
    make_global_grid();
    int sz;
    sz = c->n; set_global_x(&sz, g_gx);
    sz = c->m; set_global_y(&sz, g_gy);
    sz = c->l; set_global_z(&sz, g_gz);
    sz = c->n + 1; set_global_xu(&sz, g_gxu);
    sz = c->m + 1; set_global_yv(&sz, g_gyv);
    sz = c->l + 1; set_global_zw(&sz, g_gzw);
}

/* THCM::insert_graph_entry (THCM.C:2583-2598) */
static void graph_entry(int *idx, int *pos, int i, int j, int k, int xx)
{
    int N = g_n, M = g_m, L = g_l, ii = i;
    if (g_periodic) ii = ((i % N) + N) % N;
    if (ii >= 0 && j >= 0 && k >= 0 && ii < N && j < M && k < L)
        idx[(*pos)++] = ROW(ii, j, k, xx);
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int sort_unique(int *v, int len)
{
    qsort(v, len, sizeof(int), cmp_int);
    int w = 0;
    for (int r = 0; r < len; r++)
        if (w == 0 || v[w - 1] != v[r]) v[w++] = v[r];
    return w;
}

/* THCM::CreateMaximalGraph (THCM.C:2300-2580), one rank. */
static void build_graph(void)
{
    int N = g_n, M = g_m, L = g_l, ndim = g_ndim;
    int cap = ndim * 24 + N * M * L + 16;
    free(g_grow); free(g_gcol);
    g_grow = calloc(ndim + 1, sizeof(int));
    g_gcol = malloc(sizeof(int) * cap);
    int nz = 0, idx[24], pos;
    for (int k = 0; k < L; k++)
        for (int j = 0; j < M; j++)
            for (int i = 0; i < N; i++) {
                int gid0 = ROW(i, j, k, 1) - 1;
                /* U */
                pos = 0;
                graph_entry(idx, &pos, i, j, k, 1); graph_entry(idx, &pos, i - 1, j, k, 1);
                graph_entry(idx, &pos, i + 1, j, k, 1); graph_entry(idx, &pos, i, j - 1, k, 1);
                graph_entry(idx, &pos, i, j + 1, k, 1); graph_entry(idx, &pos, i, j, k - 1, 1);
                graph_entry(idx, &pos, i, j, k + 1, 1);
                graph_entry(idx, &pos, i, j, k, 2); graph_entry(idx, &pos, i - 1, j, k, 2);
                graph_entry(idx, &pos, i + 1, j, k, 2); graph_entry(idx, &pos, i, j - 1, k, 2);
                graph_entry(idx, &pos, i, j + 1, k, 2);
                graph_entry(idx, &pos, i, j, k, 3); graph_entry(idx, &pos, i + 1, j, k, 3);
                graph_entry(idx, &pos, i + 1, j + 1, k, 3); graph_entry(idx, &pos, i, j + 1, k, 3);
                graph_entry(idx, &pos, i, j, k - 1, 3); graph_entry(idx, &pos, i + 1, j, k - 1, 3);
                graph_entry(idx, &pos, i + 1, j + 1, k - 1, 3); graph_entry(idx, &pos, i, j + 1, k - 1, 3);
                graph_entry(idx, &pos, i, j, k, 4); graph_entry(idx, &pos, i + 1, j, k, 4);
                graph_entry(idx, &pos, i, j + 1, k, 4); graph_entry(idx, &pos, i + 1, j + 1, k, 4);
                pos = sort_unique(idx, pos);
                g_grow[gid0 + 2] = pos; memcpy(g_gcol + nz, idx, pos * sizeof(int)); nz += pos;
                /* V */
                pos = 0;
                graph_entry(idx, &pos, i, j, k, 2); graph_entry(idx, &pos, i - 1, j, k, 2);
                graph_entry(idx, &pos, i + 1, j, k, 2); graph_entry(idx, &pos, i, j - 1, k, 2);
                graph_entry(idx, &pos, i, j + 1, k, 2); graph_entry(idx, &pos, i, j, k - 1, 2);
                graph_entry(idx, &pos, i, j, k + 1, 2);
                graph_entry(idx, &pos, i, j, k, 1); graph_entry(idx, &pos, i - 1, j, k, 1);
                graph_entry(idx, &pos, i + 1, j, k, 1);
                graph_entry(idx, &pos, i, j, k, 3); graph_entry(idx, &pos, i + 1, j, k, 3);
                graph_entry(idx, &pos, i + 1, j + 1, k, 3); graph_entry(idx, &pos, i, j + 1, k, 3);
                graph_entry(idx, &pos, i, j, k - 1, 3); graph_entry(idx, &pos, i + 1, j, k - 1, 3);
                graph_entry(idx, &pos, i + 1, j + 1, k - 1, 3); graph_entry(idx, &pos, i, j + 1, k - 1, 3);
                graph_entry(idx, &pos, i, j, k, 4); graph_entry(idx, &pos, i + 1, j, k, 4);
                graph_entry(idx, &pos, i, j + 1, k, 4); graph_entry(idx, &pos, i + 1, j + 1, k, 4);
                pos = sort_unique(idx, pos);
                g_grow[gid0 + 3] = pos; memcpy(g_gcol + nz, idx, pos * sizeof(int)); nz += pos;
                /* W */
                pos = 0;
                graph_entry(idx, &pos, i, j, k, 3);
                graph_entry(idx, &pos, i, j, k, 4); graph_entry(idx, &pos, i, j, k + 1, 4);
                graph_entry(idx, &pos, i, j, k, 5); graph_entry(idx, &pos, i, j, k + 1, 5);
                graph_entry(idx, &pos, i, j, k, 6); graph_entry(idx, &pos, i, j, k + 1, 6);
                pos = sort_unique(idx, pos);
                g_grow[gid0 + 4] = pos; memcpy(g_gcol + nz, idx, pos * sizeof(int)); nz += pos;
                /* P */
                pos = 0;
                graph_entry(idx, &pos, i, j, k, 4);
                graph_entry(idx, &pos, i, j, k, 1); graph_entry(idx, &pos, i - 1, j, k, 1);
                graph_entry(idx, &pos, i, j - 1, k, 1); graph_entry(idx, &pos, i - 1, j - 1, k, 1);
                graph_entry(idx, &pos, i, j, k, 2); graph_entry(idx, &pos, i - 1, j, k, 2);
                graph_entry(idx, &pos, i, j - 1, k, 2); graph_entry(idx, &pos, i - 1, j - 1, k, 2);
                graph_entry(idx, &pos, i, j, k, 3); graph_entry(idx, &pos, i, j, k - 1, 3);
                pos = sort_unique(idx, pos);
                g_grow[gid0 + 5] = pos; memcpy(g_gcol + nz, idx, pos * sizeof(int)); nz += pos;
                /* T and S */
                for (int xx = 5; xx <= 6; xx++) {
                    int yy = (xx == 5) ? 6 : 5;
                    pos = 0;
                    graph_entry(idx, &pos, i, j, k, xx); graph_entry(idx, &pos, i - 1, j, k, xx);
                    graph_entry(idx, &pos, i + 1, j, k, xx); graph_entry(idx, &pos, i, j - 1, k, xx);
                    graph_entry(idx, &pos, i, j + 1, k, xx); graph_entry(idx, &pos, i, j, k - 1, xx);
                    graph_entry(idx, &pos, i, j, k + 1, xx);
                    graph_entry(idx, &pos, i, j, k, 1); graph_entry(idx, &pos, i - 1, j, k, 1);
                    graph_entry(idx, &pos, i - 1, j - 1, k, 1); graph_entry(idx, &pos, i, j - 1, k, 1);
                    graph_entry(idx, &pos, i, j, k, 2); graph_entry(idx, &pos, i - 1, j, k, 2);
                    graph_entry(idx, &pos, i - 1, j - 1, k, 2); graph_entry(idx, &pos, i, j - 1, k, 2);
                    graph_entry(idx, &pos, i, j, k, 3); graph_entry(idx, &pos, i, j, k - 1, 3);
                    graph_entry(idx, &pos, i, j, k, yy); graph_entry(idx, &pos, i, j, k - 1, yy);
                    graph_entry(idx, &pos, i, j, k + 1, yy);
                    int row = gid0 + xx;
                    if (xx == 6 && g_cfg.sres == 0 && row == g_rowintcon) {
                        /* dense integral row over every S column (THCM.C:2553-2570) */
                        int p2 = 0;
                        for (int ii = 0; ii < N; ii++)
                            for (int jj = 0; jj < M; jj++)
                                for (int kk = 0; kk < L; kk++) g_gcol[nz + p2++] = ROW(ii, jj, kk, 6);
                        p2 = sort_unique(g_gcol + nz, p2);
                        g_grow[row + 1] = p2; nz += p2;
                        continue;
                    }
                    pos = sort_unique(idx, pos);
                    g_grow[row + 1] = pos; memcpy(g_gcol + nz, idx, pos * sizeof(int)); nz += pos;
                }
            }
    for (int r = 0; r < ndim; r++) g_grow[r + 1] += g_grow[r];
    g_gnnz = nz;
}

/* ---- public entry points (called through ctypes by the tests) ---------- */

/* m_global::get_landm path of THCM::THCM (THCM.C:420-440): the reference's
 * own topography (itopo) or mask-file reader.  out: (n+2)(m+2)(l+2) ints. */
int thcmref_reference_landm(const thcmref_config *cfg, int itopo, int rd_mask,
                            const char *maskfile, int *out)
{
    g_cfg = *cfg;
    global_initialize(itopo, rd_mask, maskfile);
    _QMm_globalPget_landm(out);
    return 0;
}

int thcmref_init(const thcmref_config *cfg, const int *landm_in, const double *taux,
                 const double *tauy, const double *tatm, const double *emip,
                 const double *spert)
{
    if (g_inited) finalize_();
    g_cfg = *cfg;
    g_n = cfg->n; g_m = cfg->m; g_l = cfg->l; g_periodic = cfg->periodic;
    global_initialize(1, 0, NULL);

    int n = cfg->n, m = cfg->m, l = cfg->l;
    int nml = n * m * l, nland = (n + 2) * (m + 2) * (l + 2);
    int *landm = malloc(sizeof(int) * nland);
    memcpy(landm, landm_in, sizeof(int) * nland);
    double *f[5];
    const double *src[5] = {taux, tauy, tatm, emip, spert};
    for (int q = 0; q < 5; q++) {
        f[q] = calloc(n * m, sizeof(double));
        if (src[q]) memcpy(f[q], src[q], sizeof(double) * n * m);
    }
    double xmin = cfg->xmin_deg * 3.14159265358979323846 / 180.0;
    double xmax = cfg->xmax_deg * 3.14159265358979323846 / 180.0;
    double ymin = cfg->ymin_deg * 3.14159265358979323846 / 180.0;
    double ymax = cfg->ymax_deg * 3.14159265358979323846 / 180.0;
    double aT = cfg->alphaT, aS = cfg->alphaS;
    int ih = cfg->ih, vmix = cfg->vmix, tap = cfg->tap, rho = cfg->rho_mixing,
        cor = cfg->coriolis_on, per = cfg->periodic;
    /* the mixing pattern of a fresh process (THCM is one per process in the
     * reference; m_mix keeps vmix_dim across init_ calls of this harness) */
    _QMm_mixEvmix_dim = 0;
    init_(&n, &m, &l, &nml, &xmin, &xmax, &ymin, &ymax, &aT, &aS, &ih, &vmix, &tap, &rho,
          &cor, &per, landm, f[0], f[1], f[2], f[3], f[4]);
    for (int q = 0; q < 5; q++) free(f[q]);
    free(landm);

    int nrows, nnz;
    _QMm_matPget_array_sizes(&nrows, &nnz);
    g_ndim = nrows; g_nnzmax = nnz;
    free(g_begA); free(g_jcoA); free(g_coA); free(g_coB); free(g_begF); free(g_jcoF); free(g_coF);
    g_begA = calloc(nrows + 1, sizeof(int)); g_jcoA = calloc(nnz, sizeof(int));
    g_coA = calloc(nnz, sizeof(double)); g_coB = calloc(nrows, sizeof(double));
    g_begF = calloc(nrows + 1, sizeof(int)); g_jcoF = calloc(nrows, sizeof(int));
    g_coF = calloc(nrows, sizeof(double));
    _QMm_matPset_pointers(&nrows, &nnz, g_begA, g_jcoA, g_coA, g_coB, g_begF, g_jcoF, g_coF);

    /* integral condition row (THCM.C:663-698) */
    g_rowintcon = -1;
    if (cfg->sres == 0) {
        int nic = cfg->nic < 0 ? n - 1 : cfg->nic;
        int mic = cfg->mic < 0 ? m - 1 : cfg->mic;
        g_rowintcon = ROW(nic, mic, l - 1, 6);
    }
    /* intcondCoeff_ (THCM::getIntCondCoeff, THCM.C:2608-2636) */
    free(g_coef);
    g_coef = calloc(nrows, sizeof(double));
    double *val = malloc(sizeof(double) * nml);
    int *ind = malloc(sizeof(int) * nml), len = 0;
    _QMm_thcm_utilsPintcond_scaling(val, ind, &len);
    for (int q = 0; q < len; q++) g_coef[ind[q] - 1] = val[q];
    free(val); free(ind);
    build_graph();
    g_inited = 1;
    return 0;
}

int thcmref_set_par(int idx, double v) { setparcs_(&idx, &v); return 0; }
double thcmref_get_par(int idx) { double v = 0; getparcs_(&idx, &v); return v; }
int thcmref_ndim(void) { return g_ndim; }
int thcmref_rowintcon(void) { return g_rowintcon; }
int thcmref_graph_nnz(void) { return g_gnnz; }
void thcmref_intcond_coef(double *out) { memcpy(out, g_coef, sizeof(double) * g_ndim); }

/* THCM::evaluate, RHS branch (THCM.C:957-1042).  intcorr = intCorrection_. */
int thcmref_rhs(const double *x, double intcorr, double *F)
{
    double *xx = malloc(sizeof(double) * g_ndim);
    memcpy(xx, x, sizeof(double) * g_ndim);
    /* Ocean::computeRHS calls THCM::fixMixing(0) first (Ocean.C:1281,
     * THCM.C:2640-2647: only under Mixing = 2) */
    if (g_cfg.vmix == 2) { int z = 0; _QMm_mixPset_vmix_fix(&z); }
    rhs_(xx, F);
    for (int r = 0; r < g_ndim; r++) F[r] = -F[r];
    if (g_cfg.sres == 0) {
        double dot = 0.0;
        for (int r = 0; r < g_ndim; r++) dot += g_coef[r] * x[r];
        F[g_rowintcon] = g_cfg.int_sign * (dot - intcorr);
    }
    free(xx);
    return 0;
}

/* OceanGrid::recomputePsiM on one rank (OceanGrid.C:264-345): the state
 * unpacked onto the grid by the reference's usol1D (thcm_utils.F90:190-254,
 * usol), v summed over i = 1..n times dx = (xmax - xmin) / n for
 * j = 0..m, k = 1..l, then the reference's compute_psim
 * (thcm_utils.F90:95-118).  psim receives psiM(0:m, 0:l), j fastest; the
 * return value is the maximum over it, started at psiM(0,0) as the
 * reference's scan is. */
extern void _QMm_thcm_utilsPusol1d(double *un, double *u, double *v, double *w, double *p, double *t,
                                   double *s);
extern void _QMm_thcm_utilsPcompute_psim(double *vs, double *psim);
double thcmref_psim(const double *x, double *psim)
{
    const double PI_ = 3.14159265358979323846;
    int n = g_cfg.n, m = g_cfg.m, l = g_cfg.l;
    size_t nuv = (size_t)(n + 1) * (m + 1) * (l + 2), nw = (size_t)(n + 2) * (m + 2) * (l + 1);
    size_t np = (size_t)(n + 2) * (m + 2) * (l + 2);
    double *xx = malloc(sizeof(double) * g_ndim);
    double *u = calloc(nuv, sizeof(double)), *v = calloc(nuv, sizeof(double));
    double *w = calloc(nw, sizeof(double)), *p = calloc(np, sizeof(double));
    double *t = calloc(np, sizeof(double)), *sa = calloc(np, sizeof(double));
    double *vs = calloc((size_t)(m + 1) * l, sizeof(double));
    memcpy(xx, x, sizeof(double) * g_ndim);
    _QMm_thcm_utilsPusol1d(xx, u, v, w, p, t, sa);
    double xmin = g_cfg.xmin_deg * PI_ / 180.0, xmax = g_cfg.xmax_deg * PI_ / 180.0;
    double dx = (xmax - xmin) / n;
    for (int k = 1; k <= l; k++)
        for (int j = 0; j <= m; j++) {
            double sum = 0.0;
            for (int i = 1; i <= n; i++) sum += v[((size_t)(m + 1) * k + j) * (n + 1) + i];
            vs[(size_t)(m + 1) * (k - 1) + j] = sum * dx;
        }
    _QMm_thcm_utilsPcompute_psim(vs, psim);
    double mx = psim[0];
    for (int j = 0; j <= m; j++)
        for (int k = 0; k <= l; k++)
            if (psim[(size_t)(m + 1) * k + j] > mx) mx = psim[(size_t)(m + 1) * k + j];
    free(xx); free(u); free(v); free(w); free(p); free(t); free(sa); free(vs);
    return mx;
}

/* rhs_ itself (usrc.F90:523-603): the Fortran right-hand side B, not negated
 * and without the integral row, as the C ABI of THCM.C:49-176 receives it. */
int thcmref_fortran_rhs(const double *x, double *B)
{
    double *xx = malloc(sizeof(double) * g_ndim);
    memcpy(xx, x, sizeof(double) * g_ndim);
    if (g_cfg.vmix == 2) { int z = 0; _QMm_mixPset_vmix_fix(&z); }
    rhs_(xx, B);
    free(xx);
    return 0;
}

/* coB of the last matrix_ call (fillcolB, assemble.F90:18-54), as written
 * into the caller's buffer by set_pointers. */
void thcmref_fortran_coB(double *out) { memcpy(out, g_coB, sizeof(double) * g_ndim); }

/* The Fortran CSR of the last rhs_/matrix_ call (1-based), for inspection. */
int thcmref_fortran_csr_nnz(void) { return g_begA[g_ndim] - 1; }
void thcmref_fortran_csr(int *beg, int *jco, double *co)
{
    int nz = g_begA[g_ndim] - 1;
    memcpy(beg, g_begA, sizeof(int) * (g_ndim + 1));
    memcpy(jco, g_jcoA, sizeof(int) * nz);
    memcpy(co, g_coA, sizeof(double) * nz);
}

/* THCM::evaluate, Jacobian branch (THCM.C:1043-1180) into the fixed graph.
 * Returns the number of Fortran entries that fell outside the graph. */
int thcmref_jacobian(const double *x, int *rowptr, int *col, double *val, double *diagB)
{
    double *xx = malloc(sizeof(double) * g_ndim);
    memcpy(xx, x, sizeof(double) * g_ndim);
    if (g_cfg.vmix == 2) { int z = 0; _QMm_mixPset_vmix_fix(&z); }   /* Ocean.C:1302 */
    matrix_(xx);
    free(xx);
    memcpy(rowptr, g_grow, sizeof(int) * (g_ndim + 1));
    memcpy(col, g_gcol, sizeof(int) * g_gnnz);
    memset(val, 0, sizeof(double) * g_gnnz);
    int dropped = 0;
    for (int r = 0; r < g_ndim; r++) {
        diagB[r] = g_coB[r];
        if (g_cfg.sres == 0 && r == g_rowintcon) continue;
        for (int p = g_begA[r] - 1; p < g_begA[r + 1] - 1; p++) {
            int c = g_jcoA[p] - 1, lo = g_grow[r], hi = g_grow[r + 1] - 1, hit = -1;
            while (lo <= hi) {
                int mid = (lo + hi) / 2;
                if (g_gcol[mid] == c) { hit = mid; break; }
                if (g_gcol[mid] < c) lo = mid + 1; else hi = mid - 1;
            }
            if (hit >= 0) val[hit] = g_coA[p]; else dropped++;
        }
    }
    if (g_cfg.sres == 0) {
        /* THCM::intcond_S (THCM.C:2180-2257) */
        diagB[g_rowintcon] = 0.0;
        for (int p = g_grow[g_rowintcon]; p < g_grow[g_rowintcon + 1]; p++)
            val[p] = g_cfg.int_sign * g_coef[g_gcol[p]];
    }
    return dropped;
}
