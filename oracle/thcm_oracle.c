/*
 * THCM assembly oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in plain C, of the reference algorithm that builds the
 * residual F(x) and the Jacobian J(x) of the THCM ocean model behind
 * Ocean::computeRHS / Ocean::computeJacobian (one rank).  Only tests/,
 * bench.py's cpu_baseline leg and __graft_entry__.smoke() may load it, and
 * only as the checker.  The product never links or calls it.
 *
 * It keeps the reference's own structure on purpose: a dense 27-point x 6 x 6
 * "atom" array per cell (src/ocean/mat.F90:38-40), linear atoms built once per
 * parameter change (usrc.F90:605-789 lin), nonlinear atoms added per
 * evaluation (usrc.F90:792-1012), boundary edits (boundary.F90:2-393),
 * assembly into a CSR with |a|<=1e-10 dropped (assemble.F90:57-139), and the
 * C++ glue of THCM.C (negation, integral row, fixed graph).  Every floating
 * point expression keeps the reference's operand order so that results round
 * the same way; compile without FMA contraction.
 *
 * Scope: uncoupled ocean (coupled_T = coupled_S = 0), Mixing 0/1/2 with the
 * consistent and implicit vertical mixing terms (MIXP = MKAP = 0: neutral
 * physics and Gent-McWilliams are rejected with an error code).
 *
 * Indices: Fortran-style 1-based i,j,k in the comments; unknown XX=1..6 for
 * u,v,w,p,T,S; global row 6*((k-1)*n*m + (j-1)*n + i-1) + XX-1 (0-based),
 * matetc.F90:123-144.  Stencil position kk=1..27 (mat.F90 / assemble.F90:66):
 *   kk-1 = 9*lev + 3*(di+1) + (dj+1), lev 0 centre / 1 below / 2 above.
 */
#include "thcm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { UU = 1, VV, WW, PP, TT, SS };
enum { OCEAN = 0, LAND = 1, WATER = 2, PERIO = 3 };
enum { AL_T = 1, RAYL, EK_V, EK_H, ROSB, MIXP, RESC, SPL1, HMTP, SUNP, PE_H, PE_V, P_VC,
       LAMB, SALT, WIND, TEMP, BIOT, COMB, ARCL, NLES, IFRICB, CONT, ENER, ALPC, CMPR,
       FPER, SPER, MKAP, SPL2 };

/* usr.F90:131-160, par.F90:18 */
static const double pi_ = 3.14159265358979323846;
static const double omegadim = 7.292e-05, r0dim = 6.37e+06, udim = 0.1e+00;
static const double gdim = 9.8e+00, rhodim = 1.024e+03;
static const double alpt1 = 2.93, alpt2 = 8.3e-02, alpt3 = 6.6e-04;
static const double ah = 2.5e+05, av = 1.0e-03, kappah = 1.0e+03, kappav = 1.0e-04;

struct tho_model {
    tho_config c;
    int n, m, l, ndim, ncell, periodic;
    double xmin, xmax, ymin, ymax, dx, dy, dz;
    double *x, *y, *z, *xu, *yv, *zw, *dfzT, *dfzW;   /* Fortran lower bounds kept */
    double *y_mem, *xu_mem, *yv_mem, *zw_mem, *dfzW_mem;
    int *landm;                                      /* (0:n+1,0:m+1,0:l+1)      */
    double par[31];
    double *taux, *tauy, *tatm, *emip, *spert, *aemip; /* (n,m)                  */
    double *Frc;                                      /* ndim                     */
    double *Al, *An;                                  /* 27*36 per cell           */
    int vmix_flag, vmix_temp, vmix_salt, vmix_fix;
    /* the mixing Jacobian's pattern (mix_imp.f:168-227 vmix_part with
     * vmix_el_1 / vmix_el_2): built at all (vmix_dim > 0), and with the
     * salinity columns and rows (vmix_el_2 under vmix_salt = 1) */
    int vmix_part_done, vmix_pat_salt;
    int rowintcon;
    double *coef;
    int *grow, *gcol, gnnz;
    /* scratch: padded solution fields (usrc.F90:1014-1121 layout) */
    double *u, *v, *w, *p, *t, *s;
    /* Fortran CSR of the last assembly */
    int *begA, *jcoA; double *coA;
};

/* ---------------------------------------------------------------- indexing */
#define LM(h, i, j, k) ((h)->landm[(i) + ((h)->n + 2) * ((j) + ((h)->m + 2) * (k))])
/* u,v: (0:n,0:m,0:l+1); w: (0:n+1,0:m+1,0:l); p,t,s: (0:n+1,0:m+1,0:l+1) */
#define UV(h, a, i, j, k) ((a)[(i) + ((h)->n + 1) * ((j) + ((h)->m + 1) * (k))])
#define WA(h, i, j, k) ((h)->w[(i) + ((h)->n + 2) * ((j) + ((h)->m + 2) * (k))])
#define TS(h, a, i, j, k) ((a)[(i) + ((h)->n + 2) * ((j) + ((h)->m + 2) * (k))])
#define CELL(h, i, j, k) (((i) - 1) + (h)->n * (((j) - 1) + (h)->m * ((k) - 1)))
#define ROW(h, i, j, k, xx) (6 * CELL(h, i, j, k) + (xx) - 1)
/* atom array: one 27-vector per (cell, eq, var) */
#define AT(A, cell, ii, jj, kk) ((A)[(((size_t)(cell) * 6 + (ii) - 1) * 6 + (jj) - 1) * 27 + (kk) - 1])
#define ATOM(a, cell, kk) ((a)[(size_t)(cell) * 27 + (kk) - 1])

static double fz(double z, double qz)   /* grid.F90:58-73 */
{
    double th = tanh(qz * (z + 1)), tth = tanh(qz);
    if (qz > 1.0) return -1 + th / tth;
    return z + (1. - qz) * z * (1 - z);
}

static double dfdz(double z, double qz) /* grid.F90:75-91 */
{
    double ch = cosh(qz * (z + 1)), tth = tanh(qz);
    if (qz > 1.0) return qz / (tth * ch * ch);
    return 1.0 + (1. - qz) * (1. - 2. * z);
}

/* spf.F90:792-851: inhomogeneous mixing profiles */
static double amh(double y, int ih) { return ih == 0 ? 1.0 : 1. + 10.0 * exp(-5 * y * y); }
static double bmh(double y, int ih) { return ih == 0 ? 1.0 : 1.0 + 10.0 * exp(-5 * y * y); }
static double bmhy(double y, int ih) { return ih == 0 ? 0.0 : -10. * 10.0 * y * exp(-5 * y * y); }

/* ---------------------------------------------------------------- grid */
static void make_grid(tho_model *h)     /* grid.F90:2-56 */
{
    int n = h->n, m = h->m, l = h->l;
    double qz = h->c.qz, zmin = -1.0, zmax = 0.0;
    h->dx = (h->xmax - h->xmin) / n;
    h->dy = (h->ymax - h->ymin) / m;
    h->dz = (zmax - zmin) / l;
    for (int i = 1; i <= n; i++) {
        h->x[i] = ((double)i - 0.5) * h->dx + h->xmin;
        h->xu[i] = ((double)i) * h->dx + h->xmin;
    }
    h->xu[0] = h->xmin;
    for (int j = 1; j <= m; j++) {
        h->y[j] = ((double)j - 0.5) * h->dy + h->ymin;
        h->yv[j] = ((double)j) * h->dy + h->ymin;
    }
    h->y[0] = h->y[1] - h->dy;
    h->y[m + 1] = h->y[m] + h->dy;
    h->yv[0] = h->ymin;
    for (int k = 1; k <= l; k++) {
        double ze = ((double)k - 0.5) * h->dz + zmin;
        double zwe = ((double)k) * h->dz + zmin;
        h->z[k] = fz(ze, qz);
        h->zw[k] = fz(zwe, qz);
        h->dfzT[k] = dfdz(ze, qz);
        h->dfzW[k] = dfdz(zwe, qz);
    }
    h->zw[0] = zmin;
    h->dfzW[0] = dfdz(zmin, qz);
}

/* usrc.F90:1123-1160 stpnt + mix_imp.f:142-160 vmix_par */
static void stpnt(tho_model *h)
{
    double *par = h->par, hdim = h->c.hdim;
    double dz = h->dz, dfl = h->dfzT[h->l];
    par[AL_T] = 0.1 / (2 * omegadim * rhodim * hdim * udim * dz * dfl);
    par[RAYL] = h->c.alphaT * gdim * hdim / (2 * omegadim * udim * r0dim);
    par[EK_V] = av / (2 * omegadim * hdim * hdim);
    par[EK_H] = ah / (2 * omegadim * r0dim * r0dim);
    par[ROSB] = udim / (2 * omegadim * r0dim);
    par[HMTP] = 0.0;
    par[SUNP] = 0.0;
    par[PE_H] = kappah / (udim * r0dim);
    par[PE_V] = kappav * r0dim / (udim * hdim * hdim);
    par[P_VC] = 2.5e+04 * par[PE_V];
    par[LAMB] = h->c.alphaS / h->c.alphaT;
    par[SALT] = 0.0;
    par[WIND] = 0.0;
    par[TEMP] = 0.0;
    par[BIOT] = r0dim / (75. * 3600. * 24. * udim);
    par[COMB] = 0.0;
    par[NLES] = 0.0;
    par[CMPR] = 0.0;
    par[ALPC] = 1.0;
    par[ENER] = 1.0e+02;
    par[MIXP] = 0.0;
    par[MKAP] = 0.0;
    par[SPL1] = 2.0e+03;
    par[SPL2] = 0.01;
    if (h->c.vmix == 0) {
        par[MIXP] = 0.0; par[P_VC] = 0.0; par[ALPC] = 1.0; par[ENER] = 1.0e+2; par[MKAP] = 0.0;
    }
}

/* ---------------------------------------------------------------- forcing */
/* THCM.C:2653-2686 (qint -> thcm_forcing_integral_), one rank */
static double qint(const tho_model *h, const double *field)
{
    double lsint = 0.0, lfsint = 0.0;
    for (int j = 1; j <= h->m; j++)
        for (int i = 1; i <= h->n; i++) {
            int lm = LM(h, i, j, h->l);
            lfsint = field[(i - 1) + h->n * (j - 1)] * cos(h->y[j]) * (1 - lm) + lfsint;
            lsint = cos(h->y[j]) * (1 - lm) + lsint;
        }
    return lfsint / lsint;
}

static double temfun(const tho_model *h, double yy)   /* forcing.F90:417-427 */
{
    if (h->c.forcing_type == 2) return cos(pi_ * (yy - h->ymin) / (h->ymax - h->ymin));
    return cos(pi_ * yy / h->ymax) + h->par[CMPR] * sin(pi_ * yy / h->ymax);
}

static double salfun(const tho_model *h, double yy)   /* forcing.F90:430-442 */
{
    if (h->c.forcing_type == 2) return cos(pi_ * (yy - h->ymin) / (h->ymax - h->ymin));
    if (h->c.forcing_type == 1) return (cos(pi_ * yy / h->ymax) + h->par[FPER] * yy / h->ymax) / cos(yy);
    return cos(pi_ * yy / h->ymax) + h->par[FPER] * yy / h->ymax;
}

static double wfun(double yy)                          /* forcing.F90:402-414 */
{
    return 0.2 - 0.8 * sin(6 * fabs(yy)) - 0.5 * (1 - tanh(10 * fabs(yy)))
           - 0.5 * (1 - tanh(10 * (pi_ / 2 - fabs(yy))));
}

/* forcing.F90:4-218, uncoupled branch, then the Frc zeroing that
 * boundary.F90 applies (lines 222-225, 265-272, 296-307, 331-340, 349-368). */
static void forcing(tho_model *h)
{
    int n = h->n, m = h->m, l = h->l;
    double *par = h->par, *Frc = h->Frc;
    int TRES = h->c.tres, SRES = h->c.sres;
    memset(Frc, 0, sizeof(double) * h->ndim);
    double sigma = par[COMB] * par[WIND] * par[AL_T];
    if (h->c.iza == 2)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                h->taux[(i - 1) + n * (j - 1)] = wfun(h->yv[j]);
                h->tauy[(i - 1) + n * (j - 1)] = 0.0;
            }
    for (int j = 1; j <= m - 1; j++)
        for (int i = 1; i <= n; i++) {
            Frc[ROW(h, i, j, l, UU)] = sigma * h->taux[(i - 1) + n * (j - 1)];
            Frc[ROW(h, i, j, l, VV)] = sigma * h->tauy[(i - 1) + n * (j - 1)];
        }
    double etabi = par[COMB] * par[TEMP] * ((double)(1 - TRES) + TRES * par[BIOT]);
    double temcor = 0.0;
    if (h->c.ite == 1) {
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) h->tatm[(i - 1) + n * (j - 1)] = temfun(h, h->y[j]);
        if (TRES == 0) temcor = qint(h, h->tatm);
    }
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++)
            Frc[ROW(h, i, j, l, TT)] = etabi * (h->tatm[(i - 1) + n * (j - 1)] - temcor);
    double gamma = par[COMB] * par[SALT] * ((double)(1 - SRES) + SRES * par[BIOT]);
    double salcor = 0.0, asalcor = 0.0, spertcor = 0.0;
    if (h->c.its == 1) {
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                h->emip[(i - 1) + n * (j - 1)] = salfun(h, h->y[j]) * (1 - LM(h, i, j, l));
        if (SRES == 0) salcor = qint(h, h->emip);
    }
    if (SRES == 0) {
        asalcor = qint(h, h->aemip);
        spertcor = qint(h, h->spert);
    }
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++) {
            int q = (i - 1) + n * (j - 1);
            Frc[ROW(h, i, j, l, SS)] = gamma * (1 - par[HMTP]) * (h->emip[q] - salcor)
                                       + gamma * par[HMTP] * (h->aemip[q] - asalcor)
                                       + par[SPER] * ((double)(1 - SRES) + SRES * par[BIOT]) * (h->spert[q] - spertcor);
        }
    /* internal w forcing (forcing.F90:184-193) with zero internal T/S fields */
    for (int k = 1; k <= l - 1; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                Frc[ROW(h, i, j, k, WW)] = -par[COMB] * (1 - LM(h, i, j, k)) * par[RAYL]
                                           * (par[LAMB] * (0.0 + 0.0) / 2. - (0.0 + 0.0) / 2.);
    /* boundary.F90 zeroes these rows of Frc on every call */
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= m; j++)
            for (int k = 1; k <= l; k++) {
                if (LM(h, i, j, k) == OCEAN) {
                    if (LM(h, i, j, k + 1) == LAND) Frc[ROW(h, i, j, k, WW)] = 0.0;
                    if (LM(h, i, j + 1, k) == LAND || LM(h, i + 1, j, k) == LAND
                        || LM(h, i + 1, j + 1, k) == LAND) {
                        Frc[ROW(h, i, j, k, UU)] = 0.0;
                        Frc[ROW(h, i, j, k, VV)] = 0.0;
                    }
                } else {
                    for (int xx = 1; xx <= 6; xx++) Frc[ROW(h, i, j, k, xx)] = 0.0;
                }
            }
}

/* ---------------------------------------------------------------- atoms */
/* Each atom builder fills a (27, n, m, l) array; spf.F90 line ranges cited. */
static void zero_atom(const tho_model *h, double *a) { memset(a, 0, sizeof(double) * 27 * h->ncell); }

static void uderiv(const tho_model *h, int type, double *a)   /* spf.F90:13-67 */
{
    int n = h->n, m = h->m, l = h->l, ih = h->c.ih;
    double dx = h->dx, dy = h->dy, dz = h->dz;
    zero_atom(h, a);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                int c = CELL(h, i, j, k);
                double yv = h->yv[j];
                if (type == 2 && j <= m - 1) {
                    double q = 1.0 / (cos(yv) * dx);
                    double cosdx2i = q * q;
                    ATOM(a, c, 2) = amh(yv, ih) * cosdx2i;
                    ATOM(a, c, 8) = amh(yv, ih) * cosdx2i;
                    ATOM(a, c, 5) = -(ATOM(a, c, 2) + ATOM(a, c, 8));
                } else if (type == 3 && j <= m - 1) {
                    double q = 1.0 / dy, rdy2i = q * q;
                    ATOM(a, c, 4) = rdy2i * bmh(h->y[j], ih) * cos(h->y[j]) / cos(yv);
                    ATOM(a, c, 6) = rdy2i * bmh(h->y[j + 1], ih) * cos(h->y[j + 1]) / cos(yv);
                    ATOM(a, c, 5) = -(ATOM(a, c, 4) + ATOM(a, c, 6));
                } else if (type == 4) {
                    double q = 1.0 / dz, rdz2i = q * q;
                    double h1 = 1. / (h->dfzT[k] * h->dfzW[k]);
                    double h2 = 1. / (h->dfzT[k] * h->dfzW[k - 1]);
                    ATOM(a, c, 14) = h2 * rdz2i;
                    ATOM(a, c, 23) = h1 * rdz2i;
                    ATOM(a, c, 5) = -(ATOM(a, c, 14) + ATOM(a, c, 23));
                } else if (type == 5 && j <= m - 1) {
                    double tand2 = 1 - tan(yv) * tan(yv);
                    ATOM(a, c, 5) = bmh(yv, ih) * tand2 + tan(yv) * bmhy(yv, ih);
                } else if (type == 6 && j <= m - 1) {
                    double num = bmhy(yv, ih) - (amh(yv, ih) + bmh(yv, ih)) * tan(yv);
                    ATOM(a, c, 2) = num / (dx * cos(yv));
                    ATOM(a, c, 8) = -num / (dx * cos(yv));
                }
            }
}

static void vderiv(const tho_model *h, int type, double *a)   /* spf.F90:69-121 */
{
    int n = h->n, m = h->m, l = h->l, ih = h->c.ih;
    double dx = h->dx, dy = h->dy, dz = h->dz;
    zero_atom(h, a);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                int c = CELL(h, i, j, k);
                double yv = h->yv[j];
                if (type == 2 && j <= m - 1) {
                    double q = 1.0 / (cos(yv) * dx), cosdx2i = q * q;
                    ATOM(a, c, 2) = bmh(yv, ih) * cosdx2i;
                    ATOM(a, c, 5) = -2 * bmh(yv, ih) * cosdx2i;
                    ATOM(a, c, 8) = bmh(yv, ih) * cosdx2i;
                } else if (type == 3 && j <= m - 1) {
                    double q = 1.0 / dy, dy2i = q * q;
                    ATOM(a, c, 4) = dy2i * amh(h->y[j], ih) * cos(h->y[j]) / cos(yv);
                    ATOM(a, c, 6) = dy2i * amh(h->y[j + 1], ih) * cos(h->y[j + 1]) / cos(yv);
                    ATOM(a, c, 5) = -(ATOM(a, c, 4) + ATOM(a, c, 6));
                } else if (type == 4) {
                    double q = 1.0 / dz, rdz2i = q * q;
                    double h1 = 1. / (h->dfzT[k] * h->dfzW[k]);
                    double h2 = 1. / (h->dfzT[k] * h->dfzW[k - 1]);
                    ATOM(a, c, 14) = h2 * rdz2i;
                    ATOM(a, c, 23) = h1 * rdz2i;
                    ATOM(a, c, 5) = -(ATOM(a, c, 14) + ATOM(a, c, 23));
                } else if (type == 5 && j <= m - 1) {
                    ATOM(a, c, 5) = bmh(yv, ih) - amh(yv, ih) * tan(yv) * tan(yv) + bmhy(yv, ih) * tan(yv);
                } else if (type == 6 && j <= m - 1) {
                    double num = (amh(yv, ih) + bmh(yv, ih)) * tan(yv) - bmhy(yv, ih);
                    ATOM(a, c, 2) = -num / (dx * cos(yv));
                    ATOM(a, c, 8) = num / (dx * cos(yv));
                }
            }
}

static void pderiv(const tho_model *h, int type, double *a)   /* spf.F90:123-171 */
{
    int n = h->n, m = h->m, l = h->l;
    zero_atom(h, a);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                int c = CELL(h, i, j, k);
                if (type == 1) {
                    double cos2i = 1.0 / (2 * cos(h->y[j]) * h->dx);
                    ATOM(a, c, 2) = -cos2i; ATOM(a, c, 4) = cos2i;
                    ATOM(a, c, 1) = -cos2i; ATOM(a, c, 5) = cos2i;
                } else if (type == 2) {
                    double cos2i = 1. / (2 * cos(h->y[j]) * h->dy);
                    ATOM(a, c, 4) = -cos(h->yv[j - 1]) * cos2i;
                    ATOM(a, c, 2) = cos(h->yv[j]) * cos2i;
                    ATOM(a, c, 1) = -cos(h->yv[j - 1]) * cos2i;
                    ATOM(a, c, 5) = cos(h->yv[j]) * cos2i;
                } else if (type == 3) {
                    double dzi = 1.0 / h->dz;
                    ATOM(a, c, 5) = dzi / h->dfzT[k];
                    ATOM(a, c, 14) = -dzi / h->dfzT[k];
                }
            }
}

static void tderiv(const tho_model *h, int type, double *a)   /* spf.F90:173-259 */
{
    int n = h->n, m = h->m, l = h->l;
    zero_atom(h, a);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                int c = CELL(h, i, j, k);
                int om = 1 - LM(h, i, j, l);
                if ((type == 1 || type == 2) && k == l) {
                    ATOM(a, c, 5) = 1.0;
                } else if (type == 3) {
                    double q = 1.0 / (cos(h->y[j]) * h->dx), cosdx2i = q * q;
                    ATOM(a, c, 2) = cosdx2i * om;
                    ATOM(a, c, 5) = -2 * cosdx2i * om;
                    ATOM(a, c, 8) = cosdx2i * om;
                } else if (type == 4) {
                    double q = 1.0 / h->dy, dy2i = q * q;
                    ATOM(a, c, 4) = (dy2i * cos(h->yv[j - 1]) / cos(h->y[j])) * om;
                    ATOM(a, c, 6) = (dy2i * cos(h->yv[j]) / cos(h->y[j])) * om;
                    ATOM(a, c, 5) = -(ATOM(a, c, 4) + ATOM(a, c, 6));
                } else if (type == 5) {
                    double q = 1.0 / h->dz, dz2i = q * q;
                    double h1 = 1. / (h->dfzT[k] * h->dfzW[k]);
                    double h2 = 1. / (h->dfzT[k] * h->dfzW[k - 1]);
                    ATOM(a, c, 14) = h2 * dz2i * om;
                    ATOM(a, c, 23) = (k < l) ? h1 * dz2i * om : 0.0;
                    ATOM(a, c, 5) = -(ATOM(a, c, 14) + ATOM(a, c, 23));
                } else if (type == 6) {
                    ATOM(a, c, 23) = 1.0 * om;
                    ATOM(a, c, 5) = 1.0 * om;
                }
            }
}

static void coriolis(const tho_model *h, double *a)            /* spf.F90:262-293 */
{
    zero_atom(h, a);
    for (int k = 1; k <= h->l; k++)
        for (int j = 1; j <= h->m - 1; j++)
            for (int i = 1; i <= h->n; i++)
                ATOM(a, CELL(h, i, j, k), 5) = sin(h->yv[j]) * h->c.coriolis_on;
}

static void gradp(const tho_model *h, int type, double *a)     /* spf.F90:296-335 */
{
    zero_atom(h, a);
    for (int k = 1; k <= h->l; k++)
        for (int j = 1; j <= h->m; j++)
            for (int i = 1; i <= h->n; i++) {
                int c = CELL(h, i, j, k);
                if (type == 1 && j <= h->m - 1) {
                    double cosdxi = 1. / (2 * cos(h->yv[j]) * h->dx);
                    ATOM(a, c, 5) = -cosdxi; ATOM(a, c, 6) = -cosdxi;
                    ATOM(a, c, 8) = cosdxi; ATOM(a, c, 9) = cosdxi;
                } else if (type == 2 && j <= h->m - 1) {
                    double dyi = 1. / (2 * h->dy);
                    ATOM(a, c, 5) = -dyi; ATOM(a, c, 8) = -dyi;
                    ATOM(a, c, 6) = dyi; ATOM(a, c, 9) = dyi;
                } else if (type == 3) {
                    double dzi = 1. / h->dz;
                    ATOM(a, c, 5) = -dzi / h->dfzW[k];
                    ATOM(a, c, 23) = dzi / h->dfzW[k];
                }
            }
}

/* usrc.F90:605-789 lin, uncoupled branches */
static void lin(tho_model *h)
{
    size_t na = (size_t)27 * h->ncell;
    double *A[8];
    for (int q = 0; q < 8; q++) A[q] = malloc(sizeof(double) * na);
    double *par = h->par;
    double EV = par[EK_V], EH = par[EK_H];
    double ph = (1 - par[MIXP]) * par[PE_H], pv = par[PE_V];
    double lambda = par[LAMB], xes = par[NLES], bi = par[BIOT], Ra = par[RAYL];
    memset(h->Al, 0, sizeof(double) * 36 * na);
    double *Al = h->Al;
    /* u-equation */
    uderiv(h, 2, A[0]); uderiv(h, 3, A[1]); uderiv(h, 4, A[2]); uderiv(h, 5, A[3]);
    uderiv(h, 6, A[4]); coriolis(h, A[5]); gradp(h, 1, A[6]);
    for (int c = 0; c < h->ncell; c++)
        for (int kk = 1; kk <= 27; kk++) {
            AT(Al, c, UU, UU, kk) = -EH * (ATOM(A[0], c, kk) + ATOM(A[1], c, kk) + ATOM(A[3], c, kk))
                                    - EV * ATOM(A[2], c, kk);
            AT(Al, c, UU, VV, kk) = -ATOM(A[5], c, kk) - EH * ATOM(A[4], c, kk);
            AT(Al, c, UU, PP, kk) = ATOM(A[6], c, kk);
        }
    /* v-equation */
    vderiv(h, 2, A[0]); vderiv(h, 3, A[1]); vderiv(h, 4, A[2]); vderiv(h, 5, A[3]);
    vderiv(h, 6, A[4]); coriolis(h, A[5]); gradp(h, 2, A[6]);
    for (int c = 0; c < h->ncell; c++)
        for (int kk = 1; kk <= 27; kk++) {
            AT(Al, c, VV, UU, kk) = ATOM(A[5], c, kk) - EH * ATOM(A[4], c, kk);
            AT(Al, c, VV, VV, kk) = -EH * (ATOM(A[0], c, kk) + ATOM(A[1], c, kk) + ATOM(A[3], c, kk))
                                    - EV * ATOM(A[2], c, kk);
            AT(Al, c, VV, PP, kk) = ATOM(A[6], c, kk);
        }
    /* w-equation */
    gradp(h, 3, A[0]); tderiv(h, 6, A[1]);
    for (int c = 0; c < h->ncell; c++)
        for (int kk = 1; kk <= 27; kk++) {
            AT(Al, c, WW, PP, kk) = ATOM(A[0], c, kk);
            AT(Al, c, WW, TT, kk) = -Ra * (1. + xes * alpt1) * ATOM(A[1], c, kk) / 2.;
            AT(Al, c, WW, SS, kk) = lambda * Ra * ATOM(A[1], c, kk) / 2.;
        }
    /* continuity */
    pderiv(h, 1, A[0]); pderiv(h, 2, A[1]); pderiv(h, 3, A[2]);
    for (int c = 0; c < h->ncell; c++)
        for (int kk = 1; kk <= 27; kk++) {
            AT(Al, c, PP, UU, kk) = ATOM(A[0], c, kk);
            AT(Al, c, PP, VV, kk) = ATOM(A[1], c, kk);
            AT(Al, c, PP, WW, kk) = ATOM(A[2], c, kk);
        }
    /* T and S */
    tderiv(h, 1, A[0]); tderiv(h, 2, A[1]); tderiv(h, 3, A[2]); tderiv(h, 4, A[3]);
    tderiv(h, 5, A[4]);
    for (int c = 0; c < h->ncell; c++)
        for (int kk = 1; kk <= 27; kk++) {
            double txx = ATOM(A[2], c, kk), tyy = ATOM(A[3], c, kk), tzz = ATOM(A[4], c, kk);
            AT(Al, c, TT, TT, kk) = -ph * (txx + tyy) - pv * tzz + h->c.tres * bi * ATOM(A[0], c, kk);
            AT(Al, c, SS, SS, kk) = -ph * (txx + tyy) - pv * tzz + h->c.sres * bi * ATOM(A[1], c, kk);
        }
    for (int q = 0; q < 8; q++) free(A[q]);
}

/* usrc.F90:1014-1105 usol: x -> padded staggered fields */
static void usol(tho_model *h, const double *un)
{
    int n = h->n, m = h->m, l = h->l;
    size_t nuv = (size_t)(n + 1) * (m + 1) * (l + 2), nts = (size_t)(n + 2) * (m + 2) * (l + 2);
    size_t nw = (size_t)(n + 2) * (m + 2) * (l + 1);
    memset(h->u, 0, sizeof(double) * nuv); memset(h->v, 0, sizeof(double) * nuv);
    memset(h->w, 0, sizeof(double) * nw);
    memset(h->p, 0, sizeof(double) * nts); memset(h->t, 0, sizeof(double) * nts);
    memset(h->s, 0, sizeof(double) * nts);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                UV(h, h->u, i, j, k) = un[ROW(h, i, j, k, UU)];
                UV(h, h->v, i, j, k) = un[ROW(h, i, j, k, VV)];
                WA(h, i, j, k) = un[ROW(h, i, j, k, WW)];
                TS(h, h->p, i, j, k) = un[ROW(h, i, j, k, PP)];
                TS(h, h->t, i, j, k) = un[ROW(h, i, j, k, TT)];
                TS(h, h->s, i, j, k) = un[ROW(h, i, j, k, SS)];
            }
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++) {
            if (h->periodic) {
                UV(h, h->u, 0, j, k) = UV(h, h->u, n, j, k);
                UV(h, h->v, 0, j, k) = UV(h, h->v, n, j, k);
                WA(h, n + 1, j, k) = WA(h, 1, j, k);
                WA(h, 0, j, k) = WA(h, n, j, k);
                TS(h, h->p, n + 1, j, k) = TS(h, h->p, 1, j, k);
                TS(h, h->p, 0, j, k) = TS(h, h->p, n, j, k);
                TS(h, h->t, n + 1, j, k) = TS(h, h->t, 1, j, k);
                TS(h, h->t, 0, j, k) = TS(h, h->t, n, j, k);
                TS(h, h->s, n + 1, j, k) = TS(h, h->s, 1, j, k);
                TS(h, h->s, 0, j, k) = TS(h, h->s, n, j, k);
            } else {
                UV(h, h->u, 0, j, k) = 0.0; UV(h, h->u, n, j, k) = 0.0;
                UV(h, h->v, 0, j, k) = 0.0; UV(h, h->v, n, j, k) = 0.0;
                TS(h, h->p, 0, j, k) = 0.0; TS(h, h->p, n + 1, j, k) = 0.0;
                TS(h, h->t, 0, j, k) = TS(h, h->t, 1, j, k);
                TS(h, h->t, n + 1, j, k) = TS(h, h->t, n, j, k);
                TS(h, h->s, 0, j, k) = TS(h, h->s, 1, j, k);
                TS(h, h->s, n + 1, j, k) = TS(h, h->s, n, j, k);
            }
        }
    for (int k = 1; k <= l; k++)
        for (int i = 1; i <= n; i++) {
            UV(h, h->u, i, 0, k) = 0.0; UV(h, h->u, i, m, k) = 0.0;
            UV(h, h->v, i, 0, k) = 0.0; UV(h, h->v, i, m, k) = 0.0;
            TS(h, h->p, i, 0, k) = 0.0; TS(h, h->p, i, m + 1, k) = 0.0;
            TS(h, h->t, i, 0, k) = TS(h, h->t, i, 1, k);
            TS(h, h->t, i, m + 1, k) = TS(h, h->t, i, m, k);
            TS(h, h->s, i, 0, k) = TS(h, h->s, i, 1, k);
            TS(h, h->s, i, m + 1, k) = TS(h, h->s, i, m, k);
        }
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++) {
            UV(h, h->u, i, j, 0) = UV(h, h->u, i, j, 1);
            UV(h, h->u, i, j, l + 1) = UV(h, h->u, i, j, l);
            UV(h, h->v, i, j, 0) = UV(h, h->v, i, j, 1);
            UV(h, h->v, i, j, l + 1) = UV(h, h->v, i, j, l);
            WA(h, i, j, l) = 0.0;
            WA(h, i, j, 0) = 0.0;
            TS(h, h->p, i, j, l + 1) = 0.0; TS(h, h->p, i, j, 0) = 0.0;
            TS(h, h->t, i, j, l + 1) = TS(h, h->t, i, j, l);
            TS(h, h->t, i, j, 0) = TS(h, h->t, i, j, 1);
            TS(h, h->s, i, j, l + 1) = TS(h, h->s, i, j, l);
            TS(h, h->s, i, j, 0) = TS(h, h->s, i, j, 1);
        }
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= m; j++)
            for (int k = 1; k <= l; k++)
                if (LM(h, i, j, k) == 1) {
                    UV(h, h->u, i, j, k) = 0.0; UV(h, h->v, i, j, k) = 0.0;
                    UV(h, h->u, i - 1, j, k) = 0.0; UV(h, h->v, i - 1, j, k) = 0.0;
                    UV(h, h->u, i, j - 1, k) = 0.0; UV(h, h->v, i, j - 1, k) = 0.0;
                    UV(h, h->u, i - 1, j - 1, k) = 0.0; UV(h, h->v, i - 1, j - 1, k) = 0.0;
                }
}

/* Nonlinear atoms (spf.F90:362-790) evaluated for one cell into a local
 * 27-vector; 'kind' names the spf routine and type. */
enum { UNLIN = 0, VNLIN = 10, TNLIN_T = 20, TNLIN_S = 30, WNLIN = 40 };

static void nl_atom(const tho_model *h, int kind, int i, int j, int k, double *a)
{
    int n = h->n, m = h->m, l = h->l;
    const double *u = h->u, *v = h->v;
    double dx = h->dx, dy = h->dy, dz = h->dz;
    for (int q = 0; q < 27; q++) a[q] = 0.0;
#define A(kk) a[(kk) - 1]
#define U(ii, jj, kk) UV(h, u, ii, jj, kk)
#define V(ii, jj, kk) UV(h, v, ii, jj, kk)
#define W(ii, jj, kk) WA(h, ii, jj, kk)
    int type = kind % 10, fam = kind - type;
    if (fam == UNLIN || fam == VNLIN) {
        const double *vel1 = (fam == UNLIN) ? u : v;   /* the 'own' velocity */
        double cx = 1.0 / (2 * cos(h->yv[j]) * dx);
        double cy = 1.0 / (2 * cos(h->yv[j]) * dy);
        double tdzi = 1.0 / (8 * h->dfzT[k] * dz);
        double tanr = tan(h->yv[j]);
        switch (type) {
        case 1: /* unlin uux / vnlin uvx */
            if (i <= n - 1) A(8) = U(i + 1, j, k) * cx;
            if (i >= 2) A(2) = -U(i - 1, j, k) * cx;
            break;
        case 2: /* unlin Urux / vnlin uVrx */
            if (fam == UNLIN) {
                if (i <= n - 1) A(8) = 2 * U(i + 1, j, k) * cx;
                if (i >= 2) A(2) = -2 * U(i - 1, j, k) * cx;
            } else {
                if (i <= n - 1) A(8) = V(i + 1, j, k) * cx;
                if (i >= 2) A(2) = -V(i - 1, j, k) * cx;
            }
            break;
        case 3: /* unlin uvy1 / vnlin vvry */
            if (j >= 2) A(4) = -V(i, j - 1, k) * cos(h->yv[j - 1]) * cy;
            if (j <= m - 1) A(6) = V(i, j + 1, k) * cos(h->yv[j + 1]) * cy;
            break;
        case 4: /* unlin Urvy1 / vnlin Vrvy */
            if (fam == UNLIN) {
                if (j >= 2) A(4) = -U(i, j - 1, k) * cos(h->yv[j - 1]) * cy;
                if (j <= m - 1) A(6) = U(i, j + 1, k) * cos(h->yv[j + 1]) * cy;
            } else {
                if (j <= m - 1) A(6) = 2 * V(i, j + 1, k) * cos(h->yv[j + 1]) * cy;
                if (j >= 2) A(4) = -2 * V(i, j - 1, k) * cos(h->yv[j - 1]) * cy;
            }
            break;
        case 5: /* uwz / vwz */
            A(23) = (W(i, j, k) + W(i, j + 1, k) + W(i + 1, j, k) + W(i + 1, j + 1, k)) * tdzi;
            A(14) = -(W(i, j, k - 1) + W(i, j + 1, k - 1) + W(i + 1, j, k - 1) + W(i + 1, j + 1, k - 1)) * tdzi;
            A(5) = A(14) + A(23);
            break;
        case 6: { /* Urwz / Vrwz */
            double up = (UV(h, vel1, i, j, k) + UV(h, vel1, i, j, k + 1)) * tdzi;
            double dn = -(UV(h, vel1, i, j, k) + UV(h, vel1, i, j, k - 1)) * tdzi;
            A(5) = up; A(6) = up; A(8) = up; A(9) = up;
            A(14) = dn; A(15) = dn; A(17) = dn; A(18) = dn;
            break;
        }
        case 7: /* unlin uvy2: v*tan ; vnlin wvrz: u*tan */
            A(5) = ((fam == UNLIN) ? V(i, j, k) : U(i, j, k)) * tanr;
            break;
        case 8: /* unlin Urvy2: u*tan ; vnlin Urt2: 2u*tan */
            A(5) = (fam == UNLIN) ? U(i, j, k) * tanr : 2 * U(i, j, k) * tanr;
            break;
        }
    } else if (fam == TNLIN_T || fam == TNLIN_S) {
        const double *t = (fam == TNLIN_T) ? h->t : h->s;
#define T(ii, jj, kk) TS(h, t, ii, jj, kk)
        int om = 1 - LM(h, i, j, l);
        double cx = 1.0 / (4 * cos(h->y[j]) * dx);
        double cy = 1.0 / (4 * cos(h->y[j]) * dy);
        double tdzi = 1.0 / (2 * dz);
        switch (type) {
        case 2: /* urTx */
            A(2) = -(T(i, j, k) + T(i - 1, j, k)) * cx * om;
            A(4) = (T(i + 1, j, k) + T(i, j, k)) * cx * om;
            A(1) = -(T(i, j, k) + T(i - 1, j, k)) * cx * om;
            A(5) = (T(i + 1, j, k) + T(i, j, k)) * cx * om;
            break;
        case 3: /* Utrx */
            A(2) = -(U(i - 1, j, k) + U(i - 1, j - 1, k)) * cx * om;
            A(8) = (U(i, j, k) + U(i, j - 1, k)) * cx * om;
            A(5) = A(2) + A(8);
            break;
        case 4: /* vrTy */
            A(4) = -cy * (T(i, j, k) + T(i, j - 1, k)) * cos(h->yv[j - 1]) * om;
            A(1) = -cy * (T(i, j, k) + T(i, j - 1, k)) * cos(h->yv[j - 1]) * om;
            A(5) = cy * (T(i, j + 1, k) + T(i, j, k)) * cos(h->yv[j]) * om;
            A(2) = cy * (T(i, j + 1, k) + T(i, j, k)) * cos(h->yv[j]) * om;
            break;
        case 5: /* Vtry */
            A(4) = -(V(i, j - 1, k) + V(i - 1, j - 1, k)) * cy * cos(h->yv[j - 1]) * om;
            A(6) = (V(i, j, k) + V(i - 1, j, k)) * cy * cos(h->yv[j]) * om;
            A(5) = A(4) + A(6);
            break;
        case 6: /* wrTz */
            A(14) = -tdzi * om * (T(i, j, k) + T(i, j, k - 1)) / h->dfzT[k];
            A(5) = (k <= l - 1) ? tdzi * om * (T(i, j, k + 1) + T(i, j, k)) / h->dfzT[k] : 0.0;
            break;
        case 7: /* Wtrz */
            A(14) = -W(i, j, k - 1) * om * tdzi / h->dfzT[k];
            A(23) = W(i, j, k) * om * tdzi / h->dfzT[k];
            A(5) = A(14) + A(23);
            break;
        }
#undef T
    } else if (fam == WNLIN && k <= l - 1) {
        const double *t = h->t;
        double t0 = TS(h, t, i, j, k), t1 = TS(h, t, i, j, k + 1);
        switch (type) {
        case 1: A(23) = (t0 + t1) / 2.; A(5) = (t0 + t1) / 2.; break;
        case 2: A(23) = t1 / 4.; A(5) = (t0 + 2 * t1) / 4.; break;
        case 3: A(5) = 0.375 * ((t0 + t1) * (t0 + t1)); A(23) = 0.375 * ((t0 + t1) * (t0 + t1)); break;
        case 4:
            A(5) = 0.125 * (t0 * t0 + 3 * t1 * t0 + 3 * t1 * t1);
            A(23) = 0.125 * t1 * t1;
            break;
        }
    }
#undef A
#undef U
#undef V
#undef W
}

/* usrc.F90:792-887 nlin_rhs (jac=0) and 890-1012 nlin_jac (jac=1), per cell */
static void nlin(tho_model *h, int jac)
{
    double epsr = h->par[ROSB], Ra = h->par[RAYL], xes = h->par[NLES];
    double a1[27], a2[27], a3[27], a4[27], a5[27], a6[27], a7[27];
    double *An = h->An;
    for (int k = 1; k <= h->l; k++)
        for (int j = 1; j <= h->m; j++)
            for (int i = 1; i <= h->n; i++) {
                int c = CELL(h, i, j, k);
                if (!jac) {
                    nl_atom(h, UNLIN + 1, i, j, k, a1); nl_atom(h, UNLIN + 3, i, j, k, a2);
                    nl_atom(h, UNLIN + 5, i, j, k, a3); nl_atom(h, UNLIN + 7, i, j, k, a4);
                    for (int q = 1; q <= 27; q++)
                        AT(An, c, UU, UU, q) = AT(An, c, UU, UU, q) + epsr * (a1[q - 1] + a2[q - 1] + a3[q - 1] + a4[q - 1]);
                    nl_atom(h, VNLIN + 1, i, j, k, a1); nl_atom(h, VNLIN + 3, i, j, k, a2);
                    nl_atom(h, VNLIN + 5, i, j, k, a3); nl_atom(h, VNLIN + 7, i, j, k, a4);
                    for (int q = 1; q <= 27; q++) {
                        AT(An, c, VV, UU, q) = AT(An, c, VV, UU, q) + epsr * a4[q - 1];
                        AT(An, c, VV, VV, q) = AT(An, c, VV, VV, q) + epsr * (a1[q - 1] + a2[q - 1] + a3[q - 1]);
                    }
                    nl_atom(h, WNLIN + 2, i, j, k, a1); nl_atom(h, WNLIN + 4, i, j, k, a2);
                    for (int q = 1; q <= 27; q++)
                        AT(An, c, WW, TT, q) = AT(An, c, WW, TT, q) - Ra * xes * alpt2 * a1[q - 1]
                                               + Ra * xes * alpt3 * a2[q - 1];
                    for (int f = 0; f < 2; f++) {
                        int fam = f ? TNLIN_S : TNLIN_T, xx = f ? SS : TT;
                        nl_atom(h, fam + 3, i, j, k, a1); nl_atom(h, fam + 5, i, j, k, a2);
                        nl_atom(h, fam + 7, i, j, k, a3);
                        for (int q = 1; q <= 27; q++)  /* ((An + utx) + vty) + wtz */
                            AT(An, c, xx, xx, q) = AT(An, c, xx, xx, q) + a1[q - 1] + a2[q - 1] + a3[q - 1];
                    }
                } else {
                    nl_atom(h, UNLIN + 2, i, j, k, a1); nl_atom(h, UNLIN + 3, i, j, k, a2);
                    nl_atom(h, UNLIN + 4, i, j, k, a3); nl_atom(h, UNLIN + 5, i, j, k, a4);
                    nl_atom(h, UNLIN + 6, i, j, k, a5); nl_atom(h, UNLIN + 7, i, j, k, a6);
                    nl_atom(h, UNLIN + 8, i, j, k, a7);
                    for (int q = 1; q <= 27; q++) {
                        int z = q - 1;
                        AT(An, c, UU, UU, q) = AT(An, c, UU, UU, q) + epsr * (a1[z] + a2[z] + a4[z] + a6[z]);
                        AT(An, c, UU, VV, q) = AT(An, c, UU, VV, q) + epsr * (a3[z] + a7[z]);
                        AT(An, c, UU, WW, q) = AT(An, c, UU, WW, q) + epsr * a5[z];
                    }
                    nl_atom(h, VNLIN + 1, i, j, k, a1); nl_atom(h, VNLIN + 2, i, j, k, a2);
                    nl_atom(h, VNLIN + 4, i, j, k, a3); nl_atom(h, VNLIN + 5, i, j, k, a4);
                    nl_atom(h, VNLIN + 6, i, j, k, a5); nl_atom(h, VNLIN + 8, i, j, k, a6);
                    for (int q = 1; q <= 27; q++) {
                        int z = q - 1;
                        AT(An, c, VV, UU, q) = AT(An, c, VV, UU, q) + epsr * (a6[z] + a2[z]);
                        AT(An, c, VV, VV, q) = AT(An, c, VV, VV, q) + epsr * (a1[z] + a3[z] + a4[z]);
                        AT(An, c, VV, WW, q) = AT(An, c, VV, WW, q) + epsr * a5[z];
                    }
                    nl_atom(h, WNLIN + 1, i, j, k, a1); nl_atom(h, WNLIN + 3, i, j, k, a2);
                    for (int q = 1; q <= 27; q++)
                        AT(An, c, WW, TT, q) = AT(An, c, WW, TT, q) - Ra * xes * alpt2 * a1[q - 1]
                                               + Ra * xes * alpt3 * a2[q - 1];
                    for (int f = 0; f < 2; f++) {
                        int fam = f ? TNLIN_S : TNLIN_T, xx = f ? SS : TT;
                        nl_atom(h, fam + 2, i, j, k, a1); nl_atom(h, fam + 3, i, j, k, a2);
                        nl_atom(h, fam + 4, i, j, k, a3); nl_atom(h, fam + 5, i, j, k, a4);
                        nl_atom(h, fam + 6, i, j, k, a5); nl_atom(h, fam + 7, i, j, k, a6);
                        for (int q = 1; q <= 27; q++) {
                            int z = q - 1;
                            AT(An, c, xx, UU, q) = AT(An, c, xx, UU, q) + a1[z];
                            AT(An, c, xx, VV, q) = AT(An, c, xx, VV, q) + a3[z];
                            AT(An, c, xx, WW, q) = AT(An, c, xx, WW, q) + a5[z];
                            AT(An, c, xx, xx, q) = AT(An, c, xx, xx, q) + a2[z] + a4[z] + a6[z];
                        }
                    }
                }
            }
}

/* ---------------------------------------------------------------- mixing */
static double isoc(const tho_model *h, int i, int j, int k)   /* mix_imp.f:820-835 */
{
    int lm = LM(h, i, j, k);
    return (lm == OCEAN || lm == PERIO) ? 1.0 : 0.0;
}

static double tprstb(double grad, double spl, double alphaT)  /* mix_imp.f:837-856 */
{
    double fac = alphaT * spl;
    double q = -grad * fac;
    double r = tanh(q * q * q);
    return r > 0.0 ? r : 0.0;
}

/* T/S-row value of vmix_fun (mix_imp.f:231-562) at cell (i,j,k) for the
 * vertical terms only (no neutral physics / GM), reading the padded t and s
 * fields; (pk, pv, pd) optionally perturbs t (pv=0) or s (pv=1) at level pk of
 * this column by pd before evaluation (forward-difference columns). */
static void mix_rows(const tho_model *h, int i, int j, int k, int pk, int pvar, double pd,
                     double *mixT, double *mixS)
{
    const double *par = h->par;
    double xes = par[NLES], lambda = par[LAMB];
    double eps = (1.0 - par[ALPC]) * par[ENER] * par[PE_V];
    double kvc = par[P_VC], sp1 = par[SPL1];
    const double epsln = 1.0e-20;
    double Ftz[2], Fsz[2], Fti[2], Fsi[2];   /* faces k-1 (0) and k (1) */
    for (int f = 0; f < 2; f++) {
        int kf = k - 1 + f;                  /* top face of level kf */
        Ftz[f] = Fsz[f] = Fti[f] = Fsi[f] = 0.0;
        if (kf < 1) continue;                /* Ftzt/Ftimp(:,:,0) = 0 */
        double tt[2], ss[2];
        for (int q = 0; q < 2; q++) {
            int kk = kf + q;
            tt[q] = TS(h, h->t, i, j, kk);
            ss[q] = TS(h, h->s, i, j, kk);
            if (kk == pk) {
                if (pvar == 0) tt[q] = tt[q] + pd; else ss[q] = ss[q] + pd;
            }
            /* ghost levels copy level 1 / l (usrc.F90:1086-1096) */
            if (pk == 1 && kk == 0) { if (pvar == 0) tt[q] = tt[q] + pd; else ss[q] = ss[q] + pd; }
            if (pk == h->l && kk == h->l + 1) { if (pvar == 0) tt[q] = tt[q] + pd; else ss[q] = ss[q] + pd; }
        }
        double iso = isoc(h, i, j, kf + 1) * isoc(h, i, j, kf);
        double den = h->dz * h->dfzW[kf];
        double dtdz = iso * (tt[1] - tt[0]) / den;
        double dsdz = iso * (ss[1] - ss[0]) / den;
        double rho0 = lambda * ss[0] - tt[0] - xes * (alpt1 * tt[0] + alpt2 * tt[0] * tt[0] - alpt3 * tt[0] * tt[0] * tt[0]);
        double rho1 = lambda * ss[1] - tt[1] - xes * (alpt1 * tt[1] + alpt2 * tt[1] * tt[1] - alpt3 * tt[1] * tt[1] * tt[1]);
        double drdz = iso * (rho1 - rho0) / den;
        if (eps != 0.0) {
            double tp = tprstb(drdz, sp1, h->c.alphaT);
            Ftz[f] = 0.0 + tp * eps * dtdz / (drdz - epsln);
            Fsz[f] = 0.0 + tp * eps * dsdz / (drdz - epsln);
        }
        if (kvc != 0.0) {
            double tp = tprstb(-drdz, sp1, h->c.alphaT);
            Fti[f] = -tp * kvc * dtdz;
            Fsi[f] = -tp * kvc * dsdz;
        }
    }
    double dxc = h->dx * cos(h->y[j]), dyc = h->dy * cos(h->y[j]), dzk = h->dz * h->dfzT[k];
    double zero = (0.0 - 0.0);
    double mt = 0.0, ms = 0.0;
    if (h->vmix_temp == 1) {
        mt = zero / dxc + mt;
        mt = (0.0 * cos(h->yv[j]) - 0.0 * cos(h->yv[j - 1])) / dyc + mt;
        mt = (Ftz[1] - Ftz[0]) / dzk + mt;
        if (h->c.rho_mixing && xes == 0.0)
            mt = ((Fti[1] - Fti[0]) - (Fsi[1] - Fsi[0]) * lambda) / (2.0 * h->dz * h->dfzT[k]) + mt;
        else
            mt = (Fti[1] - Fti[0]) / dzk + mt;
    }
    if (h->vmix_salt == 1) {
        ms = zero / dxc + ms;
        ms = (0.0 * cos(h->yv[j]) - 0.0 * cos(h->yv[j - 1])) / dyc + ms;
        ms = (Fsz[1] - Fsz[0]) / dzk + ms;
        if (h->c.rho_mixing && xes == 0.0)
            ms = ((Fsi[1] - Fsi[0]) - (Fti[1] - Fti[0]) / lambda) / (2.0 * h->dz * h->dfzT[k]) + ms;
        else
            ms = (Fsi[1] - Fsi[0]) / dzk + ms;
    }
    *mixT = mt;
    *mixS = ms;
}

/* Neutral physics / Gent-McWilliams (MIXP, MKAP) are outside this
 * restatement. */
static int mixing_supported(const tho_model *h)
{
    return h->par[MIXP] * h->par[PE_H] == 0.0 && h->par[MKAP] * h->par[PE_H] == 0.0;
}

/* mix_imp.f:131-166 vmix_control (Mixing = 2): the T / S mixing switches
 * follow whether the T / S fields are nonzero (l2nrm > 1e-12,
 * matetc.F90:22-36); when they change and T is on, vmix_part rebuilds the
 * pattern from vmix_el_2 (mix_imp.f:934-1049) with the new switches (the
 * reference tests vmix_temp twice there, so a change that leaves T off keeps
 * the old pattern). */
static void vmix_control(tho_model *h, const double *un)
{
    double st = 0.0, ss = 0.0;
    for (int r = 0; r < h->ndim; r += 6) {
        st = un[r + TT - 1] * un[r + TT - 1] + st;
        ss = un[r + SS - 1] * un[r + SS - 1] + ss;
    }
    const int tt = sqrt(st) > 1.0e-12 ? 1 : 0, ts = sqrt(ss) > 1.0e-12 ? 1 : 0;
    if (h->vmix_temp != tt || h->vmix_salt != ts) {
        h->vmix_temp = tt;
        h->vmix_salt = ts;
        if (h->vmix_temp != 0) {
            h->vmix_part_done = 1;
            h->vmix_pat_salt = h->vmix_salt;
        }
    }
    h->vmix_fix = 1;
}

/* mix(x) for all rows; usol must have been called for x. */
static void vmix_fun(tho_model *h, double *mix)
{
    memset(mix, 0, sizeof(double) * h->ndim);
    for (int k = 1; k <= h->l; k++)
        for (int j = 1; j <= h->m; j++)
            for (int i = 1; i <= h->n; i++) {
                double mt, ms;
                mix_rows(h, i, j, k, -1, 0, 0.0, &mt, &ms);
                mix[ROW(h, i, j, k, TT)] = mt;
                mix[ROW(h, i, j, k, SS)] = ms;
            }
}

/* mix_imp.f:729-815 vmix_jac, forward differences (eps = 1e-8).  The entry
 * for row r and pattern column c is (mix_r(x + eps e_c) - mix_r(x)) / eps;
 * only the T,S entries of the cell's own column can be nonzero, the others
 * evaluate to exactly 0 and are not formed.  The pattern holds the T rows
 * against T (and against S when it has the salinity part) and, with the
 * salinity part, the S rows against T and S (vmix_el_1: all four). */
static void vmix_jac(tho_model *h)
{
    const double eps = 1.0e-08;
    double *An = h->An;
    for (int k = 1; k <= h->l; k++)
        for (int j = 1; j <= h->m; j++)
            for (int i = 1; i <= h->n; i++) {
                if (LM(h, i, j, k) != OCEAN) continue;
                double bT, bS;
                mix_rows(h, i, j, k, -1, 0, 0.0, &bT, &bS);
                for (int dk = -1; dk <= 1; dk++) {
                    int kk = k + dk, lm = LM(h, i, j, kk);
                    if (lm != OCEAN && lm != PERIO) continue;
                    int pos = (dk == 0) ? 5 : (dk < 0 ? 14 : 23);
                    for (int var = 0; var < 2; var++) {
                        if (var == 1 && !h->vmix_pat_salt) continue;   /* S columns not in the pattern */
                        double pT, pS;
                        mix_rows(h, i, j, k, kk, var, eps, &pT, &pS);
                        double fT = (pT - bT) / eps, fS = (pS - bS) / eps;
                        int col = var ? SS : TT, c = CELL(h, i, j, k);
                        AT(An, c, TT, col, pos) = AT(An, c, TT, col, pos) + fT;
                        if (h->vmix_pat_salt) AT(An, c, SS, col, pos) = AT(An, c, SS, col, pos) + fS;
                    }
                }
            }
}

/* ---------------------------------------------------------------- boundaries */
/* boundary.F90:2-393 on the atom array (Frc zeroing is folded into forcing). */
static void boundaries(tho_model *h)
{
    int n = h->n, m = h->m;
    double *An = h->An;
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= m; j++)
            for (int k = 1; k <= h->l; k++) {
                int c = CELL(h, i, j, k);
#define L(di, dj, dk) LM(h, i + (di), j + (dj), k + (dk))
#define ADDCOL(dst, src, col) for (int ii = 1; ii <= 6; ii++) AT(An, c, ii, col, dst) = AT(An, c, ii, col, dst) + AT(An, c, ii, col, src)
#define ZCOL(pos, col) for (int ii = 1; ii <= 6; ii++) AT(An, c, ii, col, pos) = 0.0
#define ZPOS(pos) for (int ii = 1; ii <= 6; ii++) for (int jj = 1; jj <= 6; jj++) AT(An, c, ii, jj, pos) = 0.0
#define ZROW(row) for (int q = 1; q <= 27; q++) for (int jj = 1; jj <= 6; jj++) AT(An, c, row, jj, q) = 0.0
                int southw = L(-1, -1, 0), west = L(-1, 0, 0), nwest = L(-1, 1, 0);
                int south = L(0, -1, 0), center = L(0, 0, 0), north = L(0, 1, 0);
                int southe = L(1, -1, 0), east = L(1, 0, 0), neast = L(1, 1, 0);
                int southwb = L(-1, -1, -1), westb = L(-1, 0, -1), nwestb = L(-1, 1, -1);
                int southb = L(0, -1, -1), bottom = L(0, 0, -1), northb = L(0, 1, -1);
                int southeb = L(1, -1, -1), eastb = L(1, 0, -1), neastb = L(1, 1, -1);
                int southwt = L(-1, -1, 1), westt = L(-1, 0, 1), nwestt = L(-1, 1, 1);
                int southt = L(0, -1, 1), top = L(0, 0, 1), northt = L(0, 1, 1);
                int southet = L(1, -1, 1), eastt = L(1, 0, 1), neastt = L(1, 1, 1);
                int southee = 0, easteast = 0, northee = 0, nnorthee = 0, nnwest = 0, nnorth = 0, nneast = 0;
                if (i < n) {
                    southee = L(2, -1, 0); easteast = L(2, 0, 0); northee = L(2, 1, 0);
                    if (j < m) nnorthee = L(2, 2, 0);
                }
                if (j < m) { nnwest = L(0, 2, 0); nnorth = L(0, 2, 0); nneast = L(0, 2, 0); }
                (void)nnwest; (void)nneast;
                if (center == OCEAN) {
                    if (bottom == LAND) {
                        if (westb == LAND && southwb == LAND && southb == LAND) { ADDCOL(1, 10, UU); ADDCOL(1, 10, VV); }
                        ZCOL(10, UU); ZCOL(10, VV);
                        if (westb == LAND && neastb == LAND && northb == LAND) { ADDCOL(2, 11, UU); ADDCOL(2, 11, VV); }
                        ZCOL(11, UU); ZCOL(11, VV);
                        if (eastb == LAND && southeb == LAND && southb == LAND) { ADDCOL(4, 13, UU); ADDCOL(4, 13, VV); }
                        ZCOL(13, UU); ZCOL(13, VV);
                        if (eastb == LAND && neastb == LAND && northb == LAND) { ADDCOL(5, 14, UU); ADDCOL(5, 14, VV); }
                        ADDCOL(5, 14, TT); ADDCOL(5, 14, SS);
                        ZPOS(14);
                    }
                    if (southwb == LAND) ZPOS(10);
                    if (westb == LAND) ZPOS(11);
                    if (nwestb == LAND) ZPOS(12);
                    if (southb == LAND) ZPOS(13);
                    if (northb == LAND) ZPOS(15);
                    if (southeb == LAND) ZPOS(16);
                    if (eastb == LAND) ZPOS(17);
                    if (neastb == LAND) ZPOS(18);
                    if (top == LAND) {
                        if (westt == LAND && southwt == LAND && southt == LAND) { ADDCOL(1, 19, UU); ADDCOL(1, 19, VV); }
                        ZCOL(19, UU); ZCOL(19, VV);
                        if (westt == LAND && nwestt == LAND && northt == LAND) { ADDCOL(2, 20, UU); ADDCOL(2, 20, VV); }
                        ZCOL(20, UU); ZCOL(20, VV);
                        if (eastt == LAND && southet == LAND && southt == LAND) { ADDCOL(4, 22, UU); ADDCOL(4, 22, VV); }
                        ZCOL(22, UU); ZCOL(22, VV);
                        if (eastt == LAND && neastt == LAND && northt == LAND) { ADDCOL(5, 23, UU); ADDCOL(5, 23, VV); }
                        ADDCOL(5, 23, TT); ADDCOL(5, 23, SS);
                        ZPOS(23);
                        ZROW(WW);
                        for (int ii = 1; ii <= 6; ii++) {
                            AT(An, c, ii, WW, 5) = 1.0e-10; AT(An, c, ii, WW, 6) = 1.0e-10;
                            AT(An, c, ii, WW, 8) = 1.0e-10; AT(An, c, ii, WW, 9) = 1.0e-10;
                        }
                        AT(An, c, WW, WW, 5) = 1.0;
                    }
                    if (southwt == LAND) ZPOS(19);
                    if (westt == LAND) ZPOS(20);
                    if (nwestt == LAND) ZPOS(21);
                    if (southt == LAND) ZPOS(22);
                    if (northt == LAND) ZPOS(24);
                    if (southet == LAND) ZPOS(25);
                    if (eastt == LAND) ZPOS(26);
                    if (neastt == LAND) ZPOS(27);
                    if (southw == LAND) { ZCOL(1, UU); ZCOL(1, VV); }
                    if (west == LAND) {
                        ADDCOL(5, 2, TT); ADDCOL(5, 2, SS);
                        ZPOS(2); ZCOL(1, UU); ZCOL(1, VV);
                    }
                    if (nwest == LAND) {
                        ZCOL(2, UU); ZCOL(2, VV); ZCOL(3, UU); ZCOL(3, VV);
                    } else if (j < m) {
                        if (nnwest == LAND) { ZCOL(3, UU); ZCOL(3, VV); }
                    }
                    if (south == LAND) {
                        ADDCOL(5, 4, SS); ADDCOL(5, 4, TT);
                        ZPOS(4); ZCOL(1, UU); ZCOL(1, VV);
                    }
                    if (north == LAND) {
                        ZCOL(2, UU); ZCOL(2, VV);
                        AT(An, c, PP, UU, 2) = 0.0; AT(An, c, PP, VV, 2) = 0.0;
                        AT(An, c, PP, UU, 5) = 0.0; AT(An, c, PP, VV, 5) = 0.0;
                        ZROW(VV); ZCOL(5, VV); AT(An, c, VV, VV, 5) = 1.0;
                        ZROW(UU); ZCOL(5, UU); AT(An, c, UU, UU, 5) = 1.0;
                        ADDCOL(5, 6, SS); ADDCOL(5, 6, TT);
                        ZPOS(6);
                    } else if (j < m) {
                        if (nnorth == LAND) { ZCOL(3, UU); ZCOL(3, VV); ZCOL(6, UU); ZCOL(6, VV); }
                    }
                    if (southe == LAND) {
                        ZCOL(4, UU); ZCOL(4, VV); ZCOL(7, UU); ZCOL(7, VV);
                    } else if (i < n) {
                        if (southee == LAND) { ZCOL(7, UU); ZCOL(7, VV); }
                    }
                    if (east == LAND) {
                        ZCOL(4, UU); ZCOL(4, VV);
                        AT(An, c, PP, UU, 4) = 0.0; AT(An, c, PP, VV, 4) = 0.0;
                        AT(An, c, PP, UU, 5) = 0.0; AT(An, c, PP, VV, 5) = 0.0;
                        ZROW(UU); ZCOL(5, UU); AT(An, c, UU, UU, 5) = 1.0;
                        ZROW(VV); ZCOL(5, VV); AT(An, c, VV, VV, 5) = 1.0;
                        ADDCOL(5, 8, SS); ADDCOL(5, 8, TT);
                        ZPOS(8); ZCOL(7, UU); ZCOL(7, VV);
                    } else if (i < n) {
                        if (easteast == LAND) { ZCOL(7, UU); ZCOL(7, VV); ZCOL(8, UU); ZCOL(8, VV); }
                    }
                    if (neast == LAND) {
                        ZROW(UU); ZCOL(5, UU); AT(An, c, UU, UU, 5) = 1.0;
                        ZROW(VV); ZCOL(5, VV); AT(An, c, VV, VV, 5) = 1.0;
                        ZCOL(7, UU); ZCOL(7, VV);
                    } else if (i < n || j < m) {
                        if (i < n) {
                            if (northee == LAND) { ZCOL(8, UU); ZCOL(8, VV); ZCOL(9, UU); ZCOL(9, VV); }
                            else if (j < m) {
                                if (nnorthee == LAND) { ZCOL(9, UU); ZCOL(9, VV); }
                            }
                        }
                        if (j < m) {
                            if (nneast == LAND) { ZCOL(6, UU); ZCOL(6, VV); ZCOL(9, UU); ZCOL(9, VV); }
                        }
                    }
                } else {
                    for (int q = 1; q <= 27; q++) ZPOS(q);
                    for (int ii = 1; ii <= 6; ii++) AT(An, c, ii, ii, 5) = 1.0;
                }
#undef L
#undef ADDCOL
#undef ZCOL
#undef ZPOS
#undef ZROW
            }
}

/* ---------------------------------------------------------------- assembly */
/* assemble.F90:57-177 fillcolA + shift: atoms -> 1-based CSR (Fortran order,
 * |a| <= 1e-10 dropped).  Returns -1 if a kept entry points outside the grid. */
static int fillcolA(tho_model *h)
{
    int v = 1, row = 1, bad = 0;
    for (int k = 1; k <= h->l; k++)
        for (int j = 1; j <= h->m; j++)
            for (int i = 1; i <= h->n; i++) {
                int c = CELL(h, i, j, k);
                for (int ii = 1; ii <= 6; ii++) {
                    h->begA[row - 1] = v;
                    for (int kk = 1; kk <= 27; kk++)
                        for (int jj = 1; jj <= 6; jj++) {
                            double a = AT(h->An, c, ii, jj, kk);
                            if (fabs(a) > 1.0e-10) {
                                int lev = (kk - 1) / 9, rem = (kk - 1) % 9;
                                int k2 = k + (lev == 0 ? 0 : (lev == 1 ? -1 : 1));
                                int i2 = i - 1 + rem / 3, j2 = j - 1 + rem % 3;
                                if (h->periodic) {
                                    if (i2 == 0) i2 = h->n; else if (i2 == h->n + 1) i2 = 1;
                                }
                                if (i2 < 1 || i2 > h->n || j2 < 1 || j2 > h->m || k2 < 1 || k2 > h->l) bad = 1;
                                h->coA[v - 1] = a;
                                h->jcoA[v - 1] = 6 * ((k2 - 1) * h->n * h->m + h->n * (j2 - 1) + i2 - 1) + jj;
                                v++;
                            }
                        }
                    row++;
                }
            }
    h->begA[h->ndim] = v;
    return bad ? -1 : 0;
}

/* ---------------------------------------------------------------- graph */
static void gentry(const tho_model *h, int *idx, int *pos, int i, int j, int k, int xx)
{   /* THCM.C:2583-2598, 0-based i,j,k */
    int N = h->n, ii = i;
    if (h->periodic) ii = ((i % N) + N) % N;
    if (ii >= 0 && j >= 0 && k >= 0 && ii < N && j < h->m && k < h->l)
        idx[(*pos)++] = 6 * ((k * h->m + j) * N + ii) + xx - 1;
}

static int cmpi(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int sortu(int *v, int len)
{
    qsort(v, len, sizeof(int), cmpi);
    int w = 0;
    for (int r = 0; r < len; r++)
        if (w == 0 || v[w - 1] != v[r]) v[w++] = v[r];
    return w;
}

/* THCM.C:2300-2580 CreateMaximalGraph; neighbour lists per equation as
 * (di, dj, dk, var) tuples. */
static const signed char G_U[24][4] = {
    {0,0,0,1},{-1,0,0,1},{1,0,0,1},{0,-1,0,1},{0,1,0,1},{0,0,-1,1},{0,0,1,1},
    {0,0,0,2},{-1,0,0,2},{1,0,0,2},{0,-1,0,2},{0,1,0,2},
    {0,0,0,3},{1,0,0,3},{1,1,0,3},{0,1,0,3},{0,0,-1,3},{1,0,-1,3},{1,1,-1,3},{0,1,-1,3},
    {0,0,0,4},{1,0,0,4},{0,1,0,4},{1,1,0,4}};
static const signed char G_V[22][4] = {
    {0,0,0,2},{-1,0,0,2},{1,0,0,2},{0,-1,0,2},{0,1,0,2},{0,0,-1,2},{0,0,1,2},
    {0,0,0,1},{-1,0,0,1},{1,0,0,1},
    {0,0,0,3},{1,0,0,3},{1,1,0,3},{0,1,0,3},{0,0,-1,3},{1,0,-1,3},{1,1,-1,3},{0,1,-1,3},
    {0,0,0,4},{1,0,0,4},{0,1,0,4},{1,1,0,4}};
static const signed char G_W[7][4] = {
    {0,0,0,3},{0,0,0,4},{0,0,1,4},{0,0,0,5},{0,0,1,5},{0,0,0,6},{0,0,1,6}};
static const signed char G_P[11][4] = {
    {0,0,0,4},{0,0,0,1},{-1,0,0,1},{0,-1,0,1},{-1,-1,0,1},
    {0,0,0,2},{-1,0,0,2},{0,-1,0,2},{-1,-1,0,2},{0,0,0,3},{0,0,-1,3}};
static const signed char G_T[20][4] = {
    {0,0,0,5},{-1,0,0,5},{1,0,0,5},{0,-1,0,5},{0,1,0,5},{0,0,-1,5},{0,0,1,5},
    {0,0,0,1},{-1,0,0,1},{-1,-1,0,1},{0,-1,0,1},
    {0,0,0,2},{-1,0,0,2},{-1,-1,0,2},{0,-1,0,2},{0,0,0,3},{0,0,-1,3},
    {0,0,0,6},{0,0,-1,6},{0,0,1,6}};

static void build_graph(tho_model *h)
{
    int N = h->n, M = h->m, L = h->l, ndim = h->ndim;
    h->grow = calloc(ndim + 1, sizeof(int));
    h->gcol = malloc(sizeof(int) * ((size_t)ndim * 24 + (size_t)N * M * L + 16));
    int nz = 0, idx[24];
    const signed char (*lists[6])[4] = {G_U, G_V, G_W, G_P, G_T, G_T};
    const int lens[6] = {24, 22, 7, 11, 20, 20};
    for (int k = 0; k < L; k++)
        for (int j = 0; j < M; j++)
            for (int i = 0; i < N; i++)
                for (int xx = 1; xx <= 6; xx++) {
                    int row = 6 * ((k * M + j) * N + i) + xx - 1, pos = 0;
                    if (xx == SS && h->c.sres == 0 && row == h->rowintcon) {
                        int p2 = 0;
                        for (int a = 0; a < N; a++)
                            for (int b = 0; b < M; b++)
                                for (int q = 0; q < L; q++)
                                    h->gcol[nz + p2++] = 6 * ((q * M + b) * N + a) + SS - 1;
                        p2 = sortu(h->gcol + nz, p2);
                        h->grow[row + 1] = p2; nz += p2;
                        continue;
                    }
                    for (int e = 0; e < lens[xx - 1]; e++) {
                        const signed char *g = lists[xx - 1][e];
                        int var = g[3];
                        if (xx == SS) var = (var == 5) ? 6 : (var == 6 ? 5 : var);
                        gentry(h, idx, &pos, i + g[0], j + g[1], k + g[2], var);
                    }
                    pos = sortu(idx, pos);
                    h->grow[row + 1] = pos;
                    memcpy(h->gcol + nz, idx, sizeof(int) * pos);
                    nz += pos;
                }
    for (int r = 0; r < ndim; r++) h->grow[r + 1] += h->grow[r];
    h->gnnz = nz;
}

/* ---------------------------------------------------------------- API */
tho_model *tho_create(const tho_config *cfg, const int *landm, const double *taux,
                      const double *tauy, const double *tatm, const double *emip,
                      const double *spert)
{
    if (cfg->coupled_t || cfg->coupled_s) return NULL;
    tho_model *h = calloc(1, sizeof(tho_model));
    h->c = *cfg;
    int n = cfg->n, m = cfg->m, l = cfg->l;
    h->n = n; h->m = m; h->l = l;
    h->ncell = n * m * l;
    h->ndim = 6 * h->ncell;
    h->periodic = cfg->periodic;
    h->xmin = cfg->xmin_deg * pi_ / 180.0; h->xmax = cfg->xmax_deg * pi_ / 180.0;
    h->ymin = cfg->ymin_deg * pi_ / 180.0; h->ymax = cfg->ymax_deg * pi_ / 180.0;
    h->x = calloc(n + 1, sizeof(double));
    h->y = calloc(m + 2, sizeof(double));
    h->z = calloc(l + 1, sizeof(double));
    h->xu = calloc(n + 1, sizeof(double));
    h->yv = calloc(m + 1, sizeof(double));
    h->zw = calloc(l + 1, sizeof(double));
    h->dfzT = calloc(l + 1, sizeof(double));
    h->dfzW = calloc(l + 1, sizeof(double));
    /* landmask handling of usrc.F90:71-101 */
    size_t nl = (size_t)(n + 2) * (m + 2) * (l + 2);
    h->landm = malloc(sizeof(int) * nl);
    memcpy(h->landm, landm, sizeof(int) * nl);
    for (int k = 0; k <= l + 1; k++)
        for (int j = 0; j <= m + 1; j++)
            for (int i = 0; i <= n + 1; i++) {
                int *q = &LM(h, i, j, k);
                if (!h->periodic && *q == PERIO) *q = OCEAN;
                if ((!h->periodic && (i == 0 || i == n + 1)) || j == 0 || j == m + 1 || k == 0 || k == l + 1)
                    *q = LAND;
            }
    size_t nm = (size_t)n * m;
    double **f[6] = {&h->taux, &h->tauy, &h->tatm, &h->emip, &h->spert, &h->aemip};
    const double *src[6] = {taux, tauy, tatm, emip, spert, NULL};
    for (int q = 0; q < 6; q++) {
        *f[q] = calloc(nm, sizeof(double));
        if (src[q]) memcpy(*f[q], src[q], sizeof(double) * nm);
    }
    h->Frc = calloc(h->ndim, sizeof(double));
    h->Al = calloc((size_t)36 * 27 * h->ncell, sizeof(double));
    h->An = calloc((size_t)36 * 27 * h->ncell, sizeof(double));
    h->u = calloc((size_t)(n + 1) * (m + 1) * (l + 2), sizeof(double));
    h->v = calloc((size_t)(n + 1) * (m + 1) * (l + 2), sizeof(double));
    h->w = calloc((size_t)(n + 2) * (m + 2) * (l + 1), sizeof(double));
    h->p = calloc(nl, sizeof(double));
    h->t = calloc(nl, sizeof(double));
    h->s = calloc(nl, sizeof(double));
    h->begA = calloc(h->ndim + 1, sizeof(int));
    h->jcoA = calloc((size_t)h->ndim * 163, sizeof(int));
    h->coA = calloc((size_t)h->ndim * 163, sizeof(double));
    make_grid(h);
    stpnt(h);
    /* mix_imp.f:69-129 vmix_init */
    h->vmix_flag = (cfg->vmix == 0) ? 0 : (cfg->vmix == 1 ? 1 : (cfg->vmix == 2 ? 2 : -1));
    h->vmix_fix = (cfg->vmix == 2) ? 0 : 1;
    h->vmix_temp = h->vmix_salt = (h->vmix_flag == 1) ? 1 : 0;
    h->vmix_part_done = h->vmix_pat_salt = (h->vmix_flag == 1) ? 1 : 0;
    forcing(h);
    lin(h);
    h->rowintcon = -1;
    if (cfg->sres == 0) {
        int nic = cfg->nic < 0 ? n - 1 : cfg->nic, mic = cfg->mic < 0 ? m - 1 : cfg->mic;
        h->rowintcon = 6 * (((l - 1) * m + mic) * n + nic) + SS - 1;
    }
    /* thcm_utils.F90:285-312 intcond_scaling */
    h->coef = calloc(h->ndim, sizeof(double));
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                if (LM(h, i, j, k) == OCEAN) h->coef[ROW(h, i, j, k, SS)] = cos(h->y[j]) * h->dfzT[k];
    build_graph(h);
    return h;
}

void tho_destroy(tho_model *h)
{
    if (!h) return;
    void *ptrs[] = {h->x, h->y, h->z, h->xu, h->yv, h->zw, h->dfzT, h->dfzW, h->landm,
                    h->taux, h->tauy, h->tatm, h->emip, h->spert, h->aemip, h->Frc, h->Al,
                    h->An, h->coef, h->grow, h->gcol, h->u, h->v, h->w, h->p, h->t, h->s,
                    h->begA, h->jcoA, h->coA};
    for (size_t q = 0; q < sizeof ptrs / sizeof ptrs[0]; q++) free(ptrs[q]);
    free(h);
}

int tho_set_par(tho_model *h, int idx, double v)   /* usrc.F90:143-162 setparcs */
{
    if (idx < 1 || idx > 30) return -1;
    h->par[idx] = v;
    forcing(h);
    lin(h);
    return 0;
}

double tho_get_par(const tho_model *h, int idx) { return (idx >= 1 && idx <= 30) ? h->par[idx] : 0.0; }
int tho_ndim(const tho_model *h) { return h->ndim; }
int tho_rowintcon(const tho_model *h) { return h->rowintcon; }
int tho_graph_nnz(const tho_model *h) { return h->gnnz; }
void tho_graph(const tho_model *h, int *rowptr, int *col)
{
    memcpy(rowptr, h->grow, sizeof(int) * (h->ndim + 1));
    memcpy(col, h->gcol, sizeof(int) * h->gnnz);
}
void tho_intcond_coef(const tho_model *h, double *out) { memcpy(out, h->coef, sizeof(double) * h->ndim); }
void tho_landm(const tho_model *h, int *out)
{
    memcpy(out, h->landm, sizeof(int) * (size_t)(h->n + 2) * (h->m + 2) * (h->l + 2));
}
void tho_forcing(const tho_model *h, double *out) { memcpy(out, h->Frc, sizeof(double) * h->ndim); }

int tho_mix(tho_model *h, const double *x, double *mix)
{
    if (!mixing_supported(h)) return -2;
    usol(h, x);
    if (h->vmix_flag == 2) h->vmix_fix = 0;
    if (h->vmix_flag >= 1) {
        if (h->vmix_fix == 0 && h->vmix_flag >= 2) vmix_control(h, x);
        if ((h->vmix_temp == 1 || h->vmix_salt == 1) && h->vmix_part_done) vmix_fun(h, mix);
        else memset(mix, 0, sizeof(double) * h->ndim);
    } else {
        memset(mix, 0, sizeof(double) * h->ndim);
    }
    return 0;
}

/* usrc.F90:523-603 rhs + THCM.C:1002-1027 */
int tho_rhs(tho_model *h, const double *x, double intcorr, double *F)
{
    if (!mixing_supported(h)) return -2;
    int nd = h->ndim;
    double *mix = calloc(nd, sizeof(double));
    memcpy(h->An, h->Al, sizeof(double) * (size_t)36 * 27 * h->ncell);
    usol(h, x);
    nlin(h, 0);
    boundaries(h);
    int bad = fillcolA(h);
    if (h->vmix_flag == 2) h->vmix_fix = 0;   /* THCM::fixMixing(0) of Ocean::computeRHS (Ocean.C:1281) */
    if (h->vmix_flag >= 1) {
        if (h->vmix_fix == 0 && h->vmix_flag >= 2) vmix_control(h, x);
        if ((h->vmix_temp == 1 || h->vmix_salt == 1) && h->vmix_part_done) vmix_fun(h, mix);
    }
    for (int r = 0; r < nd; r++) {
        double au = 0.0;
        for (int q = h->begA[r] - 1; q < h->begA[r + 1] - 1; q++) {
            int col = h->jcoA[q] - 1;
            double xv = (col >= 0 && col < nd) ? x[col] : 0.0;
            au = h->coA[q] * xv + au;
        }
        double b = -au - mix[r] + h->Frc[r];
        int cell = r / 6, i = cell % h->n + 1, j = (cell / h->n) % h->m + 1, k = cell / (h->n * h->m) + 1;
        b = b * (1 - LM(h, i, j, k));
        F[r] = -b;
    }
    if (h->c.sres == 0) {
        double dot = 0.0;
        for (int r = 0; r < nd; r++) dot += h->coef[r] * x[r];
        F[h->rowintcon] = h->c.int_sign * (dot - intcorr);
    }
    free(mix);
    return bad;
}

/* usrc.F90:449-521 matrix + THCM.C:1043-1169 (fixed graph, integral row) */
int tho_jacobian(tho_model *h, const double *x, double *val, double *diagB)
{
    if (!mixing_supported(h)) return -2;
    int nd = h->ndim, n = h->n, m = h->m, l = h->l;
    memcpy(h->An, h->Al, sizeof(double) * (size_t)36 * 27 * h->ncell);
    /* assemble.F90:18-54 fillcolB */
    memset(diagB, 0, sizeof(double) * nd);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                if (LM(h, i, j, k) == OCEAN) {
                    if (LM(h, i + 1, j, k) != LAND) diagB[ROW(h, i, j, k, UU)] = -h->par[ROSB];
                    if (LM(h, i, j + 1, k) != LAND) diagB[ROW(h, i, j, k, VV)] = -h->par[ROSB];
                    diagB[ROW(h, i, j, k, TT)] = -1.0;
                    diagB[ROW(h, i, j, k, SS)] = -1.0;
                }
    usol(h, x);
    nlin(h, 1);
    if (h->vmix_flag == 2) h->vmix_fix = 0;   /* THCM::fixMixing(0) of Ocean::computeJacobian (Ocean.C:1302) */
    if (h->vmix_flag >= 1) {
        if (h->vmix_fix == 0 && h->vmix_flag >= 2) vmix_control(h, x);
        /* A pattern built with a zero S field leaves the S columns out; the
         * reference perturbs them with whichever T column group its DSM
         * colouring puts them in (mix_sup.F90), and the T-row entries then
         * depend on that colouring, which this restatement does not follow. */
        if (h->vmix_temp == 1 && h->vmix_part_done && !h->vmix_pat_salt) return -2;
        if ((h->vmix_temp == 1 || h->vmix_salt == 1) && h->vmix_part_done) vmix_jac(h);
    }
    boundaries(h);
    int bad = fillcolA(h);
    memset(val, 0, sizeof(double) * h->gnnz);
    int dropped = 0;
    for (int r = 0; r < nd; r++) {
        if (h->c.sres == 0 && r == h->rowintcon) continue;
        for (int q = h->begA[r] - 1; q < h->begA[r + 1] - 1; q++) {
            int col = h->jcoA[q] - 1, lo = h->grow[r], hi = h->grow[r + 1] - 1, hit = -1;
            while (lo <= hi) {
                int mid = (lo + hi) / 2;
                if (h->gcol[mid] == col) { hit = mid; break; }
                if (h->gcol[mid] < col) lo = mid + 1; else hi = mid - 1;
            }
            if (hit >= 0) val[hit] = h->coA[q]; else dropped++;
        }
    }
    if (h->c.sres == 0) {
        diagB[h->rowintcon] = 0.0;
        for (int q = h->grow[h->rowintcon]; q < h->grow[h->rowintcon + 1]; q++)
            val[q] = h->c.int_sign * h->coef[h->gcol[q]];
    }
    if (bad) return -1;
    return dropped;
}
