/*
 * Test infrastructure only (checker, never the product): CPU restatement of the
 * reference THCM assembly of F(x) and J(x) on one rank.  See thcm_oracle.c.
 */
#ifndef THCM_ORACLE_H
#define THCM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order as the reference harness configuration. */
typedef struct {
    int n, m, l;                 /* grid cells in x, y, z                      */
    double xmin_deg, xmax_deg;   /* domain bounds in degrees                   */
    double ymin_deg, ymax_deg;
    int periodic;                /* periodic in x                              */
    double hdim, qz;             /* depth [m], vertical stretching             */
    double alphaT, alphaS;       /* linear EOS coefficients                    */
    int ih, vmix, tap, rho_mixing, coriolis_on;
    int tres, sres, iza, ite, its, coupled_t, coupled_s, forcing_type;
    int int_sign, nic, mic;      /* integral row sign and location (-1: last)  */
} tho_config;

typedef struct tho_model tho_model;

/* landm: (n+2)(m+2)(l+2) ints, i fastest, values OCEAN=0 LAND=1 PERIO=3.
 * taux..spert: n*m surface fields (i fastest) or NULL for zero. */
tho_model *tho_create(const tho_config *cfg, const int *landm, const double *taux,
                      const double *tauy, const double *tatm, const double *emip,
                      const double *spert);
void tho_destroy(tho_model *h);
int tho_set_par(tho_model *h, int idx, double v);       /* 1..30, recomputes forcing+lin */
double tho_get_par(const tho_model *h, int idx);
int tho_ndim(const tho_model *h);
int tho_rowintcon(const tho_model *h);
int tho_graph_nnz(const tho_model *h);
void tho_graph(const tho_model *h, int *rowptr, int *col);
void tho_intcond_coef(const tho_model *h, double *out);
void tho_landm(const tho_model *h, int *out);          /* mask as used (after init) */
void tho_forcing(const tho_model *h, double *out);     /* Frc after boundary zeroing */
/* F(x) as Ocean::computeRHS returns it; returns 0 or a negative error code. */
int tho_rhs(tho_model *h, const double *x, double intcorr, double *F);
/* J(x) on the fixed graph (values only; pattern from tho_graph) and diag(B). */
int tho_jacobian(tho_model *h, const double *x, double *val, double *diagB);
/* mix(x) of the tracer equations (vmix_fun), length ndim. */
int tho_mix(tho_model *h, const double *x, double *mix);

#ifdef __cplusplus
}
#endif
#endif
