/*
 * iemic_amd -- MI355X (gfx950) implementation of the I-EMIC implicit solve
 * path: THCM residual F(x) and Jacobian J(x) assembly, matrix/mass-matrix
 * application, preconditioner and right-preconditioned FGMRES.
 *
 * C ABI: plain pointers and sizes only.  One context = one THCM ocean model
 * (the reference keeps THCM as a process-wide singleton, THCM.H:76-84; here
 * any number of contexts may coexist, one model instance per context).
 *
 * Conventions shared with the reference:
 *  - unknown ordering: row = 6*((k*M + j)*N + i) + XX, XX = 0..5 for
 *    u,v,w,p,T,S, i fastest (THCMdefs.H FIND_ROW2, 0-based here);
 *  - F is the negated THCM right-hand side with the salinity integral
 *    condition in row `rowintcon` (THCM.C:1011-1027);
 *  - J lives on the fixed maximal graph of THCM::CreateMaximalGraph
 *    (THCM.C:2300-2580); iemic_jacobian_csr exports it in that row/column
 *    order;
 *  - land mask: (N+2)(M+2)(L+2) ints, i fastest, OCEAN=0 LAND=1 PERIO=3, as
 *    handed to the Fortran init_ (usrc.F90:6-31).
 * Functions ending in _dev take device pointers (hipMalloc'ed, resident in
 * HBM); the others take host pointers and copy.
 * Every function returns 0 on success or a negative IEMIC_E* code; the
 * message of the last error is available from iemic_last_error().
 */
#ifndef IEMIC_AMD_H
#define IEMIC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    IEMIC_OK = 0,
    IEMIC_EINVAL = -1,     /* bad argument or configuration            */
    IEMIC_EUNSUPPORTED = -2, /* option outside this implementation     */
    IEMIC_EHIP = -3,       /* HIP runtime error                         */
    IEMIC_ENOMEM = -4,     /* device allocation failed                  */
    IEMIC_ESTATE = -5,     /* call order (e.g. no Jacobian yet)         */
    IEMIC_ENOCONV = -6,    /* solver did not reach its tolerance        */
    IEMIC_EIO = -7         /* a file could not be read or written       */
};

/* THCM initialisation parameters (THCM::THCM, THCM.C:180-260, and the
 * Fortran init_ arguments, usrc.F90:6-31).  Degrees for the bounds. */
typedef struct {
    int n, m, l;
    double xmin_deg, xmax_deg, ymin_deg, ymax_deg;
    int periodic;
    double hdim, qz;
    double alphaT, alphaS;
    int ih, vmix, tap, rho_mixing, coriolis_on;
    int tres, sres, iza, ite, its, coupled_t, coupled_s, forcing_type;
    int int_sign, nic, mic;
} iemic_config;

typedef struct iemic_ctx iemic_ctx;

/* Preconditioners for the slot of Ocean::buildPreconditioner (Ocean.C:1377-1391),
 * where the reference plugs in TRIOS::BlockPreconditioner. */
enum {
    IEMIC_PREC_ROW_LU = 0,      /* block-tridiagonal LU over zonal rows (default) */
    IEMIC_PREC_COLUMN_ILU = 1   /* water-column block ILU(0)                      */
};

typedef struct {
    int kind;             /* IEMIC_PREC_*                                           */
    int blocks;           /* zonal rows / active water columns                      */
    int max_block;        /* largest dense block (unknowns)                         */
    int pins;             /* pressure basins (continuity rows replaced by p = 0)   */
    int levels;           /* dependency levels of the sweeps                        */
    double factor_bytes;  /* FP64 factor storage                                    */
} iemic_precond_info;

/* Solver controls (Ocean::getDefaultInitParameters, Ocean.C:2264-2301, and
 * GMRESSolver::setParameters, GMRESSolver.H:52-68). */
typedef struct {
    double tol;          /* relative residual ||b - J x|| / ||b||            */
    int restart;         /* Krylov space size (Belos "Num Blocks")           */
    int maxit;           /* total iteration cap                              */
    int precision;       /* 0: FP64 throughout; 1: preconditioner factors and  */
                         /*    sweeps in FP32 inside the FP64 FGMRES (basis,   */
                         /*    SpMV and residuals stay FP64)                   */
    int verbose;
} iemic_solver_params;

typedef struct {
    int iterations;       /* total FGMRES iterations                          */
    int restarts;         /* FGMRES cycles                                    */
    double rel_residual;  /* true FP64 ||b - J x|| / ||b||                    */
    double t_precon_setup;/* seconds                                          */
    double t_solve;       /* seconds                                          */
    int converged;
} iemic_solve_info;

const char *iemic_last_error(void);
int iemic_device_count(int *count);

/* ---- lifecycle: THCM::THCM / Fortran init_ (usrc.F90:6-139) ------------ */
/* taux..spert: N*M surface fields, i fastest (NULL = zero), the arrays the
 * reference passes to init_ / m_inserts (inserts.F90:12-265). */
int iemic_create(const iemic_config *cfg, const int *landm, const double *taux,
                 const double *tauy, const double *tatm, const double *emip,
                 const double *spert, iemic_ctx **out);
int iemic_destroy(iemic_ctx *ctx);

/* ---- parameters: setparcs_/getparcs_ (usrc.F90:143-180), 1..30 -------- */
int iemic_set_par(iemic_ctx *ctx, int idx, double value);
int iemic_get_par(const iemic_ctx *ctx, int idx, double *value);

/* ---- sizes and fixed structure ----------------------------------------- */
int iemic_ndim(const iemic_ctx *ctx, int *ndim);
int iemic_graph_nnz(const iemic_ctx *ctx, int *nnz);         /* CreateMaximalGraph */
int iemic_rowintcon(const iemic_ctx *ctx, int *row);         /* THCM.C:695, -1 if SRES=1 */
int iemic_intcond_coef(const iemic_ctx *ctx, double *coef);  /* getIntCondCoeff, THCM.C:2608 */
int iemic_forcing(const iemic_ctx *ctx, double *frc);        /* THCM::getForcing, THCM.C:944 */
/* THCM::setIntCondCorrection (THCM.C:2079-2097): intCorrection = coef . x */
int iemic_set_intcond_correction(iemic_ctx *ctx, const double *x);
int iemic_set_intcond_correction_value(iemic_ctx *ctx, double value);

/* ---- Ocean::computeRHS (Ocean.C:1277-1284) ------------------------------ */
int iemic_rhs(iemic_ctx *ctx, const double *x, double *F);
int iemic_rhs_dev(iemic_ctx *ctx, const double *d_x, double *d_F);

/* F(x) as iemic_rhs_dev, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * integral row's reduction partials (one per 256 unknowns) live in a
 * library-owned device array of IEMIC_NOALLOC_MAX_NDIM / 256 doubles
 * (128 KiB on each GPU the library runs on, held until the process ends)
 * instead of a buffer of the context; ndim above IEMIC_NOALLOC_MAX_NDIM is
 * rejected with IEMIC_EINVAL.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other contexts, streams or threads) race.
 * F equals iemic_rhs_dev's in every row; the integral row sums its terms in
 * another order.  The context itself is allocated by iemic_create. */
#define IEMIC_NOALLOC_MAX_NDIM 4194304
int iemic_rhs_dev_noalloc(iemic_ctx *ctx, const double *d_x, double *d_F);

/* ---- Ocean::computeJacobian (Ocean.C:1297-1309); J stays on the device -- */
int iemic_jacobian(iemic_ctx *ctx, const double *x);
int iemic_jacobian_dev(iemic_ctx *ctx, const double *d_x);
/* export J in the reference's CSR (row order, sorted columns of the fixed
 * graph) and diag(B) of THCM::evaluate (THCM.C:1157-1169); arrays of size
 * ndim+1, graph_nnz, graph_nnz, ndim. */
int iemic_jacobian_csr(iemic_ctx *ctx, int *rowptr, int *col, double *val, double *diagB);

/* ---- the Fortran interface THCM.C binds (THCM.C:49-176; usrc.F90, mat.F90)
 * A THCM.C that links this library in place of the THCM Fortran calls these
 * where it calls m_mat::get_array_sizes, m_mat::set_pointers, rhs_ and
 * matrix_ (THCM.C:617-638, 1002, 1056); host arrays, 1-based CSR. */
/* m_mat::get_array_sizes (mat.F90:58-70): nrows = ndim, nnz = ndim (6 27 + 1),
 * the capacity THCM.C allocates (THCM.C:617-630). */
int iemic_thcm_get_array_sizes(const iemic_ctx *ctx, int *nrows, int *nnz);
/* m_mat::set_pointers (mat.F90:75-104): caller-allocated buffers that
 * iemic_thcm_matrix fills: begA (nrows + 1), jcoA and coA (nnz), coB (nrows).
 * The stochastic-forcing matrix begF / jcoF / coF (get_stochastic_forcing,
 * forcing.F90:236-280) is accepted and not written (may be NULL). */
int iemic_thcm_set_pointers(iemic_ctx *ctx, int nrows, int nnz, int *begA, int *jcoA, double *coA, double *coB,
                            int *begF, int *jcoF, double *coF);
/* rhs_ (usrc.F90:523-603): B = -A(u) u - mix + Frc with land rows zeroed, the
 * Fortran's own sign and no integral condition (iemic_rhs returns F = -B with
 * it, THCM.C:1011-1027).  un and B: ndim doubles. */
int iemic_thcm_rhs(iemic_ctx *ctx, const double *un, double *B);
/* matrix_ (usrc.F90:449-521): J(un) into the set_pointers buffers as
 * fillcolA writes it (assemble.F90:57-139): rows in order, each row's entries
 * by stencil position then unknown, |a| > 1e-10 kept, 1-based begA / jcoA;
 * the integral-condition row holds its S equation.  coB = fillcolB
 * (assemble.F90:18-54).  J also becomes the context's Jacobian, as
 * iemic_jacobian would make it. */
int iemic_thcm_matrix(iemic_ctx *ctx, const double *un);

/* ---- Ocean::applyMatrix / applyMassMat (Ocean.C:1369-1374, 1448-1457) -- */
int iemic_apply_matrix(iemic_ctx *ctx, const double *v, double *out);
int iemic_apply_matrix_dev(iemic_ctx *ctx, const double *d_v, double *d_out);
/* transpose product out = J^T v (no counterpart in the reference: its
 * TRIOS preconditioner rejects SetUseTranspose, TRIOS_BlockPreconditioner.H:122;
 * used by the adjoint solve). */
int iemic_apply_matrix_transpose_dev(iemic_ctx *ctx, const double *d_v, double *d_out);
int iemic_apply_mass(iemic_ctx *ctx, const double *v, double *out);

/* ---- Ocean::pressureProjection (Ocean.C:2180-2235) ---------------------- */
/* v -= (v.s1) s1 + (v.s2) s2 in place, s1 / s2 the normalised checkerboard
 * pressure modes (wet cells with i + j even / odd, THCM.C:2802-2841), the
 * null vectors of J on basins whose wet cells connect through the B-grid
 * diagonals. */
int iemic_pressure_projection(iemic_ctx *ctx, double *v);
int iemic_pressure_projection_dev(iemic_ctx *ctx, double *d_v);

/* ---- Ocean::buildPreconditioner / applyPrecon (Ocean.C:1377-1416) ------ */
/* launch shape the row-LU apply tuned for nrhs right-hand sides (nrhs = 1:
 * the single apply; negative: FP32 factors): the number of blocks its dense
 * products aim at, chosen by timing candidates on the first apply; 0 if
 * not tuned yet (IEMIC_NO_TUNE in the environment keeps the default). */
int iemic_apply_tuning(iemic_ctx *ctx, int nrhs, int *blocks);
/* columns per thread the SpMM J X chose for nrhs columns (1, 2 or 4, timed
 * at the first product; 0 if none yet) */
int iemic_spmm_tuning(iemic_ctx *ctx, int nrhs, int *columns);

/* reuse of row-LU factors across Jacobians (the reference's Newton-Chord
 * hybrid, "enable Newton Chord hybrid solve", Continuation.H:1360): the
 * factors are rebuilt only when they are more than max_age Jacobians old,
 * or when a solve with older factors misses its tolerance (then once, and
 * the solve is repeated).  0 (default): rebuilt for every Jacobian. */
int iemic_set_precond_reuse(iemic_ctx *ctx, int max_age);

/* choose the preconditioner (IEMIC_PREC_*); drops existing factors */
int iemic_set_preconditioner(iemic_ctx *ctx, int kind);
int iemic_preconditioner_info(iemic_ctx *ctx, iemic_precond_info *info);
int iemic_build_preconditioner(iemic_ctx *ctx);
int iemic_apply_precon(iemic_ctx *ctx, const double *v, double *out);
int iemic_apply_precon_dev(iemic_ctx *ctx, const double *d_v, double *d_out);
/* the same with the factors computed and applied in FP32 (vectors stay FP64);
 * the preconditioner of the mixed-precision solve */
int iemic_apply_precon_fp32(iemic_ctx *ctx, const double *v, double *out);

/* ---- Ocean::solve (Ocean.C:1070-1147): J x = b, x0 = 0 ------------------ */
void iemic_solver_defaults(iemic_solver_params *p);
int iemic_solve(iemic_ctx *ctx, const iemic_solver_params *p, const double *b, double *x,
                iemic_solve_info *info);
int iemic_solve_dev(iemic_ctx *ctx, const iemic_solver_params *p, const double *d_b,
                    double *d_x, iemic_solve_info *info);
/* adjoint solve J^T x = b (no counterpart in the reference, whose
 * preconditioner rejects transposition, TRIOS_BlockPreconditioner.H:122):
 * FGMRES with J^T products and the transposed row-LU preconditioner; for
 * sensitivities dg/dp = -x^T dF/dp with b = dg/dx */
int iemic_solve_transpose(iemic_ctx *ctx, const iemic_solver_params *p, const double *b,
                          double *x, iemic_solve_info *info);
int iemic_solve_transpose_dev(iemic_ctx *ctx, const iemic_solver_params *p, const double *d_b,
                              double *d_x, iemic_solve_info *info);
int iemic_apply_precon_transpose_dev(iemic_ctx *ctx, const double *d_v, double *d_out);

/* ---- several right-hand sides (no single counterpart in the reference:
 * Ocean::solve takes one vector, Ocean.C:1070-1147; the Belos block solver
 * it wraps accepts a multivector) --------------------------------------- */
/* J X = B for nrhs columns stored one after another (column q at B + q
 * ndim).  FGMRES per column in lock step; the preconditioner is applied to
 * all columns in one batched sweep (row LU: the dense products become
 * GEMMs) and J in one SpMM.  infos (length nrhs, may be NULL) receives each
 * column's result; returns IEMIC_ENOCONV if any column missed the
 * tolerance. */
int iemic_solve_multi(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                      iemic_solve_info *infos);
int iemic_solve_multi_dev(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *d_B,
                          double *d_X, iemic_solve_info *infos);
/* J^T X = B for nrhs columns (adjoint batches: several diagnostics or
 * states), the same lock-step FGMRES with the transposed batched sweep */
int iemic_solve_multi_transpose(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                                iemic_solve_info *infos);
int iemic_solve_multi_transpose_dev(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *d_B,
                                    double *d_X, iemic_solve_info *infos);
/* J X = B (and J^T X = B) for 1 to 32 columns by block FGMRES: one Krylov
 * space shared by all columns (the Belos block GMRES that Ocean::solve's
 * solver manager can run on a multivector, generalising the Gram-Schmidt
 * loop of GMRESSolver.H:177-181 to blocks).  The block Gram-Schmidt and
 * basis updates run on the matrix cores; directions already in the space
 * are deflated, so after the first iteration the block shrinks to the few
 * directions the preconditioner misses.  restart = block iterations per
 * cycle; iterations in infos count the block iterations each column took
 * part in.  Same contract as iemic_solve_multi otherwise. */
int iemic_solve_block(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                      iemic_solve_info *infos);
int iemic_solve_block_dev(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *d_B,
                          double *d_X, iemic_solve_info *infos);
int iemic_solve_block_transpose(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *B, double *X,
                                iemic_solve_info *infos);
int iemic_solve_block_transpose_dev(iemic_ctx *ctx, const iemic_solver_params *p, int nrhs, const double *d_B,
                                    double *d_X, iemic_solve_info *infos);
/* the tall-skinny products of the block solver on their own (no counterpart
 * in the reference; exposed so that they can be checked against a host
 * product): G = A^T B (a x b) and C = beta C + alpha A H (H: a x b), column-
 * major operands with leading dimension n (G and H: a), b <= 32; C may be A
 * itself when b <= a.  Null stream; returns when done. */
int iemic_block_gram_dev(int n, int a, int b, const double *d_A, const double *d_B, double *d_G);
int iemic_block_update_dev(int n, int a, int b, double alpha, const double *d_A, const double *d_H,
                           double beta, double *d_C);
/* out = M^{-1} in on nrhs columns (FP64 factors) */
int iemic_apply_precon_multi_dev(iemic_ctx *ctx, int nrhs, const double *d_in, double *d_out);

/* ---- the preconditioner plug-in of the reference's MRILU interface -------
 * (mrilucpp.F90:120-554, declared for C++ in Ifpack_MRILU.cpp:25-37): an
 * integer handle per instance (id = 0 on create: a new one; id > 0: that
 * instance overwritten), the matrix handed over as 0-based CSR (copied),
 * set_params, compute (once per matrix), apply (before compute: sol = rhs,
 * with a warning), destroy (sets id = 0).  Behind it runs a GPU block ILU(0)
 * with level-scheduled sweeps over blocks of `blocksize` (1..8) rows, not
 * MRILU's multilevel renumbering: blocksize and lutol (pivot floor relative
 * to the block's largest entry) are used, the other parameters are kept and
 * have no effect.  Host vectors. */
int iemic_mrilu_create(int *id, int n, int nnz, const int *beg, const int *jco, const double *co);
int iemic_mrilu_set_params(int id, int blocksize, int cutmck, int scarow, int xactelm, int clsonce, double nlsfctr,
                           double epsw, double elmfctr, int gusmod, double gusfctr, double redfctr, double schtol,
                           double denslim, double globfrac, double locfrac, double sparslim, int ilutype,
                           double droptol, double compfct, double cpivtol, double lutol, int singlu, int outlev);
int iemic_mrilu_compute(int id);
int iemic_mrilu_apply(int id, int n, const double *rhs, double *sol);
/* block entries, forward dependency levels and floored pivots of the factors */
int iemic_mrilu_info(int id, int *blocks, int *levels, int *pivots_floored);
int iemic_mrilu_destroy(int *id);
/* The same under the reference's own names and argument passing (every
 * scalar by address, Ifpack_MRILU.cpp:25-37), so that Ifpack_MRILU.cpp links
 * unchanged; as the Fortran's `stop`, an error prints its message and ends
 * the process. */
void mrilucpp_create(int *id, int *n, int *nnz, int *beg, int *jco, double *co);
void mrilucpp_destroy(int *id);
void mrilucpp_set_params(const int *id, int *blocksize, int *cutmck, int *scarow, int *xactelm, int *clsonce,
                         double *nlsfctr, double *epsw, double *elmfctr, int *gusmod, double *gusfctr, double *redfctr,
                         double *schtol, double *denslim, double *globfrac, double *locfrac, double *sparslim,
                         int *ilutype, double *droptol, double *compfct, double *cpivtol, double *lutol, int *singlu,
                         int *outlev);
void mrilucpp_compute(const int *id);
void mrilucpp_apply(const int *id, int *dim, const double *rhs, double *sol);

/* ---- callers of the path on device state -------------------------------
 * Damped Newton on F(x) = 0 (the corrector of Continuation::newtonCorrector,
 * Continuation.H:611-790, without the arclength row): x -= t dx, J dx = F,
 * the step's pressure-mode part removed, t halved while ||F|| does not fall.
 * Stops at ||F|| <= tol or a step with ||t dx||_inf <= step_tol; x (device,
 * ndim) holds the last iterate.  Returns IEMIC_ENOCONV if neither test was
 * met within maxit iterations. */
int iemic_newton_dev(iemic_ctx *ctx, const iemic_solver_params *p, double tol, double step_tol, int maxit,
                     double *d_x, int *iterations, double *fnorm);

/* Pseudo-arclength continuation in one parameter (Continuation::run / step,
 * Continuation.H:200-1075): Euler initial tangent, secant tangents, the
 * quadratic normalisation zeta ||x - x0||^2 + (p - p0)^2 = ds^2 with
 * zeta = tan_scaling / ndim, a Newton corrector whose two solves per
 * iteration (J y = dF/dp, J z = -F) are one two-column solve, residual test
 * max(|dp|, ||dx||_inf) < newton_tol, step control by Newton iterations
 * (adjustStep) and halving after a failure (reset).  The run ends when the
 * parameter passes target (a Newton solve at p = target finishes it) or
 * after max_steps accepted steps.  par_hist / norm_hist (may be NULL) get
 * the parameter and ||x|| at the start and after every accepted step
 * (max_steps + 1 entries).  Defaults follow getDefaultInitParameters
 * (Continuation.H:1330-1385). */
typedef struct {
    int par;                 /* parameter index 1..30 (COMB = 19) */
    double target;
    double ds, ds_min, ds_max;
    int max_steps;
    int newton_maxit;        /* 7 */
    double newton_tol;       /* 1e-4 */
    double eps;              /* dF/dp forward-difference step, 1e-5 */
    double tan_scaling;      /* 1 */
    double opt_newton;       /* 3.5 */
    double predictor_bound;  /* 1e3 */
} iemic_continuation_params;
typedef struct {
    int steps, rejected, newton_iterations, linear_iterations, reached;
    double par, seconds;
} iemic_continuation_info;
void iemic_continuation_defaults(iemic_continuation_params *p);
int iemic_continuation_dev(iemic_ctx *ctx, const iemic_solver_params *p, const iemic_continuation_params *cp,
                           double *d_x, double *par_hist, double *norm_hist, iemic_continuation_info *info);

/* ---- steady-state diagnostics and their state gradients (host arrays) --- */
/* which = IEMIC_DIAG_MEAN_SST: area-mean surface T (anomaly);
 * which = IEMIC_DIAG_MAX_PSIM: maximum meridional overturning streamfunction,
 * as OceanGrid::psimMax (OceanGrid.C:270-345, compute_psim thcm_utils.F90:95-118),
 * non-dimensional.  grad (length ndim, may be NULL) = dg/dx, the right-hand
 * side of the adjoint solve. */
enum { IEMIC_DIAG_MEAN_SST = 0, IEMIC_DIAG_MAX_PSIM = 1 };
int iemic_diagnostic(const iemic_ctx *ctx, int which, const double *x, double *value, double *grad);

/* ---- several independent models (no counterpart in the reference: one
 * THCM instance per process, THCM.H:76-84) ----------------------------------
 * One Newton step per context: F(x_q), J(x_q), the preconditioner and the
 * solve J dx_q = F(x_q), with every context on its own host thread and its
 * own stream when concurrent != 0 (the pivoted factorisation of one model
 * leaves most of the GPU idle; several models fill it), else one after the
 * other.  Contexts must be distinct.  infos (length count, may be NULL) get
 * each solve's result. */
int iemic_newton_step_many(iemic_ctx *const *ctx, int count, const double *const *d_x, double *const *d_dx,
                           const iemic_solver_params *p, iemic_solve_info *infos, int concurrent);

/* ---- on-disk land masks (readmask, topo.F90:41-134) --------------------- */
/* Reads a data/mkmask text mask into landm ((n+2)(m+2)(l+2) ints, i fastest,
 * the layout iemic_create takes), with readmask's land-inversion fix.  Host
 * only (no GPU needed).  Rows may hold at most 362 values, as the reference
 * format does. */
int iemic_read_mask(const char *path, int n, int m, int l, int *landm);

/* ---- state files (EpetraExt-HDF5, Model.C:15-160) ------------------------ */
/* Model::loadStateFromFile (Model.C:15-110): reads /State/Values of a state
 * file in the reference's layout (/State: Values [1][n], GlobalLength,
 * NumVectors, __type__; /Parameters: one double per THCM::int2par label,
 * THCM.C:1895-1937) into x (n doubles, reference ordering).  par and found
 * (31 entries each, index 1..30 as par.F90; either may be NULL) receive the
 * parameters the file holds and 1/0 for present/absent (the reference skips
 * absent ones).  Fails with IEMIC_EINVAL if the file's length differs from n.
 * Host only; the HDF5 C library is loaded at the first call (libhdf5.so,
 * /opt/conda/lib/libhdf5.so, or $IEMIC_HDF5_LIB), IEMIC_EUNSUPPORTED if none
 * is found. */
int iemic_read_state(const char *path, double *x, int64_t n, double *par, int *found);
/* Model::saveStateToFile (Model.C:112-160): writes x (n doubles) and the 30
 * parameters par[1..30] (may be NULL: no /Parameters group) in that layout,
 * creating or truncating path.  The /Grid group the reference adds for its
 * plotting scripts is not written. */
int iemic_write_state(const char *path, const double *x, int64_t n, const double *par);

/* ---- device memory helpers for callers without their own allocator ----- */
/* y = a x + b y (a = 0: x is not read; b = 0: y is not read) and
 * *out = x . y on device vectors
 * of length n; the null stream, and the call returns when done.  The
 * level-1 operations of the GMRES template contract's Vector
 * (GMRESSolverDecl.H:11-14: update, norm; GMRESSolver.H:178-186: dot,
 * scale), wrapped for C++ in include/iemic_gmres.hpp. */
int iemic_vec_axpby_dev(long long n, double a, const double *d_x, double b, double *d_y);
int iemic_vec_dot_dev(long long n, const double *d_x, const double *d_y, double *out);
int iemic_dev_alloc(void **ptr, long long bytes);
int iemic_dev_free(void *ptr);
int iemic_memcpy_h2d(void *dst, const void *src, long long bytes);
int iemic_memcpy_d2h(void *dst, const void *src, long long bytes);
int iemic_synchronize(iemic_ctx *ctx);

/* ---- timing of the last device launches (HIP events, milliseconds) ----- */
/* kernel: 0 = F assembly, 1 = J assembly, 2 = SpMV, 3 = preconditioner
 * apply (FP64), 4 = preconditioner build (FP64), 5 = build (FP32),
 * 6 = apply (FP32), 7 = batched apply (several right-hand sides); SpMM
 * launches are timed under 2 */
int iemic_last_kernel_ms(const iemic_ctx *ctx, int kernel, double *ms);

#ifdef __cplusplus
}
#endif
#endif
