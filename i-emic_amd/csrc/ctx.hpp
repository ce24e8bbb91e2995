// Context of one ocean model: host-side set-up data and device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/iemic_amd.h"
#include "model.hpp"

namespace iemic {

struct Precond;   // preconditioner (precond.hip)
struct Krylov;    // solver workspace (solver.hip)
struct RowLU;     // row-block LU preconditioner (rowlu.hip)

struct Host {
    // grid (grid.F90), Fortran lower bounds kept: x[1..n], y[0..m+1],
    // yv[0..m], z/dfzT[1..l], dfzW[0..l]
    std::vector<double> x, y, xu, yv, z, zw, dfzT, dfzW;
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0, dx = 0, dy = 0, dz = 0;
    std::vector<int> landm;          // (n+2)(m+2)(l+2), after init_ border rules
    double par[31] = {0};
    std::vector<double> taux, tauy, tatm, emip, spert;  // n*m
    std::vector<double> frc;         // ndim, after boundary zeroing
    std::vector<double> coef;        // ndim, integral-condition coefficients
    std::vector<double> diagB;       // ndim, THCM::evaluate diag(B)
    std::vector<double> al;          // m*l*104 linear slot values per (j,k)
    std::vector<double> jt, kt;      // NJT*(m+2), NKT*(l+2)
};

struct Graph {  // reference CSR of the fixed graph and its slot map
    std::vector<int> rowptr, col;    // ndim+1, nnz
    std::vector<int> src;            // nnz: slot-format index (s*ncell+c), -1 for
                                     // the integral row entries
    int nnz = 0;
};

}  // namespace iemic

struct iemic_ctx {
    iemic_config cfg{};
    iemic::Scalars sc{};
    iemic::Host h;
    iemic::Graph graph;
    bool graph_built = false;
    double intcorr = 0.0;
    int vmix_flag = 1;
    // Mixing = 2 (mix_imp.f:131-227): current T / S switches, and the pattern
    // of the finite-difference mixing Jacobian (built at all; with salinity)
    int mix2_temp = 0, mix2_salt = 0, mix2_part = 0, mix2_pat_salt = 0;

    hipStream_t stream = nullptr;
    // device data
    uint8_t *d_landm = nullptr;
    double *d_al = nullptr, *d_jt = nullptr, *d_kt = nullptr;
    double *d_frc = nullptr, *d_coef = nullptr, *d_diagB = nullptr;
    double *d_J = nullptr;           // 104*ncell slot values
    double *d_x = nullptr, *d_y = nullptr;  // staging for host-pointer entry points
    double *d_red = nullptr;         // reduction partials
    int nred = 0;
    int *d_src = nullptr;            // CSR export map (slot index per entry)
    int *d_col_export = nullptr;     // CSR export columns
    double *d_icrow = nullptr;       // the integral row's S-equation entries (Fortran-layout export)
    // caller buffers of the Fortran-layout interface (m_mat::set_pointers)
    int f_nrows = 0, f_nnz = 0;
    int *f_begA = nullptr, *f_jcoA = nullptr;
    double *f_coA = nullptr, *f_coB = nullptr;
    bool have_jac = false;
    unsigned jac_version = 0;        // bumped by every Jacobian assembly
    int prec_kind = IEMIC_PREC_ROW_LU;
    unsigned prec_max_age = 0;       // factors may be this many Jacobians old
    std::map<int, int> spmm_nc;      // tuned columns per thread of the SpMM, per column count
    iemic::Precond *prec = nullptr;
    iemic::RowLU *rowlu = nullptr;
    iemic::Krylov *kry = nullptr;
    // timing
    hipEvent_t ev[8][2] = {};
    double last_ms[8] = {0};
};

namespace iemic {
// host_setup.cpp
int host_init(iemic_ctx *c, const int *landm, const double *taux, const double *tauy,
              const double *tatm, const double *emip, const double *spert);
void host_parameters_changed(iemic_ctx *c);   // forcing + linear tables (setparcs_)
void build_graph(iemic_ctx *c);
void set_mix_switches(iemic_ctx *c);   // mixing switches of the kernels (Mixing = 1 or 2)
int diagnostic(const iemic_ctx *c, int which, const double *x, double *value, double *grad);
int read_mask_text(const char *path, int n, int m, int l, int *landm);
int set_error(int code, const std::string &msg);
// assembly.hip
int upload_tables(iemic_ctx *c);
int launch_rhs(iemic_ctx *c, const double *d_x, double *d_F);
int launch_rhs_partials(iemic_ctx *c, const double *d_x, double *d_F, double *d_part, int cap);
int mixing_control(iemic_ctx *c, const double *d_x, double *d_part, int cap);
int launch_jacobian(iemic_ctx *c, const double *d_x);
int launch_rhs_fortran(iemic_ctx *c, const double *d_x, double *d_B);
int launch_fortran_count(iemic_ctx *c, int *d_cnt);
int launch_fortran_fill(iemic_ctx *c, const int *d_beg, int *d_col, double *d_val, int *d_bad);
int launch_spmv(iemic_ctx *c, const double *d_x, double *d_y);
int launch_spmv_t(iemic_ctx *c, const double *d_x, double *d_y);
int launch_spmm(iemic_ctx *c, int nrhs, const double *d_X, size_t ldx, double *d_Y, size_t ldy);
int launch_spmm_t(iemic_ctx *c, int nrhs, const double *d_X, size_t ldx, double *d_Y, size_t ldy);
int launch_export(iemic_ctx *c, double *d_val);
int launch_pressure_projection(iemic_ctx *c, double *d_v);
double device_dot(iemic_ctx *c, const double *d_a, const double *d_b, int n);
void timed_begin_ext(iemic_ctx *c, int which);
void timed_end_ext(iemic_ctx *c, int which);
}  // namespace iemic

#define HIPCHK(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return iemic::set_error(IEMIC_EHIP, std::string(#expr) + ": " +       \
                                                    hipGetErrorString(e_));       \
    } while (0)
