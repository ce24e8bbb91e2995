// Host-side set-up of one ocean model: grid, continuation parameters,
// forcing, the state-independent ("linear") part of the stencil, integral
// condition, mass matrix and the reference CSR map.  This is the part of the
// reference that runs once per parameter change (usrc.F90 init/setparcs_:
// grid, stpnt, forcing, lin); the per-state work runs in assembly.hip.
// Every floating-point expression keeps the reference's operand order so the
// tables round exactly like the Fortran arrays they replace.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "ctx.hpp"

namespace iemic {

static thread_local std::string g_err;
int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}
const char *last_error() { return g_err.c_str(); }

namespace {

enum Par { AL_T = 1, RAYL, EK_V, EK_H, ROSB, MIXP, RESC, SPL1, HMTP, SUNP, PE_H, PE_V, P_VC,
           LAMB, SALT, WIND, TEMP, BIOT, COMB, ARCL, NLES, IFRICB, CONT, ENER, ALPC, CMPR,
           FPER, SPER, MKAP, SPL2 };

// usr.F90:131-160, par.F90:18
const double kPi = 3.14159265358979323846;
const double kOmega = 7.292e-05, kR0 = 6.37e+06, kUdim = 0.1e+00, kG = 9.8e+00;
const double kRho = 1.024e+03, kAh = 2.5e+05, kAv = 1.0e-03, kKappaH = 1.0e+03;
const double kKappaV = 1.0e-04;

struct Dims {
    int n, m, l;
};

inline int lm(const iemic_ctx *c, int i, int j, int k)
{
    return c->h.landm[i + (c->cfg.n + 2) * (j + (c->cfg.m + 2) * k)];
}
inline int row(const iemic_ctx *c, int i, int j, int k, int xx)  // 1-based i,j,k; xx 0..5
{
    return 6 * ((i - 1) + c->cfg.n * ((j - 1) + c->cfg.m * (k - 1))) + xx;
}

double fz(double z, double qz)  // grid.F90:58-73
{
    double th = std::tanh(qz * (z + 1)), tth = std::tanh(qz);
    if (qz > 1.0) return -1 + th / tth;
    return z + (1. - qz) * z * (1 - z);
}
double dfdz(double z, double qz)  // grid.F90:75-91
{
    double ch = std::cosh(qz * (z + 1)), tth = std::tanh(qz);
    if (qz > 1.0) return qz / (tth * ch * ch);
    return 1.0 + (1. - qz) * (1. - 2. * z);
}
// spf.F90:792-851
double amh(double y, int ih) { return ih == 0 ? 1.0 : 1. + 10.0 * std::exp(-5 * y * y); }
double bmh(double y, int ih) { return ih == 0 ? 1.0 : 1.0 + 10.0 * std::exp(-5 * y * y); }
double bmhy(double y, int ih) { return ih == 0 ? 0.0 : -10. * 10.0 * y * std::exp(-5 * y * y); }

void make_grid(iemic_ctx *c)  // grid.F90:2-56
{
    Host &h = c->h;
    int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    double qz = c->cfg.qz, zmin = -1.0, zmax = 0.0;
    h.x.assign(n + 1, 0.0); h.xu.assign(n + 1, 0.0); h.y.assign(m + 2, 0.0);
    h.yv.assign(m + 1, 0.0); h.z.assign(l + 1, 0.0); h.zw.assign(l + 1, 0.0);
    h.dfzT.assign(l + 1, 0.0); h.dfzW.assign(l + 1, 0.0);
    h.dx = (h.xmax - h.xmin) / n;
    h.dy = (h.ymax - h.ymin) / m;
    h.dz = (zmax - zmin) / l;
    for (int i = 1; i <= n; i++) {
        h.x[i] = ((double)i - 0.5) * h.dx + h.xmin;
        h.xu[i] = ((double)i) * h.dx + h.xmin;
    }
    h.xu[0] = h.xmin;
    for (int j = 1; j <= m; j++) {
        h.y[j] = ((double)j - 0.5) * h.dy + h.ymin;
        h.yv[j] = ((double)j) * h.dy + h.ymin;
    }
    h.y[0] = h.y[1] - h.dy;
    h.y[m + 1] = h.y[m] + h.dy;
    h.yv[0] = h.ymin;
    for (int k = 1; k <= l; k++) {
        double ze = ((double)k - 0.5) * h.dz + zmin, zwe = ((double)k) * h.dz + zmin;
        h.z[k] = fz(ze, qz);
        h.zw[k] = fz(zwe, qz);
        h.dfzT[k] = dfdz(ze, qz);
        h.dfzW[k] = dfdz(zwe, qz);
    }
    h.zw[0] = zmin;
    h.dfzW[0] = dfdz(zmin, qz);
}

void stpnt(iemic_ctx *c)  // usrc.F90:1123-1160 and mix_imp.f:142-160
{
    double *par = c->h.par, hdim = c->cfg.hdim, dz = c->h.dz, dfl = c->h.dfzT[c->cfg.l];
    for (int q = 0; q <= 30; q++) par[q] = 0.0;
    par[AL_T] = 0.1 / (2 * kOmega * kRho * hdim * kUdim * dz * dfl);
    par[RAYL] = c->cfg.alphaT * kG * hdim / (2 * kOmega * kUdim * kR0);
    par[EK_V] = kAv / (2 * kOmega * hdim * hdim);
    par[EK_H] = kAh / (2 * kOmega * kR0 * kR0);
    par[ROSB] = kUdim / (2 * kOmega * kR0);
    par[PE_H] = kKappaH / (kUdim * kR0);
    par[PE_V] = kKappaV * kR0 / (kUdim * hdim * hdim);
    par[P_VC] = 2.5e+04 * par[PE_V];
    par[LAMB] = c->cfg.alphaS / c->cfg.alphaT;
    par[BIOT] = kR0 / (75. * 3600. * 24. * kUdim);
    par[ALPC] = 1.0;
    par[ENER] = 1.0e+02;
    par[SPL1] = 2.0e+03;
    par[SPL2] = 0.01;
    if (c->cfg.vmix == 0) {
        par[MIXP] = 0.0; par[P_VC] = 0.0; par[ALPC] = 1.0; par[ENER] = 1.0e+2; par[MKAP] = 0.0;
    }
}

// THCM.C:2653-2686 (Fortran qint): area-weighted surface mean
double qint(const iemic_ctx *c, const std::vector<double> &f)
{
    int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    double lsint = 0.0, lfsint = 0.0;
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++) {
            int om = 1 - lm(c, i, j, l);
            lfsint = f[(i - 1) + n * (j - 1)] * std::cos(c->h.y[j]) * om + lfsint;
            lsint = std::cos(c->h.y[j]) * om + lsint;
        }
    return lfsint / lsint;
}

double temfun(const iemic_ctx *c, double yy)  // forcing.F90:417-427
{
    const Host &h = c->h;
    if (c->cfg.forcing_type == 2) return std::cos(kPi * (yy - h.ymin) / (h.ymax - h.ymin));
    return std::cos(kPi * yy / h.ymax) + h.par[CMPR] * std::sin(kPi * yy / h.ymax);
}
double salfun(const iemic_ctx *c, double yy)  // forcing.F90:430-442
{
    const Host &h = c->h;
    if (c->cfg.forcing_type == 2) return std::cos(kPi * (yy - h.ymin) / (h.ymax - h.ymin));
    if (c->cfg.forcing_type == 1)
        return (std::cos(kPi * yy / h.ymax) + h.par[FPER] * yy / h.ymax) / std::cos(yy);
    return std::cos(kPi * yy / h.ymax) + h.par[FPER] * yy / h.ymax;
}
double wfun(double yy)  // forcing.F90:402-414
{
    return 0.2 - 0.8 * std::sin(6 * std::fabs(yy)) - 0.5 * (1 - std::tanh(10 * std::fabs(yy)))
           - 0.5 * (1 - std::tanh(10 * (kPi / 2 - std::fabs(yy))));
}

// forcing.F90:4-218 (uncoupled) followed by the Frc rows boundary.F90 zeroes
void forcing(iemic_ctx *c)
{
    Host &h = c->h;
    int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l, TRES = c->cfg.tres, SRES = c->cfg.sres;
    double *par = h.par;
    std::vector<double> &F = h.frc;
    F.assign(c->sc.ndim, 0.0);
    double sigma = par[COMB] * par[WIND] * par[AL_T];
    if (c->cfg.iza == 2)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                h.taux[(i - 1) + n * (j - 1)] = wfun(h.yv[j]);
                h.tauy[(i - 1) + n * (j - 1)] = 0.0;
            }
    for (int j = 1; j <= m - 1; j++)
        for (int i = 1; i <= n; i++) {
            F[row(c, i, j, l, U)] = sigma * h.taux[(i - 1) + n * (j - 1)];
            F[row(c, i, j, l, V)] = sigma * h.tauy[(i - 1) + n * (j - 1)];
        }
    double etabi = par[COMB] * par[TEMP] * ((double)(1 - TRES) + TRES * par[BIOT]);
    double temcor = 0.0;
    if (c->cfg.ite == 1) {
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) h.tatm[(i - 1) + n * (j - 1)] = temfun(c, h.y[j]);
        if (TRES == 0) temcor = qint(c, h.tatm);
    }
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++)
            F[row(c, i, j, l, T)] = etabi * (h.tatm[(i - 1) + n * (j - 1)] - temcor);
    double gamma = par[COMB] * par[SALT] * ((double)(1 - SRES) + SRES * par[BIOT]);
    double salcor = 0.0, asalcor = 0.0, spertcor = 0.0;
    std::vector<double> aemip(n * m, 0.0);  // adapted_emip (never inserted here)
    if (c->cfg.its == 1) {
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                h.emip[(i - 1) + n * (j - 1)] = salfun(c, h.y[j]) * (1 - lm(c, i, j, l));
        if (SRES == 0) salcor = qint(c, h.emip);
    }
    if (SRES == 0) {
        asalcor = qint(c, aemip);
        spertcor = qint(c, h.spert);
    }
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++) {
            int q = (i - 1) + n * (j - 1);
            F[row(c, i, j, l, S)] = gamma * (1 - par[HMTP]) * (h.emip[q] - salcor)
                                    + gamma * par[HMTP] * (aemip[q] - asalcor)
                                    + par[SPER] * ((double)(1 - SRES) + SRES * par[BIOT])
                                          * (h.spert[q] - spertcor);
        }
    // internal w forcing with zero internal T/S fields (forcing.F90:184-193)
    for (int k = 1; k <= l - 1; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                F[row(c, i, j, k, W)] = -par[COMB] * (1 - lm(c, i, j, k)) * par[RAYL]
                                        * (par[LAMB] * (0.0 + 0.0) / 2. - (0.0 + 0.0) / 2.);
    // rows of Frc that boundary.F90 zeroes on every evaluation
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                if (lm(c, i, j, k) == OCEAN) {
                    if (lm(c, i, j, k + 1) == LAND) F[row(c, i, j, k, W)] = 0.0;
                    if (lm(c, i, j + 1, k) == LAND || lm(c, i + 1, j, k) == LAND ||
                        lm(c, i + 1, j + 1, k) == LAND) {
                        F[row(c, i, j, k, U)] = 0.0;
                        F[row(c, i, j, k, V)] = 0.0;
                    }
                } else {
                    for (int xx = 0; xx < 6; xx++) F[row(c, i, j, k, xx)] = 0.0;
                }
            }
}

// The linear part of the stencil for an ocean cell at latitude row j and level
// k (usrc.F90:605-789 lin with the spf.F90 atom builders), as the 104 slot
// values.  Atoms are 27-vectors indexed by kk-1.
void linear_slots(const iemic_ctx *c, int j, int k, double *out)
{
    const Host &h = c->h;
    const double *par = h.par;
    int m = c->cfg.m, l = c->cfg.l, ih = c->cfg.ih;
    double dx = h.dx, dy = h.dy, dz = h.dz;
    double yv = h.yv[j];
    enum { UXX, UYY, UZZ, UCSI, VXS, FV, PX, VXX, VYY, VZZ, VCSI, UXS, FU, PY, PZ, TBC, UXC, VYC,
           WZC, TC, SC, TXX, TYY, TZZ, NA };
    double a[NA][28];
    std::memset(a, 0, sizeof a);
#define A(t, kk) a[t][(kk)]
    bool jin = (j <= m - 1);
    // uderiv (spf.F90:13-67) and vderiv (69-121)
    if (jin) {
        double q = 1.0 / (std::cos(yv) * dx), cosdx2i = q * q;
        A(UXX, 2) = amh(yv, ih) * cosdx2i;
        A(UXX, 8) = amh(yv, ih) * cosdx2i;
        A(UXX, 5) = -(A(UXX, 2) + A(UXX, 8));
        A(VXX, 2) = bmh(yv, ih) * cosdx2i;
        A(VXX, 5) = -2 * bmh(yv, ih) * cosdx2i;
        A(VXX, 8) = bmh(yv, ih) * cosdx2i;
        double r = 1.0 / dy, rdy2i = r * r;
        A(UYY, 4) = rdy2i * bmh(h.y[j], ih) * std::cos(h.y[j]) / std::cos(yv);
        A(UYY, 6) = rdy2i * bmh(h.y[j + 1], ih) * std::cos(h.y[j + 1]) / std::cos(yv);
        A(UYY, 5) = -(A(UYY, 4) + A(UYY, 6));
        A(VYY, 4) = rdy2i * amh(h.y[j], ih) * std::cos(h.y[j]) / std::cos(yv);
        A(VYY, 6) = rdy2i * amh(h.y[j + 1], ih) * std::cos(h.y[j + 1]) / std::cos(yv);
        A(VYY, 5) = -(A(VYY, 4) + A(VYY, 6));
        double tand2 = 1 - std::tan(yv) * std::tan(yv);
        A(UCSI, 5) = bmh(yv, ih) * tand2 + std::tan(yv) * bmhy(yv, ih);
        A(VCSI, 5) = bmh(yv, ih) - amh(yv, ih) * std::tan(yv) * std::tan(yv) + bmhy(yv, ih) * std::tan(yv);
        double num_u = bmhy(yv, ih) - (amh(yv, ih) + bmh(yv, ih)) * std::tan(yv);
        A(VXS, 2) = num_u / (dx * std::cos(yv));
        A(VXS, 8) = -num_u / (dx * std::cos(yv));
        double num_v = (amh(yv, ih) + bmh(yv, ih)) * std::tan(yv) - bmhy(yv, ih);
        A(UXS, 2) = -num_v / (dx * std::cos(yv));
        A(UXS, 8) = num_v / (dx * std::cos(yv));
        // coriolis (spf.F90:262-293)
        A(FV, 5) = std::sin(yv) * c->cfg.coriolis_on;
        A(FU, 5) = std::sin(yv) * c->cfg.coriolis_on;
        // gradp types 1, 2 (spf.F90:296-321)
        double cosdxi = 1. / (2 * std::cos(yv) * dx);
        A(PX, 5) = -cosdxi; A(PX, 6) = -cosdxi; A(PX, 8) = cosdxi; A(PX, 9) = cosdxi;
        double dyi = 1. / (2 * dy);
        A(PY, 5) = -dyi; A(PY, 8) = -dyi; A(PY, 6) = dyi; A(PY, 9) = dyi;
    }
    {
        double q = 1.0 / dz, rdz2i = q * q;
        double h1 = 1. / (h.dfzT[k] * h.dfzW[k]), h2 = 1. / (h.dfzT[k] * h.dfzW[k - 1]);
        A(UZZ, 14) = h2 * rdz2i; A(UZZ, 23) = h1 * rdz2i;
        A(UZZ, 5) = -(A(UZZ, 14) + A(UZZ, 23));
        A(VZZ, 14) = h2 * rdz2i; A(VZZ, 23) = h1 * rdz2i;
        A(VZZ, 5) = -(A(VZZ, 14) + A(VZZ, 23));
        // tderiv type 5 (spf.F90:225-249), ocean column: mask factor 1
        double dz2i = q * q;
        A(TZZ, 14) = h2 * dz2i * 1;
        A(TZZ, 23) = (k < l) ? h1 * dz2i * 1 : 0.0;
        A(TZZ, 5) = -(A(TZZ, 14) + A(TZZ, 23));
        // gradp type 3
        double dzi = 1. / dz;
        A(PZ, 5) = -dzi / h.dfzW[k];
        A(PZ, 23) = dzi / h.dfzW[k];
        // pderiv type 3
        A(WZC, 5) = dzi / h.dfzT[k];
        A(WZC, 14) = -dzi / h.dfzT[k];
    }
    // pderiv types 1, 2 (spf.F90:123-160)
    {
        double cos2i = 1.0 / (2 * std::cos(h.y[j]) * dx);
        A(UXC, 2) = -cos2i; A(UXC, 4) = cos2i; A(UXC, 1) = -cos2i; A(UXC, 5) = cos2i;
        double cos2iy = 1. / (2 * std::cos(h.y[j]) * dy);
        A(VYC, 4) = -std::cos(h.yv[j - 1]) * cos2iy;
        A(VYC, 2) = std::cos(h.yv[j]) * cos2iy;
        A(VYC, 1) = -std::cos(h.yv[j - 1]) * cos2iy;
        A(VYC, 5) = std::cos(h.yv[j]) * cos2iy;
    }
    // tderiv types 1, 2, 3, 4, 6 (spf.F90:173-259), ocean column
    if (k == l) { A(TC, 5) = 1.0; A(SC, 5) = 1.0; }
    {
        double q = 1.0 / (std::cos(h.y[j]) * dx), cosdx2i = q * q;
        A(TXX, 2) = cosdx2i * 1; A(TXX, 5) = -2 * cosdx2i * 1; A(TXX, 8) = cosdx2i * 1;
        double r = 1.0 / dy, dy2i = r * r;
        A(TYY, 4) = (dy2i * std::cos(h.yv[j - 1]) / std::cos(h.y[j])) * 1;
        A(TYY, 6) = (dy2i * std::cos(h.yv[j]) / std::cos(h.y[j])) * 1;
        A(TYY, 5) = -(A(TYY, 4) + A(TYY, 6));
        A(TBC, 23) = 1.0 * 1; A(TBC, 5) = 1.0 * 1;
    }
    double EV = par[EK_V], EH = par[EK_H], ph = (1 - par[MIXP]) * par[PE_H], pv = par[PE_V];
    double lambda = par[LAMB], xes = par[NLES], bi = par[BIOT], Ra = par[RAYL];
    const double alpt1 = 2.93;
    for (int s = 0; s < NSLOT; s++) {
        int kk = SLOTS[s].kk, eq = SLOTS[s].eq, var = SLOTS[s].var;
        double v = 0.0;
        switch (eq) {
        case U:
            if (var == U) v = -EH * (A(UXX, kk) + A(UYY, kk) + A(UCSI, kk)) - EV * A(UZZ, kk);
            else if (var == V) v = -A(FV, kk) - EH * A(VXS, kk);
            else if (var == P) v = A(PX, kk);
            break;
        case V:
            if (var == U) v = A(FU, kk) - EH * A(UXS, kk);
            else if (var == V) v = -EH * (A(VXX, kk) + A(VYY, kk) + A(VCSI, kk)) - EV * A(VZZ, kk);
            else if (var == P) v = A(PY, kk);
            break;
        case W:
            if (var == P) v = A(PZ, kk);
            else if (var == T) v = -Ra * (1. + xes * alpt1) * A(TBC, kk) / 2.;
            else if (var == S) v = lambda * Ra * A(TBC, kk) / 2.;
            break;
        case P:
            if (var == U) v = A(UXC, kk);
            else if (var == V) v = A(VYC, kk);
            else if (var == W) v = A(WZC, kk);
            break;
        case T:
            if (var == T)
                v = -ph * (A(TXX, kk) + A(TYY, kk)) - pv * A(TZZ, kk) + c->cfg.tres * bi * A(TC, kk);
            break;
        case S:
            if (var == S)
                v = -ph * (A(TXX, kk) + A(TYY, kk)) - pv * A(TZZ, kk) + c->cfg.sres * bi * A(SC, kk);
            break;
        }
        out[s] = v;
    }
#undef A
}

void build_tables(iemic_ctx *c)
{
    Host &h = c->h;
    int m = c->cfg.m, l = c->cfg.l;
    h.al.assign((size_t)m * l * NSLOT, 0.0);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++) linear_slots(c, j, k, &h.al[((size_t)(k - 1) * m + (j - 1)) * NSLOT]);
    h.jt.assign((size_t)NJT * (m + 2), 0.0);
    for (int j = 0; j <= m + 1; j++) {
        double *t = &h.jt[0];
        int stride = m + 2;
        if (j <= m) {
            t[JT_CX_V * stride + j] = 1.0 / (2 * std::cos(h.yv[j]) * h.dx);
            t[JT_CY_V * stride + j] = 1.0 / (2 * std::cos(h.yv[j]) * h.dy);
            t[JT_COS_YV * stride + j] = std::cos(h.yv[j]);
            t[JT_TAN_YV * stride + j] = std::tan(h.yv[j]);
        }
        t[JT_CX_T * stride + j] = 1.0 / (4 * std::cos(h.y[j]) * h.dx);
        t[JT_CY_T * stride + j] = 1.0 / (4 * std::cos(h.y[j]) * h.dy);
    }
    h.kt.assign((size_t)NKT * (l + 2), 0.0);
    for (int k = 0; k <= l; k++) {
        double *t = &h.kt[0];
        int stride = l + 2;
        t[KT_DZDFZW * stride + k] = h.dz * h.dfzW[k];
        if (k >= 1) {
            t[KT_TDZI8 * stride + k] = 1.0 / (8 * h.dfzT[k] * h.dz);
            t[KT_DFZT * stride + k] = h.dfzT[k];
            t[KT_DZDFZT * stride + k] = h.dz * h.dfzT[k];
            t[KT_2DZDFZT * stride + k] = 2.0 * h.dz * h.dfzT[k];
        }
    }
}

void build_scalars(iemic_ctx *c)
{
    Scalars &s = c->sc;
    const double *par = c->h.par;
    s.epsr = par[ROSB];
    s.Ra = par[RAYL];
    s.xes = par[NLES];
    s.lambda = par[LAMB];
    s.alpt1 = 2.93;
    s.alpt2 = 8.3e-02;
    s.alpt3 = 6.6e-04;
    s.tdzi2 = 1.0 / (2 * c->h.dz);
    s.kvc = par[P_VC];
    s.eps_cons = (1.0 - par[ALPC]) * par[ENER] * par[PE_V];
    s.fac = c->cfg.alphaT * par[SPL1];
    s.rho_mixing = c->cfg.rho_mixing ? 1 : 0;
    s.vmix_on = (c->vmix_flag >= 1) ? 1 : 0;
    set_mix_switches(c);
}

// assemble.F90:18-54 fillcolB and THCM.C:1157-1160, 2197-2198
void build_diagB(iemic_ctx *c)
{
    Host &h = c->h;
    int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    h.diagB.assign(c->sc.ndim, 0.0);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                if (lm(c, i, j, k) == OCEAN) {
                    if (lm(c, i + 1, j, k) != LAND) h.diagB[row(c, i, j, k, U)] = -h.par[ROSB];
                    if (lm(c, i, j + 1, k) != LAND) h.diagB[row(c, i, j, k, V)] = -h.par[ROSB];
                    h.diagB[row(c, i, j, k, T)] = -1.0;
                    h.diagB[row(c, i, j, k, S)] = -1.0;
                }
    if (c->sc.rowintcon >= 0) h.diagB[c->sc.rowintcon] = 0.0;
}

}  // namespace

// The mixing switches of the kernels.  Mixing = 1: every term.  Mixing = 2:
// vmix_fun adds the T (S) rows' mixing only while the T (S) field is
// nonzero and once a pattern exists (mix_imp.f:505-556, usrc.F90:551-563);
// vmix_jac differences only the pattern's entries (vmix_el_2,
// mix_imp.f:934-1049): T rows against T, and against S and the S rows when
// the pattern was built with a nonzero S field.
void set_mix_switches(iemic_ctx *c)
{
    Scalars &s = c->sc;
    if (c->vmix_flag == 2) {
        s.mix_t = c->mix2_part && c->mix2_temp;
        s.mix_ts = s.mix_t && c->mix2_pat_salt;
        s.mix_sf = c->mix2_part && c->mix2_salt;
        s.mix_sj = s.mix_sf && c->mix2_pat_salt;
    } else {
        s.mix_t = s.mix_ts = s.mix_sf = s.mix_sj = 1;
    }
}


void host_parameters_changed(iemic_ctx *c)
{
    forcing(c);
    build_tables(c);
    build_scalars(c);
    build_diagB(c);
}

int host_init(iemic_ctx *c, const int *landm, const double *taux, const double *tauy,
              const double *tatm, const double *emip, const double *spert)
{
    const iemic_config &g = c->cfg;
    int n = g.n, m = g.m, l = g.l;
    if (n < 3 || m < 2 || l < 2) return set_error(IEMIC_EINVAL, "grid must be at least 3x2x2");
    if (g.coupled_t || g.coupled_s)
        return set_error(IEMIC_EUNSUPPORTED, "coupled ocean (Coupled Temperature/Salinity) is not supported");
    if (g.vmix != 0 && g.vmix != 1 && g.vmix != 2)
        return set_error(IEMIC_EUNSUPPORTED, "Mixing must be 0, 1 or 2");
    if (g.int_sign != 1 && g.int_sign != -1) return set_error(IEMIC_EINVAL, "Salinity Integral Sign must be +-1");
    Host &h = c->h;
    h.xmin = g.xmin_deg * kPi / 180.0; h.xmax = g.xmax_deg * kPi / 180.0;
    h.ymin = g.ymin_deg * kPi / 180.0; h.ymax = g.ymax_deg * kPi / 180.0;
    Scalars &s = c->sc;
    s.n = n; s.m = m; s.l = l; s.periodic = g.periodic ? 1 : 0;
    s.ncell = n * m * l; s.ndim = 6 * s.ncell;
    // land mask rules of usrc.F90:71-101
    size_t nl = (size_t)(n + 2) * (m + 2) * (l + 2);
    h.landm.assign(landm, landm + nl);
    for (int k = 0; k <= l + 1; k++)
        for (int j = 0; j <= m + 1; j++)
            for (int i = 0; i <= n + 1; i++) {
                int &q = h.landm[i + (n + 2) * (j + (m + 2) * k)];
                if (!g.periodic && q == PERIO) q = OCEAN;
                if ((!g.periodic && (i == 0 || i == n + 1)) || j == 0 || j == m + 1 || k == 0 || k == l + 1)
                    q = LAND;
            }
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                int v = lm(c, i, j, k);
                if (v != OCEAN && v != LAND)
                    return set_error(IEMIC_EINVAL, "interior mask values must be OCEAN (0) or LAND (1)");
                if (k < l && v == OCEAN && lm(c, i, j, k + 1) == LAND)
                    return set_error(IEMIC_EINVAL, "land above ocean in the mask (apply the land-inversion fix of topo.F90 first)");
            }
    size_t nm = (size_t)n * m;
    auto take = [nm](std::vector<double> &dst, const double *src) {
        dst.assign(nm, 0.0);
        if (src) std::memcpy(dst.data(), src, nm * sizeof(double));
    };
    take(h.taux, taux); take(h.tauy, tauy); take(h.tatm, tatm); take(h.emip, emip); take(h.spert, spert);
    make_grid(c);
    stpnt(c);
    c->vmix_flag = g.vmix;
    s.rowintcon = -1;
    if (g.sres == 0) {
        int nic = g.nic < 0 ? n - 1 : g.nic, mic = g.mic < 0 ? m - 1 : g.mic;
        if (nic >= n || mic >= m) return set_error(IEMIC_EINVAL, "integral row coordinates out of range");
        if (lm(c, nic + 1, mic + 1, l) != OCEAN)
            return set_error(IEMIC_EINVAL, "integral row coordinates give a land point");
        s.rowintcon = 6 * (((l - 1) * m + mic) * n + nic) + S;
    }
    // thcm_utils.F90:285-312 intcond_scaling
    h.coef.assign(s.ndim, 0.0);
    for (int k = 1; k <= l; k++)
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                if (lm(c, i, j, k) == OCEAN) h.coef[row(c, i, j, k, S)] = std::cos(h.y[j]) * h.dfzT[k];
    host_parameters_changed(c);
    return IEMIC_OK;
}

// THCM::CreateMaximalGraph (THCM.C:2300-2580) with rows sorted as
// Epetra_CrsGraph::FillComplete leaves them, plus the slot of each entry.
void build_graph(iemic_ctx *c)
{
    int N = c->cfg.n, M = c->cfg.m, L = c->cfg.l, ncell = c->sc.ncell, ndim = c->sc.ndim;
    bool per = c->cfg.periodic != 0;
    Graph &gr = c->graph;
    gr.rowptr.assign(ndim + 1, 0);
    gr.col.clear(); gr.src.clear();
    gr.col.reserve((size_t)ncell * NSLOT + (size_t)ncell);
    gr.src.reserve(gr.col.capacity());
    std::vector<std::pair<int, int>> ent;
    for (int k = 0; k < L; k++)
        for (int j = 0; j < M; j++)
            for (int i = 0; i < N; i++) {
                int cell = i + N * (j + M * k);
                for (int eq = 0; eq < 6; eq++) {
                    int r = 6 * cell + eq;
                    ent.clear();
                    if (r == c->sc.rowintcon) {
                        for (int cc = 0; cc < ncell; cc++) ent.push_back({6 * cc + S, -1});
                    } else {
                        for (int s = ROW0[eq]; s < ROW0[eq + 1]; s++) {
                            int kk = SLOTS[s].kk;
                            int ii = i + kk_di(kk), jj = j + kk_dj(kk), kq = k + kk_dk(kk);
                            if (per) ii = ((ii % N) + N) % N;
                            if (ii < 0 || ii >= N || jj < 0 || jj >= M || kq < 0 || kq >= L) continue;
                            int colv = 6 * (ii + N * (jj + M * kq)) + SLOTS[s].var;
                            ent.push_back({colv, s * ncell + cell});
                        }
                        std::sort(ent.begin(), ent.end());
                        // a duplicate column can only arise for N < 3 periodic grids,
                        // which host_init rejects
                    }
                    for (auto &e : ent) {
                        gr.col.push_back(e.first);
                        gr.src.push_back(e.second);
                    }
                    gr.rowptr[r + 1] = (int)gr.col.size();
                }
            }
    gr.nnz = (int)gr.col.size();
}

}  // namespace iemic

namespace iemic {

// Steady-state diagnostics for sensitivity studies, with their gradients
// with respect to the state.
//  0  area-mean SST: sum over ocean surface cells of cos(y_j) T / sum cos(y_j)
//     (T is the anomaly unknown; a constant offset does not change gradients)
//  1  maximum of the meridional overturning streamfunction, the quantity
//     OceanGrid::psimMax reports (OceanGrid.C:264-345, compute_psim in
//     thcm_utils.F90:95-118): vs(j,k) = dx sum_i v(i,j,k) over the v field
//     that usol (usrc.F90:1014-1121) unpacks, which is zero on the north
//     wall (j = m), on the east wall of a closed basin (i = n) and at every
//     velocity point next to a land cell;
//     psi(j,k) = -cos(yv_j) vs(j,k) dz dfzT(k) + psi(j,k-1) below 500 m;
//     the gradient is that of psi at its (first) maximiser.
int diagnostic(const iemic_ctx *c, int which, const double *x, double *value, double *grad)
{
    const int n = c->cfg.n, m = c->cfg.m, l = c->cfg.l;
    const Host &h = c->h;
    const size_t ndim = (size_t)6 * n * m * l;
    auto row = [&](int i, int j, int k, int var) {   // 1-based i, j, k
        return (size_t)6 * ((size_t)(i - 1) + (size_t)n * ((j - 1) + (size_t)m * (k - 1))) + var;
    };
    if (grad) std::fill(grad, grad + ndim, 0.0);
    if (which == 0) {
        double sw = 0.0, s = 0.0;
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++)
                if (lm(c, i, j, l) == OCEAN) sw += std::cos(h.y[j]);
        if (sw == 0.0) return set_error(IEMIC_EINVAL, "no ocean surface cells");
        for (int j = 1; j <= m; j++)
            for (int i = 1; i <= n; i++) {
                if (lm(c, i, j, l) != OCEAN) continue;
                const double w = std::cos(h.y[j]) / sw;
                s += w * x[row(i, j, l, T)];
                if (grad) grad[row(i, j, l, T)] = w;
            }
        *value = s;
        return IEMIC_OK;
    }
    if (which == 1) {
        // v(i, j, k) as usol leaves it: the cells (i, j), (i + 1, j),
        // (i, j + 1), (i + 1, j + 1) inside 1..n x 1..m that are land zero it
        auto vlive = [&](int i, int j, int k) {
            if (j >= m) return false;
            if (i >= n && !c->cfg.periodic) return false;
            if (lm(c, i, j, k) == LAND || lm(c, i, j + 1, k) == LAND) return false;
            if (i < n && (lm(c, i + 1, j, k) == LAND || lm(c, i + 1, j + 1, k) == LAND)) return false;
            return true;
        };
        double best = 0.0;
        int bj = -1, bk = -1;
        std::vector<double> psi((size_t)(m + 1) * (l + 1), 0.0);
        for (int j = 1; j <= m; j++) {
            const double cs = std::cos(h.yv[j]);
            for (int k = 1; k <= l; k++) {
                if (!(h.z[k] * c->cfg.hdim < -500.0)) continue;
                double vs = 0.0;
                for (int i = 1; i <= n; i++) vs += vlive(i, j, k) ? x[row(i, j, k, V)] : 0.0;
                vs *= h.dx;
                psi[j + (size_t)(m + 1) * k] = -cs * vs * h.dz * h.dfzT[k] + psi[j + (size_t)(m + 1) * (k - 1)];
            }
        }
        // maximum over (j, k) as the reference scans it, psi(0,0) = 0 included
        for (int j = 0; j <= m; j++)
            for (int k = 0; k <= l; k++)
                if (psi[j + (size_t)(m + 1) * k] > best) {
                    best = psi[j + (size_t)(m + 1) * k];
                    bj = j;
                    bk = k;
                }
        *value = best;
        if (grad && bj >= 1)
            for (int k = 1; k <= bk; k++) {
                if (!(h.z[k] * c->cfg.hdim < -500.0)) continue;
                const double w = -std::cos(h.yv[bj]) * h.dx * h.dz * h.dfzT[k];
                for (int i = 1; i <= n; i++)
                    if (vlive(i, bj, k)) grad[row(i, bj, k, V)] = w;
            }
        return IEMIC_OK;
    }
    return set_error(IEMIC_EINVAL, "unknown diagnostic");
}

}  // namespace iemic

namespace iemic {

// The land-mask text files of data/mkmask, as readmask reads them
// (topo.F90:41-134): for each level k = 0 .. l+1 one header line, then the
// rows j = m+1 down to 0, each n+2 one-digit values for i = 0 .. n+1
// (format 362i1: a short record is padded with blanks, which read as 0).
// Afterwards a land cell above an ocean cell turns the ocean cell below it
// into land (the land-inversion fix, topo.F90:97-105).  Output: the bordered
// (n+2)(m+2)(l+2) mask, i fastest.  A file that ends early is an error (the
// header read of the reference has no error branch).
int read_mask_text(const char *path, int n, int m, int l, int *landm)
{
    if (!path || !landm || n < 1 || m < 1 || l < 1) return set_error(IEMIC_EINVAL, "read_mask: bad arguments");
    if (n + 2 > 362) return set_error(IEMIC_EUNSUPPORTED, "read_mask: rows longer than 362 values (format 362i1)");
    std::FILE *f = std::fopen(path, "r");
    if (!f) return set_error(IEMIC_EINVAL, std::string("read_mask: cannot open ") + path);
    const int N = n + 2, M = m + 2;
    auto at = [&](int i, int j, int k) -> int & { return landm[i + N * (j + M * k)]; };
    std::string line;
    auto getline = [&](std::string &out) -> bool {
        out.clear();
        int ch;
        bool any = false;
        while ((ch = std::fgetc(f)) != EOF) {
            any = true;
            if (ch == '\n') break;
            if (ch != '\r') out.push_back((char)ch);
        }
        return any;
    };
    int rc = IEMIC_OK;
    for (int k = 0; k <= l + 1 && rc == IEMIC_OK; k++) {
        if (!getline(line)) rc = set_error(IEMIC_EINVAL, "read_mask: file ends before level " + std::to_string(k));
        for (int j = m + 1; j >= 0 && rc == IEMIC_OK; j--) {
            if (!getline(line)) {
                rc = set_error(IEMIC_EINVAL, "read_mask: file ends in level " + std::to_string(k));
                break;
            }
            for (int i = 0; i < N; i++) {
                const char ch = i < (int)line.size() ? line[i] : ' ';
                if (ch == ' ') at(i, j, k) = 0;
                else if (ch >= '0' && ch <= '9') at(i, j, k) = ch - '0';
                else {
                    rc = set_error(IEMIC_EINVAL, "read_mask: not a digit in level " + std::to_string(k) + ", row " +
                                                     std::to_string(j));
                    break;
                }
            }
        }
    }
    std::fclose(f);
    if (rc) return rc;
    for (int i = 1; i <= n; i++)
        for (int j = 1; j <= m; j++)
            for (int k = l; k >= 2; k--)
                if (at(i, j, k) == LAND && at(i, j, k - 1) == OCEAN) at(i, j, k - 1) = LAND;
    return IEMIC_OK;
}

}  // namespace iemic
