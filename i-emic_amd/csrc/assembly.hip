// THCM residual F(x) and Jacobian J(x) assembly on gfx950.
//
// One thread per grid cell (i fastest, like the reference's unknown order).
// A thread builds the six rows of its cell directly in slot form: linear
// coefficients from the per-(j,k) table, plus the state-dependent terms of
// nlin_rhs / nlin_jac (usrc.F90:792-1012, spf.F90:362-790), plus the
// forward-difference convective-mixing block (mix_imp.f:729-815), then the
// mask-dependent edits of boundary.F90:2-393, then either stores the values
// (J, |a| <= 1e-10 stored as 0, assemble.F90:115) or forms
// F = -(-A(x)x - mix + Frc) in the reference's summation order (F).
//
// The solution fields are read straight from x with the ghost/boundary rules
// of usol (usrc.F90:1014-1105) restated as per-index functions, so no padded
// copies are written to HBM.  This file is compiled with -ffp-contract=off:
// every product and sum rounds exactly as in the reference build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace iemic {

struct KArgs {
    Scalars s;
    const uint8_t *lm;
    const double *al, *jt, *kt, *frc;
    double *icrow;   // the integral row's own S-equation entries (20 slots), or null
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int LMv(const KArgs &a, int i, int j, int k)
{
    return a.lm[i + (a.s.n + 2) * (j + (a.s.m + 2) * k)];
}
__device__ __forceinline__ double Xv(const double *x, const KArgs &a, int i, int j, int k, int var)
{
    return x[6 * ((i - 1) + a.s.n * ((j - 1) + a.s.m * (k - 1))) + var];
}
__device__ __forceinline__ double JT(const KArgs &a, int t, int j) { return a.jt[t * (a.s.m + 2) + j]; }
__device__ __forceinline__ double KT(const KArgs &a, int t, int k) { return a.kt[t * (a.s.l + 2) + k]; }

// u (var 0) or v (var 1) at padded index a in 0..n, b in 0..m, kk in 0..l+1
__device__ double uv_at(const double *x, const KArgs &g, int var, int a, int b, int kk)
{
    const int n = g.s.n, m = g.s.m, l = g.s.l;
    if (a < 0 || a > n || b < 1 || b > m) return 0.0;
    if (kk == 0 || kk == l + 1) {  // copies of level 1 / l made before the land zeroing
        if (a < 1) return 0.0;
        if ((!g.s.periodic && a == n) || b == m) return 0.0;
        return Xv(x, g, a, b, kk == 0 ? 1 : l, var);
    }
    double v;
    if (a >= 1) {
        v = ((!g.s.periodic && a == n) || b == m) ? 0.0 : Xv(x, g, a, b, kk, var);
    } else {
        if (!g.s.periodic) return 0.0;
        v = Xv(x, g, n, b, kk, var);  // periodic copy taken before u(n,m)=0
    }
#pragma unroll
    for (int da = 0; da <= 1; da++)
#pragma unroll
        for (int db = 0; db <= 1; db++) {
            int ca = a + da, cb = b + db;
            if (ca >= 1 && ca <= n && cb >= 1 && cb <= m && LMv(g, ca, cb, kk) == LAND) return 0.0;
        }
    return v;
}

// w at a in 0..n+1, b in 0..m+1, kk in 0..l
__device__ double w_at(const double *x, const KArgs &g, int a, int b, int kk)
{
    const int n = g.s.n, m = g.s.m, l = g.s.l;
    if (b < 1 || b > m || kk <= 0) return 0.0;
    int ai = a;
    if (a < 1 || a > n) {
        if (!g.s.periodic) return 0.0;
        ai = (a == 0) ? n : 1;  // ghost copies keep the level-l value
    } else if (kk == l) {
        return 0.0;
    }
    return Xv(x, g, ai, b, kk, W);
}

// T (var 4) or S (var 5) at a in 0..n+1, b in 0..m+1, kk in 0..l+1 (no corners)
__device__ double ts_at(const double *x, const KArgs &g, int var, int a, int b, int kk)
{
    const int n = g.s.n, m = g.s.m, l = g.s.l;
    if (kk == 0 || kk == l + 1) {
        if (a < 1 || a > n || b < 1 || b > m) return 0.0;
        return Xv(x, g, a, b, kk == 0 ? 1 : l, var);
    }
    if (b == 0 || b == m + 1) {
        if (a < 1 || a > n) return 0.0;
        return Xv(x, g, a, b == 0 ? 1 : m, kk, var);
    }
    if (a == 0) return Xv(x, g, g.s.periodic ? n : 1, b, kk, var);
    if (a == n + 1) return Xv(x, g, g.s.periodic ? 1 : n, b, kk, var);
    return Xv(x, g, a, b, kk, var);
}

// tanh by the fdlibm method (Sun Microsystems s_tanh.c / s_expm1.c) with the
// polynomial of expm1 evaluated in the split form used by GNU libm, so that
// with IEEE +,-,*,/ and no contraction the result equals the host libm's.
__device__ __forceinline__ unsigned hiword(double v) { return (unsigned)(__double_as_longlong(v) >> 32); }
__device__ __forceinline__ double with_hiword(double v, unsigned hw)
{
    long long u = __double_as_longlong(v);
    u = (u & 0xffffffffLL) | ((long long)hw << 32);
    return __longlong_as_double(u);
}
__device__ double expm1_fd(double x)
{
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10,
                 invln2 = 1.44269504088896338700e+00, Q1 = -3.33333333333331316428e-02,
                 Q2 = 1.58730158725481460165e-03, Q3 = -7.93650757867487942473e-05,
                 Q4 = 4.00821782732936239552e-06, Q5 = -2.01099218183624371326e-07;
    double y, hi, lo, c = 0.0, t, e, hxs, hfx, r1;
    int k;
    unsigned hx = hiword(x);
    int xsb = (hx & 0x80000000u) != 0;
    hx &= 0x7fffffffu;
    if (hx >= 0x4043687Au && xsb) return -1.0;
    if (hx > 0x3fd62e42u) {
        if (hx < 0x3FF0A2B2u) {
            if (!xsb) { hi = x - ln2_hi; lo = ln2_lo; k = 1; }
            else { hi = x + ln2_hi; lo = -ln2_lo; k = -1; }
        } else {
            k = (int)(invln2 * x + (xsb ? -0.5 : 0.5));
            t = k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
        c = (hi - x) - lo;
    } else if (hx < 0x3c900000u) {
        t = 1.0e300 + x;
        return x - (t - (1.0e300 + x));
    } else {
        k = 0;
    }
    hfx = 0.5 * x;
    hxs = x * hfx;
    {
        double R1 = 1.0 + hxs * Q1, h2 = hxs * hxs, R2 = Q2 + hxs * Q3, h4 = h2 * h2, R3 = Q4 + hxs * Q5;
        r1 = R1 + h2 * R2 + h4 * R3;
    }
    t = 3.0 - r1 * hfx;
    e = hxs * ((r1 - t) / (6.0 - x * t));
    if (k == 0) return x - (x * e - hxs);
    e = (x * (e - c) - c);
    e -= hxs;
    if (k == -1) return 0.5 * (x - e) - 0.5;
    if (k == 1) {
        if (x < -0.25) return -2.0 * (e - (x + 0.5));
        return 1.0 + 2.0 * (x - e);
    }
    if (k <= -2 || k > 56) {
        y = 1.0 - (e - x);
        y = with_hiword(y, hiword(y) + ((unsigned)k << 20));
        return y - 1.0;
    }
    t = 1.0;
    if (k < 20) {
        t = with_hiword(t, 0x3ff00000u - (0x200000u >> k));
        y = t - (e - x);
        y = with_hiword(y, hiword(y) + ((unsigned)k << 20));
    } else {
        t = with_hiword(t, (unsigned)(0x3ff - k) << 20);
        y = x - (e + t);
        y += 1.0;
        y = with_hiword(y, hiword(y) + ((unsigned)k << 20));
    }
    return y;
}
__device__ double tanh_fd(double x)
{
    unsigned jx = hiword(x), ix = jx & 0x7fffffffu;
    double t, z;
    if (ix >= 0x7ff00000u) return (jx >> 31) ? 1.0 / x - 1.0 : 1.0 / x + 1.0;
    if (ix < 0x40360000u) {
        if (x == 0.0) return x;
        if (ix < 0x3c800000u) return x * (1.0 + x);
        if (ix >= 0x3ff00000u) {
            t = expm1_fd(2.0 * fabs(x));
            z = 1.0 - 2.0 / (t + 2.0);
        } else {
            t = expm1_fd(-2.0 * fabs(x));
            z = -t / (t + 2.0);
        }
    } else {
        z = 1.0 - 1.0e-300;
    }
    return (jx >> 31) ? -z : z;
}

// mix_imp.f:837-856
__device__ __forceinline__ double tprstb(double grad, double fac)
{
    double q = -grad * fac;
    double r = tanh_fd(q * q * q);
    return r > 0.0 ? r : 0.0;
}

// Vertical fluxes through the top face of level kf (mix_imp.f:328-489 with
// neutral physics off): consistent (Ftz) and implicit (Fti) mixing.
struct Face {
    double Ftz, Fsz, Fti, Fsi;
};
__device__ Face face_flux(const KArgs &g, double t0, double t1, double s0, double s1, double iso,
                          double den)
{
    const Scalars &s = g.s;
    Face f;
    double dtdz = iso * (t1 - t0) / den, dsdz = iso * (s1 - s0) / den;
    double rho0 = s.lambda * s0 - t0 - s.xes * (s.alpt1 * t0 + s.alpt2 * t0 * t0 - s.alpt3 * t0 * t0 * t0);
    double rho1 = s.lambda * s1 - t1 - s.xes * (s.alpt1 * t1 + s.alpt2 * t1 * t1 - s.alpt3 * t1 * t1 * t1);
    double drdz = iso * (rho1 - rho0) / den;
    f.Ftz = f.Fsz = f.Fti = f.Fsi = 0.0;
    if (s.eps_cons != 0.0) {
        double tp = tprstb(drdz, s.fac);
        f.Ftz = 0.0 + tp * s.eps_cons * dtdz / (drdz - 1.0e-20);
        f.Fsz = 0.0 + tp * s.eps_cons * dsdz / (drdz - 1.0e-20);
    }
    if (s.kvc != 0.0) {
        double tp = tprstb(-drdz, s.fac);
        f.Fti = -tp * s.kvc * dtdz;
        f.Fsi = -tp * s.kvc * dsdz;
    }
    return f;
}

// vmix_fun divergence for the T and S rows from the faces below (f0) and
// above (f1) the cell
__device__ __forceinline__ void mix_div(const KArgs &g, int k, const Face &f0, const Face &f1,
                                        double &mt, double &ms)
{
    const Scalars &s = g.s;
    double dzk = KT(g, KT_DZDFZT, k), dzk2 = KT(g, KT_2DZDFZT, k);
    mt = 0.0;
    ms = 0.0;
    mt = (f1.Ftz - f0.Ftz) / dzk + mt;
    ms = (f1.Fsz - f0.Fsz) / dzk + ms;
    if (s.rho_mixing && s.xes == 0.0) {
        mt = ((f1.Fti - f0.Fti) - (f1.Fsi - f0.Fsi) * s.lambda) / dzk2 + mt;
        ms = ((f1.Fsi - f0.Fsi) - (f1.Fti - f0.Fti) / s.lambda) / dzk2 + ms;
    } else {
        mt = (f1.Fti - f0.Fti) / dzk + mt;
        ms = (f1.Fsi - f0.Fsi) / dzk + ms;
    }
}

__device__ __forceinline__ double isoc_v(const KArgs &g, int i, int j, int k)
{
    int v = LMv(g, i, j, k);
    return (v == OCEAN || v == PERIO) ? 1.0 : 0.0;
}

// Column data for the mixing terms of cell (i,j,k): T,S at k-1..k+1 and the
// face geometry.
struct MixCol {
    double t[3], s[3];   // levels k-1, k, k+1 (ghost copies at 0 / l+1)
    double iso[2], den[2];
    bool face[2];        // face exists (level k-1+f >= 1)
};
__device__ void mix_col(const double *x, const KArgs &g, int i, int j, int k, MixCol &c)
{
#pragma unroll
    for (int q = 0; q < 3; q++) {
        c.t[q] = ts_at(x, g, T, i, j, k - 1 + q);
        c.s[q] = ts_at(x, g, S, i, j, k - 1 + q);
    }
#pragma unroll
    for (int f = 0; f < 2; f++) {
        int kf = k - 1 + f;
        c.face[f] = kf >= 1;
        c.iso[f] = c.face[f] ? isoc_v(g, i, j, kf + 1) * isoc_v(g, i, j, kf) : 0.0;
        c.den[f] = KT(g, KT_DZDFZW, kf < 0 ? 0 : kf);
    }
}
__device__ __forceinline__ Face face_of(const KArgs &g, const MixCol &c, int f, const double *t,
                                        const double *s)
{
    if (!c.face[f]) return Face{0.0, 0.0, 0.0, 0.0};
    return face_flux(g, t[f], t[f + 1], s[f], s[f + 1], c.iso[f], c.den[f]);
}

// ------------------------------------------------------------------ masks
struct Nb {  // boundary.F90:34-78 neighbour flags (true = LAND)
    bool southw, west, nwest, south, north, southe, east, neast;
    bool southwb, westb, nwestb, southb, bottom, northb, southeb, eastb, neastb;
    bool southwt, westt, nwestt, southt, top, northt, southet, eastt, neastt;
    bool southee, easteast, northee, nnorthee, nnorth;
    bool ilt, jlt;  // i < n, j < m
};
__device__ Nb neighbours(const KArgs &g, int i, int j, int k)
{
    Nb b;
    auto L = [&](int di, int dj, int dk) { return LMv(g, i + di, j + dj, k + dk) == LAND; };
    b.southw = L(-1, -1, 0); b.west = L(-1, 0, 0); b.nwest = L(-1, 1, 0);
    b.south = L(0, -1, 0); b.north = L(0, 1, 0);
    b.southe = L(1, -1, 0); b.east = L(1, 0, 0); b.neast = L(1, 1, 0);
    b.southwb = L(-1, -1, -1); b.westb = L(-1, 0, -1); b.nwestb = L(-1, 1, -1);
    b.southb = L(0, -1, -1); b.bottom = L(0, 0, -1); b.northb = L(0, 1, -1);
    b.southeb = L(1, -1, -1); b.eastb = L(1, 0, -1); b.neastb = L(1, 1, -1);
    b.southwt = L(-1, -1, 1); b.westt = L(-1, 0, 1); b.nwestt = L(-1, 1, 1);
    b.southt = L(0, -1, 1); b.top = L(0, 0, 1); b.northt = L(0, 1, 1);
    b.southet = L(1, -1, 1); b.eastt = L(1, 0, 1); b.neastt = L(1, 1, 1);
    b.ilt = i < g.s.n;
    b.jlt = j < g.s.m;
    b.southee = b.ilt && L(2, -1, 0);
    b.easteast = b.ilt && L(2, 0, 0);
    b.northee = b.ilt && L(2, 1, 0);
    b.nnorthee = b.ilt && b.jlt && L(2, 2, 0);
    b.nnorth = b.jlt && L(0, 2, 0);  // boundary.F90 reads (i,j+2) for nnwest/nnorth/nneast
    return b;
}

// ------------------------------------------------------------------ boundary edits per row
// The edits of boundary.F90 for an OCEAN cell, restricted to the slots of one
// equation row (all edits act row-wise; additions keep the reference order).
// Slot indices are those of SLOTS within the row.
__device__ __forceinline__ void bnd_u(const Nb &b, double *a)
{   // u: 0(2U)1(2V)2(4U)3(4V)4(5U)5(5V)6(5W)7(5P)8(6U)9(6V)10(6W)11(6P)12(8U)13(8V)14(8W)15(8P)16(9W)17(9P)18(14U)19(14W)20(15W)21(17W)22(18W)23(23U)
    if (b.bottom) {
        if (b.eastb && b.neastb && b.northb) a[4] = a[4] + a[18];
        a[18] = 0.0; a[19] = 0.0;
    }
    if (b.northb) a[20] = 0.0;
    if (b.eastb) a[21] = 0.0;
    if (b.neastb) a[22] = 0.0;
    if (b.top) {
        if (b.eastt && b.neastt && b.northt) a[4] = a[4] + a[23];
        a[23] = 0.0;
        a[6] = 1.0e-10; a[10] = 1.0e-10; a[14] = 1.0e-10; a[16] = 1.0e-10;
    }
    if (b.west) { a[0] = 0.0; a[1] = 0.0; }
    if (b.nwest) { a[0] = 0.0; a[1] = 0.0; }
    if (b.south) { a[2] = 0.0; a[3] = 0.0; }
    bool ident = false;
    if (b.north) ident = true;
    else if (b.jlt && b.nnorth) { a[8] = 0.0; a[9] = 0.0; }
    if (b.southe) { a[2] = 0.0; a[3] = 0.0; }
    if (b.east) ident = true;
    else if (b.ilt && b.easteast) { a[12] = 0.0; a[13] = 0.0; }
    if (b.neast) ident = true;
    else if (b.ilt || b.jlt) {
        if (b.ilt && b.northee) { a[12] = 0.0; a[13] = 0.0; }
        if (b.jlt && b.nnorth) { a[8] = 0.0; a[9] = 0.0; }
    }
    if (ident) {
#pragma unroll
        for (int q = 0; q < 24; q++) a[q] = 0.0;
        a[4] = 1.0;
    }
}
__device__ __forceinline__ void bnd_v(const Nb &b, double *a)
{   // v: 0(2U)1(2V)2(4V)3(5U)4(5V)5(5W)6(5P)7(6V)8(6W)9(6P)10(8U)11(8V)12(8W)13(8P)14(9W)15(9P)16(14V)17(14W)18(15W)19(17W)20(18W)21(23V)
    if (b.bottom) {
        if (b.eastb && b.neastb && b.northb) a[4] = a[4] + a[16];
        a[16] = 0.0; a[17] = 0.0;
    }
    if (b.northb) a[18] = 0.0;
    if (b.eastb) a[19] = 0.0;
    if (b.neastb) a[20] = 0.0;
    if (b.top) {
        if (b.eastt && b.neastt && b.northt) a[4] = a[4] + a[21];
        a[21] = 0.0;
        a[5] = 1.0e-10; a[8] = 1.0e-10; a[12] = 1.0e-10; a[14] = 1.0e-10;
    }
    if (b.west) { a[0] = 0.0; a[1] = 0.0; }
    if (b.nwest) { a[0] = 0.0; a[1] = 0.0; }
    if (b.south) a[2] = 0.0;
    bool ident = false;
    if (b.north) ident = true;
    else if (b.jlt && b.nnorth) a[7] = 0.0;
    if (b.southe) a[2] = 0.0;
    if (b.east) ident = true;
    else if (b.ilt && b.easteast) { a[10] = 0.0; a[11] = 0.0; }
    if (b.neast) ident = true;
    else if (b.ilt || b.jlt) {
        if (b.ilt && b.northee) { a[10] = 0.0; a[11] = 0.0; }
        if (b.jlt && b.nnorth) a[7] = 0.0;
    }
    if (ident) {
#pragma unroll
        for (int q = 0; q < 22; q++) a[q] = 0.0;
        a[4] = 1.0;
    }
}
__device__ __forceinline__ void bnd_w(const Nb &b, double *a)
{   // w: 0(5W)1(5P)2(5T)3(5S)4(23P)5(23T)6(23S)
    if (b.top) {
#pragma unroll
        for (int q = 0; q < 7; q++) a[q] = 0.0;
        a[0] = 1.0;
    }
}
__device__ __forceinline__ void bnd_p(const Nb &b, double *a)
{   // p: 0(1U)1(1V)2(2U)3(2V)4(4U)5(4V)6(5U)7(5V)8(5W)9(5P)10(14W)
    if (b.bottom) a[10] = 0.0;
    if (b.top) a[8] = 1.0e-10;
    if (b.southw) { a[0] = 0.0; a[1] = 0.0; }
    if (b.west) { a[2] = 0.0; a[3] = 0.0; a[0] = 0.0; a[1] = 0.0; }
    if (b.nwest) { a[2] = 0.0; a[3] = 0.0; }
    if (b.south) { a[4] = 0.0; a[5] = 0.0; a[0] = 0.0; a[1] = 0.0; }
    if (b.north) { a[2] = 0.0; a[3] = 0.0; a[6] = 0.0; a[7] = 0.0; }
    if (b.southe) { a[4] = 0.0; a[5] = 0.0; }
    if (b.east) { a[4] = 0.0; a[5] = 0.0; a[6] = 0.0; a[7] = 0.0; }
    if (b.neast) { a[6] = 0.0; a[7] = 0.0; }
}
// T and S rows share one layout: slots 4/7/13/14 hold the row's own tracer
// at stencil positions 2/4/6/8, slot 11 is (5,T), 12 is (5,S); own_centre is
// 11 for the T row and 12 for the S row.  The centre additions follow the
// reference order: bottom, top, west, south, north, east.
__device__ __forceinline__ void bnd_ts(const Nb &b, double *a, int own_centre)
{   // 0(1U)1(1V)2(2U)3(2V)4(2own)5(4U)6(4V)7(4own)8(5U)9(5V)10(5W)11(5T)12(5S)13(6own)14(8own)15(14W)16(14T)17(14S)18(23T)19(23S)
    if (b.bottom) {
        a[11] = a[11] + a[16];
        a[12] = a[12] + a[17];
        a[15] = 0.0; a[16] = 0.0; a[17] = 0.0;
    }
    if (b.top) {
        a[11] = a[11] + a[18];
        a[12] = a[12] + a[19];
        a[18] = 0.0; a[19] = 0.0;
        a[10] = 1.0e-10;
    }
    if (b.southw) { a[0] = 0.0; a[1] = 0.0; }
    if (b.west) {
        a[own_centre] = a[own_centre] + a[4];
        a[2] = 0.0; a[3] = 0.0; a[4] = 0.0; a[0] = 0.0; a[1] = 0.0;
    }
    if (b.nwest) { a[2] = 0.0; a[3] = 0.0; }
    if (b.south) {
        a[own_centre] = a[own_centre] + a[7];
        a[5] = 0.0; a[6] = 0.0; a[7] = 0.0; a[0] = 0.0; a[1] = 0.0;
    }
    if (b.north) {
        a[2] = 0.0; a[3] = 0.0; a[8] = 0.0; a[9] = 0.0;
        a[own_centre] = a[own_centre] + a[13];
        a[13] = 0.0;
    }
    if (b.southe) { a[5] = 0.0; a[6] = 0.0; }
    if (b.east) {
        a[5] = 0.0; a[6] = 0.0; a[8] = 0.0; a[9] = 0.0;
        a[own_centre] = a[own_centre] + a[14];
        a[14] = 0.0;
    }
    if (b.neast) { a[8] = 0.0; a[9] = 0.0; }
}

// ------------------------------------------------------------------ cell state
// Every state value the six rows of cell (i,j,k) read, fetched once.
struct CellState {
    // u, v at (di,dj) in {-1,0,1}^2 and level k; u/v at k-1, k+1 of the column
    double u[3][3], v[3][3];
    double uk[2], vk[2];          // level k-1, k+1 at (i,j)
    double w0[2][2], w1[2][2];    // w at (i..i+1, j..j+1), levels k-1 (w0) and k (w1)
    double t[3][3], s[3][3];      // T,S at (di,dj) level k, cross-shaped use only
    double tk[2], sk[2];          // T,S at k-1, k+1 of the column
};
__device__ void load_state(const double *x, const KArgs &g, int i, int j, int k, CellState &q)
{
#pragma unroll
    for (int di = -1; di <= 1; di++)
#pragma unroll
        for (int dj = -1; dj <= 1; dj++) {
            bool need_uv = !(di == 1 && dj == 1);
            q.u[di + 1][dj + 1] = need_uv ? uv_at(x, g, U, i + di, j + dj, k) : 0.0;
            q.v[di + 1][dj + 1] = need_uv ? uv_at(x, g, V, i + di, j + dj, k) : 0.0;
            bool cross = (di == 0 || dj == 0);
            q.t[di + 1][dj + 1] = cross ? ts_at(x, g, T, i + di, j + dj, k) : 0.0;
            q.s[di + 1][dj + 1] = cross ? ts_at(x, g, S, i + di, j + dj, k) : 0.0;
        }
    q.uk[0] = uv_at(x, g, U, i, j, k - 1); q.uk[1] = uv_at(x, g, U, i, j, k + 1);
    q.vk[0] = uv_at(x, g, V, i, j, k - 1); q.vk[1] = uv_at(x, g, V, i, j, k + 1);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            q.w0[a][b] = w_at(x, g, i + a, j + b, k - 1);
            q.w1[a][b] = w_at(x, g, i + a, j + b, k);
        }
    q.tk[0] = ts_at(x, g, T, i, j, k - 1); q.tk[1] = ts_at(x, g, T, i, j, k + 1);
    q.sk[0] = ts_at(x, g, S, i, j, k - 1); q.sk[1] = ts_at(x, g, S, i, j, k + 1);
}
#define UQ(di, dj) q.u[(di) + 1][(dj) + 1]
#define VQ(di, dj) q.v[(di) + 1][(dj) + 1]

// ------------------------------------------------------------------ rows
// JAC selects nlin_jac (Newton linearisation) or nlin_rhs (Picard form).
template <bool JAC>
__device__ void row_u(const KArgs &g, const CellState &q, const double *al, int i, int j, int k,
                      double *a)
{
    const int n = g.s.n, m = g.s.m;
    const double epsr = g.s.epsr;
    double cx = JT(g, JT_CX_V, j), cy = JT(g, JT_CY_V, j), tdzi = KT(g, KT_TDZI8, k);
    double tanr = JT(g, JT_TAN_YV, j);
    // unlin types 1/2 (uux / Urux), 3 (uvy1), 5 (uwz), 7 (uvy2), 4 (Urvy1), 6 (Urwz), 8 (Urvy2)
    double x2 = 0.0, x8 = 0.0;
    if (i >= 2) x2 = JAC ? -2 * UQ(-1, 0) * cx : -UQ(-1, 0) * cx;
    if (i <= n - 1) x8 = JAC ? 2 * UQ(1, 0) * cx : UQ(1, 0) * cx;
    double vy4 = (j >= 2) ? -VQ(0, -1) * JT(g, JT_COS_YV, j - 1) * cy : 0.0;
    double vy6 = (j <= m - 1) ? VQ(0, 1) * JT(g, JT_COS_YV, j + 1) * cy : 0.0;
    double wz23 = (q.w1[0][0] + q.w1[0][1] + q.w1[1][0] + q.w1[1][1]) * tdzi;
    double wz14 = -(q.w0[0][0] + q.w0[0][1] + q.w0[1][0] + q.w0[1][1]) * tdzi;
    double wz5 = wz14 + wz23;
    double vy2 = VQ(0, 0) * tanr;
#pragma unroll
    for (int s = 0; s < 24; s++) a[s] = al[s];
    a[0] = al[0] + epsr * x2;
    a[2] = al[2] + epsr * vy4;
    a[4] = al[4] + epsr * (wz5 + vy2);
    a[8] = al[8] + epsr * vy6;
    a[12] = al[12] + epsr * x8;
    a[18] = al[18] + epsr * wz14;
    a[23] = al[23] + epsr * wz23;
    if (JAC) {
        double uy4 = (j >= 2) ? -UQ(0, -1) * JT(g, JT_COS_YV, j - 1) * cy : 0.0;
        double uy6 = (j <= m - 1) ? UQ(0, 1) * JT(g, JT_COS_YV, j + 1) * cy : 0.0;
        double uy2 = UQ(0, 0) * tanr;
        a[3] = al[3] + epsr * uy4;
        a[5] = al[5] + epsr * uy2;
        a[9] = al[9] + epsr * uy6;
        double up = (UQ(0, 0) + q.uk[1]) * tdzi, dn = -(UQ(0, 0) + q.uk[0]) * tdzi;
        a[6] = al[6] + epsr * up; a[10] = al[10] + epsr * up;
        a[14] = al[14] + epsr * up; a[16] = al[16] + epsr * up;
        a[19] = al[19] + epsr * dn; a[20] = al[20] + epsr * dn;
        a[21] = al[21] + epsr * dn; a[22] = al[22] + epsr * dn;
    }
}

template <bool JAC>
__device__ void row_v(const KArgs &g, const CellState &q, const double *al, int i, int j, int k,
                      double *a)
{
    const int n = g.s.n, m = g.s.m;
    const double epsr = g.s.epsr;
    double cx = JT(g, JT_CX_V, j), cy = JT(g, JT_CY_V, j), tdzi = KT(g, KT_TDZI8, k);
    double tanr = JT(g, JT_TAN_YV, j);
    double ux2 = (i >= 2) ? -UQ(-1, 0) * cx : 0.0;           // vnlin 1 uvx
    double ux8 = (i <= n - 1) ? UQ(1, 0) * cx : 0.0;
    double vy4, vy6;                                          // vnlin 3 vvry / 4 Vrvy
    if (JAC) {
        vy6 = (j <= m - 1) ? 2 * VQ(0, 1) * JT(g, JT_COS_YV, j + 1) * cy : 0.0;
        vy4 = (j >= 2) ? -2 * VQ(0, -1) * JT(g, JT_COS_YV, j - 1) * cy : 0.0;
    } else {
        vy6 = (j <= m - 1) ? VQ(0, 1) * JT(g, JT_COS_YV, j + 1) * cy : 0.0;
        vy4 = (j >= 2) ? -VQ(0, -1) * JT(g, JT_COS_YV, j - 1) * cy : 0.0;
    }
    double wz23 = (q.w1[0][0] + q.w1[0][1] + q.w1[1][0] + q.w1[1][1]) * tdzi;   // vnlin 5 vwz
    double wz14 = -(q.w0[0][0] + q.w0[0][1] + q.w0[1][0] + q.w0[1][1]) * tdzi;
    double wz5 = wz14 + wz23;
#pragma unroll
    for (int s = 0; s < 22; s++) a[s] = al[s];
    if (JAC) {  // VV,UU += epsr (Urt2 + uVrx)
        double vx2 = (i >= 2) ? -VQ(-1, 0) * cx : 0.0, vx8 = (i <= n - 1) ? VQ(1, 0) * cx : 0.0;
        a[0] = al[0] + epsr * vx2;
        a[3] = al[3] + epsr * (2 * UQ(0, 0) * tanr);
        a[10] = al[10] + epsr * vx8;
        double up = (VQ(0, 0) + q.vk[1]) * tdzi, dn = -(VQ(0, 0) + q.vk[0]) * tdzi;
        a[5] = al[5] + epsr * up; a[8] = al[8] + epsr * up;
        a[12] = al[12] + epsr * up; a[14] = al[14] + epsr * up;
        a[17] = al[17] + epsr * dn; a[18] = al[18] + epsr * dn;
        a[19] = al[19] + epsr * dn; a[20] = al[20] + epsr * dn;
    } else {    // VV,UU += epsr ut2 (vnlin 7: u tan)
        a[3] = al[3] + epsr * (UQ(0, 0) * tanr);
    }
    a[1] = al[1] + epsr * ux2;
    a[2] = al[2] + epsr * vy4;
    a[4] = al[4] + epsr * wz5;
    a[7] = al[7] + epsr * vy6;
    a[11] = al[11] + epsr * ux8;
    a[16] = al[16] + epsr * wz14;
    a[21] = al[21] + epsr * wz23;
}

template <bool JAC>
__device__ void row_w(const KArgs &g, const CellState &q, const double *al, int k, double *a)
{
#pragma unroll
    for (int s = 0; s < 7; s++) a[s] = al[s];
    double t0 = q.t[1][1], t1 = q.tk[1];
    double r5 = 0.0, r23 = 0.0, c5 = 0.0, c23 = 0.0;
    if (k <= g.s.l - 1) {
        if (JAC) {   // wnlin 1, 3
            r23 = (t0 + t1) / 2.; r5 = (t0 + t1) / 2.;
            c5 = 0.375 * ((t0 + t1) * (t0 + t1)); c23 = 0.375 * ((t0 + t1) * (t0 + t1));
        } else {     // wnlin 2, 4
            r23 = t1 / 4.; r5 = (t0 + 2 * t1) / 4.;
            c5 = 0.125 * (t0 * t0 + 3 * t1 * t0 + 3 * t1 * t1);
            c23 = 0.125 * t1 * t1;
        }
    }
    const Scalars &s = g.s;
    a[2] = al[2] - s.Ra * s.xes * s.alpt2 * r5 + s.Ra * s.xes * s.alpt3 * c5;
    a[5] = al[5] - s.Ra * s.xes * s.alpt2 * r23 + s.Ra * s.xes * s.alpt3 * c23;
}

// T (OWN = T) or S (OWN = S) row, before the mixing block
template <bool JAC, int OWN>
__device__ void row_ts(const KArgs &g, const CellState &q, const double *al, int j, int k, double *a)
{
    const double(&tq)[3][3] = (OWN == T) ? q.t : q.s;
    const double *tk = (OWN == T) ? q.tk : q.sk;
    double cx = JT(g, JT_CX_T, j), cy = JT(g, JT_CY_T, j);
    double cvm = JT(g, JT_COS_YV, j - 1), cv = JT(g, JT_COS_YV, j);
    double tdzi = g.s.tdzi2, dfz = KT(g, KT_DFZT, k);
#pragma unroll
    for (int s = 0; s < 20; s++) a[s] = al[s];
    // tnlin 3 Utrx, 5 Vtry, 7 Wtrz (T-T); tnlin 2 urTx, 4 vrTy, 6 wrTz (Jacobian only)
    double U2 = -(UQ(-1, 0) + UQ(-1, -1)) * cx, U8 = (UQ(0, 0) + UQ(0, -1)) * cx, U5 = U2 + U8;
    double V4 = -(VQ(0, -1) + VQ(-1, -1)) * cy * cvm, V6 = (VQ(0, 0) + VQ(-1, 0)) * cy * cv, V5 = V4 + V6;
    double W14 = -q.w0[0][0] * tdzi / dfz, W23 = q.w1[0][0] * tdzi / dfz, W5 = W14 + W23;
    const int o2 = 4, o4 = 7, o6 = 13, o8 = 14, oc = (OWN == T) ? 11 : 12;
    const int o14 = (OWN == T) ? 16 : 17, o23 = (OWN == T) ? 18 : 19;
    a[o2] = al[o2] + U2;
    a[o4] = al[o4] + V4;
    a[oc] = al[oc] + U5 + V5 + W5;
    a[o6] = al[o6] + V6;
    a[o8] = al[o8] + U8;
    a[o14] = al[o14] + W14;
    a[o23] = al[o23] + W23;
    if (JAC) {
        double tc = tq[1][1];
        double r2 = -(tc + tq[0][1]) * cx, r4 = (tq[2][1] + tc) * cx;          // urTx
        a[2] = al[2] + r2; a[0] = al[0] + r2; a[5] = al[5] + r4; a[8] = al[8] + r4;
        double s4 = -cy * (tc + tq[1][0]) * cvm, s5 = cy * (tq[1][2] + tc) * cv; // vrTy
        a[6] = al[6] + s4; a[1] = al[1] + s4; a[9] = al[9] + s5; a[3] = al[3] + s5;
        double w14 = -tdzi * (tc + tk[0]) / dfz;                              // wrTz
        double w5 = (k <= g.s.l - 1) ? tdzi * (tk[1] + tc) / dfz : 0.0;
        a[15] = al[15] + w14; a[10] = al[10] + w5;
    }
}
#undef UQ
#undef VQ

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ bool nb_cell(const Scalars &s, int i, int j, int k, int kk, int &cell)
{
    int ni = i + kk_di(kk), nj = j + kk_dj(kk), nk = k + kk_dk(kk);
    if (s.periodic) ni = (ni < 1) ? ni + s.n : (ni > s.n ? ni - s.n : ni);
    if (ni < 1 || ni > s.n || nj < 1 || nj > s.m || nk < 1 || nk > s.l) return false;
    cell = (ni - 1) + s.n * ((nj - 1) + s.m * (nk - 1));
    return true;
}

template <int R0, int LEN>
__device__ __forceinline__ void store_row(double *__restrict__ J, size_t C, int cell, const double *a)
{
#pragma unroll
    for (int q = 0; q < LEN; q++) {
        double v = a[q];
        J[(size_t)(R0 + q) * C + cell] = (fabs(v) > 1.0e-10) ? v : 0.0;
    }
}

// sum A(x)x over one row in slot order, entries |a| <= 1e-10 left out
// (assemble.F90:115, matetc.F90:147-166)
template <int R0, int LEN>
__device__ __forceinline__ double row_dot(const Scalars &s, const double *__restrict__ x, int i, int j,
                                          int k, const double *a)
{
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < LEN; q++) {
        double v = a[q];
        if (fabs(v) > 1.0e-10) {
            int cell;
            if (nb_cell(s, i, j, k, SLOTS[R0 + q].kk, cell)) acc = v * x[6 * cell + SLOTS[R0 + q].var] + acc;
        }
    }
    return acc;
}

struct FdMix {   // forward-difference entries of the T and S rows (mix_imp.f:750-813)
    double T[3][2], S[3][2];   // [level k-1,k,k+1][column T,S]
    bool lev[3];
};
__device__ void mixing_block(const double *x, const KArgs &g, int i, int j, int k, FdMix &fd,
                             double &bT, double &bS)
{
    MixCol mc;
    mix_col(x, g, i, j, k, mc);
    Face f0 = face_of(g, mc, 0, mc.t, mc.s), f1 = face_of(g, mc, 1, mc.t, mc.s);
    mix_div(g, k, f0, f1, bT, bS);
    const double eps = 1.0e-08;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int kk = k - 1 + d, v = LMv(g, i, j, kk);
        fd.lev[d] = (v == OCEAN || v == PERIO);
#pragma unroll
        for (int var = 0; var < 2; var++) {
            fd.T[d][var] = 0.0;
            fd.S[d][var] = 0.0;
            if (!fd.lev[d]) continue;
            double tt[3] = {mc.t[0], mc.t[1], mc.t[2]}, ss[3] = {mc.s[0], mc.s[1], mc.s[2]};
            if (var == 0) tt[d] = tt[d] + eps; else ss[d] = ss[d] + eps;
            Face p0 = (d <= 1) ? face_of(g, mc, 0, tt, ss) : f0;
            Face p1 = (d >= 1) ? face_of(g, mc, 1, tt, ss) : f1;
            double pT, pS;
            mix_div(g, k, p0, p1, pT, pS);
            fd.T[d][var] = (pT - bT) / eps;
            fd.S[d][var] = (pS - bS) / eps;
        }
    }
}

__global__ void __launch_bounds__(256) k_jacobian(KArgs g, const double *__restrict__ x,
                                                  double *__restrict__ J)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.s.ncell) return;
    const int n = g.s.n, m = g.s.m;
    const int i = cell % n + 1, j = (cell / n) % m + 1, k = cell / (n * m) + 1;
    const size_t C = (size_t)g.s.ncell;
    if (LMv(g, i, j, k) != OCEAN) {   // boundary.F90:379-386: identity rows
#pragma unroll
        for (int s = 0; s < NSLOT; s++)
            J[(size_t)s * C + cell] = (SLOTS[s].eq == SLOTS[s].var && SLOTS[s].kk == 5) ? 1.0 : 0.0;
        return;
    }
    const Nb b = neighbours(g, i, j, k);
    CellState q;
    load_state(x, g, i, j, k, q);
    const double *al = g.al + ((size_t)(k - 1) * m + (j - 1)) * NSLOT;
    double a[24];
    row_u<true>(g, q, al + ROW0[U], i, j, k, a); bnd_u(b, a); store_row<0, 24>(J, C, cell, a);
    row_v<true>(g, q, al + ROW0[V], i, j, k, a); bnd_v(b, a); store_row<24, 22>(J, C, cell, a);
    row_w<true>(g, q, al + ROW0[W], k, a);       bnd_w(b, a); store_row<46, 7>(J, C, cell, a);
#pragma unroll
    for (int s = 0; s < 11; s++) a[s] = al[ROW0[P] + s];
    bnd_p(b, a); store_row<53, 11>(J, C, cell, a);
    FdMix fd;
    double bT = 0.0, bS = 0.0;
    const bool mixj = g.s.vmix_on && (g.s.mix_t || g.s.mix_sj);
    if (mixj) mixing_block(x, g, i, j, k, fd, bT, bS);
    // T row: mixing entries at (14|5|23, T|S) = slots 16|11|18 and 17|12|19
    const int sl[3][2] = {{16, 17}, {11, 12}, {18, 19}};
    row_ts<true, T>(g, q, al + ROW0[T], j, k, a);
    if (mixj && g.s.mix_t)
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (fd.lev[d]) {
                a[sl[d][0]] = a[sl[d][0]] + fd.T[d][0];
                if (g.s.mix_ts) a[sl[d][1]] = a[sl[d][1]] + fd.T[d][1];
            }
    bnd_ts(b, a, 11); store_row<64, 20>(J, C, cell, a);
    row_ts<true, S>(g, q, al + ROW0[S], j, k, a);
    if (mixj && g.s.mix_sj)
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (fd.lev[d]) {
                a[sl[d][0]] = a[sl[d][0]] + fd.S[d][0];
                a[sl[d][1]] = a[sl[d][1]] + fd.S[d][1];
            }
    bnd_ts(b, a, 12);
    if (6 * cell + S == g.s.rowintcon) {   // replaced by the integral row
        if (g.icrow)   // the Fortran S equation of this row, kept for the Fortran-layout export
#pragma unroll
            for (int s = 0; s < 20; s++) g.icrow[s] = (fabs(a[s]) > 1.0e-10) ? a[s] : 0.0;
#pragma unroll
        for (int s = 0; s < 20; s++) a[s] = 0.0;
    }
    store_row<84, 20>(J, C, cell, a);
}

__global__ void __launch_bounds__(256) k_rhs(KArgs g, const double *__restrict__ x,
                                             double *__restrict__ F)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.s.ncell) return;
    const int n = g.s.n, m = g.s.m;
    const int i = cell % n + 1, j = (cell / n) % m + 1, k = cell / (n * m) + 1;
    double *Fc = F + 6 * (size_t)cell;
    if (LMv(g, i, j, k) != OCEAN) {   // B(row) * (1 - landm) = 0 (usrc.F90:580-591)
#pragma unroll
        for (int e = 0; e < 6; e++) Fc[e] = 0.0;
        return;
    }
    const Nb b = neighbours(g, i, j, k);
    CellState q;
    load_state(x, g, i, j, k, q);
    const double *al = g.al + ((size_t)(k - 1) * m + (j - 1)) * NSLOT;
    const double *frc = g.frc + 6 * (size_t)cell;
    double mT = 0.0, mS = 0.0;
    if (g.s.vmix_on && (g.s.mix_t || g.s.mix_sf)) {
        MixCol mc;
        mix_col(x, g, i, j, k, mc);
        Face f0 = face_of(g, mc, 0, mc.t, mc.s), f1 = face_of(g, mc, 1, mc.t, mc.s);
        mix_div(g, k, f0, f1, mT, mS);
        if (!g.s.mix_t) mT = 0.0;   // vmix_fun's rows are 0.0 while switched off
        if (!g.s.mix_sf) mS = 0.0;
    }
    double a[24], au;
    // B = -Au - mix + Frc (usrc.F90:576), F = -B (THCM.C:1011)
    row_u<false>(g, q, al + ROW0[U], i, j, k, a); bnd_u(b, a);
    au = row_dot<0, 24>(g.s, x, i, j, k, a);
    Fc[U] = -((-au - 0.0) + frc[U]);
    row_v<false>(g, q, al + ROW0[V], i, j, k, a); bnd_v(b, a);
    au = row_dot<24, 22>(g.s, x, i, j, k, a);
    Fc[V] = -((-au - 0.0) + frc[V]);
    row_w<false>(g, q, al + ROW0[W], k, a); bnd_w(b, a);
    au = row_dot<46, 7>(g.s, x, i, j, k, a);
    Fc[W] = -((-au - 0.0) + frc[W]);
#pragma unroll
    for (int s = 0; s < 11; s++) a[s] = al[ROW0[P] + s];
    bnd_p(b, a);
    au = row_dot<53, 11>(g.s, x, i, j, k, a);
    Fc[P] = -((-au - 0.0) + frc[P]);
    row_ts<false, T>(g, q, al + ROW0[T], j, k, a); bnd_ts(b, a, 11);
    au = row_dot<64, 20>(g.s, x, i, j, k, a);
    Fc[T] = -((-au - mT) + frc[T]);
    row_ts<false, S>(g, q, al + ROW0[S], j, k, a); bnd_ts(b, a, 12);
    au = row_dot<84, 20>(g.s, x, i, j, k, a);
    Fc[S] = -((-au - mS) + frc[S]);
}

// F[rowintcon] = intSign (coef . x - intCorrection) (THCM.C:1013-1027); the
// dot product is formed by a fixed-order block reduction.
__global__ void __launch_bounds__(256) k_dot_partial(const double *__restrict__ a,
                                                     const double *__restrict__ b, int n,
                                                     double *__restrict__ part)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        acc += a[q] * b[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_dot_final(const double *__restrict__ part, int np, double *__restrict__ out,
                            double scale, double shift, int slot)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int q = threadIdx.x; q < np; q += blockDim.x) acc += part[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = scale * (sh[0] - shift);
}

// y = J x in slot format; the integral row is added by k_dot_* afterwards.
__global__ void __launch_bounds__(256) k_spmv(Scalars s, const double *__restrict__ J,
                                              const double *__restrict__ x, double *__restrict__ y)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const int i = cell % s.n + 1, j = (cell / s.n) % s.m + 1, k = cell / (s.n * s.m) + 1;
    const size_t C = (size_t)s.ncell;
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        int nb;
        if (nb_cell(s, i, j, k, SLOTS[sl].kk, nb))
            acc[SLOTS[sl].eq] += J[(size_t)sl * C + cell] * x[6 * nb + SLOTS[sl].var];
    }
#pragma unroll
    for (int e = 0; e < 6; e++) y[6 * (size_t)cell + e] = acc[e];
}

// Y = J X on nrhs columns (column q at X + q ldx): each thread reads the
// slots of its cell once for NC columns.  The integral row is added per
// column afterwards, as in k_spmv.
template <int NC>
__global__ void __launch_bounds__(256) k_spmm(Scalars s, const double *__restrict__ J, int nrhs,
                                              const double *__restrict__ X, size_t ldx, double *__restrict__ Y,
                                              size_t ldy)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const int q0 = blockIdx.y * NC;
    const int i = cell % s.n + 1, j = (cell / s.n) % s.m + 1, k = cell / (s.n * s.m) + 1;
    const size_t C = (size_t)s.ncell;
    double acc[6][NC];
#pragma unroll
    for (int e = 0; e < 6; e++)
#pragma unroll
        for (int q = 0; q < NC; q++) acc[e][q] = 0.0;
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        int nb;
        if (nb_cell(s, i, j, k, SLOTS[sl].kk, nb)) {
            const double a = J[(size_t)sl * C + cell];
            const size_t col = 6 * (size_t)nb + SLOTS[sl].var;
#pragma unroll
            for (int q = 0; q < NC; q++)
                if (q0 + q < nrhs) acc[SLOTS[sl].eq][q] += a * X[(size_t)(q0 + q) * ldx + col];
        }
    }
#pragma unroll
    for (int q = 0; q < NC; q++)
        if (q0 + q < nrhs)
#pragma unroll
            for (int e = 0; e < 6; e++) Y[(size_t)(q0 + q) * ldy + 6 * (size_t)cell + e] = acc[e][q];
}

// y = J^T x: each cell gathers the entries of its six columns from the
// neighbours that reference them; the integral row contributes
// intSign coef(c) x[rowintcon] to every S column.
__global__ void __launch_bounds__(256) k_spmv_t(Scalars s, const double *__restrict__ J,
                                                const double *__restrict__ coef, double sign,
                                                const double *__restrict__ x, double *__restrict__ y)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= s.ncell) return;
    const int i = cell % s.n + 1, j = (cell / s.n) % s.m + 1, k = cell / (s.n * s.m) + 1;
    const size_t C = (size_t)s.ncell;
    double acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int sl = 0; sl < NSLOT; sl++) {
        int kk = SLOTS[sl].kk;
        // the row owner is the cell at -offset(kk) from this column cell
        int oi = i - kk_di(kk), oj = j - kk_dj(kk), ok = k - kk_dk(kk);
        if (s.periodic) oi = (oi < 1) ? oi + s.n : (oi > s.n ? oi - s.n : oi);
        if (oi < 1 || oi > s.n || oj < 1 || oj > s.m || ok < 1 || ok > s.l) continue;
        int owner = (oi - 1) + s.n * ((oj - 1) + s.m * (ok - 1));
        acc[SLOTS[sl].var] += J[(size_t)sl * C + owner] * x[6 * owner + SLOTS[sl].eq];
    }
    if (s.rowintcon >= 0) acc[S] += sign * coef[6 * (size_t)cell + S] * x[s.rowintcon];
#pragma unroll
    for (int e = 0; e < 6; e++) y[6 * (size_t)cell + e] = acc[e];
}

__global__ void k_export(const int *__restrict__ src, const int *__restrict__ col, int nnz,
                         const double *__restrict__ J, const double *__restrict__ coef,
                         double sign, double *__restrict__ val)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    int s = src[p];
    val[p] = (s >= 0) ? J[s] : sign * coef[col[p]];
}

// Ocean::pressureProjection (Ocean.C:2180-2235): v -= (v.s1) s1 + (v.s2) s2,
// s1 / s2 the normalised indicators of the pressure unknowns of the wet
// cells with i + j even / odd (the checkerboard modes, THCM.C:2802-2841).
// With s = e / sqrt(count), (v.s) s is the mean of v over the class, so the
// sums and counts of the two classes are reduced (fixed order) and each
// class mean subtracted.
__global__ void __launch_bounds__(256) k_pressure_partial(Scalars s, const uint8_t *__restrict__ lm,
                                                          const double *__restrict__ v,
                                                          double *__restrict__ part)
{
    __shared__ double sh[4][256];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.ncell; c += gridDim.x * blockDim.x) {
        const int i = c % s.n, j = (c / s.n) % s.m, k = c / (s.n * s.m);
        if (lm[(i + 1) + (s.n + 2) * ((j + 1) + (s.m + 2) * (k + 1))] != OCEAN) continue;
        const int cls = (i + j) & 1;
        acc[cls] += v[6 * c + 3];
        acc[2 + cls] += 1.0;
    }
    for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}
__global__ void __launch_bounds__(256) k_pressure_final(const double *__restrict__ part, int np,
                                                        double *__restrict__ mean)
{
    __shared__ double sh[4][256];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < np; b += blockDim.x)
        for (int q = 0; q < 4; q++) acc[q] += part[4 * b + q];
    for (int q = 0; q < 4; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 2) mean[threadIdx.x] = sh[2 + threadIdx.x][0] > 0.0 ? sh[threadIdx.x][0] / sh[2 + threadIdx.x][0] : 0.0;
}
__global__ void __launch_bounds__(256) k_pressure_update(Scalars s, const uint8_t *__restrict__ lm,
                                                         const double *__restrict__ mean, double *__restrict__ v)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= s.ncell) return;
    const int i = c % s.n, j = (c / s.n) % s.m, k = c / (s.n * s.m);
    if (lm[(i + 1) + (s.n + 2) * ((j + 1) + (s.m + 2) * (k + 1))] != OCEAN) return;
    v[6 * c + 3] -= mean[(i + j) & 1];
}

// ------------------------------------------------------------------ host side
static KArgs kargs(iemic_ctx *c)
{
    KArgs g;
    g.s = c->sc;
    g.lm = c->d_landm;
    g.al = c->d_al;
    g.jt = c->d_jt;
    g.kt = c->d_kt;
    g.frc = c->d_frc;
    g.icrow = c->d_icrow;
    return g;
}

int upload_tables(iemic_ctx *c)
{
    const Host &h = c->h;
    size_t nl = h.landm.size();
    if (!c->d_landm) {
        HIPCHK(hipMalloc(&c->d_landm, nl));
        std::vector<uint8_t> lm8(nl);
        for (size_t q = 0; q < nl; q++) lm8[q] = (uint8_t)h.landm[q];
        HIPCHK(hipMemcpy(c->d_landm, lm8.data(), nl, hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&c->d_al, h.al.size() * sizeof(double)));
        HIPCHK(hipMalloc(&c->d_jt, h.jt.size() * sizeof(double)));
        HIPCHK(hipMalloc(&c->d_kt, h.kt.size() * sizeof(double)));
        HIPCHK(hipMalloc(&c->d_frc, h.frc.size() * sizeof(double)));
        HIPCHK(hipMalloc(&c->d_coef, h.coef.size() * sizeof(double)));
        HIPCHK(hipMalloc(&c->d_diagB, h.diagB.size() * sizeof(double)));
        c->nred = 1024;
        HIPCHK(hipMalloc(&c->d_red, (c->nred + 8) * sizeof(double)));
        if (c->sc.rowintcon >= 0) HIPCHK(hipMalloc(&c->d_icrow, 20 * sizeof(double)));
        HIPCHK(hipMemcpy(c->d_coef, h.coef.data(), h.coef.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->d_al, h.al.data(), h.al.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_jt, h.jt.data(), h.jt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_kt, h.kt.data(), h.kt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_frc, h.frc.data(), h.frc.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_diagB, h.diagB.data(), h.diagB.size() * sizeof(double), hipMemcpyHostToDevice));
    return IEMIC_OK;
}

static inline int nblk(int n, int b) { return (n + b - 1) / b; }

static void timed_begin(iemic_ctx *c, int which)
{
    if (!c->ev[which][0]) {
        (void)hipEventCreate(&c->ev[which][0]);
        (void)hipEventCreate(&c->ev[which][1]);
    }
    (void)hipEventRecord(c->ev[which][0], c->stream);
}
static void timed_end(iemic_ctx *c, int which) { (void)hipEventRecord(c->ev[which][1], c->stream); }
void timed_begin_ext(iemic_ctx *c, int which) { timed_begin(c, which); }
void timed_end_ext(iemic_ctx *c, int which) { timed_end(c, which); }

// dot product of two device vectors on the context stream; result into
// out[slot] as scale*(dot - shift)
static int dot_into(iemic_ctx *c, const double *a, const double *b, int n, double *out, double scale,
                    double shift, int slot)
{
    int nb = std::min(c->nred, nblk(n, 256));
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, c->stream, a, b, n, c->d_red);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, c->stream, c->d_red, nb, out, scale, shift, slot);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

// Sums of squares of the T and S unknowns (l2nrm of vmix_control,
// matetc.F90:22-36), one pair of partials per block.
__global__ void __launch_bounds__(256) k_ts_sumsq(const double *__restrict__ x, int ncell, double *__restrict__ part)
{
    __shared__ double st[256], ss[256];
    double at = 0.0, as = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        const double t = x[6 * (size_t)c + T], s = x[6 * (size_t)c + S];
        at += t * t;
        as += s * s;
    }
    st[threadIdx.x] = at;
    ss[threadIdx.x] = as;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            st[threadIdx.x] += st[threadIdx.x + w];
            ss[threadIdx.x] += ss[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = st[0];
        part[2 * blockIdx.x + 1] = ss[0];
    }
}

// Mixing = 2: the control that Ocean::computeRHS / computeJacobian run before
// every evaluation (THCM::fixMixing(0), then vmix_control, mix_imp.f:131-166):
// the T / S switches follow whether the T / S fields are nonzero (2-norm
// above 1e-12); when they change and T is on, the pattern is rebuilt with the
// new switches (vmix_part with vmix_el_2, mix_imp.f:168-227, 934-1049).
// d_part holds cap doubles of device scratch; one host synchronisation.
int mixing_control(iemic_ctx *c, const double *d_x, double *d_part, int cap)
{
    if (c->vmix_flag != 2) return IEMIC_OK;
    constexpr int MAXB = 512;
    const int nb = std::min(std::min(MAXB, cap / 2), nblk(c->sc.ncell, 256));
    if (nb < 1) return set_error(IEMIC_EINVAL, "mixing control: no scratch for the field norms");
    double h[2 * MAXB];
    hipLaunchKernelGGL(k_ts_sumsq, dim3(nb), dim3(256), 0, c->stream, d_x, c->sc.ncell, d_part);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, d_part, 2 * (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double st = 0.0, ss = 0.0;
    for (int b = 0; b < nb; b++) {
        st += h[2 * b];
        ss += h[2 * b + 1];
    }
    const int tt = std::sqrt(st) > 1.0e-12 ? 1 : 0, ts = std::sqrt(ss) > 1.0e-12 ? 1 : 0;
    if (tt != c->mix2_temp || ts != c->mix2_salt) {
        c->mix2_temp = tt;
        c->mix2_salt = ts;
        if (tt != 0) {
            c->mix2_part = 1;
            c->mix2_pat_salt = ts;
        }
    }
    set_mix_switches(c);
    return IEMIC_OK;
}

int launch_rhs(iemic_ctx *c, const double *d_x, double *d_F)
{
    if (int rc = mixing_control(c, d_x, c->d_red, c->nred)) return rc;
    KArgs g = kargs(c);
    timed_begin(c, 0);
    hipLaunchKernelGGL(k_rhs, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, g, d_x, d_F);
    timed_end(c, 0);
    HIPCHK(hipGetLastError());
    if (c->sc.rowintcon >= 0)
        return dot_into(c, c->d_coef, d_x, c->sc.ndim, d_F, (double)c->cfg.int_sign, c->intcorr,
                        c->sc.rowintcon);
    return IEMIC_OK;
}

// F(x) with the integral row's reduction partials in a caller-provided
// device buffer of cap doubles (one partial per 256 unknowns), no timing
// events and no allocation: the kernels of launch_rhs, one partial per block.
int launch_rhs_partials(iemic_ctx *c, const double *d_x, double *d_F, double *d_part, int cap)
{
    if (int rc = mixing_control(c, d_x, d_part, cap)) return rc;
    const int nb = nblk(c->sc.ndim, 256);
    if (c->sc.rowintcon >= 0 && nb > cap)
        return set_error(IEMIC_EINVAL, "reduction needs " + std::to_string(nb) + " partials, the buffer holds " +
                                           std::to_string(cap));
    KArgs g = kargs(c);
    hipLaunchKernelGGL(k_rhs, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, g, d_x, d_F);
    HIPCHK(hipGetLastError());
    if (c->sc.rowintcon >= 0) {
        hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(256), 0, c->stream, c->d_coef, d_x, c->sc.ndim, d_part);
        hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, c->stream, d_part, nb, d_F, (double)c->cfg.int_sign,
                           c->intcorr, c->sc.rowintcon);
        HIPCHK(hipGetLastError());
    }
    return IEMIC_OK;
}

// rhs_ (usrc.F90:523-603) in the Fortran's own convention: B = -F, every
// row the model equation (the integral row keeps its S equation).
int launch_rhs_fortran(iemic_ctx *c, const double *d_x, double *d_B)
{
    if (int rc = mixing_control(c, d_x, c->d_red, c->nred)) return rc;
    KArgs g = kargs(c);
    hipLaunchKernelGGL(k_rhs, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, g, d_x, d_B);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

// The Jacobian in the layout fillcolA writes (assemble.F90:57-139): rows in
// order, each row's entries in stencil position then unknown order -- the
// order of its slots -- keeping |a| > 1e-10 (every nonzero slot value),
// 1-based.  The integral row holds its S equation (the Fortran matrix;
// THCM.C replaces that row only in its own copy).
__device__ __forceinline__ double fort_value(const Scalars &s, const double *J, const double *icrow, int cell, int sl)
{
    if (icrow && 6 * cell + S == s.rowintcon && SLOTS[sl].eq == S) return icrow[sl - ROW0[S]];
    return J[(size_t)sl * s.ncell + cell];
}
__global__ void __launch_bounds__(256) k_fort_count(Scalars s, const double *__restrict__ J,
                                                    const double *__restrict__ icrow, int *__restrict__ cnt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.ndim) return;
    const int cell = r / 6, e = r % 6;
    int c = 0;
    for (int sl = ROW0[e]; sl < ROW0[e + 1]; sl++) c += fort_value(s, J, icrow, cell, sl) != 0.0;
    cnt[r] = c;
}
__global__ void __launch_bounds__(256) k_fort_fill(Scalars s, const double *__restrict__ J,
                                                   const double *__restrict__ icrow, const int *__restrict__ beg,
                                                   int *__restrict__ col, double *__restrict__ val,
                                                   int *__restrict__ bad)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= s.ndim) return;
    const int cell = r / 6, e = r % 6;
    const int i = cell % s.n + 1, j = (cell / s.n) % s.m + 1, k = cell / (s.n * s.m) + 1;
    int q = beg[r];
    for (int sl = ROW0[e]; sl < ROW0[e + 1]; sl++) {
        const double v = fort_value(s, J, icrow, cell, sl);
        if (v == 0.0) continue;
        int nb;
        if (!nb_cell(s, i, j, k, SLOTS[sl].kk, nb)) {   // a value outside the grid cannot be placed
            atomicAdd(bad, 1);
            nb = cell;
        }
        col[q] = 6 * nb + SLOTS[sl].var + 1;
        val[q] = v;
        q++;
    }
}

int launch_fortran_count(iemic_ctx *c, int *d_cnt)
{
    hipLaunchKernelGGL(k_fort_count, dim3(nblk(c->sc.ndim, 256)), dim3(256), 0, c->stream, c->sc, c->d_J,
                       c->d_icrow, d_cnt);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}
int launch_fortran_fill(iemic_ctx *c, const int *d_beg, int *d_col, double *d_val, int *d_bad)
{
    hipLaunchKernelGGL(k_fort_fill, dim3(nblk(c->sc.ndim, 256)), dim3(256), 0, c->stream, c->sc, c->d_J,
                       c->d_icrow, d_beg, d_col, d_val, d_bad);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

int launch_jacobian(iemic_ctx *c, const double *d_x)
{
    if (int rc = mixing_control(c, d_x, c->d_red, c->nred)) return rc;
    if (c->sc.mix_t && !c->sc.mix_ts)
        // the reference differences the T columns in the column groups that
        // its DSM colouring builds for this pattern, and the salinity columns
        // it leaves out of the pattern are perturbed with those groups: the
        // entries then depend on that colouring, which is not restated here
        return set_error(IEMIC_EUNSUPPORTED,
                         "Mixing = 2 Jacobian at a state whose mixing pattern was built with a zero salinity field "
                         "(its entries depend on the reference's DSM column grouping)");
    if (!c->d_J) HIPCHK(hipMalloc(&c->d_J, (size_t)NSLOT * c->sc.ncell * sizeof(double)));
    KArgs g = kargs(c);
    timed_begin(c, 1);
    hipLaunchKernelGGL(k_jacobian, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, g, d_x, c->d_J);
    timed_end(c, 1);
    HIPCHK(hipGetLastError());
    c->have_jac = true;
    c->jac_version++;
    return IEMIC_OK;
}

int launch_spmv(iemic_ctx *c, const double *d_x, double *d_y)
{
    timed_begin(c, 2);
    hipLaunchKernelGGL(k_spmv, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc, c->d_J, d_x, d_y);
    timed_end(c, 2);
    HIPCHK(hipGetLastError());
    if (c->sc.rowintcon >= 0)
        return dot_into(c, c->d_coef, d_x, c->sc.ndim, d_y, (double)c->cfg.int_sign, 0.0, c->sc.rowintcon);
    return IEMIC_OK;
}

static void spmm_launch(iemic_ctx *c, int nc, int nrhs, const double *d_X, size_t ldx, double *d_Y, size_t ldy)
{
    const dim3 g(nblk(c->sc.ncell, 256), (nrhs + nc - 1) / nc);
    if (nc == 1) hipLaunchKernelGGL(k_spmm<1>, g, dim3(256), 0, c->stream, c->sc, c->d_J, nrhs, d_X, ldx, d_Y, ldy);
    else if (nc == 2) hipLaunchKernelGGL(k_spmm<2>, g, dim3(256), 0, c->stream, c->sc, c->d_J, nrhs, d_X, ldx, d_Y, ldy);
    else hipLaunchKernelGGL(k_spmm<4>, g, dim3(256), 0, c->stream, c->sc, c->d_J, nrhs, d_X, ldx, d_Y, ldy);
}

// Columns per thread of the SpMM (1, 2 or 4: more columns read the slots
// fewer times but hold 6 NC accumulators over the unrolled stencil), timed
// on the real operands at the first product with a given column count (Y
// is simply rewritten) unless IEMIC_NO_TUNE is set.
static int spmm_tuned_nc(iemic_ctx *c, int nrhs, const double *d_X, size_t ldx, double *d_Y, size_t ldy)
{
    auto it = c->spmm_nc.find(nrhs);
    if (it != c->spmm_nc.end()) return it->second;
    if (nrhs == 1 || std::getenv("IEMIC_NO_TUNE") || d_X == d_Y) return c->spmm_nc[nrhs] = 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 1;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 1;
    }
    int best = 1;
    float tbest = 1e30f;
    for (int nc : {1, 2, 4}) {
        float tmin = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            (void)hipEventRecord(e0, c->stream);
            spmm_launch(c, nc, nrhs, d_X, ldx, d_Y, ldy);
            (void)hipEventRecord(e1, c->stream);
            float ms = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) break;
            tmin = std::min(tmin, ms);
        }
        if (tmin < tbest) {
            tbest = tmin;
            best = nc;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return c->spmm_nc[nrhs] = best;
}

int launch_spmm(iemic_ctx *c, int nrhs, const double *d_X, size_t ldx, double *d_Y, size_t ldy)
{
    const int nc = spmm_tuned_nc(c, nrhs, d_X, ldx, d_Y, ldy);
    timed_begin(c, 2);
    spmm_launch(c, nc, nrhs, d_X, ldx, d_Y, ldy);
    timed_end(c, 2);
    HIPCHK(hipGetLastError());
    if (c->sc.rowintcon >= 0)
        for (int q = 0; q < nrhs; q++) {
            int rc = dot_into(c, c->d_coef, d_X + (size_t)q * ldx, c->sc.ndim, d_Y + (size_t)q * ldy,
                              (double)c->cfg.int_sign, 0.0, c->sc.rowintcon);
            if (rc) return rc;
        }
    return IEMIC_OK;
}

// Y = J^T X on nrhs columns, one transposed product per column
int launch_spmm_t(iemic_ctx *c, int nrhs, const double *d_X, size_t ldx, double *d_Y, size_t ldy)
{
    for (int q = 0; q < nrhs; q++) {
        int rc = launch_spmv_t(c, d_X + (size_t)q * ldx, d_Y + (size_t)q * ldy);
        if (rc) return rc;
    }
    return IEMIC_OK;
}

int launch_spmv_t(iemic_ctx *c, const double *d_x, double *d_y)
{
    hipLaunchKernelGGL(k_spmv_t, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc, c->d_J,
                       c->d_coef, (double)c->cfg.int_sign, d_x, d_y);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

int launch_pressure_projection(iemic_ctx *c, double *d_v)
{
    // partials: 4 per block in d_red[0 .. 4 nb); class means after them
    const int nb = std::min(c->nred / 4 - 1, nblk(c->sc.ncell, 256));
    double *mean = c->d_red + 4 * nb;
    hipLaunchKernelGGL(k_pressure_partial, dim3(nb), dim3(256), 0, c->stream, c->sc, c->d_landm, d_v, c->d_red);
    hipLaunchKernelGGL(k_pressure_final, dim3(1), dim3(256), 0, c->stream, c->d_red, nb, mean);
    hipLaunchKernelGGL(k_pressure_update, dim3(nblk(c->sc.ncell, 256)), dim3(256), 0, c->stream, c->sc,
                       c->d_landm, mean, d_v);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

int launch_export(iemic_ctx *c, double *d_val)
{
    hipLaunchKernelGGL(k_export, dim3(nblk(c->graph.nnz, 256)), dim3(256), 0, c->stream, c->d_src,
                       c->d_col_export, c->graph.nnz, c->d_J, c->d_coef, (double)c->cfg.int_sign, d_val);
    HIPCHK(hipGetLastError());
    return IEMIC_OK;
}

double device_dot(iemic_ctx *c, const double *d_a, const double *d_b, int n)
{
    if (dot_into(c, d_a, d_b, n, c->d_red + c->nred, 1.0, 0.0, 0) != IEMIC_OK) return NAN;
    double r = 0.0;
    if (hipMemcpyAsync(&r, c->d_red + c->nred, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return NAN;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return NAN;
    return r;
}

}  // namespace iemic
