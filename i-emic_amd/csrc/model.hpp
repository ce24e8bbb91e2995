// Internal definitions shared by the host set-up code and the HIP kernels.
//
// Jacobian storage ("slot format").  Every grid cell owns the 104 entries of
// its six rows on the fixed maximal graph of THCM::CreateMaximalGraph
// (THCM.C:2300-2580): 24 (u), 22 (v), 7 (w), 11 (p), 20 (T), 20 (S).  Within a
// row the slots are ordered like the Fortran assembly visits them
// (assemble.F90:107-133: stencil position kk ascending, then unknown jj
// ascending), so the residual kernel sums A(u)u in the reference's order.
// Values are stored slot-major, val[s * ncell + cell], so a wavefront of
// consecutive cells reads and writes each slot with unit stride.  Columns are
// implicit: slot s of cell (i,j,k) refers to unknown var(s) of the neighbour
// (i+di, j+dj, k+dk) given by stencil position kk (periodic wrap in i).
#pragma once
#include <cstdint>

namespace iemic {

constexpr int NSLOT = 104;
enum Var { U = 0, V = 1, W = 2, P = 3, T = 4, S = 5 };
enum Mask : uint8_t { OCEAN = 0, LAND = 1, WATER = 2, PERIO = 3 };

struct SlotDef {
    int8_t eq, var, kk;  // equation row, unknown, stencil position 1..27
};

// first slot of each equation row inside a cell
constexpr int ROW0[7] = {0, 24, 46, 53, 64, 84, 104};

// stencil position kk (1..27) -> neighbour offset (mat.F90 / assemble.F90:142-177)
__host__ __device__ constexpr int kk_di(int kk) { return ((kk - 1) % 9) / 3 - 1; }
__host__ __device__ constexpr int kk_dj(int kk) { return ((kk - 1) % 9) % 3 - 1; }
__host__ __device__ constexpr int kk_dk(int kk) { return (kk - 1) / 9 == 0 ? 0 : ((kk - 1) / 9 == 1 ? -1 : 1); }

constexpr SlotDef SLOTS[NSLOT] = {
    // u-equation (24)
    {0, U, 2}, {0, V, 2}, {0, U, 4}, {0, V, 4}, {0, U, 5}, {0, V, 5}, {0, W, 5}, {0, P, 5},
    {0, U, 6}, {0, V, 6}, {0, W, 6}, {0, P, 6}, {0, U, 8}, {0, V, 8}, {0, W, 8}, {0, P, 8},
    {0, W, 9}, {0, P, 9}, {0, U, 14}, {0, W, 14}, {0, W, 15}, {0, W, 17}, {0, W, 18}, {0, U, 23},
    // v-equation (22)
    {1, U, 2}, {1, V, 2}, {1, V, 4}, {1, U, 5}, {1, V, 5}, {1, W, 5}, {1, P, 5}, {1, V, 6},
    {1, W, 6}, {1, P, 6}, {1, U, 8}, {1, V, 8}, {1, W, 8}, {1, P, 8}, {1, W, 9}, {1, P, 9},
    {1, V, 14}, {1, W, 14}, {1, W, 15}, {1, W, 17}, {1, W, 18}, {1, V, 23},
    // w-equation (7)
    {2, W, 5}, {2, P, 5}, {2, T, 5}, {2, S, 5}, {2, P, 23}, {2, T, 23}, {2, S, 23},
    // continuity (11)
    {3, U, 1}, {3, V, 1}, {3, U, 2}, {3, V, 2}, {3, U, 4}, {3, V, 4}, {3, U, 5}, {3, V, 5},
    {3, W, 5}, {3, P, 5}, {3, W, 14},
    // T-equation (20)
    {4, U, 1}, {4, V, 1}, {4, U, 2}, {4, V, 2}, {4, T, 2}, {4, U, 4}, {4, V, 4}, {4, T, 4},
    {4, U, 5}, {4, V, 5}, {4, W, 5}, {4, T, 5}, {4, S, 5}, {4, T, 6}, {4, T, 8}, {4, W, 14},
    {4, T, 14}, {4, S, 14}, {4, T, 23}, {4, S, 23},
    // S-equation (20)
    {5, U, 1}, {5, V, 1}, {5, U, 2}, {5, V, 2}, {5, S, 2}, {5, U, 4}, {5, V, 4}, {5, S, 4},
    {5, U, 5}, {5, V, 5}, {5, W, 5}, {5, T, 5}, {5, S, 5}, {5, S, 6}, {5, S, 8}, {5, W, 14},
    {5, T, 14}, {5, S, 14}, {5, T, 23}, {5, S, 23},
};

// Per-latitude coefficient tables handed to the kernels (index j = 0..m+1,
// Fortran j).  Each entry is formed on the host with the expression of the
// reference routine named in its comment, so device products round alike.
enum JTab {
    JT_CX_V = 0,   // 1/(2 cos(yv) dx)            spf.F90 unlin/vnlin types 1,2
    JT_CY_V,       // 1/(2 cos(yv) dy)            unlin/vnlin types 3,4
    JT_COS_YV,     // cos(yv)                     (j = 0..m)
    JT_TAN_YV,     // tan(yv)                     unlin/vnlin types 7,8
    JT_CX_T,       // 1/(4 cos(y) dx)             tnlin types 2,3
    JT_CY_T,       // 1/(4 cos(y) dy)             tnlin types 4,5
    NJT
};
enum KTab {
    KT_TDZI8 = 0,  // 1/(8 dfzT dz)               unlin/vnlin types 5,6
    KT_DFZT,       // dfzT                        tnlin types 6,7
    KT_DZDFZW,     // dz dfzW                     mix_imp.f dCdzt
    KT_DZDFZT,     // dz dfzT                     vmix_fun divergence
    KT_2DZDFZT,    // 2 dz dfzT                   vmix_fun (rho mixing)
    NKT
};

// Scalars of one evaluation (continuation parameters and derived constants).
struct Scalars {
    int n, m, l, periodic;
    int ncell, ndim;
    int rowintcon;        // -1 if no integral condition
    int vmix_on;          // vmix_flag >= 1 with T and S active
    // switches of the mixing terms (all 1 under Mixing = 1; under Mixing = 2
    // set before every F and J by the state-dependent control, mix_imp.f:131-227):
    int mix_t;            //   T rows: mixing in F, and J's entries against T
    int mix_ts;           //   T rows: J's entries against S
    int mix_sf;           //   S rows: mixing in F
    int mix_sj;           //   S rows: J's entries against T and S
    int rho_mixing;
    double epsr, Ra, xes, lambda;
    double alpt2, alpt3;
    double tdzi2;         // 1/(2 dz)                 tnlin types 6,7
    double kvc;           // P_VC implicit vertical mixing
    double eps_cons;      // (1-ALPC) ENER PE_V consistent mixing
    double fac;           // alphaT * SPL1 (tprstb)
    double alpt1;
};

}  // namespace iemic
