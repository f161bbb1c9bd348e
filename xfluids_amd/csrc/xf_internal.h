// Internal structures shared by the host launcher (xf_api.hip) and the
// kernels (xf_kernels.hip).  Device layout is structure-of-arrays over the
// padded grid: field n of cell id lives at ptr[n * ncell + id].
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include "../../include/xf_api.h"

namespace xf {

constexpr int MAXNS = XF_MAXNS;
constexpr int MAXNR = XF_MAXNR;
constexpr int MAXE = XF_MAXNS + 4;

// Thermal tables (reference layout, src/include/global_setup.h:202-208)
struct DevThermo {
    double Hia[MAXNS * 21];
    double Hib[MAXNS * 6];
    // the same coefficients regrouped per (fit range r, species n): Hia's
    // seven then Hib's two, Hc[(r * MAXNS + n) * 9 + m], so that one range of
    // one species is 72 contiguous bytes (two scalar loads when r is
    // wave-uniform, instead of nine strided ones)
    double Hc[3 * MAXNS * 9];
    double Wi[MAXNS], _Wi[MAXNS], Ri[MAXNS];
    int thermo; // 1 NASA-9, 0 JANAF
};

// One reaction as the rate kernel reads it: Arrhenius A, B and
// Ea * 4.184 / Ru (the reference's left-to-right product, hoisted), bit masks
// of the species with non-zero Nu_d / third-body efficiency / Nu_f / Nu_b,
// and those coefficients packed in 4-bit fields (Nu_d stored + 8).  One
// scalar load of 64 bytes per reaction replaces NS table reads per list.
struct DevRxn {
    double A, B, EaR;
    unsigned md, mt, mf, mb;
    unsigned long long pd, pf, pb;
    int third_ind, pad;
};

// Reaction tables (reference layout, src/include/global_setup.h:210-215)
// plus the per-reaction records built from them
struct DevMech {
    double Rargus[MAXNR * 6];
    double third[MAXNR * MAXNS];
    int Nu_f[MAXNR * MAXNS], Nu_b[MAXNR * MAXNS], Nu_d[MAXNR * MAXNS];
    int third_ind[MAXNR];
    int NR;
    DevRxn rx[MAXNR];
};

// Geometry passed by value to every kernel
struct Geom {
    int Xmax, Ymax, Zmax;
    int X_inner, Y_inner, Z_inner;
    int Bx, By, Bz;
    int DimX, DimY, DimZ;
    double _dx, _dy, _dz, CFL;
    long long N; // padded cells
    double dx, dy, dz; // cell sizes (WENO-CU6 epsilon = 1e-8 dl^2)
    int positivity;    // positivity-preserving flux limiter on
    int visc;          // viscous wall fluxes subtracted from the face fluxes
    int ghost;         // GhostSpecies mixture: mass fractions from the y planes
    // viscous terms: heat conduction and species diffusion switches, the
    // diffusion limiters, and whether the sweeps compute each face's wall
    // flux themselves (1) or read it from the Fv planes (0)
    int visc_heat, visc_diffu, visc_fused;
    double Yil_limiter, Dim_limiter;
};

// Device pointers of one case
struct Fields {
    double *U, *U1, *LU;               // E * N each
    double *rho, *u, *v, *w, *p, *H, *T;
    double *c, *gamma, *e;             // full primitives (download / checks)
    double *hi;                        // NS * N species enthalpy at T (COP)
    double *Cp, *gcv, *R;              // mixture Cp, Cp/Cv, R (COP)
    double *red;                       // reduction slots, see RED_*
    double *dt;                        // device-resident dt and clock, see DT_*
    int *err;                          // error and progress words, see ERR_*
    int *substeps;                     // chemistry predictor passes per interior cell
    unsigned int *queue;               // chemistry work-queue counter
    // viscous terms: Wilke viscosity and conductivity, mixture-averaged
    // D_km (NS planes), centre velocity gradients (9 planes), and the
    // viscous face flux of the direction being swept (E planes)
    double *mu, *kap, *Dkm, *Vde, *Fv;
    double *ypart; // per-block interior min / max of each mass fraction (2 NS per block)
    // GhostSpecies mixtures only: the renormalised mass fractions of the
    // first NS - 1 species (NS - 1 planes) as the last UpdateStates formed
    // them; the ghost species' own is 0
    double *y;
    double *spart; // per-block GLF and CFL maxima of k_states (12 per block)
};

// One case of an ensemble as the ensemble's launches see it: geometry,
// fields, the stage-2 buffer, the case's first block in a states launch and
// in a sweep launch of each direction, its tile grid per direction and its
// packed ghost-fill codes (bit 28 set, as k_states takes them).
struct EnsCase {
    Geom g;
    Fields f;
    double *U2;
    long long state_off;
    long long tile_off[3];
    int tiles[3][3];
    int bcs;
    // 0 = U, 1 = U1, 2 = U2
    __device__ __forceinline__ double *buf(int sel) const { return sel == 0 ? f.U : (sel == 1 ? f.U1 : U2); }
};

// the case whose block range holds block b of a states launch (kind 3) or
// of a sweep launch in direction kind (cases in ascending block order)
__device__ __forceinline__ int ens_find(const EnsCase *c, int n, long long b, int kind)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const long long o = kind < 3 ? c[mid].tile_off[kind] : c[mid].state_off;
        if (o <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// reduction slot layout (doubles, compared as uint64 bit patterns)
// GLF running maxima per direction d: RED_GLF + 3*d + {0: |u-c|, 1: |u|, 2: |u+c|}
//   (never reset: the reference zeroes eigen_block once, src/Fluids.cpp:427,
//   and every later sycl::reduction combines with the stored value)
// GLF maxima of the last UpdateStates pass: RED_GLFP + 3*d + {...}; folded
//   into RED_GLF when GetLU runs (the Reaction's UpdateStates does not feed GLF)
// CFL maxima of the last UpdateStates pass per direction d: RED_DT + d
constexpr int RED_GLF = 0;
constexpr int RED_GLFP = 9;
constexpr int RED_DT = 18;
// uvw_c_max of the last GetDt (the CFL maxima k_dt used), read by the
// positivity-preserving limiter: RED_LAM + d
constexpr int RED_LAM = 21;
// interior min / max of each mass fraction for the diffusion limiter:
// RED_YMIN + k, RED_YMAX + k
constexpr int RED_YMIN = 32;
constexpr int RED_YMAX = RED_YMIN + MAXNS;
constexpr int RED_SIZE = RED_YMAX + MAXNS;

// error code bits in err[ERR_BITS]
constexpr int EBIT_RHO_Y = 1;
constexpr int EBIT_PRIM = 2;
constexpr int EBIT_NAN = 4;
constexpr int EBIT_CHEM = 8;

// error and progress words (ints).  A pass that finds a bad cell sets its
// bit in ERR_BITS, the lowest failing cell id of that bit in ERR_ID + b
// (b = 0 rho/Y, 1 primitives, 2 NaN check, 3 chemistry), and the step and
// stage it ran in.  Every kernel of a step returns at entry once an earlier
// launch has failed (ERR_HALT, latched by the one-lane and fold kernels
// between passes, or a bit that the kernel itself never sets), so the state
// stays as the failing pass left it, as the reference stops its time loop
// there (src/XFLUIDS.cpp:211-225, 285-287).  A kernel never consults bits its
// own launch may set, so the recorded cell does not depend on block order.
constexpr int ERR_BITS = 0;
constexpr int ERR_ID = 1;    // 4 slots
constexpr int ERR_STEP = 5;  // step counter (Iteration) of the failing pass
constexpr int ERR_WHERE = 6; // 1..3 RK stage, 4 Reaction
constexpr int ERR_HALT = 7;  // the step kernels return at entry
constexpr int ERR_COUNT = 8; // steps begun, counted by k_dt in a step
constexpr int ERR_DONE = 9;  // the target time was reached
constexpr int ERR_SIZE = 16;

// device-resident dt and clock (doubles)
constexpr int DT_STEP = 0;   // dt of the current step
constexpr int DT_CHEM = 1;   // chemistry step (dt/2 Strang, dt Lie)
constexpr int DT_STAGED = 2; // dt of xf_rk3_stage
constexpr int DT_TIME = 4;   // physical time after the current step's dt
constexpr int DT_TARGET = 5; // output time the step's dt is clamped to (+inf: none)
constexpr int DT_SIZE = 8;

// Getrhoyi's mass fractions of cell id (Update_device.hpp:7-29) for the
// kernels that follow an UpdateStates: Y_k = U_{5+k} / rho with the last
// species the remainder 1 - sum; for a GhostSpecies mixture the renormalised
// values that UpdateStates stored, and 0 for the ghost.  u5(k) gives U_{5+k}
// of the cell, r1 = 1 / rho.
template <int NS, class U5>
__device__ __forceinline__ void cell_y(const Geom &g, const Fields &f, long long id, double r1, U5 u5, double *y)
{
    constexpr int NCOP = NS - 1;
    if (g.ghost) {
        y[NCOP] = 0.0;
#pragma unroll
        for (int k = 0; k < NCOP; k++) y[k] = f.y[k * g.N + id];
    } else {
        y[NCOP] = 1.0;
#pragma unroll
        for (int k = 0; k < NCOP; k++) y[k] = u5(k) * r1, y[NCOP] += -y[k];
    }
}

// the step kernels' entry test: a halted step, or a failure recorded by an
// earlier launch (bits in mask are those the calling kernel never sets)
__device__ __forceinline__ bool step_halted(const int *err, int mask)
{
    return err[ERR_HALT] != 0 || (err[ERR_BITS] & mask) != 0;
}

// record a failing cell: the bit, the lowest id per bit, the step and stage
__device__ __forceinline__ void report_cell(int *err, int bit, long long id)
{
    const int slot = bit == EBIT_RHO_Y ? 0 : (bit == EBIT_PRIM ? 1 : (bit == EBIT_NAN ? 2 : 3));
    atomicMin(&err[ERR_ID + slot], (int)id);
}
__device__ __forceinline__ void report_pass(int *err, int bits, int where)
{
    atomicOr(&err[ERR_BITS], bits);
    atomicExch(&err[ERR_STEP], err[ERR_COUNT]);
    atomicExch(&err[ERR_WHERE], where);
}

// universal gas constant (src/include/global_setup.h:48-54), folded at
// compile time exactly as the reference's const expressions
constexpr double R0 = 6.02214076e26 * 1.380649e-23;
constexpr double Ru = 6.02214076e26 * 1.380649e-23 * 1.0E-3;

} // namespace xf
