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
    int visc;          // viscous wall fluxes (Fv) subtracted from the face fluxes
};

// Device pointers of one case
struct Fields {
    double *U, *U1, *LU;               // E * N each
    double *rho, *u, *v, *w, *p, *H, *T;
    double *c, *gamma, *e;             // full primitives (download / checks)
    double *hi;                        // NS * N species enthalpy at T (COP)
    double *Cp, *gcv, *R;              // mixture Cp, Cp/Cv, R (COP)
    double *red;                       // reduction slots, see RED_*
    double *dt;                        // device-resident dt
    int *err;                          // [0] code bits, [1] min cell id
    int *substeps;                     // chemistry predictor passes per interior cell
    unsigned int *queue;               // chemistry work-queue counter
    // viscous terms: Wilke viscosity and conductivity, mixture-averaged
    // D_km (NS planes), centre velocity gradients (9 planes), and the
    // viscous face flux of the direction being swept (E planes)
    double *mu, *kap, *Dkm, *Vde, *Fv;
    double *ypart; // per-block interior min / max of each mass fraction (2 NS per block)
    double *spart; // per-block GLF and CFL maxima of k_states (12 per block)
};

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

// error code bits in err[0]
constexpr int EBIT_RHO_Y = 1;
constexpr int EBIT_PRIM = 2;
constexpr int EBIT_NAN = 4;
constexpr int EBIT_CHEM = 8;

// universal gas constant (src/include/global_setup.h:48-54), folded at
// compile time exactly as the reference's const expressions
constexpr double R0 = 6.02214076e26 * 1.380649e-23;
constexpr double Ru = 6.02214076e26 * 1.380649e-23 * 1.0E-3;

} // namespace xf
