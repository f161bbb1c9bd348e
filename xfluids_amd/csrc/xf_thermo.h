// Species and mixture thermodynamics on the device (FP64).
//
// Same formulas and association order as the reference's
// src/solver_Ini/Thermo_device.h (NASA-9 three-range and JANAF/NASA-7
// two-range fits, T >= 200 K clamp, linear enthalpy extrapolation below
// 200 K) and src/solver_Ini/Mixing_device.h (mixture sums, damped Newton
// temperature inversion).  Restructured for the GPU: the fit range, 1/T and
// log(T) depend only on T and are evaluated once per temperature instead of
// once per species (identical values), and loops over species are unrolled
// for a compile-time species count NS.
#pragma once
#include "xf_internal.h"
#include "xf_dev_math.h"

namespace xf {

constexpr double OT = 1.0 / 3.0; // _OT, src/include/global_setup.h:35

// Per-temperature quantities shared by all species.
struct TPre {
    double TT;   // unclamped temperature
    double T;    // max(T, 200)
    double iT;   // 1/T
    double lT;   // log(T) (only when requested)
    int r;       // fit range index (Hia[.. * 3 + r], Hc[(r * MAXNS + n) * 9 + ..])
};

template <bool NEED_LOG>
__device__ __forceinline__ TPre tpre(double T0, int thermo)
{
    TPre t;
    t.TT = T0;
    t.T = dmax(T0, 200.0);
    t.iT = 1.0 / t.T;
    t.lT = NEED_LOG ? dlog(t.T) : 0.0;
    if (thermo) {
        // NASA-9: 200-1000 -> 0, 1000-6000 -> 1, >= 6000 -> 2 (Thermo_device.h:15-20)
        t.r = (t.T >= 1000.0 && t.T < 6000.0) ? 1 : ((t.T < 1000.0) ? 0 : ((t.T >= 6000.0) ? 2 : -1));
    } else {
        // JANAF: > 1000 -> 0 (high fit), else 1 (Thermo_device.h:34-37)
        t.r = (t.T > 1000.0) ? 0 : 1;
    }
    return t;
}

// The fit formulas on the nine coefficients c of one fit range of one
// species: Hia[n * 21 + m * 3 + r] (m = 0..6) then Hib[n * 6 + m * 3 + r]
// (m = 0, 1), as the regrouped copy Hc[(r * MAXNS + n) * 9 + ...] holds them.
// c comes from the table (cp_k, h_k, s_k) or from registers a wave loaded
// once (the one-wave-per-cell chemistry): the same expressions either way.

// HeatCapacity_NASA / HeatCapacity_JANAF (Thermo_device.h:10-40)
__device__ __forceinline__ double cp_fit(const double *c, int thermo, const TPre &t, double Ri)
{
    const double T = t.T;
    if (thermo) {
        if (t.r < 0) return 0.0;
        return Ri * ((c[0] * t.iT + c[1]) * t.iT + c[2] + (c[3] + (c[4] + (c[5] + c[6] * T) * T) * T) * T);
    }
    return Ri * (c[0] + (c[1] + (c[2] + (c[3] + c[4] * T) * T) * T) * T);
}

// get_Enthalpy_NASA / get_Enthalpy_JANAF (Thermo_device.h:62-102) without the
// extrapolation below 200 K (h_k adds it)
__device__ __forceinline__ double h_fit(const double *c, int thermo, const TPre &t, double Ri)
{
    const double T = t.T;
    double hi = 0.0;
    if (thermo) {
        if (t.r >= 0)
            hi = Ri * (-c[0] / T + c[1] * t.lT + (c[2] + (0.5 * c[3] + (c[4] * OT + (0.25 * c[5] + 0.2 * c[6] * T) * T) * T) * T) * T +
                       c[7]);
    } else {
        hi = Ri * (T * (c[0] + T * (c[1] * 0.5 + T * (c[2] * OT + T * (c[3] * 0.25 + c[4] * T * 0.2)))) + c[5]);
    }
    return hi;
}

// get_Entropy_NASA / get_Entropy_JANAF (Thermo_device.h:132-161)
__device__ __forceinline__ double s_fit(const double *c, int thermo, const TPre &t, double Ri)
{
    const double T = t.T;
    if (thermo) {
        if (t.r < 0) return 0.0;
        return Ri * ((-0.5 * c[0] * t.iT - c[1]) * t.iT + c[2] * t.lT +
                     (c[3] + (0.5 * c[4] + (c[5] * OT + c[6] * 0.25 * T) * T) * T) * T + c[8]);
    }
    return Ri * (c[0] * t.lT + (c[1] + (0.5 * c[2] + (c[3] * OT + c[4] * 0.25 * T) * T) * T) * T + c[6]);
}

// species n's coefficients of range t.r in the table (range 0 when t.r < 0,
// where the NASA formulas return before reading them)
__device__ __forceinline__ const double *fit_coeffs(const DevThermo *__restrict__ th, int r, int n)
{
    return th->Hc + ((r < 0 ? 0 : r) * MAXNS + n) * 9;
}

__device__ __forceinline__ double cp_k(const DevThermo *__restrict__ th, const TPre &t, double Ri, int n)
{
    return cp_fit(fit_coeffs(th, t.r, n), th->thermo, t, Ri);
}

// Cp of species n at exactly 200 K (used by the low-temperature extrapolation)
__device__ __forceinline__ double cp200_k(const DevThermo *__restrict__ th, double Ri, int n)
{
    TPre t;
    t.TT = 200.0;
    t.T = 200.0;
    t.iT = 1.0 / t.T;
    t.lT = 0.0;
    t.r = th->thermo ? 0 : 1;
    return cp_k(th, t, Ri, n);
}

__device__ __forceinline__ double h_k(const DevThermo *__restrict__ th, const TPre &t, double Ri, int n)
{
    double hi = h_fit(fit_coeffs(th, t.r, n), th->thermo, t, Ri);
    if (t.TT < 200.0) {
        const double Cpi = cp200_k(th, Ri, n);
        hi += Cpi * (t.TT - 200.0);
    }
    return hi;
}

__device__ __forceinline__ double s_k(const DevThermo *__restrict__ th, const TPre &t, double Ri, int n)
{
    return s_fit(fit_coeffs(th, t.r, n), th->thermo, t, Ri);
}

// get_CopR (Mixing_device.h:49-56)
template <int NS>
__device__ __forceinline__ double mix_R(const DevThermo *__restrict__ th, const double *yi)
{
    double R = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) R += yi[n] * th->_Wi[n];
    return R * Ru;
}

// get_CopCp (Mixing_device.h:61-68)
template <int NS>
__device__ __forceinline__ double mix_cp(const DevThermo *__restrict__ th, const double *yi, const TPre &t)
{
    double cp = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) cp += yi[n] * cp_k(th, t, th->Ri[n], n);
    return cp;
}

// get_CopCv (Mixing_device.h:73-86)
template <int NS>
__device__ __forceinline__ double mix_cv(const DevThermo *__restrict__ th, const double *yi, const TPre &t)
{
    double cv = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) cv += yi[n] * cp_k(th, t, th->Ri[n], n);
    double _W = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) _W += yi[n] * th->_Wi[n];
    cv -= Ru * _W;
    return cv;
}

// get_CopGamma(thermal, yi, Cp, T) (Mixing_device.h:114-126)
template <int NS>
__device__ __forceinline__ double mix_gamma_cp(const DevThermo *__restrict__ th, const double *yi, double Cp)
{
    double _W = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) _W += yi[n] * th->_Wi[n];
    const double CopW = 1.0 / _W;
    const double g = Cp / (Cp - Ru / CopW);
    return (g > 1.0) ? g : -1.0;
}

// sub_FuncT + get_T (Mixing_device.h:159-193).  R = get_CopR is loop
// invariant and hoisted (same value every iteration).
template <int NS>
__device__ __forceinline__ double get_T(const DevThermo *__restrict__ th, const double *yi, double e, double T0)
{
    double T = T0;
    const double R = mix_R<NS>(th, yi);
    const int thermo = th->thermo;
    for (int it = 1; it < 101; it++) {
        const TPre t = thermo ? tpre<true>(T, 1) : tpre<false>(T, 0);
        double h = 0.0;
        double Cp = 0.0;
        // the same sums with the fit range as a wave-uniform value when every
        // active lane uses one range (scalar coefficient loads)
        auto sums = [&](const TPre &tt) {
#pragma unroll
            for (int n = 0; n < NS; n++) {
                const double Ri = th->Ri[n];
                h += h_k(th, tt, Ri, n) * yi[n];
            }
#pragma unroll
            for (int n = 0; n < NS; n++) Cp += yi[n] * cp_k(th, tt, th->Ri[n], n);
        };
        const int r0 = __builtin_amdgcn_readfirstlane(t.r);
        if (__all(t.r == r0)) {
            TPre tu = t;
            tu.r = r0;
            sums(tu);
        } else {
            sums(t);
        }
        const double func_T = h - R * T - e;
        const double dfunc_T = Cp - R;
        double df = dmin(func_T / (dfunc_T + 1.0e-30), 1e-3 * T);
        df = dmax(df, -1e-2 * T);
        T = T - df;
        if (fabs(df) <= 1.0e-6) break;
    }
    return T;
}

} // namespace xf
