/*
 * Species and mixture thermodynamics (test infrastructure only).
 *
 * Restates src/solver_Ini/Thermo_device.h and src/solver_Ini/Mixing_device.h
 * operation by operation (same association order, same clamps).
 *
 * Pins (tests/test_oracle_thermo.py):
 *  - Cp/R, H/RT, S/R against a direct numpy evaluation of the published
 *    NASA-9 and NASA-7 polynomial forms at many temperatures in each range;
 *  - Cp = dH/dT (central difference) and Cp continuity of the fits;
 *  - below 200 K: H(T) = H(200) + Cp(200) (T - 200) (linear extrapolation);
 *  - g = S - H/T identity, e = h - R T identity;
 *  - get_T round trip: T(e(T)) = T to the Newton tolerance over 250..4000 K;
 *  - gamma > 1 and gamma = Cp/(Cp - R) for the mixture.
 */
#include "ora_internal.h"

/* universal gas constant: src/include/global_setup.h:48-54 */
#define ORA_AVOGADRO 6.02214076e26
#define ORA_BOLTZMANN 1.380649e-23
static const double ora_R0 = ORA_AVOGADRO * ORA_BOLTZMANN;          /* universal_gas_const */
static const double ora_Ru = ORA_AVOGADRO * ORA_BOLTZMANN * 1.0E-3; /* Ru */
static const double ora_OT = 1.0 / 3.0;                              /* _OT */

double ora_universal_gas_const(void) { return ora_R0; }
double ora_Ru_const(void) { return ora_Ru; }

#define HA(m, r) Hia[n * 7 * 3 + (m) * 3 + (r)]
#define HB(m, r) Hib[n * 2 * 3 + (m) * 3 + (r)]

/* HeatCapacity_NASA: src/solver_Ini/Thermo_device.h:10-23 */
static double cp_nasa(const double *Hia, double T0, double Ri, int n)
{
    double T = ora_max(T0, 200.0);
    double Cpi = 0.0, _T = 1.0 / T;
    int r = -1;
    if (T >= 1000.0 && T < 6000.0) r = 1;
    else if (T < 1000.0) r = 0;
    else if (T >= 6000.0) r = 2;
    if (r >= 0)
        Cpi = Ri * ((HA(0, r) * _T + HA(1, r)) * _T + HA(2, r) +
                    (HA(3, r) + (HA(4, r) + (HA(5, r) + HA(6, r) * T) * T) * T) * T);
    return Cpi;
}

/* HeatCapacity_JANAF: src/solver_Ini/Thermo_device.h:29-40 */
static double cp_janaf(const double *Hia, double T0, double Ri, int n)
{
    double T = ora_max(T0, 200.0);
    double Cpi;
    int r = (T > 1000.0) ? 0 : 1;
    Cpi = Ri * (HA(0, r) + (HA(1, r) + (HA(2, r) + (HA(3, r) + HA(4, r) * T) * T) * T) * T);
    return Cpi;
}

/* HeatCapacity: src/solver_Ini/Thermo_device.h:46-56 */
static double cp_i(const ora_config *c, double T0, double Ri, int n)
{
    return c->thermo ? cp_nasa(c->Hia, T0, Ri, n) : cp_janaf(c->Hia, T0, Ri, n);
}

/* get_Enthalpy_NASA: src/solver_Ini/Thermo_device.h:62-80 */
static double h_nasa(const ora_config *c, double T0, double Ri, int n)
{
    const double *Hia = c->Hia, *Hib = c->Hib;
    double hi = 0.0, TT = T0, T = ora_max(T0, 200.0);
    int r = -1;
    if (T >= 1000.0 && T < 6000.0) r = 1;
    else if (T < 1000.0) r = 0;
    else if (T >= 6000.0) r = 2;
    if (r >= 0)
        hi = Ri * (-HA(0, r) / T + HA(1, r) * ora_logc(c, T) +
                   (HA(2, r) + (0.5 * HA(3, r) + (HA(4, r) * ora_OT + (0.25 * HA(5, r) + 0.2 * HA(6, r) * T) * T) * T) * T) * T +
                   HB(0, r));
    if (TT < 200.0) {
        double Cpi = cp_nasa(Hia, 200.0, Ri, n);
        hi += Cpi * (TT - 200.0);
    }
    return hi;
}

/* get_Enthalpy_JANAF: src/solver_Ini/Thermo_device.h:86-102 */
static double h_janaf(const ora_config *c, double T0, double Ri, int n)
{
    const double *Hia = c->Hia;
    double hi, TT = T0, T = ora_max(T0, 200.0);
    int r = (T > 1000.0) ? 0 : 1;
    hi = Ri * (T * (HA(0, r) + T * (HA(1, r) * 0.5 + T * (HA(2, r) * ora_OT + T * (HA(3, r) * 0.25 + HA(4, r) * T * 0.2)))) + HA(5, r));
    if (TT < 200.0) {
        double Cpi = cp_janaf(Hia, 200.0, Ri, n);
        hi += Cpi * (TT - 200.0);
    }
    return hi;
}

/* get_Enthalpy: src/solver_Ini/Thermo_device.h:107-117 */
static double h_i(const ora_config *c, double T0, double Ri, int n)
{
    return c->thermo ? h_nasa(c, T0, Ri, n) : h_janaf(c, T0, Ri, n);
}

/* get_Entropy_NASA: src/solver_Ini/Thermo_device.h:132-145 */
static double s_nasa(const ora_config *c, double T0, double Ri, int n)
{
    const double *Hia = c->Hia, *Hib = c->Hib;
    double T = ora_max(T0, 200.0);
    double S = 0.0, _T = 1.0 / T;
    int r = -1;
    if (T >= 1000.0 && T < 6000.0) r = 1;
    else if (T < 1000.0) r = 0;
    else if (T >= 6000.0) r = 2;
    if (r >= 0)
        S = Ri * ((-0.5 * HA(0, r) * _T - HA(1, r)) * _T + HA(2, r) * ora_logc(c, T) +
                  (HA(3, r) + (0.5 * HA(4, r) + (HA(5, r) * ora_OT + HA(6, r) * 0.25 * T) * T) * T) * T + HB(1, r));
    return S;
}

/* get_Entropy_JANAF: src/solver_Ini/Thermo_device.h:150-161 */
static double s_janaf(const ora_config *c, double T0, double Ri, int n)
{
    const double *Hia = c->Hia;
    double T = ora_max(T0, 200.0);
    double S;
    int r = (T > 1000.0) ? 0 : 1;
    S = Ri * (HA(0, r) * ora_logc(c, T) + (HA(1, r) + (0.5 * HA(2, r) + (HA(3, r) * ora_OT + HA(4, r) * 0.25 * T) * T) * T) * T + HA(6, r));
    return S;
}

/* get_Entropy: src/solver_Ini/Thermo_device.h:166-177 */
static double s_i(const ora_config *c, double T0, double Ri, int n)
{
    return c->thermo ? s_nasa(c, T0, Ri, n) : s_janaf(c, T0, Ri, n);
}

/* get_Gibson: src/solver_Ini/Thermo_device.h:182-189 (Ri argument unused) */
static double g_i(const ora_config *c, double T, int n)
{
    double temp = 0.0;
    temp += s_i(c, T, 1.0, n);
    temp += (-h_i(c, T, 1.0, n) / T);
    return temp;
}

/* get_Internale: src/solver_Ini/Thermo_device.h:122-127 */
static double e_i(const ora_config *c, double T0, double Ri, int n)
{
    return (h_i(c, T0, 1.0, n) - T0) * Ri;
}

/* ---- exported per-species entry points (test pins) ---- */
double ora_cp_i(const ora_config *c, double T, int n) { return cp_i(c, T, c->Ri[n], n); }
double ora_h_i(const ora_config *c, double T, int n) { return h_i(c, T, c->Ri[n], n); }
double ora_s_i(const ora_config *c, double T, int n) { return s_i(c, T, c->Ri[n], n); }
double ora_g_i(const ora_config *c, double T, int n) { return g_i(c, T, n); }
double ora_e_i(const ora_config *c, double T, int n) { return e_i(c, T, c->Ri[n], n); }

/* get_CopR: src/solver_Ini/Mixing_device.h:49-56 */
double ora_mix_R(const ora_config *c, const double *yi)
{
    double R = 0.0;
    for (int n = 0; n < c->NS; n++)
        R += yi[n] * c->_Wi[n];
    return R * ora_Ru;
}

/* get_CopCp: src/solver_Ini/Mixing_device.h:61-68 */
double ora_mix_cp(const ora_config *c, const double *yi, double T)
{
    double cp = 0.0;
    for (int ii = 0; ii < c->NS; ii++)
        cp += yi[ii] * cp_i(c, T, c->Ri[ii], ii);
    return cp;
}

/* get_CopCv: src/solver_Ini/Mixing_device.h:73-86 */
double ora_mix_cv(const ora_config *c, const double *yi, double T)
{
    double cv = 0.0;
    for (int ii = 0; ii < c->NS; ii++)
        cv += yi[ii] * cp_i(c, T, c->Ri[ii], ii);
    double _W = 0.0;
    for (int ii = 0; ii < c->NS; ii++)
        _W += yi[ii] * c->_Wi[ii];
    cv -= ora_Ru * _W;
    return cv;
}

/* get_CopW: src/solver_Ini/Mixing_device.h:91-97 */
static double mix_W(const ora_config *c, const double *yi)
{
    double _W = 0.0;
    for (int ii = 0; ii < c->NS; ii++)
        _W += yi[ii] * c->_Wi[ii];
    return 1.0 / _W;
}

/* get_CopGamma(thermal, yi, T): src/solver_Ini/Mixing_device.h:102-109 */
double ora_mix_gamma_cv(const ora_config *c, const double *yi, double T)
{
    double Cp = ora_mix_cp(c, yi, T);
    double Cv = ora_mix_cv(c, yi, T);
    return Cp / Cv;
}

/* get_CopGamma(thermal, yi, Cp, T): src/solver_Ini/Mixing_device.h:114-126 */
double ora_mix_gamma_cp(const ora_config *c, const double *yi, double Cp)
{
    double CopW = mix_W(c, yi);
    double g = Cp / (Cp - ora_Ru / CopW);
    if (g > 1.0)
        return g;
    return -1.0;
}

/* get_Coph: src/solver_Ini/Mixing_device.h:131-140 */
double ora_mix_h(const ora_config *c, const double *yi, double T)
{
    double h = 0.0;
    for (int i = 0; i < c->NS; i++) {
        double hi = h_i(c, T, c->Ri[i], i);
        h += hi * yi[i];
    }
    return h;
}

/* get_Cope: src/solver_Ini/Mixing_device.h:145-154 */
double ora_mix_e(const ora_config *c, const double *yi, double T)
{
    double e = 0.0;
    for (int i = 0; i < c->NS; i++) {
        double ei = e_i(c, T, c->Ri[i], i);
        e += ei * yi[i];
    }
    return e;
}

/* sub_FuncT + get_T: src/solver_Ini/Mixing_device.h:159-193 */
double ora_get_T(const ora_config *c, const double *yi, double e, double T0)
{
    double T = T0;
    double tol = 1.0e-6;
    for (int i = 1; i < 101; i++) {
        double h = ora_mix_h(c, yi, T);
        double R = ora_mix_R(c, yi);
        double Cp = ora_mix_cp(c, yi, T);
        double func_T = h - R * T - e;
        double dfunc_T = Cp - R;
        double df = ora_min(func_T / (dfunc_T + 1.0e-30), 1e-3 * T);
        df = ora_max(df, -1e-2 * T);
        T = T - df;
        if (fabs(df) <= tol)
            break;
    }
    return T;
}
