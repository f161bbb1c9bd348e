// Transport-property fits of the host: Lennard-Jones species data, the
// collision-integral tables and the least-squares fits of ln(mu_k),
// ln(lambda_k) and ln(p D_kj) as cubics in ln T that the viscous kernels
// evaluate.  Restates Setup::GetFitCoefficient and its helpers
// (src/read_ini/src/viscfit.cpp:148-193, :318-555), Setup::ReadThermal's
// transport block (src/read_ini/src/thermal.cpp:130-164) and
// GetQuadraticInterCoeff / ZrotFunc / Solve_Overdeter_equations
// (src/read_ini/mixture.hpp:43-142), in the same operation order.
#include "xf_setup.h"
#include <cmath>
#include <cstring>
#include <fstream>

namespace xfh {
namespace {

// constants of src/include/global_setup.h:40-54
const double kPi = 3.1415926535897932384626433832795;
const double kB = 1.3806549 * 1.0e-16; // erg/K
const double kNA = 6.02214129 * 1.0e23;
const double kR0 = 6.02214076e26 * 1.380649e-23; // universal_gas_const, J/(kmol K)
const double kRu_ = kR0 * 1.0E-3;
const int kOrder = 4; // order_polynominal_fitted

// Specie_chara (global_setup.h:136-147)
enum { GEO = 0, EPS = 1, DIA = 2, MUE = 3, ALPHA = 4, ZROT = 5, WI = 6, SID = 8 };

// The collision tables as the reference's Setup holds them: Omega_table[2][37][8],
// then delta_star[8], then T_star[37].  Omega_interpolated reads one row
// past a table when the reduced temperature lies in (T_star[35], T_star[36])
// (and one column past a row for a reduced dipole in (delta_star[6],
// delta_star[7])); those reads land in the next member, so they are kept on
// this flat layout.  A read past T_star meets a pointer member, whose bits
// are a subnormal double (0 for the arithmetic here).
struct Omega {
    std::vector<double> flat = std::vector<double>(2 * 37 * 8 + 8 + 37, 0.0);
    double at(long i) const { return (i >= 0 && i < (long)flat.size()) ? flat[i] : 0.0; }
    double table(int idx, int ti, int tj) const { return at(idx * 296L + ti * 8L + tj); }
    double dstar(int j) const { return at(592 + j); }
    double tstar(int i) const { return at(600 + i); }
};

// Setup::ReadOmega_table (viscfit.cpp:348-365)
bool read_omega(const std::string &path, Omega &o)
{
    std::ifstream fin(path);
    if (!fin) return false;
    for (int n = 0; n < 8; n++) fin >> o.flat[592 + n];
    for (int i = 0; i < 37; i++) {
        fin >> o.flat[600 + i];
        for (int j = 0; j < 8; j++) fin >> o.flat[296 + i * 8 + j];
    }
    for (int p = 0; p < 37; p++)
        for (int q = 0; q < 8; q++) fin >> o.flat[p * 8 + q];
    return (bool)fin;
}

// GetQuadraticInterCoeff (mixture.hpp:43-48)
void quad_coeff(double x1, double x2, double x3, double f1, double f2, double f3, double *aa)
{
    aa[2] = ((f1 - f2) / (x1 - x2) - (f2 - f3) / (x2 - x3)) / (x1 - x3);
    aa[1] = (f1 - f2) / (x1 - x2) - aa[2] * (x1 + x2);
    aa[0] = f1 - aa[1] * x1 - aa[2] * x1 * x1;
}

// Setup::Omega_interpolated (viscfit.cpp:372-438)
double omega_interp(const Omega &o, double Tstar, double deltastar, int index)
{
    int ti1 = 0, ti2 = 1, ti3 = 2;
    if (Tstar > o.tstar(0) && Tstar < o.tstar(36)) {
        int ii = 1;
        while (Tstar > o.tstar(ii)) ii = ii + 1;
        ti1 = ii - 1, ti2 = ii, ti3 = ii + 1;
    } else if (Tstar <= o.tstar(0)) {
        ti1 = 0, ti2 = 1, ti3 = 2;
    } else if (Tstar >= o.tstar(36)) {
        ti1 = 34, ti2 = 35, ti3 = 36;
    }
    int tj1 = 0, tj2 = 1, tj3 = 2;
    if (deltastar > o.dstar(0) && deltastar < o.dstar(7)) {
        int jj = 1;
        while (deltastar > o.dstar(jj)) jj = jj + 1;
        tj1 = jj - 1, tj2 = jj, tj3 = jj + 1;
    } else if (deltastar <= o.dstar(0)) {
        tj1 = 0, tj2 = 1, tj3 = 2;
    } else if (deltastar >= o.dstar(7)) {
        tj1 = 5, tj2 = 6, tj3 = 7;
    }
    double aa[3];
    const double T1 = o.tstar(ti1), T2 = o.tstar(ti2), T3 = o.tstar(ti3);
    quad_coeff(T1, T2, T3, o.table(index, ti1, tj1), o.table(index, ti2, tj1), o.table(index, ti3, tj1), aa);
    const double temp1 = aa[0] + aa[1] * Tstar + aa[2] * Tstar * Tstar;
    quad_coeff(T1, T2, T3, o.table(index, ti1, tj2), o.table(index, ti2, tj2), o.table(index, ti3, tj2), aa);
    const double temp2 = aa[0] + aa[1] * Tstar + aa[2] * Tstar * Tstar;
    quad_coeff(T1, T2, T3, o.table(index, ti1, tj3), o.table(index, ti2, tj3), o.table(index, ti3, tj3), aa);
    const double temp3 = aa[0] + aa[1] * Tstar + aa[2] * Tstar * Tstar;
    quad_coeff(o.dstar(tj1), o.dstar(tj2), o.dstar(tj3), temp1, temp2, temp3, aa);
    return aa[0] + aa[1] * deltastar + aa[2] * deltastar * deltastar;
}

// HeatCapacity_NASA / HeatCapacity_JANAF (src/solver_Ini/Thermo_device.h:10-40),
// operation order as written there
double heat_capacity(const Setup &s, double T0, int n)
{
    const double *H = &s.Hia[(size_t)n * 21];
    const double Ri = kRu_ / s.Wi[n];
    const double T = std::max(T0, 200.0), _T = 1.0 / T;
    auto a = [&](int m, int r) { return H[m * 3 + r]; };
    if (s.desc.thermo) {
        int r = -1;
        if (T >= 1000.0 && T < 6000.0) r = 1;
        else if (T < 1000.0) r = 0;
        else if (T >= 6000.0) r = 2;
        if (r < 0) return 0.0;
        return Ri * ((a(0, r) * _T + a(1, r)) * _T + a(2, r) + (a(3, r) + (a(4, r) + (a(5, r) + a(6, r) * T) * T) * T) * T);
    }
    const int r = T > 1000.0 ? 0 : 1;
    return Ri * (a(0, r) + (a(1, r) + (a(2, r) + (a(3, r) + a(4, r) * T) * T) * T) * T);
}

// ZrotFunc (mixture.hpp:76-80)
double zrot_func(double x)
{
    return 1.0 + std::sqrt(kPi * x) * kPi / 2.0 + x * (0.25 * kPi * kPi + 2.0) + std::pow(kPi * x, 1.5);
}

struct Fitter {
    const Setup &s;
    const Omega &o;

    // Setup::viscosity (viscfit.cpp:444-451), Pa s
    double viscosity(const double *sp, double T) const
    {
        const double Tstar = T / sp[EPS];
        const double deltastar = 0.5 * sp[MUE] * sp[MUE] / sp[EPS] / kB / (std::pow(sp[DIA], 3)) * 1.0e-12;
        const double Omega2 = omega_interp(o, Tstar, deltastar, 0);
        double visc = 5 * 1.0e16 * std::sqrt(kPi * (sp[WI] * 1e3) / kNA * kB * T) / (16 * kPi * sp[DIA] * sp[DIA] * Omega2);
        return visc = 0.1 * visc;
    }

    // Setup::Dkj (viscfit.cpp:508-555), cm^2/s at pressure PP [Pa]
    double dkj(const double *sk, const double *sj, double T, double PP) const
    {
        double epsilon_jk_kB, d_jk, mue_jk_sqr;
        if ((sj[MUE] > 0 && sk[MUE] > 0) || (sj[MUE] == 0 && sk[MUE] == 0)) {
            epsilon_jk_kB = std::sqrt(sj[EPS] * sk[EPS]);
            d_jk = (sj[DIA] + sk[DIA]) / 2.0;
            mue_jk_sqr = sj[MUE] * sk[MUE];
        } else {
            double epsilon_n_kB = 0, epsilon_p_kB = 0, alpha_n = 0, mue_p = 0, d_n = 0, d_p = 0;
            if (sk[MUE] > 0 && sj[MUE] == 0) {
                epsilon_n_kB = sj[EPS], epsilon_p_kB = sk[EPS], alpha_n = sj[ALPHA];
                d_n = sj[DIA], d_p = sk[DIA], mue_p = sk[MUE];
            }
            if (sj[MUE] > 0 && sk[MUE] == 0) {
                epsilon_n_kB = sk[EPS], epsilon_p_kB = sj[EPS], alpha_n = sk[ALPHA];
                d_n = sk[DIA], d_p = sj[DIA], mue_p = sj[MUE];
            }
            const double alpha_n_star = alpha_n / std::pow(d_n, 3.0);
            const double mue_p_star = mue_p / std::pow(epsilon_p_kB * kB, 0.5) / std::pow(d_p, 1.5) * 1.0e-6;
            const double ksi = 1.0 + 0.25 * alpha_n_star * mue_p_star * std::sqrt(epsilon_p_kB / epsilon_n_kB);
            epsilon_jk_kB = ksi * ksi * std::sqrt(epsilon_n_kB * epsilon_p_kB);
            d_jk = std::pow(ksi, -1.0 / 6.0) * (sj[DIA] + sk[DIA]) / 2.0;
            mue_jk_sqr = 0.0;
        }
        const double T_jk_star = T / epsilon_jk_kB;
        const double delta_jk_star = 0.5 * mue_jk_sqr / d_jk / d_jk / d_jk / epsilon_jk_kB / kB * 1.0e-12;
        const double W_jk = sk[WI] * sj[WI] / (sk[WI] + sj[WI]) / kNA * 1.0e3;
        const double Omega1 = omega_interp(o, T_jk_star, delta_jk_star, 1);
        const double PPP = PP * 10.0;
        return 3.0 * std::sqrt(2.0 * kPi * std::pow(T * kB, 3.0) / W_jk) / (16.0 * PPP * kPi * d_jk * d_jk * Omega1) *
               1.0e16;
    }

    // Setup::thermal_conductivities (viscfit.cpp:460-500), W/(m K)
    double conductivity(const double *sp, double T, double PP) const
    {
        const double Cv_trans = 1.5 * kR0;
        double Cv_rot = 0.0, Cv_vib = 0.0;
        const int id = int(sp[SID]);
        const double Cpi = heat_capacity(s, T, id);
        const double Cv = Cpi * sp[WI] * 1.0e3 - kR0;
        switch (int(sp[GEO])) {
        case 0: Cv_rot = 0.0, Cv_vib = 0.0; break;
        case 1: Cv_rot = 1.0 * kR0, Cv_vib = Cv - 2.5 * kR0; break;
        case 2: Cv_rot = 1.5 * kR0, Cv_vib = Cv - 3.0 * kR0; break;
        }
        const double rho = PP * sp[WI] / T / kR0;
        const double Dkk = dkj(sp, sp, T, PP);
        const double visc = viscosity(sp, T);
        const double f_vib = rho * Dkk / (visc * 10.0);
        const double Zrot = sp[ZROT] * zrot_func(sp[EPS] / 298.0) / zrot_func(sp[EPS] / T);
        const double Aa = 2.5 - f_vib, Bb = Zrot + 2.0 * (5.0 * Cv_rot / 3.0 / (kR0) + f_vib) / kPi;
        const double f_trans = 2.5 * (1.0 - 2.0 * Cv_rot * Aa / kPi / Cv_trans / Bb);
        const double f_rot = f_vib * (1.0 + 2.0 * Aa / kPi / Bb);
        return visc * (f_trans * Cv_trans + f_rot * Cv_rot + f_vib * Cv_vib) / sp[WI] * 1.0e-3;
    }

    // Solve_Overdeter_equations (mixture.hpp:89-142): normal equations and a
    // Cholesky factorisation
    static void solve(const std::vector<double> &AA, const std::vector<double> &b, int mm, double *xx)
    {
        const int nn = kOrder;
        double CC[kOrder][kOrder], dd[kOrder];
        for (int i = 0; i < nn; i++) {
            dd[i] = 0.0;
            for (int j = 0; j < nn; j++) CC[i][j] = 0.0;
        }
        for (int i = 0; i < nn; i++) {
            for (int q = 0; q < mm; q++) dd[i] = dd[i] + AA[q * nn + i] * b[q];
            for (int j = 0; j < nn; j++)
                for (int k = 0; k < mm; k++) CC[i][j] = CC[i][j] + AA[k * nn + i] * AA[k * nn + j];
        }
        CC[0][0] = std::sqrt(CC[0][0]);
        for (int p = 1; p < nn; p++) CC[p][0] = CC[p][0] / CC[0][0];
        for (int k = 1; k < nn; k++) {
            for (int m = 0; m < k; m++) CC[k][k] = CC[k][k] - CC[k][m] * CC[k][m];
            CC[k][k] = std::sqrt(CC[k][k]);
            for (int i = k + 1; i < nn; i++) {
                for (int m = 0; m < k; m++) CC[i][k] = CC[i][k] - CC[i][m] * CC[k][m];
                CC[i][k] = CC[i][k] / CC[k][k];
            }
        }
        xx[0] = dd[0] / CC[0][0];
        for (int j = 1; j < nn; j++) {
            for (int q = 0; q < j; q++) dd[j] = dd[j] - CC[j][q] * xx[q];
            xx[j] = dd[j] / CC[j][j];
        }
        dd[nn - 1] = xx[nn - 1] / CC[nn - 1][nn - 1];
        for (int i = nn - 2; i >= 0; i--) {
            for (int p = nn - 1; p > i; p--) xx[i] = xx[i] - CC[p][i] * dd[p];
            dd[i] = xx[i] / CC[i][i];
        }
        for (int m = 0; m < nn; m++) xx[m] = dd[m];
    }

    // Setup::Fitting (viscfit.cpp:318-341): indicator 0 viscosity, 1
    // conductivity, 2 binary diffusion
    void fit(const std::vector<double> &TT, const double *sk, const double *sj, double *aa, int indicator) const
    {
        const int mm = (int)TT.size();
        std::vector<double> b(mm), AA((size_t)mm * kOrder);
        for (int ii = 0; ii < mm; ii++) {
            switch (indicator) {
            case 0: b[ii] = std::log(viscosity(sk, TT[ii])); break;
            case 1: b[ii] = std::log(conductivity(sk, TT[ii], 1.0)); break;
            case 2: b[ii] = std::log(dkj(sk, sj, TT[ii], 1.0)); break;
            }
            for (int jj = 0; jj < kOrder; jj++) AA[(size_t)ii * kOrder + jj] = std::pow(std::log(TT[ii]), jj);
        }
        solve(AA, b, mm, aa);
    }
};

} // namespace

bool transport_load(const std::string &root, Setup &s, std::string &err)
{
    const int NS = (int)s.species.size();
    // Setup::ReadThermal, transport block (thermal.cpp:130-164): the first
    // matching token of transport_data.dat is followed by geo, eps/kB, sigma,
    // mue, alpha, Zrot(298 K)
    std::vector<std::string> tok;
    {
        std::ifstream f(root + "/transport_data.dat");
        if (!f) {
            err = "cannot open " + root + "/transport_data.dat";
            return false;
        }
        std::string t;
        while (f >> t) tok.push_back(t);
    }
    s.chara.assign((size_t)NS * 9, 0.0);
    for (int i = 0; i < NS; i++) {
        double *sp = &s.chara[(size_t)i * 9];
        bool found = false;
        for (size_t q = 0; q < tok.size(); q++) {
            if (tok[q] == "*END") break;
            if (tok[q] == s.species[i] && q + 6 < tok.size()) {
                for (int m = 0; m < 6; m++) sp[m] = std::strtod(tok[q + 1 + m].c_str(), nullptr);
                found = true;
                break;
            }
        }
        if (!found) {
            err = "no transport data for species " + s.species[i];
            return false;
        }
        sp[WI] = s.Wi[i]; // kg/mol
        sp[7] = 0.0;
        sp[SID] = i;
    }
    Omega o;
    if (!read_omega(root + "/collision_integral.dat", o)) {
        err = "cannot read " + root + "/collision_integral.dat";
        return false;
    }
    Fitter F{s, o};
    s.fcv.assign((size_t)NS * kOrder, 0.0);
    s.fct.assign((size_t)NS * kOrder, 0.0);
    s.Dkj.assign((size_t)NS * NS * kOrder, 0.0);
    for (int k = 0; k < NS; k++) {
        const double *sk = &s.chara[(size_t)k * 9];
        F.fit(s.Tnode, sk, sk, &s.fcv[(size_t)k * kOrder], 0);
        F.fit(s.Tnode, sk, sk, &s.fct[(size_t)k * kOrder], 1);
        for (int j = 0; j < NS; j++) {
            const double *sj = &s.chara[(size_t)j * 9];
            if (k <= j)
                F.fit(s.Tnode, sk, sj, &s.Dkj[((size_t)k * NS + j) * kOrder], 2);
            else
                for (int n = 0; n < kOrder; n++)
                    s.Dkj[((size_t)k * NS + j) * kOrder + n] = s.Dkj[((size_t)j * NS + k) * kOrder + n];
        }
    }
    return true;
}

double transport_exact(const Setup &s, const std::string &root, int what, int k, int j, double T, double PP)
{
    Omega o;
    if (!read_omega(root + "/collision_integral.dat", o)) return NAN;
    Fitter F{s, o};
    const double *sk = &s.chara[(size_t)k * 9], *sj = &s.chara[(size_t)j * 9];
    if (what == 0) return F.viscosity(sk, T);
    if (what == 1) return F.conductivity(sk, T, PP);
    return F.dkj(sk, sj, T, PP);
}

} // namespace xfh
