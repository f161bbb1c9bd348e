// Seeded synthetic initial states for the benchmark workloads.
//
//   isentropic-vortex    exactly the reference's 2D Euler vortex
//                        (src/solver_Ini/sample/2D-EulerVortex/ini_sample.hpp:14-89)
//   shock-bubble         Mach-M planar shock in the "outside" mixture hitting
//                        a circular (spherical in 3-D) bubble of the "inside"
//                        mixture; radius init.bubble_shape_x or 0.2 of the
//                        domain height, tanh interface over
//                        init.bubble_boundary_cells or 3 cells
//   reactive-shock-tube  the reference's 1D reactive shock tube states
//                        (src/solver_Ini/sample/1D-X-Y-Z/reactive-st/ini_sample.hpp:46-69)
//   mixing-layer         tanh shear layer, velocity ratio 3, temperature ratio 2,
//                        seeded random-phase perturbation of 1 % amplitude
//   reactive-channel     periodic premixed H2/O2/Ar channel with a hot kernel;
//                        T0 += U[-50, 50] K and equivalence ratio U[0.8, 1.2] by seed
// Conserved state: U = [rho, rho u, rho v, rho w, rho (h + |u|^2/2) - p, rho Y_k (k < NS-1)]
// as the reference's samples build it (e.g. reactive-st/ini_sample.hpp:69).
#include "xf_setup.h"
#include <cmath>
#include <random>

namespace xfh {

namespace {

const double kPi = 3.1415926535897932384626433832795;

struct Writer {
    const Setup &s;
    double *U, *T;
    int E, NS;
    void put(long long id, double rho, double u, double v, double w, double p, double Tk, const double *Y)
    {
        double *q = U + (size_t)E * id;
        q[0] = rho;
        q[1] = rho * u;
        q[2] = rho * v;
        q[3] = rho * w;
        if (s.desc.cop) {
            const double h = host_mix_h(s, Y, Tk);
            q[4] = rho * (h + 0.5 * (u * u + v * v + w * w)) - p;
            for (int k = 0; k < NS - 1; k++) q[5 + k] = rho * Y[k];
        } else {
            q[4] = p / (s.desc.ncop_gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
        }
        T[id] = Tk;
    }
};

inline double coord(int i, int B, double d, double xmin) { return (i - B) * d + 0.5 * d + xmin; }

} // namespace

bool setup_initial(const Setup &s, const std::string &wl, unsigned long long seed, double *U, double *T, std::string &err)
{
    const xf_case_desc &d = s.desc;
    const int NS = d.NS, E = d.cop ? NS + 4 : 5;
    Writer W{s, U, T, E, NS};
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::vector<double> Y(NS), Ya(NS), Yb(NS);

    if (wl == "isentropic-vortex") {
        if (d.cop) {
            err = "isentropic-vortex needs the single-species (NO-COP) mixture";
            return false;
        }
        for (int k = 0; k < d.Zmax; k++)
            for (int j = 0; j < d.Ymax; j++)
                for (int i = 0; i < d.Xmax; i++) {
                    const long long id = (long long)d.Xmax * d.Ymax * k + (long long)d.Xmax * j + i;
                    const double x = d.DimX ? coord(i, d.Bwidth_X, s.dx, s.Domain_xmin) : 0.0;
                    const double y = d.DimY ? coord(j, d.Bwidth_Y, s.dy, s.Domain_ymin) : 0.0;
                    const double beta = 5.0, beta2 = beta * beta;
                    const double xv = x - 0.5 * s.Domain_length, yv = y - 0.5 * s.Domain_width;
                    const double r2 = xv * xv + yv * yv;
                    const double inter1 = beta2 * std::exp(1.0 - r2) / 28.0 / kPi / kPi;
                    const double inter2 = 0.5 * beta * std::exp(0.5 * (1.0 - r2)) / kPi;
                    const double rho = std::pow(1.0 - inter1, 2.5);
                    const double p = std::pow(rho, 1.4);
                    const double u = 1.0 - inter2 * yv, v = 1.0 + inter2 * xv, w = 0.0;
                    double *q = U + (size_t)E * id;
                    q[0] = rho, q[1] = rho * u, q[2] = rho * v, q[3] = rho * w;
                    q[4] = p / (1.4 - 1.0) + 0.5 * rho * (u * u + v * v + w * w);
                    T[id] = 0.0;
                }
        return true;
    }
    if (!d.cop) {
        err = "workload " + wl + " needs a multi-component mixture";
        return false;
    }

    if (wl == "shock-bubble") {
        const double Ma = s.blast_mach > 1.0 ? s.blast_mach : 1.22;
        const double p1 = s.blast_down[1] > 0 ? s.blast_down[1] : 101325.0;
        const double T1 = s.blast_down[2] > 0 ? s.blast_down[2] : 295.0;
        const double xs = s.blast_center[0];
        const double xb = s.cop_center[0], yb = s.cop_center[1];
        const double R = s.bubble_r > 0.0 ? s.bubble_r : 0.2 * (d.DimY ? s.Domain_width : s.Domain_length);
        const double delta = (s.bubble_cells > 0.0 ? s.bubble_cells : 3.0) * s.dx;
        // Rankine-Hugoniot jump with the air's gamma at T1
        const double Rair = host_mix_R(s, s.Y_out.data());
        const double cp = host_mix_cp(s, s.Y_out.data(), T1);
        const double g = cp / (cp - Rair);
        const double rho1 = p1 / (Rair * T1), c1 = std::sqrt(g * p1 / rho1);
        const double p2 = p1 * (1.0 + 2.0 * g / (g + 1.0) * (Ma * Ma - 1.0));
        const double rho2 = rho1 * (g + 1.0) * Ma * Ma / ((g - 1.0) * Ma * Ma + 2.0);
        const double u2 = Ma * c1 * (1.0 - rho1 / rho2);
        const double T2 = p2 / (Rair * rho2);
        for (int k = 0; k < d.Zmax; k++)
            for (int j = 0; j < d.Ymax; j++)
                for (int i = 0; i < d.Xmax; i++) {
                    const long long id = (long long)d.Xmax * d.Ymax * k + (long long)d.Xmax * j + i;
                    const double x = coord(i, d.Bwidth_X, s.dx, s.Domain_xmin);
                    const double y = d.DimY ? coord(j, d.Bwidth_Y, s.dy, s.Domain_ymin) : yb;
                    const double z = d.DimZ ? coord(k, d.Bwidth_Z, s.dz, s.Domain_zmin) : s.cop_center[2];
                    if (x < xs) {
                        W.put(id, rho2, u2, 0.0, 0.0, p2, T2, s.Y_out.data());
                        continue;
                    }
                    const double r = std::sqrt((x - xb) * (x - xb) + (y - yb) * (y - yb) +
                                               (d.DimZ ? (z - s.cop_center[2]) * (z - s.cop_center[2]) : 0.0));
                    const double phi = 0.5 * (1.0 - std::tanh((r - R) / delta));
                    for (int n = 0; n < NS; n++) Y[n] = s.Y_out[n] * (1.0 - phi) + s.Y_in[n] * phi;
                    const double rho = p1 / (host_mix_R(s, Y.data()) * T1);
                    W.put(id, rho, 0.0, 0.0, 0.0, p1, T1, Y.data());
                }
        return true;
    }

    if (wl == "reactive-shock-tube") {
        for (int k = 0; k < d.Zmax; k++)
            for (int j = 0; j < d.Ymax; j++)
                for (int i = 0; i < d.Xmax; i++) {
                    const long long id = (long long)d.Xmax * d.Ymax * k + (long long)d.Xmax * j + i;
                    const double x = coord(i, d.Bwidth_X, s.dx, s.Domain_xmin);
                    const bool left = x < 0.06;
                    const double rho = left ? 0.072 : 0.18075, u = left ? 0.0 : -487.34, p = left ? 7173.0 : 35594.0;
                    const double Tk = p / host_mix_R(s, s.Y_out.data()) / rho;
                    W.put(id, rho, u, 0.0, 0.0, p, Tk, s.Y_out.data());
                }
        return true;
    }

    if (wl == "mixing-layer") {
        // upper stream: the "outside" mixture, lower stream: the "inside" mixture
        const double p0 = 101325.0, T1 = 1200.0, T2 = 600.0, u1 = 300.0, u2 = 100.0;
        const double delta = 0.05 * s.Domain_width;
        const double y0 = s.Domain_ymin + 0.5 * s.Domain_width;
        double ph[4], ps[4];
        for (int m = 0; m < 4; m++) ph[m] = 2.0 * kPi * uni(rng), ps[m] = 2.0 * kPi * uni(rng);
        for (int k = 0; k < d.Zmax; k++)
            for (int j = 0; j < d.Ymax; j++)
                for (int i = 0; i < d.Xmax; i++) {
                    const long long id = (long long)d.Xmax * d.Ymax * k + (long long)d.Xmax * j + i;
                    const double x = coord(i, d.Bwidth_X, s.dx, s.Domain_xmin);
                    const double y = coord(j, d.Bwidth_Y, s.dy, s.Domain_ymin) - y0;
                    const double z = d.DimZ ? coord(k, d.Bwidth_Z, s.dz, s.Domain_zmin) : 0.0;
                    const double phi = 0.5 * (1.0 + std::tanh(y / delta));
                    for (int n = 0; n < NS; n++) Y[n] = s.Y_in[n] * (1.0 - phi) + s.Y_out[n] * phi;
                    const double Tk = T2 + (T1 - T2) * phi;
                    const double u = u2 + (u1 - u2) * phi;
                    double pert = 0.0;
                    for (int m = 0; m < 4; m++)
                        pert += std::sin(2.0 * kPi * (m + 1) * x / s.Domain_length + ph[m]) *
                                std::cos(2.0 * kPi * (m + 1) * z / s.Domain_height + ps[m]);
                    const double v = 0.01 * (u1 - u2) * std::exp(-(y / delta) * (y / delta)) * pert * 0.25;
                    const double rho = p0 / (host_mix_R(s, Y.data()) * Tk);
                    W.put(id, rho, u, v, 0.0, p0, Tk, Y.data());
                }
        return true;
    }

    if (wl == "reactive-channel") {
        const double dT = -50.0 + 100.0 * uni(rng);
        const double eqr = 0.8 + 0.4 * uni(rng);
        // H2 : O2 : rest = 2 phi : 1 : (the outside mixture's remainder)
        std::vector<double> x(s.xi_out);
        for (int n = 0; n < NS; n++)
            if (s.species[n] == "H2") x[n] = x[n] * eqr;
        host_x_to_y(s, x.data(), Ya.data());
        const double p0 = 101325.0, T0 = 1100.0 + dT;
        const double xc = s.Domain_xmin + 0.5 * s.Domain_length, yc = s.Domain_ymin + 0.5 * s.Domain_width;
        const double rk = 0.1 * s.Domain_length;
        for (int k = 0; k < d.Zmax; k++)
            for (int j = 0; j < d.Ymax; j++)
                for (int i = 0; i < d.Xmax; i++) {
                    const long long id = (long long)d.Xmax * d.Ymax * k + (long long)d.Xmax * j + i;
                    const double xx = coord(i, d.Bwidth_X, s.dx, s.Domain_xmin) - xc;
                    const double yy = d.DimY ? coord(j, d.Bwidth_Y, s.dy, s.Domain_ymin) - yc : 0.0;
                    const double Tk = T0 + 300.0 * std::exp(-(xx * xx + yy * yy) / (rk * rk));
                    const double rho = p0 / (host_mix_R(s, Ya.data()) * Tk);
                    const double u = 1.0 * (uni(rng) - 0.5), v = 1.0 * (uni(rng) - 0.5);
                    W.put(id, rho, u, v, 0.0, p0, Tk, Ya.data());
                }
        return true;
    }
    err = "unknown workload " + wl;
    return false;
}

} // namespace xfh
