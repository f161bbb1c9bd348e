/*
 * Viscous wall fluxes of the oracle (test infrastructure only): velocity
 * gradients at cell centres, mixture-averaged transport coefficients and the
 * fourth-order viscous, heat-conduction and species-diffusion face fluxes that
 * GetLU subtracts from the convective face fluxes before the divergence.
 *
 * Restates, operation by operation:
 *   src/solver_Reconstruction/viscosity/Visc_block.hpp:8-84 (GetCellCenterDerivative)
 *   src/solver_Reconstruction/viscosity/Fourth_Order/Visc_Order_kernels.hpp
 *       :16-71 (GetInnerCellCenterDerivativeKernel), :73-342 (GetWallViscousFlux{X,Y,Z})
 *   src/solver_Reconstruction/viscosity/Fourth_Order/Flux_discrete.h:4-93 (MARCO_*VISC*)
 *   src/solver_Reconstruction/viscosity/Visc_kernels.hpp:7-235 (CenterDerivativeBCKernel{X,Y,Z},
 *       Gettransport_coeff_aver)
 *   src/solver_Reconstruction/viscosity/Visc_device.h:9-177 (Viscosity, PHI,
 *       Thermal_conductivity, GetDkj, Get_transport_coeff_aver)
 *   src/solver_Ini/Mixing_device.h:31-44 (get_xi)
 *   src/solver_Reconstruction/FDM_Method/ConVenction_block.hpp:424-575 (the driver)
 *
 * Choices where the reference has no defined value:
 *   - sycl::pow(x, 0.25), pow(x, 0.5) and pow(x, -0.5) in PHI are taken as
 *     sqrt(sqrt(x)), sqrt(x) and 1/sqrt(x) (IEEE square roots);
 *   - a centre derivative along an inactive direction reads outside the
 *     arrays in the reference (times DimX_t = 0); it is +0 here;
 *   - the z no-slip wall of CenterDerivativeBCKernelZ indexes k - offset
 *     (outside the array); the mirror offset - k of the x and y kernels is
 *     used.
 * Kept as written: the diffusion limits of the single-process build (yi_min
 * and yi_max reduce into values initialised to 0; Dim_max stays 0, so the
 * diffusion flux is clamped to a signed zero), the species fluxes built
 * from the last species' clamped Dim_Yil, and the heat conductivity
 * averaged with the viscosity's Wilke denominators.
 *
 * Pins (tests/test_oracle_visc.py): centre and face differences are exact
 * on quartic and cubic fields; a linear shear gives tau = mu du/dy and a
 * linear temperature profile the flux kappa dT/dx exactly; the Wilke and
 * mixture-averaged formulas reduce to the pure-species values for one
 * species and to (1 - Y_k) / sum_j X_j / D_jk for a binary mixture; the
 * fitted pure-species viscosities, conductivities and binary diffusion
 * coefficients are checked against published measurements.
 */
#include "ora_internal.h"
#include <stdlib.h>
#include <string.h>

#define IDX(f, i, j, k) ((size_t)(f)->cfg.Xmax * (f)->cfg.Ymax * (k) + (size_t)(f)->cfg.Xmax * (j) + (i))

enum { DUCX = 0, DVCX = 1, DWCX = 2, DUCY = 3, DVCY = 4, DWCY = 5, DUCZ = 6, DVCZ = 7, DWCZ = 8 };

static const double OT = 1.0 / 3.0;    /* _OT   global_setup.h:35 */
static const double SXTN = 1.0 / 16.0; /* _sxtn global_setup.h:36 */
static const double TWFR = 1.0 / 24.0; /* _twfr global_setup.h:37 */
static const double TWLE = 1.0 / 12.0; /* _twle global_setup.h:38 */

/* sycl::min / sycl::max on doubles: min(x, y) = y < x ? y : x,
   max(x, y) = x < y ? y : x */
static inline double smin(double x, double y) { return y < x ? y : x; }
static inline double smax(double x, double y) { return x < y ? y : x; }

/* log-polynomial fit of order 4 evaluated as Viscosity / Thermal_conductivity /
   GetDkj do (Visc_device.h:9-20, 41-52, 77-90) */
static double fit_poly(const ora_config *c, const double *a, double T)
{
    double v = a[3];
    for (int i = 2; i >= 0; i--) v = v * ora_logc(c, T) + a[i];
    return v;
}

void ora_visc_alloc(ora_fluid *f)
{
    if (!f->cfg.visc) return;
    size_t n = f->ncell;
    f->Vde = calloc(9 * n, 8);
    f->va = calloc(n, 8);
    f->tca = calloc(n, 8);
    f->Dkm = calloc(n * f->cfg.NS, 8);
    f->hi = calloc(n * f->cfg.NS, 8);
}

void ora_visc_free(ora_fluid *f)
{
    free(f->Vde), free(f->va), free(f->tca), free(f->Dkm), free(f->hi);
}

/* GetInnerCellCenterDerivativeKernel (Visc_Order_kernels.hpp:16-71) over
   i in [Bwidth_X - 2, Xmax - Bwidth_X + 1] per active direction */
static void centre_derivatives(ora_fluid *f)
{
    const ora_config *c = &f->cfg;
    const double DX = c->DimX, DY = c->DimY, DZ = c->DimZ;
    const int i0 = c->DimX ? c->Bwidth_X - 2 : 0, i1 = c->DimX ? c->Xmax - c->Bwidth_X + 1 : 0;
    const int j0 = c->DimY ? c->Bwidth_Y - 2 : 0, j1 = c->DimY ? c->Ymax - c->Bwidth_Y + 1 : 0;
    const int k0 = c->DimZ ? c->Bwidth_Z - 2 : 0, k1 = c->DimZ ? c->Zmax - c->Bwidth_Z + 1 : 0;
    const size_t n = f->ncell, sy = (size_t)c->Xmax, sz = (size_t)c->Xmax * c->Ymax;
    const double *u = f->u, *v = f->v, *w = f->w;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                const size_t id = IDX(f, i, j, k);
                double D[9];
                for (int q = 0; q < 9; q++) D[q] = 0.0;
                if (c->DimX) {
                    const size_t m1 = id - 1, m2 = id - 2, p1 = id + 1, p2 = id + 2;
                    D[DUCX] = (8.0 * (u[p1] - u[m1]) - (u[p2] - u[m2])) * c->_dx * TWLE * DX;
                    D[DVCX] = (8.0 * (v[p1] - v[m1]) - (v[p2] - v[m2])) * c->_dx * TWLE * DX * DY;
                    D[DWCX] = (8.0 * (w[p1] - w[m1]) - (w[p2] - w[m2])) * c->_dx * TWLE * DX * DZ;
                }
                if (c->DimY) {
                    const size_t m1 = id - sy, m2 = id - 2 * sy, p1 = id + sy, p2 = id + 2 * sy;
                    D[DUCY] = (8.0 * (u[p1] - u[m1]) - (u[p2] - u[m2])) * c->_dy * TWLE * DY * DX;
                    D[DVCY] = (8.0 * (v[p1] - v[m1]) - (v[p2] - v[m2])) * c->_dy * TWLE * DY;
                    D[DWCY] = (8.0 * (w[p1] - w[m1]) - (w[p2] - w[m2])) * c->_dy * TWLE * DY * DZ;
                }
                if (c->DimZ) {
                    const size_t m1 = id - sz, m2 = id - 2 * sz, p1 = id + sz, p2 = id + 2 * sz;
                    D[DUCZ] = (8.0 * (u[p1] - u[m1]) - (u[p2] - u[m2])) * c->_dz * TWLE * DZ * DX;
                    D[DVCZ] = (8.0 * (v[p1] - v[m1]) - (v[p2] - v[m2])) * c->_dz * TWLE * DZ * DY;
                    D[DWCZ] = (8.0 * (w[p1] - w[m1]) - (w[p2] - w[m2])) * c->_dz * TWLE * DZ;
                }
                for (int q = 0; q < 9; q++) f->Vde[q * n + id] = D[q];
            }
}

/* CenterDerivativeBCKernel{X,Y,Z} (Visc_kernels.hpp:7-203) on one ghost cell
   of direction d; comp[4] are the components that direction's wall flux
   reads across the boundary.  lo: the low side (mirror_offset 0,
   index_inner = B, sign +1) or the high side. */
static void deriv_bc_cell(ora_fluid *f, int d, int bc, int lo, int i, int j, int k, const int *comp)
{
    const ora_config *c = &f->cfg;
    const int B = d == 0 ? c->Bwidth_X : (d == 1 ? c->Bwidth_Y : c->Bwidth_Z);
    const int inner = d == 0 ? c->X_inner : (d == 1 ? c->Y_inner : c->Z_inner);
    const int L = d == 0 ? c->Xmax : (d == 1 ? c->Ymax : c->Zmax);
    const int s = d == 0 ? i : (d == 1 ? j : k);
    const int mirror_offset = lo ? 0 : inner, index_inner = lo ? B : L - B - 1, sign = lo ? 1 : -1;
    const size_t n = f->ncell, id = IDX(f, i, j, k);
    int t = -1, neg = 0;
    switch (bc) {
    case ORA_SYMMETRY: t = 2 * (B + mirror_offset) - 1 - s; break;
    case ORA_PERIODIC: t = s + sign * inner; break;
    case ORA_OUTFLOW: t = index_inner; break;
    case ORA_NSLIPWALL: t = 2 * (B + mirror_offset) - 1 - s, neg = 1; break;
    case ORA_INFLOW:
        for (int q = 0; q < 4; q++) f->Vde[comp[q] * n + id] = 0.0;
        return;
    default: return; /* viscWall, slipWall, innerBlock: unchanged */
    }
    const size_t tid = d == 0 ? IDX(f, t, j, k) : (d == 1 ? IDX(f, i, t, k) : IDX(f, i, j, t));
    for (int q = 0; q < 4; q++) {
        const double v = f->Vde[comp[q] * n + tid];
        f->Vde[comp[q] * n + id] = neg ? -v : v;
    }
}

/* GetCellCenterDerivative (Visc_block.hpp:8-84): inner kernel, then the x,
   y and z ghost kernels (each over every cell of the other two axes) */
static void cell_centre_derivative(ora_fluid *f)
{
    const ora_config *c = &f->cfg;
    static const int cx[4] = {DUCY, DUCZ, DVCY, DWCZ}, cy[4] = {DVCX, DVCZ, DUCX, DWCZ},
                     cz[4] = {DWCX, DWCY, DUCX, DVCY};
    centre_derivatives(f);
    if (c->DimX)
        for (int k = 0; k < c->Zmax; k++)
            for (int j = 0; j < c->Ymax; j++)
                for (int g = 0; g < c->Bwidth_X; g++) {
                    deriv_bc_cell(f, 0, c->BC[0], 1, g, j, k, cx);
                    deriv_bc_cell(f, 0, c->BC[1], 0, g + c->Xmax - c->Bwidth_X, j, k, cx);
                }
    if (c->DimY)
        for (int k = 0; k < c->Zmax; k++)
            for (int g = 0; g < c->Bwidth_Y; g++)
                for (int i = 0; i < c->Xmax; i++) {
                    deriv_bc_cell(f, 1, c->BC[2], 1, i, g, k, cy);
                    deriv_bc_cell(f, 1, c->BC[3], 0, i, g + c->Ymax - c->Bwidth_Y, k, cy);
                }
    if (c->DimZ)
        for (int g = 0; g < c->Bwidth_Z; g++)
            for (int j = 0; j < c->Ymax; j++)
                for (int i = 0; i < c->Xmax; i++) {
                    deriv_bc_cell(f, 2, c->BC[4], 1, i, j, g, cz);
                    deriv_bc_cell(f, 2, c->BC[5], 0, i, j, g + c->Zmax - c->Bwidth_Z, cz);
                }
}

/* Gettransport_coeff_aver + Get_transport_coeff_aver (Visc_kernels.hpp:205-235,
   Visc_device.h:99-177) for cell id: species enthalpies (Visc_Diffu), mole
   fractions (get_xi), Wilke viscosity and conductivity, mixture-averaged
   diffusion coefficients */
static void transport_cell(ora_fluid *f, size_t id)
{
    const ora_config *c = &f->cfg;
    const int NS = c->NS;
    const size_t n = f->ncell;
    const double T = f->T[id], rho = f->rho[id], p = f->p[id];
    const double *yi = f->y + (size_t)NS * id;
    if (c->visc_diffu)
        for (int ii = 0; ii < NS; ii++) f->hi[(size_t)NS * id + ii] = ora_h_i(c, T, ii);
    double X[ORA_MAXNS], C_total = 0.0;
    for (int i = 0; i < NS; i++) {
        X[i] = yi[i] * c->_Wi[i] * 1e-3 * rho;
        C_total = C_total + X[i];
    }
    const double _C_total = 1.0 / C_total;
    for (int i = 0; i < NS; i++) X[i] = X[i] * _C_total;
    /* the species viscosities; PHI re-evaluates them per pair, with the same value */
    double mu[ORA_MAXNS];
    for (int k = 0; k < NS; k++) mu[k] = ora_expc(c, fit_poly(c, c->fcv + 4 * k, T));
    double va = 0.0, tca = 0.0;
    for (int k = 0; k < NS; k++) {
        double den = 0.0;
        for (int i = 0; i < NS; i++) {
            /* PHI(specie[k], specie[i]) */
            double phi = sqrt(sqrt(c->Wi[i] / c->Wi[k])) * sqrt(mu[k] / mu[i]);
            phi = (phi + 1.0) * (phi + 1.0) * 0.5 / sqrt(2.0);
            phi = phi * (1.0 / sqrt(1.0 + c->Wi[k] / c->Wi[i]));
            den = den + X[i] * phi;
        }
        const double _den = 1.0 / den;
        va = va + X[k] * mu[k] * _den;
        if (c->visc_heat) tca = tca + X[k] * ora_expc(c, fit_poly(c, c->fct + 4 * k, T)) * _den;
    }
    f->va[id] = va;
    if (c->visc_heat) f->tca[id] = tca;
    if (c->visc_diffu) {
        for (int k = 0; k < NS; k++) {
            double temp1 = 0.0, temp2 = 1.0e-20;
            for (int i = 0; i < NS; i++) {
                if (i == k) continue;
                temp1 += (X[i] + 1.0e-40) * c->Wi[i];
                /* GetDkj(specie[i], specie[k]) = exp(fit(log T)) / p */
                const double Dik = ora_expc(c, fit_poly(c, c->Dkj + 4 * ((size_t)i * NS + k), T)) / p;
                temp2 += (X[i] + 1.0e-40) / (Dik + 1.0e-40);
            }
            double D;
            if (temp1 <= 0.0) /* sycl::step(ceil(temp1), 0): ceil(t) <= 0 exactly when t <= 0 */
                D = ora_expc(c, fit_poly(c, c->Dkj + 4 * ((size_t)k * NS + k), T)) / p;
            else
                D = temp1 / temp2 / rho * C_total;
            D *= 1.0e-1;
            f->Dkm[k * n + id] = smax(D, 1.0e-10);
        }
    }
}

/* GetWallViscousFlux{X,Y,Z} with MARCO_PREVISCFLUX / MARCO_VISCFLUX
   (Visc_Order_kernels.hpp:73-342, Flux_discrete.h:4-93) on the face to the
   right of cell id along d; subtracts the viscous flux from Fw */
static void wall_flux(ora_fluid *f, int d, size_t id, const double *yil_lim, const double *diffu_lim, double *Fwall)
{
    const ora_config *c = &f->cfg;
    const int NS = c->NS, E = f->E;
    const size_t n = f->ncell;
    const size_t st = d == 0 ? 1 : (d == 1 ? (size_t)c->Xmax : (size_t)c->Xmax * c->Ymax);
    const size_t m1 = id - st, p1 = id + st, p2 = id + 2 * st;
    const double DX = c->DimX, DY = c->DimY, DZ = c->DimZ;
    const double _dl = d == 0 ? c->_dx : (d == 1 ? c->_dy : c->_dz);
    const double *u = f->u, *v = f->v, *w = f->w, *va = f->va;
    const double *Vd = f->Vde;
#define HALF(a) ((9.0 * ((a)[p1] + (a)[id]) - ((a)[p2] + (a)[m1])) * SXTN)
#define GRAD(a) ((27.0 * ((a)[p1] - (a)[id]) - ((a)[p2] - (a)[m1])) * _dl * TWFR)
    const double mue = (9.0 * (va[p1] + va[id]) - (va[p2] + va[m1])) * SXTN;
    const double lamada = -2.0 * OT * mue;
    double f_x, f_y, f_z, u_hlf, v_hlf, w_hlf;
    const double *A, *Bv;
    if (d == 0) {
        A = Vd + DVCY * n, Bv = Vd + DWCZ * n;
        f_x = (2.0 * mue + lamada) * (27.0 * (u[p1] - u[id]) - (u[p2] - u[m1])) * _dl * TWFR;
        f_x += lamada * (9.0 * (A[p1] + A[id]) - (A[p2] + A[m1]) + 9.0 * (Bv[p1] + Bv[id]) - (Bv[p2] + Bv[m1])) * SXTN;
        const double *Cy = Vd + DUCY * n, *Cz = Vd + DUCZ * n;
        f_y = mue * (27.0 * (v[p1] - v[id]) - (v[p2] - v[m1])) * _dl * TWFR * DY;
        f_y += mue * (9.0 * (Cy[p1] + Cy[id]) - (Cy[p2] + Cy[m1])) * SXTN * DY;
        f_z = mue * (27.0 * (w[p1] - w[id]) - (w[p2] - w[m1])) * _dl * TWFR * DZ;
        f_z += mue * (9.0 * (Cz[p1] + Cz[id]) - (Cz[p2] + Cz[m1])) * SXTN * DZ;
        u_hlf = (9.0 * (u[p1] + u[id]) - (u[p2] + u[m1])) * SXTN;
        v_hlf = (9.0 * (v[p1] + v[id]) - (v[p2] + v[m1])) * SXTN * DY;
        w_hlf = (9.0 * (w[p1] + w[id]) - (w[p2] + w[m1])) * SXTN * DZ;
    } else if (d == 1) {
        const double *Dvcx = Vd + DVCX * n, *Dvcz = Vd + DVCZ * n, *Ducx = Vd + DUCX * n, *Dwcz = Vd + DWCZ * n;
        f_x = mue * (27.0 * (u[p1] - u[id]) - (u[p2] - u[m1])) * _dl * TWFR * DX;
        f_x += mue * (9.0 * (Dvcx[p1] + Dvcx[id]) - (Dvcx[p2] + Dvcx[m1])) * SXTN * DX;
        f_y = (2.0 * mue + lamada) * (27.0 * (v[p1] - v[id]) - (v[p2] - v[m1])) * _dl * TWFR;
        f_y += lamada * (9.0 * (Ducx[p1] + Ducx[id]) - (Ducx[p2] + Ducx[m1]) + 9.0 * (Dwcz[p1] + Dwcz[id]) -
                         (Dwcz[p2] + Dwcz[m1])) * SXTN;
        f_z = mue * (27.0 * (w[p1] - w[id]) - (w[p2] - w[m1])) * _dl * TWFR * DZ;
        f_z += mue * (9.0 * (Dvcz[p1] + Dvcz[id]) - (Dvcz[p2] + Dvcz[m1])) * SXTN * DZ;
        u_hlf = (9.0 * (u[p1] + u[id]) - (u[p2] + u[m1])) * SXTN * DX;
        v_hlf = (9.0 * (v[p1] + v[id]) - (v[p2] + v[m1])) * SXTN;
        w_hlf = (9.0 * (w[p1] + w[id]) - (w[p2] + w[m1])) * SXTN * DZ;
    } else {
        const double *Dwcx = Vd + DWCX * n, *Dwcy = Vd + DWCY * n, *Ducx = Vd + DUCX * n, *Dvcy = Vd + DVCY * n;
        f_x = mue * (27.0 * (u[p1] - u[id]) - (u[p2] - u[m1])) * _dl * TWFR * DX;
        f_x += mue * (9.0 * (Dwcx[p1] + Dwcx[id]) - (Dwcx[p2] + Dwcx[m1])) * SXTN * DX;
        f_y = mue * (27.0 * (v[p1] - v[id]) - (v[p2] - v[m1])) * _dl * TWFR * DY;
        f_y += mue * (9.0 * (Dwcy[p1] + Dwcy[id]) - (Dwcy[p2] + Dwcy[m1])) * SXTN * DY;
        f_z = (2.0 * mue + lamada) * (27.0 * (w[p1] - w[id]) - (w[p2] - w[m1])) * _dl * TWFR;
        f_z += lamada * (9.0 * (Ducx[p1] + Ducx[id]) - (Ducx[p2] + Ducx[m1]) + 9.0 * (Dvcy[p1] + Dvcy[id]) -
                         (Dvcy[p2] + Dvcy[m1])) * SXTN;
        u_hlf = (9.0 * (u[p1] + u[id]) - (u[p2] + u[m1])) * SXTN * DX;
        v_hlf = (9.0 * (v[p1] + v[id]) - (v[p2] + v[m1])) * SXTN * DY;
        w_hlf = (9.0 * (w[p1] + w[id]) - (w[p2] + w[m1])) * SXTN;
    }
    double F[ORA_MAXE];
    F[0] = 0.0;
    F[1] = f_x;
    F[2] = f_y;
    F[3] = f_z;
    F[4] = f_x * u_hlf + f_y * v_hlf + f_z * w_hlf;
    if (c->visc_heat) {
        double kk = HALF(f->tca);
        kk *= GRAD(f->T);
        F[4] += kk;
    }
    if (c->visc_diffu) {
        const double rho_wall = HALF(f->rho);
        double CorrectTerm = 0.0, Dim_Yil = 1.0e-20, Yi_wall[ORA_MAXNS];
        for (int l = 0; l < NS; l++) {
            const double *hl = f->hi, *Dl = f->Dkm + (size_t)l * n, *Y = f->y;
            /* hi and y are species-interleaved (NS * id + l) */
            const double hi_wall =
                (9.0 * (hl[l + NS * p1] + hl[l + NS * id]) - (hl[l + NS * p2] + hl[l + NS * m1])) * SXTN;
            const double Dim_wall = (9.0 * (Dl[p1] + Dl[id]) - (Dl[p2] + Dl[m1])) * SXTN;
            const double Yil_wall =
                smin(smax((27.0 * (Y[l + NS * p1] - Y[l + NS * id]) - (Y[l + NS * p2] - Y[l + NS * m1])) * _dl * TWFR,
                          -yil_lim[l]),
                     yil_lim[l]);
            Yi_wall[l] =
                smin(smax((9.0 * (Y[l + NS * p1] + Y[l + NS * id]) - (Y[l + NS * p2] + Y[l + NS * m1])) * SXTN, 1.0E-20),
                     1.0);
            Dim_Yil = smin(smax(Dim_wall * Yil_wall, -diffu_lim[l]), diffu_lim[l]);
            CorrectTerm += Dim_Yil;
            F[4] += rho_wall * hi_wall * Dim_Yil;
        }
        CorrectTerm *= rho_wall;
        for (int q = 5; q < E; q++) F[q] = rho_wall * Dim_Yil - Yi_wall[q - 5] * CorrectTerm;
    } else {
        for (int q = 5; q < E; q++) F[q] = 0.0;
    }
    for (int q = 0; q < E; q++) Fwall[q + (size_t)E * id] -= F[q];
#undef HALF
#undef GRAD
}

/* the viscous part of GetLU (ConVenction_block.hpp:424-575; limits :462-503) for the state
   whose primitives the last UpdateStates left in f */
void ora_viscous_fluxes(ora_fluid *f)
{
    const ora_config *c = &f->cfg;
    const int NS = c->NS;
    const long n = (long)f->ncell;
    cell_centre_derivative(f);
#pragma omp parallel for schedule(static)
    for (long id = 0; id < n; id++) transport_cell(f, (size_t)id);
    double yil_lim[ORA_MAXNS], diffu_lim[ORA_MAXNS];
    if (c->visc_diffu) {
        for (int nn = 0; nn < NS; nn++) {
            /* sycl reductions combining with the stored 0.0 */
            double ymin = 0.0, ymax = 0.0, dmax = 0.0;
            for (int k = c->Bwidth_Z; k < c->Bwidth_Z + c->Z_inner; k++)
                for (int j = c->Bwidth_Y; j < c->Bwidth_Y + c->Y_inner; j++)
                    for (int i = c->Bwidth_X; i < c->Bwidth_X + c->X_inner; i++) {
                        const double y = f->y[(size_t)NS * IDX(f, i, j, k) + nn];
                        ymin = smin(ymin, y), ymax = smax(ymax, y);
                    }
            ymax -= ymin;
            ymax *= c->Yil_limiter;
            dmax *= c->Dim_limiter * ymax;
            yil_lim[nn] = ymax, diffu_lim[nn] = dmax;
        }
    }
    const int dims[3] = {c->DimX, c->DimY, c->DimZ};
    double *Fw[3] = {f->FluxFw, f->FluxGw, f->FluxHw};
    for (int d = 0; d < 3; d++) {
        if (!dims[d]) continue;
        const int i0 = c->Bwidth_X - (d == 0), j0 = c->Bwidth_Y - (d == 1), k0 = c->Bwidth_Z - (d == 2);
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = k0; k < c->Bwidth_Z + c->Z_inner; k++)
            for (int j = j0; j < c->Bwidth_Y + c->Y_inner; j++)
                for (int i = i0; i < c->Bwidth_X + c->X_inner; i++)
                    wall_flux(f, d, IDX(f, i, j, k), yil_lim, diffu_lim, Fw[d]);
    }
}
