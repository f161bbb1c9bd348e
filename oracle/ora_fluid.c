/*
 * Fluid stages of the oracle (test infrastructure only): conservative to
 * primitive conversion, physical fluxes, ghost fill, CFL dt, characteristic
 * WENO face fluxes, flux divergence and the SSP-RK3 update.
 *
 * Restates, operation by operation:
 *   src/solver_UpdateStates/Update_device.hpp, Update_kernels.hpp,
 *   UpdateStates_block.cpp, Estimate_kernels.hpp;
 *   src/solver_GetDt/GlobalDt_block.hpp;
 *   src/solver_BCs/BCs_kernels.hpp, BCs_block.cpp;
 *   src/include/global_marco.h (MARCO_ROE);
 *   src/solver_Reconstruction/FDM_Method/ConVenction_block.hpp,
 *   Reconstruction_kernels.hpp, positive-definite_eigen/{Eigen_value.hpp,
 *   Eigen_callback.h, Eigen_matrix.hpp, Utils_device.hpp};
 *   src/solver_Reconstruction/schemes/WENO5s_schemes.hpp.
 *
 * Pins (tests/test_oracle_fluid.py):
 *   - L R = I for every Roe system built here (NO-COP and COP, x/y/z);
 *   - WENO5-JS / WENO5-Z: ideal weights on smooth data reduce to the linear
 *     5th-order upwind stencil; exact on quadratics;
 *   - isentropic vortex: error vs the analytic translated vortex decreases
 *     at the scheme's order under refinement;
 *   - conservation of sum(U) to round-off on periodic grids;
 *   - dt: brute-force max over the padded grid (numpy);
 *   - ghost fill: numpy restatement of each BC type.
 *   ReconstructSoundSpeed / SoundSpeedMultiSpecies (multi-species Roe c^2):
 *   between equal states it is (gamma - 1)(q^2/2 + p/rho + c_p T - h), and it
 *   is symmetric in the two cells (tests/test_oracle_recon.py); its
 *   dp/drho, dp/drho_n and dp/d(rho E) equal central finite differences of
 *   the equation of state through the Newton temperature, and on faces with
 *   real jumps the corrected derivatives reproduce the pressure jump, with
 *   the exact residual the energy term's rho_bar carries
 *   (tests/test_oracle_pins.py; planted mistakes in either fail).
 */
#include "ora_internal.h"
#include <stdlib.h>
#include <string.h>

double ora_mix_gamma_cv(const ora_config *c, const double *yi, double T);
double ora_mix_gamma_cp(const ora_config *c, const double *yi, double Cp);

#define IDX(f, i, j, k) ((size_t)(f)->cfg.Xmax * (f)->cfg.Ymax * (k) + (size_t)(f)->cfg.Xmax * (j) + (i))

/* ------------------------------------------------------------------ */
/* lifecycle                                                          */
/* ------------------------------------------------------------------ */
ora_fluid *ora_create(const ora_config *cfg)
{
    ora_fluid *f = (ora_fluid *)calloc(1, sizeof(ora_fluid));
    f->cfg = *cfg;
    f->E = cfg->cop ? cfg->NS + 4 : 5;
    f->NCOP = cfg->cop ? cfg->NS - 1 : 0;
    f->ncell = (size_t)cfg->Xmax * cfg->Ymax * cfg->Zmax;
    size_t nE = f->ncell * f->E;
    f->U = calloc(nE, 8); f->U1 = calloc(nE, 8); f->LU = calloc(nE, 8);
    f->FluxF = calloc(nE, 8); f->FluxG = calloc(nE, 8); f->FluxH = calloc(nE, 8);
    f->FluxFw = calloc(nE, 8); f->FluxGw = calloc(nE, 8); f->FluxHw = calloc(nE, 8);
    f->eigen_local_x = calloc(nE, 8); f->eigen_local_y = calloc(nE, 8); f->eigen_local_z = calloc(nE, 8);
    size_t n = f->ncell;
    f->rho = calloc(n, 8); f->p = calloc(n, 8); f->u = calloc(n, 8); f->v = calloc(n, 8);
    f->w = calloc(n, 8); f->e = calloc(n, 8); f->gamma = calloc(n, 8); f->T = calloc(n, 8);
    f->H = calloc(n, 8); f->c = calloc(n, 8); f->Cp = calloc(n, 8); f->Ri = calloc(n, 8);
    f->y = calloc(n * cfg->NS, 8);
    /* eigen_block zeroed once at allocation: src/Fluids.cpp:421-429 */
    memset(f->eigen_block, 0, sizeof(f->eigen_block));
    memset(f->uvw_c_max, 0, sizeof(f->uvw_c_max));
    f->target = HUGE_VAL;
    ora_visc_alloc(f);
    return f;
}

void ora_destroy(ora_fluid *f)
{
    if (!f) return;
    double *ptrs[] = {f->U, f->U1, f->LU, f->FluxF, f->FluxG, f->FluxH, f->FluxFw, f->FluxGw, f->FluxHw,
                      f->eigen_local_x, f->eigen_local_y, f->eigen_local_z, f->rho, f->p, f->u, f->v, f->w,
                      f->e, f->gamma, f->T, f->H, f->c, f->Cp, f->Ri, f->y};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    ora_visc_free(f);
    free(f);
}

int ora_num_eqn(const ora_fluid *f) { return f->E; }
size_t ora_num_cells(const ora_fluid *f) { return f->ncell; }

double *ora_field(ora_fluid *f, const char *name)
{
#define F(nm) if (!strcmp(name, #nm)) return f->nm;
    F(U) F(U1) F(LU) F(rho) F(p) F(u) F(v) F(w) F(e) F(y) F(gamma) F(T) F(H) F(c) F(Cp) F(Ri)
    F(FluxF) F(FluxG) F(FluxH) F(FluxFw) F(FluxGw) F(FluxHw)
    F(eigen_local_x) F(eigen_local_y) F(eigen_local_z) F(eigen_block) F(uvw_c_max)
    F(Vde) F(va) F(tca) F(Dkm) F(hi)
#undef F
    return NULL;
}

/* ------------------------------------------------------------------ */
/* UpdateStates                                                       */
/* ------------------------------------------------------------------ */

/* Getrhoyi: src/solver_UpdateStates/Update_device.hpp:7-29.  Default: the
   last species is the remainder 1 - sum.  GhostSpecies (:15-21): the last
   species is a ghost with Y = 0, the others are renormalised to sum 1 and
   written back into the state. */
static void getrhoyi(const ora_fluid *f, double *UI, double *rho, double *yi)
{
    *rho = UI[0];
    double rho1 = 1.0 / *rho;
    if (f->cfg.cop && f->cfg.ghost) {
        yi[f->NCOP] = 0.0;
        double sum_yi = 0.0;
        for (int ii = 0; ii < f->NCOP; ii++)
            yi[ii] = UI[ii + 5] * rho1, sum_yi += yi[ii];
        sum_yi = 1.0 / sum_yi;
        for (int ii = 0; ii < f->NCOP; ii++)
            yi[ii] *= sum_yi, UI[ii + 5] = *rho * yi[ii];
    } else if (f->cfg.cop) {
        yi[f->NCOP] = 1.0;
        for (int ii = 5; ii < f->E; ii++)
            yi[ii - 5] = UI[ii] * rho1, yi[f->NCOP] += -yi[ii - 5];
    }
}

/* GetStates: src/solver_UpdateStates/Update_device.hpp:33-54 */
static void getstates(const ora_fluid *f, const double *UI, double rho, double *u, double *v, double *w,
                      double *p, double *H, double *c, double *gamma, double *T, double *e, double *Cp,
                      double *R, const double *yi)
{
    const ora_config *cf = &f->cfg;
    double rho1 = 1.0 / rho;
    *u = UI[1] * rho1, *v = UI[2] * rho1, *w = UI[3] * rho1;
    double tme = UI[4] * rho1 - 0.5 * (*u * *u + *v * *v + *w * *w);
    if (cf->cop) {
        double R_ = ora_mix_R(cf, yi);
        *T = ora_get_T(cf, yi, tme, *T);
        *p = rho * R_ * *T, *R = R_;
        *Cp = ora_mix_cp(cf, yi, *T);
        *gamma = ora_mix_gamma_cp(cf, yi, *Cp);
    } else {
        *gamma = cf->ncop_gamma;
        *p = (cf->ncop_gamma - 1.0) * rho * tme;
    }
    *H = (UI[4] + *p) * rho1;
    *c = sqrt(*gamma * *p * rho1);
    *e = tme;
}

/* GetPhysFlux: src/solver_UpdateStates/Update_device.hpp:81-110 */
static void getphysflux(const ora_fluid *f, const double *UI, const double *yi, double *FF, double *FG,
                        double *FH, double u, double v, double w, double p)
{
    FF[0] = UI[1];
    FF[1] = UI[1] * u + p;
    FF[2] = UI[1] * v;
    FF[3] = UI[1] * w;
    FF[4] = (UI[4] + p) * u;
    FG[0] = UI[2];
    FG[1] = UI[2] * u;
    FG[2] = UI[2] * v + p;
    FG[3] = UI[2] * w;
    FG[4] = (UI[4] + p) * v;
    FH[0] = UI[3];
    FH[1] = UI[3] * u;
    FH[2] = UI[3] * v;
    FH[3] = UI[3] * w + p;
    FH[4] = (UI[4] + p) * w;
    if (f->cfg.cop)
        for (int ii = 5; ii < f->E; ii++) {
            FF[ii] = UI[1] * yi[ii - 5];
            FG[ii] = UI[2] * yi[ii - 5];
            FH[ii] = UI[3] * yi[ii - 5];
        }
}

/* UpdateFluidStateFlux: src/solver_UpdateStates/UpdateStates_block.cpp:7-191
   (ESTIM_NAN on, no ERROR_PATCH).  Returns 0, or 1 for a rho/Y error
   (EstimateYiKernel, Estimate_kernels.hpp:5-102), or 2 for a rho/p/T error
   (EstimatePrimitiveVarKernel, Estimate_kernels.hpp:104-162). */
int ora_update_states(ora_fluid *f, int which)
{
    const ora_config *cf = &f->cfg;
    double *UI = which ? f->U1 : f->U;
    const int E = f->E, NS = cf->NS;
    long n = (long)f->ncell;
#pragma omp parallel for schedule(static)
    for (long id = 0; id < n; id++)
        getrhoyi(f, UI + E * id, &f->rho[id], f->y + NS * id);

    int err = 0;
    /* rho / Y validation over interior cells */
    for (int k = cf->Bwidth_Z; k < cf->Zmax - cf->Bwidth_Z; k++)
        for (int j = cf->Bwidth_Y; j < cf->Ymax - cf->Bwidth_Y; j++)
            for (int i = cf->Bwidth_X; i < cf->Xmax - cf->Bwidth_X; i++) {
                size_t id = IDX(f, i, j, k);
                double r = f->rho[id];
                int bad = (r < 0) || isnan(r) || isinf(r);
                if (cf->cop)
                    for (int s = 0; s < NS; s++) {
                        double yv = f->y[NS * id + s];
                        bad = bad || isnan(yv) || isinf(yv);
                    }
                if (bad && !err) err = 1, f->err_code = 1, f->err_id = id; /* ascending ids: the lowest */
            }
    if (err) return 1;

#pragma omp parallel for schedule(static)
    for (long id = 0; id < n; id++) {
        const double *U = UI + E * id;
        double *yi = f->y + NS * id;
        getstates(f, U, f->rho[id], &f->u[id], &f->v[id], &f->w[id], &f->p[id], &f->H[id], &f->c[id],
                  &f->gamma[id], &f->T[id], &f->e[id], &f->Cp[id], &f->Ri[id], yi);
        getphysflux(f, U, yi, f->FluxF + E * id, f->FluxG + E * id, f->FluxH + E * id, f->u[id], f->v[id],
                    f->w[id], f->p[id]);
    }

    for (int k = cf->Bwidth_Z; k < cf->Zmax - cf->Bwidth_Z; k++)
        for (int j = cf->Bwidth_Y; j < cf->Ymax - cf->Bwidth_Y; j++)
            for (int i = cf->Bwidth_X; i < cf->Xmax - cf->Bwidth_X; i++) {
                size_t id = IDX(f, i, j, k);
                double vs[3] = {f->rho[id], f->p[id], f->T[id]};
                for (int q = 0; q < 3; q++)
                    if ((vs[q] < 0 || isnan(vs[q]) || isinf(vs[q])) && !err) err = 2, f->err_code = 2, f->err_id = id;
            }
    return err;
}

/* ------------------------------------------------------------------ */
/* Boundary conditions: src/solver_BCs/BCs_kernels.hpp:9-258,          */
/* launch ranges and ordering src/solver_BCs/BCs_block.cpp:50-212      */
/* ------------------------------------------------------------------ */
static void bc_apply(ora_fluid *f, double *U, size_t id, size_t tid, int bc, int mom)
{
    const int E = f->E;
    switch (bc) {
    case ORA_SYMMETRY:
        for (int n = 0; n < E; n++) U[E * id + n] = U[E * tid + n];
        U[E * id + mom] = -U[E * tid + mom];
        break;
    case ORA_PERIODIC:
    case ORA_OUTFLOW:
        for (int n = 0; n < E; n++) U[E * id + n] = U[E * tid + n];
        break;
    case ORA_NSLIPWALL:
    case ORA_VISCWALL:
    case ORA_SLIPWALL:
        U[E * id + 0] = U[E * tid + 0];
        U[E * id + 1] = -U[E * tid + 1];
        U[E * id + 2] = -U[E * tid + 2];
        U[E * id + 3] = -U[E * tid + 3];
        U[E * id + 4] = U[E * tid + 4];
        for (int n = E - f->NCOP; n < E; n++) U[E * id + n] = U[E * tid + n];
        break;
    default: /* Inflow, innerBlock: ghosts keep their values */
        break;
    }
}

/* target index along one axis for ghost coordinate g (FluidBCKernel{X,Y,Z}) */
static int bc_target(int bc, int g, int B, int inner, int mirror_offset, int index_inner, int sign)
{
    switch (bc) {
    case ORA_SYMMETRY:
    case ORA_NSLIPWALL:
    case ORA_VISCWALL:
    case ORA_SLIPWALL:
        return 2 * (B + mirror_offset) - 1 - g;
    case ORA_PERIODIC:
        return g + sign * inner;
    case ORA_OUTFLOW:
        return index_inner;
    default:
        return g;
    }
}

void ora_boundary(ora_fluid *f, int which)
{
    const ora_config *c = &f->cfg;
    double *U = which ? f->U1 : f->U;
    if (c->DimX) {
        for (int k = 0; k < c->Zmax; k++)
            for (int j = 0; j < c->Ymax; j++)
                for (int g = 0; g < c->Bwidth_X; g++) {
                    int i0 = g, i1 = g + c->Xmax - c->Bwidth_X;
                    int t0 = bc_target(c->BC[0], i0, c->Bwidth_X, c->X_inner, 0, c->Bwidth_X, 1);
                    int t1 = bc_target(c->BC[1], i1, c->Bwidth_X, c->X_inner, c->X_inner, c->Xmax - c->Bwidth_X - 1, -1);
                    bc_apply(f, U, IDX(f, i0, j, k), IDX(f, t0, j, k), c->BC[0], 1);
                    bc_apply(f, U, IDX(f, i1, j, k), IDX(f, t1, j, k), c->BC[1], 1);
                }
    }
    if (c->DimY) {
        for (int k = 0; k < c->Zmax; k++)
            for (int g = 0; g < c->Bwidth_Y; g++)
                for (int i = 0; i < c->Xmax; i++) {
                    int j0 = g, j1 = g + c->Ymax - c->Bwidth_Y;
                    int t0 = bc_target(c->BC[2], j0, c->Bwidth_Y, c->Y_inner, 0, c->Bwidth_Y, 1);
                    int t1 = bc_target(c->BC[3], j1, c->Bwidth_Y, c->Y_inner, c->Y_inner, c->Ymax - c->Bwidth_Y - 1, -1);
                    bc_apply(f, U, IDX(f, i, j0, k), IDX(f, i, t0, k), c->BC[2], 2);
                    bc_apply(f, U, IDX(f, i, j1, k), IDX(f, i, t1, k), c->BC[3], 2);
                }
    }
    if (c->DimZ) {
        for (int g = 0; g < c->Bwidth_Z; g++)
            for (int j = 0; j < c->Ymax; j++)
                for (int i = 0; i < c->Xmax; i++) {
                    int k0 = g, k1 = g + c->Zmax - c->Bwidth_Z;
                    int t0 = bc_target(c->BC[4], k0, c->Bwidth_Z, c->Z_inner, 0, c->Bwidth_Z, 1);
                    int t1 = bc_target(c->BC[5], k1, c->Bwidth_Z, c->Z_inner, c->Z_inner, c->Zmax - c->Bwidth_Z - 1, -1);
                    bc_apply(f, U, IDX(f, i, j, k0), IDX(f, i, j, t0), c->BC[4], 3);
                    bc_apply(f, U, IDX(f, i, j, k1), IDX(f, i, j, t1), c->BC[5], 3);
                }
    }
}

/* ------------------------------------------------------------------ */
/* GetDt: src/solver_GetDt/GlobalDt_block.hpp:6-114                    */
/* ------------------------------------------------------------------ */
double ora_get_dt(ora_fluid *f)
{
    const ora_config *c = &f->cfg;
    for (int n = 0; n < 6; n++) f->uvw_c_max[n] = 0.0;
    const double *vel[3] = {f->u, f->v, f->w};
    const int dims[3] = {c->DimX, c->DimY, c->DimZ};
    long n = (long)f->ncell;
    for (int d = 0; d < 3; d++) {
        if (!dims[d]) continue;
        double m = 0.0;
#pragma omp parallel for reduction(max : m)
        for (long id = 0; id < n; id++) {
            double c_local = sqrt(1.4 * f->p[id] / f->rho[id]);
            double v = fabs(vel[d][id]) + c_local;
            if (v > m) m = v;
        }
        f->uvw_c_max[d] = m;
    }
    double dtref = f->uvw_c_max[0] * c->_dx + f->uvw_c_max[1] * c->_dy + f->uvw_c_max[2] * c->_dz;
    return c->CFLnumber / dtref;
}

/* ------------------------------------------------------------------ */
/* SSP-RK3: src/solver_UpdateStates/Update_kernels.hpp:64-94           */
/* ------------------------------------------------------------------ */
void ora_rk3(ora_fluid *f, int flag, double dt)
{
    const ora_config *c = &f->cfg;
    const int E = f->E;
    const double OT = 1.0 / 3.0;
    for (int k = c->Bwidth_Z; k < c->Bwidth_Z + c->Z_inner; k++)
        for (int j = c->Bwidth_Y; j < c->Bwidth_Y + c->Y_inner; j++)
            for (int i = c->Bwidth_X; i < c->Bwidth_X + c->X_inner; i++) {
                size_t id = IDX(f, i, j, k);
                double *U = f->U + E * id, *U1 = f->U1 + E * id, *LU = f->LU + E * id;
                for (int n = 0; n < E; n++) {
                    if (flag == 1) U1[n] = U[n] + dt * LU[n];
                    else if (flag == 2) U1[n] = 0.75 * U[n] + 0.25 * U1[n] + 0.25 * dt * LU[n];
                    else U[n] = (U[n] + 2.0 * U1[n] + 2.0 * dt * LU[n]) * OT;
                }
            }
}

/* ------------------------------------------------------------------ */
/* Reconstruction                                                     */
/* ------------------------------------------------------------------ */

/* weno5old_BODY (USE_DOUBLE): src/solver_Reconstruction/schemes/WENO5s_schemes.hpp:12-69 */
static double weno5old_body(double v1, double v2, double v3, double v4, double v5)
{
    double a1, a2, a3;
    double dtwo = 2.0, dtre = 3.0;
    a1 = v1 - dtwo * v2 + v3;
    double s1 = 13.0 * a1 * a1;
    a1 = v1 - 4.0 * v2 + dtre * v3;
    s1 += dtre * a1 * a1;
    a1 = v2 - dtwo * v3 + v4;
    double s2 = 13.0 * a1 * a1;
    a1 = v2 - v4;
    s2 += dtre * a1 * a1;
    a1 = v3 - dtwo * v4 + v5;
    double s3 = 13.0 * a1 * a1;
    a1 = dtre * v3 - 4.0 * v4 + v5;
    s3 += dtre * a1 * a1;
    double tol = 1.0E-6;
    s1 += tol, s2 += tol, s3 += tol;
    a1 = 0.1 * s2 * s2 * s3 * s3;
    a2 = 0.6 * s1 * s1 * s3 * s3;
    a3 = 0.3 * s1 * s1 * s2 * s2;
    double tw1 = 1.0 / (a1 + a2 + a3);
    a1 = a1 * tw1, a2 = a2 * tw1, a3 = a3 * tw1;
    s1 = a1 * (dtwo * v1 - 7.0 * v2 + 11.0 * v3);
    s2 = a2 * (-v2 + 5.0 * v3 + dtwo * v4);
    s3 = a3 * (dtwo * v3 + 5.0 * v4 - v5);
    return (s1 + s2 + s3);
}

/* weno5old_GPU(&pp[3], &mm[3]): WENO5s_schemes.hpp:71-95 */
double ora_weno5js(const double *pp, const double *mm)
{
    static const double six = 1.0 / 6.0; /* _six, Utils_schemes.hpp:5 */
    double temf = weno5old_body(pp[1], pp[2], pp[3], pp[4], pp[5]);
    double temm = weno5old_body(mm[6], mm[5], mm[4], mm[3], mm[2]);
    return (temf + temm) * six;
}

/* weno5Z_P / weno5Z_M body: WENO5s_schemes.hpp:246-316, in the reference's
   expression order (each product rounded, then each sum).  The reference
   writes every indicator difference twice, 13.0 * (d) * (d); both copies
   round to the same value, so it is formed once. */
static double weno5z_body(double v1, double v2, double v3, double v4, double v5)
{
    double s1, s2, s3, a1, a2, a3, w1, w2, w3;
    double epsilon = 1.0e-6;
    double d1 = v1 - 2.0 * v2 + v3, e1 = v1 - 4.0 * v2 + 3.0 * v3;
    double d2 = v2 - 2.0 * v3 + v4, e2 = v2 - v4;
    double d3 = v3 - 2.0 * v4 + v5, e3 = 3.0 * v3 - 4.0 * v4 + v5;
    s1 = 13.0 * d1 * d1 + 3.0 * e1 * e1;
    s2 = 13.0 * d2 * d2 + 3.0 * e2 * e2;
    s3 = 13.0 * d3 * d3 + 3.0 * e3 * e3;
    double s5 = fabs(s1 - s3);
    a1 = 0.1 * (1.0 + s5 / (s1 + epsilon));
    a2 = 0.6 * (1.0 + s5 / (s2 + epsilon));
    a3 = 0.3 * (1.0 + s5 / (s3 + epsilon));
    w1 = a1 / (a1 + a2 + a3);
    w2 = a2 / (a1 + a2 + a3);
    w3 = a3 / (a1 + a2 + a3);
    return w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3) / 6.0 + w2 * (-v2 + 5.0 * v3 + 2.0 * v4) / 6.0 +
           w3 * (2.0 * v3 + 5.0 * v4 - v5) / 6.0;
}

/* The same body with each multiply-add pair of the indicators and candidate
   polynomials fused (a*b + c -> fma(a, b, c), same association): the
   arithmetic of the fused FP64 device form, not the reference's. */
static double weno5z_body_fused(double v1, double v2, double v3, double v4, double v5)
{
    double s1, s2, s3, a1, a2, a3, w1, w2, w3;
    double epsilon = 1.0e-6;
    double d1 = fma(-2.0, v2, v1) + v3, e1 = fma(3.0, v3, fma(-4.0, v2, v1)); /* v1 - 2v2 + v3, v1 - 4v2 + 3v3 */
    double d2 = fma(-2.0, v3, v2) + v4, e2 = v2 - v4;
    double d3 = fma(-2.0, v4, v3) + v5, e3 = fma(-4.0, v4, 3.0 * v3) + v5;  /* 3v3 - 4v4 + v5 */
    s1 = fma(13.0 * d1, d1, (3.0 * e1) * e1);                               /* 13 d^2 + 3 e^2 */
    s2 = fma(13.0 * d2, d2, (3.0 * e2) * e2);
    s3 = fma(13.0 * d3, d3, (3.0 * e3) * e3);
    double s5 = fabs(s1 - s3);
    a1 = 0.1 * (1.0 + s5 / (s1 + epsilon));
    a2 = 0.6 * (1.0 + s5 / (s2 + epsilon));
    a3 = 0.3 * (1.0 + s5 / (s3 + epsilon));
    w1 = a1 / (a1 + a2 + a3);
    w2 = a2 / (a1 + a2 + a3);
    w3 = a3 / (a1 + a2 + a3);
    return w1 * fma(11.0, v3, fma(-7.0, v2, 2.0 * v1)) / 6.0 + w2 * fma(2.0, v4, fma(5.0, v3, -v2)) / 6.0 +
           w3 * (fma(5.0, v4, 2.0 * v3) - v5) / 6.0;
}

double ora_weno5z(const double *pp, const double *mm)
{
    return weno5z_body(pp[1], pp[2], pp[3], pp[4], pp[5]) + weno5z_body(mm[6], mm[5], mm[4], mm[3], mm[2]);
}

double ora_weno5z_fused(const double *pp, const double *mm)
{
    return weno5z_body_fused(pp[1], pp[2], pp[3], pp[4], pp[5]) +
           weno5z_body_fused(mm[6], mm[5], mm[4], mm[3], mm[2]);
}

/* WENOCU6_BODYGPU: src/solver_Reconstruction/schemes/WENO6s_schemes.hpp:5-51,
   constants Utils_schemes.hpp:5-15 and src/include/global_setup.h:36-37 */
static double wenocu6_body(double v1, double v2, double v3, double v4, double v5, double v6, double epsilon)
{
    const double ohtz = 1.0 / 120.0, ohff = 1.0 / 144.0, ftss = 1.0 / 5760.0, sxtn = 1.0 / 16.0, twfr = 1.0 / 24.0;
    const double six = 1.0 / 6.0;
    const double wa2a2 = 13.0 / 3.0, wa3a3 = 3129.0 / 80.0, wa4a4 = 87617.0 / 140.0, wa3a5 = 14127.0 / 224.0,
                 wa5a5 = 252337135.0 / 16128.0;
    double s11 = v1 - 2.0 * v2 + v3;
    double s12 = v1 - 4.0 * v2 + 3.0 * v3;
    double s1 = 13.0 * s11 * s11 + 3.0 * s12 * s12;
    double s21 = v2 - 2.0 * v3 + v4;
    double s22 = v2 - v4;
    double s2 = 13.0 * s21 * s21 + 3.0 * s22 * s22;
    double s31 = v3 - 2.0 * v4 + v5;
    double s32 = 3.0 * v3 - 4.0 * v4 + v5;
    double s3 = 13.0 * s31 * s31 + 3.0 * s32 * s32;
    double tau61 = (259.0 * v6 - 1895.0 * v5 + 6670.0 * v4 - 2590.0 * v3 - 2785.0 * v2 + 341.0 * v1) * ftss;
    double tau62 = -(v5 - 12.0 * v4 + 22.0 * v3 - 12.0 * v2 + v1) * sxtn;
    double tau63 = -(7.0 * v6 - 47.0 * v5 + 94.0 * v4 - 70.0 * v3 + 11.0 * v2 + 5.0 * v1) * ohff;
    double tau64 = (v5 - 4.0 * v4 + 6.0 * v3 - 4.0 * v2 + v1) * twfr;
    double tau65 = -(-v6 + 5.0 * v5 - 10.0 * v4 + 10.0 * v3 - 5.0 * v2 + v1) * ohtz;
    double a1a1 = 1.0, a2a2 = wa2a2, a1a3 = 0.5, a3a3 = wa3a3, a2a4 = 4.2;
    double a1a5 = 0.125, a4a4 = wa4a4, a3a5 = wa3a5, a5a5 = wa5a5;
    double s6 = (tau61 * tau61 * a1a1 + tau62 * tau62 * a2a2 + tau61 * tau63 * a1a3 + tau63 * tau63 * a3a3 +
                 tau62 * tau64 * a2a4 + tau61 * tau65 * a1a5 + tau64 * tau64 * a4a4 + tau63 * tau65 * a3a5 +
                 tau65 * tau65 * a5a5) * 12.0;
    double s55 = (s1 + s3 + 4.0 * s2) * six;
    double s5 = fabs(s6 - s55);
    double r0 = 20.0;
    double r1 = r0 + s5 / (s1 + epsilon);
    double r2 = r0 + s5 / (s2 + epsilon);
    double r3 = r0 + s5 / (s3 + epsilon);
    double r4 = r0 + s5 / (s6 + epsilon);
    double a1 = 0.05 * r1;
    double a2 = 0.45 * r2;
    double a3 = 0.45 * r3;
    double a4 = 0.05 * r4;
    double tw1 = 1.0 / (a1 + a2 + a3 + a4);
    double w1 = a1 * tw1;
    double w2 = a2 * tw1;
    double w3 = a3 * tw1;
    double w4 = a4 * tw1;
    double temp = 0.0;
    temp += w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3);
    temp += w2 * (-v2 + 5.0 * v3 + 2.0 * v4);
    temp += w3 * (2.0 * v3 + 5.0 * v4 - v5);
    temp += w4 * (11.0 * v4 - 7.0 * v5 + 2.0 * v6);
    return temp;
}

/* WENOCU6_GPU(&pp[3], &mm[3], dl): WENO6s_schemes.hpp:54-80; epsilon = 1e-8 dl^2 */
double ora_wenocu6(const double *pp, const double *mm, double delta)
{
    const double six = 1.0 / 6.0;
    const double epsilon = 1.e-8 * delta * delta;
    double temf = wenocu6_body(pp[1], pp[2], pp[3], pp[4], pp[5], pp[6], epsilon);
    double temm = wenocu6_body(mm[6], mm[5], mm[4], mm[3], mm[2], mm[1], epsilon);
    return (temf + temm) * six;
}

/* RoeAverageLeft_{x,y,z} / RoeAverageRight_{x,y,z} (EIGEN_ALLOC == 0):
   positive-definite_eigen/Eigen_matrix.hpp:7-455, with MARCO_PREEIGEN
   (Eigen_callback.h:102-106) evaluated per call as in the reference. */
void ora_eigen_lr(const ora_config *c, int dir, int n, double *L, double *R, double *eigen_value,
                  const double *z, const double *yi, double c2, double _rho, double _u, double _v, double _w,
                  double _H, double b1, double b3)
{
    const int E = c->cop ? c->NS + 4 : 5;
    const int NCOP = c->cop ? c->NS - 1 : 0;
    (void)_rho;
    /* MARCO_PREEIGEN */
    double q2 = _u * _u + _v * _v + _w * _w;
    double _c = sqrt(c2);
    double b2 = 1.0 + b1 * q2 - b1 * _H;
    double _c1 = 1.0 / _c;
    /* normal velocity and the slots of the three momentum components: (a)
       normal, (t1, t2) tangential in the order they appear in the rows */
    double un = dir == 0 ? _u : (dir == 1 ? _v : _w);
    double un_c = un * _c1;
    if (L) {
        for (int m = 0; m < E; m++) L[m] = 0.0;
        if (0 == n) {
            L[0] = 0.5 * (b2 + un_c + b3);
            L[1] = -0.5 * (b1 * _u + (dir == 0 ? _c1 : 0.0));
            L[2] = -0.5 * (b1 * _v + (dir == 1 ? _c1 : 0.0));
            L[3] = -0.5 * (b1 * _w + (dir == 2 ? _c1 : 0.0));
            if (dir != 0) L[1] = -0.5 * (b1 * _u);
            if (dir != 1) L[2] = -0.5 * (b1 * _v);
            if (dir != 2) L[3] = -0.5 * (b1 * _w);
            L[4] = 0.5 * b1;
            *eigen_value = fabs(un - _c);
            for (int m = 0; m < NCOP; m++) L[m + E - NCOP] = -0.5 * b1 * z[m];
        } else if (E - 1 == n) {
            L[0] = 0.5 * (b2 - un_c + b3);
            L[1] = dir == 0 ? 0.5 * (-b1 * _u + _c1) : 0.5 * (-b1 * _u);
            L[2] = dir == 1 ? 0.5 * (-b1 * _v + _c1) : 0.5 * (-b1 * _v);
            L[3] = dir == 2 ? 0.5 * (-b1 * _w + _c1) : 0.5 * (-b1 * _w);
            L[4] = 0.5 * b1;
            *eigen_value = fabs(un + _c);
            for (int m = 0; m < NCOP; m++) L[m + E - NCOP] = -0.5 * b1 * z[m];
        } else if (n == 1 + dir) {
            /* the (1 - b2 - b3)/b1 row sits at n = 1 (x), 2 (y), 3 (z) */
            L[0] = (1.0 - b2 - b3) / b1;
            L[1] = _u;
            L[2] = _v;
            L[3] = _w;
            L[4] = -1.0;
            *eigen_value = fabs(un);
            for (int m = 0; m < NCOP; m++) L[m + E - NCOP] = z[m];
        } else if (n >= 1 && n <= 3) {
            *eigen_value = fabs(un);
            if (dir == 0) {
                if (n == 2) { L[0] = _v; L[2] = -1.0; }
                else { L[0] = -_w; L[3] = 1.0; }
            } else if (dir == 1) {
                if (n == 1) { L[0] = -_u; L[1] = 1.0; }
                else { L[0] = _w; L[3] = -1.0; }
            } else {
                if (n == 1) { L[0] = _u; L[1] = -1.0; }
                else { L[0] = -_v; L[2] = 1.0; }
            }
        } else {
            /* species rows (COP): n = 4 .. E-2 */
            L[0] = -yi[n + c->NS - E];
            *eigen_value = fabs(un);
            for (int m = 0; m < NCOP; m++) L[m + E - NCOP] = (n + c->NS - E == m) ? 1.0 : 0.0;
        }
    }
    if (R) {
        for (int m = 0; m < E; m++) R[m] = 0.0;
        if (0 == n || E - 1 == n) {
            double sg = (0 == n) ? -1.0 : 1.0;
            R[0] = 1.0;
            R[1] = _u, R[2] = _v, R[3] = _w;
            if (0 == n) {
                if (dir == 0) R[1] = _u - _c; else if (dir == 1) R[2] = _v - _c; else R[3] = _w - _c;
                R[4] = _H - un * _c;
            } else {
                if (dir == 0) R[1] = _u + _c; else if (dir == 1) R[2] = _v + _c; else R[3] = _w + _c;
                R[4] = _H + un * _c;
            }
            (void)sg;
            for (int m = 0; m < NCOP; m++) R[m + E - NCOP] = yi[m];
        } else if (n == 1 + dir) {
            R[0] = b1;
            R[1] = _u * b1;
            R[2] = _v * b1;
            R[3] = _w * b1;
            R[4] = _H * b1 - 1.0;
            for (int m = 0; m < NCOP; m++) R[m + E - NCOP] = b1 * yi[m];
        } else if (n >= 1 && n <= 3) {
            if (dir == 0) {
                if (n == 2) { R[2] = -1.0; R[4] = -_v; }
                else { R[3] = 1.0; R[4] = _w; }
            } else if (dir == 1) {
                if (n == 1) { R[1] = 1.0; R[4] = _u; }
                else { R[3] = -1.0; R[4] = -_w; }
            } else {
                if (n == 1) { R[1] = -1.0; R[4] = -_u; }
                else { R[2] = 1.0; R[4] = _v; }
            }
        } else {
            R[4] = z[n + c->NS - E];
            for (int m = 0; m < NCOP; m++) R[m + E - NCOP] = (m == n + c->NS - E) ? 1.0 : 0.0;
        }
    }
}

/* get_DpDrho / get_DpDrhoi: positive-definite_eigen/Utils_device.hpp:14-35 */
static double dpdrho(double hN, double RN, double q2, double Cp, double R, double T, double gamma)
{
    double RNT = RN * T;
    return (gamma - 1.0) * (0.5 * q2 - hN + Cp * RNT / R);
}
#ifndef ORA_MUTANT
#define ORA_MUTANT 0 /* > 0: a planted mistake for the pin tests' mutation checks */
#endif
static double dpdrhoi(double hin, double Rin, double hiN, double RiN, double T, double Cp, double R, double gamma)
{
    double hN_minus_hi = -hin + hiN;
    double Ri_minus_RN = ORA_MUTANT == 7 ? (RiN - Rin) : (Rin - RiN);
    return (gamma - 1.0) * (hN_minus_hi + Cp * Ri_minus_RN * T / R);
}

/* ReconstructSoundSpeed + SoundSpeedMultiSpecies:
   positive-definite_eigen/Utils_device.hpp:42-79, 102-140 */
/* terms (may be NULL) receives the intermediate quantities, for the pin
   tests: see ora_face_c2_terms */
static double reconstruct_c2(const ora_fluid *f, size_t id_l, size_t id_r, double D, double D1, double _rho,
                             double _P, double *_yi, double *z, double *b1, double *b3, double *Gamma,
                             double *terms)
{
    const ora_config *c = &f->cfg;
    const int NS = c->NS, NCOP = f->NCOP;
    const double *yi_l = f->y + NS * id_l, *yi_r = f->y + NS * id_r;
    double hi_l[ORA_MAXNS], hi_r[ORA_MAXNS], _dpdrhoi[ORA_MAXNS], drhoi[ORA_MAXNS];
    const double T_l = f->T[id_l], T_r = f->T[id_r];
    for (int n = 0; n < NS; n++) {
        hi_l[n] = ora_h_i(c, T_l, n);
        hi_r[n] = ora_h_i(c, T_r, n);
        _yi[n] = (yi_l[n] + D * yi_r[n]) * D1;
    }
    double gamma_l = ora_mix_gamma_cv(c, yi_l, T_l);
    double gamma_r = ora_mix_gamma_cv(c, yi_r, T_r);
    *Gamma = (gamma_l + D * gamma_r) * D1;
    const double ul = f->u[id_l], vl = f->v[id_l], wl = f->w[id_l];
    const double ur = f->u[id_r], vr = f->v[id_r], wr = f->w[id_r];
    double q2_l = ul * ul + vl * vl + wl * wl;
    double q2_r = ur * ur + vr * vr + wr * wr;
    double R_l = ora_mix_R(c, yi_l), R_r = ora_mix_R(c, yi_r);
    double Cp_l = ora_mix_cp(c, yi_l, T_l), Cp_r = ora_mix_cp(c, yi_r, T_r);
    const double pl = f->p[id_l], pr = f->p[id_r], rl = f->rho[id_l], rr = f->rho[id_r];
    double e_l = f->H[id_l] - 0.5 * q2_l - pl / rl, e_r = f->H[id_r] - 0.5 * q2_r - pr / rr;
    double _dpdrho = (dpdrho(hi_l[NCOP], c->Ri[NCOP], q2_l, Cp_l, R_l, T_l, gamma_l) +
                      D * dpdrho(hi_r[NCOP], c->Ri[NCOP], q2_r, Cp_r, R_r, T_r, gamma_r)) * D1;
    for (int nn = 0; nn < NCOP; nn++) {
        _dpdrhoi[nn] = (dpdrhoi(hi_l[nn], c->Ri[nn], hi_l[NCOP], c->Ri[NCOP], T_l, Cp_l, R_l, gamma_l) +
                        D * dpdrhoi(hi_r[nn], c->Ri[nn], hi_r[NCOP], c->Ri[NCOP], T_r, Cp_r, R_r, gamma_r)) * D1;
        drhoi[nn] = rr * yi_r[nn] - rl * yi_l[nn];
    }
    double _prho = (pl / rl + D * (pr / rr)) * D1 +
                   0.5 * D * D1 * D1 * ((ur - ul) * (ur - ul) + (vr - vl) * (vr - vl) + (wr - wl) * (wr - wl));
    double _dpdE = ((gamma_l - 1.0) + D * (gamma_r - 1.0)) * D1;
    double _dpde = ((gamma_l - 1.0) * rl + D * ((gamma_r - 1.0) * rr)) * D1;
    /* SoundSpeedMultiSpecies */
    double dp = pr - pl, drho = rr - rl, de = e_r - e_l;
    double _dpdrhoi_new[ORA_MAXNS], Sum_dpdrhoi = 0.0, Sum_drhoi = 0.0, Sum_dpdrhoi2 = 0.0, Sum_Yidpdrhoi = 0.0;
    for (int n = 0; n < NCOP; n++) {
        Sum_dpdrhoi += _dpdrhoi[n] * drhoi[n];
        Sum_dpdrhoi2 += _dpdrhoi[n] * drhoi[n] * _dpdrhoi[n] * drhoi[n];
    }
    double temp1 = dp - (_dpdrho * drho + _dpde * de + Sum_dpdrhoi);
    double temp = temp1 / (_dpdrho * _dpdrho * drho * drho + _dpde * de * _dpde * de + Sum_dpdrhoi2 + 1e-19);
    for (int n = 0; n < NCOP; n++) {
        Sum_drhoi += drhoi[n] * drhoi[n];
        Sum_Yidpdrhoi += _yi[n] * _dpdrhoi[n];
    }
    double _dpdE_new = ORA_MUTANT == 6 ? _dpdE + _dpdE * _dpdE * de * temp : _dpdE + _dpdE * _dpdE * de * _rho * temp;
    double _dpdrho_new = _dpdrho + _dpdrho * _dpdrho * drho * temp;
    for (int n = 0; n < NCOP; n++)
        _dpdrhoi_new[n] = _dpdrhoi[n] + _dpdrhoi[n] * _dpdrhoi[n] * drhoi[n] * temp;
    double csqr = _dpdrho_new + _dpdE_new * _prho + Sum_Yidpdrhoi;
    *b1 = _dpdE_new / csqr;
    for (int n = 0; n < NCOP; n++) {
        z[n] = -_dpdrhoi_new[n] / _dpdE_new;
        *b3 += _yi[n] * z[n];
    }
    *b3 *= *b1;
    if (terms) {
        const double t[14] = {_dpdrho, _dpde, _dpdE, dp, drho, de, _rho, _prho, temp1, temp, _dpdrho_new, _dpdE_new,
                              csqr, (double)NCOP};
        for (int q = 0; q < 14; q++) terms[q] = t[q];
        for (int n = 0; n < NCOP; n++) {
            terms[14 + n] = _dpdrhoi[n];
            terms[14 + NCOP + n] = drhoi[n];
            terms[14 + 2 * NCOP + n] = _dpdrhoi_new[n];
            terms[14 + 3 * NCOP + n] = _yi[n];
        }
    }
    /* c2 fallback when c2 <= 0 (sycl::step(c2, 0)) */
    double c2 = csqr;
    double c2w = (c2 <= 0.0) ? 1.0 : 0.0;
    c2 = *Gamma * _P * _rho * c2w + (1.0 - c2w) * c2;
    return c2;
}

/* The multi-species Roe sound speed of face (id_l, id_r) for the primitives
   of the last UpdateStates (MARCO_ROE + reconstruct_c2), with b1 and b3 */
double ora_face_c2(ora_fluid *f, size_t id_l, size_t id_r, double *b1_out, double *b3_out)
{
    double D = sqrt(f->rho[id_r] / f->rho[id_l]);
    double D1 = 1.0 / (D + 1.0);
    double _P = (f->p[id_l] + D * f->p[id_r]) * D1;
    double _rho = sqrt(f->rho[id_r] * f->rho[id_l]);
    double _yi[ORA_MAXNS] = {0}, z[ORA_MAXNS] = {0}, b1 = 0.0, b3 = 0.0, Gamma0;
    double c2 = reconstruct_c2(f, id_l, id_r, D, D1, _rho, _P, _yi, z, &b1, &b3, &Gamma0, NULL);
    if (b1_out) *b1_out = b1;
    if (b3_out) *b3_out = b3;
    return c2;
}

/* The same face, with the quantities of ReconstructSoundSpeed and
   SoundSpeedMultiSpecies in terms[14 + 4 NCOP]: the Roe-averaged dp/drho,
   dp/de, dp/dE, the jumps dp, drho, de, rho_bar, (p/rho)_bar, temp1, temp,
   the corrected dp/drho and dp/dE, c^2 before the fallback, NCOP; then per
   species n < NCOP the averaged dp/drho_n, the jump drho_n, the corrected
   dp/drho_n and the averaged Y_n. */
double ora_face_c2_terms(ora_fluid *f, size_t id_l, size_t id_r, double *terms)
{
    double D = sqrt(f->rho[id_r] / f->rho[id_l]);
    double D1 = 1.0 / (D + 1.0);
    double _P = (f->p[id_l] + D * f->p[id_r]) * D1;
    double _rho = sqrt(f->rho[id_r] * f->rho[id_l]);
    double _yi[ORA_MAXNS] = {0}, z[ORA_MAXNS] = {0}, b1 = 0.0, b3 = 0.0, Gamma0;
    return reconstruct_c2(f, id_l, id_r, D, D1, _rho, _P, _yi, z, &b1, &b3, &Gamma0, terms);
}

/* One characteristic face flux, MARCO_FLUXWALL_WENO5 (EIGEN_ALLOC == 0):
   positive-definite_eigen/Eigen_callback.h:179-230, inside
   ReconstructFlux{X,Y,Z} (FDM_Method/Reconstruction_kernels.hpp:8-199). */
static void face_flux(ora_fluid *f, int dir, size_t id_l, size_t id_r, long stride, const double *UI,
                      const double *Fl, const double *eigen_local, const double *eigen_block, double *Fwall)
{
    const ora_config *c = &f->cfg;
    const int E = f->E;
    /* MARCO_ROE: src/include/global_marco.h:26-34 */
    double D = sqrt(f->rho[id_r] / f->rho[id_l]);
    double D1 = 1.0 / (D + 1.0);
    double _u = (f->u[id_l] + D * f->u[id_r]) * D1;
    double _v = (f->v[id_l] + D * f->v[id_r]) * D1;
    double _w = (f->w[id_l] + D * f->w[id_r]) * D1;
    double _H = (f->H[id_l] + D * f->H[id_r]) * D1;
    double _P = (f->p[id_l] + D * f->p[id_r]) * D1;
    double _rho = sqrt(f->rho[id_r] * f->rho[id_l]);
    double _yi[ORA_MAXNS] = {0}, z[ORA_MAXNS] = {0}, b1 = 0.0, b3 = 0.0, c2, Gamma0;
    if (c->cop) {
        /* MARCO_COPC2: Eigen_callback.h:7-11 */
        Gamma0 = 1.4;
        c2 = reconstruct_c2(f, id_l, id_r, D, D1, _rho, _P, _yi, z, &b1, &b3, &Gamma0, NULL);
    } else {
        /* MARCO_NOCOPC2: Eigen_callback.h:87-91 */
        _yi[0] = 1.0;
        Gamma0 = c->ncop_gamma;
        c2 = Gamma0 * _P / _rho;
        b1 = (Gamma0 - 1.0) / c2;
    }
    double uf[10], ff[10], pp[10], mm[10], f_flux[ORA_MAXE], _r[ORA_MAXE][ORA_MAXE], eig[ORA_MAXE],
        eigen_value, artificial_viscosity;
    const double Roe_type = 0.0, LLF_type = 0.0, GLF_type = 1.0; /* Artificial_type 3 (GLF) */
    for (int n = 0; n < E; n++) {
        double eigen_local_max = 0.0;
        ora_eigen_lr(c, dir, n, eig, NULL, &eigen_value, z, _yi, c2, _rho, _u, _v, _w, _H, b1, b3);
        for (int m = -2; m < 4; m++) {
            long id_local_1 = (long)id_l + m * stride;
            eigen_local_max = ora_max(eigen_local_max, fabs(eigen_local[E * id_local_1 + n]));
        }
        artificial_viscosity = Roe_type * eigen_value + LLF_type * eigen_local_max + GLF_type * eigen_block[n];
        for (int m = -3; m <= 4; m++) {
            long id_local = (long)id_l + m * stride;
            uf[m + 3] = 0.0;
            ff[m + 3] = 0.0;
            if (c->fused) {
                /* fused FP64 form: each multiply-add pair rounded once */
                for (int n1 = 0; n1 < E; n1++) {
                    uf[m + 3] = fma(UI[E * id_local + n1], eig[n1], uf[m + 3]);
                    ff[m + 3] = fma(Fl[E * id_local + n1], eig[n1], ff[m + 3]);
                }
                pp[m + 3] = 0.5 * fma(artificial_viscosity, uf[m + 3], ff[m + 3]);
                mm[m + 3] = 0.5 * fma(-artificial_viscosity, uf[m + 3], ff[m + 3]);
            } else {
                for (int n1 = 0; n1 < E; n1++) {
                    uf[m + 3] = uf[m + 3] + UI[E * id_local + n1] * eig[n1];
                    ff[m + 3] = ff[m + 3] + Fl[E * id_local + n1] * eig[n1];
                }
                pp[m + 3] = 0.5 * (ff[m + 3] + artificial_viscosity * uf[m + 3]);
                mm[m + 3] = 0.5 * (ff[m + 3] - artificial_viscosity * uf[m + 3]);
            }
        }
        if (c->scheme == ORA_WENOCU6)
            f_flux[n] = ora_wenocu6(pp, mm, dir == 0 ? c->dx : (dir == 1 ? c->dy : c->dz));
        else if (c->scheme == ORA_WENO5Z)
            f_flux[n] = c->fused ? ora_weno5z_fused(pp, mm) : ora_weno5z(pp, mm);
        else
            f_flux[n] = ora_weno5js(pp, mm);
        ora_eigen_lr(c, dir, n, NULL, _r[n], &eigen_value, z, _yi, c2, _rho, _u, _v, _w, _H, b1, b3);
    }
    /* F-flux reconstruction: _p[n1][n] = f_flux[n1] * R[n1][n], summed over
       n1 in increasing order (fused form: one fma per term) */
    for (int n = 0; n < E; n++) {
        double fluxl = 0.0;
        for (int n1 = 0; n1 < E; n1++) {
            if (c->fused)
                fluxl = fma(f_flux[n1], _r[n1][n], fluxl);
            else
                fluxl += f_flux[n1] * _r[n1][n];
        }
        Fwall[E * id_l + n] = fluxl;
    }
}

/* PositivityPreservingKernel: positive-definite_eigen/PositivityPreserving_kernels.hpp:5-76,
   face (id_l, id_r).  Blends the high-order face flux toward the
   Lax-Friedrichs flux until density and each of the Ns-1 mass fractions of
   the two adjacent cells stay above epsilon after a forward-Euler step of
   size lambda.  Kept as written, including the species loop's update of
   FF[n] (indexed by the species counter), whose values are never read
   again. */
static void positivity_face(ora_fluid *f, size_t id_l, size_t id_r, const double *UI, const double *Fl, double *Fwall,
                            double lambda_0, double lambda, const double *epsilon)
{
    const int E = f->E, NCOP = f->NCOP;
    double theta, theta_u, theta_p, rho_min, F_LF[ORA_MAXE], FF_LF[ORA_MAXE], FF[ORA_MAXE];
    const double *UU = UI + (size_t)E * id_l, *UP = UI + (size_t)E * id_r;
    double *Fw = Fwall + (size_t)E * id_l;
    for (int n = 0; n < E; n++) {
        F_LF[n] = 0.5 * (Fl[n + (size_t)E * id_l] + Fl[n + (size_t)E * id_r] + lambda_0 * (UU[n] - UP[n]));
        FF_LF[n] = 2.0 * lambda * F_LF[n];
        FF[n] = 2.0 * lambda * Fw[n];
    }
    theta_u = 1.0, theta_p = 1.0;
    rho_min = ora_min(UU[0], epsilon[0]);
    if (UU[0] - FF[0] < rho_min)
        theta_u = (UU[0] - FF_LF[0] - rho_min + 1.0e-40) / (FF[0] - FF_LF[0] + 1.0e-40);
    rho_min = ora_min(UP[0], epsilon[0]);
    if (UP[0] + FF[0] < rho_min)
        theta_p = (UP[0] + FF_LF[0] - rho_min + 1.0e-40) / (FF_LF[0] - FF[0] + 1.0e-40);
    theta = ora_min(ora_max(ora_min(theta_u, theta_p), 0.0), 1.0);
    for (int n = 0; n < E; n++) {
        FF[n] = (1.0 - theta) * FF_LF[n] + theta * FF[n];
        Fw[n] = (1.0 - theta) * F_LF[n] + theta * Fw[n];
    }
    double yi_q[ORA_MAXNS], yi_u[ORA_MAXNS], yi_qp[ORA_MAXNS], yi_up[ORA_MAXNS];
    const double _rhoq = 1.0 / (UU[0] - FF[0]), _rhou = 1.0 / (UU[0] - FF_LF[0]);
    const double _rhoqp = 1.0 / (UP[0] + FF[0]), _rhoup = 1.0 / (UP[0] + FF_LF[0]);
    yi_q[NCOP] = 1.0, yi_u[NCOP] = 1.0, yi_qp[NCOP] = 1.0, yi_up[NCOP] = 1.0;
    for (int n = 0; n < NCOP; n++) {
        const int tid = n + 5;
        yi_q[n] = (UU[tid] - FF[tid]) * _rhoq, yi_q[NCOP] -= yi_q[n];
        yi_u[n] = (UU[tid] - FF_LF[tid]) * _rhou, yi_u[NCOP] -= yi_u[n];
        yi_qp[n] = (UP[tid] + FF[tid]) * _rhoqp, yi_qp[NCOP] -= yi_qp[n];
        yi_up[n] = (UP[tid] + FF_LF[tid]) * _rhoup, yi_up[NCOP] -= yi_up[n];
        theta_u = 1.0, theta_p = 1.0;
        const double temp = epsilon[n + 2];
        if (yi_q[n] < temp) {
            const double yi_min = ora_min(yi_u[n], temp);
            theta_u = (yi_u[n] - yi_min + 1.0e-40) / (yi_u[n] - yi_q[n] + 1.0e-40);
        }
        if (yi_qp[n] < temp) {
            const double yi_min = ora_min(yi_up[n], temp);
            theta_p = (yi_up[n] - yi_min + 1.0e-40) / (yi_up[n] - yi_qp[n] + 1.0e-40);
        }
        theta = ora_min(ora_max(ora_min(theta_u, theta_p), 0.0), 1.0);
        for (int nn = 0; nn < E; nn++) {
            FF[n] = (1.0 - theta) * FF_LF[n] + theta * FF[n];
            Fw[nn] = (1.0 - theta) * F_LF[nn] + theta * Fw[nn];
        }
    }
}

/* GetLU: FDM_Method/ConVenction_block.hpp:10-617 (no viscous terms).  eigen_block keeps its running maximum across calls:
   it is zeroed once (src/Fluids.cpp:427) and each sycl::reduction without
   initialize_to_identity combines with the stored value. */
void ora_compute_lu(ora_fluid *f, int which)
{
    const ora_config *c = &f->cfg;
    const int E = f->E;
    const double *UI = which ? f->U1 : f->U;
    const int dims[3] = {c->DimX, c->DimY, c->DimZ};
    double *el[3] = {f->eigen_local_x, f->eigen_local_y, f->eigen_local_z};
    const double *vel[3] = {f->u, f->v, f->w};
    long n = (long)f->ncell;
    /* GetLocalEigen: positive-definite_eigen/Eigen_value.hpp:5-33 */
    for (int d = 0; d < 3; d++) {
        if (!dims[d]) continue;
#pragma omp parallel for schedule(static)
        for (long id = 0; id < n; id++) {
            double uu = vel[d][id];
            double uuPc = uu + f->c[id];
            double uuMc = uu - f->c[id];
            el[d][E * id + 0] = uuMc;
            for (int ii = 1; ii < E - 1; ii++) el[d][E * id + ii] = uu;
            el[d][E * id + E - 1] = uuPc;
        }
    }
    /* GLF maxima (running): ConVenction_block.hpp:115-172 */
    for (int d = 0; d < 3; d++) {
        if (!dims[d]) continue;
        for (int nn = 0; nn < E; nn++) {
            double m = f->eigen_block[d * ORA_MAXE + nn];
            for (long id = 0; id < n; id++) {
                double v = fabs(el[d][E * id + nn]);
                if (v > m) m = v;
            }
            f->eigen_block[d * ORA_MAXE + nn] = m;
        }
    }
    /* face fluxes */
    const double *Fl[3] = {f->FluxF, f->FluxG, f->FluxH};
    double *Fw[3] = {f->FluxFw, f->FluxGw, f->FluxHw};
    const long strides[3] = {1, c->Xmax, (long)c->Xmax * c->Ymax};
    for (int d = 0; d < 3; d++) {
        if (!dims[d]) continue;
        int i0 = c->Bwidth_X - (d == 0), j0 = c->Bwidth_Y - (d == 1), k0 = c->Bwidth_Z - (d == 2);
        int i1 = c->Bwidth_X + c->X_inner, j1 = c->Bwidth_Y + c->Y_inner, k1 = c->Bwidth_Z + c->Z_inner;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = k0; k < k1; k++)
            for (int j = j0; j < j1; j++)
                for (int i = i0; i < i1; i++) {
                    size_t id_l = IDX(f, i, j, k);
                    size_t id_r = id_l + strides[d];
                    face_flux(f, d, id_l, id_r, strides[d], UI, Fl[d], el[d], f->eigen_block + d * ORA_MAXE, Fw[d]);
                }
    }
    /* UpdateFluidLU: Reconstruction_kernels.hpp:201-234 */
    /* positivity-preserving limiter, x then y then z, on the faces to the
       right of each interior cell (ConVenction_block.hpp:328-413); lambda_0 is
       uvw_c_max of the last GetDt */
    if (c->positivity) {
        double epsilon[ORA_MAXNS + 2];
        epsilon[0] = 1.0e-13, epsilon[1] = 1.0e-13;
        for (int ii = 2; ii < c->NS + 2; ii++) epsilon[ii] = 0.0;
        for (int d = 0; d < 3; d++) {
            if (!dims[d]) continue;
            const double lambda_0 = f->uvw_c_max[d];
            const double lambda = c->CFLnumber / lambda_0;
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = c->Bwidth_Z; k < c->Bwidth_Z + c->Z_inner; k++)
                for (int j = c->Bwidth_Y; j < c->Bwidth_Y + c->Y_inner; j++)
                    for (int i = c->Bwidth_X; i < c->Bwidth_X + c->X_inner; i++) {
                        size_t id_l = IDX(f, i, j, k);
                        positivity_face(f, id_l, id_l + strides[d], UI, Fl[d], Fw[d], lambda_0, lambda, epsilon);
                    }
        }
    }
    /* viscous, heat-conduction and diffusion wall fluxes
       (ConVenction_block.hpp:424-575), after the limiter */
    if (c->visc) ora_viscous_fluxes(f);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = c->Bwidth_Z; k < c->Bwidth_Z + c->Z_inner; k++)
        for (int j = c->Bwidth_Y; j < c->Bwidth_Y + c->Y_inner; j++)
            for (int i = c->Bwidth_X; i < c->Bwidth_X + c->X_inner; i++) {
                size_t id = IDX(f, i, j, k);
                for (int nq = 0; nq < E; nq++) {
                    double LU0 = 0.0;
                    if (c->DimX) LU0 += (f->FluxFw[E * (id - 1) + nq] - f->FluxFw[E * id + nq]) * c->_dx;
                    if (c->DimY) LU0 += (f->FluxGw[E * (id - strides[1]) + nq] - f->FluxGw[E * id + nq]) * c->_dy;
                    if (c->DimZ) LU0 += (f->FluxHw[E * (id - strides[2]) + nq] - f->FluxHw[E * id + nq]) * c->_dz;
                    f->LU[E * id + nq] = LU0;
                }
            }
}

/* EstimateFluidNANKernel: src/Fluids.cpp:47-87 via Fluid::EstimateFluidNAN
   (src/Fluids.cpp:963-1040); flag 1,2 check U1, flag 3,4 check U. */
int ora_estimate_nan(ora_fluid *f, int flag)
{
    const ora_config *c = &f->cfg;
    const int E = f->E;
    const double *UI = (flag <= 2) ? f->U1 : f->U;
    int err = 0;
    for (int k = c->Bwidth_Z; k < c->Zmax - c->Bwidth_Z; k++)
        for (int j = c->Bwidth_Y; j < c->Ymax - c->Bwidth_Y; j++)
            for (int i = c->Bwidth_X; i < c->Xmax - c->Bwidth_X; i++) {
                size_t id = IDX(f, i, j, k) * E;
                int bad = UI[id] < 0;
                for (int ii = 0; ii < E; ii++)
                    bad = bad || isnan(UI[id + ii]) || isinf(UI[id + ii]) || isnan(f->LU[id + ii]) || isinf(f->LU[id + ii]);
                if (bad && !err) err = 3, f->err_code = 3, f->err_id = id / E;
            }
    return err;
}
