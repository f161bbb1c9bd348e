/*
 * Finite-rate chemistry of the oracle (test infrastructure only):
 * constant-volume production/loss rates and the CHEMEQ2 quasi-steady-state
 * integrator with temperature carried in the ODE vector.
 *
 * Restates src/solver_Reaction/ProductionRates.h:117-176 (evalcore, q/p form),
 * :227-244 (xfrsign, Alpha), :459-479 (evalcv, y[0] = T form), :488-638
 * (Chemeq2), and src/solver_Reaction/Reaction_kernels.hpp:46-73
 * (ChemeODEQ2SolverKernel1), launched over interior cells as
 * src/solver_Reaction/Reaction_block.hpp:4-38.
 *
 * Deviation for safety only: the substep loop is bounded (ORA_CHEM_MAXSUB);
 * a cell that reaches the bound is reported, never silently accepted.  The
 * reference loop has no bound.
 *
 * Pins (tests/test_oracle_chem.py): Arrhenius closed form; Kc against the
 * Gibbs-energy closed form; mass (sum Y = 1) and element (H, O) conservation;
 * constant-volume energy conservation e(T, Y) = const; 0-D H2/O2/Ar ignition
 * against scipy solve_ivp(Radau) of an independent numpy restatement of the
 * RHS (committed fixture).  The integrator body and its step control
 * (tests/test_oracle_pins.py), run on model systems with closed-form
 * solutions through ora_chemeq2_model: alpha(r) against the exact
 * quasi-steady weight (fifth order), the first substep sqreps / k, exact
 * substep products on linear decay, the relaxation step against the exact
 * exponential, the Riccati error band of epscl, and the step-size law on
 * every pass of stiff Riccati runs (accept / reject, 1/sqrt(eps) + 0.005 by
 * two Heron iterations, the 0.34 cap); each constant changed on its own
 * (ORA_MUTANT) fails one of them.
 */
#include "ora_internal.h"

#define ORA_CHEM_MAXSUB 200000

static const double ora_refP = 1.0 / 101325.0;                         /* _refP */
static const double ora_R0c = 6.02214076e26 * 1.380649e-23;           /* universal_gas_const */
static const double ora_Ruc = 6.02214076e26 * 1.380649e-23 * 1.0E-3;   /* Ru */

/* sycl::pown for the non-negative stoichiometric exponents of the tables */
static double ora_pown(double x, int n)
{
    if (n == 0) return 1.0;
    double r = x;
    for (int k = 1; k < n; k++) r *= x;
    return r;
}

/* evalcore (q/p form): ProductionRates.h:117-176 */
static void evalcore(const ora_config *c, const double *yi, double *q, double *p, double m_dens,
                     double m_p, double m_tmp)
{
    const int NS = c->NS, NR = c->NR;
    double Kf[ORA_MAXNR], _Kck[ORA_MAXNR];
    double logStandConc = ora_logc(c, m_p / (ora_R0c * m_tmp));
    for (int m = 0; m < NR; m++) {
        double DeltaGibbs = 0.0;
        double A = c->Rargus[m * 6 + 0], B = c->Rargus[m * 6 + 1], E = c->Rargus[m * 6 + 2];
        Kf[m] = A * ora_expc(c, B * ora_logc(c, m_tmp) - E * 4.184 / ora_Ruc / m_tmp);
        const int *Nu_dm_ = c->Nu_d + m * NS;
        int m_dn = 0;
        for (int n = 0; n < NS; n++) {
            DeltaGibbs += Nu_dm_[n] * (ora_logc(c, m_p * ora_refP) - ora_g_i(c, m_tmp, n));
            m_dn += Nu_dm_[n];
        }
        _Kck[m] = ora_min(ora_expc(c, DeltaGibbs - m_dn * logStandConc), 1.0E+40);
    }
    double m_cm[ORA_MAXNS];
    for (int n = 0; n < NS; n++)
        m_cm[n] = yi[n] * c->_Wi[n] * m_dens * 1.0E-3;
    for (int r = 0; r < NR; r++) {
        double tb = 0.0;
        if (1 == c->third_ind[r]) {
            for (int it = 0; it < NS; it++)
                tb += c->React_ThirdCoef[r * NS + it] * m_cm[it];
        } else
            tb = 1.0;
        Kf[r] *= tb;
        const int *nu_f = c->Nu_f + r * NS;
        _Kck[r] *= Kf[r];
        const int *nu_b = c->Nu_b + r * NS;
        for (int it = 0; it < NS; it++)
            Kf[r] *= ora_pown(m_cm[it], nu_f[it]);
        for (int it = 0; it < NS; it++)
            _Kck[r] *= ora_pown(m_cm[it], nu_b[it]);
    }
    for (int n = 0; n < NS; n++) {
        q[n] = 0.0, p[n] = 0.0;
        for (int r = 0; r < NR; r++) {
            const int *nu_f = c->Nu_f + r * NS;
            const int *nu_b = c->Nu_b + r * NS;
            q[n] += nu_b[n] * Kf[r] + nu_f[n] * _Kck[r];
            p[n] += nu_f[n] * Kf[r] + nu_b[n] * _Kck[r];
        }
    }
}

/* evalcv (y[0] = T form): ProductionRates.h:459-479 */
void ora_evalcv(const ora_config *c, const double *y, double *q, double *p, double m_dens, double m_p)
{
    q[0] = 0.0, p[0] = 0.0;
    const double *yi = y + 1;
    double *qi = q + 1, *pi = p + 1;
    double m_tmp = y[0], _dens = 1.0 / m_dens;
    evalcore(c, yi, qi, pi, m_dens, m_p, m_tmp);
    for (int n = 0; n < c->NS; n++) {
        qi[n] *= c->Wi[n] * _dens;
        pi[n] *= c->Wi[n] * _dens;
        double ei = ora_e_i(c, m_tmp, n);
        q[0] -= qi[n] * ei, p[0] -= pi[n] * ei;
    }
    double _Cv = 1.0 / ora_mix_cv(c, yi, m_tmp);
    q[0] *= _Cv, p[0] *= _Cv;
}

/* xfrsign(a, b): ProductionRates.h:227-241 */
static double xfrsign2(double a, double b) { return (b >= 0.0 ? 1.0 : -1.0) * fabs(a); }

/* A planted mistake for the pin tests' mutation checks: each ORA_MUTANT
   value > 0 changes one constant (the default build has none). */
#ifndef ORA_MUTANT
#define ORA_MUTANT 0
#endif

/* Alpha(rtaui): ProductionRates.h:244 */
static double alpha_of(double r)
{
    const double a2 = ORA_MUTANT == 1 ? 12.0 : 11.0;
    return (180.0 + r * (60.0 + r * (a2 + r))) / (360.0 + r * (60.0 + r * (12.0 + r)));
}
double ora_alpha(double r) { return alpha_of(r); }

/* The right-hand side of the CHEMEQ2 system: production q and loss d of
   every component at y */
typedef void (*chem_rhs)(const void *ctx, const double *y, double *q, double *d);

/* The body of Chemeq2 (ProductionRates.h:488-638) after the mass-fraction
   normalisation: the initial step, predictor, corrector, error measure and
   step control over dtg, for NL components with right-hand side rhs.
   Returns the predictor passes (accepted + rejected substeps), or -1 at
   the substep bound; dt_first (may be NULL) receives the first substep, and
   trace (may be NULL) for each of the first cap passes its substep, its
   error measure times epscl, the threshold it is compared with and the
   time the substep ends at. */
static int chemeq2_core(int NL, double *y, double dtg, chem_rhs rhs, const void *ctx, double *dt_first,
                        double *trace, int cap)
{
    double ymin = 1.0e-20, dtmin = 1.0e-7;
    double eps, epsmin = 1.0e-4, scrtch = 1e-25;
    const double sqreps = ORA_MUTANT == 2 ? 0.06 : 0.05;
    const double epscl = ORA_MUTANT == 3 ? 1.0e3 : 1.0e4;
    const double shrink = ORA_MUTANT == 4 ? 0.5 : 0.34;
    const double grow = ORA_MUTANT == 5 ? 0.05 : 0.005;
    double tfd = 1.0 + 1.0e-10, epsmax = 1.0;
    double ys[ORA_MAXNS + 1], y1[ORA_MAXNS + 1];
    double rtau[ORA_MAXNS + 1], scrarray[ORA_MAXNS + 1];
    double qs[ORA_MAXNS + 1], ds[ORA_MAXNS + 1], q[ORA_MAXNS + 1], d[ORA_MAXNS + 1];
    double dt = 0.0, ts, tn = 0.0;
    int passes = 0;

    rhs(ctx, y, q, d);
    for (int i = 0; i < NL; i++) {
        const double ascr = fabs(q[i]);
        const double scr2 = xfrsign2(1.0 / y[i], 0.1 * epsmin * ascr - d[i]);
        const double scr1 = scr2 * d[i];
        scrtch = ora_max(-fabs(ascr - d[i]) * scr2, ora_max(scr1, scrtch));
    }
    dt = ora_min(sqreps / scrtch, dtg);
    dtmin *= dt;
    if (dt_first) *dt_first = dt;

    for (;;) {
        ts = tn;
        for (int i = 0; i < NL; i++) {
            ys[i] = y[i];
            qs[i] = q[i], ds[i] = d[i];
        }
    apredictor:
        if (++passes > ORA_CHEM_MAXSUB) return -1;
        for (int i = 0; i < NL; i++) {
            rtau[i] = dt * ds[i] / ys[i];
            double alpha = alpha_of(rtau[i]);
            scrarray[i] = dt * (qs[i] - ds[i]) / (1.0 + alpha * rtau[i]);
        }
        for (int i = 0; i < NL; i++)
            y[i] = ora_max(ys[i] + scrarray[i], ymin);
        tn = ts + dt;
        for (int i = 0; i < NL; i++)
            y1[i] = y[i];
        rhs(ctx, y, q, d);
        for (int i = 0; i < NL; i++) {
            const double rtaub = 0.5 * (rtau[i] + dt * d[i] / y[i]);
            const double alpha = alpha_of(rtaub);
            const double qt = (1.0 - alpha) * qs[i] + alpha * q[i];
            scrarray[i] = (qt * dt - rtaub * ys[i]) / (1.0 + alpha * rtaub);
        }
        eps = 1e-10;
        for (int i = 0; i < NL; i++) {
            const double scr2 = ora_max(ys[i] + scrarray[i], 0.0);
            double scr1 = fabs(scr2 - y1[i]);
            y[i] = ora_max(scr2, ymin);
            if ((0.25 * (ys[i] + y[i])) > ymin) {
                scr1 = scr1 / y[i];
                eps = ora_max(0.5 * (scr1 + ora_min(fabs(q[i] - d[i]) / (q[i] + d[i] + 1.0e-30), scr1)), eps);
            }
        }
        eps = eps * epscl;
        if (trace && passes <= cap) {
            double *t = trace + 4 * (passes - 1);
            t[0] = dt, t[1] = eps, t[2] = epsmax, t[3] = tn;
        }
        if (eps < epsmax) {
            if (dtg <= (tn * tfd))
                return passes;
        } else {
            tn = ts;
        }
        double dto = dt;
        double rteps = 0.5 * (eps + 1.0);
        rteps = 0.5 * (rteps + eps / rteps);
        rteps = 0.5 * (rteps + eps / rteps);
        dt = ora_min(dt * (1.0 / rteps + grow), tfd * (dtg - tn));
        if (eps > epsmax) {
            dt = ora_min(dt, shrink * dto);
            if (dt <= dtmin)
                epsmax *= 10.0;
            goto apredictor;
        }
        epsmax = 1.0;
    }
}

struct mech_ctx {
    const ora_config *c;
    double rho, p;
};
static void mech_rhs(const void *v, const double *y, double *q, double *d)
{
    const struct mech_ctx *m = (const struct mech_ctx *)v;
    ora_evalcv(m->c, y, q, d, m->rho, m->p);
}

/* Chemeq2<NS,NR> (y[0] = T form): ProductionRates.h:488-638: the mass
   fractions floored at ymin and normalised, then the integration with the
   mechanism's rates (evalcv).  Returns the predictor passes (accepted +
   rejected substeps), or -1 when the substep bound is reached. */
int ora_chemeq2(const ora_config *c, double *y, double dtg, double rho, double m_p)
{
    const int NL = c->NS + 1;
    const double ymin = 1.0e-20;
    double sumy = 0.0;
    for (int i = 1; i < NL; i++)
        y[i] = ora_max(y[i], ymin), sumy += y[i];
    sumy = 1.0 / sumy;
    for (int i = 1; i < NL; i++)
        y[i] *= sumy;
    struct mech_ctx m = {c, rho, m_p};
    return chemeq2_core(NL, y, dtg, mech_rhs, &m, NULL, NULL, 0);
}

/* A model system for the pin tests: constant production q_i and loss
   d_i = k_i * y_i^power (power 1: dy/dt = q - k y, exponential relaxation;
   power 2 with q = 0: dy/dt = -k y^2), integrated by the same CHEMEQ2
   body as the mechanism. */
struct model_ctx {
    const double *q, *k;
    int NL, power;
};
static void model_rhs(const void *v, const double *y, double *q, double *d)
{
    const struct model_ctx *m = (const struct model_ctx *)v;
    for (int i = 0; i < m->NL; i++) {
        q[i] = m->q[i];
        d[i] = m->power == 2 ? m->k[i] * y[i] * y[i] : m->k[i] * y[i];
    }
}
int ora_chemeq2_model(int NL, double *y, double dtg, const double *q, const double *k, int power, double *dt_first,
                      double *trace, int cap)
{
    if (NL < 1 || NL > ORA_MAXNS + 1) return -2;
    struct model_ctx m = {q, k, NL, power};
    return chemeq2_core(NL, y, dtg, model_rhs, &m, dt_first, trace, cap);
}

/* ChemeODEQ2SolverKernel1 over interior cells: Reaction_kernels.hpp:46-73,
   Reaction_block.hpp:17-34.  Uses rho, y, T, p of the last UpdateStates. */
static int chem_status;
void ora_chemistry(ora_fluid *f, double dt)
{
    const ora_config *c = &f->cfg;
    const int E = f->E, NS = c->NS;
    int bad = 0;
    long bad_id = -1;
#pragma omp parallel for collapse(2) schedule(dynamic, 1) reduction(| : bad) reduction(max : bad_id)
    for (int k = c->Bwidth_Z; k < c->Zmax - c->Bwidth_Z; k++)
        for (int j = c->Bwidth_Y; j < c->Ymax - c->Bwidth_Y; j++)
            for (int i = c->Bwidth_X; i < c->Xmax - c->Bwidth_X; i++) {
                size_t id = (size_t)c->Xmax * c->Ymax * k + (size_t)c->Xmax * j + i;
                double _y[ORA_MAXNS + 1] = {0};
                _y[0] = f->T[id];
                double *yi = f->y + NS * id, *_yi = _y + 1, _rho = f->rho[id];
                for (int n = 0; n < NS; n++)
                    _yi[n] = yi[n];
                double e0 = ora_mix_e(c, _yi, _y[0]);
                if (ora_chemeq2(c, _y, dt, _rho, f->p[id]) < 0) {
                    bad = 1;
                    /* the lowest failing id, kept as the largest -id */
                    if (-(long)id - 1 > bad_id) bad_id = -(long)id - 1;
                }
                double *_U = f->U + E * id + 5;
                double e1 = ora_mix_e(c, _yi, _y[0]);
                for (int n = 0; n < NS - 1; n++)
                    _U[n] = _yi[n] * _rho;
                f->T[id] = _y[0], *(--_U) += (e1 - e0) * _rho;
            }
    chem_status = bad;
    if (bad) f->err_code = 5, f->err_id = (size_t)(-bad_id - 1);
}

int ora_chemistry_status(void) { return chem_status; }
