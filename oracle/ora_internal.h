/* Internal definitions of the CPU oracle (test infrastructure only). */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H

#include "xf_oracle.h"
#include <math.h>

struct ora_fluid {
    ora_config cfg;
    int E;          /* Emax: NS + 4 with COP (Eigen_global_definition.h:3-4), 5 without */
    int NCOP;       /* NUM_COP = NS - 1 */
    size_t ncell;   /* Xmax * Ymax * Zmax */
    /* conservative fields (AoS) */
    double *U, *U1, *LU;
    double *FluxF, *FluxG, *FluxH, *FluxFw, *FluxGw, *FluxHw;
    /* primitives (FlowData, global_setup.h:231-250) */
    double *rho, *p, *u, *v, *w, *e, *y, *gamma, *T, *H, *c, *Cp, *Ri;
    double *eigen_local_x, *eigen_local_y, *eigen_local_z;
    double eigen_block[3 * ORA_MAXE];
    double uvw_c_max[6];
    /* viscous terms (cfg.visc): velocity gradients Vde[9][n], Wilke viscosity
       and conductivity, mixture-averaged D_km (species-major), species
       enthalpies (AoS, NS * id + k) */
    double *Vde, *va, *tca, *Dkm, *hi;
    /* the last physical error: code, lowest failing cell id of that code,
       and the step (Iteration) and stage (1..3 RK, 4 Reaction) it came in */
    int err_code, err_step, err_where;
    size_t err_id;
    /* the clock of XFLUIDS::Evolution (src/XFLUIDS.cpp:170-198):
       physicalTime, the output time dt is clamped to, Iteration, and
       whether the time has reached the target */
    double time, target;
    int step, done;
};

void ora_visc_alloc(ora_fluid *f);
void ora_visc_free(ora_fluid *f);
void ora_viscous_fluxes(ora_fluid *f);

/* transcendental dispatch: libm or the deterministic restatement */
double ora_xexp(double x);
double ora_xlog(double x);
static inline double ora_expc(const ora_config *c, double x) { return c->use_libm ? exp(x) : ora_xexp(x); }
static inline double ora_logc(const ora_config *c, double x) { return c->use_libm ? log(x) : ora_xlog(x); }

static inline double ora_max(double a, double b) { return a > b ? a : b; }   /* sycl::max on non-NaN */
static inline double ora_min(double a, double b) { return a < b ? a : b; }

#endif
