/*
 * xf_oracle.h -- CPU oracle for the XFluids per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, array-of-structures
 * restatement of the reference's per-cell/per-face functions, written to check
 * the HIP implementation.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; no product code path links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the XFluids source tree).  Layout follows the reference: U[Emax*id + n],
 * y[NUM_SPECIES*id + k], id = Xmax*Ymax*k + Xmax*j + i, primitives one array
 * per variable.  Build with -ffp-contract=off (no compiler FMA contraction):
 * every product and sum rounds as the reference writes it.  The one
 * exception is the fused FP64 variant of the face flux (cfg.fused = 1), which
 * rounds each multiply-add pair of the characteristic projections, GLF
 * split, back-projection and WENO5-Z once (explicit fma() calls) to check
 * the device form with the same arithmetic.
 *
 * Parity status per function (what pins it) is listed in the header comment
 * of each source file; functions with no closed-form / invariant / library
 * pin are marked "parity unpinned" there.
 */
#ifndef XF_ORACLE_H
#define XF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_MAXNS 16
#define ORA_MAXNR 40
#define ORA_MAXE (ORA_MAXNS + 4)

/* Boundary condition codes: src/include/global_setup.h:96-110 */
enum { ORA_INFLOW = 0, ORA_OUTFLOW = 1, ORA_SYMMETRY = 2, ORA_PERIODIC = 3,
       ORA_NSLIPWALL = 4, ORA_VISCWALL = 5, ORA_SLIPWALL = 6, ORA_INNERBLOCK = 7 };

/* Scheme codes for the characteristic face flux */
enum { ORA_WENO5JS = 0, ORA_WENO5Z = 1, ORA_WENOCU6 = 2 };

typedef struct {
    /* geometry: Block / MeshSize (src/include/global_setup.h:162-200),
       derived exactly as Setup::init (src/read_ini/src/iniset.cpp:290-369) */
    int Xmax, Ymax, Zmax;
    int X_inner, Y_inner, Z_inner;
    int Bwidth_X, Bwidth_Y, Bwidth_Z;
    int DimX, DimY, DimZ;
    double dx, dy, dz, _dx, _dy, _dz, CFLnumber;
    int BC[6];
    /* compile-time switches of the reference made runtime */
    int NS;          /* NUM_SPECIES */
    int NR;          /* NUM_REA */
    int cop;         /* COP defined (multi-component) */
    int thermo;      /* Thermo: 1 = NASA-9, 0 = JANAF/NASA-7 */
    int scheme;      /* ORA_WENO5JS (weno5old_GPU), ORA_WENO5Z or ORA_WENOCU6 */
    int react;       /* ReactSources */
    int strang;      /* SlipOrder == "Strang" */
    double ncop_gamma; /* NCOP_Gamma */
    /* Thermal tables (src/include/global_setup.h:202-208; filled as
       src/read_ini/src/thermal.cpp:36-175) */
    double Hia[ORA_MAXNS * 21];
    double Hib[ORA_MAXNS * 6];
    double Wi[ORA_MAXNS], _Wi[ORA_MAXNS], Ri[ORA_MAXNS];
    /* Reaction tables (src/include/global_setup.h:210-215; filled as
       src/read_ini/src/reaction.cpp:6-172) */
    int Nu_f[ORA_MAXNR * ORA_MAXNS], Nu_b[ORA_MAXNR * ORA_MAXNS], Nu_d[ORA_MAXNR * ORA_MAXNS];
    int third_ind[ORA_MAXNR];
    double React_ThirdCoef[ORA_MAXNR * ORA_MAXNS];
    double Rargus[ORA_MAXNR * 6];
    /* 1: transcendental functions from libm (the reference CPU build's
       behaviour); 0: the deterministic exp/log shared by the HIP kernels */
    int use_libm;
    /* PositivityPreserving (settings equations block): the flux limiter of
       PositivityPreserving_kernels.hpp:5-76 on the face fluxes */
    int positivity;
    /* Visc, Visc_Heat, Visc_Diffu (CMakeLists.txt:55-57, compile-time in the
       reference): viscous stresses, heat conduction and species diffusion */
    int visc, visc_heat, visc_diffu;
    /* BlSz.Yil_limiter and BlSz.Dim_limiter as Setup::init forms them
       (src/read_ini/src/iniset.cpp:358-359) */
    double Yil_limiter, Dim_limiter;
    /* transport fits of Setup::GetFitCoefficient (src/read_ini/src/viscfit.cpp:148-193):
       ln(mu_k), ln(lambda_k) and ln(p D_kj) as cubics in ln T */
    double fcv[ORA_MAXNS * 4], fct[ORA_MAXNS * 4], Dkj[ORA_MAXNS * ORA_MAXNS * 4];
    /* 0: the face flux in the reference's expression order (default);
       1: its multiply-add pairs fused, the arithmetic of the fused FP64
       device form */
    int fused;
    /* GhostSpecies (a mixture's case_setup.h, e.g. runtime.dat/Inert-SBI):
       the last species is a ghost with Y = 0 and Getrhoyi renormalises the
       others, writing them back into the state */
    int ghost;
} ora_config;

typedef struct ora_fluid ora_fluid;

/* lifecycle */
ora_fluid *ora_create(const ora_config *cfg);
void ora_destroy(ora_fluid *f);
int ora_num_eqn(const ora_fluid *f);
size_t ora_num_cells(const ora_fluid *f);
/* named field pointers: "U","U1","LU","rho","p","u","v","w","e","y","gamma",
   "T","H","c","Cp","Ri","FluxF","FluxG","FluxH","FluxFw","FluxGw","FluxHw",
   "eigen_local_x","eigen_local_y","eigen_local_z","eigen_block","uvw_c_max" */
double *ora_field(ora_fluid *f, const char *name);

/* stage seam (mirrors the reference's stage functions) */
void ora_boundary(ora_fluid *f, int which);                 /* FluidBoundaryCondition */
int ora_update_states(ora_fluid *f, int which);             /* UpdateFluidStateFlux   */
void ora_compute_lu(ora_fluid *f, int which);               /* GetLU                  */
/* Roe-averaged multi-species c^2 of face (id_l, id_r) (ReconstructSoundSpeed) */
double ora_face_c2(ora_fluid *f, size_t id_l, size_t id_r, double *b1, double *b3);
void ora_rk3(ora_fluid *f, int stage, double dt);           /* UpdateURK3rd           */
double ora_get_dt(ora_fluid *f);                            /* GetDt                  */
void ora_chemistry(ora_fluid *f, double dt);                /* ChemeODEQ2Solver       */
int ora_reaction(ora_fluid *f, double dt);                  /* XFLUIDS::Reaction      */
int ora_estimate_nan(ora_fluid *f, int flag);               /* Fluid::EstimateFluidNAN */
/* whole steps in the reference's Evolution ordering; returns 0 or error */
int ora_initialize(ora_fluid *f);                           /* main: BC + UpdateStates */
int ora_steps(ora_fluid *f, int nsteps, double *dt_out);
/* the clock of Evolution: physicalTime, Iteration, the output time each dt
   is clamped to (default +inf), and whether the time has reached it */
void ora_set_time(ora_fluid *f, double time, int step);
void ora_set_target_time(ora_fluid *f, double target);
void ora_get_time(const ora_fluid *f, double *time, int *step, int *reached);
/* the last physical error code (0 none), its lowest failing cell and the
   step and stage (1..3 RK, 4 Reaction) of ora_steps it came in */
int ora_last_error(const ora_fluid *f, int *i, int *j, int *k, int *step, int *where);

/* per-function entry points used by the pin tests */
double ora_exp(double x);
double ora_log(double x);
double ora_cp_i(const ora_config *c, double T, int n);      /* HeatCapacity    */
double ora_h_i(const ora_config *c, double T, int n);       /* get_Enthalpy    */
double ora_s_i(const ora_config *c, double T, int n);       /* get_Entropy     */
double ora_g_i(const ora_config *c, double T, int n);       /* get_Gibson      */
double ora_e_i(const ora_config *c, double T, int n);       /* get_Internale   */
double ora_mix_cp(const ora_config *c, const double *yi, double T);
double ora_mix_cv(const ora_config *c, const double *yi, double T);
double ora_mix_h(const ora_config *c, const double *yi, double T);
double ora_mix_e(const ora_config *c, const double *yi, double T);
double ora_mix_R(const ora_config *c, const double *yi);
double ora_get_T(const ora_config *c, const double *yi, double e, double T0);
double ora_weno5js(const double *pp, const double *mm);     /* weno5old_GPU(&pp[3],&mm[3]) */
double ora_weno5z(const double *pp, const double *mm);      /* weno5Z_P + weno5Z_M        */
double ora_weno5z_fused(const double *pp, const double *mm); /* the same, multiply-adds fused */
double ora_wenocu6(const double *pp, const double *mm, double delta); /* WENOCU6_GPU(&pp[3],&mm[3],dl) */
/* left row n / right column n of the Roe eigen-system in direction dir
   (0,1,2); z, yi of length NS-1 */
void ora_eigen_lr(const ora_config *c, int dir, int n, double *eig_l, double *eig_r,
                  double *eigen_value, const double *z, const double *yi, double c2,
                  double rho, double u, double v, double w, double H, double b1,
                  double b3);
/* the face sound speed with its intermediate quantities (pin tests) */
double ora_face_c2_terms(ora_fluid *f, size_t id_l, size_t id_r, double *terms);
/* CHEMEQ2's alpha(r) and its integrator on a model system (pin tests):
   constant production q, loss k y^power; trace (NULL ok): per pass the
   substep, error measure * epscl, threshold and end time, cap passes */
double ora_alpha(double r);
int ora_chemeq2_model(int NL, double *y, double dtg, const double *q, const double *k, int power, double *dt_first,
                      double *trace, int cap);
/* production/loss rates and the CHEMEQ2 integrator on one cell */
void ora_evalcv(const ora_config *c, const double *y, double *q, double *d,
                double rho, double p);
int ora_chemeq2(const ora_config *c, double *y, double dtg, double rho, double p);

#ifdef __cplusplus
}
#endif
#endif
