/*
 * xf_api.h -- C-ABI of the MI355X (gfx950) implementation of the XFluids
 * per-step hot path.
 *
 * Plain C types only: pointers, sizes, POD structs.  Every call returns an
 * int status: 0 = ok, > 0 = physical error code (XF_ERR_*; the offending cell
 * is available from xf_last_error), < 0 = HIP/runtime error (XF_EHIP...).
 * No exception crosses the ABI.  One context per host thread; all calls
 * enqueue on the context's HIP stream and return after the stream drained,
 * except the *_async variants.
 *
 * Layout contract: host arrays exchanged through xf_upload_state /
 * xf_download_* use the reference's array-of-structures order
 * U[Emax*id + n], y[NS*id + k], id = Xmax*Ymax*k + Xmax*j + i
 * (src/solver_UpdateStates/Update_kernels.hpp:14-15).  On the device the
 * fields are structure-of-arrays.
 */
#ifndef XF_API_H
#define XF_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XF_MAXNS 16
#define XF_MAXNR 40

/* status codes */
#define XF_OK 0
#define XF_ERR_RHO_Y 1     /* EstimateYiKernel (Estimate_kernels.hpp:5-102)            */
#define XF_ERR_PRIM 2      /* EstimatePrimitiveVarKernel (Estimate_kernels.hpp:104-162) */
#define XF_ERR_NAN 3       /* EstimateFluidNANKernel (src/Fluids.cpp:47-87)             */
#define XF_ERR_CHEM 5      /* chemistry substep bound reached                          */
#define XF_EINVAL (-1)     /* invalid argument / unsupported configuration             */
#define XF_EHIP (-2)       /* HIP runtime error                                        */
#define XF_ENOKERNEL (-3)  /* no kernel instance for this (NS, scheme, precision)      */

/* boundary codes: BConditions, src/include/global_setup.h:96-110 */
#define XF_BC_INFLOW 0
#define XF_BC_OUTFLOW 1
#define XF_BC_SYMMETRY 2
#define XF_BC_PERIODIC 3
#define XF_BC_NSLIPWALL 4
#define XF_BC_VISCWALL 5
#define XF_BC_SLIPWALL 6

/* face-flux schemes */
#define XF_WENO5JS 0 /* weno5old_GPU, schemes/WENO5s_schemes.hpp:71-95 */
#define XF_WENO5Z 1  /* weno5Z_P + weno5Z_M, schemes/WENO5s_schemes.hpp:246-316 */
#define XF_WENOCU6 2 /* WENOCU6_GPU, schemes/WENO6s_schemes.hpp:5-80 (SCHEME_ORDER 6) */

/* precision forms */
#define XF_PLAIN_FP64 0 /* everything FP64 in the reference's expression order, each
                           product and sum rounded as written (bitwise vs the CPU
                           restatement) */
#define XF_MIXED 1      /* FP32 smoothness indicators, weights and projections */
#define XF_FUSED_FP64 2 /* FP64 with the multiply-add pairs of the characteristic
                           projections, GLF split, back-projection and WENO5-Z
                           rounded once (fma); not the reference's rounding */

/* chemistry drivers */
#define XF_CHEM_PER_CELL 0 /* one lane per cell */
#define XF_CHEM_QUEUE 1    /* persistent lanes draining a device work queue */
#define XF_CHEM_WAVE 2     /* one wave per cell, each pass spread over its lanes (few
                              cells, e.g. a 1-D tube: the stiffest cell's chain of passes
                              sets the launch time) */

/* Case descriptor: the runtime form of the reference's Block/MeshSize
   (src/include/global_setup.h:162-200, derived as Setup::init,
   src/read_ini/src/iniset.cpp:290-369) plus the compile-time switches
   (NUM_SPECIES, NUM_REA, COP, Thermo, SCHEME_ORDER, CHEME_SPLITTING). */
typedef struct xf_case_desc {
    int Xmax, Ymax, Zmax;
    int X_inner, Y_inner, Z_inner;
    int Bwidth_X, Bwidth_Y, Bwidth_Z;
    int DimX, DimY, DimZ;
    double dx, dy, dz, _dx, _dy, _dz;
    double CFLnumber;
    int BC[6];          /* XMIN, XMAX, YMIN, YMAX, ZMIN, ZMAX */
    int NS;             /* species (1 = single-species ideal gas, NO-COP) */
    int NR;             /* reactions */
    int cop;            /* multi-component (COP) */
    int thermo;         /* 1 = NASA-9 (Thermo=1), 0 = JANAF/NASA-7 (Thermo=0) */
    int scheme;         /* XF_WENO5JS, XF_WENO5Z or XF_WENOCU6 */
    int react;          /* Sources_React */
    int strang;         /* 1 = Strang splitting, 0 = Lie */
    double ncop_gamma;  /* NCOP_Gamma */
    int precision;      /* XF_PLAIN_FP64, XF_MIXED or XF_FUSED_FP64 */
    int chem_driver;    /* XF_CHEM_PER_CELL, XF_CHEM_QUEUE or XF_CHEM_WAVE */
    int positivity;     /* PositivityPreserving flux limiter on the face fluxes
                           (PositivityPreserving_kernels.hpp:5-76) */
    int visc;           /* Visc: viscous stresses in the wall fluxes (CMakeLists.txt:55) */
    int visc_heat;      /* Visc_Heat: Fourier heat conduction (CMakeLists.txt:56) */
    int visc_diffu;     /* Visc_Diffu: species diffusion (CMakeLists.txt:57) */
    double Yil_limiter; /* BlSz.Yil_limiter and BlSz.Dim_limiter */
    double Dim_limiter; /*   (src/read_ini/src/iniset.cpp:358-359) */
    int ghost;          /* GhostSpecies: the last species is a ghost with Y = 0; UpdateStates
                           renormalises the others and writes them back into the state
                           (Getrhoyi, src/solver_UpdateStates/Update_device.hpp:15-21) */
} xf_case_desc;

typedef struct xf_ctx xf_ctx;

/* ---- context ----------------------------------------------------------- */
/* Allocates all device fields (U, U1, LU, primitives, scratch) for one case.
   Replaces Fluid::AllocateFluidMemory (src/Fluids.cpp:270-583). */
int xf_create(const xf_case_desc *desc, xf_ctx **out);
void xf_destroy(xf_ctx *ctx);
int xf_num_eqn(const xf_ctx *ctx);
/* The descriptor the context was created with. */
int xf_get_desc(const xf_ctx *ctx, xf_case_desc *out);
size_t xf_num_cells(const xf_ctx *ctx);
const char *xf_error_string(int status);

/* Thermal tables in the reference's layout: Hia[NS*7*3], Hib[NS*2*3],
   Wi[NS] in kg/mol.  Ri = Ru/Wi and _Wi = 1/Wi are derived as
   src/read_ini/src/thermal.cpp:158-161.  Replaces Setup::ReadThermal +
   CpyToGPU (src/read_ini/src/gpucopy.cpp:4-36). */
int xf_set_thermo(xf_ctx *ctx, const double *Hia, const double *Hib, const double *Wi);

/* Mechanism tables in the reference's layout: Nu_f[NR*NS], Nu_b[NR*NS],
   React_ThirdCoef[NR*NS], Rargus[NR*6]; Nu_d and third_ind derived as
   src/read_ini/src/reaction.cpp:31-40,142-172.  Replaces
   Setup::ReadReactions (src/read_ini/src/reaction.cpp:6-138). */
int xf_set_mechanism(xf_ctx *ctx, const int *Nu_f, const int *Nu_b, const double *React_ThirdCoef,
                     const double *Rargus);

/* Compiled mechanisms.  The library is built with the records of the
   mechanisms shipped under xfluids_amd/data as compile-time constants; when
   the tables given to xf_set_mechanism equal one of them field by field, the
   chemistry kernels run a rate evaluation specialised to it (same arithmetic,
   bit-identical results, no per-reaction table reads or mask tests).  No
   counterpart in the reference, whose kernels read the tables.
   xf_mech_specialization_index: the compiled mechanism matching these tables
   (same arguments as xf_set_mechanism plus NS, NR), or -1; host only.
   xf_set_chem_specialization: on = 0 keeps the generic table loop for this
   context (the default is on); drops a captured step graph.
   xf_chem_specialization: the specialisation in use, or -1. */
int xf_mech_specialization_index(int NS, int NR, const int *Nu_f, const int *Nu_b, const double *React_ThirdCoef,
                                 const double *Rargus);
int xf_set_chem_specialization(xf_ctx *ctx, int on);
int xf_chem_specialization(const xf_ctx *ctx);
/* Queue chemistry (XF_CHEM_QUEUE): a cell still running after `passes`
   CHEMEQ2 passes is handed from its lane to a whole wave, which spreads each
   rate evaluation over its 64 lanes (species and reactions in parallel; the
   same operations on every value, bit-identical results).  A few very stiff
   cells then no longer keep single lanes busy long after the rest.  At most
   65 536 cells per launch are handed over (the rest stay on their lanes);
   passes = 0 turns it off.  Ensembles use their first case's setting.  No
   counterpart in the reference (one work-item per cell).
   Currently switched off: its first GPU run did not finish and the cause has
   not been found, so the default is 0 and any passes > 0 returns XF_EINVAL. */
int xf_set_chem_handover(xf_ctx *ctx, int passes);

/* Transport fits in the reference's layout, each a cubic in ln T:
   fitted_coefficients_visc[NS][4] (ln mu_k, Pa s), fitted_coefficients_therm[NS][4]
   (ln lambda_k, W/(m K)) and Dkj_matrix[NS*NS][4] (ln(p D_kj), D in cm^2/s,
   p in Pa).  Replaces the device copies of Setup::GetFitCoefficient
   (src/read_ini/src/viscfit.cpp:148-193).  Needs xf_set_thermo first (Wi). */
int xf_set_transport(xf_ctx *ctx, const double *fcv, const double *fct, const double *Dkj);

/* ---- state exchange (AoS host <-> SoA device) ---------------------------- */
/* which: 0 = U, 1 = U1, 2 = U2 (the stage-2 result the fused step keeps apart
   from U1).  T is the per-cell temperature (Newton initial guess).
   Uploading U (which = 0) also sets U1 = U, as the reference's initial-
   condition kernels do (e.g. 2D-EulerVortex/ini_sample.hpp:84-88). */
int xf_upload_state(xf_ctx *ctx, int which, const double *U_aos, const double *T);
int xf_download_state(xf_ctx *ctx, int which, double *U_aos, double *T);
/* Primitives of the last xf_update_states call (all padded cells); any
   pointer may be NULL.  y_aos is NS per cell. */
int xf_download_prims(xf_ctx *ctx, double *rho, double *p, double *u, double *v, double *w, double *H,
                      double *c, double *gamma, double *e, double *y_aos);
/* LU (flux divergence) of the last xf_compute_lu, AoS, interior values. */
int xf_download_lu(xf_ctx *ctx, double *LU_aos);
/* GLF maxima eigen_block[3][XF_MAXNS+4] and uvw_c_max[3] of the last dt. */
int xf_download_maxima(xf_ctx *ctx, double *eigen_block, double *uvw_c_max);

/* ---- the stage seam ---------------------------------------------------- */
/* FluidBoundaryCondition (src/solver_BCs/BCs_block.cpp:4-218) on U (0), U1 (1)
   or U2 (2); the same which codes hold for every stage function below. */
int xf_boundary(xf_ctx *ctx, int which);
/* UpdateFluidStateFlux (src/solver_UpdateStates/UpdateStates_block.cpp:7-191):
   rho, Y, u, v, w, p, H, c, gamma, T (Newton), validation.  Also folds the
   GLF maxima (GetLU's eigen/GLF passes, ConVenction_block.hpp:42-172) and the
   CFL maxima (GetDt, GlobalDt_block.hpp:6-114) into the same pass. */
int xf_update_states(xf_ctx *ctx, int which);
/* GetLU (src/solver_Reconstruction/FDM_Method/ConVenction_block.hpp:10-617):
   characteristic WENO face fluxes, the positivity-preserving limiter when the
   case enables it (ConVenction_block.hpp:328-413), the viscous wall fluxes
   when Visc is on (ConVenction_block.hpp:424-575), and LU = -div F. */
int xf_compute_lu(xf_ctx *ctx, int which);
/* UpdateURK3rd (src/solver_UpdateStates/UpdateStates_block.cpp:193-211). */
int xf_rk3_stage(xf_ctx *ctx, int stage, double dt);
/* GetDt (src/solver_GetDt/GlobalDt_block.hpp:6-114) from the maxima of the
   last xf_update_states. */
int xf_get_dt(xf_ctx *ctx, double *dt_out);
/* GetDt (src/solver_GetDt/GlobalDt_block.hpp:6-114) with the reference's
   own arguments instead of a context: the FlowData arrays rho, p, u, v, w
   (device pointers, one double per padded cell, Xmax*Ymax*Zmax of them, the
   FlowData layout of src/include/global_setup.h:231-250; the array of an
   inactive direction may be NULL) and the Block geometry and CFLnumber of
   desc.  uvw_c_max[3] (device, may be NULL) receives max |u_d| +
   sqrt(1.4 p / rho) over all padded cells per active direction (0 for the
   others) and dt_out (device) CFLnumber / (sum_d max_d / Delta_d).  Runs on
   stream (a hipStream_t; NULL = the default stream) with a workspace of one
   maximum triple per reduction block, and returns after the stream drained. */
int xf_get_dt_flow(const xf_case_desc *desc, const double *rho, const double *p, const double *u, const double *v,
                   const double *w, double *uvw_c_max, double *dt_out, void *stream);
/* xf_get_dt_flow without any allocation, for callers that cannot allocate
   device memory (code under HIP graph capture, tight serving loops).  No
   counterpart in the reference.  The same kernels keep their per-block
   maxima in a fixed 96 KiB device array of the library (4096 triples; the
   reduction never uses more blocks), which each GPU holds until the process
   ends.  Cells: at most INT_MAX (Xmax*Ymax*Zmax); a larger grid returns
   XF_EINVAL.  Enqueues on stream and returns without waiting: results land
   in uvw_c_max and dt_out when the stream reaches them.  Calls that overlap
   on one GPU (several streams or host threads) share that array and race;
   serialise them. */
int xf_get_dt_flow_noalloc(const xf_case_desc *desc, const double *rho, const double *p, const double *u,
                           const double *v, const double *w, double *uvw_c_max, double *dt_out, void *stream);
/* ChemeODEQ2Solver (src/solver_Reaction/Reaction_block.hpp:4-38): CHEMEQ2
   over dt_chem on every interior cell, using the primitives of the last
   xf_update_states(0).  substeps (optional, NULL ok) receives the per-cell
   predictor-pass counts (interior cells, row-major i fastest). */
int xf_chemistry(xf_ctx *ctx, double dt_chem, int *substeps);
/* XFLUIDS::Reaction (src/XFLUIDS.cpp:625-642): UpdateStates(U), CHEMEQ2 over
   dt/2 (Strang) or dt (Lie), NaN check. */
int xf_reaction(xf_ctx *ctx, double dt);
/* EstimateFluidNAN (src/Fluids.cpp:963-1040). */
int xf_estimate_nan(xf_ctx *ctx, int flag);

/* ---- whole steps --------------------------------------------------------- */
/* main.cpp:44-47: BoundaryCondition(U) + UpdateStates(U). */
int xf_initialize(xf_ctx *ctx);
/* nsteps of XFLUIDS::Evolution's ordering (src/XFLUIDS.cpp:172-225) with a
   device-resident dt and fused stage kernels; dt values written to dt_out
   (may be NULL).  use_graph = 1 captures one step as a hipGraph and replays it. */
int xf_steps(xf_ctx *ctx, int nsteps, int use_graph, double *dt_out);
/* Enqueue nsteps without waiting (for timing with events on the stream). */
int xf_steps_async(xf_ctx *ctx, int nsteps, int use_graph);
int xf_synchronize(xf_ctx *ctx);
/* The context's HIP stream (hipStream_t as void*). */
void *xf_stream(xf_ctx *ctx);

/* ---- the clock of the time loop ------------------------------------------
   XFLUIDS::Evolution keeps physicalTime and Iteration and clamps each dt so
   that physicalTime lands on the next output time (src/XFLUIDS.cpp:170-198:
   while (physicalTime < target_t) { Iteration++; dt = ComputeTimeStep;
   if (physicalTime + dt > target_t) dt = target_t - physicalTime;
   physicalTime += dt; ... }).  Here the clock lives on the device, so a
   captured step needs no host round trip: each step of xf_steps counts
   itself, clamps its dt to the target and advances the time; once the time
   has reached the target, further steps do nothing (reached = 1) until a
   new target is set.  The default target is +infinity (no clamp).
   xf_set_time sets the time and the step count (e.g. from a checkpoint
   header); xf_set_target_time sets the output time and clears reached. */
int xf_set_time(xf_ctx *ctx, double time, int step);
int xf_set_target_time(xf_ctx *ctx, double target);
int xf_get_time(xf_ctx *ctx, double *time, int *step, int *reached);

/* ---- diagnostics --------------------------------------------------------- */
/* The last physical error (XF_ERR_*) and its cell: the lowest padded-grid
   (i, j, k) whose failure raised that code, in the first pass that failed.
   The reference prints the cell of its error check (UpdateStates_block.cpp:
   80-90, 152-160; Fluids.cpp:1015-1029) and stops the time loop at that
   stage (src/XFLUIDS.cpp:211-225, 285-287, 377-389, 441-525).  In xf_steps
   the failing pass halts the rest of its step and every later step on the
   device, so the state stays as that pass left it (a failure in the fused
   third-stage update has already written U there).  Codes 1 and 2 come
   from UpdateStates, 3 from EstimateNAN (before each stage's update, and
   after each Reaction), 5 from the chemistry. */
int xf_last_error(xf_ctx *ctx, int *code, int *i, int *j, int *k);
/* Step and stage of the last physical error: step = the Iteration counter of
   the device clock (1 for the first step after xf_set_time(ctx, t, 0));
   where = 1, 2, 3 for the RK stage (its UpdateStates or its EstimateNAN,
   the reference's "_RK1".."_RK3" and flags 1-3), 4 for a Reaction ("_React",
   the chemistry and EstimateNAN flag 4).  Zeros for a failure in a stage
   function called on its own. */
int xf_last_error_step(xf_ctx *ctx, int *step, int *where);
/* Average device time (ms) per launch of each kernel class since the last
   reset, measured with HIP events on the context stream when profiling is on. */
int xf_set_profiling(xf_ctx *ctx, int on);
/* Restrict profiling events to the comma-separated kernel classes in csv
   ("states,sweep_x,sweep_y,sweep_z,boundary,rk3,chem"); NULL or "" = all. */
int xf_set_profile_filter(xf_ctx *ctx, const char *csv);
int xf_kernel_times(xf_ctx *ctx, const char **names, double *avg_ms, int *counts, int cap);
/* Average device time (ms) of one kernel class, launched reps times back to
   back on the context stream between two HIP events, in the configuration
   of RK stage 1 of the fused step: "states" (conservative->primitive on U)
   or "sweep_x"/"sweep_y"/"sweep_z" (face fluxes; the last active direction
   carries the fused RK update U1 = U + dt LU); with viscous terms also
   "visc_prep" (transport coefficients and centre gradients) and
   "visc_x"/"visc_y"/"visc_z" (wall fluxes).  Overwrites only LU, U1, the
   primitives and the viscous work arrays, which every stage recomputes. */
int xf_time_kernel(xf_ctx *ctx, const char *name, int reps, double *avg_ms);
/* Sweep tile shape per direction (0 = x, 1 = y, 2 = z): x 0 = 64 faces x 4
   pencils, 1 = 32 x 8; y, z 0 = 16 x 16, 1 = 32 x 8, 2 = 8 x 32.  Every shape
   computes the same bits.  Replaces the fixed work-group sizes of the
   settings file (run.blockSize_x/y/z, src/read_ini/settings/read_json.cpp:112-114). */
int xf_set_tile_shape(xf_ctx *ctx, int dir, int shape);
int xf_get_tile_shape(const xf_ctx *ctx, int *shape3);
/* Viscous wall fluxes (GetWallViscousFlux{X,Y,Z},
   src/solver_Reconstruction/FDM_Method/ConVenction_block.hpp:506-575):
   on = 1 (the default) computes each face's wall flux inside the sweep of
   its direction, for mixtures of four or more species (with fewer, the
   sweep runs more waves per SIMD and keeps its registers for them, and the
   wall flux stays a launch of its own); on = 0 runs a wall-flux launch
   before each sweep, which writes the E flux planes the sweep then reads
   (the reference's structure).  Both give the same bits.  XF_EINVAL
   without viscous terms. */
int xf_set_visc_fusion(xf_ctx *ctx, int on);
/* Autotuner: times every tile shape of every active sweep direction on the
   context's current state (reps back-to-back launches each, stage-1
   configuration) and keeps the fastest.  shape3[3] receives the choice,
   ms[9] (optional) the mean ms per launch, ms[3*dir + shape].  Leaves the
   running GLF maxima unchanged; overwrites LU, U1 and the primitives, which
   every stage recomputes.  Call after xf_upload_state + xf_initialize. */
int xf_autotune(xf_ctx *ctx, int reps, int *shape3, double *ms);
/* ---- ensembles ------------------------------------------------------------
   Independent cases (any grid sizes; one species set, mechanism and thermo
   tables) advanced together.  One ensemble step is every case's own step
   (dt, Strang Reaction, three RK stages, Reaction).  Cases that agree in
   their kernel configuration run each kernel class once per stage over the
   blocks of all cases (a device table maps each block to its case); other
   ensembles run each case's stages on its own stream.  The chemistry of all
   cases is one work-queue launch at each Reaction either way;
   xf_ensemble_steps(..., use_graph = 1) captures the whole step as one
   hipGraph.  Every case's result is bit-identical to stepping it alone.  The contexts stay owned by the caller and must outlive the
   ensemble; read results with xf_download_state on each.  No counterpart in
   the reference, which runs one case per process. */
typedef struct xf_ensemble xf_ensemble;
int xf_ensemble_create(xf_ctx *const *cases, int ncases, xf_ensemble **out);
int xf_ensemble_steps(xf_ensemble *ens, int nsteps, int use_graph);
/* Waits for every case; returns the first case's physical error code. */
int xf_ensemble_synchronize(xf_ensemble *ens);
long long xf_ensemble_cells(const xf_ensemble *ens);
/* 1 when each ensemble step is one launch per kernel class over the cells of
   every case (the cases agree in species count, scheme, precision, active
   directions and the limiter, and none is viscous; the tile shapes are the
   first case's at creation), 0 when each case's stages run on its own
   stream with one shared chemistry launch. */
int xf_ensemble_one_launch(const xf_ensemble *ens);
void xf_ensemble_destroy(xf_ensemble *ens);
/* Evaluate the device exp and log (the fixed-sequence FP64 functions every
   kernel uses) on n host values; for bit-level checks against the CPU. */
int xf_debug_math(const double *x, double *exp_out, double *log_out, int n);
/* Streaming copy of nbytes (a multiple of 16) between two fresh device
   buffers, reps times after one warm-up launch on the null stream; average
   ms per launch.  A plain copy of known size: the HBM bandwidth reference
   and the scale for the memory-traffic counters.  No reference counterpart. */
int xf_copy_probe(size_t nbytes, int reps, double *avg_ms);
/* FP64 fused multiply-add throughput: eight independent fma chains per lane
   on 2048 blocks of 256 lanes, reps launches after one warm-up on the null
   stream; TFLOP/s with two flops per fma.  The measured counterpart of the
   datasheet FP64 peak the roofline quotes.  No reference counterpart. */
int xf_fma_probe(int reps, double *tflops);

#ifdef __cplusplus
}
#endif
#endif
