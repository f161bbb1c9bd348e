/*
 * xf_host.h -- C-ABI of the host layer above the kernels: case loading
 * from the reference's settings/*.json + runtime.dat formats, synthetic
 * initial states, and a driver that sequences the stage seam exactly as
 * XFLUIDS::Evolution does.
 */
#ifndef XF_HOST_H
#define XF_HOST_H

#include "xf_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xf_setup xf_setup;

/* Setup::Setup (src/read_ini/src/constructor.cpp:12-77): read the settings
   JSON and the runtime.dat tables; geometry as Setup::init
   (src/read_ini/src/iniset.cpp:290-369).  err (may be NULL) receives a
   message of at most errcap bytes. */
int xf_setup_load(const char *settings_json, const char *runtime_root, xf_setup **out, char *err, size_t errcap);
void xf_setup_free(xf_setup *s);
/* Override mesh resolution / CFL / scheme / thermo before use (recomputes
   the geometry as Setup::init); pass 0 (or < 0 for cfl) to keep a value. */
int xf_setup_override(xf_setup *s, int nx, int ny, int nz, double cfl, int scheme, int thermo);
/* Switch the positivity-preserving flux limiter on (1) or off (0), as the
   settings key equations.PositivityPreserving does (src/read_ini/settings/read_json.cpp:68). */
int xf_setup_set_positivity(xf_setup *s, int on);
/* Switch the Reaction stage on (1, needs a mechanism) or off (0), as the
   settings key equations.Sources_React does (Sources_React, src/XFLUIDS.cpp:211-225):
   off gives the inert (fluid-only) step of a reacting case. */
int xf_setup_set_reaction(xf_setup *s, int on);
/* Switch the viscous terms (Visc, Visc_Heat, Visc_Diffu of CMakeLists.txt:55-57,
   settings keys equations.Visc / Visc_Heat / Visc_Diffu here); fits the
   transport properties (Setup::GetFitCoefficient, viscfit.cpp:148-193) when
   visc is switched on.  Needs a multi-component mixture. */
int xf_setup_set_viscous(xf_setup *s, int visc, int heat, int diffu);
/* The fitted transport tables of xf_set_transport: fcv[NS*4], fct[NS*4],
   Dkj[NS*NS*4] (zeros when the case has no viscous terms). */
int xf_setup_transport(const xf_setup *s, double *fcv, double *fct, double *Dkj);
/* The kinetic-theory values the fits are made from (Setup::viscosity,
   thermal_conductivities and Dkj, viscfit.cpp:444-555): what = 0 mu_k [Pa s],
   1 lambda_k [W/(m K)] at pressure p, 2 D_kj [cm^2/s] at pressure p [Pa].
   NaN on error. */
double xf_setup_transport_exact(const xf_setup *s, int what, int k, int j, double T, double p);
int xf_setup_desc(const xf_setup *s, xf_case_desc *out);
/* Thermal tables Hia[NS*21], Hib[NS*6], Wi[NS] (kg/mol). */
int xf_setup_thermo(const xf_setup *s, double *Hia, double *Hib, double *Wi);
/* Mechanism tables Nu_f[NR*NS], Nu_b[NR*NS], third[NR*NS], Rargus[NR*6]. */
int xf_setup_mechanism(const xf_setup *s, int *Nu_f, int *Nu_b, double *third, double *Rargus);
/* Mass fractions outside / inside the bubble (species_list.dat ratios). */
int xf_setup_mass_fractions(const xf_setup *s, double *Y_out, double *Y_in);
/* Species name n. */
const char *xf_setup_species(const xf_setup *s, int n);
/* Domain origin and size [xmin, ymin, zmin, Lx, Ly, Lz]. */
int xf_setup_domain(const xf_setup *s, double *dom6);
/* Synthetic initial state (AoS U over padded cells, T per cell). */
int xf_setup_initial(const xf_setup *s, const char *workload, unsigned long long seed, double *U_aos, double *T);
/* Create a device context for the setup and upload thermo + mechanism. */
int xf_setup_context(const xf_setup *s, int precision, int chem_driver, xf_ctx **out);

/* XFLUIDS::Evolution (src/XFLUIDS.cpp:172-225) through the stage seam, one
   call per reference stage: ComputeTimeStep, [Reaction], 3 x
   (BoundaryCondition, UpdateStates, ComputeLU, EstimateNAN, UpdateU),
   [Reaction].  Call xf_initialize first (main.cpp:44-47). */
int xf_evolve_staged(xf_ctx *ctx, int nsteps, double *dt_out);

/* ---- checkpoints ----------------------------------------------------------
   The reference's checkpoint file (XFLUIDS::Output_Ubak / Read_Ubak,
   src/XFLUIDS.cpp:658-724): int Step, double Time, float Time_consumption
   (4 + 8 + 4 bytes, no padding), then the conserved state with ghost cells
   as the device holds it in the reference, AoS U[Emax*id + n] over
   Xmax*Ymax*Zmax cells (Emax*Xmax*Ymax*Zmax doubles, iniset.cpp:345).  The
   temperature is not part of the format: a restart starts the Newton
   inversion from the caller's T.  Host-only functions; the file size must
   match the descriptor exactly. */
int xf_checkpoint_write(const char *path, const xf_case_desc *desc, const double *U_aos, int step, double time,
                        float time_consumption);
int xf_checkpoint_read(const char *path, const xf_case_desc *desc, double *U_aos, int *step, double *time,
                       float *time_consumption);
/* CopyToUbak + Output_Ubak / Read_Ubak + CopyToU for a context: download U and
   write it; read and upload it (which = 0, so U1 = U as after
   initialisation), keeping the context's current T as the Newton guess; the
   load also sets the device clock (xf_set_time) from the header. */
int xf_save_checkpoint(xf_ctx *ctx, const char *path, int step, double time, float time_consumption);
int xf_load_checkpoint(xf_ctx *ctx, const char *path, int *step, double *time, float *time_consumption);

#ifdef __cplusplus
}
#endif
#endif
