// Host-side case setup: the runtime form of the reference's Setup
// (src/read_ini/setupini.h:19-83).  Reads the settings/*.json schema
// (keys of src/read_ini/settings/read_json.cpp:30-148) and the runtime.dat
// files (species_list.dat, thermal_dynamics.dat, thermal_dynamics_janaf.dat,
// reaction_list.dat), and derives the block geometry exactly as
// Setup::init (src/read_ini/src/iniset.cpp:290-369).
//
// The reference fixes the mixture, thermo fit and scheme at compile time
// (CMake MIXTURE_MODEL, THERMAL, WENO_ORDER).  Here they are runtime keys
// of the "equations" section: "Mixture" (directory under runtime.dat, a
// string or an array of path components),
// "Thermo" ("NASA" | "JANAF"), "Scheme" ("WENO5" | "WENO5Z").
#pragma once
#include "../../include/xf_api.h"
#include "xf_json.h"
#include <string>
#include <vector>

namespace xfh {

struct Setup {
    Json conf;
    xf_case_desc desc{};
    // domain (Block)
    double Domain_xmin = 0, Domain_ymin = 0, Domain_zmin = 0;
    double Domain_length = 1, Domain_width = 1, Domain_height = 1;
    double dx = 1, dy = 1, dz = 1;
    int nStepMax = 10;
    std::string mixture, slip_order = "Strang";
    // init section (read_json.cpp:132-148)
    double blast_mach = 0.0;
    // init.bubble_shape_x (bubble radius) and init.bubble_boundary_cells
    // (interface width in cells) of settings/shock-bubble.json; 0 = defaults
    double bubble_r = 0.0, bubble_cells = 0.0;
    std::vector<double> blast_center{0.5, 0.5, 0.5}, cop_center{0, 0, 0};
    std::vector<double> blast_up{0, 0, 298.15, 0, 0, 0}, blast_down{0, 0, 298.15, 0, 0, 0}, cop_in;
    // species (thermal.cpp:6-32): names, molar and mass fractions outside / inside
    std::vector<std::string> species;
    std::vector<double> xi_out, xi_in, Y_out, Y_in;
    // thermal tables (reference layout, Wi in kg/mol)
    std::vector<double> Hia, Hib, Wi;
    // mechanism tables (reference layout)
    int NR = 0;
    std::vector<int> Nu_f, Nu_b;
    std::vector<double> third, Rargus;
    // transport (viscous terms): species characteristics (Specie_chara
    // layout, 9 per species), fitting temperatures, fits
    std::vector<double> chara, Tnode, fcv, fct, Dkj;
    double Yil_limiter_json = 1.0E10, Dim_limiter_json = 2.0E-3;
    std::string runtime_root;
};

// settings: path of the case JSON; runtime_root: directory holding
// thermal_dynamics*.dat and the mixture sub-directories.
// thermo_override: -1 keeps the settings file's "Thermo", 1 NASA-9, 0 JANAF
bool setup_load(const std::string &settings, const std::string &runtime_root, Setup &s, std::string &err,
                int thermo_override = -1);

// Lennard-Jones data and transport fits (Setup::GetFitCoefficient); needs
// species, Wi, the thermo tables and Tnode
bool transport_load(const std::string &root, Setup &s, std::string &err);
// kinetic-theory value behind a fit: what 0 viscosity, 1 conductivity, 2 D_kj
double transport_exact(const Setup &s, const std::string &root, int what, int k, int j, double T, double PP);
// BlSz.Dim_limiter and BlSz.Yil_limiter from the settings values and _dx, _dy, _dz
void set_visc_limiters(Setup &s);

// host thermodynamics for initial conditions (plain libm; not the hot path)
double host_h_species(const Setup &s, double T, int n);          // J/kg
double host_cp_species(const Setup &s, double T, int n);         // J/kg/K
double host_mix_R(const Setup &s, const double *Y);               // J/kg/K
double host_mix_h(const Setup &s, const double *Y, double T);     // J/kg
double host_mix_cp(const Setup &s, const double *Y, double T);    // J/kg/K
void host_x_to_y(const Setup &s, const double *x, double *y);     // get_yi (Mixing_device.h:18-26)

// Synthetic initial states (AoS U over padded cells, T per cell).
// workload: "isentropic-vortex", "shock-bubble", "reactive-shock-tube",
// "mixing-layer", "reactive-channel".  Returns false with err on failure.
bool setup_initial(const Setup &s, const std::string &workload, unsigned long long seed, double *U, double *T,
                   std::string &err);

} // namespace xfh
