#include "xf_setup.h"
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>

namespace xfh {

namespace {

const double kRu = 6.02214076e26 * 1.380649e-23 * 1.0E-3; // Ru, global_setup.h:48-54

bool read_tokens(const std::string &path, std::vector<std::string> &tok)
{
    std::ifstream in(path);
    if (!in) return false;
    std::string w;
    while (in >> w) tok.push_back(w);
    return true;
}

// species_list.dat: names on the first line, molar ratios outside and inside
// (Setup::ReadSpecies, src/read_ini/src/thermal.cpp:6-32)
bool read_species(const std::string &path, Setup &s, std::string &err)
{
    std::ifstream in(path);
    if (!in) {
        err = "cannot open " + path;
        return false;
    }
    std::string line;
    std::getline(in, line);
    std::istringstream ls(line);
    std::string w;
    while (ls >> w) s.species.push_back(w);
    const size_t NS = s.species.size();
    if (NS == 0) {
        err = "no species in " + path;
        return false;
    }
    s.xi_out.assign(NS, 0.0);
    s.xi_in.assign(NS, 0.0);
    for (size_t n = 0; n < NS; n++)
        if (!(in >> s.xi_out[n])) s.xi_out[n] = 0.0;
    for (size_t n = 0; n < NS; n++)
        if (!(in >> s.xi_in[n])) s.xi_in[n] = s.xi_out[n];
    return true;
}

// thermal_dynamics.dat (NASA-9: 3 x (7 a + 2 b) + W) and
// thermal_dynamics_janaf.dat (NASA-7: 7 high + 7 low + W), located by the
// "*NAME" keyword (Setup::ReadThermal, thermal.cpp:36-127)
bool read_thermo(const std::string &root, int thermo, Setup &s, std::string &err)
{
    const size_t NS = s.species.size();
    s.Hia.assign(NS * 21, 0.0);
    s.Hib.assign(NS * 6, 0.0);
    s.Wi.assign(NS, 0.0);
    std::vector<std::string> t9, t7;
    if (!read_tokens(root + "/thermal_dynamics.dat", t9)) {
        err = "cannot open " + root + "/thermal_dynamics.dat";
        return false;
    }
    if (!read_tokens(root + "/thermal_dynamics_janaf.dat", t7)) {
        err = "cannot open " + root + "/thermal_dynamics_janaf.dat";
        return false;
    }
    for (size_t n = 0; n < NS; n++) {
        const std::string key = "*" + s.species[n];
        std::vector<double> nasa9, janaf;
        for (size_t q = 0; q < t9.size(); q++) {
            if (t9[q] == "*END") break;
            if (t9[q] == key) {
                for (size_t r = 1; r <= 28 && q + r < t9.size(); r++) nasa9.push_back(std::atof(t9[q + r].c_str()));
                break;
            }
        }
        for (size_t q = 0; q < t7.size(); q++) {
            if (t7[q] == "*END") break;
            if (t7[q] == key) {
                for (size_t r = 1; r <= 15 && q + r < t7.size(); r++) janaf.push_back(std::atof(t7[q + r].c_str()));
                break;
            }
        }
        if (thermo == 1) {
            if (nasa9.size() != 28) {
                err = "species " + s.species[n] + " not in thermal_dynamics.dat";
                return false;
            }
            for (int rg = 0; rg < 3; rg++) {
                for (int m = 0; m < 7; m++) s.Hia[n * 21 + m * 3 + rg] = nasa9[rg * 9 + m];
                for (int m = 0; m < 2; m++) s.Hib[n * 6 + m * 3 + rg] = nasa9[rg * 9 + 7 + m];
            }
            s.Wi[n] = nasa9[27] * 1e-3;
        } else {
            if (janaf.size() != 15) {
                err = "species " + s.species[n] + " not in thermal_dynamics_janaf.dat";
                return false;
            }
            for (int m = 0; m < 7; m++) s.Hia[n * 21 + m * 3 + 0] = janaf[m];
            for (int m = 0; m < 7; m++) s.Hia[n * 21 + m * 3 + 1] = janaf[7 + m];
            s.Wi[n] = janaf[14] * 1e-3;
        }
    }
    return true;
}

// reaction_list.dat sections (Setup::ReadReactions, reaction.cpp:6-58)
bool read_reactions(const std::string &path, Setup &s, std::string &err)
{
    std::vector<std::string> tok;
    if (!read_tokens(path, tok)) {
        s.NR = 0;
        return true; // inert mixture
    }
    const size_t NS = s.species.size();
    std::vector<double> fw, bw, tb, ar;
    std::vector<double> *cur = nullptr;
    for (const std::string &w : tok) {
        if (w == "*forward_reaction") cur = &fw;
        else if (w == "*backward_reaction") cur = &bw;
        else if (w == "*third_body") cur = &tb;
        else if (w == "*Arrhenius") cur = &ar;
        else if (!w.empty() && w[0] == '*') cur = nullptr;
        else if (cur) cur->push_back(std::atof(w.c_str()));
    }
    if (fw.size() % NS != 0) {
        err = "malformed *forward_reaction in " + path;
        return false;
    }
    s.NR = (int)(fw.size() / NS);
    if (bw.size() < fw.size() || tb.size() < fw.size() || ar.size() < (size_t)s.NR * 3) {
        err = "incomplete mechanism in " + path;
        return false;
    }
    s.Nu_f.resize(fw.size());
    s.Nu_b.resize(fw.size());
    s.third.assign(fw.begin(), fw.end());
    for (size_t q = 0; q < fw.size(); q++) {
        s.Nu_f[q] = (int)fw[q];
        s.Nu_b[q] = (int)bw[q];
        s.third[q] = tb[q];
    }
    s.Rargus.assign((size_t)s.NR * 6, 0.0);
    for (int r = 0; r < s.NR; r++)
        for (int q = 0; q < 3; q++) s.Rargus[r * 6 + q] = ar[r * 3 + q];
    return true;
}

int bc_of(const std::vector<double> &b, int q) { return q < (int)b.size() ? (int)b[q] : 2; }

} // namespace

void host_x_to_y(const Setup &s, const double *x, double *y)
{
    const size_t NS = s.species.size();
    double W = 0.0;
    for (size_t i = 0; i < NS; i++) W += x[i] * s.Wi[i];
    const double iW = 1.0 / W;
    for (size_t n = 0; n < NS; n++) y[n] = x[n] * s.Wi[n] * iW;
}

bool setup_load(const std::string &settings, const std::string &runtime_root, Setup &s, std::string &err,
                int thermo_override)
{
    s = Setup();
    if (!json_read_settings(settings, s.conf, err)) return false;
    static const Json empty = [] {
        Json j;
        j.kind = Json::Object;
        return j;
    }();
    const Json *run = s.conf.get("run"), *eq = s.conf.get("equations"), *mesh = s.conf.get("mesh"),
               *ini = s.conf.get("init");
    if (!run) run = &empty;
    if (!eq) eq = &empty;
    if (!mesh) mesh = &empty;
    if (!ini) ini = &empty;

    xf_case_desc &d = s.desc;
    std::memset(&d, 0, sizeof(d));
    d.CFLnumber = run->value("CFLnumber", 0.4);
    s.nStepMax = (int)run->value("nStepMax", 10.0);

    // mixture, thermo fit, scheme
    // "Mixture": directory under runtime.dat, given as a string or as an array
    // of path components (the settings reader cuts lines at the first '/')
    s.mixture = eq->value("Mixture", std::string("NO-COP"));
    if (const Json *mx = eq->get("Mixture"); mx && mx->kind == Json::Array) {
        s.mixture.clear();
        for (const Json &e : mx->arr) s.mixture += (s.mixture.empty() ? "" : "/") + e.str;
    }
    const std::string thermo = eq->value("Thermo", std::string("NASA"));
    const std::string scheme = eq->value("Scheme", std::string("WENO5"));
    d.thermo = (thermo == "JANAF") ? 0 : 1;
    if (thermo_override >= 0) d.thermo = thermo_override ? 1 : 0;
    d.scheme = (scheme == "WENO5Z") ? XF_WENO5Z : (scheme == "WENOCU6" ? XF_WENOCU6 : XF_WENO5JS);
    s.slip_order = eq->value("SlipOrder", std::string("Strang"));
    d.strang = s.slip_order == "Strang";
    d.precision = XF_PLAIN_FP64;
    d.chem_driver = XF_CHEM_PER_CELL;
    d.ncop_gamma = 1.4;

    const std::string mixdir = runtime_root + "/" + s.mixture;
    if (!read_species(mixdir + "/species_list.dat", s, err)) return false;
    const int NS = (int)s.species.size();
    d.NS = NS;
    d.cop = !(NS == 1 && s.species[0] == "NCOP");
    if (d.cop) {
        if (!read_thermo(runtime_root, d.thermo, s, err)) return false;
        if (!read_reactions(mixdir + "/reaction_list.dat", s, err)) return false;
    } else {
        s.Hia.assign(21, 0.0);
        s.Hib.assign(6, 0.0);
        s.Wi.assign(1, 1.0);
        s.NR = 0;
    }
    d.NR = s.NR;
    // GhostSpecies: compile-time in the reference, set by a mixture's
    // case_setup.h (runtime.dat/Inert-SBI, Reaction/RSBI-19REA, ...): the last
    // species of the list is a ghost copy that carries no mass
    d.ghost = eq->value("GhostSpecies", 0.0) != 0.0 ? 1 : 0;
    if (d.ghost && (!d.cop || NS < 2)) {
        err = "GhostSpecies needs a multi-component mixture";
        return false;
    }
    // Sources_React defaults to COP_CHEME: on for reacting mixtures
    d.react = (int)eq->value("Sources_React", (d.cop && s.NR > 0) ? 1.0 : 0.0) && s.NR > 0;
    // equations.PositivityPreserving (settings/*.json): the flux limiter
    d.positivity = eq->value("PositivityPreserving", 0.0) != 0.0 ? 1 : 0;
    // viscous terms: compile-time Visc / Visc_Heat / Visc_Diffu in the
    // reference (CMakeLists.txt:55-57, init_options.cmake:81-91); fitting
    // temperatures and limiters as read_json.cpp:70,104-105
    d.visc = eq->value("Visc", 0.0) != 0.0 ? 1 : 0;
    d.visc_heat = d.visc && eq->value("Visc_Heat", 0.0) != 0.0 ? 1 : 0;
    d.visc_diffu = d.visc && eq->value("Visc_Diffu", 0.0) != 0.0 ? 1 : 0;
    s.Tnode = eq->values("ViscosityFittingTnode", {273.15, 500.0, 750.0, 1000.0, 1250.0, 1500.0, 1750.0, 2000.0,
                                                   2250.0, 2500.0, 2750.0, 3000.0, 5000.0});
    s.Yil_limiter_json = mesh->value("Yil_limiter", 1.0E10);
    s.Dim_limiter_json = mesh->value("Dim_limiter", 2.0E-3);
    s.runtime_root = runtime_root;
    s.Y_out.assign(NS, 1.0);
    s.Y_in.assign(NS, 1.0);
    if (d.cop) {
        host_x_to_y(s, s.xi_out.data(), s.Y_out.data());
        host_x_to_y(s, s.xi_in.data(), s.Y_in.data());
    }

    // geometry: Setup::ReadIni + Setup::init
    const std::vector<double> res = mesh->values("Resolution", {1, 0, 0});
    const std::vector<double> gw = mesh->values("Ghost_width", {4, 4, 4});
    const std::vector<double> size = mesh->values("DOMAIN_Size", {1.0, 1.0, 1.0});
    const std::vector<double> medg = mesh->values("DOMAIN_Medg", {0.0, 0.0, 0.0});
    const std::vector<double> bcs = mesh->values("Boundarys", {2, 2, 2, 2, 2, 2});
    d.DimX = res.size() > 0 && res[0] != 0;
    d.DimY = res.size() > 1 && res[1] != 0;
    d.DimZ = res.size() > 2 && res[2] != 0;
    d.X_inner = d.DimX ? (int)res[0] : 1;
    d.Y_inner = d.DimY ? (int)res[1] : 1;
    d.Z_inner = d.DimZ ? (int)res[2] : 1;
    d.Bwidth_X = d.DimX ? (int)gw[0] : 0;
    d.Bwidth_Y = d.DimY ? (int)gw[1] : 0;
    d.Bwidth_Z = d.DimZ ? (int)gw[2] : 0;
    s.Domain_xmin = medg[0], s.Domain_ymin = medg[1], s.Domain_zmin = medg[2];
    s.Domain_length = d.DimX ? size[0] : 1.0;
    s.Domain_width = d.DimY ? size[1] : 1.0;
    s.Domain_height = d.DimZ ? size[2] : 1.0;
    s.dx = d.DimX ? s.Domain_length / double(1 * d.X_inner) : 1.0;
    s.dy = d.DimY ? s.Domain_width / double(1 * d.Y_inner) : 1.0;
    s.dz = d.DimZ ? s.Domain_height / double(1 * d.Z_inner) : 1.0;
    d.Xmax = d.DimX ? d.X_inner + 2 * d.Bwidth_X : 1;
    d.Ymax = d.DimY ? d.Y_inner + 2 * d.Bwidth_Y : 1;
    d.Zmax = d.DimZ ? d.Z_inner + 2 * d.Bwidth_Z : 1;
    d.dx = s.dx, d.dy = s.dy, d.dz = s.dz;
    d._dx = 1.0 / s.dx, d._dy = 1.0 / s.dy, d._dz = 1.0 / s.dz;
    for (int q = 0; q < 6; q++) d.BC[q] = bc_of(bcs, q);
    set_visc_limiters(s);
    if (d.visc) {
        if (!d.cop) {
            err = "viscous terms need a multi-component mixture";
            return false;
        }
        if (!transport_load(runtime_root, s, err)) return false;
    }

    // init section
    s.blast_mach = ini->value("blast_mach", 0.0);
    s.bubble_r = ini->value("bubble_shape_x", 0.0);
    s.bubble_cells = ini->value("bubble_boundary_cells", 0.0);
    s.blast_center = ini->values("blast_center", {0.5, 0.5, 0.5});
    s.cop_center = ini->values("cop_center", {0.0, 0.0, 0.0});
    s.blast_up = ini->values("blast_upstream", {0, 0, 298.15, 0, 0, 0});
    s.blast_down = ini->values("blast_downstream", {0, 0, 298.15, 0, 0, 0});
    s.cop_in = ini->values("cop_inside", {s.blast_down[0], s.blast_down[1], s.blast_down[2], 0, 0, 0});
    return true;
}

void set_visc_limiters(Setup &s)
{
    // Setup::init (src/read_ini/src/iniset.cpp:358-359), std::min / std::max
    xf_case_desc &d = s.desc;
    d.Dim_limiter = std::min(std::max(s.Dim_limiter_json, 0.0), 1.0);
    d.Yil_limiter = std::max(std::max(d._dx, d._dy), d._dz) * std::min(std::max(s.Yil_limiter_json, 0.0), 1.0);
}

// ---- host thermodynamics (initial conditions only) -------------------------
static int nasa_range(double T) { return (T >= 1000.0 && T < 6000.0) ? 1 : (T < 1000.0 ? 0 : 2); }

double host_cp_species(const Setup &s, double T0, int n)
{
    const double T = std::max(T0, 200.0), Ri = kRu / s.Wi[n];
    const double *a = &s.Hia[n * 21];
    if (s.desc.thermo) {
        const int r = nasa_range(T);
        return Ri * (a[0 + r] / (T * T) + a[3 + r] / T + a[6 + r] + a[9 + r] * T + a[12 + r] * T * T +
                     a[15 + r] * T * T * T + a[18 + r] * T * T * T * T);
    }
    const int r = T > 1000.0 ? 0 : 1;
    return Ri * (a[0 + r] + a[3 + r] * T + a[6 + r] * T * T + a[9 + r] * T * T * T + a[12 + r] * T * T * T * T);
}

double host_h_species(const Setup &s, double T0, int n)
{
    const double T = std::max(T0, 200.0), Ri = kRu / s.Wi[n];
    const double *a = &s.Hia[n * 21];
    double h;
    if (s.desc.thermo) {
        const int r = nasa_range(T);
        h = Ri * (-a[0 + r] / T + a[3 + r] * std::log(T) + a[6 + r] * T + a[9 + r] * T * T / 2 +
                  a[12 + r] * T * T * T / 3 + a[15 + r] * T * T * T * T / 4 + a[18 + r] * T * T * T * T * T / 5 +
                  s.Hib[n * 6 + r]);
    } else {
        const int r = T > 1000.0 ? 0 : 1;
        h = Ri * (a[0 + r] * T + a[3 + r] * T * T / 2 + a[6 + r] * T * T * T / 3 + a[9 + r] * T * T * T * T / 4 +
                  a[12 + r] * T * T * T * T * T / 5 + a[15 + r]);
    }
    if (T0 < 200.0) h += host_cp_species(s, 200.0, n) * (T0 - 200.0);
    return h;
}

double host_mix_R(const Setup &s, const double *Y)
{
    double r = 0.0;
    for (size_t n = 0; n < s.species.size(); n++) r += Y[n] / s.Wi[n];
    return r * kRu;
}

double host_mix_h(const Setup &s, const double *Y, double T)
{
    double h = 0.0;
    for (size_t n = 0; n < s.species.size(); n++) h += Y[n] * host_h_species(s, T, (int)n);
    return h;
}

double host_mix_cp(const Setup &s, const double *Y, double T)
{
    double c = 0.0;
    for (size_t n = 0; n < s.species.size(); n++) c += Y[n] * host_cp_species(s, T, (int)n);
    return c;
}

} // namespace xfh
