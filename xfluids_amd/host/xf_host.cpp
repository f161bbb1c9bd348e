// Host C-ABI (include/xf_host.h): case loading, initial states, context
// creation, and the stage-by-stage driver mirroring the reference's
// Fluid / XFLUIDS classes (src/Fluids.cpp:897-1045, src/XFLUIDS.cpp:105-642).
#include "../../include/xf_host.h"
#include "xf_setup.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

struct xf_setup {
    xfh::Setup s;
    std::string settings, root;
};

namespace {

void put_err(char *err, size_t cap, const std::string &m)
{
    if (err && cap) {
        std::snprintf(err, cap, "%s", m.c_str());
    }
}

// Fluid (src/global_class.h:14-54): the stage methods the time loop calls
struct Fluid {
    xf_ctx *c;
    int UpdateFluidStates(int flag) { return xf_update_states(c, flag); }   // Fluids.cpp:935-944
    int ComputeFluidLU(int flag) { return xf_compute_lu(c, flag); }         // Fluids.cpp:951-961
    int UpdateFluidURK3(int flag, double dt) { return xf_rk3_stage(c, flag, dt); } // Fluids.cpp:946-949
    int GetFluidDt(double *dt) { return xf_get_dt(c, dt); }                // Fluids.cpp:897-920
    int BoundaryCondition(int flag) { return xf_boundary(c, flag); }       // Fluids.cpp:922-933
    int EstimateFluidNAN(int flag) { return xf_estimate_nan(c, flag); }    // Fluids.cpp:963-1040
    int ODESolver(double dt) { return xf_chemistry(c, dt, nullptr); }      // Fluids.cpp:1042-1045
};

// XFLUIDS::RungeKuttaSP3rd (src/XFLUIDS.cpp:441-525)
int rk_stage(Fluid &fl, int flag, double dt)
{
    const int which = (flag == 1) ? 0 : 1;
    int e = fl.BoundaryCondition(which);
    if (!e) e = fl.UpdateFluidStates(which);
    if (!e) e = fl.ComputeFluidLU(which);
    if (!e) e = fl.EstimateFluidNAN(flag);
    if (!e) e = fl.UpdateFluidURK3(flag, dt);
    return e;
}

} // namespace

extern "C" {

int xf_setup_load(const char *settings_json, const char *runtime_root, xf_setup **out, char *err, size_t errcap)
{
    if (!settings_json || !runtime_root || !out) return XF_EINVAL;
    xf_setup *h = new xf_setup();
    std::string m;
    if (!xfh::setup_load(settings_json, runtime_root, h->s, m)) {
        put_err(err, errcap, m);
        delete h;
        *out = nullptr;
        return XF_EINVAL;
    }
    h->settings = settings_json;
    h->root = runtime_root;
    *out = h;
    return XF_OK;
}

void xf_setup_free(xf_setup *s) { delete s; }

int xf_setup_override(xf_setup *h, int nx, int ny, int nz, double cfl, int scheme, int thermo)
{
    if (!h) return XF_EINVAL;
    xfh::Setup &s = h->s;
    if (cfl > 0) s.desc.CFLnumber = cfl;
    if (scheme >= 0) s.desc.scheme = scheme;
    if (thermo >= 0 && thermo != s.desc.thermo) {
        // the fit decides which table file is read (thermal.cpp:81-127): reload
        xfh::Setup t;
        std::string m;
        if (!xfh::setup_load(h->settings, h->root, t, m, thermo)) return XF_EINVAL;
        t.desc.CFLnumber = s.desc.CFLnumber;
        t.desc.scheme = s.desc.scheme;
        s = t;
    }
    xf_case_desc &d = s.desc;
    const int res[3] = {nx, ny, nz};
    int *inner[3] = {&d.X_inner, &d.Y_inner, &d.Z_inner};
    const int dims[3] = {d.DimX, d.DimY, d.DimZ};
    double *L[3] = {&s.Domain_length, &s.Domain_width, &s.Domain_height};
    double *dl[3] = {&s.dx, &s.dy, &s.dz};
    int *Bw[3] = {&d.Bwidth_X, &d.Bwidth_Y, &d.Bwidth_Z};
    int *mx[3] = {&d.Xmax, &d.Ymax, &d.Zmax};
    for (int q = 0; q < 3; q++) {
        if (res[q] <= 0 || !dims[q]) continue;
        *inner[q] = res[q];
        *dl[q] = *L[q] / double(res[q]);
        *mx[q] = res[q] + 2 * *Bw[q];
    }
    d.dx = s.dx, d.dy = s.dy, d.dz = s.dz;
    d._dx = 1.0 / s.dx, d._dy = 1.0 / s.dy, d._dz = 1.0 / s.dz;
    xfh::set_visc_limiters(s);
    return XF_OK;
}

int xf_setup_set_positivity(xf_setup *h, int on)
{
    if (!h || (on != 0 && on != 1)) return XF_EINVAL;
    h->s.desc.positivity = on;
    return XF_OK;
}

int xf_setup_set_reaction(xf_setup *h, int on)
{
    if (!h || (on != 0 && on != 1)) return XF_EINVAL;
    if (on && (h->s.NR <= 0 || !h->s.desc.cop)) return XF_EINVAL;
    h->s.desc.react = on;
    return XF_OK;
}

int xf_setup_set_viscous(xf_setup *h, int visc, int heat, int diffu)
{
    if (!h || visc < 0 || visc > 1 || heat < 0 || heat > 1 || diffu < 0 || diffu > 1) return XF_EINVAL;
    xfh::Setup &s = h->s;
    if (visc && !s.desc.cop) return XF_EINVAL;
    s.desc.visc = visc;
    s.desc.visc_heat = visc && heat;
    s.desc.visc_diffu = visc && diffu;
    if (visc && s.fcv.empty()) {
        std::string m;
        if (!xfh::transport_load(s.runtime_root, s, m)) {
            fprintf(stderr, "xfluids_amd: %s\n", m.c_str());
            s.desc.visc = s.desc.visc_heat = s.desc.visc_diffu = 0;
            return XF_EINVAL;
        }
    }
    return XF_OK;
}

int xf_setup_transport(const xf_setup *h, double *fcv, double *fct, double *Dkj)
{
    if (!h || !fcv || !fct || !Dkj) return XF_EINVAL;
    const xfh::Setup &s = h->s;
    const size_t NS = s.species.size();
    std::fill(fcv, fcv + NS * 4, 0.0);
    std::fill(fct, fct + NS * 4, 0.0);
    std::fill(Dkj, Dkj + NS * NS * 4, 0.0);
    if (!s.fcv.empty()) {
        std::copy(s.fcv.begin(), s.fcv.end(), fcv);
        std::copy(s.fct.begin(), s.fct.end(), fct);
        std::copy(s.Dkj.begin(), s.Dkj.end(), Dkj);
    }
    return XF_OK;
}

double xf_setup_transport_exact(const xf_setup *h, int what, int k, int j, double T, double p)
{
    if (!h || h->s.chara.empty()) return NAN;
    const int NS = (int)h->s.species.size();
    if (what < 0 || what > 2 || k < 0 || k >= NS || j < 0 || j >= NS) return NAN;
    return xfh::transport_exact(h->s, h->s.runtime_root, what, k, j, T, p);
}

int xf_setup_desc(const xf_setup *h, xf_case_desc *out)
{
    if (!h || !out) return XF_EINVAL;
    *out = h->s.desc;
    return XF_OK;
}

int xf_setup_thermo(const xf_setup *h, double *Hia, double *Hib, double *Wi)
{
    if (!h) return XF_EINVAL;
    const xfh::Setup &s = h->s;
    if (Hia) std::memcpy(Hia, s.Hia.data(), s.Hia.size() * sizeof(double));
    if (Hib) std::memcpy(Hib, s.Hib.data(), s.Hib.size() * sizeof(double));
    if (Wi) std::memcpy(Wi, s.Wi.data(), s.Wi.size() * sizeof(double));
    return XF_OK;
}

int xf_setup_mechanism(const xf_setup *h, int *Nu_f, int *Nu_b, double *third, double *Rargus)
{
    if (!h) return XF_EINVAL;
    const xfh::Setup &s = h->s;
    if (Nu_f && !s.Nu_f.empty()) std::memcpy(Nu_f, s.Nu_f.data(), s.Nu_f.size() * sizeof(int));
    if (Nu_b && !s.Nu_b.empty()) std::memcpy(Nu_b, s.Nu_b.data(), s.Nu_b.size() * sizeof(int));
    if (third && !s.third.empty()) std::memcpy(third, s.third.data(), s.third.size() * sizeof(double));
    if (Rargus && !s.Rargus.empty()) std::memcpy(Rargus, s.Rargus.data(), s.Rargus.size() * sizeof(double));
    return XF_OK;
}

int xf_setup_mass_fractions(const xf_setup *h, double *Y_out, double *Y_in)
{
    if (!h) return XF_EINVAL;
    const xfh::Setup &s = h->s;
    if (Y_out) std::memcpy(Y_out, s.Y_out.data(), s.Y_out.size() * sizeof(double));
    if (Y_in) std::memcpy(Y_in, s.Y_in.data(), s.Y_in.size() * sizeof(double));
    return XF_OK;
}

const char *xf_setup_species(const xf_setup *h, int n)
{
    if (!h || n < 0 || n >= (int)h->s.species.size()) return nullptr;
    return h->s.species[n].c_str();
}

int xf_setup_domain(const xf_setup *h, double *dom6)
{
    if (!h || !dom6) return XF_EINVAL;
    const xfh::Setup &s = h->s;
    const double v[6] = {s.Domain_xmin, s.Domain_ymin, s.Domain_zmin, s.Domain_length, s.Domain_width, s.Domain_height};
    std::memcpy(dom6, v, sizeof(v));
    return XF_OK;
}

int xf_setup_initial(const xf_setup *h, const char *workload, unsigned long long seed, double *U_aos, double *T)
{
    if (!h || !workload || !U_aos || !T) return XF_EINVAL;
    std::string m;
    if (!xfh::setup_initial(h->s, workload, seed, U_aos, T, m)) {
        std::fprintf(stderr, "xfluids_amd: %s\n", m.c_str());
        return XF_EINVAL;
    }
    return XF_OK;
}

int xf_setup_context(const xf_setup *h, int precision, int chem_driver, xf_ctx **out)
{
    if (!h || !out) return XF_EINVAL;
    xf_case_desc d = h->s.desc;
    d.precision = precision;
    d.chem_driver = chem_driver;
    int e = xf_create(&d, out);
    if (e) return e;
    if (d.cop) {
        e = xf_set_thermo(*out, h->s.Hia.data(), h->s.Hib.data(), h->s.Wi.data());
        if (!e && d.NR > 0)
            e = xf_set_mechanism(*out, h->s.Nu_f.data(), h->s.Nu_b.data(), h->s.third.data(), h->s.Rargus.data());
        if (!e && d.visc) e = xf_set_transport(*out, h->s.fcv.data(), h->s.fct.data(), h->s.Dkj.data());
    }
    if (e) {
        xf_destroy(*out);
        *out = nullptr;
    }
    return e;
}

int xf_evolve_staged(xf_ctx *ctx, int nsteps, double *dt_out)
{
    if (!ctx || nsteps < 0) return XF_EINVAL;
    xf_case_desc d;
    if (xf_get_desc(ctx, &d)) return XF_EINVAL;
    Fluid fl{ctx};
    for (int s = 0; s < nsteps; s++) {
        double dt = 0.0;
        int e = fl.GetFluidDt(&dt); // ComputeTimeStep (XFLUIDS.cpp:196, 527-544)
        if (e) return e;
        if (dt_out) dt_out[s] = dt;
        if (d.react && d.strang)
            if ((e = xf_reaction(ctx, dt))) return e; // XFLUIDS.cpp:211-216
        for (int flag = 1; flag <= 3; flag++)
            if ((e = rk_stage(fl, flag, dt))) return e; // SinglePhaseSolverRK3rd
        if (d.react)
            if ((e = xf_reaction(ctx, dt))) return e; // XFLUIDS.cpp:220-225
    }
    return XF_OK;
}

namespace {
size_t state_doubles(const xf_case_desc &d)
{
    const size_t E = d.cop ? (size_t)d.NS + 4 : 5;
    return E * (size_t)d.Xmax * (size_t)d.Ymax * (size_t)d.Zmax;
}
} // namespace

int xf_checkpoint_write(const char *path, const xf_case_desc *d, const double *U_aos, int step, double time,
                        float time_consumption)
{
    if (!path || !d || !U_aos) return XF_EINVAL;
    std::FILE *f = std::fopen(path, "wb");
    if (!f) return XF_EINVAL;
    const size_t n = state_doubles(*d);
    bool ok = std::fwrite(&step, sizeof(int), 1, f) == 1 && std::fwrite(&time, sizeof(double), 1, f) == 1 &&
              std::fwrite(&time_consumption, sizeof(float), 1, f) == 1 && std::fwrite(U_aos, sizeof(double), n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? XF_OK : XF_EINVAL;
}

int xf_checkpoint_read(const char *path, const xf_case_desc *d, double *U_aos, int *step, double *time,
                       float *time_consumption)
{
    if (!path || !d || !U_aos) return XF_EINVAL;
    std::FILE *f = std::fopen(path, "rb");
    if (!f) return XF_EINVAL;
    const size_t n = state_doubles(*d);
    int st = 0;
    double tm = 0.0;
    float tc = 0.0f;
    bool ok = std::fread(&st, sizeof(int), 1, f) == 1 && std::fread(&tm, sizeof(double), 1, f) == 1 &&
              std::fread(&tc, sizeof(float), 1, f) == 1 && std::fread(U_aos, sizeof(double), n, f) == n;
    char extra;
    ok = ok && std::fread(&extra, 1, 1, f) == 0; // nothing may follow the state
    std::fclose(f);
    if (!ok) return XF_EINVAL;
    if (step) *step = st;
    if (time) *time = tm;
    if (time_consumption) *time_consumption = tc;
    return XF_OK;
}

int xf_save_checkpoint(xf_ctx *ctx, const char *path, int step, double time, float time_consumption)
{
    xf_case_desc d;
    if (!ctx || xf_get_desc(ctx, &d)) return XF_EINVAL;
    std::vector<double> U(state_doubles(d));
    if (int e = xf_download_state(ctx, 0, U.data(), nullptr)) return e;
    return xf_checkpoint_write(path, &d, U.data(), step, time, time_consumption);
}

int xf_load_checkpoint(xf_ctx *ctx, const char *path, int *step, double *time, float *time_consumption)
{
    xf_case_desc d;
    if (!ctx || xf_get_desc(ctx, &d)) return XF_EINVAL;
    std::vector<double> U(state_doubles(d));
    int st = 0;
    double tm = 0.0;
    if (int e = xf_checkpoint_read(path, &d, U.data(), &st, &tm, time_consumption)) return e;
    if (step) *step = st;
    if (time) *time = tm;
    if (int e = xf_upload_state(ctx, 0, U.data(), nullptr)) return e;
    // the device clock continues from the header (Read_Ubak restores
    // Iteration and physicalTime, src/XFLUIDS.cpp:616-622)
    return xf_set_time(ctx, tm, st);
}

} // extern "C"
