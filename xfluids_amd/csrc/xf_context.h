// Context of one case and the per-species-count kernel dispatch table,
// shared by the C-ABI translation unit (xf_api.hip) and the translation units
// that instantiate the kernels for each compile-time species count
// (xf_ops_ns*.hip), so that those compile in parallel.
#pragma once
#include "xf_sweep_kernels.h"
#include "xf_chem_kernels.h"
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace xf;

namespace xf {

struct EvPair {
    const char *name;
    hipEvent_t a, b;
};

struct Ops {
    // UpdateStates on U; bcs != 0 (packed BC codes) also fills the ghost cells;
    // where: the stage a failure is reported in (1..3 RK stage, 4 Reaction)
    void (*states)(xf_ctx *, double *U, int full, int bcs, int where);
    void (*sweep)(xf_ctx *, int dir, const double *Ug, int mode, int stage, const double *Ua, const double *Ub,
                  double *Uo);
    // check = 1: the post-Reaction NaN check of a time step on each cell
    void (*chem)(xf_ctx *, int record, int check);
    void (*prepare)(xf_ctx *);
    // shared queue over the cells of several cases (ensembles)
    // spec: compiled mechanism whose records the tables hold, or -1
    void (*chem_multi)(hipStream_t, const ChemCase *, int ncases, long long total, const DevThermo *,
                       const DevMech *, unsigned long long *queue, int blocks, int record, int check, int spec,
                       void *spill, int spill_cap, int spill_after);
    // bytes of one hand-over record (ChemSpill<NS>)
    size_t (*spill_bytes)();
    int (*multi_blocks)(long long total, int spec);
    // viscous terms of GetLU: transport coefficients and centre gradients of
    // the state Ug, then the wall flux of one direction (before its sweep)
    void (*visc_prep)(xf_ctx *, const double *Ug);
    void (*visc_flux)(xf_ctx *, int dir, const double *Ug);
    // one launch over every case of an ensemble (case table on the device;
    // sel codes pick U, U1 or U2 per case; the scheme, precision and tile
    // shape are those of the template case c0)
    void (*states_multi)(const xf_ctx *c0, hipStream_t, const EnsCase *, int ncases, long long blocks, int full,
                         int fill, int where, int sel);
    void (*sweep_multi)(const xf_ctx *c0, hipStream_t, const EnsCase *, int ncases, long long blocks, int dir, int shape,
                        int mode, int stage, int sel_g, int sel_a, int sel_b, int sel_o);
    // the sweeps compute the wall flux themselves when Geom::visc_fused is set
    bool visc_in_sweep;
};

// dispatch tables, one translation unit each
Ops ops_ns1();
Ops ops_ns2();
Ops ops_ns3();
Ops ops_ns4();
Ops ops_ns5();
Ops ops_ns9();
Ops ops_ns10();
Ops ops_ns11();

} // namespace xf

struct xf_ctx {
    xf_case_desc d;
    Geom g;
    int E, NCOP;
    long long N, Ninner;
    hipStream_t stream = nullptr;
    Fields f{};
    double *U2 = nullptr; // RK stage-2 output of the fused path
    DevThermo h_th{};
    DevThermo *d_th = nullptr;
    DevMech h_mc{};
    DevMech *d_mc = nullptr;
    DevTransport *d_tr = nullptr;
    bool thermo_set = false, mech_set = false, transport_set = false;
    Ops ops{};
    bool has_ops = false;
    int chem_blocks = 0; // resident one-wave blocks of the work-queue chemistry kernel
    // cells the queue hands to whole waves after spill_after passes (at most
    // spill_cap of them per launch; 0 = never).  Switched off: the one run of
    // the hand-over on the GPU did not finish and no cause has been shown, so
    // xf_set_chem_handover refuses any other value and nothing is allocated.
    void *spill = nullptr;
    int spill_cap = 0, spill_after = 0;
    // compiled mechanism the device tables match (MechSpec<mech_spec>), or -1;
    // spec_allowed = false keeps the generic loop (xf_set_chem_specialization)
    int mech_spec = -1;
    bool spec_allowed = true;
    int shape[3] = {0, 0, 0}; // sweep tile shape per direction (SweepShape)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int last_code = 0, last_id = -1;
    int last_step = 0, last_where = 0; // step (Iteration) and stage of the last failure
    int step_count = 0, target_reached = 0; // device clock as of the last error read
    bool prof = false;
    std::string prof_filter; // ",name,name," or empty for all
    std::vector<EvPair> evs;
    std::map<std::string, std::pair<double, int>> times;
};

#define XF_HIP(call)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) {                                                                                        \
            fprintf(stderr, "xfluids_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__);       \
            return XF_EHIP;                                                                                            \
        }                                                                                                              \
    } while (0)

namespace xf {

inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// tile grid of a sweep in direction dir with tile shape `shape` (TS faces by
// TW = 256 / TS pencils, the shapes of SweepShape): x tiles along the sweep
// axis of x, pencil rows, third axis
inline void sweep_grid(const Geom &g, int dir, int shape, unsigned out[3])
{
    const int TS = dir == 0 ? (shape == 0 ? 64 : 32) : (shape == 0 ? 16 : (shape == 1 ? 32 : 8));
    const int TW = 256 / TS;
    if (dir == 0) {
        out[0] = (unsigned)((g.X_inner + TS - 2) / (TS - 1));
        out[1] = (unsigned)((g.Y_inner * g.Z_inner + TW - 1) / TW);
        out[2] = 1;
    } else if (dir == 1) {
        out[0] = (unsigned)((g.X_inner + TW - 1) / TW);
        out[1] = (unsigned)((g.Y_inner + TS - 2) / (TS - 1));
        out[2] = (unsigned)g.Z_inner;
    } else {
        out[0] = (unsigned)((g.X_inner + TW - 1) / TW);
        out[1] = (unsigned)((g.Z_inner + TS - 2) / (TS - 1));
        out[2] = (unsigned)g.Y_inner;
    }
}

struct ProfScope {
    xf_ctx *c;
    const char *name;
    hipEvent_t a{}, b{};
    bool on;
    ProfScope(xf_ctx *c_, const char *n) : c(c_), name(n)
    {
        on = c->prof && (c->prof_filter.empty() || c->prof_filter.find("," + std::string(n) + ",") != std::string::npos);
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope()
    {
        if (on) {
            (void)hipEventRecord(b, c->stream);
            c->evs.push_back({name, a, b});
        }
    }
};

template <int NS, bool COP> struct Launch {
    static void states(xf_ctx *c, double *U, int full, int bcs, int where)
    {
        ProfScope ps(c, "states");
        hipLaunchKernelGGL((k_states<NS, COP>), dim3(nblk(c->N)), dim3(256), 0, c->stream, c->g, c->d_th, U, c->f,
                           c->d.ncop_gamma, full, bcs, where);
        if constexpr (COP) {
            if (c->g.ghost)
                hipLaunchKernelGGL((k_rhoyi_writeback<NS>), dim3(nblk(c->N)), dim3(256), 0, c->stream, c->g, U,
                                   c->f);
        }
        hipLaunchKernelGGL(k_states_fold, dim3(12), dim3(1024), 0, c->stream, c->f.spart, (int)nblk(c->N), c->f.red,
                           c->f.err);
    }
    template <int DIR, int SCH, int PREC, int SHAPE>
    static void sweep_t(xf_ctx *c, const double *Ug, int mode, int stage, const double *Ua, const double *Ub, double *Uo)
    {
        constexpr int E = COP ? NS + 4 : 5;
        static_assert(SweepShape<DIR, E, SHAPE>::TS == (DIR == 0 ? (SHAPE == 0 ? 64 : 32)
                                                                 : (SHAPE == 0 ? 16 : (SHAPE == 1 ? 32 : 8))),
                      "sweep_grid's tile sizes");
        unsigned gr[3];
        sweep_grid(c->g, DIR, SHAPE, gr);
        const dim3 grid(gr[0], gr[1], gr[2]);
        static const char *names[3] = {"sweep_x", "sweep_y", "sweep_z"};
        ProfScope ps(c, names[DIR]);
        hipLaunchKernelGGL((k_sweep<NS, COP, DIR, SCH, PREC, SHAPE>), grid, dim3(256), 0, c->stream, c->g, c->d_th,
                           Ug, c->f, c->d.ncop_gamma, mode, stage, Ua, Ub, Uo);
    }
    template <int DIR, int SCH, int PREC>
    static void sweep_s_(xf_ctx *c, const double *Ug, int mode, int stage, const double *Ua, const double *Ub, double *Uo)
    {
        switch (c->shape[DIR]) {
        case 1: sweep_t<DIR, SCH, PREC, 1>(c, Ug, mode, stage, Ua, Ub, Uo); break;
        case 2:
            if constexpr (SweepShapes<DIR>::count > 2) {
                sweep_t<DIR, SCH, PREC, 2>(c, Ug, mode, stage, Ua, Ub, Uo);
                break;
            }
            [[fallthrough]];
        default: sweep_t<DIR, SCH, PREC, 0>(c, Ug, mode, stage, Ua, Ub, Uo);
        }
    }
    template <int DIR, int PREC>
    static void sweep_p(xf_ctx *c, const double *Ug, int mode, int stage, const double *Ua, const double *Ub, double *Uo)
    {
        if (c->d.scheme == XF_WENO5Z) sweep_s_<DIR, XF_WENO5Z, PREC>(c, Ug, mode, stage, Ua, Ub, Uo);
        else if (c->d.scheme == XF_WENOCU6) sweep_s_<DIR, XF_WENOCU6, PREC>(c, Ug, mode, stage, Ua, Ub, Uo);
        else sweep_s_<DIR, XF_WENO5JS, PREC>(c, Ug, mode, stage, Ua, Ub, Uo);
    }
    template <int DIR>
    static void sweep_d(xf_ctx *c, const double *Ug, int mode, int stage, const double *Ua, const double *Ub, double *Uo)
    {
        if (c->d.precision == XF_MIXED) sweep_p<DIR, XF_MIXED>(c, Ug, mode, stage, Ua, Ub, Uo);
        else if (c->d.precision == XF_FUSED_FP64) sweep_p<DIR, XF_FUSED_FP64>(c, Ug, mode, stage, Ua, Ub, Uo);
        else sweep_p<DIR, XF_PLAIN_FP64>(c, Ug, mode, stage, Ua, Ub, Uo);
    }
    static void sweep(xf_ctx *c, int dir, const double *Ug, int mode, int stage, const double *Ua, const double *Ub,
                      double *Uo)
    {
        if (dir == 0) sweep_d<0>(c, Ug, mode, stage, Ua, Ub, Uo);
        if (dir == 1) sweep_d<1>(c, Ug, mode, stage, Ua, Ub, Uo);
        if (dir == 2) sweep_d<2>(c, Ug, mode, stage, Ua, Ub, Uo);
    }
    // calls f(std::integral_constant<int, K>) for the compiled mechanism K
    // == spec if one exists for this species count, else with K = -1 (the
    // generic loop over the device tables)
    template <int K, class F> static void spec_switch(int spec, F &f)
    {
        if constexpr (K < XF_NMECH_SPEC) {
            if constexpr (COP && MechSpec<K>::NS == NS) {
                if (spec == K) {
                    f(std::integral_constant<int, K>{});
                    return;
                }
            }
            spec_switch<K + 1>(spec, f);
        } else {
            f(std::integral_constant<int, -1>{});
        }
    }
    static void chem(xf_ctx *c, int record, int check)
    {
        if constexpr (COP) {
            ProfScope ps(c, "chem");
            auto launch = [&](auto K) {
                constexpr int S = decltype(K)::value;
                if (c->d.chem_driver == XF_CHEM_QUEUE) {
                    const int cap = c->spill_after > 0 ? c->spill_cap : 0;
                    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, c->stream, c->f.queue);
                    hipLaunchKernelGGL((k_chem_queue<NS, S>), dim3(c->chem_blocks), dim3(64), 0, c->stream, c->g,
                                       c->d_th, c->d_mc, c->f, record, check, c->spill, cap, c->spill_after);
                    if (cap > 0)
                        hipLaunchKernelGGL((k_chem_coop<NS>), dim3(coop_blocks(cap)), dim3(64), 0, c->stream, c->g,
                                           c->d_th, c->d_mc, c->f, record, check, c->spill, cap);
                } else if (c->d.chem_driver == XF_CHEM_WAVE) {
                    if (c->d.thermo)
                        hipLaunchKernelGGL((k_chem_wave<NS, 1>), dim3((unsigned)c->Ninner), dim3(64), 0, c->stream,
                                           c->g, c->d_th, c->d_mc, c->f, record, check);
                    else
                        hipLaunchKernelGGL((k_chem_wave<NS, 0>), dim3((unsigned)c->Ninner), dim3(64), 0, c->stream,
                                           c->g, c->d_th, c->d_mc, c->f, record, check);
                } else {
                    hipLaunchKernelGGL((k_chem<NS, S>), dim3(nblk(c->Ninner, 64)), dim3(64), 0, c->stream, c->g,
                                       c->d_th, c->d_mc, c->f, record, check);
                }
            };
            spec_switch<0>(c->mech_spec, launch);
        }
    }
    // persistent lanes of the queue driver: as many one-wave blocks as stay
    // resident, never more than the cells need (queried outside any stream
    // capture, again once the mechanism has picked the kernel)
    static void prepare(xf_ctx *c)
    {
        if constexpr (COP) {
            int per_cu = 0, ncu = 0, dev = 0;
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            auto query = [&](auto K) {
                constexpr int S = decltype(K)::value;
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_chem_queue<NS, S>, 64, 0);
            };
            spec_switch<0>(c->mech_spec, query);
            const long long want = (c->Ninner + 63) / 64;
            const long long nb = (long long)std::max(per_cu, 1) * std::max(ncu, 1);
            c->chem_blocks = (int)std::max(1LL, std::min(nb, want));
        }
    }
    static void chem_multi(hipStream_t st, const ChemCase *cases, int ncases, long long total, const DevThermo *th,
                           const DevMech *mc, unsigned long long *queue, int blocks, int record, int check, int spec,
                           void *spill, int spill_cap, int spill_after)
    {
        if constexpr (COP) {
            const int cap = spill_after > 0 ? spill_cap : 0;
            auto launch = [&](auto K) {
                constexpr int S = decltype(K)::value;
                hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, (unsigned int *)queue);
                hipLaunchKernelGGL((k_chem_queue_multi<NS, S>), dim3(blocks), dim3(64), 0, st, cases, ncases, total,
                                   th, mc, queue, record, check, spill, cap, spill_after);
            };
            spec_switch<0>(spec, launch);
            if (cap > 0)
                hipLaunchKernelGGL((k_chem_coop_multi<NS>), dim3(coop_blocks(cap)), dim3(64), 0, st, cases, ncases, th,
                                   mc, queue, record, check, spill, cap);
        }
    }
    // persistent one-wave blocks of the hand-over kernel: a wave per cell
    // while the list lasts, no more than 2048
    static unsigned coop_blocks(int cap) { return (unsigned)std::min(cap, 2048); }
    static size_t spill_bytes() { return COP ? sizeof(ChemSpill<NS>) : 0; }
    static int multi_blocks(long long total, int spec)
    {
        int per_cu = 0, ncu = 0, dev = 0;
        if constexpr (COP) {
            (void)hipGetDevice(&dev);
            (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            auto query = [&](auto K) {
                constexpr int S = decltype(K)::value;
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_chem_queue_multi<NS, S>,
                                                                   64, 0);
            };
            spec_switch<0>(spec, query);
        }
        const long long want = (total + 63) / 64;
        const long long nb = (long long)std::max(per_cu, 1) * std::max(ncu, 1);
        return (int)std::max(1LL, std::min(nb, want));
    }
    static void visc_prep(xf_ctx *c, const double *Ug)
    {
        if constexpr (COP) {
            ProfScope ps(c, "visc_prep");
            const Geom &g = c->g;
            if (c->d.precision == XF_MIXED)
                hipLaunchKernelGGL((k_transport<NS, 1>), dim3(nblk(c->N)), dim3(256), 0, c->stream, g, c->d_th,
                                   c->d_tr, Ug, c->f, c->d.visc_heat, c->d.visc_diffu);
            else
                hipLaunchKernelGGL((k_transport<NS, 0>), dim3(nblk(c->N)), dim3(256), 0, c->stream, g, c->d_th,
                                   c->d_tr, Ug, c->f, c->d.visc_heat, c->d.visc_diffu);
            if (c->d.visc_diffu)
                hipLaunchKernelGGL((k_yreduce<NS>), dim3(2 * NS), dim3(1024), 0, c->stream, c->f.ypart,
                                   (int)nblk(c->N), c->f.red);
            const int i0 = g.DimX ? g.Bx - 2 : 0, nx = g.DimX ? g.X_inner + 4 : 1;
            const int j0 = g.DimY ? g.By - 2 : 0, ny = g.DimY ? g.Y_inner + 4 : 1;
            const int k0 = g.DimZ ? g.Bz - 2 : 0, nz = g.DimZ ? g.Z_inner + 4 : 1;
            const int *bc = c->d.BC;
            // the mixed form stores the gradients (and the transport
            // coefficients) as FP32 planes
            if (c->d.precision == XF_MIXED)
                hipLaunchKernelGGL((k_vel_deriv<XF_MIXED>), dim3(nblk((long long)nx * ny * nz)), dim3(256), 0,
                                   c->stream, g, c->f, i0, nx, j0, ny, k0, nz, bc[0], bc[1], bc[2], bc[3], bc[4],
                                   bc[5]);
            else
                hipLaunchKernelGGL((k_vel_deriv<XF_PLAIN_FP64>), dim3(nblk((long long)nx * ny * nz)), dim3(256), 0,
                                   c->stream, g, c->f, i0, nx, j0, ny, k0, nz, bc[0], bc[1], bc[2], bc[3], bc[4],
                                   bc[5]);
        }
    }
    template <int DIR> static void visc_flux_d(xf_ctx *c, const double *Ug)
    {
        const Geom &g = c->g;
        const long long plane = (long long)(g.X_inner + (DIR == 0)) * (g.Y_inner + (DIR == 1));
        const long long nz = g.Z_inner + (DIR == 2);
        // z: (chunks of 256 faces of a plane) x (face planes), as the kernel maps them
        const unsigned blocks = DIR == 2 ? nblk(plane) * (unsigned)nz : nblk(plane * nz);
        static const char *names[3] = {"visc_x", "visc_y", "visc_z"};
        ProfScope ps(c, names[DIR]);
        if (c->d.precision == XF_MIXED)
            hipLaunchKernelGGL((k_visc_flux<NS, DIR, XF_MIXED>), dim3(blocks), dim3(256), 0, c->stream, g, Ug, c->f);
        else
            hipLaunchKernelGGL((k_visc_flux<NS, DIR, XF_PLAIN_FP64>), dim3(blocks), dim3(256), 0, c->stream, g, Ug,
                               c->f);
    }
    static void visc_flux(xf_ctx *c, int dir, const double *Ug)
    {
        if constexpr (COP) {
            if (dir == 0) visc_flux_d<0>(c, Ug);
            if (dir == 1) visc_flux_d<1>(c, Ug);
            if (dir == 2) visc_flux_d<2>(c, Ug);
        }
    }
    static void states_multi(const xf_ctx *c0, hipStream_t st, const EnsCase *cases, int ncases, long long blocks,
                             int full, int fill, int where, int sel)
    {
        hipLaunchKernelGGL((k_states_multi<NS, COP>), dim3((unsigned)blocks), dim3(256), 0, st, cases, ncases, c0->d_th,
                           c0->d.ncop_gamma, full, fill, where, sel);
        hipLaunchKernelGGL(k_states_fold_multi, dim3((unsigned)(12 * ncases)), dim3(1024), 0, st, cases);
    }
    template <int DIR, int SCH, int PREC, int SHAPE>
    static void sweep_multi_t(const xf_ctx *c0, hipStream_t st, const EnsCase *cases, int ncases, long long blocks,
                              int mode, int stage, int sg, int sa, int sb, int so)
    {
        hipLaunchKernelGGL((k_sweep_multi<NS, COP, DIR, SCH, PREC, SHAPE>), dim3((unsigned)blocks), dim3(256), 0, st,
                           cases, ncases, c0->d_th, c0->d.ncop_gamma, mode, stage, sg, sa, sb, so);
    }
    template <int DIR, int SCH, int PREC>
    static void sweep_multi_s(const xf_ctx *c0, hipStream_t st, const EnsCase *cases, int ncases, long long blocks,
                              int shape, int mode, int stage, int sg, int sa, int sb, int so)
    {
        switch (shape) {
        case 1: sweep_multi_t<DIR, SCH, PREC, 1>(c0, st, cases, ncases, blocks, mode, stage, sg, sa, sb, so); break;
        case 2:
            if constexpr (SweepShapes<DIR>::count > 2) {
                sweep_multi_t<DIR, SCH, PREC, 2>(c0, st, cases, ncases, blocks, mode, stage, sg, sa, sb, so);
                break;
            }
            [[fallthrough]];
        default: sweep_multi_t<DIR, SCH, PREC, 0>(c0, st, cases, ncases, blocks, mode, stage, sg, sa, sb, so);
        }
    }
    template <int DIR, int PREC>
    static void sweep_multi_p(const xf_ctx *c0, hipStream_t st, const EnsCase *cases, int ncases, long long blocks,
                              int shape, int mode, int stage, int sg, int sa, int sb, int so)
    {
        if (c0->d.scheme == XF_WENO5Z)
            sweep_multi_s<DIR, XF_WENO5Z, PREC>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
        else if (c0->d.scheme == XF_WENOCU6)
            sweep_multi_s<DIR, XF_WENOCU6, PREC>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
        else
            sweep_multi_s<DIR, XF_WENO5JS, PREC>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
    }
    template <int DIR>
    static void sweep_multi_d(const xf_ctx *c0, hipStream_t st, const EnsCase *cases, int ncases, long long blocks,
                              int shape, int mode, int stage, int sg, int sa, int sb, int so)
    {
        if (c0->d.precision == XF_MIXED)
            sweep_multi_p<DIR, XF_MIXED>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
        else if (c0->d.precision == XF_FUSED_FP64)
            sweep_multi_p<DIR, XF_FUSED_FP64>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
        else
            sweep_multi_p<DIR, XF_PLAIN_FP64>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
    }
    static void sweep_multi(const xf_ctx *c0, hipStream_t st, const EnsCase *cases, int ncases, long long blocks,
                            int dir, int shape, int mode, int stage, int sg, int sa, int sb, int so)
    {
        if (dir == 0) sweep_multi_d<0>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
        if (dir == 1) sweep_multi_d<1>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
        if (dir == 2) sweep_multi_d<2>(c0, st, cases, ncases, blocks, shape, mode, stage, sg, sa, sb, so);
    }
    static Ops ops()
    {
        return Ops{&states,    &sweep,     &chem,         &prepare,     &chem_multi,         &spill_bytes,
                   &multi_blocks, &visc_prep, &visc_flux, &states_multi, &sweep_multi, visc_in_sweep<NS>()};
    }
};

} // namespace xf
