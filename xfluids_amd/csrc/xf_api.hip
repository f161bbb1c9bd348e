// Host side of the C-ABI (include/xf_api.h): context, device buffers,
// kernel dispatch over compile-time species counts, the stage seam and the
// fused device-resident time-step loop (optionally one hipGraph per step).
#include "xf_context.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <memory>
#include <vector>

using namespace xf;

// The whole-wave hand-over of stiff chemistry cells (k_chem_coop): off.  Its
// one GPU run (reactive channel 24 x 24, 2e-7 s, hand-over after 8 passes)
// did not finish within 120 s, and reading the queue, the hand-over list and
// the whole-wave loop has not shown why; until a cause is found no context
// allocates the list and xf_set_chem_handover accepts only 0.
constexpr bool CHEM_HANDOVER_ON = false;

namespace {

bool select_ops(int NS, int cop, Ops &o)
{
    if (!cop) {
        if (NS != 1) return false;
        o = ops_ns1();
        return true;
    }
    switch (NS) {
    case 2: o = ops_ns2(); return true;
    case 3: o = ops_ns3(); return true;
    case 4: o = ops_ns4(); return true;
    case 5: o = ops_ns5(); return true;
    case 9: o = ops_ns9(); return true;
    case 10: o = ops_ns10(); return true;
    case 11: o = ops_ns11(); return true;
    default: return false;
    }
}

__global__ void k_zero_int(int *p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = (t == 1) ? 0x7fffffff : 0;
}

__global__ void k_debug_math(const double *x, double *ex, double *lg, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        ex[t] = dexp(x[t]);
        lg[t] = dlog(x[t]);
    }
}

// streaming copy of n double2 (16-byte loads and stores, grid-stride)
__global__ void __launch_bounds__(256) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) dst[t] = src[t];
}

// FP64 issue-rate probe: 8 independent fused multiply-add chains per lane,
// iters x 8 fmas each, one result per lane written so nothing is dropped
__global__ void __launch_bounds__(256) k_fma_probe(double *out, int iters, double a, double b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double x[8];
#pragma unroll
    for (int q = 0; q < 8; q++) x[q] = 1e-3 * (double)(t + q);
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int q = 0; q < 8; q++) x[q] = fma(x[q], a, b);
    }
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) sum += x[q];
    out[t] = sum;
}

__global__ void k_set_double(double *p, double v)
{
    if (threadIdx.x == 0) p[0] = v;
}

int check_launch()
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "xfluids_amd: launch error %s\n", hipGetErrorString(e));
        return XF_EHIP;
    }
    return 0;
}

// clear the error words (bits, cell ids, step, stage, halt); the step
// counter and the target-reached flag stay
__global__ void k_err_reset(int *err)
{
    const int t = threadIdx.x;
    if (t < ERR_COUNT) err[t] = (t >= ERR_ID && t < ERR_ID + 4) ? 0x7fffffff : 0;
}

// the clock: op 0 sets the time and the step count, op 1 the target time
// (and clears the target-reached flag)
__global__ void k_clock(double *dt, int *err, int op, double t, int step)
{
    if (threadIdx.x != 0) return;
    if (op == 0) {
        dt[DT_TIME] = t;
        err[ERR_COUNT] = step;
    } else {
        dt[DT_TARGET] = t;
        err[ERR_DONE] = 0;
        err[ERR_HALT] = 0;
    }
}

// read and clear the device error words; returns an XF_ERR code or 0.  The
// pass that failed first halted the step, so every bit comes from that one
// pass; a pass that finds both bad densities/mass fractions and bad
// primitives reports code 1, as the reference checks those first
// (UpdateStates_block.cpp:58-115).  The cell is the lowest id with the
// reported code's failure.
int take_error(xf_ctx *c)
{
    int h[ERR_SIZE];
    if (hipMemcpyAsync(h, c->f.err, sizeof(h), hipMemcpyDeviceToHost, c->stream) != hipSuccess) return XF_EHIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return XF_EHIP;
    c->step_count = h[ERR_COUNT];
    c->target_reached = h[ERR_DONE];
    if (h[ERR_BITS] == 0 && h[ERR_HALT] == 0) return 0;
    hipLaunchKernelGGL(k_err_reset, dim3(1), dim3(64), 0, c->stream, c->f.err);
    if (h[ERR_BITS] == 0) return 0; // halted at the target time
    const int b = h[ERR_BITS];
    const int code = (b & EBIT_RHO_Y) ? XF_ERR_RHO_Y : (b & EBIT_PRIM) ? XF_ERR_PRIM : (b & EBIT_CHEM) ? XF_ERR_CHEM : XF_ERR_NAN;
    const int slot = code == XF_ERR_RHO_Y ? 0 : (code == XF_ERR_PRIM ? 1 : (code == XF_ERR_NAN ? 2 : 3));
    c->last_code = code;
    c->last_id = h[ERR_ID + slot];
    c->last_step = h[ERR_STEP];
    c->last_where = h[ERR_WHERE];
    return code;
}

// FluidBoundaryCondition: all ghost cells in one launch; reset_red also
// zeroes the per-pass maxima for the UpdateStates pass that follows
void boundary_async(xf_ctx *c, double *U, int reset_red)
{
    ProfScope ps(c, "boundary");
    const Geom &g = c->g;
    const long long n = (g.DimZ ? 2LL * g.Bz * g.Ymax * g.Xmax : 0) + (g.DimY ? 2LL * g.By * g.Xmax * g.Z_inner : 0) +
                        (g.DimX ? 2LL * g.Bx * g.Y_inner * g.Z_inner : 0);
    hipLaunchKernelGGL(k_boundary_all, dim3(std::max(1u, nblk(n))), dim3(256), 0, c->stream, g, U, c->E, c->d.BC[0],
                       c->d.BC[1], c->d.BC[2], c->d.BC[3], c->d.BC[4], c->d.BC[5], c->f.red, reset_red);
}

// UpdateStates (the pass's GLF and CFL maxima replace the stored ones);
// fill = 1 also runs FluidBoundaryCondition first, fused into the same pass;
// where: the stage a failure is reported in (1..3 RK stage, 4 Reaction)
void states_async(xf_ctx *c, double *U, int full, int fill, int where)
{
    int bcs = 0;
    if (fill) {
        for (int q = 0; q < 6; q++) bcs |= (c->d.BC[q] & 15) << (4 * q);
        bcs |= 1 << 28;
    }
    c->ops.states(c, U, full, bcs, where);
}

// GetLU: sweeps in x, y, z order (each folds its direction's GLF maxima).
// rk_stage > 0 fuses the SSP-RK3 stage into the last sweep (Ua, Ub inputs,
// Uo output).
void lu_async(xf_ctx *c, const double *Ug, int rk_stage, const double *Ua, const double *Ub, double *Uo)
{
    int dirs[3], nd = 0;
    if (c->g.DimX) dirs[nd++] = 0;
    if (c->g.DimY) dirs[nd++] = 1;
    if (c->g.DimZ) dirs[nd++] = 2;
    // viscous terms (ConVenction_block.hpp:424-575): transport coefficients
    // and centre gradients once; each sweep subtracts its direction's wall
    // flux from the face flux, computing it itself, or (visc_fused 0) from
    // the planes a wall-flux launch right before it writes
    if (c->d.visc) c->ops.visc_prep(c, Ug);
    for (int q = 0; q < nd; q++) {
        int mode = (q > 0) ? 1 : 0;
        if (rk_stage > 0 && q == nd - 1) mode |= 2;
        if (c->d.visc && !(c->g.visc_fused && c->ops.visc_in_sweep)) c->ops.visc_flux(c, dirs[q], Ug);
        c->ops.sweep(c, dirs[q], Ug, mode, rk_stage, Ua, Ub, Uo);
    }
}

void reaction_prep_async(xf_ctx *c);

// XFLUIDS::Reaction before its chemistry launch: UpdateStates(U) and the
// device-side chemistry step (dt/2 for Strang)
void reaction_prep_async(xf_ctx *c)
{
    states_async(c, c->f.U, 0, 0, 4);
    hipLaunchKernelGGL(k_dt_chem, dim3(1), dim3(64), 0, c->stream, c->f.dt, c->d.strang, c->f.err);
}

// A step is: dt; [Strang: Reaction]; three RK stages; [Reaction].  The parts
// between chemistry launches are step_pre and step_mid, so that an ensemble
// can launch one chemistry kernel for all its cases at each Reaction.
void step_pre_async(xf_ctx *c)
{
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, c->stream, c->g, c->f.red, c->f.dt, c->f.err, 1);
    if (c->d.react && c->d.strang) reaction_prep_async(c);
}

void step_mid_async(xf_ctx *c)
{
    Fields &f = c->f;
    // stage 1: U -> U1 ; stage 2: (U, U1) -> U2 ; stage 3: (U, U2) -> U.
    // At stage 1 the sweep's Ub is U2, the previous step's second-stage
    // result, which the reference's stage-1 NaN check reads as its U1.
    double *in[3] = {f.U, f.U1, c->U2};
    for (int s = 1; s <= 3; s++) {
        double *Ug = in[s - 1];
        states_async(c, Ug, 0, 1, s); // ghost fill + UpdateStates
        double *Uo = (s == 1) ? f.U1 : (s == 2 ? c->U2 : f.U);
        lu_async(c, Ug, s, f.U, (s == 1) ? c->U2 : Ug, Uo);
    }
    if (c->d.react) reaction_prep_async(c);
}

void step_async(xf_ctx *c)
{
    step_pre_async(c);
    if (c->d.react && c->d.strang) c->ops.chem(c, 0, 1);
    step_mid_async(c);
    if (c->d.react) c->ops.chem(c, 0, 1);
}

} // namespace

extern "C" {

const char *xf_error_string(int s)
{
    switch (s) {
    case XF_OK: return "ok";
    case XF_ERR_RHO_Y: return "negative/NaN density or NaN mass fraction";
    case XF_ERR_PRIM: return "negative/NaN density, pressure or temperature";
    case XF_ERR_NAN: return "NaN/Inf in conserved variables or flux divergence";
    case XF_ERR_CHEM: return "chemistry substep bound reached";
    case XF_EINVAL: return "invalid argument";
    case XF_EHIP: return "HIP runtime error";
    case XF_ENOKERNEL: return "no kernel instance for this configuration";
    default: return "unknown";
    }
}

int xf_create(const xf_case_desc *desc, xf_ctx **out)
{
    if (!desc || !out) return XF_EINVAL;
    *out = nullptr;
    const xf_case_desc &d = *desc;
    if (d.NS < 1 || d.NS > XF_MAXNS || d.NR < 0 || d.NR > XF_MAXNR) return XF_EINVAL;
    if (d.Xmax <= 0 || d.Ymax <= 0 || d.Zmax <= 0) return XF_EINVAL;
    if ((d.DimX && d.Bwidth_X < 3) || (d.DimY && d.Bwidth_Y < 3) || (d.DimZ && d.Bwidth_Z < 3)) return XF_EINVAL;
    if (d.Xmax != d.X_inner + 2 * d.Bwidth_X || d.Ymax != d.Y_inner + 2 * d.Bwidth_Y ||
        d.Zmax != d.Z_inner + 2 * d.Bwidth_Z)
        return XF_EINVAL;
    if (d.precision != XF_PLAIN_FP64 && d.precision != XF_MIXED && d.precision != XF_FUSED_FP64) return XF_EINVAL;
    if (d.chem_driver != XF_CHEM_PER_CELL && d.chem_driver != XF_CHEM_QUEUE && d.chem_driver != XF_CHEM_WAVE)
        return XF_EINVAL;
    if (d.scheme != XF_WENO5JS && d.scheme != XF_WENO5Z && d.scheme != XF_WENOCU6) return XF_EINVAL;
    // the viscous terms need the species transport data of a mixture, and
    // their centre derivatives reach 4 cells into the ghost layers
    if (d.visc && !d.cop) return XF_EINVAL;
    if (d.visc && ((d.DimX && d.Bwidth_X < 4) || (d.DimY && d.Bwidth_Y < 4) || (d.DimZ && d.Bwidth_Z < 4)))
        return XF_EINVAL;
    Ops ops;
    if (!select_ops(d.NS, d.cop, ops)) return XF_ENOKERNEL;
    xf_ctx *c = new xf_ctx();
    c->d = d;
    c->ops = ops;
    c->has_ops = true;
    c->E = d.cop ? d.NS + 4 : 5;
    c->NCOP = d.cop ? d.NS - 1 : 0;
    c->N = (long long)d.Xmax * d.Ymax * d.Zmax;
    c->Ninner = (long long)d.X_inner * d.Y_inner * d.Z_inner;
    // one 64-lane block per cell: the grid must stay below 2^32 work-items
    if (d.chem_driver == XF_CHEM_WAVE && c->Ninner * 64 > 0xffffffffLL) {
        delete c;
        return XF_EINVAL;
    }
    Geom &g = c->g;
    g.Xmax = d.Xmax, g.Ymax = d.Ymax, g.Zmax = d.Zmax;
    g.X_inner = d.X_inner, g.Y_inner = d.Y_inner, g.Z_inner = d.Z_inner;
    g.Bx = d.Bwidth_X, g.By = d.Bwidth_Y, g.Bz = d.Bwidth_Z;
    g.DimX = d.DimX, g.DimY = d.DimY, g.DimZ = d.DimZ;
    g._dx = d._dx, g._dy = d._dy, g._dz = d._dz, g.CFL = d.CFLnumber;
    g.dx = d.dx, g.dy = d.dy, g.dz = d.dz;
    g.positivity = d.positivity ? 1 : 0;
    g.visc = d.visc ? 1 : 0;
    g.ghost = d.ghost ? 1 : 0;
    g.visc_heat = d.visc_heat ? 1 : 0;
    g.visc_diffu = d.visc_diffu ? 1 : 0;
    g.visc_fused = 1;
    g.Yil_limiter = d.Yil_limiter;
    g.Dim_limiter = d.Dim_limiter;
    g.N = c->N;
    auto fail = [&](int code) {
        xf_destroy(c);
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(XF_EHIP);
    const size_t nE = (size_t)c->E * c->N * sizeof(double), n1 = (size_t)c->N * sizeof(double);
    Fields &f = c->f;
    double **dbl[] = {&f.rho, &f.u, &f.v, &f.w, &f.p, &f.H, &f.T, &f.c, &f.gamma, &f.e, &f.Cp, &f.gcv, &f.R};
    if (hipMalloc(&f.U, nE) || hipMalloc(&f.U1, nE) || hipMalloc(&f.LU, nE) || hipMalloc(&c->U2, nE)) return fail(XF_EHIP);
    for (auto pp : dbl)
        if (hipMalloc(pp, n1)) return fail(XF_EHIP);
    if (hipMalloc(&f.hi, (size_t)d.NS * n1) || hipMalloc(&f.red, RED_SIZE * sizeof(double)) ||
        hipMalloc(&f.dt, DT_SIZE * sizeof(double)) || hipMalloc(&f.err, ERR_SIZE * sizeof(int)) ||
        hipMalloc(&f.substeps, (size_t)c->Ninner * sizeof(int)) || hipMalloc(&f.queue, 4 * sizeof(unsigned)) ||
        hipMalloc(&c->d_th, sizeof(DevThermo)) || hipMalloc(&c->d_mc, sizeof(DevMech)) ||
        hipMalloc(&f.spart, (size_t)nblk(c->N) * 12 * sizeof(double)))
        return fail(XF_EHIP);
    // zero everything once (eigen_block starts at 0, src/Fluids.cpp:427);
    // plain memsets at creation, never inside a captured step
    if (hipMemset(f.U, 0, nE) || hipMemset(f.U1, 0, nE) || hipMemset(f.LU, 0, nE) || hipMemset(c->U2, 0, nE))
        return fail(XF_EHIP);
    for (auto pp : dbl)
        if (hipMemset(*pp, 0, n1)) return fail(XF_EHIP);
    if (hipMemset(f.hi, 0, (size_t)d.NS * n1) || hipMemset(f.red, 0, RED_SIZE * sizeof(double)) ||
        hipMemset(f.substeps, 0, (size_t)c->Ninner * sizeof(int)))
        return fail(XF_EHIP);
    if (hipDeviceSynchronize()) return fail(XF_EHIP);
    int e0[ERR_SIZE] = {0};
    for (int q = 0; q < 4; q++) e0[ERR_ID + q] = 0x7fffffff;
    double t0[DT_SIZE] = {0.0};
    t0[DT_TARGET] = HUGE_VAL; // no output time to land on
    if (hipMemcpy(f.err, e0, sizeof(e0), hipMemcpyHostToDevice) || hipMemcpy(f.dt, t0, sizeof(t0), hipMemcpyHostToDevice))
        return fail(XF_EHIP);
    c->h_th.thermo = d.thermo;
    c->h_mc.NR = d.NR;
    if (hipMemcpy(c->d_th, &c->h_th, sizeof(DevThermo), hipMemcpyHostToDevice) ||
        hipMemcpy(c->d_mc, &c->h_mc, sizeof(DevMech), hipMemcpyHostToDevice))
        return fail(XF_EHIP);
    unsigned qz[4] = {0, 0, 0, 0};
    if (hipMemcpy(f.queue, qz, sizeof(qz), hipMemcpyHostToDevice)) return fail(XF_EHIP);
    // GhostSpecies: the renormalised mass fractions of UpdateStates
    if (d.ghost) {
        if (!d.cop || d.NS < 2) return fail(XF_EINVAL);
        if (hipMalloc(&f.y, (size_t)(d.NS - 1) * n1) || hipMemset(f.y, 0, (size_t)(d.NS - 1) * n1))
            return fail(XF_EHIP);
    }
    // the queue chemistry's hand-over list (cells that go on as whole waves)
    if (CHEM_HANDOVER_ON && d.cop && d.react && c->ops.spill_bytes() > 0) {
        c->spill_cap = (int)std::min<long long>(c->Ninner, 65536);
        if (hipMalloc(&c->spill, (size_t)c->spill_cap * c->ops.spill_bytes())) return fail(XF_EHIP);
    }
    if (d.visc) {
        const size_t nv[6] = {n1, n1, (size_t)d.NS * n1, 9 * n1, nE,
                              (size_t)nblk(c->N) * 2 * d.NS * sizeof(double)};
        double **pv[6] = {&f.mu, &f.kap, &f.Dkm, &f.Vde, &f.Fv, &f.ypart};
        for (int q = 0; q < 6; q++)
            if (hipMalloc(pv[q], nv[q]) || hipMemset(*pv[q], 0, nv[q])) return fail(XF_EHIP);
        if (hipMalloc(&c->d_tr, sizeof(DevTransport))) return fail(XF_EHIP);
        if (hipDeviceSynchronize()) return fail(XF_EHIP);
    }
    c->ops.prepare(c);
    *out = c;
    return XF_OK;
}

void xf_destroy(xf_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    for (auto &e : c->evs) (void)hipEventDestroy(e.a), (void)hipEventDestroy(e.b);
    Fields &f = c->f;
    void *ptrs[] = {f.U,   f.U1,    f.LU,       c->U2,   f.rho,   f.u,     f.v,   f.w,   f.p,   f.H,
                    f.T,   f.c,     f.gamma,    f.e,     f.Cp,    f.gcv,   f.R,   f.hi,  f.red, f.dt,
                    f.err, f.substeps, f.queue, c->d_th, c->d_mc, c->d_tr, f.mu,  f.kap, f.Dkm, f.Vde,
                    f.Fv,  f.ypart, f.spart, c->spill, f.y};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int xf_num_eqn(const xf_ctx *c) { return c ? c->E : 0; }
int xf_get_desc(const xf_ctx *c, xf_case_desc *out)
{
    if (!c || !out) return XF_EINVAL;
    *out = c->d;
    return XF_OK;
}
size_t xf_num_cells(const xf_ctx *c) { return c ? (size_t)c->N : 0; }
void *xf_stream(xf_ctx *c) { return c ? (void *)c->stream : nullptr; }

int xf_set_thermo(xf_ctx *c, const double *Hia, const double *Hib, const double *Wi)
{
    if (!c || !Hia || !Hib || !Wi) return XF_EINVAL;
    const int NS = c->d.NS;
    std::memcpy(c->h_th.Hia, Hia, sizeof(double) * NS * 21);
    std::memcpy(c->h_th.Hib, Hib, sizeof(double) * NS * 6);
    for (int r = 0; r < 3; r++)
        for (int n = 0; n < NS; n++) {
            double *h = c->h_th.Hc + (r * MAXNS + n) * 9;
            for (int m = 0; m < 7; m++) h[m] = Hia[n * 21 + m * 3 + r];
            for (int m = 0; m < 2; m++) h[7 + m] = Hib[n * 6 + m * 3 + r];
        }
    for (int n = 0; n < NS; n++) {
        // src/read_ini/src/thermal.cpp:158-161
        c->h_th.Wi[n] = Wi[n];
        c->h_th._Wi[n] = 1.0 / Wi[n];
        c->h_th.Ri[n] = Ru / Wi[n];
    }
    c->h_th.thermo = c->d.thermo;
    XF_HIP(hipMemcpy(c->d_th, &c->h_th, sizeof(DevThermo), hipMemcpyHostToDevice));
    c->thermo_set = true;
    return XF_OK;
}

int xf_set_transport(xf_ctx *c, const double *fcv, const double *fct, const double *Dkj)
{
    if (!c || !fcv || !fct || !Dkj || !c->d.visc || !c->d_tr || !c->thermo_set) return XF_EINVAL;
    const int NS = c->d.NS;
    DevTransport h{};
    for (int k = 0; k < NS; k++)
        for (int q = 0; q < 4; q++) h.fcv[4 * k + q] = fcv[4 * k + q], h.fct[4 * k + q] = fct[4 * k + q];
    for (int k = 0; k < NS; k++)
        for (int j = 0; j < NS; j++)
            for (int q = 0; q < 4; q++) {
                // the kernel evaluates each species pair once for both
                // orders, as the symmetric table Setup builds allows
                if (Dkj[(k * NS + j) * 4 + q] != Dkj[(j * NS + k) * 4 + q]) return XF_EINVAL;
                h.Dkj[(k * MAXNS + j) * 4 + q] = Dkj[(k * NS + j) * 4 + q];
            }
    // PHI's species-pair factors (Visc_device.h:27-35): pow(W_j/W_k, 1/4) and
    // pow(1 + W_k/W_j, -1/2) as IEEE square roots
    for (int k = 0; k < NS; k++)
        for (int j = 0; j < NS; j++) {
            const double Wk = c->h_th.Wi[k], Wj = c->h_th.Wi[j];
            h.phiA[k * MAXNS + j] = std::sqrt(std::sqrt(Wj / Wk));
            h.phiB[k * MAXNS + j] = 1.0 / std::sqrt(1.0 + Wk / Wj);
            // the mixed form's FP32 factors, phiB with the constant 0.5 / sqrt(2)
            h.phiAf[k * MAXNS + j] = (float)h.phiA[k * MAXNS + j];
            h.phiBcf[k * MAXNS + j] = (float)(h.phiB[k * MAXNS + j] * (0.5 / std::sqrt(2.0)));
        }
    XF_HIP(hipMemcpy(c->d_tr, &h, sizeof(DevTransport), hipMemcpyHostToDevice));
    c->transport_set = true;
    return XF_OK;
}

// The reference's reaction tables plus the per-reaction records of the rate
// kernel, from the host arrays (reaction.cpp:6-58, 170-172)
static int build_mech(DevMech &m, int NS, int NR, const int *Nu_f, const int *Nu_b, const double *third,
                      const double *Rargus)
{
    m = DevMech{};
    m.NR = NR;
    for (int r = 0; r < NR; r++) {
        double sum = 0.0;
        for (int j = 0; j < NS; j++) {
            m.Nu_f[r * NS + j] = Nu_f[r * NS + j];
            m.Nu_b[r * NS + j] = Nu_b[r * NS + j];
            m.Nu_d[r * NS + j] = Nu_b[r * NS + j] - Nu_f[r * NS + j]; // reaction.cpp:38
            m.third[r * NS + j] = third[r * NS + j];
            sum += third[r * NS + j];
        }
        m.third_ind[r] = (sum > 0.0) ? 1 : 0; // reaction.cpp:170-172
        for (int q = 0; q < 6; q++) m.Rargus[r * 6 + q] = Rargus[r * 6 + q];
        DevRxn &x = m.rx[r];
        x = DevRxn{};
        x.A = Rargus[r * 6 + 0];
        x.B = Rargus[r * 6 + 1];
        x.EaR = Rargus[r * 6 + 2] * 4.184 / Ru;
        x.third_ind = m.third_ind[r];
        for (int j = 0; j < NS; j++) {
            const int nf = m.Nu_f[r * NS + j], nb = m.Nu_b[r * NS + j], nd = m.Nu_d[r * NS + j];
            if (nf < 0 || nf > 15 || nb < 0 || nb > 15 || nd < -8 || nd > 7) return XF_EINVAL;
            if (nd != 0) x.md |= 1u << j, x.pd |= (unsigned long long)(nd + 8) << (4 * j);
            if (m.third[r * NS + j] != 0.0) x.mt |= 1u << j;
            if (nf != 0) x.mf |= 1u << j, x.pf |= (unsigned long long)nf << (4 * j);
            if (nb != 0) x.mb |= 1u << j, x.pb |= (unsigned long long)nb << (4 * j);
        }
    }
    return XF_OK;
}

int xf_set_mechanism(xf_ctx *c, const int *Nu_f, const int *Nu_b, const double *third, const double *Rargus)
{
    if (!c) return XF_EINVAL;
    const int NS = c->d.NS, NR = c->d.NR;
    if (NR > 0 && (!Nu_f || !Nu_b || !third || !Rargus)) return XF_EINVAL;
    const int rc = build_mech(c->h_mc, NS, NR, Nu_f, Nu_b, third, Rargus);
    if (rc != XF_OK) return rc;
    XF_HIP(hipMemcpy(c->d_mc, &c->h_mc, sizeof(DevMech), hipMemcpyHostToDevice));
    c->mech_set = true;
    // a compiled specialisation runs only on exactly its own records
    c->mech_spec = c->spec_allowed ? mech_spec_find(NS, NR, c->h_mc.rx, c->h_mc.third) : -1;
    c->ops.prepare(c);
    // a captured step graph may hold the previous chemistry kernel
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec), c->gexec = nullptr;
    if (c->graph) (void)hipGraphDestroy(c->graph), c->graph = nullptr;
    return XF_OK;
}

int xf_mech_specialization_index(int NS, int NR, const int *Nu_f, const int *Nu_b, const double *third,
                                 const double *Rargus)
{
    if (NS < 1 || NS > MAXNS || NR < 1 || NR > MAXNR || !Nu_f || !Nu_b || !third || !Rargus) return -1;
    // host memory only: the record build does not touch the GPU
    std::unique_ptr<DevMech> m(new DevMech());
    if (build_mech(*m, NS, NR, Nu_f, Nu_b, third, Rargus) != XF_OK) return -1;
    return mech_spec_find(NS, NR, m->rx, m->third);
}

int xf_set_chem_specialization(xf_ctx *c, int on)
{
    if (!c) return XF_EINVAL;
    c->spec_allowed = on != 0;
    c->mech_spec = (c->spec_allowed && c->mech_set) ? mech_spec_find(c->d.NS, c->d.NR, c->h_mc.rx, c->h_mc.third) : -1;
    c->ops.prepare(c);
    // a captured step graph holds the previous chemistry kernel
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec), c->gexec = nullptr;
    if (c->graph) (void)hipGraphDestroy(c->graph), c->graph = nullptr;
    return XF_OK;
}

int xf_chem_specialization(const xf_ctx *c) { return c ? c->mech_spec : -1; }

int xf_set_chem_handover(xf_ctx *c, int passes)
{
    if (!c || passes < 0) return XF_EINVAL;
    if (passes > 0 && !CHEM_HANDOVER_ON) {
        fprintf(stderr, "xfluids_amd: the whole-wave chemistry hand-over is switched off in this build (its GPU "
                        "run did not finish and the cause has not been found); only passes = 0 is accepted\n");
        return XF_EINVAL;
    }
    if (c->spill_after != passes) {
        c->spill_after = passes;
        // a captured step graph holds the old launch arguments
        if (c->gexec) (void)hipGraphExecDestroy(c->gexec), c->gexec = nullptr;
        if (c->graph) (void)hipGraphDestroy(c->graph), c->graph = nullptr;
    }
    return XF_OK;
}

static double *which_ptr(xf_ctx *c, int which)
{
    return which == 0 ? c->f.U : (which == 1 ? c->f.U1 : c->U2);
}

int xf_upload_state(xf_ctx *c, int which, const double *U_aos, const double *T)
{
    if (!c || which < 0 || which > 2) return XF_EINVAL;
    const long long N = c->N;
    const int E = c->E;
    if (U_aos) {
        std::vector<double> soa((size_t)E * N);
        for (long long id = 0; id < N; id++)
            for (int n = 0; n < E; n++) soa[(size_t)n * N + id] = U_aos[(size_t)E * id + n];
        XF_HIP(hipMemcpy(which_ptr(c, which), soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
        if (which == 0) {
            // initial state: U1 = U as the reference's initial-condition kernels
            // set it (e.g. 2D-EulerVortex/ini_sample.hpp:84-88); ghost cells of
            // an Inflow boundary keep these values for the whole run
            XF_HIP(hipMemcpy(c->f.U1, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
            XF_HIP(hipMemcpy(c->U2, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    if (T) XF_HIP(hipMemcpy(c->f.T, T, (size_t)N * sizeof(double), hipMemcpyHostToDevice));
    return XF_OK;
}

int xf_download_state(xf_ctx *c, int which, double *U_aos, double *T)
{
    if (!c || which < 0 || which > 2) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(c->stream));
    const long long N = c->N;
    const int E = c->E;
    if (U_aos) {
        std::vector<double> soa((size_t)E * N);
        XF_HIP(hipMemcpy(soa.data(), which_ptr(c, which), soa.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (long long id = 0; id < N; id++)
            for (int n = 0; n < E; n++) U_aos[(size_t)E * id + n] = soa[(size_t)n * N + id];
    }
    if (T) XF_HIP(hipMemcpy(T, c->f.T, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
    return XF_OK;
}

int xf_download_prims(xf_ctx *c, double *rho, double *p, double *u, double *v, double *w, double *H, double *cc,
                      double *gamma, double *e, double *y_aos)
{
    if (!c) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(c->stream));
    const size_t n1 = (size_t)c->N * sizeof(double);
    Fields &f = c->f;
    double *dst[] = {rho, p, u, v, w, H, cc, gamma, e};
    double *src[] = {f.rho, f.p, f.u, f.v, f.w, f.H, f.c, f.gamma, f.e};
    for (int q = 0; q < 9; q++)
        if (dst[q]) XF_HIP(hipMemcpy(dst[q], src[q], n1, hipMemcpyDeviceToHost));
    if (y_aos && c->d.ghost) {
        // GhostSpecies: the renormalised fractions UpdateStates stored, the
        // ghost's 0 (Update_device.hpp:15-21)
        const int NS = c->d.NS;
        const long long N = c->N;
        std::vector<double> pl((size_t)(NS - 1) * N);
        XF_HIP(hipMemcpy(pl.data(), f.y, pl.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (long long id = 0; id < N; id++) {
            double *y = y_aos + (size_t)NS * id;
            for (int k = 0; k < NS - 1; k++) y[k] = pl[(size_t)k * N + id];
            y[NS - 1] = 0.0;
        }
    } else if (y_aos) {
        // Y of every padded cell from U, exactly as Getrhoyi (Update_device.hpp:7-29)
        const int E = c->E, NS = c->d.NS;
        const long long N = c->N;
        std::vector<double> soa((size_t)E * N);
        XF_HIP(hipMemcpy(soa.data(), f.U, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (long long id = 0; id < N; id++) {
            if (!c->d.cop) {
                y_aos[id] = 1.0;
                continue;
            }
            const double rho1 = 1.0 / soa[id];
            double *y = y_aos + (size_t)NS * id;
            y[NS - 1] = 1.0;
            for (int ii = 5; ii < E; ii++) y[ii - 5] = soa[(size_t)ii * N + id] * rho1, y[NS - 1] += -y[ii - 5];
        }
    }
    return XF_OK;
}

int xf_download_lu(xf_ctx *c, double *LU_aos)
{
    if (!c || !LU_aos) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(c->stream));
    const long long N = c->N;
    const int E = c->E;
    std::vector<double> soa((size_t)E * N);
    XF_HIP(hipMemcpy(soa.data(), c->f.LU, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long long id = 0; id < N; id++)
        for (int n = 0; n < E; n++) LU_aos[(size_t)E * id + n] = soa[(size_t)n * N + id];
    return XF_OK;
}

int xf_download_maxima(xf_ctx *c, double *eigen_block, double *uvw_c_max)
{
    if (!c) return XF_EINVAL;
    double red[RED_SIZE];
    XF_HIP(hipStreamSynchronize(c->stream));
    XF_HIP(hipMemcpy(red, c->f.red, sizeof(red), hipMemcpyDeviceToHost));
    const int E = c->E;
    if (eigen_block)
        for (int d = 0; d < 3; d++)
            for (int n = 0; n < XF_MAXNS + 4; n++)
                eigen_block[d * (XF_MAXNS + 4) + n] =
                    n < E ? red[RED_GLF + 3 * d + (n == 0 ? 0 : (n == E - 1 ? 2 : 1))] : 0.0;
    if (uvw_c_max)
        for (int d = 0; d < 3; d++) uvw_c_max[d] = red[RED_DT + d];
    return XF_OK;
}

int xf_boundary(xf_ctx *c, int which)
{
    if (!c || which < 0 || which > 2) return XF_EINVAL;
    boundary_async(c, which_ptr(c, which), 0);
    if (int e = check_launch()) return e;
    XF_HIP(hipStreamSynchronize(c->stream));
    return XF_OK;
}

int xf_update_states(xf_ctx *c, int which)
{
    if (!c || which < 0 || which > 2) return XF_EINVAL;
    if (c->d.cop && !c->thermo_set) return XF_EINVAL;
    states_async(c, which_ptr(c, which), 1, 0, 0);
    if (int e = check_launch()) return e;
    return take_error(c);
}

int xf_compute_lu(xf_ctx *c, int which)
{
    if (!c || which < 0 || which > 2) return XF_EINVAL;
    if (c->d.visc && !c->transport_set) return XF_EINVAL;
    lu_async(c, which_ptr(c, which), 0, nullptr, nullptr, nullptr);
    if (int e = check_launch()) return e;
    return take_error(c);
}

int xf_rk3_stage(xf_ctx *c, int stage, double dt)
{
    if (!c || stage < 1 || stage > 3) return XF_EINVAL;
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, c->stream, c->f.dt + DT_STAGED, dt);
    {
        ProfScope ps(c, "rk3");
        hipLaunchKernelGGL(k_rk3, dim3(nblk(c->Ninner)), dim3(256), 0, c->stream, c->g, c->E, c->f.U, c->f.U1, c->f.LU,
                           c->f.dt + DT_STAGED, stage);
    }
    if (int e = check_launch()) return e;
    XF_HIP(hipStreamSynchronize(c->stream));
    return XF_OK;
}

int xf_get_dt(xf_ctx *c, double *dt_out)
{
    if (!c || !dt_out) return XF_EINVAL;
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(64), 0, c->stream, c->g, c->f.red, c->f.dt, c->f.err, 0);
    if (int e = check_launch()) return e;
    XF_HIP(hipMemcpyAsync(dt_out, c->f.dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    return XF_OK;
}

int xf_chemistry(xf_ctx *c, double dt_chem, int *substeps)
{
    if (!c || !c->d.cop || c->d.NR <= 0) return XF_EINVAL;
    if (!c->thermo_set || !c->mech_set) return XF_EINVAL;
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, c->stream, c->f.dt + DT_CHEM, dt_chem);
    c->ops.chem(c, substeps ? 1 : 0, 0);
    if (int e = check_launch()) return e;
    if (substeps)
        XF_HIP(hipMemcpyAsync(substeps, c->f.substeps, (size_t)c->Ninner * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    return take_error(c);
}

int xf_estimate_nan(xf_ctx *c, int flag)
{
    if (!c || flag < 1 || flag > 4) return XF_EINVAL;
    const double *UI = (flag <= 2) ? c->f.U1 : c->f.U;
    hipLaunchKernelGGL(k_estimate_nan, dim3(nblk(c->Ninner)), dim3(256), 0, c->stream, c->g, c->E, UI, c->f.LU,
                       c->f.err, flag, 0);
    if (int e = check_launch()) return e;
    return take_error(c);
}

int xf_reaction(xf_ctx *c, double dt)
{
    if (!c) return XF_EINVAL;
    int e = xf_update_states(c, 0);
    if (e) return e;
    double tem_dt = dt;
    if (c->d.strang) tem_dt *= 0.5;
    e = xf_chemistry(c, tem_dt, nullptr);
    if (e) return e;
    return xf_estimate_nan(c, 4);
}

int xf_initialize(xf_ctx *c)
{
    if (!c) return XF_EINVAL;
    int e = xf_boundary(c, 0);
    if (e) return e;
    return xf_update_states(c, 0);
}

static int steps_enqueue(xf_ctx *c, int nsteps, int use_graph)
{
    if (c->d.cop && !c->thermo_set) return XF_EINVAL;
    if (c->d.visc && !c->transport_set) return XF_EINVAL;
    if (c->d.react && (!c->mech_set || c->d.NR <= 0)) return XF_EINVAL;
    if (use_graph) {
        if (!c->gexec) {
            const bool prof = c->prof;
            c->prof = false;
            XF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            step_async(c);
            XF_HIP(hipStreamEndCapture(c->stream, &c->graph));
            c->prof = prof;
            XF_HIP(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        }
        for (int s = 0; s < nsteps; s++) XF_HIP(hipGraphLaunch(c->gexec, c->stream));
    } else {
        for (int s = 0; s < nsteps; s++) step_async(c);
    }
    return check_launch();
}

int xf_steps(xf_ctx *c, int nsteps, int use_graph, double *dt_out)
{
    if (!c || nsteps < 0) return XF_EINVAL;
    for (int s = 0; s < nsteps; s++) {
        if (dt_out) {
            int e = steps_enqueue(c, 1, use_graph);
            if (e) return e;
            XF_HIP(hipMemcpyAsync(dt_out + s, c->f.dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            e = take_error(c);
            if (e) return e;
        }
    }
    if (!dt_out) {
        int e = steps_enqueue(c, nsteps, use_graph);
        if (e) return e;
    }
    return take_error(c);
}

int xf_steps_async(xf_ctx *c, int nsteps, int use_graph)
{
    if (!c || nsteps < 0) return XF_EINVAL;
    return steps_enqueue(c, nsteps, use_graph);
}

int xf_synchronize(xf_ctx *c)
{
    if (!c) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(c->stream));
    return take_error(c);
}

int xf_last_error(xf_ctx *c, int *code, int *i, int *j, int *k)
{
    if (!c) return XF_EINVAL;
    if (code) *code = c->last_code;
    const int id = c->last_id;
    const int X = c->g.Xmax, Y = c->g.Ymax;
    if (i) *i = id >= 0 ? id % X : -1;
    if (j) *j = id >= 0 ? (id / X) % Y : -1;
    if (k) *k = id >= 0 ? id / (X * Y) : -1;
    return XF_OK;
}

int xf_last_error_step(xf_ctx *c, int *step, int *where)
{
    if (!c) return XF_EINVAL;
    if (step) *step = c->last_code ? c->last_step : 0;
    if (where) *where = c->last_code ? c->last_where : 0;
    return XF_OK;
}

int xf_set_time(xf_ctx *c, double time, int step)
{
    if (!c || step < 0 || !(time == time)) return XF_EINVAL;
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, c->stream, c->f.dt, c->f.err, 0, time, step);
    if (int e = check_launch()) return e;
    XF_HIP(hipStreamSynchronize(c->stream));
    return XF_OK;
}

int xf_set_target_time(xf_ctx *c, double target)
{
    if (!c || !(target == target)) return XF_EINVAL;
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, c->stream, c->f.dt, c->f.err, 1, target, 0);
    if (int e = check_launch()) return e;
    XF_HIP(hipStreamSynchronize(c->stream));
    return XF_OK;
}

int xf_get_time(xf_ctx *c, double *time, int *step, int *reached)
{
    if (!c) return XF_EINVAL;
    double t[DT_SIZE];
    int e[ERR_SIZE];
    XF_HIP(hipMemcpyAsync(t, c->f.dt, sizeof(t), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipMemcpyAsync(e, c->f.err, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (time) *time = t[DT_TIME];
    if (step) *step = e[ERR_COUNT];
    if (reached) *reached = e[ERR_DONE];
    return XF_OK;
}

int xf_set_profiling(xf_ctx *c, int on)
{
    if (!c) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(c->stream));
    for (auto &e : c->evs) (void)hipEventDestroy(e.a), (void)hipEventDestroy(e.b);
    c->evs.clear();
    c->times.clear();
    c->prof = on != 0;
    return XF_OK;
}

namespace {
// the GLF/CFL reduction slots, saved around measurement launches so that a
// timing or tuning run leaves the running GLF maxima exactly as it found them
struct RedGuard {
    xf_ctx *c;
    double h[RED_SIZE];
    bool ok;
    explicit RedGuard(xf_ctx *c_) : c(c_)
    {
        ok = hipMemcpyAsync(h, c->f.red, sizeof(h), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
    }
    ~RedGuard()
    {
        if (ok) {
            (void)hipMemcpyAsync(c->f.red, h, sizeof(h), hipMemcpyHostToDevice, c->stream);
            (void)hipStreamSynchronize(c->stream);
        }
    }
};

// average device ms of `reps` back-to-back launches of one kernel class in
// its stage-1 configuration (dir < 0: UpdateStates on U)
int time_launches(xf_ctx *c, int dir, int reps, double *avg_ms)
{
    int dirs[3], nd = 0;
    if (c->g.DimX) dirs[nd++] = 0;
    if (c->g.DimY) dirs[nd++] = 1;
    if (c->g.DimZ) dirs[nd++] = 2;
    // dir: -1 UpdateStates, 0..2 the sweep of that direction, 10 the viscous
    // preparation, 20..22 the viscous wall flux of direction dir - 20
    int q = -1;
    for (int i = 0; i < nd; i++)
        if (dirs[i] == dir || dirs[i] == dir - 20) q = i;
    if (((dir >= 0 && dir <= 2) || dir >= 20) && q < 0) return XF_EINVAL;
    if (dir >= 10 && (!c->d.visc || !c->transport_set)) return XF_EINVAL;
    hipEvent_t a, b;
    XF_HIP(hipEventCreate(&a));
    XF_HIP(hipEventCreate(&b));
    XF_HIP(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; r++) {
        if (dir < 0) {
            states_async(c, c->f.U, 0, 1, 1); // as in a step: ghost fill + UpdateStates
        } else if (dir == 10) {
            c->ops.visc_prep(c, c->f.U);
        } else if (dir >= 20) {
            c->ops.visc_flux(c, dir - 20, c->f.U);
        } else {
            int mode = (q > 0) ? 1 : 0;
            if (q == nd - 1) mode |= 2;
            c->ops.sweep(c, dir, c->f.U, mode, 1, c->f.U, c->U2, c->f.U1);
        }
    }
    XF_HIP(hipEventRecord(b, c->stream));
    XF_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    XF_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = ms / reps;
    return check_launch();
}
} // namespace

int xf_set_tile_shape(xf_ctx *c, int dir, int shape)
{
    if (!c || dir < 0 || dir > 2 || shape < 0) return XF_EINVAL;
    const int count = dir == 0 ? SweepShapes<0>::count : SweepShapes<1>::count;
    if (shape >= count) return XF_EINVAL;
    if (c->shape[dir] != shape) {
        c->shape[dir] = shape;
        // the captured step graph holds the old kernels
        if (c->gexec) (void)hipGraphExecDestroy(c->gexec), c->gexec = nullptr;
        if (c->graph) (void)hipGraphDestroy(c->graph), c->graph = nullptr;
    }
    return XF_OK;
}

int xf_set_visc_fusion(xf_ctx *c, int on)
{
    if (!c || !c->d.visc || (on != 0 && on != 1)) return XF_EINVAL;
    if (c->g.visc_fused != on) {
        c->g.visc_fused = on;
        // the captured step graph holds the old launches and arguments
        if (c->gexec) (void)hipGraphExecDestroy(c->gexec), c->gexec = nullptr;
        if (c->graph) (void)hipGraphDestroy(c->graph), c->graph = nullptr;
    }
    return XF_OK;
}

int xf_get_tile_shape(const xf_ctx *c, int *shape3)
{
    if (!c || !shape3) return XF_EINVAL;
    for (int d = 0; d < 3; d++) shape3[d] = c->shape[d];
    return XF_OK;
}

int xf_autotune(xf_ctx *c, int reps, int *shape3, double *ms)
{
    if (!c || reps <= 0) return XF_EINVAL;
    if (c->d.cop && !c->thermo_set) return XF_EINVAL;
    const bool prof = c->prof;
    c->prof = false;
    int rc = XF_OK;
    {
        RedGuard guard(c);
        if (!guard.ok) return XF_EHIP;
        const int dims[3] = {c->g.DimX, c->g.DimY, c->g.DimZ};
        for (int dir = 0; dir < 3 && rc == XF_OK; dir++) {
            const int count = dir == 0 ? SweepShapes<0>::count : SweepShapes<1>::count;
            for (int k = 0; k < 3; k++)
                if (ms) ms[3 * dir + k] = 0.0;
            if (!dims[dir]) continue;
            int best = 0;
            double best_ms = 0.0;
            for (int k = 0; k < count && rc == XF_OK; k++) {
                c->shape[dir] = k;
                double t = 0.0;
                rc = time_launches(c, dir, 2, &t); // warm-up (code load, caches)
                if (rc == XF_OK) rc = time_launches(c, dir, reps, &t);
                if (ms) ms[3 * dir + k] = t;
                if (rc == XF_OK && (k == 0 || t < best_ms)) best = k, best_ms = t;
            }
            c->shape[dir] = best;
        }
    }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec), c->gexec = nullptr;
    if (c->graph) (void)hipGraphDestroy(c->graph), c->graph = nullptr;
    c->prof = prof;
    if (shape3)
        for (int d = 0; d < 3; d++) shape3[d] = c->shape[d];
    if (rc != XF_OK) return rc;
    return take_error(c);
}

int xf_time_kernel(xf_ctx *c, const char *name, int reps, double *avg_ms)
{
    if (!c || !name || reps <= 0 || !avg_ms) return XF_EINVAL;
    if (c->d.cop && !c->thermo_set) return XF_EINVAL;
    const std::string n(name);
    int dir = -1;
    if (n == "sweep_x") dir = 0;
    else if (n == "sweep_y") dir = 1;
    else if (n == "sweep_z") dir = 2;
    else if (n == "visc_prep") dir = 10;
    else if (n == "visc_x") dir = 20;
    else if (n == "visc_y") dir = 21;
    else if (n == "visc_z") dir = 22;
    else if (n != "states") return XF_EINVAL;
    const bool prof = c->prof;
    c->prof = false;
    int rc;
    {
        RedGuard guard(c);
        if (!guard.ok) return XF_EHIP;
        rc = time_launches(c, dir, reps, avg_ms);
    }
    c->prof = prof;
    if (rc != XF_OK) return rc;
    return take_error(c);
}

int xf_set_profile_filter(xf_ctx *c, const char *csv)
{
    if (!c) return XF_EINVAL;
    c->prof_filter = (csv && *csv) ? "," + std::string(csv) + "," : std::string();
    return XF_OK;
}

int xf_kernel_times(xf_ctx *c, const char **names, double *avg_ms, int *counts, int cap)
{
    if (!c) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(c->stream));
    for (auto &e : c->evs) {
        float ms = 0.f;
        XF_HIP(hipEventElapsedTime(&ms, e.a, e.b));
        auto &t = c->times[e.name];
        t.first += ms;
        t.second += 1;
        (void)hipEventDestroy(e.a), (void)hipEventDestroy(e.b);
    }
    c->evs.clear();
    int n = 0;
    for (auto &kv : c->times) {
        if (n < cap) {
            if (names) names[n] = kv.first.c_str(); // stable: map nodes live with the context
            if (avg_ms) avg_ms[n] = kv.second.first / std::max(1, kv.second.second);
            if (counts) counts[n] = kv.second.second;
        }
        n++;
    }
    return n;
}

int xf_debug_math(const double *x, double *exp_out, double *log_out, int n)
{
    if (!x || !exp_out || !log_out || n <= 0) return XF_EINVAL;
    double *d = nullptr;
    XF_HIP(hipMalloc(&d, 3 * (size_t)n * sizeof(double)));
    hipError_t e = hipMemcpy(d, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_debug_math, dim3(nblk(n)), dim3(256), 0, 0, d, d + n, d + 2 * n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(exp_out, d + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(log_out, d + 2 * n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? XF_OK : XF_EHIP;
}

// ---------------------------------------------------------------------------
// Ensembles: independent cases advanced together, one graph launch per step
// ---------------------------------------------------------------------------
} // extern "C"

struct xf_ensemble {
    std::vector<xf_ctx *> cases;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    ChemCase *d_cases = nullptr;
    unsigned long long *d_queue = nullptr;
    long long total = 0;
    int blocks = 0;
    bool react = false, strang = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // one launch per kernel class over all cases (the cases agree in species
    // count, scheme, precision, dimensions and switches, and are inviscid)
    bool multi = false;
    EnsCase *d_ens = nullptr;
    void *spill = nullptr; // hand-over list of the shared chemistry queue
    int spill_cap = 0;
    long long state_blocks = 0, tiles[3] = {0, 0, 0};
    int shape[3] = {0, 0, 0};
};

namespace {

void ens_fork(xf_ensemble *e)
{
    (void)hipEventRecord(e->fork, e->stream);
    for (xf_ctx *c : e->cases) (void)hipStreamWaitEvent(c->stream, e->fork, 0);
}

void ens_join(xf_ensemble *e)
{
    for (size_t k = 0; k < e->cases.size(); k++) {
        (void)hipEventRecord(e->join[k], e->cases[k]->stream);
        (void)hipStreamWaitEvent(e->stream, e->join[k], 0);
    }
}

void ens_chem(xf_ensemble *e)
{
    xf_ctx *c0 = e->cases[0];
    c0->ops.chem_multi(e->stream, e->d_cases, (int)e->cases.size(), e->total, c0->d_th, c0->d_mc, e->d_queue,
                       e->blocks, 0, 1, c0->mech_spec, e->spill, e->spill_cap, c0->spill_after);
}

// one step of every case with one launch per kernel class: dt, [Reaction],
// three stages of UpdateStates (with the ghost fill) and the sweeps (the
// last fusing the RK update), [Reaction]; the launches of a single case's
// step, each over the blocks of every case
void ens_step_multi(xf_ensemble *e)
{
    xf_ctx *c0 = e->cases[0];
    const int n = (int)e->cases.size();
    hipStream_t st = e->stream;
    hipLaunchKernelGGL(k_dt_multi, dim3(1), dim3(64), 0, st, e->d_ens, n);
    auto reaction = [&]() {
        c0->ops.states_multi(c0, st, e->d_ens, n, e->state_blocks, 0, 0, 4, 0);
        hipLaunchKernelGGL(k_dt_chem_multi, dim3(1), dim3(64), 0, st, e->d_ens, n, c0->d.strang);
        ens_chem(e);
    };
    if (e->react && e->strang) reaction();
    int dirs[3], nd = 0;
    if (c0->g.DimX) dirs[nd++] = 0;
    if (c0->g.DimY) dirs[nd++] = 1;
    if (c0->g.DimZ) dirs[nd++] = 2;
    // stage 1: U -> U1; stage 2: (U, U1) -> U2; stage 3: (U, U2) -> U, as
    // step_mid_async; at stage 1 the sweep's Ub is U2 (the NaN check's U1)
    const int in_sel[3] = {0, 1, 2}, out_sel[3] = {1, 2, 0};
    for (int s = 1; s <= 3; s++) {
        c0->ops.states_multi(c0, st, e->d_ens, n, e->state_blocks, 0, 1, s, in_sel[s - 1]);
        for (int q = 0; q < nd; q++) {
            int mode = (q > 0) ? 1 : 0;
            if (q == nd - 1) mode |= 2;
            const int d = dirs[q];
            c0->ops.sweep_multi(c0, st, e->d_ens, n, e->tiles[d], d, e->shape[d], mode, s, in_sel[s - 1], 0,
                                s == 1 ? 2 : in_sel[s - 1], out_sel[s - 1]);
        }
    }
    if (e->react) reaction();
}

// one step of every case: each case's stages on its own stream, the
// chemistry of all cases in one work-queue launch at each Reaction
void ens_step(xf_ensemble *e)
{
    if (e->multi) {
        ens_step_multi(e);
        return;
    }
    ens_fork(e);
    for (xf_ctx *c : e->cases) step_pre_async(c);
    ens_join(e);
    if (e->react && e->strang) ens_chem(e);
    ens_fork(e);
    for (xf_ctx *c : e->cases) step_mid_async(c);
    ens_join(e);
    if (e->react) ens_chem(e);
}

} // namespace

extern "C" {

int xf_ensemble_create(xf_ctx *const *cases, int ncases, xf_ensemble **out)
{
    if (!cases || ncases <= 0 || !out) return XF_EINVAL;
    *out = nullptr;
    xf_ctx *c0 = cases[0];
    if (!c0) return XF_EINVAL;
    for (int k = 0; k < ncases; k++) {
        xf_ctx *c = cases[k];
        if (!c || c->d.NS != c0->d.NS || c->d.cop != c0->d.cop || c->d.NR != c0->d.NR ||
            c->d.react != c0->d.react || c->d.strang != c0->d.strang || c->d.ghost != c0->d.ghost)
            return XF_EINVAL;
        if (c->d.cop && !c->thermo_set) return XF_EINVAL;
        if (c->d.react && !c->mech_set) return XF_EINVAL;
        if (c->d.visc && !c->transport_set) return XF_EINVAL;
        // one chemistry launch serves every case: the tables must agree
        if (c->d.react && (std::memcmp(&c->h_th, &c0->h_th, sizeof(DevThermo)) != 0 ||
                           std::memcmp(&c->h_mc, &c0->h_mc, sizeof(DevMech)) != 0))
            return XF_EINVAL;
        for (int j = 0; j < k; j++)
            if (cases[j] == c) return XF_EINVAL;
    }
    xf_ensemble *e = new xf_ensemble();
    e->cases.assign(cases, cases + ncases);
    e->react = c0->d.react != 0;
    e->strang = c0->d.strang != 0;
    auto fail = [&](int code) {
        xf_ensemble_destroy(e);
        return code;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return fail(XF_EHIP);
    if (hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) return fail(XF_EHIP);
    e->join.assign(ncases, nullptr);
    for (auto &ev : e->join)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return fail(XF_EHIP);
    std::vector<ChemCase> h(ncases);
    long long off = 0;
    for (int k = 0; k < ncases; k++) {
        xf_ctx *c = cases[k];
        ChemCase &r = h[k];
        r.off = off;
        r.g = c->g;
        r.U = c->f.U, r.T = c->f.T, r.p = c->f.p, r.dt = c->f.dt;
        r.substeps = c->f.substeps, r.err = c->f.err;
        r.y = c->f.y;
        off += c->Ninner;
    }
    e->total = off;
    if (hipMalloc(&e->d_cases, ncases * sizeof(ChemCase)) != hipSuccess ||
        hipMalloc(&e->d_queue, 2 * sizeof(unsigned long long)) != hipSuccess)
        return fail(XF_EHIP);
    if (hipMemcpy(e->d_cases, h.data(), ncases * sizeof(ChemCase), hipMemcpyHostToDevice) != hipSuccess)
        return fail(XF_EHIP);
    unsigned long long z[2] = {0, 0};
    if (hipMemcpy(e->d_queue, z, sizeof(z), hipMemcpyHostToDevice) != hipSuccess) return fail(XF_EHIP);
    if (e->react) {
        e->blocks = c0->ops.multi_blocks(e->total, c0->mech_spec);
        if (CHEM_HANDOVER_ON) {
            e->spill_cap = (int)std::min<long long>(e->total, 65536);
            if (hipMalloc(&e->spill, (size_t)e->spill_cap * c0->ops.spill_bytes()) != hipSuccess)
                return fail(XF_EHIP);
        }
    }
    // the one-launch form: cases that agree in everything the launches are
    // specialised on, with the template case's tile shapes
    bool multi = !c0->d.visc && !c0->d.ghost;
    for (int k = 0; k < ncases && multi; k++) {
        const xf_case_desc &d = cases[k]->d, &d0 = c0->d;
        multi = !d.visc && !d.ghost && d.scheme == d0.scheme && d.precision == d0.precision && d.DimX == d0.DimX &&
                d.DimY == d0.DimY && d.DimZ == d0.DimZ && d.positivity == d0.positivity &&
                d.ncop_gamma == d0.ncop_gamma && cases[k]->ops.states == c0->ops.states &&
                (!d.react || cases[k]->mech_spec == c0->mech_spec);
    }
    if (multi) {
        std::vector<EnsCase> t(ncases);
        for (int q = 0; q < 3; q++) e->shape[q] = c0->shape[q];
        long long sb = 0, tb[3] = {0, 0, 0};
        for (int k = 0; k < ncases; k++) {
            xf_ctx *c = cases[k];
            EnsCase &r = t[k];
            r.g = c->g;
            r.f = c->f;
            r.U2 = c->U2;
            r.state_off = sb;
            sb += nblk(c->N);
            int bcs = 0;
            for (int q = 0; q < 6; q++) bcs |= (c->d.BC[q] & 15) << (4 * q);
            r.bcs = bcs | (1 << 28);
            for (int dd = 0; dd < 3; dd++) {
                unsigned gr[3];
                sweep_grid(c->g, dd, e->shape[dd], gr);
                for (int q = 0; q < 3; q++) r.tiles[dd][q] = (int)gr[q];
                r.tile_off[dd] = tb[dd];
                tb[dd] += (long long)gr[0] * gr[1] * gr[2];
            }
        }
        e->state_blocks = sb;
        for (int q = 0; q < 3; q++) e->tiles[q] = tb[q];
        if (hipMalloc(&e->d_ens, ncases * sizeof(EnsCase)) != hipSuccess ||
            hipMemcpy(e->d_ens, t.data(), ncases * sizeof(EnsCase), hipMemcpyHostToDevice) != hipSuccess)
            return fail(XF_EHIP);
        e->multi = true;
    }
    *out = e;
    return XF_OK;
}

int xf_ensemble_steps(xf_ensemble *e, int nsteps, int use_graph)
{
    if (!e || nsteps < 0) return XF_EINVAL;
    for (xf_ctx *c : e->cases) c->prof = false;
    if (use_graph) {
        if (!e->gexec) {
            XF_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            ens_step(e);
            XF_HIP(hipStreamEndCapture(e->stream, &e->graph));
            XF_HIP(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
        }
        for (int s = 0; s < nsteps; s++) XF_HIP(hipGraphLaunch(e->gexec, e->stream));
    } else {
        for (int s = 0; s < nsteps; s++) ens_step(e);
    }
    return check_launch();
}

int xf_ensemble_synchronize(xf_ensemble *e)
{
    if (!e) return XF_EINVAL;
    XF_HIP(hipStreamSynchronize(e->stream));
    int first = XF_OK;
    for (xf_ctx *c : e->cases) {
        XF_HIP(hipStreamSynchronize(c->stream));
        const int st = take_error(c);
        if (st && !first) first = st;
    }
    return first;
}

long long xf_ensemble_cells(const xf_ensemble *e) { return e ? e->total : 0; }

int xf_ensemble_one_launch(const xf_ensemble *e) { return e ? (e->multi ? 1 : 0) : XF_EINVAL; }

void xf_ensemble_destroy(xf_ensemble *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->fork) (void)hipEventDestroy(e->fork);
    for (auto ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->d_cases) (void)hipFree(e->d_cases);
    if (e->d_ens) (void)hipFree(e->d_ens);
    if (e->spill) (void)hipFree(e->spill);
    if (e->d_queue) (void)hipFree(e->d_queue);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int xf_fma_probe(int reps, double *tflops)
{
    if (reps <= 0 || !tflops) return XF_EINVAL;
    // 8 waves on each SIMD of 256 CUs; every lane writes its own out[t]
    const unsigned blocks = 256 * 8, threads = 256;
    const int iters = 4096;
    double *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&out, (size_t)blocks * threads * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999999, 1e-6); // warm-up
        e = hipEventRecord(e0, 0);
    }
    for (int r = 0; e == hipSuccess && r < reps; r++) {
        hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999999, 1e-6);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    // two flops per fma
    if (e == hipSuccess) *tflops = 2.0 * 8.0 * iters * (double)blocks * threads * reps / (ms * 1e-3) / 1e12;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return e == hipSuccess ? XF_OK : XF_EHIP;
}

int xf_copy_probe(size_t nbytes, int reps, double *avg_ms)
{
    if (nbytes < 16 || (nbytes & 15) || reps <= 0 || !avg_ms) return XF_EINVAL;
    double2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const long long n = (long long)(nbytes / 16);
    hipError_t e = hipMalloc(&a, nbytes);
    if (e == hipSuccess) e = hipMalloc(&b, nbytes);
    if (e == hipSuccess) e = hipMemset(a, 0, nbytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // 8 blocks of 256 lanes per CU-slot over 256 CUs, grid-stride beyond that
    const long long want = (n + 255) / 256;
    const unsigned blocks = (unsigned)(want < 256 * 32 ? want : 256 * 32);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm-up
        e = hipEventRecord(e0, 0);
    }
    for (int r = 0; e == hipSuccess && r < reps; r++) {
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) *avg_ms = (double)ms / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return e == hipSuccess ? XF_OK : XF_EHIP;
}

} // extern "C"
