// Fluid-stage kernels for gfx950: ghost fill, conservative->primitive
// conversion with fused GLF/CFL reductions, SSP-RK3 update, dt.
//
// Reference stages restated (operation order preserved so the plain-FP64
// form is bit-identical to the CPU restatement):
//   FluidBCKernel{X,Y,Z}      src/solver_BCs/BCs_kernels.hpp:9-258
//   Updaterhoyi + GetStates    src/solver_UpdateStates/Update_device.hpp:7-54
//   Estimate{Yi,PrimitiveVar}  src/solver_UpdateStates/Estimate_kernels.hpp
//   GetLocalEigen + GLF max    FDM_Method/positive-definite_eigen/Eigen_value.hpp:5-33,
//                              FDM_Method/ConVenction_block.hpp:115-172
//   GetDt                      src/solver_GetDt/GlobalDt_block.hpp:6-114
//   UpdateURK3rdKernel         src/solver_UpdateStates/Update_kernels.hpp:64-94
#pragma once
#include "xf_thermo.h"

namespace xf {

// (i, j, k) of padded cell id: 32-bit divisions when the grid has fewer than
// 2^31 cells (a uniform branch), 64-bit ones otherwise
__device__ __forceinline__ void cell_ijk(const Geom &g, long long id, int &i, int &j, int &k)
{
    if (g.N < 2147483648LL) {
        const unsigned u = (unsigned)id, X = (unsigned)g.Xmax, Y = (unsigned)g.Ymax;
        const unsigned r = u / X;
        i = (int)(u - r * X), j = (int)(r % Y), k = (int)(r / Y);
    } else {
        i = (int)(id % g.Xmax), j = (int)((id / g.Xmax) % g.Ymax), k = (int)(id / ((long long)g.Xmax * g.Ymax));
    }
}

// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs (linear id b runs on XCD b % 8), and each XCD has its own L2.  The
// tile a workgroup works on is chosen so that XCD x owns one contiguous run
// of tiles in row-major order: neighbouring tiles, which share halo cells and
// the cache lines that straddle tile edges, then meet in the same L2.  The
// map is a bijection of [0, nb) for any nb.
constexpr int kXcds = 8;
__device__ __forceinline__ void xcd_tile(int &tx, int &ty, int &tz)
{
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int nb = gx * gy * (int)gridDim.z;
    const int b = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, slot = b / kXcds;
    const int l = (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
    tx = l % gx;
    ty = (l / gx) % gy;
    tz = l / (gx * gy);
}

// ---------------------------------------------------------------------------
// wave64 max reduction (butterfly over lane permutes) of non-negative doubles
// ---------------------------------------------------------------------------
// maximum over the wave of magnitudes (>= +0, or NaN): fmax takes two
// instructions per step where a compare and two selects took three, and it
// drops a NaN, as the fold of the block maxima does (it keeps only values > 0)
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// ---------------------------------------------------------------------------
// ghost fill along one axis: one thread per (ghost layer g, transverse a, b)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bc_target(int bc, int g, int B, int inner, int mirror_offset, int index_inner, int sign)
{
    switch (bc) {
    case XF_BC_SYMMETRY:
    case XF_BC_NSLIPWALL:
    case XF_BC_VISCWALL:
    case XF_BC_SLIPWALL:
        return 2 * (B + mirror_offset) - 1 - g;
    case XF_BC_PERIODIC:
        return g + sign * inner;
    case XF_BC_OUTFLOW:
        return index_inner;
    default:
        return g;
    }
}

__device__ __forceinline__ void bc_apply(double *__restrict__ U, long long N, int E, int NCOP, long long id, long long tid,
                                         int bc, int mom)
{
    switch (bc) {
    case XF_BC_SYMMETRY:
        for (int n = 0; n < E; n++) U[n * N + id] = U[n * N + tid];
        U[mom * N + id] = -U[mom * N + tid];
        break;
    case XF_BC_PERIODIC:
    case XF_BC_OUTFLOW:
        for (int n = 0; n < E; n++) U[n * N + id] = U[n * N + tid];
        break;
    case XF_BC_NSLIPWALL:
    case XF_BC_VISCWALL:
    case XF_BC_SLIPWALL:
        U[0 * N + id] = U[0 * N + tid];
        U[1 * N + id] = -U[1 * N + tid];
        U[2 * N + id] = -U[2 * N + tid];
        U[3 * N + id] = -U[3 * N + tid];
        U[4 * N + id] = U[4 * N + tid];
        for (int n = E - NCOP; n < E; n++) U[n * N + id] = U[n * N + tid];
        break;
    default:
        break;
    }
}

// All ghost cells in one pass.  The reference fills x, then y, then z
// (BCs_block.cpp:50-212), each pass reading cells the previous one may have
// written (corners).  Composing the three axis maps per ghost cell gives the
// same values in a single launch: starting from the cell, apply the z map if
// it is a z ghost, then the y map, then the x map; the value is the source
// cell's value with the momentum sign flips of every map applied.  Sources
// are interior cells or Inflow ghosts (which no pass writes), so no thread
// reads a cell another thread writes.  Copies and negations are exact.
// Thread 0 also zeroes the per-pass reduction slots for the following
// UpdateStates pass when reset_red is set.
__device__ __forceinline__ void bc_axis(int bc, int &s, int B, int inner, int Lmax, int mom, int &flip)
{
    if (s >= B && s < Lmax - B) return;
    const bool lo = s < B;
    const int t = lo ? bc_target(bc, s, B, inner, 0, B, 1) : bc_target(bc, s, B, inner, inner, Lmax - B - 1, -1);
    if (bc == XF_BC_SYMMETRY) flip ^= 1 << mom;
    if (bc == XF_BC_NSLIPWALL || bc == XF_BC_VISCWALL || bc == XF_BC_SLIPWALL) flip ^= (1 << 1) | (1 << 2) | (1 << 3);
    s = t;
}

static __global__ void __launch_bounds__(256) k_boundary_all(Geom g, double *__restrict__ U, int E, int bx0, int bx1, int by0,
                                                      int by1, int bz0, int bz1, double *red, int reset_red)
{
    if (reset_red && blockIdx.x == 0 && threadIdx.x < 12) {
        if (threadIdx.x < 9) red[RED_GLFP + threadIdx.x] = 0.0;
        else red[RED_DT + threadIdx.x - 9] = 0.0;
    }
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    // enumerate the ghost shell: z slabs, then y slabs inside interior z,
    // then x slabs inside interior y and z
    const long long nz = g.DimZ ? 2LL * g.Bz * g.Ymax * g.Xmax : 0;
    const long long ny = g.DimY ? 2LL * g.By * g.Xmax * g.Z_inner : 0;
    const long long nx = g.DimX ? 2LL * g.Bx * g.Y_inner * g.Z_inner : 0;
    int i, j, k;
    // quotient and remainder in 32 bits when the shell is below 2^31 cells
    auto dm = [&](long long a, long long b, long long &q, long long &r) {
        if (nz + ny + nx < 2147483648LL) {
            const unsigned qa = (unsigned)a / (unsigned)b;
            q = qa, r = (long long)((unsigned)a - qa * (unsigned)b);
        } else {
            q = a / b, r = a % b;
        }
    };
    long long q, r, q2, r2;
    if (t < nz) {
        dm(t, (long long)g.Ymax * g.Xmax, q, r);
        const int gz = (int)q;
        k = gz < g.Bz ? gz : gz - g.Bz + g.Zmax - g.Bz;
        dm(r, g.Xmax, q2, r2);
        j = (int)q2, i = (int)r2;
    } else if ((t -= nz) < ny) {
        dm(t, (long long)g.Xmax * g.Z_inner, q, r);
        const int gy = (int)q;
        j = gy < g.By ? gy : gy - g.By + g.Ymax - g.By;
        dm(r, g.Xmax, q2, r2);
        k = g.Bz + (int)q2, i = (int)r2;
    } else if ((t -= ny) < nx) {
        dm(t, (long long)g.Y_inner * g.Z_inner, q, r);
        const int gx = (int)q;
        i = gx < g.Bx ? gx : gx - g.Bx + g.Xmax - g.Bx;
        dm(r, g.Y_inner, q2, r2);
        k = g.Bz + (int)q2, j = g.By + (int)r2;
    } else
        return;
    int si = i, sj = j, sk = k, flip = 0;
    if (g.DimZ) bc_axis(k < g.Bz ? bz0 : bz1, sk, g.Bz, g.Z_inner, g.Zmax, 3, flip);
    if (g.DimY) bc_axis(sj < g.By ? by0 : by1, sj, g.By, g.Y_inner, g.Ymax, 2, flip);
    if (g.DimX) bc_axis(si < g.Bx ? bx0 : bx1, si, g.Bx, g.X_inner, g.Xmax, 1, flip);
    if (si == i && sj == j && sk == k) return; // Inflow along every ghost axis: untouched
    const long long N = g.N;
    const long long id = (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
    const long long sid = (long long)g.Xmax * g.Ymax * sk + (long long)g.Xmax * sj + si;
    for (int n = 0; n < E; n++) {
        const double v = U[n * N + sid];
        U[n * N + id] = ((flip >> n) & 1) ? -v : v;
    }
}

// ---------------------------------------------------------------------------
// conservative -> primitive over all padded cells
// ---------------------------------------------------------------------------
// bcs: 0 = the ghost cells already hold their values; otherwise the six BC
// codes packed 4 bits each (XMIN in bits 0-3 ... ZMAX in bits 20-23) plus
// bit 28, and every ghost cell first takes its value from its source cell as
// k_boundary_all forms it (FluidBoundaryCondition, fused into UpdateStates:
// the sources are interior cells or Inflow ghosts, which no thread writes)
// the 256 cells of block bid of one case
template <int NS, bool COP>
__device__ __forceinline__ void states_block(const Geom &g, const DevThermo *__restrict__ th, double *Ug,
                                             const Fields &f, double ncop_gamma, int full, int bcs, int where,
                                             long long bid)
{
    constexpr int E = COP ? NS + 4 : 5;
    constexpr int NCOP = COP ? NS - 1 : 0;
    // block-uniform: the words tested are not written by this launch
    if (step_halted(f.err, ~(EBIT_RHO_Y | EBIT_PRIM))) return;
    const long long N = g.N;
    const long long id = bid * blockDim.x + threadIdx.x;
    const bool live = id < N;
    double glf[9], dtm[3];
#pragma unroll
    for (int q = 0; q < 9; q++) glf[q] = 0.0;
#pragma unroll
    for (int q = 0; q < 3; q++) dtm[q] = 0.0;
    int ebits = 0;
    if (live) {
        int i, j, k;
        cell_ijk(g, id, i, j, k);
        long long src = id;
        int flip = 0;
        bool wb = false;
        if (bcs) {
            int si = i, sj = j, sk = k;
            auto bc = [&](int q) { return (bcs >> (4 * q)) & 15; };
            if (g.DimZ) bc_axis(sk < g.Bz ? bc(4) : bc(5), sk, g.Bz, g.Z_inner, g.Zmax, 3, flip);
            if (g.DimY) bc_axis(sj < g.By ? bc(2) : bc(3), sj, g.By, g.Y_inner, g.Ymax, 2, flip);
            if (g.DimX) bc_axis(si < g.Bx ? bc(0) : bc(1), si, g.Bx, g.X_inner, g.Xmax, 1, flip);
            if (si != i || sj != j || sk != k) {
                src = (long long)g.Xmax * g.Ymax * sk + (long long)g.Xmax * sj + si;
                wb = true;
            }
        }
        double U[E];
#pragma unroll
        for (int n = 0; n < E; n++) {
            const double v = Ug[n * N + src];
            U[n] = ((flip >> n) & 1) ? -v : v;
        }
        if (wb) {
#pragma unroll
            for (int n = 0; n < E; n++) Ug[n * N + id] = U[n];
        }
        // Getrhoyi
        const double rho = U[0];
        const double rho1 = 1.0 / rho;
        double yi[COP ? NS : 1];
        if (COP && g.ghost) {
            // GhostSpecies: renormalise the first NS - 1 fractions; the state's
            // partial densities become rho * Y_k in k_rhoyi_writeback, after
            // this launch (the ghost fill above reads sources unchanged)
            yi[NCOP] = 0.0;
            double sum_yi = 0.0;
#pragma unroll
            for (int ii = 0; ii < NCOP; ii++) yi[ii] = U[ii + 5] * rho1, sum_yi += yi[ii];
            sum_yi = 1.0 / sum_yi;
#pragma unroll
            for (int ii = 0; ii < NCOP; ii++) yi[ii] *= sum_yi, f.y[ii * N + id] = yi[ii];
        } else if (COP) {
            yi[NCOP] = 1.0;
#pragma unroll
            for (int ii = 5; ii < E; ii++) yi[ii - 5] = U[ii] * rho1, yi[NCOP] += -yi[ii - 5];
        }
        // GetStates
        const double u = U[1] * rho1, v = U[2] * rho1, w = U[3] * rho1;
        const double tme = U[4] * rho1 - 0.5 * (u * u + v * v + w * w);
        double T = f.T[id], p, gamma;
        if (COP) {
            const double R_ = mix_R<NS>(th, yi);
            T = get_T<NS>(th, yi, tme, T);
            p = rho * R_ * T;
            const TPre t = th->thermo ? tpre<true>(T, 1) : tpre<false>(T, 0);
            // species Cp and h: with the fit range as a wave-uniform value when
            // every lane shares it (scalar coefficient loads)
            double cpk[NS], hk[NS];
            auto species = [&](const TPre &tt) {
#pragma unroll
                for (int n = 0; n < NS; n++) cpk[n] = cp_k(th, tt, th->Ri[n], n), hk[n] = h_k(th, tt, th->Ri[n], n);
            };
            const int r0 = __builtin_amdgcn_readfirstlane(t.r);
            if (__all(t.r == r0)) {
                TPre tu = t;
                tu.r = r0;
                species(tu);
            } else {
                species(t);
            }
            double Cp = 0.0;
#pragma unroll
            for (int n = 0; n < NS; n++) Cp += yi[n] * cpk[n];
            gamma = mix_gamma_cp<NS>(th, yi, Cp);
            // ingredients of the face sound speed (ReconstructSoundSpeed,
            // positive-definite_eigen/Utils_device.hpp:102-140): h_k(T), Cp,
            // get_CopGamma(yi, T) = Cp / Cv, R -- identical values to the
            // per-face recomputation of the reference
            double cv = 0.0, _W = 0.0;
#pragma unroll
            for (int n = 0; n < NS; n++) cv += yi[n] * cpk[n];
#pragma unroll
            for (int n = 0; n < NS; n++) _W += yi[n] * th->_Wi[n];
            cv -= Ru * _W;
#pragma unroll
            for (int n = 0; n < NS; n++) f.hi[n * N + id] = hk[n];
            f.Cp[id] = Cp;
            f.gcv[id] = Cp / cv;
            f.R[id] = R_;
            f.T[id] = T;
        } else {
            gamma = ncop_gamma;
            p = (ncop_gamma - 1.0) * rho * tme;
        }
        const double H = (U[4] + p) * rho1;
        const double c = sqrt(gamma * p * rho1);
        // the sweeps rebuild u, v, w and H from U and 1/rho with the same
        // operations, so the time-step path stores only p (and the COP
        // ingredients above); the full set is written for the stage seam
        f.p[id] = p;
        if (full) {
            f.rho[id] = rho;
            f.u[id] = u;
            f.v[id] = v;
            f.w[id] = w;
            f.H[id] = H;
            f.c[id] = c;
            f.gamma[id] = gamma;
            f.e[id] = tme;
        }
        // validation on interior cells
        const bool interior = i >= g.Bx && i < g.Xmax - g.Bx && j >= g.By && j < g.Ymax - g.By && k >= g.Bz &&
                              k < g.Zmax - g.Bz;
        if (interior) {
            bool bad = (rho < 0) || isnan(rho) || isinf(rho);
            if (COP) {
#pragma unroll
                for (int n = 0; n < NS; n++) bad = bad || isnan(yi[n]) || isinf(yi[n]);
            }
            if (bad) ebits |= EBIT_RHO_Y;
            const double Tc = COP ? T : f.T[id];
            const double vs[3] = {rho, p, Tc};
#pragma unroll
            for (int q = 0; q < 3; q++)
                if (vs[q] < 0 || isnan(vs[q]) || isinf(vs[q])) ebits |= EBIT_PRIM;
        }
        // GLF (eigen_local = u_d -+ c, u_d) and CFL (|u_d| + sqrt(1.4 p / rho))
        const double cl = sqrt(1.4 * p / rho);
        const double vel[3] = {u, v, w};
        const int dims[3] = {g.DimX, g.DimY, g.DimZ};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (!dims[d]) continue;
            const double uu = vel[d];
            glf[3 * d + 0] = fabs(uu - c);
            glf[3 * d + 1] = fabs(uu);
            glf[3 * d + 2] = fabs(uu + c);
            dtm[d] = fabs(uu) + cl;
        }
    }
    // block reduction: wave butterfly, then one lane per wave folds into LDS
    __shared__ double sred[4][12];
    __shared__ int serr;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (threadIdx.x == 0) serr = 0;
    __syncthreads();
    // the directions the case has (the others stay +0)
    const int dimv[3] = {g.DimX, g.DimY, g.DimZ};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (!dimv[d]) continue;
#pragma unroll
        for (int q = 0; q < 3; q++) glf[3 * d + q] = wave_max(glf[3 * d + q]);
        dtm[d] = wave_max(dtm[d]);
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 9; q++) sred[wv][q] = glf[q];
#pragma unroll
        for (int q = 0; q < 3; q++) sred[wv][9 + q] = dtm[q];
    }
    if (ebits) {
        atomicOr(&serr, ebits);
        if (ebits & EBIT_RHO_Y) report_cell(f.err, EBIT_RHO_Y, id);
        if (ebits & EBIT_PRIM) report_cell(f.err, EBIT_PRIM, id);
    }
    __syncthreads();
    // the block's maxima go to spart; k_states_fold combines the blocks (one
    // pass over the partials instead of same-address atomics from every block)
    if (threadIdx.x < 12) {
        double m = sred[0][threadIdx.x];
        const int nw = (blockDim.x + 63) >> 6;
        for (int q = 1; q < nw; q++) m = sred[q][threadIdx.x] > m ? sred[q][threadIdx.x] : m;
        f.spart[bid * 12 + threadIdx.x] = m;
    }
    if (threadIdx.x == 0 && serr) report_pass(f.err, serr, where);
}

template <int NS, bool COP>
__global__ void __launch_bounds__(256) k_states(Geom g, const DevThermo *__restrict__ th, double *Ug, Fields f,
                                                double ncop_gamma, int full, int bcs, int where)
{
    states_block<NS, COP>(g, th, Ug, f, ncop_gamma, full, bcs, where, (long long)blockIdx.x);
}

// GhostSpecies write-back of Getrhoyi (Update_device.hpp:20-21) on every
// padded cell: U_{5+k} = rho * Y_k with the fractions k_states stored.  A
// launch of its own after k_states, whose ghost fill reads interior cells as
// its sources; it runs exactly when k_states did (the same entry test, and
// before k_states_fold latches this pass's failures).
template <int NS>
__global__ void __launch_bounds__(256) k_rhoyi_writeback(Geom g, double *Ug, Fields f)
{
    if (step_halted(f.err, ~(EBIT_RHO_Y | EBIT_PRIM))) return;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= g.N) return;
    const double rho = Ug[id];
#pragma unroll
    for (int k = 0; k < NS - 1; k++) Ug[(5 + k) * g.N + id] = rho * f.y[k * g.N + id];
}

// UpdateStates of every case of an ensemble in one launch (fill = 1: with
// each case's ghost fill); sel picks the state (0 U, 1 U1, 2 U2)
template <int NS, bool COP>
__global__ void __launch_bounds__(256) k_states_multi(const EnsCase *__restrict__ cases, int ncases,
                                                      const DevThermo *__restrict__ th, double ncop_gamma, int full,
                                                      int fill, int where, int sel)
{
    const long long b = blockIdx.x;
    const int k = ens_find(cases, ncases, b, 3);
    const EnsCase &c = cases[k];
    states_block<NS, COP>(c.g, th, c.buf(sel), c.f, ncop_gamma, full, fill ? c.bcs : 0, where, b - c.state_off);
}

// the per-block GLF and CFL maxima of k_states into the reduction slots: one
// block per slot (12), each the maximum of +0 and every positive block value
// (the pass's maxima, combined with the 0 the pass starts from)
__device__ __forceinline__ void states_fold(const double *__restrict__ spart, int nblocks, double *red, int *err, int q)
{
    __shared__ unsigned long long s[1024];
    // latch a failure of the UpdateStates pass (or of any earlier one)
    if (q == 0 && threadIdx.x == 0 && err[ERR_BITS]) err[ERR_HALT] = 1;
    unsigned long long v = 0ULL; // +0.0
    for (int b = threadIdx.x; b < nblocks; b += 1024) {
        const double x = spart[(long long)b * 12 + q];
        const unsigned long long bx = (unsigned long long)__double_as_longlong(x);
        if (x > 0.0 && bx > v) v = bx;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if (threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) red[q < 9 ? RED_GLFP + q : RED_DT + q - 9] = __longlong_as_double((long long)s[0]);
}

static __global__ void __launch_bounds__(1024) k_states_fold(const double *__restrict__ spart, int nblocks, double *red,
                                                             int *err)
{
    states_fold(spart, nblocks, red, err, blockIdx.x);
}

// the folds of every case: 12 blocks per case
static __global__ void __launch_bounds__(1024) k_states_fold_multi(const EnsCase *__restrict__ cases)
{
    const EnsCase &c = cases[blockIdx.x / 12];
    states_fold(c.f.spart, (int)((c.g.N + 255) / 256), c.f.red, c.f.err, blockIdx.x % 12);
}

// ---------------------------------------------------------------------------
// small control kernels
// ---------------------------------------------------------------------------
// GetDt: dt = CFL / sum_d max_d / Delta_d, written to device memory.
// step = 1 is the head of a time step (XFLUIDS::Evolution,
// src/XFLUIDS.cpp:172-198): nothing happens once the step is halted or the
// clock has reached the target time (the loop condition physicalTime <
// target_t); otherwise the step is counted (Iteration++), dt is clamped so
// that the clock lands on the target (physicalTime + dt > target_t ->
// dt = target_t - physicalTime) and the clock advances.  It also latches a
// failure of the previous pass into ERR_HALT.  step = 0 is GetDt alone.
__device__ __forceinline__ void dt_case(const Geom &g, double *red, double *dt, int *err, int step)
{
    if (step) {
        if (err[ERR_BITS]) err[ERR_HALT] = 1;
        if (err[ERR_HALT]) return;
        if (!(dt[DT_TIME] < dt[DT_TARGET])) {
            err[ERR_DONE] = 1;
            err[ERR_HALT] = 1;
            dt[DT_STEP] = 0.0; // no step taken
            return;
        }
        err[ERR_COUNT] = err[ERR_COUNT] + 1;
    }
    const double mx = g.DimX ? red[RED_DT + 0] : 0.0;
    const double my = g.DimY ? red[RED_DT + 1] : 0.0;
    const double mz = g.DimZ ? red[RED_DT + 2] : 0.0;
    const double dtref = mx * g._dx + my * g._dy + mz * g._dz;
    double d = g.CFL / dtref;
    if (step) {
        const double t = dt[DT_TIME], te = dt[DT_TARGET];
        if (t + d > te) d = te - t;
        dt[DT_TIME] = t + d;
    }
    dt[DT_STEP] = d;
    // uvw_c_max as GetDt leaves it (the positivity limiter's lambda_0)
    red[RED_LAM + 0] = mx, red[RED_LAM + 1] = my, red[RED_LAM + 2] = mz;
}

static __global__ void k_dt(Geom g, double *red, double *dt, int *err, int step)
{
    if (threadIdx.x == 0) dt_case(g, red, dt, err, step);
}

// Strang half step on device: dt[1] = dt[0] * 0.5 (or dt[0] for Lie)
static __global__ void k_dt_chem(double *dt, int strang, const int *err)
{
    if (threadIdx.x == 0 && !step_halted(err, ~0)) dt[DT_CHEM] = strang ? dt[DT_STEP] * 0.5 : dt[DT_STEP];
}

// the same two for every case of an ensemble, one lane per case
static __global__ void k_dt_multi(const EnsCase *__restrict__ cases, int ncases)
{
    for (int k = threadIdx.x; k < ncases; k += blockDim.x) dt_case(cases[k].g, cases[k].f.red, cases[k].f.dt, cases[k].f.err, 1);
}
static __global__ void k_dt_chem_multi(const EnsCase *__restrict__ cases, int ncases, int strang)
{
    for (int k = threadIdx.x; k < ncases; k += blockDim.x) {
        const Fields &f = cases[k].f;
        if (!step_halted(f.err, ~0)) f.dt[DT_CHEM] = strang ? f.dt[DT_STEP] * 0.5 : f.dt[DT_STEP];
    }
}

// ---------------------------------------------------------------------------
// SSP-RK3 stage on interior cells (Update_kernels.hpp:64-94); dt from device
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rk3_point(int stage, double dt, double &U, double &U1, double LU)
{
    if (stage == 1)
        U1 = U + dt * LU;
    else if (stage == 2)
        U1 = 0.75 * U + 0.25 * U1 + 0.25 * dt * LU;
    else
        U = (U + 2.0 * U1 + 2.0 * dt * LU) * OT;
}

static __global__ void __launch_bounds__(256) k_rk3(Geom g, int E, double *__restrict__ U, double *__restrict__ U1,
                                             const double *__restrict__ LU, const double *__restrict__ dtp, int stage)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.X_inner * g.Y_inner * g.Z_inner;
    if (t >= total) return;
    const int i = g.Bx + (int)(t % g.X_inner);
    const int j = g.By + (int)((t / g.X_inner) % g.Y_inner);
    const int k = g.Bz + (int)(t / ((long long)g.X_inner * g.Y_inner));
    const long long id = (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
    const double dt = dtp[0];
    for (int n = 0; n < E; n++) {
        double a = U[n * g.N + id], b = U1[n * g.N + id];
        rk3_point(stage, dt, a, b, LU[n * g.N + id]);
        if (stage == 3) U[n * g.N + id] = a;
        else U1[n * g.N + id] = b;
    }
}

// EstimateFluidNANKernel (src/Fluids.cpp:47-87) on interior cells.  LU may
// be null: after the Reaction (flag 4) the reference checks the LU of the
// third stage again, which that stage's own check already passed and nothing
// has changed since.  step = 1: a pass of a time step (returns at entry once
// an earlier pass failed).
static __global__ void __launch_bounds__(256) k_estimate_nan(Geom g, int E, const double *__restrict__ UI,
                                                      const double *__restrict__ LU, int *err, int where, int step)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.X_inner * g.Y_inner * g.Z_inner;
    if (t >= total) return;
    if (step && step_halted(err, ~EBIT_NAN)) return;
    const int i = g.Bx + (int)(t % g.X_inner);
    const int j = g.By + (int)((t / g.X_inner) % g.Y_inner);
    const int k = g.Bz + (int)(t / ((long long)g.X_inner * g.Y_inner));
    const long long id = (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
    bool bad = UI[id] < 0;
    for (int n = 0; n < E; n++) {
        const double a = UI[n * g.N + id];
        bad = bad || isnan(a) || isinf(a);
        if (LU) {
            const double b = LU[n * g.N + id];
            bad = bad || isnan(b) || isinf(b);
        }
    }
    if (bad) {
        report_cell(err, EBIT_NAN, id);
        report_pass(err, EBIT_NAN, where);
    }
}

} // namespace xf
