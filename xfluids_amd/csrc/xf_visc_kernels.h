// Viscous wall fluxes: velocity gradients at cell centres, mixture-averaged
// transport coefficients and the fourth-order viscous, heat-conduction and
// species-diffusion face fluxes that GetLU subtracts from the convective face
// fluxes (src/solver_Reconstruction/FDM_Method/ConVenction_block.hpp:424-575).
//
// Kernels, in launch order for one GetLU:
//   k_transport<NS,PREC> every padded cell: u, v, w, Wilke mu and lambda, D_km
//                     (PREC 1: the mixed form's FP32 exponentials and factors);
//                     per-block min/max of each interior mass fraction
//   k_yreduce<NS>     one block: those partials into the reduction slots
//   k_vel_deriv<PREC> fourth-order centre derivatives of u, v, w, with the
//                     ghost-cell rules for the values each wall flux reads
//   visc_face<NS,DIR,PREC> per face of one direction, inside the sweep of
//                     that direction (NS >= 4), or k_visc_flux<NS,DIR,PREC>
//                     writing the flux planes the sweep subtracts
// All arithmetic of the plain form is FP64 in the reference's operation
// order, with the same exp/log as the oracle, so it stays bit-identical; the
// mixed form's transport coefficients take FP32 where the FP64 path spends
// its instructions (transport_mixed), and the mixed form stores the transport
// coefficients and gradients as FP32 planes that the wall flux reads back
// into its FP64 arithmetic.
#pragma once
#include "xf_fluid_kernels.h"
#include <type_traits>

namespace xf {

// Transport fits (Setup::GetFitCoefficient, src/read_ini/src/viscfit.cpp:148-193)
// and the species-pair factors of PHI (Visc_device.h:27-35):
// phiA[k][j] = (W_j/W_k)^(1/4) = sqrt(sqrt(W_j/W_k)), phiB[k][j] = (1 + W_k/W_j)^(-1/2)
struct DevTransport {
    double fcv[MAXNS * 4], fct[MAXNS * 4], Dkj[MAXNS * MAXNS * 4];
    double phiA[MAXNS * MAXNS], phiB[MAXNS * MAXNS];
    float phiAf[MAXNS * MAXNS], phiBcf[MAXNS * MAXNS]; // mixed form: phiA, phiB * 0.5 / sqrt(2)
};

// Vde component order (global_setup.h:122-133)
enum { DUCX = 0, DVCX = 1, DWCX = 2, DUCY = 3, DVCY = 4, DWCY = 5, DUCZ = 6, DVCZ = 7, DWCZ = 8 };

// sycl::min / sycl::max on doubles
__device__ __forceinline__ double smin(double x, double y) { return y < x ? y : x; }
__device__ __forceinline__ double smax(double x, double y) { return x < y ? y : x; }

// RN(a / b) from y = RN(1 / b) (Markstein's correction, exact for a and b
// in [2^-500, 2^500]); other operands take the IEEE division.  Lets quotients
// that share a divisor share one reciprocal.
__device__ __noinline__ double ieee_div(double a, double b) { return a / b; }
__device__ __forceinline__ double qdiv(double a, double b, double y)
{
    double q = fma(fma(-b, a * y, a), y, a * y);
    const double aa = fabs(a), ab = fabs(b);
    if (__builtin_expect(!(aa >= 0x1p-500 && aa <= 0x1p500 && ab >= 0x1p-500 && ab <= 0x1p500), 0))
        q = ieee_div(a, b);
    return q;
}

// cubic in ln T (Viscosity / Thermal_conductivity / GetDkj, Visc_device.h:9-90)
__device__ __forceinline__ double fit4(const double *a, double lT)
{
    double v = a[3];
#pragma unroll
    for (int i = 2; i >= 0; i--) v = v * lT + a[i];
    return v;
}

// The viscous work planes (mu, kap, D_km, the centre gradients) hold FP64
// values, or in the mixed form FP32 ones in the first half of each plane's
// allocation.  A reader sees either as doubles, so the wall flux's
// arithmetic is FP64 in both forms.
template <class T> struct VPlane {
    const T *p;
    __device__ __forceinline__ double operator[](long long i) const { return (double)p[i]; }
    __device__ __forceinline__ VPlane at(long long off) const { return VPlane{p + off}; }
};
template <int PREC> using vtype = typename std::conditional<PREC == XF_MIXED, float, double>::type;
template <int PREC> __device__ __forceinline__ VPlane<vtype<PREC>> vplane(const double *p)
{
    return VPlane<vtype<PREC>>{reinterpret_cast<const vtype<PREC> *>(p)};
}
template <int PREC> __device__ __forceinline__ vtype<PREC> *vplane_w(double *p)
{
    return reinterpret_cast<vtype<PREC> *>(p);
}

// e^x to FP32 accuracy (about 1e-7 relative) for the mixed form: the argument
// goes to base 2 in FP64, its integer part is applied exactly by the exponent
// and the fraction in [-1/2, 1/2] through the hardware base-2 exponential
__device__ __forceinline__ float exp_f32(double x)
{
    const double t = x * 0x1.71547652b82fep+0;
    const double n = rint(t);
    return ldexpf(__builtin_amdgcn_exp2f((float)(t - n)), (int)n);
}

// The mixed form's transport coefficients (the same formulas as the FP64
// path below): the fits' cubics and the sums that carry the reference's
// 1e-40 and 1e-20 guards stay FP64; the exponentials, the Wilke factors and
// their reciprocals and square roots are FP32.  About a quarter of the FP64
// path's instructions.
template <int NS>
__device__ __forceinline__ void transport_mixed(const DevThermo *__restrict__ th, const DevTransport *__restrict__ tr,
                                                Fields &f, long long id, long long N, const double *X, double C_total,
                                                double rho, double rho1, double p, double lT, int heat, int diffu)
{
    float mu[NS], rmu[NS], Xf[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        mu[k] = exp_f32(fit4(tr->fcv + 4 * k, lT));
        rmu[k] = __builtin_amdgcn_rcpf(mu[k]);
        Xf[k] = (float)X[k];
    }
    float va = 0.0f, tca = 0.0f;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        float den = 0.0f;
#pragma unroll
        for (int i = 0; i < NS; i++) {
            float phi = i == k ? tr->phiAf[k * MAXNS + i]
                               : tr->phiAf[k * MAXNS + i] * __builtin_amdgcn_sqrtf(mu[k] * rmu[i]);
            phi = (phi + 1.0f) * (phi + 1.0f) * tr->phiBcf[k * MAXNS + i];
            den += Xf[i] * phi;
        }
        const float _den = __builtin_amdgcn_rcpf(den);
        va += Xf[k] * mu[k] * _den;
        if (heat) tca += Xf[k] * exp_f32(fit4(tr->fct + 4 * k, lT)) * _den;
    }
    vplane_w<XF_MIXED>(f.mu)[id] = va;
    if (heat) vplane_w<XF_MIXED>(f.kap)[id] = tca;
    if (diffu) {
        double t2[NS];
        const float rp = __builtin_amdgcn_rcpf((float)p);
#pragma unroll
        for (int k = 0; k < NS; k++) t2[k] = 1.0e-20;
#pragma unroll
        for (int k = 0; k < NS; k++) {
#pragma unroll
            for (int i = 0; i < k; i++) {
                const double rDe = (double)__builtin_amdgcn_rcpf(exp_f32(fit4(tr->Dkj + 4 * (i * MAXNS + k), lT)) * rp);
                t2[k] += (X[i] + 1.0e-40) * rDe;
                t2[i] += (X[k] + 1.0e-40) * rDe;
            }
        }
#pragma unroll
        for (int k = 0; k < NS; k++) {
            double t1 = 0.0;
#pragma unroll
            for (int i = 0; i < NS; i++)
                if (i != k) t1 += (X[i] + 1.0e-40) * th->Wi[i];
            double D;
            if (t1 <= 0.0)
                D = (double)(exp_f32(fit4(tr->Dkj + 4 * (k * MAXNS + k), lT)) * rp);
            else
                D = t1 / t2[k] * rho1 * C_total;
            D *= 1.0e-1;
            vplane_w<XF_MIXED>(f.Dkm)[k * N + id] = (float)smax(D, 1.0e-10);
        }
    }
}

// Waves per SIMD k_transport is compiled for (1: no bound).  Unbounded, the
// FP64 form at nine species takes 194 registers and runs two waves per SIMD,
// with the FP64 divisions, square roots and exponentials waiting on each
// other; held to three waves (168 registers, 96 bytes per lane spilled) it
// takes 15 % less time, and four waves (256 bytes spilled) gain nothing more.
// At ten and eleven species a three-wave bound spills 272 and 432 bytes, and
// the mixed form already runs five waves.
template <int NS, int PREC> struct TransportOccupancy {
    static constexpr int waves = (PREC == 0 && NS == 9) ? 3 : 1;
};

// Gettransport_coeff_aver + Get_transport_coeff_aver (Visc_kernels.hpp:205-235,
// Visc_device.h:99-177) with get_xi (Mixing_device.h:31-44) on every padded
// cell; also GetStates' u, v, w (Update_device.hpp:36) and, with diffusion,
// each block's min / max of the interior mass fractions
// (ConVenction_block.hpp:462-490) into ypart[block][2 NS]
template <int NS, int PREC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TransportOccupancy<NS, PREC>::waves)))
k_transport(Geom g, const DevThermo *__restrict__ th, const DevTransport *__restrict__ tr,
            const double *__restrict__ Ug, Fields f, int heat, int diffu)
{
    constexpr int E = NS + 4;
    constexpr int NCOP = NS - 1;
    if (step_halted(f.err, ~0)) return; // block-uniform (no bit is set here)
    const long long N = g.N;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = id < N;
    // T and p, loaded ahead of the block's reduction barrier
    const double T = live ? f.T[id] : 0.0, p = live ? f.p[id] : 0.0;
    // mass fractions (Getrhoyi), velocities (GetStates) and mole fractions
    // (get_xi); the mass fractions feed the min / max and are then dropped
    double X[NS], C_total = 0.0, rho = 1.0, rho1 = 1.0;
    {
        double yi[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) yi[k] = 0.0;
        if (live) {
            double U[E];
#pragma unroll
            for (int n = 0; n < E; n++) U[n] = Ug[n * N + id];
            rho = U[0];
            rho1 = 1.0 / rho;
            cell_y<NS>(g, f, id, rho1, [&](int k) { return U[5 + k]; }, yi);
            f.u[id] = U[1] * rho1;
            f.v[id] = U[2] * rho1;
            f.w[id] = U[3] * rho1;
#pragma unroll
            for (int i = 0; i < NS; i++) {
                X[i] = yi[i] * th->_Wi[i] * 1e-3 * rho;
                C_total = C_total + X[i];
            }
            const double _C_total = 1.0 / C_total;
#pragma unroll
            for (int i = 0; i < NS; i++) X[i] = X[i] * _C_total;
        }
        if (diffu) {
            // interior cells' min / max of each mass fraction: wave shuffles,
            // then the block's four waves in LDS (other cells contribute 0.0,
            // the value the reference's reduction combines with)
            int i = 0, j = 0, kk = 0;
            if (live) cell_ijk(g, id, i, j, kk);
            const bool in = live && i >= g.Bx && i < g.Xmax - g.Bx && j >= g.By && j < g.Ymax - g.By &&
                            kk >= g.Bz && kk < g.Zmax - g.Bz;
            __shared__ double sm[4][2 * NS];
            const int wv = threadIdx.x >> 6;
#pragma unroll
            for (int k = 0; k < NS; k++) {
                double a = in ? yi[k] : 0.0, b = a;
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) {
                    a = smin(a, __shfl_xor(a, off, 64));
                    b = smax(b, __shfl_xor(b, off, 64));
                }
                if ((threadIdx.x & 63) == 0) sm[wv][k] = a, sm[wv][NS + k] = b;
            }
            __syncthreads();
            if (threadIdx.x < 2 * NS) {
                const int q = threadIdx.x;
                double v = sm[0][q];
                for (int w = 1; w < 4; w++) v = q < NS ? smin(v, sm[w][q]) : smax(v, sm[w][q]);
                f.ypart[(long long)blockIdx.x * 2 * NS + q] = v;
            }
        }
    }
    if (!live) return;
    const double lT = dlog(T);
    if constexpr (PREC == 1) {
        transport_mixed<NS>(th, tr, f, id, N, X, C_total, rho, rho1, p, lT, heat, diffu);
        return;
    }
    // Wilke mixture viscosity and conductivity (the conductivity uses the
    // viscosity's denominators, as Visc_device.h:121-134 does)
    double mu[NS], rmu[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) mu[k] = dexp(fit4(tr->fcv + 4 * k, lT)), rmu[k] = 1.0 / mu[k];
    double va = 0.0, tca = 0.0;
    // the quotients by sqrt(2) take the corrected reciprocal of that constant
    // (folded at compile time), which is the IEEE quotient bit for bit
    const double s2 = sqrt(2.0), rs2 = 1.0 / s2;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        double den = 0.0;
#pragma unroll
        for (int i = 0; i < NS; i++) {
            double phi;
            if (i == k) {
                // RN(mu_k / mu_k) = 1 wherever qdiv's correction holds (it
                // returns exactly 1 for a = b), and sqrt(1) = 1: the product
                // is phiA itself; elsewhere the IEEE quotient, as qdiv takes
                const double am = fabs(mu[k]);
                phi = tr->phiA[k * MAXNS + i];
                if (__builtin_expect(!(am >= 0x1p-500 && am <= 0x1p500), 0)) phi = phi * sqrt(ieee_div(mu[k], mu[k]));
            } else {
                phi = tr->phiA[k * MAXNS + i] * sqrt(qdiv(mu[k], mu[i], rmu[i]));
            }
            phi = qdiv((phi + 1.0) * (phi + 1.0) * 0.5, s2, rs2);
            phi = phi * tr->phiB[k * MAXNS + i];
            den = den + X[i] * phi;
        }
        const double _den = 1.0 / den;
        va = va + X[k] * mu[k] * _den;
        if (heat) tca = tca + X[k] * dexp(fit4(tr->fct + 4 * k, lT)) * _den;
    }
    f.mu[id] = va;
    if (heat) f.kap[id] = tca;
    if (diffu) {
        // mixture-averaged D_km: temp2[k] = 1e-20 + sum_{i != k, i ascending}
        // (X_i + 1e-40) / (D_ik + 1e-40), with D_ik = exp(fit(ln T)) / p.
        // The fits are symmetric (Setup copies the lower triangle from the
        // upper), so each pair's exponential serves both species.
        double t2[NS];
        const double rp = 1.0 / p;
#pragma unroll
        for (int k = 0; k < NS; k++) t2[k] = 1.0e-20;
#pragma unroll
        for (int k = 0; k < NS; k++) {
#pragma unroll
            for (int i = 0; i < k; i++) {
                const double D = qdiv(dexp(fit4(tr->Dkj + 4 * (i * MAXNS + k), lT)), p, rp);
                const double De = D + 1.0e-40, rDe = 1.0 / De;
                t2[k] += qdiv(X[i] + 1.0e-40, De, rDe);
                t2[i] += qdiv(X[k] + 1.0e-40, De, rDe);
            }
        }
#pragma unroll
        for (int k = 0; k < NS; k++) {
            double t1 = 0.0;
#pragma unroll
            for (int i = 0; i < NS; i++)
                if (i != k) t1 += (X[i] + 1.0e-40) * th->Wi[i];
            double D;
            if (t1 <= 0.0) // sycl::step(ceil(temp1), 0)
                D = dexp(fit4(tr->Dkj + 4 * (k * MAXNS + k), lT)) / p;
            else
                D = qdiv(t1 / t2[k], rho, rho1) * C_total;
            D *= 1.0e-1;
            f.Dkm[k * N + id] = smax(D, 1.0e-10);
        }
    }
}

// the per-block partials of k_transport into red[RED_YMIN + k] (min) and
// red[RED_YMAX + k] (max), each combined with 0.0 as the reference's
// sycl::reduction combines with the stored yi_min = yi_max = 0; one block per
// quantity (2 NS blocks)
template <int NS>
__global__ void __launch_bounds__(1024) k_yreduce(const double *__restrict__ ypart, int nblocks, double *red)
{
    __shared__ double s[1024];
    const int q = blockIdx.x;
    const bool mx = q >= NS;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 1024) {
        const double x = ypart[(long long)b * 2 * NS + q];
        v = mx ? smax(v, x) : smin(v, x);
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = mx ? smax(s[threadIdx.x], s[threadIdx.x + w]) : smin(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) red[mx ? RED_YMAX + q - NS : RED_YMIN + q] = s[0];
}

// fourth-order centre derivatives of u, v, w along direction a at cell id
// (GetInnerCellCenterDerivativeKernel, Fourth_Order/Visc_Order_kernels.hpp:16-71):
// D[0..2] = du/da, dv/da, dw/da with the reference's DimX_t / DimY_t / DimZ_t
// factors in its order
__device__ __forceinline__ void centre_deriv(const Geom &g, const Fields &f, long long id, int a, double *D)
{
    const long long st = a == 0 ? 1 : (a == 1 ? (long long)g.Xmax : (long long)g.Xmax * g.Ymax);
    const long long m1 = id - st, m2 = id - 2 * st, p1 = id + st, p2 = id + 2 * st;
    const double DX = g.DimX, DY = g.DimY, DZ = g.DimZ, tw = 1.0 / 12.0;
    const double _d = a == 0 ? g._dx : (a == 1 ? g._dy : g._dz);
    const double Da = a == 0 ? DX : (a == 1 ? DY : DZ);
    const double *u = f.u, *v = f.v, *w = f.w;
    const double du = (8.0 * (u[p1] - u[m1]) - (u[p2] - u[m2])) * _d * tw;
    const double dv = (8.0 * (v[p1] - v[m1]) - (v[p2] - v[m2])) * _d * tw;
    const double dw = (8.0 * (w[p1] - w[m1]) - (w[p2] - w[m2])) * _d * tw;
    if (a == 0) D[0] = du * DX, D[1] = dv * DX * DY, D[2] = dw * DX * DZ;
    else if (a == 1) D[0] = du * DY * DX, D[1] = dv * DY, D[2] = dw * DY * DZ;
    else D[0] = du * DZ * DX, D[1] = dv * DZ * DY, D[2] = dw * DZ;
    (void)Da;
}

// GetCellCenterDerivative (Visc_block.hpp:8-84) in one pass: the centre
// derivatives on i in [Bwidth_X - 2, Xmax - Bwidth_X + 1] per active
// direction (index 0 of an inactive one; a derivative along an inactive
// direction is +0), and on the ghost cells of that range the
// CenterDerivativeBCKernel{X,Y,Z} rules (Visc_kernels.hpp:7-203) for the four
// components each direction's wall flux reads across the boundary: the value
// at the mirror / periodic / outflow cell (negated for a no-slip wall, 0 for
// inflow, unchanged for the other codes) is recomputed there with the same
// operations.  Ghost cells in two directions at once (edges and corners) are
// read by no wall flux.  The z no-slip wall uses the mirror offset - k of the
// x and y kernels (the reference's k - offset indexes outside the array).
template <int PREC>
__global__ void __launch_bounds__(256) k_vel_deriv(Geom g, Fields f, int i0, int nx, int j0, int ny, int k0,
                                                          int nz, int bc0, int bc1, int bc2, int bc3, int bc4, int bc5)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nx * ny * nz) return;
    if (step_halted(f.err, ~0)) return;
    const int c[3] = {i0 + (int)(t % nx), j0 + (int)((t / nx) % ny), k0 + (int)(t / ((long long)nx * ny))};
    const long long N = g.N;
    const long long sy = g.Xmax, sz = (long long)g.Xmax * g.Ymax;
    const long long id = sz * c[2] + sy * c[1] + c[0];
    const int dims[3] = {g.DimX, g.DimY, g.DimZ};
    const int Bw[3] = {g.Bx, g.By, g.Bz}, inner[3] = {g.X_inner, g.Y_inner, g.Z_inner};
    const int Lm[3] = {g.Xmax, g.Ymax, g.Zmax};
    const int bcs[6] = {bc0, bc1, bc2, bc3, bc4, bc5};
    double D[9];
#pragma unroll
    for (int q = 0; q < 9; q++) D[q] = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (dims[a]) centre_deriv(g, f, id, a, D + 3 * a);
    // components read across a boundary of direction d (Vde order
    // ducx dvcx dwcx ducy dvcy dwcy ducz dvcz dwcz)
    const int comp[3][4] = {{DUCY, DUCZ, DVCY, DWCZ}, {DVCX, DVCZ, DUCX, DWCZ}, {DWCX, DWCY, DUCX, DVCY}};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (!dims[d]) continue;
        const int s = c[d], B = Bw[d];
        const bool lo = s < B, hi = s >= B + inner[d];
        if (!lo && !hi) continue;
        const int bc = bcs[2 * d + (hi ? 1 : 0)];
        const int mirror_offset = lo ? 0 : inner[d], index_inner = lo ? B : Lm[d] - B - 1, sign = lo ? 1 : -1;
        int tt;
        double fac = 1.0;
        if (bc == XF_BC_SYMMETRY) tt = 2 * (B + mirror_offset) - 1 - s;
        else if (bc == XF_BC_PERIODIC) tt = s + sign * inner[d];
        else if (bc == XF_BC_OUTFLOW) tt = index_inner;
        else if (bc == XF_BC_NSLIPWALL) tt = 2 * (B + mirror_offset) - 1 - s, fac = -1.0;
        else if (bc == XF_BC_INFLOW) tt = -1;
        else continue; // viscWall, slipWall, innerBlock: the centre value stays
        if (tt < 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) D[comp[d][q]] = 0.0;
            continue;
        }
        int ct[3] = {c[0], c[1], c[2]};
        ct[d] = tt;
        const long long tid = sz * ct[2] + sy * ct[1] + ct[0];
        double T9[9];
#pragma unroll
        for (int q = 0; q < 9; q++) T9[q] = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++)
            if (dims[a] && a != d) centre_deriv(g, f, tid, a, T9 + 3 * a);
#pragma unroll
        for (int q = 0; q < 4; q++) D[comp[d][q]] = fac < 0.0 ? -T9[comp[d][q]] : T9[comp[d][q]];
    }
#pragma unroll
    for (int q = 0; q < 9; q++) vplane_w<PREC>(f.Vde)[q * N + id] = (vtype<PREC>)D[q];
}

// GetWallViscousFlux{X,Y,Z} with MARCO_PREVISCFLUX / MARCO_VISCFLUX
// (Fourth_Order/Visc_Order_kernels.hpp:73-342, Fourth_Order/Flux_discrete.h:4-93):
// the viscous flux F[E] of the face right of cell id along DIR, from the
// transport coefficients, centre gradients and limits of this GetLU
// uat(c, n) is U_n and r1at(c) is 1 / U_0 (as GetStates forms it) of the
// stencil cell c = 0..3 (i - 1, i, i + 1, i + 2 along DIR): from memory, or
// from a sweep's staged tile.  The velocities are formed from them as
// GetStates (and k_transport) form u, v, w.
template <int NS, int DIR, int PREC, class UAt, class R1At>
__device__ __forceinline__ void visc_face(const Geom &g, const Fields &f, long long id, UAt uat, R1At r1at, double *F)
{
    constexpr int E = NS + 4;
    const int heat = g.visc_heat, diffu = g.visc_diffu;
    const double Yil_limiter = g.Yil_limiter, Dim_limiter = g.Dim_limiter;
    const long long N = g.N;
    const long long sy = g.Xmax, sz = (long long)g.Xmax * g.Ymax;
    const long long st = DIR == 0 ? 1 : (DIR == 1 ? sy : sz);
    const long long m1 = id - st, p1 = id + st, p2 = id + 2 * st;
    const double DX = g.DimX, DY = g.DimY, DZ = g.DimZ;
    const double _dl = DIR == 0 ? g._dx : (DIR == 1 ? g._dy : g._dz);
    const double OT = 1.0 / 3.0, SX = 1.0 / 16.0, TF = 1.0 / 24.0;
    // u, v, w of the four stencil cells (GetStates' U_{1..3} / rho)
    double uu[4], vv[4], ww[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const double r1 = r1at(c);
        uu[c] = uat(c, 1) * r1, vv[c] = uat(c, 2) * r1, ww[c] = uat(c, 3) * r1;
    }
    const auto va = vplane<PREC>(f.mu), Vd = vplane<PREC>(f.Vde);
    auto half = [&](const auto &a) { return (9.0 * (a[p1] + a[id]) - (a[p2] + a[m1])) * SX; };
    auto grad = [&](const auto &a) { return (27.0 * (a[p1] - a[id]) - (a[p2] - a[m1])) * _dl * TF; };
    auto pair2 = [&](const auto &a, const auto &b) {
        return 9.0 * (a[p1] + a[id]) - (a[p2] + a[m1]) + 9.0 * (b[p1] + b[id]) - (b[p2] + b[m1]);
    };
    const double mue = half(va);
    const double lamada = -2.0 * OT * mue;
    double f_x, f_y, f_z, u_hlf, v_hlf, w_hlf;
    if (DIR == 0) {
        const auto Dvcy = Vd.at(DVCY * N), Dwcz = Vd.at(DWCZ * N), Ducy = Vd.at(DUCY * N), Ducz = Vd.at(DUCZ * N);
        f_x = (2.0 * mue + lamada) * (27.0 * (uu[2] - uu[1]) - (uu[3] - uu[0])) * _dl * TF;
        f_x += lamada * pair2(Dvcy, Dwcz) * SX;
        f_y = mue * (27.0 * (vv[2] - vv[1]) - (vv[3] - vv[0])) * _dl * TF * DY;
        f_y += mue * (9.0 * (Ducy[p1] + Ducy[id]) - (Ducy[p2] + Ducy[m1])) * SX * DY;
        f_z = mue * (27.0 * (ww[2] - ww[1]) - (ww[3] - ww[0])) * _dl * TF * DZ;
        f_z += mue * (9.0 * (Ducz[p1] + Ducz[id]) - (Ducz[p2] + Ducz[m1])) * SX * DZ;
        u_hlf = (9.0 * (uu[2] + uu[1]) - (uu[3] + uu[0])) * SX;
        v_hlf = (9.0 * (vv[2] + vv[1]) - (vv[3] + vv[0])) * SX * DY;
        w_hlf = (9.0 * (ww[2] + ww[1]) - (ww[3] + ww[0])) * SX * DZ;
    } else if (DIR == 1) {
        const auto Dvcx = Vd.at(DVCX * N), Dvcz = Vd.at(DVCZ * N), Ducx = Vd.at(DUCX * N), Dwcz = Vd.at(DWCZ * N);
        f_x = mue * (27.0 * (uu[2] - uu[1]) - (uu[3] - uu[0])) * _dl * TF * DX;
        f_x += mue * (9.0 * (Dvcx[p1] + Dvcx[id]) - (Dvcx[p2] + Dvcx[m1])) * SX * DX;
        f_y = (2.0 * mue + lamada) * (27.0 * (vv[2] - vv[1]) - (vv[3] - vv[0])) * _dl * TF;
        f_y += lamada * pair2(Ducx, Dwcz) * SX;
        f_z = mue * (27.0 * (ww[2] - ww[1]) - (ww[3] - ww[0])) * _dl * TF * DZ;
        f_z += mue * (9.0 * (Dvcz[p1] + Dvcz[id]) - (Dvcz[p2] + Dvcz[m1])) * SX * DZ;
        u_hlf = (9.0 * (uu[2] + uu[1]) - (uu[3] + uu[0])) * SX * DX;
        v_hlf = (9.0 * (vv[2] + vv[1]) - (vv[3] + vv[0])) * SX;
        w_hlf = (9.0 * (ww[2] + ww[1]) - (ww[3] + ww[0])) * SX * DZ;
    } else {
        const auto Dwcx = Vd.at(DWCX * N), Dwcy = Vd.at(DWCY * N), Ducx = Vd.at(DUCX * N), Dvcy = Vd.at(DVCY * N);
        f_x = mue * (27.0 * (uu[2] - uu[1]) - (uu[3] - uu[0])) * _dl * TF * DX;
        f_x += mue * (9.0 * (Dwcx[p1] + Dwcx[id]) - (Dwcx[p2] + Dwcx[m1])) * SX * DX;
        f_y = mue * (27.0 * (vv[2] - vv[1]) - (vv[3] - vv[0])) * _dl * TF * DY;
        f_y += mue * (9.0 * (Dwcy[p1] + Dwcy[id]) - (Dwcy[p2] + Dwcy[m1])) * SX * DY;
        f_z = (2.0 * mue + lamada) * (27.0 * (ww[2] - ww[1]) - (ww[3] - ww[0])) * _dl * TF;
        f_z += lamada * pair2(Ducx, Dvcy) * SX;
        u_hlf = (9.0 * (uu[2] + uu[1]) - (uu[3] + uu[0])) * SX * DX;
        v_hlf = (9.0 * (vv[2] + vv[1]) - (vv[3] + vv[0])) * SX * DY;
        w_hlf = (9.0 * (ww[2] + ww[1]) - (ww[3] + ww[0])) * SX;
    }
    F[0] = 0.0;
    F[1] = f_x;
    F[2] = f_y;
    F[3] = f_z;
    F[4] = f_x * u_hlf + f_y * v_hlf + f_z * w_hlf;
    if (heat) {
        double kk = half(vplane<PREC>(f.kap));
        kk *= grad(f.T);
        F[4] += kk;
    }
    if (diffu) {
        // mass fractions of the four cells, as Getrhoyi forms them
        const long long cid[4] = {m1, id, p1, p2};
        double Y[4][NS];
        double rw[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            rw[c] = uat(c, 0);
            cell_y<NS>(g, f, cid[c], r1at(c), [&](int ii) { return uat(c, 5 + ii); }, Y[c]);
        }
        const double rho_wall = (9.0 * (rw[2] + rw[1]) - (rw[3] + rw[0])) * SX;
        double CorrectTerm = 0.0, Dim_Yil = 1.0e-20, Yi_wall[NS];
#pragma unroll
        for (int l = 0; l < NS; l++) {
            // limits of ConVenction_block.hpp:462-503 (single-process build)
            double ymax = f.red[RED_YMAX + l];
            const double ymin = f.red[RED_YMIN + l];
            double dmax = 0.0;
            ymax -= ymin;
            ymax *= Yil_limiter;
            dmax *= Dim_limiter * ymax;
            const double *hl = f.hi + l * N;
            const auto Dl = vplane<PREC>(f.Dkm).at(l * N);
            const double hi_wall = half(hl);
            const double Dim_wall = half(Dl);
            const double Yil_wall =
                smin(smax((27.0 * (Y[2][l] - Y[1][l]) - (Y[3][l] - Y[0][l])) * _dl * TF, -ymax), ymax);
            Yi_wall[l] = smin(smax((9.0 * (Y[2][l] + Y[1][l]) - (Y[3][l] + Y[0][l])) * SX, 1.0E-20), 1.0);
            Dim_Yil = smin(smax(Dim_wall * Yil_wall, -dmax), dmax);
            CorrectTerm += Dim_Yil;
            F[4] += rho_wall * hi_wall * Dim_Yil;
        }
        CorrectTerm *= rho_wall;
#pragma unroll
        for (int q = 5; q < E; q++) F[q] = rho_wall * Dim_Yil - Yi_wall[q - 5] * CorrectTerm;
    } else {
#pragma unroll
        for (int q = 5; q < E; q++) F[q] = 0.0;
    }
}

// GetWallViscousFlux{X,Y,Z} as a kernel of its own, on the faces right of
// cells i in [B - 1, B + inner - 1] along DIR (interior on the other axes):
// writes the flux the sweep subtracts, Fv[n * N + id].  With four or more
// species the sweeps compute the same faces themselves unless
// Geom::visc_fused is 0 (visc_in_sweep).
template <int NS, int DIR, int PREC>
__global__ void __launch_bounds__(256) k_visc_flux(Geom g, const double *__restrict__ Ug, Fields f)
{
    constexpr int E = NS + 4;
    if (step_halted(f.err, ~0)) return;
    const int nx = g.X_inner + (DIR == 0), ny = g.Y_inner + (DIR == 1), nz = g.Z_inner + (DIR == 2);
    // Block order.  x needs no reorder.  y faces: XCD-aware order, so the
    // stencil rows that neighbouring blocks share meet in one L2 (2.95 ->
    // 2.13 ms at 256^3).  z faces: the grid is (chunks of 256 faces of a
    // plane) x (face planes), and each XCD takes a contiguous run of it with
    // the plane index fastest.  The four blocks whose stencils read the same
    // row of every input then run together on one XCD and share its L2; in
    // plane-major order they ran a plane apart, and every row was fetched
    // about four times.
    int bx = (int)blockIdx.x, by_, bz_;
    if (DIR != 0) xcd_tile(bx, by_, bz_);
    int i, j, k;
    if (DIR == 2) {
        const long long P = (long long)nx * ny;
        const long long tp = (long long)(bx / nz) * blockDim.x + threadIdx.x;
        if (tp >= P) return;
        i = g.Bx + (int)(tp % nx);
        j = g.By + (int)(tp / nx);
        k = g.Bz - 1 + bx % nz;
    } else {
        const long long t = (long long)bx * blockDim.x + threadIdx.x;
        if (t >= (long long)nx * ny * nz) return;
        i = g.Bx - (DIR == 0) + (int)(t % nx);
        j = g.By - (DIR == 1) + (int)((t / nx) % ny);
        k = g.Bz + (int)(t / ((long long)nx * ny));
    }
    const long long id = (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
    const long long st = DIR == 0 ? 1 : (DIR == 1 ? (long long)g.Xmax : (long long)g.Xmax * g.Ymax);
    double F[E];
    visc_face<NS, DIR, PREC>(
        g, f, id, [&](int c, int n) { return Ug[n * g.N + id + (c - 1) * st]; },
        [&](int c) { return 1.0 / Ug[id + (c - 1) * st]; }, F);
#pragma unroll
    for (int q = 0; q < E; q++) f.Fv[q * g.N + id] = F[q];
}

} // namespace xf
