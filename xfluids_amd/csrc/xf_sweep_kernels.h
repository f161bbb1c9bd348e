// Characteristic WENO face fluxes and flux divergence, one sweep direction
// per launch, LDS-staged stencil pencils.
//
// Reference: ReconstructFlux{X,Y,Z} (FDM_Method/Reconstruction_kernels.hpp:8-199)
// with MARCO_ROE (src/include/global_marco.h:26-34), MARCO_COPC2 /
// ReconstructSoundSpeed (positive-definite_eigen/Utils_device.hpp:42-140),
// MARCO_NOCOPC2 (Eigen_callback.h:87-91), MARCO_PREEIGEN (Eigen_callback.h:102-106),
// RoeAverageLeft/Right_{x,y,z} (Eigen_matrix.hpp:7-455), MARCO_FLUXWALL_WENO5
// (Eigen_callback.h:179-230), weno5old_GPU / weno5Z_{P,M}
// (schemes/WENO5s_schemes.hpp), UpdateFluidLU (Reconstruction_kernels.hpp:201-234).
//
// Three arithmetic forms share the kernel (template PREC):
//  * XF_PLAIN_FP64: every product and sum rounded in the reference's
//    expression order (uf + U*l, 0.5 * (ff + a*uf), w * poly / 6, fluxl += f*r);
//  * XF_FUSED_FP64: the same FP64 operations with each multiply-add pair of
//    the characteristic projections, the GLF split, the back-projection and
//    the WENO5-Z indicators and candidate polynomials rounded once (fma);
//  * XF_MIXED: FP32 projections and WENO weights, FP64 elsewhere.
//
// GPU structure (FP64 forms):
//  * a block owns a tile of TS faces along the sweep axis x TW pencils; the
//    TS+5 cells of each pencil are staged once into LDS as conserved U and
//    physical flux F (F built from U, p and velocity exactly as GetPhysFlux);
//  * one lane per face: Roe average, sound speed, then for each of the E
//    characteristic fields the left row, the 6-point projections of U and F
//    that WENO5 reads (points i-2..i+3; the reference also projects i-3 and
//    i+4, which WENO5 never uses), the GLF split, WENO5 and the accumulation
//    with the right column;
//  * structurally-zero eigenvector entries are skipped: the reference adds
//    +/-0 products to an accumulator that starts at +0.0 and is never -0.0
//    afterwards, so skipping them leaves every sum bit-identical;
//  * the TS face fluxes go back to LDS and the TS-1 interior cells of the
//    tile form LU = (F_{i-1/2} - F_{i+1/2}) / dx in the reference's direction
//    order (x, then y, then z), optionally fused with the SSP-RK3 stage.
#pragma once
#include "xf_fluid_kernels.h"
#include "xf_visc_kernels.h"

namespace xf {

// Tile shapes of a 256-lane sweep block: TS faces along the sweep axis by
// TW = 256 / TS pencils.  Longer tiles stage fewer halo cells per interior
// cell; wider tiles read longer contiguous row segments (y and z sweeps).
// x: 0 = 64 x 4, 1 = 32 x 8.  y, z: 0 = 16 x 16, 1 = 32 x 8, 2 = 8 x 32.
template <int DIR> struct SweepShapes {
    static constexpr int count = DIR == 0 ? 2 : 3;
};
template <int DIR, int E, int SHAPE> struct SweepShape {
    static constexpr int TS = DIR == 0 ? (SHAPE == 0 ? 64 : 32) : (SHAPE == 0 ? 16 : (SHAPE == 1 ? 32 : 8));
    static constexpr int TW = 256 / TS;
    static constexpr int NC = (TS + 5) * TW; // staged cells
    static_assert(NC >= 256, "the flux-divergence pass reuses the staging area for 256 face fluxes");
};

// staged-cell LDS index: lanes adjacent in the wave hit adjacent words
template <int DIR, int TS, int TW>
__device__ __forceinline__ int lds_ix(int ls, int tcol)
{
    return DIR == 0 ? tcol * (TS + 5) + ls : ls * TW + tcol;
}

// WENO5-JS body (weno5old_BODY, USE_DOUBLE branch)
__device__ __forceinline__ double weno5js_body(double v1, double v2, double v3, double v4, double v5)
{
    double a1, a2, a3;
    const double dtwo = 2.0, dtre = 3.0;
    a1 = v1 - dtwo * v2 + v3;
    double s1 = 13.0 * a1 * a1;
    a1 = v1 - 4.0 * v2 + dtre * v3;
    s1 += dtre * a1 * a1;
    a1 = v2 - dtwo * v3 + v4;
    double s2 = 13.0 * a1 * a1;
    a1 = v2 - v4;
    s2 += dtre * a1 * a1;
    a1 = v3 - dtwo * v4 + v5;
    double s3 = 13.0 * a1 * a1;
    a1 = dtre * v3 - 4.0 * v4 + v5;
    s3 += dtre * a1 * a1;
    const double tol = 1.0E-6;
    s1 += tol, s2 += tol, s3 += tol;
    a1 = 0.1 * s2 * s2 * s3 * s3;
    a2 = 0.6 * s1 * s1 * s3 * s3;
    a3 = 0.3 * s1 * s1 * s2 * s2;
    const double tw1 = 1.0 / (a1 + a2 + a3);
    a1 = a1 * tw1, a2 = a2 * tw1, a3 = a3 * tw1;
    s1 = a1 * (dtwo * v1 - 7.0 * v2 + 11.0 * v3);
    s2 = a2 * (-v2 + 5.0 * v3 + dtwo * v4);
    s3 = a3 * (dtwo * v3 + 5.0 * v4 - v5);
    return (s1 + s2 + s3);
}

// a*b + c as the form computes it: one rounding (fused form) or the
// reference's two, product first
template <bool FUSED> __device__ __forceinline__ double madd(double a, double b, double c)
{
    if (FUSED) return fma(a, b, c);
    return c + a * b;
}

// Correctly rounded quotients without the full division sequence.
// Markstein (IBM J. Res. Dev. 34, 1990): if y = RN(1/b) and q0 = RN(a y) is
// within one ulp of a/b, the remainder r = a - b q0 is exact (one fma) and
// RN(q0 + r y) = RN(a/b).  Valid while a - b q0 does not underflow.
__device__ __forceinline__ double quot_from_recip(double a, double b, double y)
{
    const double q0 = a * y;
    return fma(fma(-b, q0, a), y, q0);
}
// RN(x / 6) for x = 0 (either sign) or 2^-960 <= |x| < inf; other x are
// flagged in `slow` and the caller takes the IEEE division for them
__device__ __forceinline__ double div6_fast(double x, bool &slow)
{
    const double q = quot_from_recip(x, 6.0, 1.0 / 6.0);
    slow = slow || (x != 0.0 && !(fabs(x) >= 0x1p-960 && fabs(x) <= 0x1p1000));
    // x / 6 carries the sign of x, zeros included (the corrected quotient of
    // -0 comes out +0): one bit insert instead of a compare and two selects
    return copysign(q, x);
}

// WENO5-Z body (weno5Z_P / weno5Z_M).  The three weights share the divisor
// S = a1 + a2 + a3 >= 1 and have numerators >= 0.1, so they come from one
// correctly rounded reciprocal RN(1/S); the divisions by 6 use div6_fast.  Every
// quotient is the IEEE-rounded one, bit for bit.  The plain form rounds the
// indicators and candidate polynomials as the reference writes them (each
// indicator difference, written twice there, is formed once: both copies
// round alike); the fused form rounds each multiply-add pair once.
template <bool FUSED>
__device__ __forceinline__ double weno5z_body(double v1, double v2, double v3, double v4, double v5)
{
    const double epsilon = 1.0e-6;
    double s1, s2, s3, p1, p2, p3;
    if (FUSED) {
        const double d1 = fma(-2.0, v2, v1) + v3, e1 = fma(3.0, v3, fma(-4.0, v2, v1));
        const double d2 = fma(-2.0, v3, v2) + v4, e2 = v2 - v4;
        const double d3 = fma(-2.0, v4, v3) + v5, e3 = fma(-4.0, v4, 3.0 * v3) + v5;
        s1 = fma(13.0 * d1, d1, (3.0 * e1) * e1);
        s2 = fma(13.0 * d2, d2, (3.0 * e2) * e2);
        s3 = fma(13.0 * d3, d3, (3.0 * e3) * e3);
        p1 = fma(11.0, v3, fma(-7.0, v2, 2.0 * v1));
        p2 = fma(2.0, v4, fma(5.0, v3, -v2));
        p3 = fma(5.0, v4, 2.0 * v3) - v5;
    } else {
        const double d1 = v1 - 2.0 * v2 + v3, e1 = v1 - 4.0 * v2 + 3.0 * v3;
        const double d2 = v2 - 2.0 * v3 + v4, e2 = v2 - v4;
        const double d3 = v3 - 2.0 * v4 + v5, e3 = 3.0 * v3 - 4.0 * v4 + v5;
        s1 = 13.0 * d1 * d1 + 3.0 * e1 * e1;
        s2 = 13.0 * d2 * d2 + 3.0 * e2 * e2;
        s3 = 13.0 * d3 * d3 + 3.0 * e3 * e3;
        p1 = 2.0 * v1 - 7.0 * v2 + 11.0 * v3;
        p2 = -v2 + 5.0 * v3 + 2.0 * v4;
        p3 = 2.0 * v3 + 5.0 * v4 - v5;
    }
    const double s5 = fabs(s1 - s3);
    const double a1 = 0.1 * (1.0 + s5 / (s1 + epsilon));
    const double a2 = 0.6 * (1.0 + s5 / (s2 + epsilon));
    const double a3 = 0.3 * (1.0 + s5 / (s3 + epsilon));
    const double S = a1 + a2 + a3;
    const double rS = 1.0 / S;
    const double w1 = quot_from_recip(a1, S, rS);
    const double w2 = quot_from_recip(a2, S, rS);
    const double w3 = quot_from_recip(a3, S, rS);
    const double x1 = w1 * p1;
    const double x2 = w2 * p2;
    const double x3 = w3 * p3;
    bool slow = false;
    double q1 = div6_fast(x1, slow), q2 = div6_fast(x2, slow), q3 = div6_fast(x3, slow);
    if (slow) q1 = x1 / 6.0, q2 = x2 / 6.0, q3 = x3 / 6.0;
    return q1 + q2 + q3;
}

// WENOCU6_BODYGPU (schemes/WENO6s_schemes.hpp:5-51): three upwind and one
// downwind candidate, the sixth-order indicator s6 from tau61..tau65, and
// weights 0.05 / 0.45 / 0.45 / 0.05 times (20 + s5 / (s_k + epsilon))
__device__ __forceinline__ double wenocu6_body(double v1, double v2, double v3, double v4, double v5, double v6,
                                               double epsilon)
{
    const double ohtz = 1.0 / 120.0, ohff = 1.0 / 144.0, ftss = 1.0 / 5760.0, sxtn = 1.0 / 16.0, twfr = 1.0 / 24.0;
    const double six = 1.0 / 6.0;
    const double wa2a2 = 13.0 / 3.0, wa3a3 = 3129.0 / 80.0, wa4a4 = 87617.0 / 140.0, wa3a5 = 14127.0 / 224.0,
                 wa5a5 = 252337135.0 / 16128.0;
    const double s11 = v1 - 2.0 * v2 + v3;
    const double s12 = v1 - 4.0 * v2 + 3.0 * v3;
    const double s1 = 13.0 * s11 * s11 + 3.0 * s12 * s12;
    const double s21 = v2 - 2.0 * v3 + v4;
    const double s22 = v2 - v4;
    const double s2 = 13.0 * s21 * s21 + 3.0 * s22 * s22;
    const double s31 = v3 - 2.0 * v4 + v5;
    const double s32 = 3.0 * v3 - 4.0 * v4 + v5;
    const double s3 = 13.0 * s31 * s31 + 3.0 * s32 * s32;
    const double tau61 = (259.0 * v6 - 1895.0 * v5 + 6670.0 * v4 - 2590.0 * v3 - 2785.0 * v2 + 341.0 * v1) * ftss;
    const double tau62 = -(v5 - 12.0 * v4 + 22.0 * v3 - 12.0 * v2 + v1) * sxtn;
    const double tau63 = -(7.0 * v6 - 47.0 * v5 + 94.0 * v4 - 70.0 * v3 + 11.0 * v2 + 5.0 * v1) * ohff;
    const double tau64 = (v5 - 4.0 * v4 + 6.0 * v3 - 4.0 * v2 + v1) * twfr;
    const double tau65 = -(-v6 + 5.0 * v5 - 10.0 * v4 + 10.0 * v3 - 5.0 * v2 + v1) * ohtz;
    const double s6 = (tau61 * tau61 * 1.0 + tau62 * tau62 * wa2a2 + tau61 * tau63 * 0.5 + tau63 * tau63 * wa3a3 +
                       tau62 * tau64 * 4.2 + tau61 * tau65 * 0.125 + tau64 * tau64 * wa4a4 + tau63 * tau65 * wa3a5 +
                       tau65 * tau65 * wa5a5) * 12.0;
    const double s55 = (s1 + s3 + 4.0 * s2) * six;
    const double s5 = fabs(s6 - s55);
    const double r0 = 20.0;
    const double a1 = 0.05 * (r0 + s5 / (s1 + epsilon));
    const double a2 = 0.45 * (r0 + s5 / (s2 + epsilon));
    const double a3 = 0.45 * (r0 + s5 / (s3 + epsilon));
    const double a4 = 0.05 * (r0 + s5 / (s6 + epsilon));
    const double tw1 = 1.0 / (a1 + a2 + a3 + a4);
    const double w1 = a1 * tw1, w2 = a2 * tw1, w3 = a3 * tw1, w4 = a4 * tw1;
    double temp = 0.0;
    temp += w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3);
    temp += w2 * (-v2 + 5.0 * v3 + 2.0 * v4);
    temp += w3 * (2.0 * v3 + 5.0 * v4 - v5);
    temp += w4 * (11.0 * v4 - 7.0 * v5 + 2.0 * v6);
    return temp;
}

// pp, mm hold the split values of points i-2..i+3 in [1..6] (WENO5 reads
// pp[1..5] and mm[2..6]; WENO-CU6 reads all six on both sides)
template <int SCHEME, bool FUSED>
__device__ __forceinline__ double weno_face(const double *pp, const double *mm, double delta)
{
    if (SCHEME == XF_WENOCU6) {
        const double epsilon = 1.e-8 * delta * delta;
        const double temf = wenocu6_body(pp[1], pp[2], pp[3], pp[4], pp[5], pp[6], epsilon);
        const double temm = wenocu6_body(mm[6], mm[5], mm[4], mm[3], mm[2], mm[1], epsilon);
        return (temf + temm) * (1.0 / 6.0);
    }
    if (SCHEME == XF_WENO5Z)
        return weno5z_body<FUSED>(pp[1], pp[2], pp[3], pp[4], pp[5]) +
               weno5z_body<FUSED>(mm[6], mm[5], mm[4], mm[3], mm[2]);
    const double six = 1.0 / 6.0;
    const double temf = weno5js_body(pp[1], pp[2], pp[3], pp[4], pp[5]);
    const double temm = weno5js_body(mm[6], mm[5], mm[4], mm[3], mm[2]);
    return (temf + temm) * six;
}

// ---- mixed-precision form: FP32 WENO5 ------------------------------------
// The smoothness indicators are evaluated on the stencil divided by its
// largest magnitude m, where each is at most ~400, and the reference's
// epsilon (1e-6 in its 12-beta units) becomes 1e-6 / m^2: the nonlinear
// weights are those of the FP64 definition up to FP32 rounding, with no
// overflow for any data magnitude.  WENO5-JS weights are normalised through
// the ratios (s_min / s_k)^2; WENO5-Z keeps tau5 = |beta_1 - beta_3|.  The
// weights are normalised before they multiply the candidate values.
// Divisions use the hardware reciprocal (1 ulp).  A stencil below 1e-30 in
// magnitude takes the ideal weights, as the FP64 form does when epsilon
// dominates every indicator.
__device__ __forceinline__ float rcp_f32(float x) { return __builtin_amdgcn_rcpf(x); }

template <int SCHEME>
__device__ __forceinline__ float weno5_body_f32(float v1, float v2, float v3, float v4, float v5)
{
    const float q1 = 2.0f * v1 - 7.0f * v2 + 11.0f * v3;
    const float q2 = -v2 + 5.0f * v3 + 2.0f * v4;
    const float q3 = 2.0f * v3 + 5.0f * v4 - v5;
    const float m = fmaxf(fmaxf(fmaxf(fabsf(v1), fabsf(v2)), fmaxf(fabsf(v3), fabsf(v4))), fabsf(v5));
    if (!(m > 1e-30f)) return (0.1f * q1 + 0.6f * q2 + 0.3f * q3) * (1.0f / 6.0f);
    const float inv = rcp_f32(m);
    const float u1 = v1 * inv, u2 = v2 * inv, u3 = v3 * inv, u4 = v4 * inv, u5 = v5 * inv;
    float t = u1 - 2.0f * u2 + u3, r = u1 - 4.0f * u2 + 3.0f * u3;
    float s1 = 13.0f * t * t + 3.0f * r * r;
    t = u2 - 2.0f * u3 + u4, r = u2 - u4;
    float s2 = 13.0f * t * t + 3.0f * r * r;
    t = u3 - 2.0f * u4 + u5, r = 3.0f * u3 - 4.0f * u4 + u5;
    float s3 = 13.0f * t * t + 3.0f * r * r;
    const float eps = fmaxf(fminf(1e-6f * inv * inv, 1e30f), 1e-30f);
    float a1, a2, a3;
    if (SCHEME == XF_WENO5Z) {
        const float tau = fabsf(s1 - s3);
        a1 = 0.1f * (1.0f + tau * rcp_f32(s1 + eps));
        a2 = 0.6f * (1.0f + tau * rcp_f32(s2 + eps));
        a3 = 0.3f * (1.0f + tau * rcp_f32(s3 + eps));
    } else {
        s1 += eps, s2 += eps, s3 += eps;
        const float smin = fminf(fminf(s1, s2), s3);
        const float r1 = smin * rcp_f32(s1), r2 = smin * rcp_f32(s2), r3 = smin * rcp_f32(s3);
        a1 = 0.1f * r1 * r1, a2 = 0.6f * r2 * r2, a3 = 0.3f * r3 * r3;
    }
    // normalised weights first: a_k reaches ~1e32 where an indicator is 0
    // and epsilon is small, and a_k * q_k would overflow FP32 at the
    // magnitudes of strong shocks
    const float ia = rcp_f32(a1 + a2 + a3);
    return ((a1 * ia) * q1 + (a2 * ia) * q2 + (a3 * ia) * q3) * (1.0f / 6.0f);
}

// WENO-CU6 in FP32 with the same rescaling: every indicator (s1..s3, s6)
// is quadratic in the data, so epsilon = 1e-8 dl^2 becomes 1e-8 dl^2 / m^2
__device__ __forceinline__ float wenocu6_body_f32(float v1, float v2, float v3, float v4, float v5, float v6,
                                                  float eps_abs)
{
    const float c1 = 2.0f * v1 - 7.0f * v2 + 11.0f * v3;
    const float c2 = -v2 + 5.0f * v3 + 2.0f * v4;
    const float c3 = 2.0f * v3 + 5.0f * v4 - v5;
    const float c4 = 11.0f * v4 - 7.0f * v5 + 2.0f * v6;
    const float m = fmaxf(fmaxf(fmaxf(fabsf(v1), fabsf(v2)), fmaxf(fabsf(v3), fabsf(v4))), fmaxf(fabsf(v5), fabsf(v6)));
    if (!(m > 1e-30f)) return (0.05f * c1 + 0.45f * c2 + 0.45f * c3 + 0.05f * c4) * (1.0f / 6.0f);
    const float inv = rcp_f32(m);
    const float u1 = v1 * inv, u2 = v2 * inv, u3 = v3 * inv, u4 = v4 * inv, u5 = v5 * inv, u6 = v6 * inv;
    float t = u1 - 2.0f * u2 + u3, r = u1 - 4.0f * u2 + 3.0f * u3;
    const float s1 = 13.0f * t * t + 3.0f * r * r;
    t = u2 - 2.0f * u3 + u4, r = u2 - u4;
    const float s2 = 13.0f * t * t + 3.0f * r * r;
    t = u3 - 2.0f * u4 + u5, r = 3.0f * u3 - 4.0f * u4 + u5;
    const float s3 = 13.0f * t * t + 3.0f * r * r;
    const float t1 = (259.0f * u6 - 1895.0f * u5 + 6670.0f * u4 - 2590.0f * u3 - 2785.0f * u2 + 341.0f * u1) *
                     (1.0f / 5760.0f);
    const float t2 = -(u5 - 12.0f * u4 + 22.0f * u3 - 12.0f * u2 + u1) * (1.0f / 16.0f);
    const float t3 = -(7.0f * u6 - 47.0f * u5 + 94.0f * u4 - 70.0f * u3 + 11.0f * u2 + 5.0f * u1) * (1.0f / 144.0f);
    const float t4 = (u5 - 4.0f * u4 + 6.0f * u3 - 4.0f * u2 + u1) * (1.0f / 24.0f);
    const float t5 = -(-u6 + 5.0f * u5 - 10.0f * u4 + 10.0f * u3 - 5.0f * u2 + u1) * (1.0f / 120.0f);
    const float s6 = (t1 * t1 + t2 * t2 * (13.0f / 3.0f) + t1 * t3 * 0.5f + t3 * t3 * (3129.0f / 80.0f) +
                      t2 * t4 * 4.2f + t1 * t5 * 0.125f + t4 * t4 * (87617.0f / 140.0f) +
                      t3 * t5 * (14127.0f / 224.0f) + t5 * t5 * (252337135.0f / 16128.0f)) * 12.0f;
    const float s5 = fabsf(s6 - (s1 + s3 + 4.0f * s2) * (1.0f / 6.0f));
    const float eps = fmaxf(fminf(eps_abs * inv * inv, 1e30f), 1e-30f);
    const float a1 = 0.05f * (20.0f + s5 * rcp_f32(s1 + eps));
    const float a2 = 0.45f * (20.0f + s5 * rcp_f32(s2 + eps));
    const float a3 = 0.45f * (20.0f + s5 * rcp_f32(s3 + eps));
    const float a4 = 0.05f * (20.0f + s5 * rcp_f32(s6 + eps));
    // normalised weights first (see weno5_body_f32): a_k up to ~1e34
    const float ia = rcp_f32(a1 + a2 + a3 + a4);
    return ((a1 * ia) * c1 + (a2 * ia) * c2 + (a3 * ia) * c3 + (a4 * ia) * c4) * (1.0f / 6.0f);
}

template <int SCHEME>
__device__ __forceinline__ float weno_face_f32(const float *pp, const float *mm, float delta)
{
    if (SCHEME == XF_WENOCU6) {
        const float eps = 1.e-8f * delta * delta;
        return wenocu6_body_f32(pp[1], pp[2], pp[3], pp[4], pp[5], pp[6], eps) +
               wenocu6_body_f32(mm[6], mm[5], mm[4], mm[3], mm[2], mm[1], eps);
    }
    return weno5_body_f32<SCHEME>(pp[1], pp[2], pp[3], pp[4], pp[5]) +
           weno5_body_f32<SCHEME>(mm[6], mm[5], mm[4], mm[3], mm[2]);
}

// get_DpDrho / get_DpDrhoi (Utils_device.hpp:14-35)
__device__ __forceinline__ double dpdrho(double hN, double RN, double q2, double Cp, double R, double T, double gamma)
{
    const double RNT = RN * T;
    return (gamma - 1.0) * (0.5 * q2 - hN + Cp * RNT / R);
}
__device__ __forceinline__ double dpdrhoi(double hin, double Rin, double hiN, double RiN, double T, double Cp, double R,
                                          double gamma)
{
    const double hN_minus_hi = -hin + hiN;
    const double Ri_minus_RN = (Rin - RiN);
    return (gamma - 1.0) * (hN_minus_hi + Cp * Ri_minus_RN * T / R);
}

// Face-local Roe state
template <int NS> struct RoeFace {
    double _u, _v, _w, _H, _P, _rho, c2, b1, b3;
    double _yi[NS];
    double z[NS > 1 ? NS - 1 : 1];
};

// ReconstructSoundSpeed + SoundSpeedMultiSpecies (Utils_device.hpp:42-140)
// from the per-cell ingredients stored by the states pass.
template <int NS>
__device__ __forceinline__ void cop_c2(const DevThermo *__restrict__ th, RoeFace<NS> &rf, double D, double D1,
                                       const double *yl, const double *yr, const double *hl, const double *hr,
                                       double T_l, double T_r, double gamma_l, double gamma_r, double R_l, double R_r,
                                       double Cp_l, double Cp_r, double rl, double rr, double ul, double vl, double wl,
                                       double ur, double vr, double wr, double pl, double pr, double Hl, double Hr)
{
    constexpr int NCOP = NS - 1;
#pragma unroll
    for (int n = 0; n < NS; n++) rf._yi[n] = (yl[n] + D * yr[n]) * D1;
    const double Gamma = (gamma_l + D * gamma_r) * D1;
    const double q2_l = ul * ul + vl * vl + wl * wl;
    const double q2_r = ur * ur + vr * vr + wr * wr;
    const double e_l = Hl - 0.5 * q2_l - pl / rl, e_r = Hr - 0.5 * q2_r - pr / rr;
    const double RN = th->Ri[NCOP];
    const double _dpdrho = (dpdrho(hl[NCOP], RN, q2_l, Cp_l, R_l, T_l, gamma_l) +
                            D * dpdrho(hr[NCOP], RN, q2_r, Cp_r, R_r, T_r, gamma_r)) * D1;
    double _dpdrhoi[NCOP > 0 ? NCOP : 1], drhoi[NCOP > 0 ? NCOP : 1];
#pragma unroll
    for (int nn = 0; nn < NCOP; nn++) {
        const double Rin = th->Ri[nn];
        _dpdrhoi[nn] = (dpdrhoi(hl[nn], Rin, hl[NCOP], RN, T_l, Cp_l, R_l, gamma_l) +
                        D * dpdrhoi(hr[nn], Rin, hr[NCOP], RN, T_r, Cp_r, R_r, gamma_r)) * D1;
        drhoi[nn] = rr * yr[nn] - rl * yl[nn];
    }
    const double _prho = (pl / rl + D * (pr / rr)) * D1 +
                         0.5 * D * D1 * D1 * ((ur - ul) * (ur - ul) + (vr - vl) * (vr - vl) + (wr - wl) * (wr - wl));
    const double _dpdE = ((gamma_l - 1.0) + D * (gamma_r - 1.0)) * D1;
    const double _dpde = ((gamma_l - 1.0) * rl + D * ((gamma_r - 1.0) * rr)) * D1;
    const double dp = pr - pl, drho = rr - rl, de = e_r - e_l;
    double Sum_dpdrhoi = 0.0, Sum_dpdrhoi2 = 0.0, Sum_Yidpdrhoi = 0.0;
#pragma unroll
    for (int n = 0; n < NCOP; n++) {
        Sum_dpdrhoi += _dpdrhoi[n] * drhoi[n];
        Sum_dpdrhoi2 += _dpdrhoi[n] * drhoi[n] * _dpdrhoi[n] * drhoi[n];
    }
    const double temp1 = dp - (_dpdrho * drho + _dpde * de + Sum_dpdrhoi);
    const double temp = temp1 / (_dpdrho * _dpdrho * drho * drho + _dpde * de * _dpde * de + Sum_dpdrhoi2 + 1e-19);
#pragma unroll
    for (int n = 0; n < NCOP; n++) Sum_Yidpdrhoi += rf._yi[n] * _dpdrhoi[n];
    const double _dpdE_new = _dpdE + _dpdE * _dpdE * de * rf._rho * temp;
    const double _dpdrho_new = _dpdrho + _dpdrho * _dpdrho * drho * temp;
    const double csqr = _dpdrho_new + _dpdE_new * _prho + Sum_Yidpdrhoi;
    rf.b1 = _dpdE_new / csqr;
    double b3 = 0.0;
#pragma unroll
    for (int n = 0; n < NCOP; n++) {
        const double nw = _dpdrhoi[n] + _dpdrhoi[n] * _dpdrhoi[n] * drhoi[n] * temp;
        rf.z[n] = -nw / _dpdE_new;
        b3 += rf._yi[n] * rf.z[n];
    }
    rf.b3 = b3 * rf.b1;
    double c2 = csqr;
    const double c2w = (c2 <= 0.0) ? 1.0 : 0.0;
    c2 = Gamma * rf._P * rf._rho * c2w + (1.0 - c2w) * c2;
    rf.c2 = c2;
}

} // namespace xf

namespace xf {


// GLF dissipation of characteristic field n (Artificial_type 3): the running
// maximum of |eigen_local[n]|; fields 1..E-2 share the |u_n| maximum.
// GetLU's GLF reduction happens here: the running maximum is folded with
// the maximum of the last UpdateStates pass.  Every block computes the same
// max(running, pass) and block 0 stores it back; readers before or after the
// store see the same value (max is idempotent).
template <int DIR, int E>
__device__ __forceinline__ double glf_alpha(const double *red, int n)
{
    const int q = 3 * DIR + (n == 0 ? 0 : (n == E - 1 ? 2 : 1));
    const double a = red[RED_GLF + q], b = red[RED_GLFP + q];
    return b > a ? b : a;
}

// PositivityPreservingKernel (positive-definite_eigen/PositivityPreserving_kernels.hpp:5-76)
// on the face flux Fw of one lane: UL/UR, FL/FR are the conservative state
// and physical flux of the left and right cells.  Same operations as the
// reference; the species loop's FF[n] updates (indexed by the species
// counter) feed nothing and are left out.
template <int E, int NCOP>
__device__ __forceinline__ void positivity_face(double *Fw, const double *UL, const double *UR, const double *FL,
                                                const double *FR, double lambda_0, double lambda)
{
    const double eps_rho = 1.0e-13, eps_y = 0.0;
    double F_LF[E], FF_LF[E], FF[E];
#pragma unroll
    for (int n = 0; n < E; n++) {
        F_LF[n] = 0.5 * (FL[n] + FR[n] + lambda_0 * (UL[n] - UR[n]));
        FF_LF[n] = 2.0 * lambda * F_LF[n];
        FF[n] = 2.0 * lambda * Fw[n];
    }
    double theta_u = 1.0, theta_p = 1.0;
    double rho_min = dmin(UL[0], eps_rho);
    if (UL[0] - FF[0] < rho_min) theta_u = (UL[0] - FF_LF[0] - rho_min + 1.0e-40) / (FF[0] - FF_LF[0] + 1.0e-40);
    rho_min = dmin(UR[0], eps_rho);
    if (UR[0] + FF[0] < rho_min) theta_p = (UR[0] + FF_LF[0] - rho_min + 1.0e-40) / (FF_LF[0] - FF[0] + 1.0e-40);
    double theta = dmin(dmax(dmin(theta_u, theta_p), 0.0), 1.0);
#pragma unroll
    for (int n = 0; n < E; n++) {
        FF[n] = (1.0 - theta) * FF_LF[n] + theta * FF[n];
        Fw[n] = (1.0 - theta) * F_LF[n] + theta * Fw[n];
    }
    if (NCOP > 0) {
        const double _rhoq = 1.0 / (UL[0] - FF[0]), _rhou = 1.0 / (UL[0] - FF_LF[0]);
        const double _rhoqp = 1.0 / (UR[0] + FF[0]), _rhoup = 1.0 / (UR[0] + FF_LF[0]);
#pragma unroll
        for (int n = 0; n < NCOP; n++) {
            const int tid = n + 5;
            const double yq = (UL[tid] - FF[tid]) * _rhoq, yu = (UL[tid] - FF_LF[tid]) * _rhou;
            const double yqp = (UR[tid] + FF[tid]) * _rhoqp, yup = (UR[tid] + FF_LF[tid]) * _rhoup;
            theta_u = 1.0, theta_p = 1.0;
            if (yq < eps_y) {
                const double yi_min = dmin(yu, eps_y);
                theta_u = (yu - yi_min + 1.0e-40) / (yu - yq + 1.0e-40);
            }
            if (yqp < eps_y) {
                const double yi_min = dmin(yup, eps_y);
                theta_p = (yup - yi_min + 1.0e-40) / (yup - yqp + 1.0e-40);
            }
            theta = dmin(dmax(dmin(theta_u, theta_p), 0.0), 1.0);
#pragma unroll
            for (int nn = 0; nn < E; nn++) Fw[nn] = (1.0 - theta) * F_LF[nn] + theta * Fw[nn];
        }
    }
}

// Waves per SIMD the sweep is compiled for (the register budget per lane is
// 512 / waves): the species count sets how much state a face lane carries
template <int NS, int PREC> struct SweepOccupancy {
    // the mixed form's FP32 projections leave room for a fourth wave at NS <= 3
    static constexpr int waves = NS == 1 ? 4 : (NS <= 3 ? (PREC == XF_MIXED ? 4 : 3) : 2);
};

// Whether the sweeps compute the viscous wall flux of their faces (when
// Geom::visc_fused asks for it): with NS >= 4, where every form of the sweep
// runs two waves per SIMD and has the registers.  With fewer species the
// sweep runs three or four waves per SIMD, the wall flux's loads would cost
// it registers (and spills) in inviscid runs too, and the wall flux stays a
// launch of its own.
template <int NS> constexpr bool visc_in_sweep() { return NS >= 4; }

// One tile of TS faces x TW pencils of one direction: (tx, ty, tz) is the
// tile, in units of tiles along x, the pencil rows and the third axis.
template <int NS, bool COP, int DIR, int SCHEME, int PREC, int SHAPE>
__device__ __forceinline__ void sweep_tile(const Geom &g, const DevThermo *__restrict__ th,
                                           const double *__restrict__ Ug, const Fields &f, double ncop_gamma,
                                           int mode, int stage, const double *Ua, const double *Ub, double *Uo,
                                           int tx, int ty, int tz)
{
    constexpr int E = COP ? NS + 4 : 5;
    constexpr int NCOP = COP ? NS - 1 : 0;
    using SS = SweepShape<DIR, E, SHAPE>;
    constexpr int TS = SS::TS, TW = SS::TW, NC = SS::NC;
    constexpr bool MIX = PREC == XF_MIXED;
    constexpr bool FUSED = PREC == XF_FUSED_FP64;
    __shared__ double sU[E * NC];
    __shared__ double sF[MIX ? 1 : E * NC];   // FP64 flux: plain form only
    __shared__ float sUf[MIX ? E * NC : 1];   // FP32 copies: mixed form only
    __shared__ float sFf[MIX ? E * NC : 1];
    __shared__ double sP[NC], sR1[NC];
    // block-uniform: the words tested are not written by this launch
    if (step_halted(f.err, ~EBIT_NAN)) return;

    const long long N = g.N;
    const int tid = threadIdx.x;
    const int fidx = DIR == 0 ? tid % TS : tid / TW;
    const int tcol = DIR == 0 ? tid / TS : tid % TW;
    const int B = DIR == 0 ? g.Bx : (DIR == 1 ? g.By : g.Bz);
    const int inner = DIR == 0 ? g.X_inner : (DIR == 1 ? g.Y_inner : g.Z_inner);
    const int Lmax = DIR == 0 ? g.Xmax : (DIR == 1 ? g.Ymax : g.Zmax);
    const int c0 = B + (TS - 1) * (DIR == 0 ? tx : ty);
    const long long stride = DIR == 0 ? 1LL : (DIR == 1 ? (long long)g.Xmax : (long long)g.Xmax * g.Ymax);
    const double *vel_n = DIR == 0 ? f.u : (DIR == 1 ? f.v : f.w);

    auto pencil_base = [&](int q, bool &ok) -> long long {
        if (DIR == 0) {
            const int r = ty * TW + q;
            ok = r < g.Y_inner * g.Z_inner;
            const int j = g.By + r % g.Y_inner, k = g.Bz + r / g.Y_inner;
            return (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j;
        } else if (DIR == 1) {
            const int i = g.Bx + tx * TW + q;
            ok = i < g.Bx + g.X_inner;
            return (long long)g.Xmax * g.Ymax * (g.Bz + tz) + i;
        } else {
            const int i = g.Bx + tx * TW + q;
            ok = i < g.Bx + g.X_inner;
            return (long long)g.Xmax * (g.By + tz) + i;
        }
    };

    // ---- stage U and the physical flux F of TS+5 cells per pencil ----------
    // A lane stages at most NIT cells; every cell's loads are issued before
    // any of them is used, so the second cell's loads overlap the first's
    constexpr int NIT = (NC + 255) / 256;
    double Us[NIT][E], ps[NIT];
    long long ids[NIT];
    int lis[NIT];
    bool oks[NIT];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
        const int q = tid + 256 * it;
        oks[it] = false;
        lis[it] = 0;
        ids[it] = 0;
        if (q < NC) {
            const int pc = DIR == 0 ? q / (TS + 5) : q % TW;
            const int ls = DIR == 0 ? q % (TS + 5) : q / TW;
            bool ok;
            const long long base = pencil_base(pc, ok);
            const int s = c0 - 3 + ls;
            lis[it] = lds_ix<DIR, TS, TW>(ls, pc);
            if (ok && s < Lmax) {
                oks[it] = true;
                ids[it] = base + (long long)s * stride;
#pragma unroll
                for (int n = 0; n < E; n++) Us[it][n] = Ug[n * N + ids[it]];
                ps[it] = f.p[ids[it]];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < NIT; it++) {
        if (tid + 256 * it >= NC) continue;
        const int li = lis[it];
        if (oks[it]) {
            const long long id = ids[it];
            const double *U = Us[it];
            // velocities exactly as GetStates: u = U1 * (1 / rho)
            const double rho1 = 1.0 / U[0];
            const double vel[3] = {U[1] * rho1, U[2] * rho1, U[3] * rho1};
            const double p = ps[it];
            const double Um = U[1 + DIR];
            sP[li] = p;
            sR1[li] = rho1;
            double F[E];
            F[0] = Um;
#pragma unroll
            for (int k = 0; k < 3; k++) F[1 + k] = (k == DIR) ? Um * vel[k] + p : Um * vel[k];
            F[4] = (U[4] + p) * vel[DIR];
            // GetPhysFlux: UI[1 + DIR] * yi (Y_k = U_{5+k} / rho, or the stored
            // fractions of a GhostSpecies mixture)
            if (g.ghost) {
#pragma unroll
                for (int ii = 5; ii < E; ii++) F[ii] = Um * f.y[(ii - 5) * N + id];
            } else {
#pragma unroll
                for (int ii = 5; ii < E; ii++) F[ii] = Um * (U[ii] * rho1);
            }
#pragma unroll
            for (int n = 0; n < E; n++) {
                sU[n * NC + li] = U[n];
                if (MIX) {
                    sUf[n * NC + li] = (float)U[n];
                    sFf[n * NC + li] = (float)F[n];
                } else {
                    sF[n * NC + li] = F[n];
                }
            }
        } else {
#pragma unroll
            for (int n = 0; n < E; n++) {
                sU[n * NC + li] = 0.0;
                if (MIX) sUf[n * NC + li] = 0.0f, sFf[n * NC + li] = 0.0f;
                else sF[n * NC + li] = 0.0;
            }
            sP[li] = 0.0;
            sR1[li] = 0.0;
        }
    }
    // the lane's face, and the thermal values of its two cells that the
    // multi-species sound speed reads from memory, loaded before the staging
    // barrier so that they arrive while the tile's other waves stage
    bool okp;
    const long long pbase = pencil_base(tcol, okp);
    const int sl = c0 + fidx - 1;
    const bool face_ok = okp && (sl <= B + inner - 1);
    constexpr int NT = COP ? NS : 1;
    double hl[NT], hr[NT], Tlr[2], gcvlr[2], Rlr[2], Cplr[2];
    if (COP && face_ok) {
        const long long id_l = pbase + (long long)sl * stride, id_r = id_l + stride;
#pragma unroll
        for (int k = 0; k < NT; k++) hl[k] = f.hi[k * N + id_l], hr[k] = f.hi[k * N + id_r];
        Tlr[0] = f.T[id_l], Tlr[1] = f.T[id_r];
        gcvlr[0] = f.gcv[id_l], gcvlr[1] = f.gcv[id_r];
        Rlr[0] = f.R[id_l], Rlr[1] = f.R[id_r];
        Cplr[0] = f.Cp[id_l], Cplr[1] = f.Cp[id_r];
    }
    __syncthreads();

    if (tx == 0 && ty == 0 && tz == 0 && tid < 3) {
        const double a = f.red[RED_GLF + 3 * DIR + tid], b = f.red[RED_GLFP + 3 * DIR + tid];
        f.red[RED_GLF + 3 * DIR + tid] = b > a ? b : a;
    }

    // ---- one characteristic face flux per lane ------------------------------
    double flux[E];
#pragma unroll
    for (int n = 0; n < E; n++) flux[n] = 0.0;
    if (face_ok) {
        const long long id_l = pbase + (long long)sl * stride;
        const long long id_r = id_l + stride;
        const int li_l = lds_ix<DIR, TS, TW>(fidx + 2, tcol);
        const int li_r = lds_ix<DIR, TS, TW>(fidx + 3, tcol);
        // primitives of the two cells, rebuilt as GetStates forms them
        const double rl = sU[li_l], rr = sU[li_r];
        const double r1l = sR1[li_l], r1r = sR1[li_r];
        const double pl = sP[li_l], pr = sP[li_r];
        const double ul = sU[NC + li_l] * r1l, vl = sU[2 * NC + li_l] * r1l, wl = sU[3 * NC + li_l] * r1l;
        const double ur = sU[NC + li_r] * r1r, vr = sU[2 * NC + li_r] * r1r, wr = sU[3 * NC + li_r] * r1r;
        const double Hl = (sU[4 * NC + li_l] + pl) * r1l, Hr = (sU[4 * NC + li_r] + pr) * r1r;
        // MARCO_ROE
        const double D = sqrt(rr / rl);
        const double D1 = 1.0 / (D + 1.0);
        RoeFace<COP ? NS : 1> rf;
        rf._u = (ul + D * ur) * D1;
        rf._v = (vl + D * vr) * D1;
        rf._w = (wl + D * wr) * D1;
        rf._H = (Hl + D * Hr) * D1;
        rf._P = (pl + D * pr) * D1;
        rf._rho = sqrt(rr * rl);
        if (COP) {
            double yl[NT], yr[NT];
            cell_y<NT>(g, f, id_l, r1l, [&](int k) { return sU[(5 + k) * NC + li_l]; }, yl);
            cell_y<NT>(g, f, id_r, r1r, [&](int k) { return sU[(5 + k) * NC + li_r]; }, yr);
            cop_c2<NT>(th, rf, D, D1, yl, yr, hl, hr, Tlr[0], Tlr[1], gcvlr[0], gcvlr[1], Rlr[0], Rlr[1], Cplr[0],
                       Cplr[1], rl, rr, ul, vl, wl, ur, vr, wr, pl, pr, Hl, Hr);
        } else {
            rf._yi[0] = 1.0;
            rf.c2 = ncop_gamma * rf._P / rf._rho;
            rf.b1 = (ncop_gamma - 1.0) / rf.c2;
            rf.b3 = 0.0;
        }
        // MARCO_PREEIGEN (identical for every row, evaluated once)
        const double _u = rf._u, _v = rf._v, _w = rf._w, _H = rf._H, b1 = rf.b1, b3 = rf.b3;
        const double q2 = _u * _u + _v * _v + _w * _w;
        const double _c = sqrt(rf.c2);
        const double b2 = 1.0 + b1 * q2 - b1 * _H;
        const double _c1 = 1.0 / _c;
        const double vr3[3] = {_u, _v, _w};
        const double un = vr3[DIR];
        const double un_c = un * _c1;

        // GLF split of the six projected points and the face value (WENO5
        // uses pp[1..5], mm[2..6]; WENO-CU6 all six on both sides)
        constexpr int PLO = SCHEME == XF_WENOCU6 ? 1 : 2, PHI = SCHEME == XF_WENOCU6 ? 6 : 5;
        const double dl = DIR == 0 ? g.dx : (DIR == 1 ? g.dy : g.dz);
        auto split_weno = [&](double av, const double *uf, const double *ff) -> double {
            double pp[7], mm[7];
#pragma unroll
            for (int m = 1; m <= PHI; m++) pp[m] = FUSED ? 0.5 * fma(av, uf[m], ff[m]) : 0.5 * (ff[m] + av * uf[m]);
#pragma unroll
            for (int m = PLO; m <= 6; m++) mm[m] = FUSED ? 0.5 * fma(-av, uf[m], ff[m]) : 0.5 * (ff[m] - av * uf[m]);
            return weno_face<SCHEME, FUSED>(pp, mm, dl);
        };
        auto split_weno_f32 = [&](float av, const float *uf, const float *ff) -> double {
            float pp[7], mm[7];
#pragma unroll
            for (int m = 1; m <= PHI; m++) pp[m] = 0.5f * fmaf(av, uf[m], ff[m]);
#pragma unroll
            for (int m = PLO; m <= 6; m++) mm[m] = 0.5f * fmaf(-av, uf[m], ff[m]);
            return (double)weno_face_f32<SCHEME>(pp, mm, (float)dl);
        };
        // dense rows: n = 0 (u-c), n = 1 + DIR (normal), n = E - 1 (u+c);
        // called with literal n, so every branch on n folds at compile time
        auto dense_row = [&](const int n) {
            double L[E];
            if (n == 0) {
                L[0] = 0.5 * (b2 + un_c + b3);
#pragma unroll
                for (int k = 0; k < 3; k++)
                    L[1 + k] = (k == DIR) ? -0.5 * (b1 * vr3[k] + _c1) : -0.5 * (b1 * vr3[k]);
                L[4] = 0.5 * b1;
#pragma unroll
                for (int m = 0; m < NCOP; m++) L[5 + m] = -0.5 * b1 * rf.z[m];
            } else if (n == E - 1) {
                L[0] = 0.5 * (b2 - un_c + b3);
#pragma unroll
                for (int k = 0; k < 3; k++)
                    L[1 + k] = (k == DIR) ? 0.5 * (-b1 * vr3[k] + _c1) : 0.5 * (-b1 * vr3[k]);
                L[4] = 0.5 * b1;
#pragma unroll
                for (int m = 0; m < NCOP; m++) L[5 + m] = -0.5 * b1 * rf.z[m];
            } else {
                L[0] = (1.0 - b2 - b3) / b1;
                L[1] = _u;
                L[2] = _v;
                L[3] = _w;
                L[4] = -1.0;
#pragma unroll
                for (int m = 0; m < NCOP; m++) L[5 + m] = rf.z[m];
            }
            // projections, component-outer: every accumulator still sums
            // over n1 in increasing order, product then sum (plain form) or
            // one fused multiply-add per term (fused form); the empty asm is a compiler-only
            // fence that keeps the LDS loads of one component from being
            // hoisted above the previous one (register pressure)
            double fl;
            if (MIX) {
                // FP32 projections (fused multiply-add) and FP32 WENO
                float uf[7], ff[7];
#pragma unroll
                for (int m = 1; m <= 6; m++) uf[m] = 0.0f, ff[m] = 0.0f;
#pragma unroll
                for (int n1 = 0; n1 < E; n1++) {
                    const float Ln = (float)L[n1];
#pragma unroll
                    for (int m = 1; m <= 6; m++) {
                        const int li = lds_ix<DIR, TS, TW>(fidx + m - 1, tcol);
                        uf[m] = fmaf(sUf[n1 * NC + li], Ln, uf[m]);
                        ff[m] = fmaf(sFf[n1 * NC + li], Ln, ff[m]);
                    }
                    asm volatile("" ::: "memory");
                }
                fl = split_weno_f32((float)glf_alpha<DIR, E>(f.red, n), uf, ff);
            } else {
                double uf[7], ff[7];
#pragma unroll
                for (int m = 1; m <= 6; m++) uf[m] = 0.0, ff[m] = 0.0;
#pragma unroll
                for (int n1 = 0; n1 < E; n1++) {
                    const double Ln = L[n1];
#pragma unroll
                    for (int m = 1; m <= 6; m++) {
                        const int li = lds_ix<DIR, TS, TW>(fidx + m - 1, tcol);
                        uf[m] = madd<FUSED>(sU[n1 * NC + li], Ln, uf[m]);
                        ff[m] = madd<FUSED>(sF[n1 * NC + li], Ln, ff[m]);
                    }
                    asm volatile("" ::: "memory");
                }
                fl = split_weno(glf_alpha<DIR, E>(f.red, n), uf, ff);
            }
            // right column n, accumulated in the reference's summation order
            // (fluxl += f_flux * r per entry, fused in the fused form)
            if (n == 0 || n == E - 1) {
                flux[0] += fl;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double Rk = (k == DIR) ? (n == 0 ? vr3[k] - _c : vr3[k] + _c) : vr3[k];
                    flux[1 + k] = madd<FUSED>(fl, Rk, flux[1 + k]);
                }
                flux[4] = madd<FUSED>(fl, n == 0 ? _H - un * _c : _H + un * _c, flux[4]);
#pragma unroll
                for (int m = 0; m < NCOP; m++) flux[5 + m] = madd<FUSED>(fl, rf._yi[m], flux[5 + m]);
            } else {
                flux[0] = madd<FUSED>(fl, b1, flux[0]);
                flux[1] = madd<FUSED>(fl, _u * b1, flux[1]);
                flux[2] = madd<FUSED>(fl, _v * b1, flux[2]);
                flux[3] = madd<FUSED>(fl, _w * b1, flux[3]);
                flux[4] = madd<FUSED>(fl, _H * b1 - 1.0, flux[4]);
#pragma unroll
                for (int m = 0; m < NCOP; m++) flux[5 + m] = madd<FUSED>(fl, b1 * rf._yi[m], flux[5 + m]);
            }
        };

        dense_row(0);
        // the E-3 rows with two structural non-zeros (L[0] and +-1 at kk; the
        // right column has +-1 at kk and an energy entry R4), in a runtime
        // loop; contributions go to flux[kk] through a masked add (+0.0 on
        // the other components is exact: no accumulator is ever -0.0).
        // A product by an exact 0 or +-1 rounds the same fused or unfused,
        // so these rows equal the oracle's full dot products in either form.
#pragma unroll 1
        for (int n = 1; n <= E - 2; n++) {
            if (n == 1 + DIR) {
                dense_row(1 + DIR);
                continue;
            }
            double L0, R4;
            int kk;
            bool neg;
            if (n <= 3) {
                kk = n;
                if (DIR == 0) { L0 = (n == 2) ? _v : -_w; neg = (n == 2); R4 = (n == 2) ? -_v : _w; }
                else if (DIR == 1) { L0 = (n == 1) ? -_u : _w; neg = (n == 3); R4 = (n == 1) ? _u : -_w; }
                else { L0 = (n == 1) ? _u : -_v; neg = (n == 1); R4 = (n == 1) ? -_u : _v; }
            } else {
                const int m = n - 4;
                double ym = 0.0, zm = 0.0;
#pragma unroll
                for (int q = 0; q < NCOP; q++) {
                    ym = (q == m) ? rf._yi[q] : ym;
                    zm = (q == m) ? rf.z[q] : zm;
                }
                kk = 5 + m;
                L0 = -ym;
                R4 = zm;
                neg = false;
            }
            double fl;
            if (MIX) {
                const float L0f = (float)L0;
                float uf[7], ff[7];
#pragma unroll
                for (int m = 1; m <= 6; m++) {
                    const int li = lds_ix<DIR, TS, TW>(fidx + m - 1, tcol);
                    const float au = sUf[kk * NC + li], bf = sFf[kk * NC + li];
                    uf[m] = fmaf(sUf[li], L0f, neg ? -au : au);
                    ff[m] = fmaf(sFf[li], L0f, neg ? -bf : bf);
                }
                fl = split_weno_f32((float)glf_alpha<DIR, E>(f.red, n), uf, ff);
            } else {
                double uf[7], ff[7];
#pragma unroll
                for (int m = 1; m <= 6; m++) {
                    const int li = lds_ix<DIR, TS, TW>(fidx + m - 1, tcol);
                    const double a0 = 0.0 + sU[li] * L0;
                    const double b0 = 0.0 + sF[li] * L0;
                    const double au = sU[kk * NC + li], bf = sF[kk * NC + li];
                    uf[m] = neg ? a0 + (-au) : a0 + au;
                    ff[m] = neg ? b0 + (-bf) : b0 + bf;
                }
                fl = split_weno(glf_alpha<DIR, E>(f.red, n), uf, ff);
            }
            const double t = neg ? -fl : fl;
#pragma unroll
            for (int k = 1; k < E; k++)
                if (k != 4) flux[k] += (k == kk) ? t : 0.0;
            flux[4] = madd<FUSED>(fl, R4, flux[4]);
        }
        dense_row(E - 1);
        // positivity-preserving limiter on the faces right of interior cells
        if (g.positivity && sl >= B) {
            const double lambda_0 = f.red[RED_LAM + DIR];
            const double lambda = g.CFL / lambda_0;
            double UL[E], UR[E], FL[E], FR[E];
            auto cell_flux = [&](int li, long long idc, double *Uc, double *Fc) {
#pragma unroll
                for (int n = 0; n < E; n++) Uc[n] = sU[n * NC + li];
                if (!MIX) {
#pragma unroll
                    for (int n = 0; n < E; n++) Fc[n] = sF[n * NC + li];
                } else {
                    // the physical flux exactly as staged for the plain form
                    const double rho1 = sR1[li], p = sP[li];
                    const double vel[3] = {Uc[1] * rho1, Uc[2] * rho1, Uc[3] * rho1};
                    const double Um = Uc[1 + DIR];
                    Fc[0] = Um;
#pragma unroll
                    for (int k = 0; k < 3; k++) Fc[1 + k] = (k == DIR) ? Um * vel[k] + p : Um * vel[k];
                    Fc[4] = (Uc[4] + p) * vel[DIR];
                    if (g.ghost) {
#pragma unroll
                        for (int ii = 5; ii < E; ii++) Fc[ii] = Um * f.y[(ii - 5) * N + idc];
                    } else {
#pragma unroll
                        for (int ii = 5; ii < E; ii++) Fc[ii] = Um * (Uc[ii] * rho1);
                    }
                }
            };
            cell_flux(li_l, id_l, UL, FL);
            cell_flux(li_r, id_r, UR, FR);
            positivity_face<E, NCOP>(flux, UL, UR, FL, FR, lambda_0, lambda);
        }
        // viscous wall flux of this face (after the limiter, as GetLU orders
        // them): the same function k_visc_flux runs, computed here while the
        // other waves of the tile still run their characteristic fluxes, or
        // (visc_fused 0) read from the planes k_visc_flux wrote
        if (g.visc) {
            double Fv[E];
            if constexpr (COP && visc_in_sweep<NS>()) {
                if (g.visc_fused) {
                    // the stencil cells' U and 1 / rho from the staged tile
                    // (the face's left cell sits at fidx + 2)
                    visc_face<NS, DIR, MIX ? XF_MIXED : XF_PLAIN_FP64>(
                        g, f, id_l, [&](int c, int n) { return sU[n * NC + lds_ix<DIR, TS, TW>(fidx + 1 + c, tcol)]; },
                        [&](int c) { return sR1[lds_ix<DIR, TS, TW>(fidx + 1 + c, tcol)]; }, Fv);
                } else {
#pragma unroll
                    for (int n = 0; n < E; n++) Fv[n] = f.Fv[n * N + id_l];
                }
            } else {
#pragma unroll
                for (int n = 0; n < E; n++) Fv[n] = f.Fv[n * N + id_l];
            }
#pragma unroll
            for (int n = 0; n < E; n++) flux[n] = flux[n] - Fv[n];
        }
    }

    // ---- flux divergence on the TS-1 cells of the tile ----------------------
    // the lane's cell, and what the divergence reads of it from memory (LU
    // of the earlier directions; U and the stage input of the RK update),
    // loaded before the barriers: the loads are in flight while the tile's
    // other waves finish their faces.  Each cell is read and written only by
    // its own lane, and no tile stages from the arrays written here.
    const bool cell_ok = fidx >= 1 && okp && sl <= B + inner - 1;
    double LUi[E], Ai[E], Bi[E];
#pragma unroll
    for (int n = 0; n < E; n++) LUi[n] = 0.0, Ai[n] = 0.0, Bi[n] = 0.0;
    if (cell_ok) {
        const long long id = pbase + (long long)sl * stride;
        if (mode & 1) {
#pragma unroll
            for (int n = 0; n < E; n++) LUi[n] = f.LU[n * N + id];
        }
        if (mode & 2) {
#pragma unroll
            for (int n = 0; n < E; n++) Ai[n] = Ua[n * N + id], Bi[n] = Ub[n * N + id];
        }
    }
    __syncthreads();
    double *sFw = sU; // reuse the staging area (E * NC >= E * 256)
    const int fw = DIR == 0 ? tcol * TS + fidx : fidx * TW + tcol;
#pragma unroll
    for (int n = 0; n < E; n++) sFw[n * 256 + fw] = flux[n];
    __syncthreads();
    if (cell_ok) {
        const long long id = pbase + (long long)sl * stride;
        const int fwm = DIR == 0 ? tcol * TS + fidx - 1 : (fidx - 1) * TW + tcol;
        const double dl = DIR == 0 ? g._dx : (DIR == 1 ? g._dy : g._dz);
        const double dt = (mode & 2) ? f.dt[0] : 0.0;
        bool bad = false;
#pragma unroll
        for (int n = 0; n < E; n++) {
            const double d = (sFw[n * 256 + fwm] - sFw[n * 256 + fw]) * dl;
            const double LUt = (mode & 1) ? LUi[n] + d : 0.0 + d;
            if (mode & 2) {
                // fused SSP-RK3 stage: Ua = U, Ub = stage input (stage 2, 3),
                // Uo = output.  Before the update, EstimateNAN (flags 1-3,
                // src/XFLUIDS.cpp:460-466, Fluids.cpp:963-1040) checks LU and
                // the stage's UI: U1 at stages 1 and 2 (at stage 1 still the
                // previous step's second-stage result, which is what Ub holds
                // then), U at stage 3
                const double a0 = Ai[n], b0 = Bi[n];
                const double ui = (stage == 3) ? a0 : b0;
                bad = bad || isnan(LUt) || isinf(LUt) || isnan(ui) || isinf(ui) || (n == 0 && ui < 0);
                double a = a0, b = (stage == 1) ? 0.0 : b0;
                rk3_point(stage, dt, a, b, LUt);
                const double o = (stage == 3) ? a : b;
                Uo[n * N + id] = o;
            } else {
                f.LU[n * N + id] = LUt;
            }
        }
        if (bad) {
            report_cell(f.err, EBIT_NAN, id);
            report_pass(f.err, EBIT_NAN, stage);
        }
    }
}

// tiles of one case: grid (tiles along the sweep's x, rows, third axis), in
// the XCD-aware order
template <int NS, bool COP, int DIR, int SCHEME, int PREC, int SHAPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<NS, PREC>::waves)))
k_sweep(Geom g, const DevThermo *__restrict__ th, const double *__restrict__ Ug, Fields f, double ncop_gamma, int mode,
        int stage, const double *Ua, const double *Ub, double *Uo)
{
    int tx, ty, tz;
    xcd_tile(tx, ty, tz);
    sweep_tile<NS, COP, DIR, SCHEME, PREC, SHAPE>(g, th, Ug, f, ncop_gamma, mode, stage, Ua, Ub, Uo, tx, ty, tz);
}

// the tiles of every case of an ensemble in one launch: block b belongs to
// the case whose block range [off, off + count) holds it; the stage's
// buffers are picked per case by code (0 U, 1 U1, 2 U2)
template <int NS, bool COP, int DIR, int SCHEME, int PREC, int SHAPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SweepOccupancy<NS, PREC>::waves)))
k_sweep_multi(const EnsCase *__restrict__ cases, int ncases, const DevThermo *__restrict__ th, double ncop_gamma,
              int mode, int stage, int sel_g, int sel_a, int sel_b, int sel_o)
{
    const long long b = blockIdx.x;
    const int k = ens_find(cases, ncases, b, DIR);
    const EnsCase &c = cases[k];
    const int l = (int)(b - c.tile_off[DIR]);
    const int gx = c.tiles[DIR][0], gy = c.tiles[DIR][1];
    sweep_tile<NS, COP, DIR, SCHEME, PREC, SHAPE>(c.g, th, c.buf(sel_g), c.f, ncop_gamma, mode, stage, c.buf(sel_a),
                                                  c.buf(sel_b), c.buf(sel_o), l % gx, (l / gx) % gy, l / (gx * gy));
}

} // namespace xf
