// Stiff finite-rate chemistry on the device: constant-volume production and
// loss rates and the CHEMEQ2 quasi-steady-state integrator, temperature
// carried as y[0].
//
// Reference: evalcore (q/p form) src/solver_Reaction/ProductionRates.h:117-176,
// evalcv :459-479, xfrsign/Alpha :227-244, Chemeq2 :488-638,
// ChemeODEQ2SolverKernel1 src/solver_Reaction/Reaction_kernels.hpp:46-73.
//
// Restructured without changing any result bit:
//  * get_Gibson(T, n) and log(p/p_atm), log(T) do not depend on the reaction
//    and are evaluated once per rate evaluation instead of NR*NS times;
//  * terms with a zero stoichiometric coefficient or third-body efficiency
//    add +0 (or multiply by pown(x, 0) = 1) in the reference and are skipped;
//  * rates are accumulated into q, d reaction by reaction; every species'
//    sum still runs over reactions in increasing order;
//  * the predictor loop writes y directly and the corrector loop is fused
//    with the convergence loop (y1 is the predicted y read before overwrite).
// Safety bound: at most CHEM_MAXSUB predictor passes per cell; a cell that
// reaches it raises XF_ERR_CHEM.  The reference loop is unbounded.
#pragma once
#include "xf_thermo.h"

namespace xf {

constexpr int CHEM_MAXSUB = 200000;

__device__ __forceinline__ double pown_i(double x, int n)
{
    if (n == 0) return 1.0;
    double r = x;
    for (int k = 1; k < n; k++) r *= x;
    return r;
}

// evalcv (y[0] = T): q, d of length NS + 1
template <int NS>
__device__ __forceinline__ void evalcv_dev(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc,
                                           const double *y, double *q, double *d, double m_dens, double m_p)
{
    const double m_tmp = y[0];
    const double _dens = 1.0 / m_dens;
    // species Gibbs energy and internal-energy ingredients at T (Ri = 1)
    const TPre t = th->thermo ? tpre<true>(m_tmp, 1) : tpre<true>(m_tmp, 0);
    double h1[NS], gk[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
        h1[n] = h_k(th, t, 1.0, n);
        double temp = 0.0;
        temp += s_k(th, t, 1.0, n);
        temp += (-h1[n] / m_tmp);
        gk[n] = temp;
    }
    const double logStandConc = dlog(m_p / (R0 * m_tmp));
    const double logT = dlog(m_tmp);
    const double logP = dlog(m_p * (1.0 / 101325.0));
    double m_cm[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) m_cm[n] = y[1 + n] * th->_Wi[n] * m_dens * 1.0E-3;
    double *qi = q + 1, *pi = d + 1;
#pragma unroll
    for (int n = 0; n < NS; n++) qi[n] = 0.0, pi[n] = 0.0;
    const int NR = mc->NR;
    for (int r = 0; r < NR; r++) {
        const double A = mc->Rargus[r * 6 + 0], B = mc->Rargus[r * 6 + 1], Ea = mc->Rargus[r * 6 + 2];
        double Kf = A * dexp(B * logT - Ea * 4.184 / Ru / m_tmp);
        double DeltaGibbs = 0.0;
        int m_dn = 0;
#pragma unroll
        for (int n = 0; n < NS; n++) {
            const int nu = mc->Nu_d[r * NS + n];
            if (nu != 0) {
                DeltaGibbs += nu * (logP - gk[n]);
                m_dn += nu;
            }
        }
        double Kck = dmin(dexp(DeltaGibbs - m_dn * logStandConc), 1.0E+40);
        double tb = 1.0;
        if (1 == mc->third_ind[r]) {
            tb = 0.0;
#pragma unroll
            for (int it = 0; it < NS; it++) {
                const double cf = mc->third[r * NS + it];
                if (cf != 0.0) tb += cf * m_cm[it];
            }
        }
        Kf *= tb;
        Kck *= Kf;
#pragma unroll
        for (int it = 0; it < NS; it++) {
            const int nf = mc->Nu_f[r * NS + it];
            if (nf != 0) Kf *= pown_i(m_cm[it], nf);
        }
#pragma unroll
        for (int it = 0; it < NS; it++) {
            const int nb = mc->Nu_b[r * NS + it];
            if (nb != 0) Kck *= pown_i(m_cm[it], nb);
        }
#pragma unroll
        for (int n = 0; n < NS; n++) {
            const int nf = mc->Nu_f[r * NS + n], nb = mc->Nu_b[r * NS + n];
            if ((nf | nb) != 0) {
                qi[n] += nb * Kf + nf * Kck;
                pi[n] += nf * Kf + nb * Kck;
            }
        }
    }
    double q0 = 0.0, d0 = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) {
        qi[n] *= th->Wi[n] * _dens;
        pi[n] *= th->Wi[n] * _dens;
        const double ei = (h1[n] - m_tmp) * th->Ri[n];
        q0 -= qi[n] * ei, d0 -= pi[n] * ei;
    }
    const double _Cv = 1.0 / mix_cv<NS>(th, y + 1, t);
    q[0] = q0 * _Cv, d[0] = d0 * _Cv;
}

__device__ __forceinline__ double alpha_of(double r)
{
    return (180.0 + r * (60.0 + r * (11.0 + r))) / (360.0 + r * (60.0 + r * (12.0 + r)));
}

// Chemeq2 on one cell.  Returns predictor passes, or -1 at the safety bound.
template <int NS>
__device__ int chemeq2_dev(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, double *y, double dtg,
                           double rho, double m_p)
{
    constexpr int NL = NS + 1;
    const double ymin = 1.0e-20, epsmin = 1.0e-4, tfd = 1.0 + 1.0e-10, sqreps = 0.05, epscl = 1.0e4;
    double dtmin = 1.0e-7, scrtch = 1e-25, epsmax = 1.0;
    double ys[NL], qs[NL], ds[NL], q[NL], d[NL], rtau[NL];
    double sumy = 0.0;
#pragma unroll
    for (int i = 1; i < NL; i++) y[i] = dmax(y[i], ymin), sumy += y[i];
    sumy = 1.0 / sumy;
#pragma unroll
    for (int i = 1; i < NL; i++) y[i] *= sumy;
    evalcv_dev<NS>(th, mc, y, q, d, rho, m_p);
#pragma unroll
    for (int i = 0; i < NL; i++) {
        const double ascr = fabs(q[i]);
        const double sb = 0.1 * epsmin * ascr - d[i];
        const double scr2 = (sb >= 0.0 ? 1.0 : -1.0) * fabs(1.0 / y[i]);
        const double scr1 = scr2 * d[i];
        scrtch = dmax(-fabs(ascr - d[i]) * scr2, dmax(scr1, scrtch));
    }
    double dt = dmin(sqreps / scrtch, dtg);
    dtmin *= dt;
    double ts = 0.0, tn = 0.0;
#pragma unroll
    for (int i = 0; i < NL; i++) ys[i] = y[i], qs[i] = q[i], ds[i] = d[i];
    int passes = 0;
    for (;;) {
        if (++passes > CHEM_MAXSUB) return -1;
        // predictor
#pragma unroll
        for (int i = 0; i < NL; i++) {
            rtau[i] = dt * ds[i] / ys[i];
            const double alpha = alpha_of(rtau[i]);
            const double scr = dt * (qs[i] - ds[i]) / (1.0 + alpha * rtau[i]);
            y[i] = dmax(ys[i] + scr, ymin);
        }
        tn = ts + dt;
        evalcv_dev<NS>(th, mc, y, q, d, rho, m_p);
        // corrector fused with the convergence test
        double eps = 1e-10;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const double yp = y[i];
            const double rtaub = 0.5 * (rtau[i] + dt * d[i] / yp);
            const double alpha = alpha_of(rtaub);
            const double qt = (1.0 - alpha) * qs[i] + alpha * q[i];
            const double scr = (qt * dt - rtaub * ys[i]) / (1.0 + alpha * rtaub);
            const double scr2 = dmax(ys[i] + scr, 0.0);
            double scr1 = fabs(scr2 - yp);
            y[i] = dmax(scr2, ymin);
            if ((0.25 * (ys[i] + y[i])) > ymin) {
                scr1 = scr1 / y[i];
                eps = dmax(0.5 * (scr1 + dmin(fabs(q[i] - d[i]) / (q[i] + d[i] + 1.0e-30), scr1)), eps);
            }
        }
        eps = eps * epscl;
        if (eps < epsmax) {
            if (dtg <= (tn * tfd)) return passes;
        } else {
            tn = ts;
        }
        const double dto = dt;
        double rteps = 0.5 * (eps + 1.0);
        rteps = 0.5 * (rteps + eps / rteps);
        rteps = 0.5 * (rteps + eps / rteps);
        dt = dmin(dt * (1.0 / rteps + 0.005), tfd * (dtg - tn));
        if (eps > epsmax) {
            dt = dmin(dt, 0.34 * dto);
            if (dt <= dtmin) epsmax *= 10.0;
            continue; // back to the predictor from the same saved state
        }
        epsmax = 1.0;
        ts = tn;
#pragma unroll
        for (int i = 0; i < NL; i++) ys[i] = y[i], qs[i] = q[i], ds[i] = d[i];
    }
}

// get_Cope (Mixing_device.h:145-154) at one temperature
template <int NS>
__device__ __forceinline__ double mix_e(const DevThermo *__restrict__ th, const double *yi, double T)
{
    const TPre t = th->thermo ? tpre<true>(T, 1) : tpre<false>(T, 0);
    double e = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) {
        const double ei = (h_k(th, t, 1.0, n) - T) * th->Ri[n];
        e += ei * yi[n];
    }
    return e;
}

// One interior cell of ChemeODEQ2SolverKernel1.
template <int NS>
__device__ __forceinline__ int chem_cell(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, Fields &f,
                                         long long N, long long id, double dtc)
{
    constexpr int E = NS + 4;
    double y[NS + 1];
    y[0] = f.T[id];
    const double rho = f.U[id]; // rho = U[0] (Getrhoyi)
    const double rho1 = 1.0 / rho;
    y[1 + NS - 1] = 1.0;
#pragma unroll
    for (int k = 0; k < NS - 1; k++) {
        y[1 + k] = f.U[(5 + k) * N + id] * rho1;
        y[NS] += -y[1 + k];
    }
    const double e0 = mix_e<NS>(th, y + 1, y[0]);
    const int passes = chemeq2_dev<NS>(th, mc, y, dtc, rho, f.p[id]);
    const double e1 = mix_e<NS>(th, y + 1, y[0]);
#pragma unroll
    for (int n = 0; n < NS - 1; n++) f.U[(5 + n) * N + id] = y[1 + n] * rho;
    f.T[id] = y[0];
    f.U[4 * N + id] += (e1 - e0) * rho;
    (void)E;
    return passes;
}

// one lane per interior cell
template <int NS>
__global__ void __launch_bounds__(256) k_chem(Geom g, const DevThermo *__restrict__ th,
                                              const DevMech *__restrict__ mc, Fields f, int record)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.X_inner * g.Y_inner * g.Z_inner;
    if (t >= total) return;
    const int i = g.Bx + (int)(t % g.X_inner);
    const int j = g.By + (int)((t / g.X_inner) % g.Y_inner);
    const int k = g.Bz + (int)(t / ((long long)g.X_inner * g.Y_inner));
    const long long id = (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
    const int passes = chem_cell<NS>(th, mc, f, g.N, id, f.dt[1]);
    if (record) f.substeps[t] = passes;
    if (passes < 0) {
        atomicOr(&f.err[0], EBIT_CHEM);
        atomicMin(&f.err[1], (int)id);
    }
}

} // namespace xf
