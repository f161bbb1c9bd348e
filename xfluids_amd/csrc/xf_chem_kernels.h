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

// Chemeq2 state of one cell, advanced one predictor/corrector pass at a time
// so that a lane can also work through a queue of cells (k_chem_queue).
template <int NS> struct ChemLane {
    static constexpr int NL = NS + 1;
    double y[NL], ys[NL], qs[NL], ds[NL], q[NL], d[NL], rtau[NL];
    double dt, dtg, dtmin, epsmax, ts, tn, rho, m_p;
    int passes;

    // y = [T, Y_1..Y_NS] on entry: clip and renormalise Y, first rate
    // evaluation, initial step (ProductionRates.h:488-530)
    __device__ __forceinline__ void start(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc,
                                          double dtg_, double rho_, double p_)
    {
        const double ymin = 1.0e-20, epsmin = 1.0e-4, sqreps = 0.05;
        dtg = dtg_, rho = rho_, m_p = p_;
        dtmin = 1.0e-7, epsmax = 1.0;
        double scrtch = 1e-25;
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
        dt = dmin(sqreps / scrtch, dtg);
        dtmin *= dt;
        ts = 0.0, tn = 0.0;
#pragma unroll
        for (int i = 0; i < NL; i++) ys[i] = y[i], qs[i] = q[i], ds[i] = d[i];
        passes = 0;
    }

    // one predictor/corrector pass (ProductionRates.h:531-638): 1 = reached
    // dtg, 0 = continue, -1 = safety bound
    __device__ __forceinline__ int pass(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc)
    {
        const double ymin = 1.0e-20, tfd = 1.0 + 1.0e-10, epscl = 1.0e4;
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
            if (dtg <= (tn * tfd)) return 1;
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
            return 0; // back to the predictor from the same saved state
        }
        epsmax = 1.0;
        ts = tn;
#pragma unroll
        for (int i = 0; i < NL; i++) ys[i] = y[i], qs[i] = q[i], ds[i] = d[i];
        return 0;
    }
};

// Chemeq2 on one cell.  Returns predictor passes, or -1 at the safety bound.
template <int NS>
__device__ int chemeq2_dev(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, double *y, double dtg,
                           double rho, double m_p)
{
    ChemLane<NS> L;
#pragma unroll
    for (int i = 0; i < NS + 1; i++) L.y[i] = y[i];
    L.start(th, mc, dtg, rho, m_p);
    int st;
    while ((st = L.pass(th, mc)) == 0) {
    }
#pragma unroll
    for (int i = 0; i < NS + 1; i++) y[i] = L.y[i];
    return st < 0 ? -1 : L.passes;
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

// Interior cell t (0 <= t < total) -> padded-grid id
__device__ __forceinline__ long long chem_cell_id(const Geom &g, long long t)
{
    const int i = g.Bx + (int)(t % g.X_inner);
    const int j = g.By + (int)((t / g.X_inner) % g.Y_inner);
    const int k = g.Bz + (int)(t / ((long long)g.X_inner * g.Y_inner));
    return (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
}

// Work-queue driver: persistent lanes draining the interior cells.  Cells
// of very different stiffness take from 1 to ~10^3 passes; with one lane per
// cell a wave runs as long as its stiffest cell while the other lanes idle.
// Here every lane that has finished a cell takes the next one from a device
// counter (one atomic per wave, handed out by lane rank), so the lanes of a
// wave stay busy with the same predictor/corrector code until the queue is
// empty.  Each cell's arithmetic is exactly that of the per-cell driver.
// Termination: the counter only grows and a wave stops once every lane has
// drawn an index >= total and holds no cell; a pass is bounded by
// CHEM_MAXSUB per cell.  The counter queue[0] must be zero at launch.
template <int NS>
__global__ void __launch_bounds__(64) k_chem_queue(Geom g, const DevThermo *__restrict__ th,
                                                   const DevMech *__restrict__ mc, Fields f, int record)
{
    const long long total = (long long)g.X_inner * g.Y_inner * g.Z_inner;
    const long long N = g.N;
    const double dtc = f.dt[1];
    const int lane = threadIdx.x & 63;
    ChemLane<NS> L;
    long long t = -1, id = 0;
    double e0 = 0.0;
    bool active = false, done = false;
    for (;;) {
        const bool need = !active && !done;
        const unsigned long long want = __ballot(need);
        if (want) {
            const int cnt = __popcll(want);
            const int leader = __ffsll((long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd((unsigned long long *)f.queue, (unsigned long long)cnt);
            base = __shfl(base, leader);
            if (need) {
                const int rank = __popcll(want & ((1ull << lane) - 1ull));
                t = (long long)(base + (unsigned long long)rank);
                if (t < total) {
                    id = chem_cell_id(g, t);
                    L.y[0] = f.T[id];
                    const double rho = f.U[id];
                    const double rho1 = 1.0 / rho;
                    L.y[NS] = 1.0;
#pragma unroll
                    for (int k = 0; k < NS - 1; k++) {
                        L.y[1 + k] = f.U[(5 + k) * N + id] * rho1;
                        L.y[NS] += -L.y[1 + k];
                    }
                    e0 = mix_e<NS>(th, L.y + 1, L.y[0]);
                    L.start(th, mc, dtc, rho, f.p[id]);
                    active = true;
                } else {
                    done = true;
                }
            }
        }
        if (!__any(active)) break;
        if (active) {
            const int st = L.pass(th, mc);
            if (st != 0) {
                const double rho = L.rho;
                const double e1 = mix_e<NS>(th, L.y + 1, L.y[0]);
#pragma unroll
                for (int n = 0; n < NS - 1; n++) f.U[(5 + n) * N + id] = L.y[1 + n] * rho;
                f.T[id] = L.y[0];
                f.U[4 * N + id] += (e1 - e0) * rho;
                if (record) f.substeps[t] = st < 0 ? -1 : L.passes;
                if (st < 0) {
                    atomicOr(&f.err[0], EBIT_CHEM);
                    atomicMin(&f.err[1], (int)id);
                }
                active = false;
            }
        }
    }
}

// zero the work-queue counter (a kernel, so the step graph holds no memset)
__global__ void k_queue_reset(unsigned int *q)
{
    if (threadIdx.x < 4) q[threadIdx.x] = 0u;
}

} // namespace xf
