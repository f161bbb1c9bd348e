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
#include "gen/xf_mech_gen.h"
#include <utility>

namespace xf {

constexpr int CHEM_MAXSUB = 200000;

__device__ __forceinline__ double pown_i(double x, int n)
{
    // the products of the loop below, written out for the usual orders
    if (n == 1) return x;
    if (n == 2) return x * x;
    if (n == 3) return (x * x) * x;
    if (n == 0) return 1.0;
    double r = x;
    for (int k = 1; k < n; k++) r *= x;
    return r;
}

// What one reaction reads besides its record: T, the shared reciprocal of T
// and whether it qualifies for the corrected quotient, the logarithms
struct RxnCtx {
    double m_tmp, rT, logT, logP, logStandConc;
    bool t_ok;
};

// One reaction's contribution to the production and loss rates (evalcore,
// ProductionRates.h:117-176).  The record R and the third-body efficiencies
// eff(s) come either from the device tables (generic loop) or from
// compile-time constants (a specialisation of a shipped mechanism): the code
// is this one function in both cases, so the results are the same bits.
// Species loops run over compile-time n and test the record's masks, which
// skips exactly the zero-coefficient terms (each adds +0 or multiplies by 1 in
// the reference, so skipping changes no bit).
template <int NS, class Eff>
__device__ __forceinline__ void rxn_rates(const DevRxn R, Eff eff, const RxnCtx &c, const double *gk,
                                          const double *m_cm, double &Kf_out, double &Kck_out)
{
    const double A = R.A, B = R.B, EaR = R.EaR;
    const unsigned md = R.md, mt = R.mt, mf = R.mf, mb = R.mb;
    const unsigned long long pd = R.pd, pf = R.pf, pb = R.pb;
    const int tind = R.third_ind;
    // EaR / T from the shared reciprocal: Markstein's corrected quotient
    // fma(fma(-T, q0, a), y, q0), q0 = a y, y = RN(1/T), is the correctly
    // rounded a / T while a, T lie in [2^-500, 2^500] (the remainder is exact
    // and nothing underflows); a = 0 keeps q0, the signed zero of a / T
    double ET;
    const double aE = fabs(EaR);
    if (c.t_ok && (EaR == 0.0 || (aE >= 0x1p-500 && aE <= 0x1p500))) {
        const double q0 = EaR * c.rT;
        ET = EaR == 0.0 ? q0 : fma(fma(-c.m_tmp, q0, EaR), c.rT, q0);
    } else {
        ET = EaR / c.m_tmp;
    }
    double Kf = A * dexp(B * c.logT - ET);
    double DeltaGibbs = 0.0;
    int m_dn = 0;
#pragma unroll
    for (int n = 0; n < NS; n++) {
        if ((md >> n) & 1u) {
            const int nu = (int)((pd >> (4 * n)) & 15ull) - 8;
            DeltaGibbs += nu * (c.logP - gk[n]);
            m_dn += nu;
        }
    }
    double Kck = dmin(dexp(DeltaGibbs - m_dn * c.logStandConc), 1.0E+40);
    double tb = 1.0;
    if (1 == tind) {
        tb = 0.0;
        // every efficiency of the row first (one burst of scalar loads when
        // they come from the tables)
        double e[NS];
#pragma unroll
        for (int it = 0; it < NS; it++) e[it] = eff(it);
#pragma unroll
        for (int it = 0; it < NS; it++)
            if ((mt >> it) & 1u) tb += e[it] * m_cm[it];
    }
    Kf *= tb;
    Kck *= Kf;
#pragma unroll
    for (int it = 0; it < NS; it++)
        if ((mf >> it) & 1u) Kf *= pown_i(m_cm[it], (int)((pf >> (4 * it)) & 15ull));
#pragma unroll
    for (int it = 0; it < NS; it++)
        if ((mb >> it) & 1u) Kck *= pown_i(m_cm[it], (int)((pb >> (4 * it)) & 15ull));
    Kf_out = Kf, Kck_out = Kck;
}

// the reaction's terms of the production and loss rates of every species
template <int NS>
__device__ __forceinline__ void rxn_accum(unsigned mf, unsigned mb, unsigned long long pf, unsigned long long pb,
                                          double Kf, double Kck, double *qi, double *pi)
{
#pragma unroll
    for (int n = 0; n < NS; n++) {
        if (((mf | mb) >> n) & 1u) {
            const int nf = (int)((pf >> (4 * n)) & 15ull), nb = (int)((pb >> (4 * n)) & 15ull);
            qi[n] += nb * Kf + nf * Kck;
            pi[n] += nf * Kf + nb * Kck;
        }
    }
}

template <int NS, class Eff>
__device__ __forceinline__ void rxn_step(const DevRxn R, Eff eff, const RxnCtx &c, const double *gk,
                                         const double *m_cm, double *qi, double *pi)
{
    double Kf, Kck;
    rxn_rates<NS>(R, eff, c, gk, m_cm, Kf, Kck);
    rxn_accum<NS>(R.mf, R.mb, R.pf, R.pb, Kf, Kck, qi, pi);
}

// Reaction r of compiled mechanism K: its record as a compile-time constant
template <int NS, int K, int r>
__device__ __forceinline__ void rxn_spec(const RxnCtx &c, const double *gk, const double *m_cm, double *qi,
                                         double *pi)
{
    constexpr DevRxn R = MechSpec<K>::rx(r);
    rxn_step<NS>(R, [](int s) { return MechSpec<K>::eff(r, s); }, c, gk, m_cm, qi, pi);
}
template <int NS, int K, int... r>
__device__ __forceinline__ void rxns_spec(std::integer_sequence<int, r...>, const RxnCtx &c, const double *gk,
                                          const double *m_cm, double *qi, double *pi)
{
    (rxn_spec<NS, K, r>(c, gk, m_cm, qi, pi), ...);
}

// evalcv (y[0] = T): q, d of length NS + 1.  SPEC >= 0: the reactions of
// compiled mechanism SPEC, unrolled with their records as constants (the
// caller has checked that the device tables hold exactly those records);
// SPEC < 0: the loop over the records in the device tables.
template <int NS, int SPEC = -1>
__device__ __forceinline__ void evalcv_dev(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc,
                                           const double *y, double *q, double *d, double m_dens, double m_p)
{
    const double m_tmp = y[0];
    const double _dens = 1.0 / m_dens;
    // species Gibbs energy and internal-energy ingredients at T (Ri = 1).
    // When every lane of the wave uses the same fit range (the usual case),
    // the coefficient reads take that range as a wave-uniform value: scalar
    // loads, shared by the wave, instead of one vector load per lane.
    const TPre t = th->thermo ? tpre<true>(m_tmp, 1) : tpre<true>(m_tmp, 0);
    const int r0 = __builtin_amdgcn_readfirstlane(t.r);
    const bool uni = __all(t.r == r0);
    TPre tu = t;
    tu.r = r0;
    double h1[NS], gk[NS];
    auto species = [&](const TPre &tt) {
#pragma unroll
        for (int n = 0; n < NS; n++) {
            h1[n] = h_k(th, tt, 1.0, n);
            double temp = 0.0;
            temp += s_k(th, tt, 1.0, n);
            temp += (-h1[n] / m_tmp);
            gk[n] = temp;
        }
    };
    if (uni) species(tu);
    else species(t);
    const double logStandConc = dlog(m_p / (R0 * m_tmp));
    // log T: the fit's log max(T, 200) when T >= 200 (the same call on the same value)
    const double logT = m_tmp >= 200.0 ? t.lT : dlog(m_tmp);
    const double logP = dlog(m_p * (1.0 / 101325.0));
    double m_cm[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) m_cm[n] = y[1 + n] * th->_Wi[n] * m_dens * 1.0E-3;
    double *qi = q + 1, *pi = d + 1;
#pragma unroll
    for (int n = 0; n < NS; n++) qi[n] = 0.0, pi[n] = 0.0;
    RxnCtx c;
    c.m_tmp = m_tmp, c.logT = logT, c.logP = logP, c.logStandConc = logStandConc;
    c.rT = 1.0 / m_tmp;
    c.t_ok = __all(m_tmp >= 0x1p-500 && m_tmp <= 0x1p500);
    if constexpr (SPEC >= 0) {
        static_assert(MechSpec<SPEC>::NS == NS, "mechanism specialisation for another species count");
        rxns_spec<NS, SPEC>(std::make_integer_sequence<int, MechSpec<SPEC>::NR>{}, c, gk, m_cm, qi, pi);
    } else {
        const int NR = mc->NR;
        for (int r = 0; r < NR; r++) {
            // the uniform record, read in one burst of scalar loads
            const DevRxn R = mc->rx[r];
            const double *er = mc->third + r * NS;
            rxn_step<NS>(R, [er](int it) { return er[it]; }, c, gk, m_cm, qi, pi);
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
    double cv;
    if (uni) cv = mix_cv<NS>(th, y + 1, tu);
    else cv = mix_cv<NS>(th, y + 1, t);
    const double _Cv = 1.0 / cv;
    q[0] = q0 * _Cv, d[0] = d0 * _Cv;
}

// evalcv with the work of one cell spread over the 64 lanes of a wave, every
// input uniform across the wave and every output left uniform: lane n < NS
// evaluates species n's enthalpy and Gibbs energy, lane r < NR reaction r's
// forward and reverse rates; both are broadcast, and every lane then sums
// the species' production and loss over the reactions in increasing order
// and finishes the temperature terms exactly as evalcv_dev.  The operations
// on every value are evalcv_dev's, so the results are its bits (the table
// form of the rates, which equals every compiled mechanism's).
template <int NS>
__device__ __forceinline__ void evalcv_coop(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc,
                                            const double *y, double *q, double *d, double m_dens, double m_p)
{
    const int ln = threadIdx.x & 63;
    const double m_tmp = y[0];
    const double _dens = 1.0 / m_dens;
    const TPre t = th->thermo ? tpre<true>(m_tmp, 1) : tpre<true>(m_tmp, 0);
    double h1[NS], gk[NS];
    {
        double h = 0.0, gv = 0.0;
        if (ln < NS) {
            h = h_k(th, t, 1.0, ln);
            double temp = 0.0;
            temp += s_k(th, t, 1.0, ln);
            temp += (-h / m_tmp);
            gv = temp;
        }
#pragma unroll
        for (int n = 0; n < NS; n++) h1[n] = __shfl(h, n), gk[n] = __shfl(gv, n);
    }
    const double logStandConc = dlog(m_p / (R0 * m_tmp));
    // log T: the fit's log max(T, 200) when T >= 200 (the same call on the same value)
    const double logT = m_tmp >= 200.0 ? t.lT : dlog(m_tmp);
    const double logP = dlog(m_p * (1.0 / 101325.0));
    double m_cm[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) m_cm[n] = y[1 + n] * th->_Wi[n] * m_dens * 1.0E-3;
    RxnCtx c;
    c.m_tmp = m_tmp, c.logT = logT, c.logP = logP, c.logStandConc = logStandConc;
    c.rT = 1.0 / m_tmp;
    c.t_ok = m_tmp >= 0x1p-500 && m_tmp <= 0x1p500;
    const int NR = mc->NR;
    double Kf = 0.0, Kck = 0.0;
    if (ln < NR) {
        const DevRxn R = mc->rx[ln];
        const double *er = mc->third + ln * NS;
        rxn_rates<NS>(R, [er](int it) { return er[it]; }, c, gk, m_cm, Kf, Kck);
    }
    double *qi = q + 1, *pi = d + 1;
#pragma unroll
    for (int n = 0; n < NS; n++) qi[n] = 0.0, pi[n] = 0.0;
    for (int r = 0; r < NR; r++) {
        const DevRxn &R = mc->rx[r];
        rxn_accum<NS>(R.mf, R.mb, R.pf, R.pb, __shfl(Kf, r), __shfl(Kck, r), qi, pi);
    }
    double q0 = 0.0, d0 = 0.0;
#pragma unroll
    for (int n = 0; n < NS; n++) {
        qi[n] *= th->Wi[n] * _dens;
        pi[n] *= th->Wi[n] * _dens;
        const double ei = (h1[n] - m_tmp) * th->Ri[n];
        q0 -= qi[n] * ei, d0 -= pi[n] * ei;
    }
    const double cv = mix_cv<NS>(th, y + 1, t);
    const double _Cv = 1.0 / cv;
    q[0] = q0 * _Cv, d[0] = d0 * _Cv;
}

__device__ __forceinline__ double alpha_of(double r)
{
    return (180.0 + r * (60.0 + r * (11.0 + r))) / (360.0 + r * (60.0 + r * (12.0 + r)));
}

// Chemeq2 state of one cell, advanced one rate evaluation at a time so that
// a lane can also work through a queue of cells (k_chem_queue).  Every
// advance is: before() -> evalcv_dev(y) -> after().  The first advance of a
// cell is the start-up (clip and renormalise Y, initial step from the rates,
// ProductionRates.h:488-530); each later one is a predictor/corrector pass
// (ProductionRates.h:531-638).  The sequence of operations per cell is the
// reference's; only the place of the rate evaluation in the code is shared.
//
// Where the saved state (ys, qs, ds) and rtau live is a policy: registers for
// the one-lane-per-cell kernel, LDS for the queue kernel, whose persistent
// one-wave blocks would otherwise spill them to scratch memory at two waves
// per SIMD.  Only the storage differs; the arithmetic is the same.
template <int NL> struct ChemRegs {
    double a[4][NL];
    __device__ __forceinline__ double &at(int k, int i) { return a[k][i]; }
};
// arrays k = 0..3 of lane `lane`: p[(k * NL + i) * 64], lane-contiguous rows
// (consecutive lanes hit consecutive banks)
template <int NL> struct ChemLds {
    double *p;
    __device__ __forceinline__ double &at(int k, int i) { return p[(k * NL + i) * 64]; }
};

template <int NS, class St = ChemRegs<NS + 1>, int SPEC = -1> struct ChemLane {
    static constexpr int NL = NS + 1;
    double y[NL], q[NL], d[NL];
    St st;
    double dt, dtg, dtmin, epsmax, ts, tn, rho, m_p;
    int passes;
    bool started;
    __device__ __forceinline__ double &ys(int i) { return st.at(0, i); }
    __device__ __forceinline__ double &qs(int i) { return st.at(1, i); }
    __device__ __forceinline__ double &ds(int i) { return st.at(2, i); }
    __device__ __forceinline__ double &rtau(int i) { return st.at(3, i); }

    __device__ __forceinline__ void begin(double dtg_, double rho_, double p_)
    {
        dtg = dtg_, rho = rho_, m_p = p_;
        dtmin = 1.0e-7, epsmax = 1.0;
        started = false;
        passes = 0;
    }

    // returns -1 when the pass count reaches the safety bound (no evaluation
    // is needed then), 0 otherwise
    __device__ __forceinline__ int before()
    {
        const double ymin = 1.0e-20;
        if (!started) {
            double sumy = 0.0;
#pragma unroll
            for (int i = 1; i < NL; i++) y[i] = dmax(y[i], ymin), sumy += y[i];
            sumy = 1.0 / sumy;
#pragma unroll
            for (int i = 1; i < NL; i++) y[i] *= sumy;
            return 0;
        }
        if (++passes > CHEM_MAXSUB) return -1;
        // predictor
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const double rt = dt * ds(i) / ys(i);
            rtau(i) = rt;
            const double alpha = alpha_of(rt);
            const double scr = dt * (qs(i) - ds(i)) / (1.0 + alpha * rt);
            y[i] = dmax(ys(i) + scr, ymin);
        }
        tn = ts + dt;
        return 0;
    }

    // after the rate evaluation at y: 1 = reached dtg, 0 = continue
    __device__ __forceinline__ int after()
    {
        const double ymin = 1.0e-20, tfd = 1.0 + 1.0e-10, epscl = 1.0e4;
        if (!started) {
            const double epsmin = 1.0e-4, sqreps = 0.05;
            double scrtch = 1e-25;
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
            for (int i = 0; i < NL; i++) ys(i) = y[i], qs(i) = q[i], ds(i) = d[i];
            started = true;
            return 0;
        }
        // corrector fused with the convergence test
        double eps = 1e-10;
#pragma unroll
        for (int i = 0; i < NL; i++) {
            const double yp = y[i];
            const double rtaub = 0.5 * (rtau(i) + dt * d[i] / yp);
            const double alpha = alpha_of(rtaub);
            const double ysi = ys(i);
            const double qt = (1.0 - alpha) * qs(i) + alpha * q[i];
            const double scr = (qt * dt - rtaub * ysi) / (1.0 + alpha * rtaub);
            const double scr2 = dmax(ysi + scr, 0.0);
            double scr1 = fabs(scr2 - yp);
            y[i] = dmax(scr2, ymin);
            if ((0.25 * (ysi + y[i])) > ymin) {
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
        for (int i = 0; i < NL; i++) ys(i) = y[i], qs(i) = q[i], ds(i) = d[i];
        return 0;
    }

    // one advance: 1 = done, 0 = continue, -1 = safety bound
    __device__ __forceinline__ int advance(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc)
    {
        if (before() < 0) return -1;
        evalcv_dev<NS, SPEC>(th, mc, y, q, d, rho, m_p);
        return after();
    }
    // the same advance by a whole wave on one cell (state uniform across
    // the wave): the rate evaluation spread over the lanes
    __device__ __forceinline__ int advance_coop(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc)
    {
        if (before() < 0) return -1;
        evalcv_coop<NS>(th, mc, y, q, d, rho, m_p);
        return after();
    }
};

// Everything of a cell's CHEMEQ2 run between two passes, so that a cell the
// queue lanes hand over continues on a whole wave (k_chem_coop) from exactly
// where it stopped: y, q, d, ys, qs, ds, rtau, the step state, the start
// energy, the cell and its case.
template <int NS> struct ChemSpill {
    static constexpr int NL = NS + 1;
    static constexpr int NV = 7 * NL + 9;
    double v[NV];
    long long id, tl;
    int passes, started, k, pad;
};

template <int NS, class St, int SPEC>
__device__ __forceinline__ void chem_spill_save(ChemSpill<NS> &r, ChemLane<NS, St, SPEC> &L, double e0, long long id,
                                                long long tl, int k)
{
    constexpr int NL = NS + 1;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        r.v[i] = L.y[i], r.v[NL + i] = L.q[i], r.v[2 * NL + i] = L.d[i];
        r.v[3 * NL + i] = L.ys(i), r.v[4 * NL + i] = L.qs(i), r.v[5 * NL + i] = L.ds(i), r.v[6 * NL + i] = L.rtau(i);
    }
    double *w = r.v + 7 * NL;
    w[0] = L.dt, w[1] = L.dtg, w[2] = L.dtmin, w[3] = L.epsmax, w[4] = L.ts, w[5] = L.tn, w[6] = L.rho, w[7] = L.m_p;
    w[8] = e0;
    r.id = id, r.tl = tl, r.passes = L.passes, r.started = L.started ? 1 : 0, r.k = k;
}

template <int NS, class St, int SPEC>
__device__ __forceinline__ double chem_spill_load(const ChemSpill<NS> &r, ChemLane<NS, St, SPEC> &L)
{
    constexpr int NL = NS + 1;
#pragma unroll
    for (int i = 0; i < NL; i++) {
        L.y[i] = r.v[i], L.q[i] = r.v[NL + i], L.d[i] = r.v[2 * NL + i];
        L.ys(i) = r.v[3 * NL + i], L.qs(i) = r.v[4 * NL + i], L.ds(i) = r.v[5 * NL + i], L.rtau(i) = r.v[6 * NL + i];
    }
    const double *w = r.v + 7 * NL;
    L.dt = w[0], L.dtg = w[1], L.dtmin = w[2], L.epsmax = w[3], L.ts = w[4], L.tn = w[5], L.rho = w[6], L.m_p = w[7];
    L.passes = r.passes, L.started = r.started != 0;
    return w[8];
}

// Chemeq2 on one cell, its saved state at lds (lane-strided, ChemLds).
// Returns predictor passes, or -1 at the safety bound.
template <int NS, int SPEC>
__device__ int chemeq2_dev(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, double *y, double dtg,
                           double rho, double m_p, double *lds)
{
    ChemLane<NS, ChemLds<NS + 1>, SPEC> L;
    L.st.p = lds;
#pragma unroll
    for (int i = 0; i < NS + 1; i++) L.y[i] = y[i];
    L.begin(dtg, rho, m_p);
    int st;
    while ((st = L.advance(th, mc)) == 0) {
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

// The mass fractions a cell's chemistry starts from: the y array of the
// Reaction's UpdateStates (ChemeODEQ2SolverKernel1 reads it,
// Reaction_kernels.hpp:58-60) -- Y_k = U_{5+k} / rho and the remainder, or for
// a GhostSpecies mixture the stored renormalised planes and 0
template <int NS>
__device__ __forceinline__ void chem_y0(const double *U, const double *ypl, int ghost, long long N, long long id,
                                        double rho1, double *yi)
{
    if (ghost) {
        yi[NS - 1] = 0.0;
#pragma unroll
        for (int k = 0; k < NS - 1; k++) yi[k] = ypl[k * N + id];
    } else {
        yi[NS - 1] = 1.0;
#pragma unroll
        for (int k = 0; k < NS - 1; k++) {
            yi[k] = U[(5 + k) * N + id] * rho1;
            yi[NS - 1] += -yi[k];
        }
    }
}

// One interior cell of ChemeODEQ2SolverKernel1.
template <int NS, int SPEC>
__device__ __forceinline__ int chem_cell(const Geom &g, const DevThermo *__restrict__ th,
                                         const DevMech *__restrict__ mc, Fields &f, long long N, long long id,
                                         double dtc, double *lds)
{
    // the cell's CHEMEQ2 state held directly (no copy of y in and out)
    ChemLane<NS, ChemLds<NS + 1>, SPEC> L;
    L.st.p = lds;
    L.y[0] = f.T[id];
    const double rho = f.U[id]; // rho = U[0] (Getrhoyi)
    const double rho1 = 1.0 / rho;
    chem_y0<NS>(f.U, f.y, g.ghost, N, id, rho1, L.y + 1);
    const double e0 = mix_e<NS>(th, L.y + 1, L.y[0]);
    L.begin(dtc, rho, f.p[id]);
    int st;
    while ((st = L.advance(th, mc)) == 0) {
    }
    const double e1 = mix_e<NS>(th, L.y + 1, L.y[0]);
#pragma unroll
    for (int n = 0; n < NS - 1; n++) f.U[(5 + n) * N + id] = L.y[1 + n] * rho;
    f.T[id] = L.y[0];
    f.U[4 * N + id] += (e1 - e0) * rho;
    return st < 0 ? -1 : L.passes;
}

// EstimateNAN after the Reaction (flag 4, src/XFLUIDS.cpp:639;
// EstimateFluidNANKernel, src/Fluids.cpp:47-87) on the cell the chemistry has
// just written: U[0] < 0 or any component NaN/Inf.  The reference also reads
// the third stage's LU there, which that stage's own check passed and nothing
// has changed since.
__device__ __forceinline__ bool cell_bad(const double *U, long long N, long long id, int E)
{
    bool bad = U[id] < 0;
    for (int n = 0; n < E; n++) {
        const double a = U[n * N + id];
        bad = bad || isnan(a) || isinf(a);
    }
    return bad;
}

// two waves per SIMD for the 9-species state (<= 256 registers per lane)
template <int NS> struct ChemOccupancy {
    static constexpr int waves = NS <= 3 ? 3 : 2;
};

// LDS of one one-wave block: ys, qs, ds and rtau of its 64 lanes, 20 KiB at
// NS = 9, so the 8 blocks of two waves per SIMD fill the CU's 160 KiB
template <int NS> struct ChemQueueLds {
    static constexpr int doubles = 4 * (NS + 1) * 64;
};

// one lane per interior cell, one-wave blocks; check = 1 (a time step's
// Reaction) also runs the post-Reaction NaN check on the cell.  The saved
// CHEMEQ2 state (ys, qs, ds, rtau) lives in LDS: in registers, with the rate
// evaluation's, it needs ~380 per lane and the kernel ran one wave per SIMD.
template <int NS, int SPEC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ChemOccupancy<NS>::waves)))
k_chem(Geom g, const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, Fields f, int record, int check)
{
    __shared__ double lds[ChemQueueLds<NS>::doubles];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)g.X_inner * g.Y_inner * g.Z_inner;
    if (t >= total) return;
    if (step_halted(f.err, ~EBIT_CHEM)) return;
    const int i = g.Bx + (int)(t % g.X_inner);
    const int j = g.By + (int)((t / g.X_inner) % g.Y_inner);
    const int k = g.Bz + (int)(t / ((long long)g.X_inner * g.Y_inner));
    const long long id = (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
    const int passes = chem_cell<NS, SPEC>(g, th, mc, f, g.N, id, f.dt[DT_CHEM], lds + (threadIdx.x & 63));
    if (record) f.substeps[t] = passes;
    int bits = passes < 0 ? EBIT_CHEM : 0;
    if (passes < 0) report_cell(f.err, EBIT_CHEM, id);
    if (check && cell_bad(f.U, g.N, id, NS + 4)) {
        report_cell(f.err, EBIT_NAN, id);
        bits |= EBIT_NAN;
    }
    if (bits) report_pass(f.err, bits, 4);
}

// Interior cell t (0 <= t < total) -> padded-grid id
__device__ __forceinline__ long long chem_cell_id(const Geom &g, long long t)
{
    const int i = g.Bx + (int)(t % g.X_inner);
    const int j = g.By + (int)((t / g.X_inner) % g.Y_inner);
    const int k = g.Bz + (int)(t / ((long long)g.X_inner * g.Y_inner));
    return (long long)g.Xmax * g.Ymax * k + (long long)g.Xmax * j + i;
}

// The part of one case the chemistry step touches.  `off` is the index of
// the case's first interior cell in the concatenated cell list of an
// ensemble (0 for a single case).
struct ChemCase {
    long long off;
    Geom g;
    double *U, *T, *p, *dt;
    int *substeps, *err;
    double *y; // GhostSpecies mass-fraction planes (g.ghost), else unused
};

// Write-back of a finished cell (ChemeODEQ2SolverKernel1's tail,
// Reaction_kernels.hpp:60-73): Ns-1 partial densities, T, the energy
// change; the pass count; the pass bound and post-Reaction NaN check.
template <int NS, class St, int SPEC>
__device__ __forceinline__ void chem_store(const DevThermo *__restrict__ th, const ChemCase &cc, long long id,
                                           long long tl, ChemLane<NS, St, SPEC> &L, double e0, int st, int record,
                                           int check)
{
    const long long N = cc.g.N;
    const double rho = L.rho;
    const double e1 = mix_e<NS>(th, L.y + 1, L.y[0]);
#pragma unroll
    for (int n = 0; n < NS - 1; n++) cc.U[(5 + n) * N + id] = L.y[1 + n] * rho;
    cc.T[id] = L.y[0];
    cc.U[4 * N + id] += (e1 - e0) * rho;
    if (record) cc.substeps[tl] = st < 0 ? -1 : L.passes;
    int bits = st < 0 ? EBIT_CHEM : 0;
    if (st < 0) report_cell(cc.err, EBIT_CHEM, id);
    if (check && cell_bad(cc.U, N, id, NS + 4)) {
        report_cell(cc.err, EBIT_NAN, id);
        bits |= EBIT_NAN;
    }
    if (bits) report_pass(cc.err, bits, 4);
}

// Cells that the queue's lanes hand to whole waves: records, their count
// (count[0], may exceed cap: the lanes that found the list full kept their
// cells) and the waves' pick counter (count[1]); a cell is handed over once
// it has run `after` passes.  cap = 0: no hand-over.
struct ChemSpillBuf {
    void *rec;
    unsigned *count;
    int cap, after;
};

// Work-queue driver: persistent lanes draining the interior cells of one or
// several cases.  Cells of very different stiffness take from 1 to ~10^3
// passes; with one lane per cell a wave runs as long as its stiffest cell
// while the other lanes idle.  Here every lane that has finished a cell takes
// the next one from a device counter (one atomic per wave, handed out by lane
// rank), so the lanes of a wave stay busy with the same code until the queue
// is empty; a cell that starts and the cells that continue share one rate
// evaluation per iteration.  A cell still running after sp.after passes is
// handed to the whole-wave kernel (k_chem_coop) while the list has
// room: a few very stiff cells would otherwise keep single lanes busy, one
// serial rate evaluation after another, long after the rest.  Each cell's
// arithmetic is exactly that of the per-cell driver.  Termination: the
// counter only grows, a wave stops once every lane has drawn an index >=
// total and holds no cell, and a cell is bounded by CHEM_MAXSUB passes.  The
// counters must be zero at launch.
template <int NS, int SPEC, class Cases>
__device__ __forceinline__ void chem_queue_run(const Cases &cases, long long total, const DevThermo *__restrict__ th,
                                               const DevMech *__restrict__ mc, unsigned long long *queue, int record,
                                               int check, double *lds, ChemSpillBuf sp)
{
    const int lane = threadIdx.x & 63;
    ChemLane<NS, ChemLds<NS + 1>, SPEC> L;
    L.st.p = lds + lane;
    // a lane keeps only its cell's case index, id and interior index across
    // passes; the case record is read again where it is needed (for an
    // ensemble it differs per lane, and holding it would cost ~40 registers)
    long long tl = -1, id = 0;
    int ck = 0;
    double e0 = 0.0;
    bool active = false, done = false, can_spill = sp.cap > 0;
    for (;;) {
        const bool need = !active && !done;
        const unsigned long long want = __ballot(need);
        if (want) {
            const int cnt = __popcll(want);
            const int leader = __ffsll((long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(queue, (unsigned long long)cnt);
            base = __shfl(base, leader);
            if (need) {
                const int rank = __popcll(want & ((1ull << lane) - 1ull));
                const long long t = (long long)(base + (unsigned long long)rank);
                ChemCase cc{};
                if (t < total) {
                    ck = cases.find_k(t);
                    cc = cases.get(ck);
                }
                // the cells of a case whose step has halted are skipped
                if (t < total && !step_halted(cc.err, ~EBIT_CHEM)) {
                    tl = t - cc.off;
                    const long long N = cc.g.N;
                    id = chem_cell_id(cc.g, tl);
                    L.y[0] = cc.T[id];
                    const double rho = cc.U[id];
                    const double rho1 = 1.0 / rho;
                    chem_y0<NS>(cc.U, cc.y, cc.g.ghost, N, id, rho1, L.y + 1);
                    e0 = mix_e<NS>(th, L.y + 1, L.y[0]);
                    L.begin(cc.dt[DT_CHEM], rho, cc.p[id]);
                    active = true;
                } else if (t >= total) {
                    done = true;
                }
            }
        }
        if (!__any(active)) break;
        if (active) {
            // starting cells and running cells share this one evaluation
            const int st = L.advance(th, mc);
            if (st != 0) {
                chem_store<NS>(th, cases.get(ck), id, tl, L, e0, st, record, check);
                active = false;
            } else if (can_spill && L.passes >= sp.after) {
                const unsigned idx = atomicAdd(sp.count, 1u);
                if (idx < (unsigned)sp.cap) {
                    chem_spill_save(((ChemSpill<NS> *)sp.rec)[idx], L, e0, id, tl, ck);
                    active = false;
                } else {
                    can_spill = false; // the list is full: keep this cell here
                }
            }
        }
    }
}

// The cells the queue handed over, one whole wave per cell (evalcv_coop):
// persistent one-wave blocks take records from the list until it is
// drained, and write each cell back as the queue would have.
template <int NS, class Cases>
__device__ __forceinline__ void chem_coop_run(const Cases &cases, const DevThermo *__restrict__ th,
                                              const DevMech *__restrict__ mc, ChemSpillBuf sp, int record, int check)
{
    const int ln = threadIdx.x & 63;
    const ChemSpill<NS> *rec = (const ChemSpill<NS> *)sp.rec;
    const unsigned n = min(sp.count[0], (unsigned)sp.cap);
    for (;;) {
        unsigned idx = 0;
        if (ln == 0) idx = atomicAdd(sp.count + 1, 1u);
        idx = __shfl(idx, 0);
        if (idx >= n) break;
        ChemLane<NS, ChemRegs<NS + 1>, -1> L;
        const double e0 = chem_spill_load(rec[idx], L);
        const ChemCase cc = cases.get(rec[idx].k);
        int st;
        while ((st = L.advance_coop(th, mc)) == 0) {
        }
        if (ln == 0) chem_store<NS>(th, cc, rec[idx].id, rec[idx].tl, L, e0, st, record, check);
    }
}

struct OneCase {
    ChemCase c;
    __device__ __forceinline__ int find_k(long long) const { return 0; }
    __device__ __forceinline__ ChemCase get(int) const { return c; }
};

// cases[k].off ascending; the case holding cell t by binary search
struct ManyCases {
    const ChemCase *cases;
    int n;
    __device__ __forceinline__ int find_k(long long t) const
    {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cases[mid].off <= t) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }
    __device__ __forceinline__ ChemCase get(int k) const { return cases[k]; }
};


__device__ __forceinline__ OneCase one_case(const Geom &g, const Fields &f)
{
    OneCase one;
    one.c.off = 0;
    one.c.g = g;
    one.c.U = f.U, one.c.T = f.T, one.c.p = f.p, one.c.dt = f.dt;
    one.c.substeps = f.substeps, one.c.err = f.err;
    one.c.y = f.y;
    return one;
}

// the queue counter is queue[0..1] (64 bits); the hand-over counters follow
// (queue[2] records, queue[3] waves' picks)
template <int NS, int SPEC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ChemOccupancy<NS>::waves)))
k_chem_queue(Geom g, const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, Fields f, int record,
             int check, void *spill, int spill_cap, int spill_after)
{
    const OneCase one = one_case(g, f);
    __shared__ double lds[ChemQueueLds<NS>::doubles];
    const ChemSpillBuf sp{spill, f.queue + 2, spill_cap, spill_after};
    chem_queue_run<NS, SPEC>(one, (long long)g.X_inner * g.Y_inner * g.Z_inner, th, mc, (unsigned long long *)f.queue,
                             record, check, lds, sp);
}

// one queue over the cells of every case of an ensemble (shared mechanism)
template <int NS, int SPEC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(ChemOccupancy<NS>::waves)))
k_chem_queue_multi(const ChemCase *__restrict__ cases, int ncases, long long total, const DevThermo *__restrict__ th,
                   const DevMech *__restrict__ mc, unsigned long long *queue, int record, int check, void *spill,
                   int spill_cap, int spill_after)
{
    ManyCases many{cases, ncases};
    __shared__ double lds[ChemQueueLds<NS>::doubles];
    const ChemSpillBuf sp{spill, (unsigned *)queue + 2, spill_cap, spill_after};
    chem_queue_run<NS, SPEC>(many, total, th, mc, queue, record, check, lds, sp);
}

// the handed-over cells of one case / of an ensemble, one wave per cell
template <int NS>
__global__ void __launch_bounds__(64) k_chem_coop(Geom g, const DevThermo *__restrict__ th,
                                                  const DevMech *__restrict__ mc, Fields f, int record, int check,
                                                  void *spill, int spill_cap)
{
    const OneCase one = one_case(g, f);
    chem_coop_run<NS>(one, th, mc, ChemSpillBuf{spill, f.queue + 2, spill_cap, 0}, record, check);
}
template <int NS>
__global__ void __launch_bounds__(64) k_chem_coop_multi(const ChemCase *__restrict__ cases, int ncases,
                                                        const DevThermo *__restrict__ th,
                                                        const DevMech *__restrict__ mc, unsigned long long *queue,
                                                        int record, int check, void *spill, int spill_cap)
{
    ManyCases many{cases, ncases};
    chem_coop_run<NS>(many, th, mc, ChemSpillBuf{spill, (unsigned *)queue + 2, spill_cap, 0}, record, check);
}

// ---------------------------------------------------------------------------
// One cell per wave (XF_CHEM_WAVE).  With few cells (a 1-D tube) the chemistry
// launch lasts as long as the stiffest cell's serial chain of passes, about
// 4500 instructions each on one lane.  Here the wave of a cell spreads every
// pass over its lanes: lane i < NS + 1 holds solution component i (0 = T,
// 1 + n = Y_n) and does that component's predictor and corrector; lane 1 + n
// evaluates species n's enthalpy, Gibbs energy and concentration, lane r < NR
// reaction r's rates, and lane 1 + n sums its species' production and loss
// over the reactions in order.  Every sum and maximum the reference forms over
// components or species (normalisation, step-size scratch, energy terms, Cv,
// the error norm) runs as an ordered chain over the lanes' values, so each
// value is rounded exactly as in ChemLane / evalcv_dev and the results are the
// per-cell driver's bits.  Requires NS + 1 <= 64 and NR <= 64.
// ---------------------------------------------------------------------------

// Cross-lane exchanges of the cell's wave go through LDS: the lanes that own
// values write them, one barrier (a single wave: no wait on other waves), and
// every lane reads them back at the same addresses (LDS broadcasts).  Sums are
// then formed on every lane in the reference's order.
template <int NS> struct WaveShared {
    double K[2 * MAXNR];    // Kf, Kck of each reaction
    double S[2 * MAXNS];    // Gibbs energy and concentration of each species
    double V[4 * 64];       // per-lane operands of the ordered sums
    // species lane ln's nu_f | nu_b << 4 in reaction r at nu[64 * r + ln]
    // (filled once per cell; each lane reads only its own entries)
    unsigned char nu[64 * MAXNR];
};

// ((init + v[lo]) + v[lo+1]) + ... of a lane-indexed LDS array with stride st
__device__ __forceinline__ double lds_sum(const double *v, int st, int lo, int hi, double init)
{
    double s = init;
    for (int k = lo; k < hi; k++) s += v[k * st];
    return s;
}
// ((init - v[lo]) - v[lo+1]) - ...
__device__ __forceinline__ double lds_sub(const double *v, int st, int lo, int hi, double init)
{
    double s = init;
    for (int k = lo; k < hi; k++) s -= v[k * st];
    return s;
}

// What a lane reads on every pass, loaded once per cell: species lane 1 + n
// keeps W, 1/W and R of species n and the fit coefficients of the current
// range (reloaded when T crosses into another range); reaction lane r keeps
// its record and third-body efficiencies.
template <int NS> struct WaveLane {
    double Wi = 0.0, iWi = 0.0, Ri = 0.0;
    int rc = -2; // range the coefficients hc belong to
    double hc[9];
    DevRxn R{};
    double eff[NS];

    __device__ __forceinline__ void load(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc, int NR)
    {
        const int ln = threadIdx.x & 63;
#pragma unroll
        for (int m = 0; m < 9; m++) hc[m] = 0.0;
        if (ln >= 1 && ln <= NS) Wi = th->Wi[ln - 1], iWi = th->_Wi[ln - 1], Ri = th->Ri[ln - 1];
#pragma unroll
        for (int it = 0; it < NS; it++) eff[it] = 0.0;
        if (ln < NR) {
            R = mc->rx[ln];
#pragma unroll
            for (int it = 0; it < NS; it++) eff[it] = mc->third[ln * NS + it];
        }
    }
    // the fit coefficients of range r (the same on every lane: T is the cell's)
    __device__ __forceinline__ void range(const DevThermo *__restrict__ th, int r)
    {
        if (r != rc) {
            const int ln = threadIdx.x & 63;
            if (ln >= 1 && ln <= NS) {
                const double *c = fit_coeffs(th, r, ln - 1);
#pragma unroll
                for (int m = 0; m < 9; m++) hc[m] = c[m];
            }
            rc = r;
        }
    }
};

// evalcv (ProductionRates.h:459-479) at the distributed state: yv is
// component ln, m_tmp = T; returns the component's q and d
template <int NS, int THERMO>
__device__ __forceinline__ void evalcv_wave(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc,
                                            WaveLane<NS> &W, WaveShared<NS> &sh, int NR, double yv, double m_tmp,
                                            double &qv, double &dv, double m_dens, double m_p, double _dens,
                                            double logP)
{
    const int ln = threadIdx.x & 63;
    const int n = ln - 1;
    const bool sp = ln >= 1 && ln <= NS;
    // THERMO: the fit family (1 NASA-9, 0 JANAF) as a constant, so that each
    // formula reads the coefficients at fixed positions (kept in registers)
    constexpr int thermo = THERMO;
    const TPre t = tpre<true>(m_tmp, thermo);
    W.range(th, t.r);
    double h1 = 0.0;
    if (sp) {
        // h_k and s_k with Ri = 1
        h1 = h_fit(W.hc, thermo, t, 1.0);
        if (t.TT < 200.0) h1 += cp200_k(th, 1.0, n) * (t.TT - 200.0);
        double temp = 0.0;
        temp += s_fit(W.hc, thermo, t, 1.0);
        temp += (-h1 / m_tmp);
        sh.S[n] = temp;
        sh.S[MAXNS + n] = yv * W.iWi * m_dens * 1.0E-3;
    }
    const double logStandConc = dlog(m_p / (R0 * m_tmp));
    // log T: the fit's log max(T, 200) when T >= 200 (the same call on the same value)
    const double logT = m_tmp >= 200.0 ? t.lT : dlog(m_tmp);
    __syncthreads();
    double Kf = 0.0, Kck = 0.0;
    if (ln < NR) {
        double gka[NS], mca[NS];
#pragma unroll
        for (int k = 0; k < NS; k++) gka[k] = sh.S[k], mca[k] = sh.S[MAXNS + k];
        RxnCtx c;
        c.m_tmp = m_tmp, c.logT = logT, c.logP = logP, c.logStandConc = logStandConc;
        c.rT = m_tmp >= 200.0 ? t.iT : 1.0 / m_tmp; // the fit's 1 / max(T, 200) when T >= 200
        c.t_ok = m_tmp >= 0x1p-500 && m_tmp <= 0x1p500;
        rxn_rates<NS>(W.R, [&W](int it) { return W.eff[it]; }, c, gka, mca, Kf, Kck);
        sh.K[ln] = Kf, sh.K[MAXNR + ln] = Kck;
    }
    __syncthreads();
    // species n's sums over the reactions, in increasing order (rxn_accum)
    double qi = 0.0, pi = 0.0;
    if (sp) {
        // a reaction adds to species n exactly when its nu_f or nu_b is non-zero
        for (int r = 0; r < NR; r++) {
            const unsigned b = sh.nu[64 * r + ln];
            if (b) {
                const int nf = (int)(b & 15u), nb = (int)(b >> 4);
                const double kf = sh.K[r], kc = sh.K[MAXNR + r];
                qi += nb * kf + nf * kc;
                pi += nf * kf + nb * kc;
            }
        }
        qi *= W.Wi * _dens;
        pi *= W.Wi * _dens;
        const double ei = (h1 - m_tmp) * W.Ri;
        sh.V[4 * ln + 0] = qi * ei;
        sh.V[4 * ln + 1] = pi * ei;
        sh.V[4 * ln + 2] = yv * cp_fit(W.hc, thermo, t, W.Ri);
        sh.V[4 * ln + 3] = yv * W.iWi;
    }
    __syncthreads();
    const double q0 = lds_sub(sh.V + 0, 4, 1, NS + 1, 0.0);
    const double d0 = lds_sub(sh.V + 1, 4, 1, NS + 1, 0.0);
    // mix_cv: sum Y_n Cp_n, sum Y_n / W_n, both in species order
    double cv = lds_sum(sh.V + 2, 4, 1, NS + 1, 0.0);
    const double _W = lds_sum(sh.V + 3, 4, 1, NS + 1, 0.0);
    __syncthreads(); // V is written again before the next read
    cv -= Ru * _W;
    const double _Cv = 1.0 / cv;
    qv = ln == 0 ? q0 * _Cv : qi;
    dv = ln == 0 ? d0 * _Cv : pi;
}

// get_Cope at the distributed state (mix_e's order)
template <int NS, int THERMO>
__device__ __forceinline__ double mix_e_wave(const DevThermo *__restrict__ th, WaveShared<NS> &sh, double yv, double T)
{
    const int ln = threadIdx.x & 63;
    if (ln >= 1 && ln <= NS) {
        const TPre t = THERMO ? tpre<true>(T, 1) : tpre<false>(T, 0);
        const double ei = (h_k(th, t, 1.0, ln - 1) - T) * th->Ri[ln - 1];
        sh.V[ln] = ei * yv;
    }
    __syncthreads();
    const double e = lds_sum(sh.V, 1, 1, NS + 1, 0.0);
    __syncthreads();
    return e;
}

// T, held by lane 0, to every lane
template <int NS> __device__ __forceinline__ double wave_T(WaveShared<NS> &sh, double yv)
{
    if ((threadIdx.x & 63) == 0) sh.V[0] = yv;
    __syncthreads();
    const double T = sh.V[0];
    __syncthreads();
    return T;
}

// ChemeODEQ2SolverKernel1 + Chemeq2 for interior cell tl of case cc on one
// wave; every branch below is uniform across the wave
template <int NS, int THERMO>
__device__ __forceinline__ void chem_wave_cell(const DevThermo *__restrict__ th, const DevMech *__restrict__ mc,
                                               const ChemCase &cc, long long tl, int record, int check,
                                               WaveShared<NS> &sh)
{
    constexpr int NL = NS + 1;
    const int ln = threadIdx.x & 63;
    const bool comp = ln < NL;
    const long long N = cc.g.N;
    const long long id = chem_cell_id(cc.g, tl);
    const double rho = cc.U[id];
    const double rho1 = 1.0 / rho;
    const double m_p = cc.p[id];
    const int NR = mc->NR;
    // evalcv's 1 / rho and log(p / p_atm): constants of the cell
    const double _dens = 1.0 / rho;
    const double logP = dlog(m_p * (1.0 / 101325.0));
    // the start values: T, then Getrhoyi's fractions (chem_y0)
    double yv = 0.0;
    if (ln == 0) yv = cc.T[id];
    if (ln >= 1 && ln < NS) yv = cc.g.ghost ? cc.y[(ln - 1) * N + id] : cc.U[(4 + ln) * N + id] * rho1;
    if (cc.g.ghost) {
        if (ln == NS) yv = 0.0;
    } else {
        sh.V[ln] = yv;
        __syncthreads();
        const double rem = lds_sub(sh.V, 1, 1, NS, 1.0); // ((1 - Y_0) - Y_1) - ...
        __syncthreads();
        if (ln == NS) yv = rem;
    }
    const double e0 = mix_e_wave<NS, THERMO>(th, sh, yv, wave_T<NS>(sh, yv));
    WaveLane<NS> W;
    W.load(th, mc, NR);
    for (int r = 0; r < NR; r++) {
        unsigned v = 0;
        if (ln >= 1 && ln <= NS) {
            const unsigned long long pf = mc->rx[r].pf, pb = mc->rx[r].pb;
            v = (unsigned)((pf >> (4 * (ln - 1))) & 15ull) | ((unsigned)((pb >> (4 * (ln - 1))) & 15ull) << 4);
        }
        sh.nu[64 * r + ln] = (unsigned char)v;
    }
    // Chemeq2 (ProductionRates.h:488-638), the steps of ChemLane
    const double ymin = 1.0e-20, tfd = 1.0 + 1.0e-10, epscl = 1.0e4;
    const double dtg = cc.dt[DT_CHEM];
    double dt = 0.0, dtmin = 1.0e-7, epsmax = 1.0, ts = 0.0, tn = 0.0;
    double qv = 0.0, dv = 0.0, ys = 0.0, qs = 0.0, ds = 0.0, rtau = 0.0;
    int passes = 0, st = 0;
    bool started = false;
    for (;;) {
        if (!started) {
            if (ln >= 1 && comp) yv = dmax(yv, ymin);
            sh.V[ln] = yv;
            __syncthreads();
            double sumy = lds_sum(sh.V, 1, 1, NL, 0.0);
            __syncthreads();
            sumy = 1.0 / sumy;
            if (ln >= 1 && comp) yv *= sumy;
        } else {
            if (++passes > CHEM_MAXSUB) {
                st = -1;
                break;
            }
            if (comp) {
                const double rt = dt * ds / ys;
                rtau = rt;
                const double alpha = alpha_of(rt);
                const double scr = dt * (qs - ds) / (1.0 + alpha * rt);
                yv = dmax(ys + scr, ymin);
            }
            tn = ts + dt;
        }
        evalcv_wave<NS, THERMO>(th, mc, W, sh, NR, yv, wave_T<NS>(sh, yv), qv, dv, rho, m_p, _dens, logP);
        if (!started) {
            const double epsmin = 1.0e-4, sqreps = 0.05;
            if (comp) {
                const double ascr = fabs(qv);
                const double sb = 0.1 * epsmin * ascr - dv;
                const double scr2 = (sb >= 0.0 ? 1.0 : -1.0) * fabs(1.0 / yv);
                sh.V[4 * ln + 0] = -fabs(ascr - dv) * scr2;
                sh.V[4 * ln + 1] = scr2 * dv;
            }
            __syncthreads();
            double scrtch = 1e-25;
            for (int i = 0; i < NL; i++) scrtch = dmax(sh.V[4 * i], dmax(sh.V[4 * i + 1], scrtch));
            __syncthreads();
            dt = dmin(sqreps / scrtch, dtg);
            dtmin *= dt;
            ts = 0.0, tn = 0.0;
            ys = yv, qs = qv, ds = dv;
            started = true;
            continue;
        }
        // corrector and the error norm (the maximum in component order)
        if (comp) {
            const double yp = yv;
            const double rtaub = 0.5 * (rtau + dt * dv / yp);
            const double alpha = alpha_of(rtaub);
            const double ysi = ys;
            const double qt = (1.0 - alpha) * qs + alpha * qv;
            const double scr = (qt * dt - rtaub * ysi) / (1.0 + alpha * rtaub);
            const double scr2 = dmax(ysi + scr, 0.0);
            double scr1 = fabs(scr2 - yp);
            yv = dmax(scr2, ymin);
            double cand = 0.0, has = 0.0;
            if ((0.25 * (ysi + yv)) > ymin) {
                scr1 = scr1 / yv;
                cand = 0.5 * (scr1 + dmin(fabs(qv - dv) / (qv + dv + 1.0e-30), scr1));
                has = 1.0;
            }
            sh.V[4 * ln + 0] = cand;
            sh.V[4 * ln + 1] = has;
        }
        __syncthreads();
        double eps = 1e-10;
        for (int i = 0; i < NL; i++)
            if (sh.V[4 * i + 1] != 0.0) eps = dmax(sh.V[4 * i], eps);
        __syncthreads();
        eps = eps * epscl;
        if (eps < epsmax) {
            if (dtg <= (tn * tfd)) {
                st = 1;
                break;
            }
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
            continue;
        }
        epsmax = 1.0;
        ts = tn;
        ys = yv, qs = qv, ds = dv;
    }
    // write-back (ChemeODEQ2SolverKernel1's tail), the pass count, the checks
    const double T1 = wave_T<NS>(sh, yv);
    const double e1 = mix_e_wave<NS, THERMO>(th, sh, yv, T1);
    bool bad = false;
    if (ln >= 1 && ln < NS) {
        const double v = yv * rho;
        cc.U[(4 + ln) * N + id] = v;
        bad = isnan(v) || isinf(v);
    }
    if (ln == 0) {
        cc.T[id] = yv;
        const double u4 = cc.U[4 * N + id] + (e1 - e0) * rho;
        cc.U[4 * N + id] = u4;
        if (record) cc.substeps[tl] = st < 0 ? -1 : passes;
        bad = rho < 0 || isnan(rho) || isinf(rho) || isnan(u4) || isinf(u4);
        for (int k = 1; k < 4; k++) bad = bad || isnan(cc.U[k * N + id]) || isinf(cc.U[k * N + id]);
    }
    const bool any_bad = __any(bad);
    if (ln == 0) {
        int bits = st < 0 ? EBIT_CHEM : 0;
        if (st < 0) report_cell(cc.err, EBIT_CHEM, id);
        if (check && any_bad) {
            report_cell(cc.err, EBIT_NAN, id);
            bits |= EBIT_NAN;
        }
        if (bits) report_pass(cc.err, bits, 4);
    }
}

// one interior cell per one-wave block; THERMO = the context's fit family
template <int NS, int THERMO>
__global__ void __launch_bounds__(64) k_chem_wave(Geom g, const DevThermo *__restrict__ th,
                                                  const DevMech *__restrict__ mc, Fields f, int record, int check)
{
    const long long tl = blockIdx.x;
    if (tl >= (long long)g.X_inner * g.Y_inner * g.Z_inner) return;
    if (step_halted(f.err, ~EBIT_CHEM)) return;
    __shared__ WaveShared<NS> sh;
    chem_wave_cell<NS, THERMO>(th, mc, one_case(g, f).c, tl, record, check, sh);
}

// zero the work-queue counter (a kernel, so the step graph holds no memset)
static __global__ void k_queue_reset(unsigned int *q)
{
    if (threadIdx.x < 4) q[threadIdx.x] = 0u;
}

} // namespace xf
