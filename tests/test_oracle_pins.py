"""Pins of the two oracle functions that no golden vector of the reference
covers: CHEMEQ2's integrator and step control (ProductionRates.h:488-638) and
the multi-species Roe sound speed (ReconstructSoundSpeed +
SoundSpeedMultiSpecies, positive-definite_eigen/Utils_device.hpp:14-140).

CHEMEQ2 runs here on model systems with closed-form solutions, through the
same integrator body the mechanism uses (ora_chemeq2_model):
  * alpha(r) approximates the exact quasi-steady weight
    1/(1 - exp(-r)) - 1/r with an error of order r^5;
  * on dy/dt = -k y the first substep is sqreps / k (sqreps = 0.05), an
    accepted step multiplies the step by 1/sqrt(eps) + 0.005 with the square
    root by two Heron iterations, and each substep of size r/k multiplies y
    by 1 - r/(1 + alpha(r) r): the result follows from those rules exactly;
  * on dy/dt = -k y^2 (exact y0/(1 + k y0 t)) the error control holds the
    relative error near 1/epscl = 1e-4 over hundreds of substeps.
The sound speed's ingredients are checked against the equation of state
itself: dp/drho, dp/drho_n and dp/d(rho E) of one cell by central finite
differences of p(rho, rho Y_n, rho E) through the damped-Newton temperature,
and the jump correction by the identity it is built to satisfy (the
corrected derivatives reproduce the pressure jump of the face).

Each planted mistake (oracle/Makefile `mut`: one constant of CHEMEQ2 or of
the sound speed changed) must fail at least one pin.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import xfluids_amd as xf
from oracle_lib import ROOT, Oracle, config_from_setup, load

P = C.POINTER(C.c_double)
MUTANTS = range(1, 8)


def _bind(L):
    L.ora_alpha.restype = C.c_double
    L.ora_alpha.argtypes = [C.c_double]
    L.ora_chemeq2_model.restype = C.c_int
    L.ora_chemeq2_model.argtypes = [C.c_int, P, C.c_double, P, P, C.c_int, P, P, C.c_int]
    L.ora_face_c2_terms.restype = C.c_double
    L.ora_face_c2_terms.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, P]
    return L


def _lib(path=None):
    return _bind(load(path or os.path.join(ROOT, "oracle", "_build", "liboracle.so")))


def model(L, y0, q, k, dtg, power=1, cap=0):
    y = np.array(y0, float)
    qq, kk = np.array(q, float), np.array(k, float)
    dtf = C.c_double()
    tr = np.zeros(4 * max(cap, 1))
    n = L.ora_chemeq2_model(len(y), y.ctypes.data_as(P), dtg, qq.ctypes.data_as(P), kk.ctypes.data_as(P), power,
                            C.byref(dtf), tr.ctypes.data_as(P), cap)
    if cap:
        return n, y, dtf.value, tr.reshape(-1, 4)[:min(max(n, 0), cap)]
    return n, y, dtf.value


def alpha_exact(r):
    return 1.0 / (-np.expm1(-r)) - 1.0 / r


# ---- CHEMEQ2 pins: each returns a list of failures (empty = pinned) -------
def pin_alpha_order(L):
    bad = []
    rs = [0.4, 0.2, 0.1, 0.05, 0.025]
    err = [L.ora_alpha(r) - alpha_exact(r) for r in rs]
    if abs(err[0]) > 2e-7:
        bad.append("alpha error %.3e at r = 0.4" % err[0])
    for a, b in zip(err, err[1:]):
        if not (26.0 < a / b < 36.0):  # 2^5 = 32: fifth order
            bad.append("alpha error ratio %.2f" % (a / b))
    return bad


def heron_rteps(eps):
    r = 0.5 * (eps + 1.0)
    r = 0.5 * (r + eps / r)
    return 0.5 * (r + eps / r)


def pin_decay_steps(L):
    """dy/dt = -k y from y = 1: first substep sqreps/k; a run of steps that
    grow by 1/sqrt(1e-10 * 1e4) + 0.005 (a linear problem's predictor and
    corrector agree, so the error measure sits at its floor 1e-10); the last
    substep ends on dtg.  y is the product of 1 - r/(1 + alpha(r) r)."""
    bad = []
    k = 2.0
    n, y, dtf = model(L, [1.0], [0.0], [k], 0.049 / k)
    if n != 1:
        bad.append("%d passes for r = 0.049" % n)
    n, y, dtf = model(L, [1.0], [0.0], [k], 0.051 / k)
    if n != 2 or abs(dtf * k - 0.05) > 1e-15:
        bad.append("first substep r = %.6f, %d passes for r = 0.051" % (dtf * k, n))
    g = 1.0 / heron_rteps(1e-10 * 1e4) + 0.005
    tfd = 1.0 + 1.0e-10
    for total in (1.0, 10.0, 40.0):
        dtg = total / k
        dt, tn, yy, steps = 0.05 / k, 0.0, 1.0, 0
        while True:
            r = dt * k
            yy = max(yy + dt * (0.0 - k * yy) / (1.0 + L.ora_alpha(r) * r), 1e-20)
            tn += dt
            steps += 1
            if dtg <= tn * tfd:
                break
            dt = min(dt * g, tfd * (dtg - tn))
        n, y, _ = model(L, [1.0], [0.0], [k], dtg)
        if n != steps or abs(y[0] - yy) > 1e-13 * max(yy, 1e-300):
            bad.append("decay to r = %g: %d passes y %.17g, rule gives %d, %.17g" % (total, n, y[0], steps, yy))
    # a single substep against the exact exponential (the quasi-steady
    # weight's accuracy), relaxation towards q/k
    for r in (0.005, 0.0025):
        k, q, y0 = 3.0, 2.0, 0.1
        n, y, _ = model(L, [y0], [q], [k], r / k)
        ex = q / k + (y0 - q / k) * np.exp(-r)
        if n != 1 or abs((y[0] - ex) / (ex - y0)) > 1e-12:
            bad.append("relaxation step r = %g: %d passes, increment error %.2e" % (r, n, (y[0] - ex) / (ex - y0)))
    return bad


def pin_riccati(L):
    """dy/dt = -k y^2: the relative error over the run stays in
    [1e-5, 2e-4], the band of CHEMEQ2's error target 1/epscl."""
    bad = []
    for T in (0.5, 2.0, 10.0):
        n, y, _ = model(L, [1.0], [0.0], [1.0], T, 2)
        ex = 1.0 / (1.0 + T)
        e = abs(y[0] - ex) / ex
        if not (1e-5 < e < 2e-4) or n < 10:
            bad.append("riccati T = %g: rel error %.2e in %d passes" % (T, e, n))
    return bad


def pin_step_control(L):
    """The step-size law of CHEMEQ2 on every pass of a stiff Riccati run
    (dy/dt = 1 - 50 y^2 from y = 0, and dy/dt = -k y^2): a pass whose
    scaled error measure eps is below the threshold is accepted; the next
    substep is dt (1/sqrt(eps) + 0.005), the square root by two Heron
    iterations from (1 + eps)/2, capped by the rest of the interval after an
    accepted pass and by 0.34 dt after a rejected one.  The run must reject
    at least once with the 0.34 cap binding."""
    bad = []
    tfd = 1.0 + 1.0e-10
    binding = 0
    for (y0, q, k, T) in ((0.0, 1.0, 50.0, 3.0), (1.0, 0.0, 1.0, 10.0)):
        n, y, dtf, tr = model(L, [y0], [q], [k], T, 2, cap=4000)
        if n <= 0 or n > 4000:
            bad.append("riccati run: %d passes" % n)
            continue
        for i in range(n - 1):
            dt, eps, epsmax, tn = tr[i]
            acc = eps < epsmax
            g = dt * (1.0 / heron_rteps(eps) + 0.005)
            if acc:
                want = min(g, tfd * (T - tn))
            else:
                want = min(min(g, tfd * (T - (tn - dt))), 0.34 * dt)
                binding += want == 0.34 * dt
            if tr[i + 1][0] != want:
                bad.append("pass %d: next substep %.17g, the law gives %.17g" % (i, tr[i + 1][0], want))
                break
    if binding == 0:
        bad.append("no rejected pass with the 0.34 cap binding")
    return bad


def chem_pins(L):
    return pin_alpha_order(L) + pin_decay_steps(L) + pin_riccati(L) + pin_step_control(L)


def test_chemeq2_alpha_is_fifth_order():
    assert pin_alpha_order(_lib()) == []


def test_chemeq2_step_rules_and_exact_steps():
    assert pin_decay_steps(_lib()) == []


def test_chemeq2_error_control_on_riccati():
    assert pin_riccati(_lib()) == []


def test_chemeq2_step_size_law():
    assert pin_step_control(_lib()) == []


# ---- sound speed pins ------------------------------------------------------
def _state(case="shock-bubble", nx=48, ny=24, steps=3, L=None):
    s = xf.Setup(xf.case_path(case), nx=nx, ny=ny)
    U, T = s.initial(case, seed=4)
    o = Oracle(config_from_setup(s), L)
    o.set_state(U, T)
    assert o.initialize() == 0
    o.steps(steps)
    o.boundary(0)
    assert o.update_states(0) == 0
    return s, o


def _terms(L, o, il, ir):
    NS = o.cfg.NS
    t = np.zeros(14 + 4 * NS)
    c2 = L.ora_face_c2_terms(o._h, il, ir, t.ctypes.data_as(P))
    ncop = int(t[13])
    return c2, t[:14], t[14:14 + ncop], t[14 + ncop:14 + 2 * ncop], t[14 + 2 * ncop:14 + 3 * ncop]


def _eos_p(o, L, rho, rho_n, rhoE, mom, T0):
    """p of one cell from its conserved variables: Y_n = rho_n / rho (n <
    NS - 1), Y_N = 1 - sum, e = rho E / rho - |m|^2 / (2 rho^2), T by the
    oracle's damped Newton (iterated to its fixed point), p = rho R T."""
    NS = o.cfg.NS
    Y = np.zeros(NS)
    Y[:NS - 1] = rho_n / rho
    Y[NS - 1] = 1.0 - Y[:NS - 1].sum()
    e = rhoE / rho - 0.5 * float(np.dot(mom, mom)) / rho ** 2
    Yp = Y.ctypes.data_as(P)
    T = T0
    for _ in range(6):
        T = L.ora_get_T(C.byref(o.cfg), Yp, e, T)
    return rho * L.ora_mix_R(C.byref(o.cfg), Yp) * T


def pin_c2_derivatives(L):
    """At a face between two copies of one cell the jumps vanish and the
    averaged derivatives are the cell's own: compare with central finite
    differences of the equation of state."""
    bad = []
    s, o = _state(L=L)
    d = s.desc
    NS, E = d.NS, d.E
    U = o.field("U")
    for (i, j) in ((10, 12), (20, 12), (30, 6)):
        idc = (d.Bwidth_Y + j) * d.Xmax + d.Bwidth_X + i
        _, t, dpdrhoi, drhoi, _ = _terms(L, o, idc, idc)
        u = U[E * idc:E * idc + E].copy()
        rho, mom, rhoE, rho_n = u[0], u[1:4], u[4], u[5:].copy()
        T0 = o.field("T")[idc]

        def p_of(rho_=rho, rho_n_=rho_n, rhoE_=rhoE):
            return _eos_p(o, L, rho_, rho_n_, rhoE_, mom, T0)

        h = 1e-4
        fd_E = (p_of(rhoE_=rhoE * (1 + h)) - p_of(rhoE_=rhoE * (1 - h))) / (2 * h * rhoE)
        fd_rho = (p_of(rho_=rho * (1 + h)) - p_of(rho_=rho * (1 - h))) / (2 * h * rho)
        if abs(t[2] - fd_E) > 1e-6 * abs(fd_E):
            bad.append("dp/d(rho E) %.10g vs %.10g" % (t[2], fd_E))
        if abs(t[0] - fd_rho) > 1e-5 * (abs(fd_rho) + abs(t[0])):
            bad.append("dp/drho %.10g vs %.10g" % (t[0], fd_rho))
        for n in range(NS - 1):
            dn = h * rho
            a, b = rho_n.copy(), rho_n.copy()
            a[n] += dn
            b[n] -= dn
            fd_n = (p_of(rho_n_=a) - p_of(rho_n_=b)) / (2 * dn)
            scale = abs(fd_n) + abs(t[0])
            if abs(dpdrhoi[n] - fd_n) > 1e-5 * scale:
                bad.append("dp/drho_%d %.10g vs %.10g" % (n, dpdrhoi[n], fd_n))
        if np.any(drhoi != 0.0) or t[8] != 0.0:
            bad.append("non-zero jump at equal states")
    return bad


def pin_c2_jump_identity(L):
    """For faces with real jumps: the correction makes the corrected
    derivatives reproduce the pressure jump, dp = dp/drho' drho +
    sum dp/drho_n' drho_n + (rho_bar dp/dE') de, except that the energy term
    is corrected through rho_bar dp/dE where the jump was measured with the
    averaged dp/de: the residual is exactly (dp/de - rho_bar dp/dE) de +
    temp ((dp/de de)^2 - (rho_bar dp/dE de)^2 + 1e-19), to rounding."""
    bad = []
    s, o = _state(L=L)
    d = s.desc
    checked = 0
    for j in range(d.Y_inner):
        for i in range(d.X_inner - 1):
            il = (d.Bwidth_Y + j) * d.Xmax + d.Bwidth_X + i
            _, t, dpdrhoi, drhoi, dpdrhoi_new = _terms(L, o, il, il + 1)
            dpdrho, dpde, dpdE, dp, drho, de, rb, _, temp1, temp, dpdrho_new, dpdE_new = t[:12]
            if abs(dp) < 1e-3 * o.field("p")[il]:
                continue
            lhs = dpdrho_new * drho + float(np.dot(dpdrhoi_new, drhoi)) + rb * dpdE_new * de
            carry = (dpde - rb * dpdE) * de + temp * ((dpde * de) ** 2 - (rb * dpdE * de) ** 2 + 1e-19)
            resid = dp - lhs - carry
            scale = abs(dp) + abs(dpdrho * drho) + abs(dpde * de) + float(np.abs(dpdrhoi * drhoi).sum())
            checked += 1
            if abs(resid) > 1e-9 * scale:
                bad.append("face (%d,%d): identity residual %.3e of %.3e" % (i, j, resid, scale))
                if len(bad) > 3:
                    return bad
    if checked < 20:
        bad.append("only %d faces with a pressure jump" % checked)
    return bad


def test_c2_derivatives_are_the_equation_of_state():
    assert pin_c2_derivatives(_lib()) == []


def test_c2_jump_correction_reproduces_the_pressure_jump():
    assert pin_c2_jump_identity(_lib()) == []


def test_each_planted_mistake_fails_a_pin():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "mut"], check=True)
    for m in MUTANTS:
        L = _lib(os.path.join(ROOT, "oracle", "_build", "liboracle_mut%d.so" % m))
        fails = chem_pins(L) if m <= 5 else (pin_c2_derivatives(L) + pin_c2_jump_identity(L))
        assert fails, "planted mistake %d passed every pin" % m
