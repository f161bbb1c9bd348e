"""Pins of the oracle's species and mixture thermodynamics.

Independent numpy evaluation of the published NASA-9 and NASA-7 (JANAF)
polynomial forms, with the coefficients parsed here from the runtime.dat
files (not through the library), against the oracle's Cp, H, S, Gibbs,
internal energy; Cp = dH/dT; the T < 200 K linear extrapolation; the Newton
temperature inversion T(e(T)) = T.
"""
import ctypes as C
import os

import numpy as np
import pytest

import xfluids_amd as xf
from oracle_lib import config_from_setup, lib

RU = 6.02214076e26 * 1.380649e-23 * 1.0e-3
DATA = xf.RUNTIME_DAT
SPECIES = ["H2", "O2", "O", "H", "OH", "HO2", "H2O2", "H2O", "AR"]


def parse(fname, nvals):
    toks = open(os.path.join(DATA, fname)).read().split()
    out = {}
    for i, t in enumerate(toks):
        if t.startswith("*") and t != "*END":
            out[t[1:]] = [float(v) for v in toks[i + 1:i + 1 + nvals]]
    return out


NASA9 = parse("thermal_dynamics.dat", 28)
JANAF = parse("thermal_dynamics_janaf.dat", 15)


def nasa9_fit(sp, T):
    c = NASA9[sp]
    r = 0 if T < 1000 else (1 if T < 6000 else 2)
    a = c[9 * r:9 * r + 7]
    b = c[9 * r + 7:9 * r + 9]
    cp = a[0] * T ** -2 + a[1] / T + a[2] + a[3] * T + a[4] * T ** 2 + a[5] * T ** 3 + a[6] * T ** 4
    h = (-a[0] / T + a[1] * np.log(T) + a[2] * T + a[3] * T ** 2 / 2 + a[4] * T ** 3 / 3 + a[5] * T ** 4 / 4 +
         a[6] * T ** 5 / 5 + b[0])
    s = (-a[0] * T ** -2 / 2 - a[1] / T + a[2] * np.log(T) + a[3] * T + a[4] * T ** 2 / 2 + a[5] * T ** 3 / 3 +
         a[6] * T ** 4 / 4 + b[1])
    return cp, h, s, c[27] * 1e-3


def janaf_fit(sp, T):
    c = JANAF[sp]
    a = c[0:7] if T > 1000 else c[7:14]
    cp = a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4
    h = (a[0] + a[1] / 2 * T + a[2] / 3 * T ** 2 + a[3] / 4 * T ** 3 + a[4] / 5 * T ** 4 + a[5] / T) * T
    s = a[0] * np.log(T) + a[1] * T + a[2] / 2 * T ** 2 + a[3] / 3 * T ** 3 + a[4] / 4 * T ** 4 + a[6]
    return cp, h, s, c[14] * 1e-3


@pytest.fixture(scope="module", params=[1, 0], ids=["nasa9", "janaf"])
def cfg(request):
    s = xf.Setup(xf.case_path("reactive-shock-tube"), thermo=request.param)
    assert s.species == SPECIES
    return config_from_setup(s), request.param


TEMPS = [210.0, 300.0, 650.0, 999.0, 1000.0, 1500.0, 2500.0, 3999.0, 5999.0, 6000.0, 8000.0]


def test_species_polynomials(cfg):
    c, thermo = cfg
    L = lib()
    for n, sp in enumerate(SPECIES):
        for T in TEMPS:
            if not thermo and T > 5000:
                continue
            cp, h, s, W = (nasa9_fit if thermo else janaf_fit)(sp, T)
            R = RU / W
            assert abs(L.ora_cp_i(C.byref(c), T, n) - R * cp) <= 1e-12 * abs(R * cp)
            assert abs(L.ora_h_i(C.byref(c), T, n) - R * h) <= 1e-11 * max(abs(R * h), R * T)
            assert abs(L.ora_s_i(C.byref(c), T, n) - R * s) <= 1e-12 * max(abs(R * s), R)
            g = L.ora_g_i(C.byref(c), T, n)  # Gibbs / R_i (Ri = 1 form)
            assert abs(g - (s - h / T)) <= 1e-11 * max(abs(s), abs(h / T))
            e = L.ora_e_i(C.byref(c), T, n)
            assert abs(e - R * (h - T)) <= 1e-11 * max(abs(R * h), R * T)


def test_cp_is_dh_dt(cfg):
    c, thermo = cfg
    L = lib()
    for n in range(len(SPECIES)):
        for T in (400.0, 900.0, 1300.0, 2800.0):
            dT = 1e-3 * T
            dh = (L.ora_h_i(C.byref(c), T + dT, n) - L.ora_h_i(C.byref(c), T - dT, n)) / (2 * dT)
            cp = L.ora_cp_i(C.byref(c), T, n)
            assert abs(dh - cp) <= 1e-6 * cp


def test_low_temperature_extrapolation(cfg):
    c, _ = cfg
    L = lib()
    for n in range(len(SPECIES)):
        h200 = L.ora_h_i(C.byref(c), 200.0, n)
        cp200 = L.ora_cp_i(C.byref(c), 200.0, n)
        for T in (50.0, 120.0, 199.0):
            assert abs(L.ora_h_i(C.byref(c), T, n) - (h200 + cp200 * (T - 200.0))) <= 1e-9 * abs(h200) + 1e-9
            assert L.ora_cp_i(C.byref(c), T, n) == cp200


def test_newton_temperature_roundtrip(cfg):
    c, _ = cfg
    L = lib()
    rng = np.random.default_rng(2)
    DP = C.POINTER(C.c_double)
    for _ in range(200):
        Y = rng.dirichlet(np.ones(len(SPECIES)))
        T = rng.uniform(250.0, 4000.0)
        y = np.ascontiguousarray(Y)
        yp = y.ctypes.data_as(DP)
        e = L.ora_mix_h(C.byref(c), yp, T) - L.ora_mix_R(C.byref(c), yp) * T
        assert abs(e - L.ora_mix_e(C.byref(c), yp, T)) <= 1e-9 * max(1.0, abs(e))
        T0 = T * rng.uniform(0.97, 1.03)
        Tr = L.ora_get_T(C.byref(c), yp, e, T0)
        assert abs(Tr - T) <= 1e-5, (T, T0, Tr)
        cp = L.ora_mix_cp(C.byref(c), yp, T)
        cv = L.ora_mix_cv(C.byref(c), yp, T)
        R = L.ora_mix_R(C.byref(c), yp)
        assert abs(cp - cv - R) <= 1e-12 * cp and cp / cv > 1.0


def test_newton_step_limits_and_iteration_cap(cfg):
    """get_T's damped Newton (Mixing_device.h:171-193): each step is clamped to
    T - 1e-3 T .. T + 1e-2 T and the loop stops after 100 iterations.  From a
    guess far below (above) the root every step hits the upper (lower) clamp,
    so the result is the guess grown by 1 % (shrunk by 0.1 %) 100 times,
    through the same floating-point recurrence T <- T - df."""
    c, _ = cfg
    L = lib()
    DP = C.POINTER(C.c_double)
    y = np.ascontiguousarray(np.full(len(SPECIES), 1.0 / len(SPECIES)))
    yp = y.ctypes.data_as(DP)
    for T_true, T0, frac in ((3500.0, 300.0, -1e-2), (300.0, 3000.0, 1e-3)):
        e = L.ora_mix_e(C.byref(c), yp, T_true)
        want = T0
        for _ in range(100):
            want = want - frac * want
        assert L.ora_get_T(C.byref(c), yp, e, T0) == want
        # one more iteration's worth of room is not taken: the cap is 100
        assert abs(want - T_true) > 50.0
