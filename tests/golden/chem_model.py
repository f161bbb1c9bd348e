"""Independent numpy statement of the constant-volume H2/O2 kinetics model.

Written from the model equations of the mechanism files (law of mass action,
modified Arrhenius forward rates A T^B exp(-E/(R_c T)) with E in cal/mol,
reverse rates from equilibrium constants of the NASA-9 Gibbs energies,
third-body efficiencies, constant-volume energy balance), with the unit
conventions the XFluids model uses: concentrations C_k = rho Y_k / W_k * 1e-3,
Kc^-1 = exp(sum nu (ln(p/p_atm) - g_k)) (p/(R0 T))^-(sum nu), capped at 1e40,
and the pressure frozen at its initial value during a chemistry call.
Test infrastructure: the reference to check CHEMEQ2 against (with scipy).
"""
import os

import numpy as np

RU = 6.02214076e26 * 1.380649e-23 * 1.0e-3
R0 = 6.02214076e26 * 1.380649e-23
P_ATM = 101325.0
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
DATA = os.path.join(ROOT, "xfluids_amd", "data", "runtime.dat")


def _tokens(path):
    return open(path).read().split()


def load(mixture=("Reaction", "H2O_19_reaction")):
    d = os.path.join(DATA, *mixture)
    lines = open(os.path.join(d, "species_list.dat")).read().splitlines()
    names = lines[0].split()
    ratio = [float(v) for v in lines[1].split()]
    ns = len(names)
    tok = _tokens(os.path.join(DATA, "thermal_dynamics.dat"))
    coef = {}
    for i, t in enumerate(tok):
        if t.startswith("*") and t != "*END":
            coef[t[1:]] = np.array([float(v) for v in tok[i + 1:i + 29]])
    nasa = np.array([coef[n] for n in names])
    W = nasa[:, 27] * 1e-3
    sec, cur = {}, None
    for t in _tokens(os.path.join(d, "reaction_list.dat")):
        if t.startswith("*"):
            cur = t
            sec[cur] = []
        elif cur:
            sec[cur].append(float(t))
    nf = np.array(sec["*forward_reaction"]).reshape(-1, ns)
    nb = np.array(sec["*backward_reaction"]).reshape(-1, ns)
    tb = np.array(sec["*third_body"]).reshape(-1, ns)
    arr = np.array(sec["*Arrhenius"]).reshape(-1, 3)
    return dict(names=names, ratio=np.array(ratio), nasa=nasa, W=W, nf=nf, nb=nb, tb=tb, arr=arr)


def _fit(m, T):
    """Cp/R, H/(R T), S/R per species at T (three-range NASA-9, T >= 200)."""
    T = max(T, 200.0)
    r = 0 if T < 1000 else (1 if T < 6000 else 2)
    a = m["nasa"][:, 9 * r:9 * r + 7]
    b = m["nasa"][:, 9 * r + 7:9 * r + 9]
    cp = a[:, 0] / T ** 2 + a[:, 1] / T + a[:, 2] + a[:, 3] * T + a[:, 4] * T ** 2 + a[:, 5] * T ** 3 + a[:, 6] * T ** 4
    h = (-a[:, 0] / T + a[:, 1] * np.log(T) + a[:, 2] * T + a[:, 3] * T ** 2 / 2 + a[:, 4] * T ** 3 / 3 +
         a[:, 5] * T ** 4 / 4 + a[:, 6] * T ** 5 / 5 + b[:, 0]) / T
    s = (-a[:, 0] / T ** 2 / 2 - a[:, 1] / T + a[:, 2] * np.log(T) + a[:, 3] * T + a[:, 4] * T ** 2 / 2 +
         a[:, 5] * T ** 3 / 3 + a[:, 6] * T ** 4 / 4 + b[:, 1])
    return cp, h, s


def rates(m, T, Y, rho, p):
    """Net mass-fraction production rates dY/dt and dT/dt (constant volume)."""
    cp, h, s = _fit(m, T)
    g = s - h                                   # G/(R T) with a minus sign: g = S/R - H/(RT)
    C = rho * Y / m["W"] * 1e-3
    kf = m["arr"][:, 0] * np.exp(m["arr"][:, 1] * np.log(T) - m["arr"][:, 2] * 4.184 / RU / T)
    nu = m["nb"] - m["nf"]
    dn = nu.sum(axis=1)
    kcinv = np.minimum(np.exp(nu @ (np.log(p / P_ATM) - g) - dn * np.log(p / (R0 * T))), 1e40)
    third = np.where(m["tb"].sum(axis=1) > 0, m["tb"] @ C, 1.0)
    rf = kf * third * np.prod(C[None, :] ** m["nf"], axis=1)
    rb = kf * third * kcinv * np.prod(C[None, :] ** m["nb"], axis=1)
    w = nu.T @ (rf - rb)                        # mol-based production per species
    dY = w * m["W"] / rho
    R = RU / m["W"]
    e = R * (h * T - T)                          # species internal energy, J/kg
    cv = float(np.dot(Y, R * cp - R))
    dT = -float(np.dot(dY, e)) / cv
    return dY, dT
