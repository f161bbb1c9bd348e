"""Write tests/golden/ignition_h2o2ar.json: 0-D constant-volume H2/O2/Ar
ignition integrated with scipy's Radau (rtol 1e-10) on the independent
numpy model in chem_model.py.

The chemistry step of XFluids freezes the pressure at the value of the
preceding UpdateStates call; the reactor is advanced in chunks of DT with
p = rho R T refreshed at every chunk, as a sequence of chemistry calls does.

    python tests/golden/make_ignition_fixture.py
"""
import json
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import chem_model as cm  # noqa: E402

T0, P0, DT, NCHUNK = 1400.0, 101325.0, 1.0e-5, 40


def initial(m):
    x = m["ratio"] / m["ratio"].sum()
    Y = x * m["W"] / np.dot(x, m["W"])
    R = cm.RU * float(np.sum(Y / m["W"]))
    rho = P0 / (R * T0)
    return Y, rho


def main():
    m = cm.load()
    Y, rho = initial(m)
    T = T0
    out = {"species": m["names"], "T0": T0, "p0": P0, "rho": rho, "Y0": Y.tolist(), "dt_chunk": DT,
           "times": [], "T": [], "Y": []}
    for c in range(NCHUNK):
        p = rho * cm.RU * float(np.sum(Y / m["W"])) * T

        def f(t, z):
            dY, dT = cm.rates(m, z[0], z[1:], rho, p)
            return np.concatenate([[dT], dY])

        sol = solve_ivp(f, (0.0, DT), np.concatenate([[T], Y]), method="Radau", rtol=1e-10, atol=1e-16)
        T, Y = float(sol.y[0, -1]), sol.y[1:, -1].copy()
        out["times"].append((c + 1) * DT)
        out["T"].append(T)
        out["Y"].append(Y.tolist())
    with open(os.path.join(HERE, "ignition_h2o2ar.json"), "w") as fo:
        json.dump(out, fo, indent=1)
    print("T(t_end) = %.2f K" % T)


if __name__ == "__main__":
    main()
