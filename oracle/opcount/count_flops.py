"""Algorithmic FP64 operation counts of the hot path, measured on the oracle.

TEST INFRASTRUCTURE.  Runs the operation-counting build of the CPU oracle
(oracle/_build/liboracle_count.so: the same C sources compiled as C++ with
`double` replaced by a counting scalar, see xf_count.hpp) on seeded workloads
and reports operations per unit of work:

  * per characteristic face flux in x and in y (ReconstructFlux, the flux
    part of GetLU), solved exactly from counts on grids of different aspect
    ratio (per x-face, per y-face, per padded cell, per interior cell);
  * per padded cell of UpdateStates (data dependent: Newton iterations),
    per interior cell of the RK3 update, per padded cell of GetDt;
  * per interior cell-update of one full step.

Classes: add/sub, mul, div, sqrt, exp, log, pow; "flops" = the sum of the
first four (each counts 1; a fused multiply-add counts as one mul and one
add, the two operations of the algorithm it stands for).

    python oracle/opcount/count_flops.py [--case shock-bubble] [--json]
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.dirname(HERE)
ROOT = os.path.dirname(ORACLE)
COUNT_SO = os.path.join(ORACLE, "_build", "liboracle_count.so")
CLASSES = ["add", "mul", "div", "sqrt", "exp", "log", "pow"]


def _load():
    if not os.path.exists(COUNT_SO):
        subprocess.run(["make", "-s", "-C", ORACLE, "count"], check=True)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, ROOT)
    import oracle_lib
    L = oracle_lib.load(COUNT_SO)
    L.ora_count_reset.restype = None
    L.ora_count_read.restype = None
    L.ora_count_read.argtypes = [C.POINTER(C.c_ulonglong)]
    return oracle_lib, L


def _read(L):
    buf = (C.c_ulonglong * 8)()
    L.ora_count_read(buf)
    return np.array(list(buf)[:7], dtype=np.float64)


def counted(L, fn):
    L.ora_count_reset()
    fn()
    return _read(L)


def measure(case="shock-bubble", workload=None, grids=((64, 32), (32, 64), (48, 40), (40, 24)), warm=2, seed=1):
    ol, L = _load()
    import xfluids_amd as xf
    workload = workload or case
    rows, rhs = [], []
    per = {"states": [], "rk3": [], "dt": []}
    E = None
    for nx, ny in grids:
        s = xf.Setup(xf.case_path(case), nx=nx, ny=ny)
        d = s.desc
        E = d.E
        U, T = s.initial(workload, seed=seed)
        o = ol.Oracle(ol.config_from_setup(s), L)
        o.set_state(U, T)
        o.initialize()
        o.steps(warm)
        padded = d.Xmax * d.Ymax * d.Zmax
        inner = d.X_inner * d.Y_inner * d.Z_inner
        fx = (d.X_inner + 1) * d.Y_inner * d.Z_inner if d.DimX else 0
        fy = d.X_inner * (d.Y_inner + 1) * d.Z_inner if d.DimY else 0
        o.boundary(0)
        per["states"].append(counted(L, lambda: o.update_states(0)) / padded)
        lu = counted(L, lambda: o.compute_lu(0))
        rows.append([fx, fy, padded, inner])
        rhs.append(lu)
        per["rk3"].append(counted(L, lambda: o.rk3(1, 1e-9)) / inner)
        per["dt"].append(counted(L, lambda: o.get_dt()) / padded)
    A = np.array(rows, dtype=np.float64)
    B = np.array(rhs)
    sol, res, rank, _ = np.linalg.lstsq(A, B, rcond=None)
    resid = float(np.abs(A @ sol - B).max())
    out = {"case": case, "E": E, "grids": [list(g) for g in grids], "classes": CLASSES,
           "face_x": sol[0].round(6).tolist(), "face_y": sol[1].round(6).tolist(),
           "lu_per_padded_cell": sol[2].round(6).tolist(), "lu_per_inner_cell": sol[3].round(6).tolist(),
           "solve_max_residual": resid,
           "states_per_padded_cell": np.mean(per["states"], axis=0).round(3).tolist(),
           "rk3_per_inner_cell": np.mean(per["rk3"], axis=0).round(6).tolist(),
           "dt_per_padded_cell": np.mean(per["dt"], axis=0).round(6).tolist()}
    for k in ("face_x", "face_y", "lu_per_padded_cell", "lu_per_inner_cell", "states_per_padded_cell",
              "rk3_per_inner_cell", "dt_per_padded_cell"):
        out[k + "_flops"] = float(np.sum(out[k][:4]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="shock-bubble")
    ap.add_argument("--workload", default=None)
    ap.add_argument("--json", action="store_true")
    a = ap.parse_args()
    r = measure(a.case, a.workload)
    if a.json:
        print(json.dumps(r))
        return
    for k, v in r.items():
        print("%-28s %s" % (k, v))


if __name__ == "__main__":
    main()
