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
  * per interior cell-update of one full step;
  * with --chem: the chemistry (ChemeODEQ2SolverKernel1 per cell), fitted
    exactly as a + b * passes over the cells of a seeded reactive state,
    where passes = the CHEMEQ2 predictor passes of the cell (the count the
    device kernels record): b is one rate evaluation plus one predictor and
    corrector, a the start (normalisation, first evaluation, step size) and
    the two mixture energies around it.

Classes: add/sub, mul, div, sqrt, exp, log, pow; "flops" = the sum of the
first four (each counts 1; a fused multiply-add counts as one mul and one
add, the two operations of the algorithm it stands for).

    python oracle/opcount/count_flops.py [--case shock-bubble] [--json]
    python oracle/opcount/count_flops.py --case mixing-layer --nz 8 --inviscid
    python oracle/opcount/count_flops.py --case mixing-layer --chem
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


GRIDS_2D = ((64, 32), (32, 64), (48, 40), (40, 24))
GRIDS_3D = ((12, 8, 8), (8, 12, 8), (8, 8, 12), (10, 10, 8), (12, 10, 10), (8, 10, 12))


def measure(case="shock-bubble", workload=None, grids=None, warm=2, seed=1, nz=0, inviscid=False):
    """Face-flux, UpdateStates, RK3 and GetDt counts.  nz > 0 counts a 3-D
    case on grids of nx x ny x nz cells (per x-, y- and z-face); inviscid
    switches the viscous terms off, so that the faces count the
    characteristic flux alone (the viscous wall fluxes are a kernel of their
    own on the device)."""
    ol, L = _load()
    import xfluids_amd as xf
    workload = workload or case
    three = nz > 0
    grids = grids or (GRIDS_3D if three else GRIDS_2D)
    rows, rhs = [], []
    per = {"states": [], "rk3": [], "dt": []}
    E = None
    for g in grids:
        kw = dict(nx=g[0], ny=g[1])
        if three:
            kw["nz"] = g[2]
        if inviscid:
            kw["viscous"] = (0, 0, 0)
        s = xf.Setup(xf.case_path(case), reaction=False, **kw)
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
        fz = d.X_inner * d.Y_inner * (d.Z_inner + 1) if d.DimZ else 0
        o.boundary(0)
        per["states"].append(counted(L, lambda: o.update_states(0)) / padded)
        lu = counted(L, lambda: o.compute_lu(0))
        rows.append([fx, fy] + ([fz] if three else []) + [padded, inner])
        rhs.append(lu)
        per["rk3"].append(counted(L, lambda: o.rk3(1, 1e-9)) / inner)
        per["dt"].append(counted(L, lambda: o.get_dt()) / padded)
    A = np.array(rows, dtype=np.float64)
    B = np.array(rhs)
    sol, res, rank, _ = np.linalg.lstsq(A, B, rcond=None)
    resid = float(np.abs(A @ sol - B).max())
    names = ["face_x", "face_y"] + (["face_z"] if three else []) + ["lu_per_padded_cell", "lu_per_inner_cell"]
    out = {"case": case, "E": E, "grids": [list(g) for g in grids], "classes": CLASSES, "inviscid": inviscid,
           "solve_max_residual": resid}
    for q, nm in enumerate(names):
        out[nm] = sol[q].round(6).tolist()
    out["states_per_padded_cell"] = np.mean(per["states"], axis=0).round(3).tolist()
    out["rk3_per_inner_cell"] = np.mean(per["rk3"], axis=0).round(6).tolist()
    out["dt_per_padded_cell"] = np.mean(per["dt"], axis=0).round(6).tolist()
    for k in names + ["states_per_padded_cell", "rk3_per_inner_cell", "dt_per_padded_cell"]:
        out[k + "_flops"] = float(np.sum(out[k][:4]))
    return out


def measure_chem(case="mixing-layer", workload=None, n=6, seed=11):
    """Chemistry counts per cell as a + b * passes (see the module comment),
    on the seeded initial field of a reactive case (n^3 cells for a 3-D case,
    n^2 otherwise): the CHEMEQ2 call of every interior cell over half the
    first step's dt (the Strang half step), with the primitives of
    UpdateStates, plus the two mixture energies of the kernel."""
    ol, L = _load()
    import xfluids_amd as xf
    workload = workload or case
    s0 = xf.Setup(xf.case_path(case))
    kw = dict(nx=n, ny=n) if s0.desc.DimZ == 0 else dict(nx=n, ny=n, nz=n)
    s = xf.Setup(xf.case_path(case), **kw)
    d = s.desc
    U, T = s.initial(workload, seed=seed)
    cfg = ol.config_from_setup(s)
    o = ol.Oracle(cfg, L)
    o.set_state(U, T)
    assert o.initialize() == 0
    dt = 0.5 * o.get_dt()
    NS = d.NS
    rho, p, Tf, y = o.field("rho"), o.field("p"), o.field("T"), o.field("y")
    P = C.POINTER(C.c_double)
    rows, cnt = [], []
    for k in range(d.Bwidth_Z, d.Bwidth_Z + d.Z_inner):
        for j in range(d.Bwidth_Y, d.Bwidth_Y + d.Y_inner):
            for i in range(d.Bwidth_X, d.Bwidth_X + d.X_inner):
                idc = (k * d.Ymax + j) * d.Xmax + i
                yy = np.zeros(NS + 1)
                yy[0] = Tf[idc]
                yy[1:] = y[NS * idc:NS * idc + NS]
                yi = np.ascontiguousarray(yy[1:])
                L.ora_count_reset()
                L.ora_mix_e(C.byref(cfg), yi.ctypes.data_as(P), yy[0])
                passes = L.ora_chemeq2(C.byref(cfg), yy.ctypes.data_as(P), dt, rho[idc], p[idc])
                yi = np.ascontiguousarray(yy[1:])
                L.ora_mix_e(C.byref(cfg), yi.ctypes.data_as(P), yy[0])
                cnt.append(_read(L))
                rows.append([1.0, float(passes)])
    A = np.array(rows)
    B = np.array(cnt)
    sol, _, _, _ = np.linalg.lstsq(A, B, rcond=None)
    fit = A @ sol
    flops = B[:, :4].sum(axis=1)
    out = {"case": case, "cells": len(rows), "NS": NS, "NR": d.NR, "classes": CLASSES,
           "passes_min": int(A[:, 1].min()), "passes_max": int(A[:, 1].max()),
           "per_cell": sol[0].round(3).tolist(), "per_pass": sol[1].round(3).tolist(),
           "fit_max_rel_residual_flops": float(np.abs(fit[:, :4].sum(axis=1) - flops).max() / flops.max())}
    out["per_cell_flops"] = float(np.sum(out["per_cell"][:4]))
    out["per_pass_flops"] = float(np.sum(out["per_pass"][:4]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="shock-bubble")
    ap.add_argument("--workload", default=None)
    ap.add_argument("--json", action="store_true")
    ap.add_argument("--nz", type=int, default=0, help="count a 3-D case (grids with nz cells in z)")
    ap.add_argument("--inviscid", action="store_true", help="viscous terms off (face flux alone)")
    ap.add_argument("--chem", action="store_true", help="chemistry per cell and per CHEMEQ2 pass")
    a = ap.parse_args()
    r = measure_chem(a.case, a.workload) if a.chem else measure(a.case, a.workload, nz=a.nz, inviscid=a.inviscid)
    if a.json:
        print(json.dumps(r))
        return
    for k, v in r.items():
        print("%-28s %s" % (k, v))


if __name__ == "__main__":
    main()
