"""Throughput of the XFluids per-step hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form both|plain|mixed] [--no-cpu]

Workload: 2D inert shock-bubble interaction, 3 species (H2, O2, N2) with
NASA-7 (JANAF) thermodynamics, 1024 x 512 cells, WENO5-Z characteristic
fluxes over Ns+4 = 7 variables, SSP-RK3, CFL 0.4 (cases/shock-bubble.json).
One step = dt reduction + 3 RK stages (ghost fill, conservative->primitive
with Newton temperature, x and y characteristic sweeps, RK update).
Inputs are resident in HBM before the timed region.  Prints one JSON line.

Two forms run by default: the plain FP64 form (the headline `value`,
bit-identical to the CPU restatement) and the mixed-precision form (FP32
smoothness indicators, weights and characteristic projections), each with its
error against the FP64 oracle on a reduced grid of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import xfluids_amd as xf  # noqa: E402

PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 (vector and matrix)
CASE = "shock-bubble"
NX, NY = 1024, 512
# Compulsory HBM bytes per cell-update of one full step for Ns = 3, E = 7:
# 8 * (8E + 6) (stage traffic of U, U1 and the T guess).
ALG_BYTES_PER_CELL = 8 * (8 * 7 + 6)
# Algorithmic FP64 operations (add/sub, mul, div, sqrt) of this workload,
# counted on the operation-counting build of the CPU restatement
# (`python oracle/opcount/count_flops.py --case shock-bubble`): per
# characteristic face flux (x and y alike).  The count includes work the
# reference performs whose result is unused (terms multiplied by zero,
# projections of the stencil points i-3 and i+4); the HIP sweep skips it.
# A fused multiply-add counts as its two operations; a subexpression the
# reference writes twice (the WENO5-Z indicator differences) counts once.
FACE_FLOPS = 3869
# The same counting run, per unit of the other stages of one step: UpdateStates
# per padded cell (data dependent: Newton iterations on the seeded field), the
# flux divergence per padded and per interior cell, the RK3 update per
# interior cell, GetDt per padded cell.
STATES_FLOPS = 246.485
LU_PADDED_FLOPS = 4.0
LU_INNER_FLOPS = 42.0
RK3_INNER_FLOPS = 14.0
DT_PADDED_FLOPS = 8.0026
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILE = os.path.join(PROFILES, "pmc-traffic-shock-bubble-1024x512.md")


def measured_traffic():
    """HBM bytes per launch per kernel class (and per tile shape, key
    "sweep_y/1") from the committed rocprofv3 counter summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel factor), or {} when none is
    committed."""
    out = {}
    if os.path.exists(TRAFFIC_FILE):
        for ln in open(TRAFFIC_FILE):
            ln = ln.strip()
            if ln.startswith("traffic[") and "=" in ln:
                k, v = ln.split("=", 1)
                out[k.strip()[8:-1]] = float(v.split()[0])
    return out


def run(args, form):
    s = xf.Setup(xf.case_path(CASE), nx=NX, ny=NY)
    d = s.desc
    precision = xf.XF_MIXED if form == "mixed" else xf.XF_PLAIN_FP64
    U, T = s.initial(CASE, seed=args.seed)
    ctx = s.context(precision=precision)
    ctx.upload(U, T)
    ctx.initialize()
    # tile shape per sweep direction, chosen by timing on this state
    tuned = ctx.autotune(reps=10) if not args.no_tune else (ctx.tile_shape(), None)
    # warm-up (also instantiates the step graph)
    ctx.steps(args.warmup, graph=True)
    ctx.synchronize()
    # timed region: K steps, one hipGraph replay per step, inputs resident
    t0 = time.perf_counter()
    ctx.steps_async(args.steps, graph=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    # per-kernel durations with HIP events on the context stream: each heavy
    # kernel in its stage-1 configuration, launched back to back
    names = ["states"] + [n for n, on in (("sweep_x", d.DimX), ("sweep_y", d.DimY), ("sweep_z", d.DimZ)) if on]
    kt = {n: (ctx.time_kernel(n, args.kernel_reps), args.kernel_reps) for n in names}
    Uf, Tf = ctx.download(0)
    ctx.close()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    cells = d.X_inner * d.Y_inner * d.Z_inner
    value = cells * args.steps / (t1 - t0) / 1e6
    finite = bool(np.all(np.isfinite(Uf)))
    return d, value, ms_per_step, kt, finite, tuned


def roofline(d, kt, shapes=None):
    # dominant kernel: every class runs three times per step, so the largest
    # average launch duration is also the largest share of the step
    name = max(kt, key=lambda k: kt[k][0])
    avg_ms, cnt = kt[name]
    tr = measured_traffic()
    traffic = tr.get(name)
    if name.startswith("sweep") and shapes is not None:
        traffic = tr.get("%s/%d" % (name, int(shapes["xyz".index(name[-1])])), traffic)
    if name.startswith("sweep"):
        faces = {"sweep_x": (d.X_inner + 1) * d.Y_inner * d.Z_inner,
                 "sweep_y": d.X_inner * (d.Y_inner + 1) * d.Z_inner,
                 "sweep_z": d.X_inner * d.Y_inner * (d.Z_inner + 1)}[name]
        flops = faces * FACE_FLOPS
        achieved = flops / (avg_ms * 1e-3) / 1e12
        return {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                "avg_launch_ms": avg_ms, "timed_launches": cnt, "alg_flops_per_launch": flops,
                "faces_per_launch": faces, "flops_per_face": FACE_FLOPS}
    bytes_ = d.Xmax * d.Ymax * d.Zmax * 8 * (d.E + 10)
    achieved = bytes_ / (avg_ms * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "avg_launch_ms": avg_ms,
            "timed_launches": cnt}


def step_roofline(d, value, fp64_tflops=None):
    """Whole-step fractions: counted FP64 operations and compulsory HBM bytes
    per cell-update of the C2 step (dt + three stages), times the measured
    cell-updates/s, against the datasheet peaks (and the measured fma rate)."""
    cells = d.X_inner * d.Y_inner * d.Z_inner
    padded = d.Xmax * d.Ymax * d.Zmax
    faces = ((d.X_inner + 1) * d.Y_inner * d.Z_inner if d.DimX else 0) + \
            (d.X_inner * (d.Y_inner + 1) * d.Z_inner if d.DimY else 0) + \
            (d.X_inner * d.Y_inner * (d.Z_inner + 1) if d.DimZ else 0)
    stage = faces * FACE_FLOPS + padded * (STATES_FLOPS + LU_PADDED_FLOPS) + cells * (LU_INNER_FLOPS + RK3_INNER_FLOPS)
    flops_cu = (3 * stage + padded * DT_PADDED_FLOPS) / cells
    tflops = flops_cu * value * 1e6 / 1e12
    gbs = ALG_BYTES_PER_CELL * value * 1e6 / 1e9
    out = {"flops_per_cell_update": flops_cu, "bytes_per_cell_update": ALG_BYTES_PER_CELL,
           "achieved_tflops": tflops, "frac_fp64_peak": tflops / PEAK_FP64_TFLOPS,
           "achieved_gbs": gbs, "frac_hbm_peak": gbs / PEAK_HBM_GBS,
           "intensity_flop_per_byte": flops_cu / ALG_BYTES_PER_CELL, "bound": "fp64"}
    if fp64_tflops:
        out["frac_measured_fma_rate"] = tflops / fp64_tflops
    return out


def vortex_leg(steps=200, n=256):
    """The isentropic vortex (single species, WENO5-JS, periodic, CFL 0.5):
    throughput of both forms, and after `steps` steps the L1 error of rho
    against the exact solution (the initial vortex advected by (t, t) modulo
    the 10 x 10 domain) and the drift of the interior sums of rho, rho u,
    rho v and E (conserved to round-off on a periodic grid)."""
    case = "isentropic-vortex"
    s = xf.Setup(xf.case_path(case), nx=n, ny=n)
    d = s.desc
    U, T = s.initial(case)
    E = d.E
    shp = (d.Zmax, d.Ymax, d.Xmax, E)
    sl = (0, slice(d.Bwidth_Y, d.Bwidth_Y + d.Y_inner), slice(d.Bwidth_X, d.Bwidth_X + d.X_inner))
    inner = lambda A: np.asarray(A).reshape(shp)[sl]
    sums0 = inner(U).sum(axis=(0, 1))
    out = {"workload": "2D isentropic vortex, single species, WENO5-JS, SSP-RK3, CFL 0.5, periodic",
           "grid": "%dx%d" % (n, n), "steps": steps, "forms": {}}
    for form in ("plain", "mixed"):
        prec = xf.XF_MIXED if form == "mixed" else xf.XF_PLAIN_FP64
        ctx = s.context(precision=prec)
        ctx.upload(U, T)
        ctx.initialize()
        ctx.steps(5, graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.steps_async(steps, graph=True)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.close()
        # accuracy run from the initial state, recording dt
        ctx = s.context(precision=prec)
        ctx.upload(U, T)
        ctx.initialize()
        dts = ctx.steps(steps, graph=True, record_dt=True)
        Ug, _ = ctx.download(0)
        ctx.close()
        t = float(np.sum(dts))
        dx = 10.0 / n
        x = (np.arange(n) + 0.5) * dx
        X, Y = np.meshgrid(x, x)
        xv = np.mod(X - t, 10.0) - 5.0
        yv = np.mod(Y - t, 10.0) - 5.0
        rho_ex = (1.0 - 25.0 * np.exp(1.0 - (xv ** 2 + yv ** 2)) / (28.0 * np.pi ** 2)) ** 2.5
        Ui = inner(Ug)
        sums1 = Ui.sum(axis=(0, 1))
        scale = np.abs(sums0).max()
        out["forms"][form] = {
            "value": n * n * steps / (t1 - t0) / 1e6, "ms_per_step": (t1 - t0) * 1e3 / steps,
            "time": t, "l1_abs_rho_vs_exact": float(np.abs(Ui[:, :, 0] - rho_ex).mean()),
            "conservation_drift": {k: float(abs(sums1[q] - sums0[q]) / scale)
                                   for k, q in (("rho", 0), ("rho_u", 1), ("rho_v", 2), ("E", 4))}}
    return out


def copy_probe(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (the library's streaming copy
    kernel): HBM copy bandwidth, and the scale for the FETCH_SIZE and
    WRITE_SIZE counters when run under rocprofv3 --pmc."""
    ms = xf.copy_probe(nbytes, reps)
    return {"bytes_read": nbytes, "bytes_written": nbytes, "reps": reps, "avg_ms": ms,
            "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}


def accuracy_check(form, steps=10):
    """A form vs the CPU oracle on a reduced grid of the same workload (the
    oracle runs the full 1024x512 grid too slowly): L1/Linf of rho, p, T and
    mass-fraction errors and ranges, bitwise fraction of U."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, config_from_setup
    from parity import field_errors
    s = xf.Setup(xf.case_path(CASE), nx=128, ny=64)
    d = s.desc
    U, T = s.initial(CASE, seed=3)
    ctx = s.context(precision=xf.XF_MIXED if form == "mixed" else xf.XF_PLAIN_FP64)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(steps, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    o = Oracle(config_from_setup(s))
    o.set_state(U, T)
    o.initialize()
    o.steps(steps)
    Uo, To = o.field("U"), o.field("T")
    out = {"vs": "cpu oracle (FP64)", "grid": "128x64", "steps": steps}
    out.update(field_errors(s, Uo, To, Ug, Tg))
    out["bitwise_fraction_U"] = float(np.mean(Ug.view(np.uint64) == Uo.view(np.uint64)))
    return out


REACTIVE = [
    # (label, case file / workload, overrides, timed steps)
    ("reactive shock tube H2/O2/Ar, Ns=9 NR=19, 8192 cells, Strang", "reactive-shock-tube", dict(nx=8192), 40),
    ("reactive channel H2/O2/Ar with hot kernel, Ns=9 NR=19, 512x512, Strang", "reactive-channel",
     dict(nx=512, ny=512), 10),
]
REACTIVE_FORMS = [("plain", "per-cell"), ("plain", "queue"), ("mixed", "queue")]


def reactive_leg(label, case, over, steps, warmup=2, seed=1):
    """Full reactive steps (dt, two Strang chemistry half-steps, three RK
    stages) for each form; chemistry kernel time from HIP events in a few
    ungraphed steps; per-cell predictor-pass statistics of one call."""
    s = xf.Setup(xf.case_path(case), **over)
    d = s.desc
    U, T = s.initial(case, seed=seed)
    cells = d.X_inner * d.Y_inner * d.Z_inner
    out = {"workload": label, "cells": cells, "steps": steps, "forms": {}}
    for prec, drv in REACTIVE_FORMS:
        ctx = s.context(precision=xf.XF_MIXED if prec == "mixed" else xf.XF_PLAIN_FP64,
                        chem_driver=xf.XF_CHEM_QUEUE if drv == "queue" else xf.XF_CHEM_PER_CELL)
        ctx.upload(U, T)
        ctx.initialize()
        # first chemistry call on the initial field: radicals start at the
        # 1e-20 floor, so every cell takes many passes (the stiff start-up)
        ctx.update_states(0)
        dt0 = ctx.get_dt()
        ctx.synchronize()
        c0 = time.perf_counter()
        sub0 = ctx.chemistry(0.5 * dt0, record=True)
        c1 = time.perf_counter()
        ctx.upload(U, T)
        ctx.initialize()
        shapes, _ = ctx.autotune(reps=5)
        ctx.steps(warmup, graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.steps_async(steps, graph=True)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.set_profiling(True, only=["chem", "sweep_x", "sweep_y", "sweep_z", "states"])
        ctx.steps(2)
        kt = ctx.kernel_times()
        ctx.set_profiling(False)
        ctx.update_states(0)
        sub = ctx.chemistry(0.5 * ctx.get_dt(), record=True)
        ctx.close()
        out["forms"]["%s/%s" % (prec, drv)] = {
            "value": cells * steps / (t1 - t0) / 1e6, "ms_per_step": (t1 - t0) * 1e3 / steps,
            "tile_shapes": [int(x) for x in shapes],
            "kernels_ms": {k: v[0] for k, v in kt.items()},
            "chem_passes": {"min": int(sub.min()), "median": float(np.median(sub)),
                            "p99": float(np.percentile(sub, 99)), "max": int(sub.max()), "mean": float(sub.mean())},
            "first_chem_call": {"ms": (c1 - c0) * 1e3, "passes_min": int(sub0.min()),
                                "passes_median": float(np.median(sub0)), "passes_max": int(sub0.max()),
                                "passes_mean": float(sub0.mean())}}
    base = out["forms"]["plain/per-cell"]["value"]
    for k, v in out["forms"].items():
        v["speedup_vs_plain_per_cell"] = v["value"] / base
    # the fluid part alone: the same case and steps with the Reaction stage
    # switched off (equations.Sources_React = 0), both precision forms
    out["fluid_only"] = {}
    si = xf.Setup(xf.case_path(case), reaction=False, **over)
    for prec in ("plain", "mixed"):
        ctx = si.context(precision=xf.XF_MIXED if prec == "mixed" else xf.XF_PLAIN_FP64)
        ctx.upload(U, T)
        ctx.initialize()
        ctx.autotune(reps=5)
        ctx.steps(warmup, graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.steps_async(steps, graph=True)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.close()
        out["fluid_only"][prec] = {"value": cells * steps / (t1 - t0) / 1e6, "ms_per_step": (t1 - t0) * 1e3 / steps}
    return out


def reactive_accuracy(steps=20):
    """Mixed form with the queue driver vs the FP64 oracle, reactive shock
    tube on 1024 cells (the oracle runs it in about a second)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, config_from_setup
    from parity import field_errors
    s = xf.Setup(xf.case_path("reactive-shock-tube"), nx=1024)
    U, T = s.initial("reactive-shock-tube", seed=1)
    o = Oracle(config_from_setup(s))
    o.set_state(U, T)
    o.initialize()
    o.steps(steps)
    out = {"vs": "cpu oracle (FP64)", "workload": "reactive shock tube, 1024 cells", "steps": steps}
    for prec, drv in (("plain", "queue"), ("mixed", "queue")):
        ctx = s.context(precision=xf.XF_MIXED if prec == "mixed" else xf.XF_PLAIN_FP64,
                        chem_driver=xf.XF_CHEM_QUEUE)
        ctx.upload(U, T)
        ctx.initialize()
        ctx.steps(steps, graph=True)
        Ug, Tg = ctx.download(0)
        ctx.close()
        out["%s/%s" % (prec, drv)] = field_errors(s, o.field("U"), o.field("T"), Ug, Tg)
    return out


def mixing_layer_leg(n, steps, warmup, log=None):
    """3-D reacting mixing layer, H2/N2 (600 K) under air (1200 K), velocity
    ratio 3, Ns=9 NR=19 (H2O-N2 mechanism), n^3 cells, periodic in x and z.
    Mixture-averaged viscous, heat-conduction and diffusion wall fluxes
    (the case file's Visc / Visc_Heat / Visc_Diffu).  Plain FP64 with per-cell
    chemistry vs the mixed form with the work-queue chemistry; sweep tile
    shapes autotuned per direction, steps replayed as hipGraphs."""
    s = xf.Setup(xf.case_path("mixing-layer"), nx=n, ny=n, nz=n)
    d = s.desc
    U, T = s.initial("mixing-layer", seed=11)
    cells = d.X_inner * d.Y_inner * d.Z_inner
    out = {"workload": "3-D reacting mixing layer %d^3, Ns=9 NR=19, viscous (mixture-averaged transport)" % n,
           "cells": cells,
           "steps": steps, "warmup": warmup, "forms": {}}
    for prec, drv in (("plain", "per-cell"), ("mixed", "queue")):
        ctx = s.context(precision=xf.XF_MIXED if prec == "mixed" else xf.XF_PLAIN_FP64,
                        chem_driver=xf.XF_CHEM_QUEUE if drv == "queue" else xf.XF_CHEM_PER_CELL)
        ctx.upload(U, T)
        ctx.initialize()
        shapes, ms = ctx.autotune(reps=3)
        ctx.steps(warmup, graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.steps_async(steps, graph=True)
        ctx.synchronize()
        t1 = time.perf_counter()
        names = ("states", "sweep_x", "sweep_y", "sweep_z") + (
            ("visc_prep", "visc_x", "visc_y", "visc_z") if d.visc else ())
        kt = {k: ctx.time_kernel(k, 5) for k in names}
        ctx.close()
        out["forms"]["%s/%s" % (prec, drv)] = {
            "value": cells * steps / (t1 - t0) / 1e6, "ms_per_step": (t1 - t0) * 1e3 / steps,
            "tile_shapes": [int(x) for x in shapes], "autotune_ms": ms.tolist(), "kernels_ms": kt}
        if log:
            log("mixing layer: %s/%s done" % (prec, drv))
    base = out["forms"]["plain/per-cell"]["value"]
    for v in out["forms"].values():
        v["speedup_vs_plain_per_cell"] = v["value"] / base
    return out


def ensemble_leg(ncases, sides, steps, warmup, seed=7, log=None):
    """Batched ensemble of independent reactive channel cases (H2/O2/Ar,
    Ns=9, NR=19), grid side per case drawn by seed from `sides`, initial
    temperature and equivalence ratio perturbed per case by its own seed.
    Forms: every case stepped alone, one after another (plain FP64, per-cell
    chemistry: what running the cases one by one costs), and the ensemble
    (one hipGraph per step over all cases, one work-queue chemistry launch for
    all cells) in the plain and the mixed form.  The mixed ensemble's error is
    measured against the plain single-case runs of the smallest cases."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity import field_errors
    rng = np.random.default_rng(seed)
    picks = [int(v) for v in rng.choice(sides, size=ncases)]
    cases = []
    for k, n in enumerate(picks):
        s = xf.Setup(xf.case_path("reactive-channel"), nx=n, ny=n)
        U, T = s.initial("reactive-channel", seed=1000 + k)
        cases.append((s, U, T))
    cells = sum(s.desc.ninner for s, _, _ in cases)
    out = {"workload": "ensemble of %d reactive channel cases, sides from %s by seed %d" % (ncases, sides, seed),
           "cells": cells, "sides": picks, "steps": steps, "warmup": warmup, "forms": {}}

    def say(msg):
        if log:
            log(msg)

    # one after another
    t_alone = 0.0
    alone = {}
    small = sorted(range(ncases), key=lambda k: picks[k])[:4]
    for k, (s, U, T) in enumerate(cases):
        ctx = s.context()
        ctx.upload(U, T)
        ctx.initialize()
        ctx.steps(warmup, graph=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.steps_async(steps, graph=True)
        ctx.synchronize()
        t_alone += time.perf_counter() - t0
        if k in small:
            alone[k] = ctx.download(0)
        ctx.close()
    say("ensemble leg: cases one by one done")
    out["forms"]["plain/one-by-one"] = {"value": cells * steps / t_alone / 1e6, "ms_per_step": t_alone * 1e3 / steps}
    for prec in ("plain", "mixed"):
        ctxs = []
        for s, U, T in cases:
            c = s.context(precision=xf.XF_MIXED if prec == "mixed" else xf.XF_PLAIN_FP64,
                          chem_driver=xf.XF_CHEM_QUEUE)
            c.upload(U, T)
            c.initialize()
            ctxs.append(c)
        ens = xf.Ensemble(ctxs)
        ens.steps(warmup, graph=True)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.steps(steps, graph=True)
        ens.synchronize()
        t1 = time.perf_counter()
        res = {"value": cells * steps / (t1 - t0) / 1e6, "ms_per_step": (t1 - t0) * 1e3 / steps}
        errs = []
        for k in small:
            Ug, Tg = ctxs[k].download(0)
            s, _, _ = cases[k]
            e = field_errors(s, alone[k][0], alone[k][1], Ug, Tg)
            e["side"] = picks[k]
            errs.append(e)
        res["error_vs_single_case_fp64"] = errs
        ens.close()
        for c in ctxs:
            c.close()
        out["forms"]["%s/ensemble" % prec] = res
        say("ensemble leg: %s ensemble done" % prec)
    base = out["forms"]["plain/one-by-one"]["value"]
    for v in out["forms"].values():
        v["speedup_vs_one_by_one"] = v["value"] / base
    return out


def cpu_baseline(sample_steps):
    """Oracle (plain C, FP64, OpenMP) on the full 1024x512 grid for a few steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, config_from_setup
    s = xf.Setup(xf.case_path(CASE), nx=NX, ny=NY)
    d = s.desc
    U, T = s.initial(CASE, seed=1)
    o = Oracle(config_from_setup(s))
    o.set_state(U, T)
    o.initialize()
    o.steps(1)
    t0 = time.perf_counter()
    o.steps(sample_steps)
    dt = time.perf_counter() - t0
    cells = d.X_inner * d.Y_inner
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"value": cells * sample_steps / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads,
           "kind": "port", "sample": "%d steps of the %dx%d workload after 1 warm-up step (%.1f s)" %
           (sample_steps, NX, NY, dt)}
    # the same oracle on one thread (its OpenMP runtime is this process's)
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        t0 = time.perf_counter()
        o.steps(2)
        dt1 = time.perf_counter() - t0
        gomp.omp_set_num_threads(threads)
        out["single_thread"] = {"value": cells * 2 / dt1 / 1e6, "cores": 1,
                                "sample": "2 further steps (%.1f s)" % dt1}
    except OSError:
        pass
    return out


def progress(msg):
    """Progress to stderr (the JSON line on stdout stays the only output)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--kernel-reps", type=int, default=60)
    ap.add_argument("--form", choices=["both", "plain", "mixed"], default="both",
                    help="forms to run; the first is the headline value (plain when both)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the reduced-grid check against the oracle")
    ap.add_argument("--no-reactive", action="store_true", help="skip the reactive workloads")
    ap.add_argument("--no-tune", action="store_true", help="keep the default sweep tile shapes")
    ap.add_argument("--mixing-layer", type=int, default=256, metavar="N", help="3-D mixing layer N^3 (0: skip)")
    ap.add_argument("--ensemble-cases", type=int, default=64, help="cases in the ensemble leg (0: skip)")
    ap.add_argument("--ensemble-sides", default="128,256,512,1024", help="grid sides the ensemble draws from")
    ap.add_argument("--ensemble-steps", type=int, default=2)
    ap.add_argument("--reactive-only", type=int, default=-1, metavar="I",
                    help="profiling aid: run only reactive workload I and print its JSON")
    ap.add_argument("--copy-probe", action="store_true",
                    help="accepted for older command lines: the 1 GiB copy and the fma probe always run")
    ap.add_argument("--no-vortex", action="store_true", help="skip the isentropic-vortex leg")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.reactive_only >= 0:
        print(json.dumps(reactive_leg(*REACTIVE[args.reactive_only][:3], steps=args.steps, warmup=args.warmup)))
        return
    forms = ["plain", "mixed"] if args.form == "both" else [args.form]
    res = {}
    for f in forms:
        res[f] = run(args, f)
        progress("%s form timed" % f)
    head = forms[0]
    d, value, ms, kt, finite, head_tuned = res[head]
    line = {
        "metric": "Mcell-updates/s per RK3 step, plain FP64 vs mixed/batched, at stated L1 err",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2D inert shock-bubble, Ns=3 (H2,O2,N2), JANAF, WENO5-Z, SSP-RK3, CFL 0.4",
                   "grid": "%dx%d" % (NX, NY), "form": head, "finite": finite},
        "roofline": roofline(d, kt, head_tuned[0]),
        "kernels_ms": {k: v[0] for k, v in kt.items()},
    }
    line["forms"] = {}
    for f in forms:
        _, v, m, k, fin, tuned = res[f]
        line["forms"][f] = {"value": v, "ms_per_step": m, "kernels_ms": {n: t[0] for n, t in k.items()},
                            "finite": fin, "tile_shapes": [int(x) for x in tuned[0]]}
        if tuned[1] is not None:
            line["forms"][f]["autotune_ms"] = {"sweep_x": tuned[1][0, :2].tolist(), "sweep_y": tuned[1][1].tolist()}
        if not args.no_accuracy:
            line["forms"][f]["accuracy"] = accuracy_check(f)
    if not args.no_reactive:
        line["reactive"] = [reactive_leg(*r) for r in REACTIVE]
        progress("reactive legs done")
        if not args.no_accuracy:
            line["reactive_accuracy"] = reactive_accuracy()
    if args.mixing_layer > 0:
        line["mixing_layer"] = mixing_layer_leg(args.mixing_layer, 2, 1, log=progress)
    if args.ensemble_cases > 0:
        sides = [int(v) for v in args.ensemble_sides.split(",")]
        line["ensemble"] = ensemble_leg(args.ensemble_cases, sides, args.ensemble_steps, 1, log=progress)
    # measured ceilings next to the datasheet peaks: streaming copy (HBM) and
    # independent FP64 fma chains (VALU)
    line["copy_probe"] = copy_probe()
    fp64 = xf.fma_probe(10)
    line["fp64_probe"] = {"tflops": fp64, "peak_datasheet": PEAK_FP64_TFLOPS}
    line["roofline"]["frac_measured_fma_rate"] = line["roofline"]["achieved"] / fp64
    line["step_roofline"] = step_roofline(d, value, fp64)
    if not args.no_vortex:
        line["vortex"] = vortex_leg()
        progress("vortex leg done")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
