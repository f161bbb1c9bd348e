"""Throughput of the XFluids per-step hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]

Headline workload: the 3-D reacting mixing layer, 256^3 interior cells
(264^3 padded), Ns = 9, NR = 19 (H2O-N2 mechanism), NASA-9, WENO5-JS
characteristic fluxes over E = 13 variables, mixture-averaged viscous,
conduction and diffusion wall fluxes, SSP-RK3 with Strang chemistry (CHEMEQ2),
CFL 0.4 (cases/mixing-layer.json).  It is the largest workload that fits one
GPU.  One step = dt + Reaction + three RK stages + Reaction.

The first steps from the seeded field are the stiff chemical start-up (radicals
at the 1e-20 floor: up to 2000 CHEMEQ2 passes per cell).  The bench runs
STARTUP steps in the plain form, keeps that state as the checkpoint every form
starts from, then per form: W untimed warm-up steps and exactly K timed steps,
one hipGraph replay per step, inputs resident in HBM.  `value` is the plain
FP64 form with per-cell chemistry (bit-identical to the CPU restatement on the
parity grids); the other forms are under `forms`, with their error against
the plain form over the same window.

Prints one JSON line: the headline fields first, then the other workloads
(C2 shock-bubble 1024x512, C1 isentropic vortex, C3 reactive shock tube in its
post-reflection window, C5 64-case ensemble), then a summary.
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
PEAK_FP64_TFLOPS = 78.6     # MI355X FP64 vector (and matrix) peak

# ---- counted work (oracle/opcount/count_flops.py on the CPU restatement;
# add/sub, mul, div, sqrt count 1 each, exp/log apart) -------------------------
# Headline (mixing layer, E = 13, WENO5-JS, NASA-9):
#   characteristic face flux, inviscid part (the sweep kernel's work;
#   `--case mixing-layer --nz 8 --inviscid`), per face in x, y and z
FACE_FLOPS_E13 = 10486
#   viscous wall flux per face (GetWallViscousFlux, computed inside the
#   sweeps): the face count of the viscous case less the inviscid one
#   (`--case mixing-layer --nz 8` against `... --inviscid`: 11 053 - 10 486;
#   the viscous solve leaves a residual of 7 operations in ~10^7 per grid,
#   from the centre-derivative range, which no column of the solve fits)
FACE_VISC_FLOPS_E13 = 567
#   chemistry per interior cell as a + b * passes (`--case mixing-layer
#   --chem`): a = start of CHEMEQ2 plus the two mixture energies, b = one
#   rate evaluation plus one predictor and corrector; exp and log apart
CHEM_CELL_FLOPS = 9113.141
CHEM_PASS_FLOPS = 10900.994
CHEM_PASS_EXP, CHEM_PASS_LOG = 38, 542
# Executed FP64 operations (ADD + MUL + 2 FMA + TRANS from the SQ counters of
# profiles/pmc-instmix-mixing-layer-256cubed.md) over the counted ones: the
# per-cell chemistry at 4.60 passes executes 24 279 per cell against
# 9113 + 4.60 * 10 901 counted (it evaluates each species' Gibbs energy once
# per pass, not once per reaction); the plain z sweep with its wall flux
# executes 5322 per face against 10 486 + 567 (it skips the unused stencil
# ends and the zero terms)
CHEM_EXECUTED_RATIO = 24279.0 / (9113.141 + 4.60 * 10900.994)
SWEEP_E13_EXECUTED_RATIO = 5322.0 / (10486.0 + 567.0)
# Compulsory HBM bytes per cell-update of a full reactive step, E = 13,
# Ns = 9: fluid 8 * (8E + 6) plus two Strang half-steps 2 * 8 * ((E+2) + (Ns+1))
ALG_BYTES_E13 = 8 * (8 * 13 + 6) + 2 * 8 * ((13 + 2) + (9 + 1))
# Chemistry launch: reads rho, Ns-1 partial densities, rho E, T, p; writes
# Ns-1 partial densities, rho E, T (per interior cell)
CHEM_BYTES_PER_CELL = 8 * ((1 + 8 + 1 + 1 + 1) + (8 + 1 + 1))
# C2 (shock-bubble, E = 7, WENO5-Z, JANAF; `--case shock-bubble`): per face,
# per padded / interior cell of the other stages of one step
FACE_FLOPS = 3869
STATES_FLOPS = 246.485
LU_PADDED_FLOPS = 4.0
LU_INNER_FLOPS = 42.0
RK3_INNER_FLOPS = 14.0
DT_PADDED_FLOPS = 8.0026
ALG_BYTES_PER_CELL = 8 * (8 * 7 + 6)
CASE = "shock-bubble"
NX, NY = 1024, 512

HEADLINE_N = 256
STARTUP = 10
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_MIXING_LAYER = os.path.join(PROFILES, "pmc-traffic-mixing-layer-256cubed.md")
TRAFFIC_SHOCK_BUBBLE = os.path.join(PROFILES, "pmc-traffic-shock-bubble-1024x512.md")


def progress(msg):
    """Progress to stderr (the JSON line on stdout stays the only output)."""
    print("[bench] %s" % msg, file=sys.stderr, flush=True)


def measured_traffic(fn):
    """HBM bytes per launch by kernel key from one workload's committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel factor)."""
    out = {}
    if os.path.exists(fn):
        for ln in open(fn):
            ln = ln.strip()
            if ln.startswith("traffic[") and "=" in ln:
                k, v = ln.split("=", 1)
                out[k.strip()[8:-1]] = float(v.split()[0])
    return out


CHEM_DRIVERS = {"per-cell": xf.XF_CHEM_PER_CELL, "queue": xf.XF_CHEM_QUEUE, "wave": xf.XF_CHEM_WAVE}


def ctx_for(s, prec, drv):
    return s.context(precision=xf.XF_MIXED if prec == "mixed" else xf.XF_PLAIN_FP64,
                     chem_driver=CHEM_DRIVERS[drv])


def timed_steps(ctx, warmup, steps):
    ctx.steps(warmup, graph=True)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.steps_async(steps, graph=True)
    ctx.synchronize()
    return time.perf_counter() - t0


def profile_steps(ctx, nsteps=2):
    """Per-class device ms per step and per launch: HIP events on the context
    stream around every launch of nsteps ungraphed steps."""
    ctx.set_profiling(True)
    ctx.steps(nsteps)
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    return {k: {"ms_per_launch": v[0], "launches_per_step": v[1] / nsteps, "ms_per_step": v[0] * v[1] / nsteps}
            for k, v in kt.items()}


def chem_sample(ctx, d, reps=4):
    """reps chemistry launches (the Strang half step of the current dt) on the
    live state, each timed with HIP events on the context stream and with its
    per-cell CHEMEQ2 pass counts: the work of each launch is counted as
    cells * CHEM_CELL_FLOPS + sum(passes) * CHEM_PASS_FLOPS."""
    ms, flops, passes = [], [], []
    for _ in range(reps):
        ctx.update_states(0)
        dt = ctx.get_dt()
        ctx.set_profiling(True, only=["chem"])
        sub = ctx.chemistry(0.5 * dt, record=True)
        kt = ctx.kernel_times()
        ctx.set_profiling(False)
        ms.append(kt["chem"][0])
        passes.append(sub)
        flops.append(d.ninner * CHEM_CELL_FLOPS + float(sub.sum()) * CHEM_PASS_FLOPS)
    allp = np.concatenate(passes)
    return {"launches": reps, "avg_launch_ms": float(np.mean(ms)), "alg_flops_per_launch": float(np.mean(flops)),
            "achieved_tflops": float(np.sum(flops) / (np.sum(ms) * 1e-3) / 1e12),
            "passes": {"mean": float(allp.mean()), "median": float(np.median(allp)),
                       "p99": float(np.percentile(allp, 99)), "max": int(allp.max())},
            "exp_per_launch": float(np.mean([p.sum() for p in passes])) * CHEM_PASS_EXP,
            "log_per_launch": float(np.mean([p.sum() for p in passes])) * CHEM_PASS_LOG}


def rel_errors(s, Ua, Ta, Ub, Tb):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity import field_errors
    e = field_errors(s, Ua, Ta, Ub, Tb)
    return {k: v for k, v in e.items() if not k.endswith("Yk")}


# ---------------------------------------------------------------------------
# headline: the 3-D reacting mixing layer
# ---------------------------------------------------------------------------
HEADLINE_FORMS = [("plain", "per-cell"), ("plain", "queue"), ("mixed", "queue")]


def mixing_layer(n, startup, warmup, steps):
    s = xf.Setup(xf.case_path("mixing-layer"), nx=n, ny=n, nz=n)
    d = s.desc
    U0, T0 = s.initial("mixing-layer", seed=11)
    # the stiff start-up, then the checkpoint every form starts from
    ctx = ctx_for(s, "plain", "per-cell")
    ctx.upload(U0, T0)
    ctx.initialize()
    t0 = time.perf_counter()
    ctx.steps(startup, graph=True)
    t_start = time.perf_counter() - t0
    Uc, Tc = ctx.download(0)
    clock = ctx.time()
    ctx.close()
    progress("mixing layer %d^3: %d start-up steps (%.1f s)" % (n, startup, t_start))
    cells = d.ninner
    out = {"workload": "3-D reacting mixing layer %d^3, Ns=9 NR=19 (H2O-N2), NASA-9, WENO5-JS, viscous "
                       "(mixture-averaged), Strang CHEMEQ2, CFL 0.4" % n,
           "cells": cells, "startup_steps": startup, "startup_s": t_start, "checkpoint_time": clock[0],
           "warmup": warmup, "steps": steps, "forms": {}}
    finals = {}
    for prec, drv in HEADLINE_FORMS:
        key = "%s/%s" % (prec, drv)
        ctx = ctx_for(s, prec, drv)
        ctx.upload(Uc, Tc)
        ctx.initialize()
        ctx.set_time(clock[0], clock[1])
        shapes, _ = ctx.autotune(reps=3)
        dt_s = timed_steps(ctx, warmup, steps)
        finals[key] = ctx.download(0)
        f = {"value": cells * steps / dt_s / 1e6, "ms_per_step": dt_s * 1e3 / steps,
             "tile_shapes": [int(x) for x in shapes], "finite": bool(np.all(np.isfinite(finals[key][0])))}
        f["kernels"] = profile_steps(ctx)
        f["chem_sample"] = chem_sample(ctx, d)
        ctx.close()
        out["forms"][key] = f
        progress("mixing layer: %s %.1f Mcell-updates/s" % (key, f["value"]))
    base = out["forms"]["plain/per-cell"]["value"]
    Ub, Tb = finals["plain/per-cell"]
    for key, f in out["forms"].items():
        f["speedup_vs_plain_per_cell"] = f["value"] / base
        if key != "plain/per-cell":
            Ua, Ta = finals[key]
            f["bitwise_vs_plain_per_cell"] = bool(np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64)))
            f["error_vs_plain_fp64"] = rel_errors(s, Ub, Tb, Ua, Ta)
    return s, out


def headline_roofline(d, form):
    """Roofline of the kernel with the largest share of the plain step."""
    ks = form["kernels"]
    name = max(ks, key=lambda k: ks[k]["ms_per_step"])
    tr = measured_traffic(TRAFFIC_MIXING_LAYER)
    if name == "chem":
        cs = form["chem_sample"]
        achieved = cs["achieved_tflops"]
        r = {"kernel": "chem (k_chem<9>, one lane per cell)", "bound": "fp64-valu", "achieved": achieved,
             "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
             "traffic": tr.get("chem"), "avg_launch_ms": cs["avg_launch_ms"], "timed_launches": cs["launches"],
             "alg_flops_per_launch": cs["alg_flops_per_launch"],
             "flops_per_cell": CHEM_CELL_FLOPS, "flops_per_pass": CHEM_PASS_FLOPS,
             "passes_mean": cs["passes"]["mean"], "exp_log_per_pass": [CHEM_PASS_EXP, CHEM_PASS_LOG],
             "alg_bytes_per_launch": d.ninner * CHEM_BYTES_PER_CELL,
             "executed_over_counted": CHEM_EXECUTED_RATIO,
             "frac_executed": achieved * CHEM_EXECUTED_RATIO / PEAK_FP64_TFLOPS}
        return r
    faces = {"sweep_x": (d.X_inner + 1) * d.Y_inner * d.Z_inner, "sweep_y": d.X_inner * (d.Y_inner + 1) * d.Z_inner,
             "sweep_z": d.X_inner * d.Y_inner * (d.Z_inner + 1)}.get(name)
    avg = ks[name]["ms_per_launch"]
    if faces:
        # the sweep computes the characteristic face flux and, inside the
        # same launch, the face's viscous wall flux
        per_face = FACE_FLOPS_E13 + FACE_VISC_FLOPS_E13
        achieved = faces * per_face / (avg * 1e-3) / 1e12
        return {"kernel": name, "bound": "fp64-valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": tr.get(name),
                "avg_launch_ms": avg, "alg_flops_per_launch": faces * per_face,
                "flops_per_face": [FACE_FLOPS_E13, FACE_VISC_FLOPS_E13],
                "executed_over_counted": SWEEP_E13_EXECUTED_RATIO,
                "frac_executed": achieved * SWEEP_E13_EXECUTED_RATIO / PEAK_FP64_TFLOPS}
    bytes_ = d.Xmax * d.Ymax * d.Zmax * 8 * (d.E + 10)
    achieved = bytes_ / (avg * 1e-3) / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": tr.get(name), "avg_launch_ms": avg}


def mixing_layer_cpu(n, startup, max_steps=3, budget_s=12.0):
    """The oracle (plain C, FP64, OpenMP on the host cores) on the same
    workload at n^3, from the GPU plain form's state after the same start-up
    steps (the oracle would take minutes for the stiff start-up itself):
    steps until budget_s has passed or max_steps are done."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, config_from_setup
    s = xf.Setup(xf.case_path("mixing-layer"), nx=n, ny=n, nz=n)
    d = s.desc
    U0, T0 = s.initial("mixing-layer", seed=11)
    ctx = ctx_for(s, "plain", "per-cell")
    ctx.upload(U0, T0)
    ctx.initialize()
    ctx.steps(startup, graph=True)
    Uc, Tc = ctx.download(0)
    ctx.close()
    o = Oracle(config_from_setup(s))
    o.set_state(Uc, Tc)
    o.initialize()
    done, t0 = 0, time.perf_counter()
    while done < max_steps and (time.perf_counter() - t0) < budget_s:
        st, _ = o.steps(1)
        done += 1
        if st:
            break
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": d.ninner * done / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads, "kind": "port",
            "sample": "%d step(s) of the mixing layer at %d^3 (%d cells) from the state after %d start-up steps, "
                      "%.1f s" % (done, n, d.ninner, startup, dt)}


# ---------------------------------------------------------------------------
# C2: 2-D inert shock-bubble, 1024 x 512
# ---------------------------------------------------------------------------
def shock_bubble(steps, warmup, kernel_reps, accuracy=True):
    s = xf.Setup(xf.case_path(CASE), nx=NX, ny=NY)
    d = s.desc
    U, T = s.initial(CASE, seed=1)
    out = {"workload": "2D inert shock-bubble, Ns=3 (H2,O2,N2), JANAF, WENO5-Z, SSP-RK3, CFL 0.4",
           "grid": "%dx%d" % (NX, NY), "steps": steps, "forms": {}}
    for prec in ("plain", "mixed", "fused"):
        p = {"plain": xf.XF_PLAIN_FP64, "mixed": xf.XF_MIXED, "fused": xf.XF_FUSED_FP64}[prec]
        ctx = s.context(precision=p)
        ctx.upload(U, T)
        ctx.initialize()
        shapes, _ = ctx.autotune(reps=10)
        dt_s = timed_steps(ctx, warmup, steps)
        kt = {n: ctx.time_kernel(n, kernel_reps) for n in ("states", "sweep_x", "sweep_y")}
        ctx.close()
        f = {"value": d.ninner * steps / dt_s / 1e6, "ms_per_step": dt_s * 1e3 / steps,
             "tile_shapes": [int(x) for x in shapes], "kernels_ms": kt}
        if accuracy and prec != "plain":
            f["accuracy_vs_oracle_128x64_10_steps"] = accuracy_check(prec)
        out["forms"][prec] = f
    kt = out["forms"]["plain"]["kernels_ms"]
    name = max(("sweep_x", "sweep_y"), key=lambda k: kt[k])
    faces = (d.X_inner + 1) * d.Y_inner if name == "sweep_x" else d.X_inner * (d.Y_inner + 1)
    ach = faces * FACE_FLOPS / (kt[name] * 1e-3) / 1e12
    out["roofline"] = {"kernel": name, "bound": "fp64-valu", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                       "frac": ach / PEAK_FP64_TFLOPS, "avg_launch_ms": kt[name], "timed_launches": kernel_reps,
                       "alg_flops_per_launch": faces * FACE_FLOPS,
                       "traffic": measured_traffic(TRAFFIC_SHOCK_BUBBLE).get(name)}
    out["step_roofline"] = step_roofline(d, out["forms"]["plain"]["value"])
    return out


def step_roofline(d, value, fp64_tflops=None):
    """Whole-step fractions of the C2 step: counted FP64 operations and
    compulsory HBM bytes per cell-update, times the measured cell-updates/s."""
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


def accuracy_check(form, steps=10):
    """A form vs the CPU oracle on a reduced grid of C2 (the oracle runs the
    full grid too slowly)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle, config_from_setup
    s = xf.Setup(xf.case_path(CASE), nx=128, ny=64)
    U, T = s.initial(CASE, seed=3)
    p = {"plain": xf.XF_PLAIN_FP64, "mixed": xf.XF_MIXED, "fused": xf.XF_FUSED_FP64}[form]
    ctx = s.context(precision=p)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(steps, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    o = Oracle(config_from_setup(s))
    o.set_state(U, T)
    o.initialize()
    o.steps(steps)
    out = rel_errors(s, o.field("U"), o.field("T"), Ug, Tg)
    out["bitwise_fraction_U"] = float(np.mean(Ug.view(np.uint64) == o.field("U").view(np.uint64)))
    return out


# ---------------------------------------------------------------------------
# C1: isentropic vortex
# ---------------------------------------------------------------------------
def vortex_leg(steps=200, n=256):
    """Both forms' throughput, and after `steps` steps the L1 error of rho
    against the exact solution (the initial vortex advected by (t, t) modulo
    the 10 x 10 domain) and the drift of the conserved interior sums."""
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
        dt_s = timed_steps(ctx, 5, steps)
        ctx.close()
        ctx = s.context(precision=prec)
        ctx.upload(U, T)
        ctx.initialize()
        dts = ctx.steps(steps, graph=True, record_dt=True)
        Ug, _ = ctx.download(0)
        ctx.close()
        t = float(np.sum(dts))
        x = (np.arange(n) + 0.5) * (10.0 / n)
        X, Y = np.meshgrid(x, x)
        xv = np.mod(X - t, 10.0) - 5.0
        yv = np.mod(Y - t, 10.0) - 5.0
        rho_ex = (1.0 - 25.0 * np.exp(1.0 - (xv ** 2 + yv ** 2)) / (28.0 * np.pi ** 2)) ** 2.5
        Ui = inner(Ug)
        sums1 = Ui.sum(axis=(0, 1))
        scale = np.abs(sums0).max()
        out["forms"][form] = {
            "value": n * n * steps / dt_s / 1e6, "ms_per_step": dt_s * 1e3 / steps, "time": t,
            "l1_abs_rho_vs_exact": float(np.abs(Ui[:, :, 0] - rho_ex).mean()),
            "conservation_drift": {k: float(abs(sums1[q] - sums0[q]) / scale)
                                   for k, q in (("rho", 0), ("rho_u", 1), ("rho_v", 2), ("E", 4))}}
    return out


# ---------------------------------------------------------------------------
# C3: reactive shock tube in its post-reflection window
# ---------------------------------------------------------------------------
# the batched forms of the tube: the queue, and one wave per cell (each pass
# spread over the lanes), the driver for few cells of uneven stiffness
C3_FORMS = [("plain", "per-cell"), ("plain", "queue"), ("mixed", "queue"), ("plain", "wave"), ("mixed", "wave")]


def shock_tube(nx, t_window, steps, warmup):
    """H2/O2/Ar 2:1:7, Ns = 9, NR = 19, nx cells on 0.12 m, left wall,
    Strang.  The plain form advances the tube on the device clock to
    t_window (the reference's output window starts at 170 us:
    settings/1d-reactive-st.json), where the reflected shock has heated the
    gas near the wall; every form is timed from that checkpoint."""
    case = "reactive-shock-tube"
    s = xf.Setup(xf.case_path(case), nx=nx)
    d = s.desc
    U0, T0 = s.initial(case, seed=1)
    ctx = ctx_for(s, "plain", "queue")
    ctx.upload(U0, T0)
    ctx.initialize()
    ctx.set_target_time(t_window)
    t0 = time.perf_counter()
    while not ctx.time()[2]:
        ctx.steps(2000, graph=True)
    t_adv = time.perf_counter() - t0
    Uc, Tc = ctx.download(0)
    clock = ctx.time()
    ctx.close()
    progress("shock tube: advanced to %.1f us in %d steps (%.1f s)" % (clock[0] * 1e6, clock[1], t_adv))
    out = {"workload": "1-D reactive shock tube H2/O2/Ar, Ns=9 NR=19, %d cells, Strang" % nx,
           "window_start_s": clock[0], "steps_to_window": clock[1], "advance_s": t_adv,
           "T_max_at_window": float(np.max(Tc)), "steps": steps, "forms": {}}
    for prec, drv in C3_FORMS:
        ctx = ctx_for(s, prec, drv)
        ctx.upload(Uc, Tc)
        ctx.initialize()
        ctx.set_time(clock[0], clock[1])
        dt_s = timed_steps(ctx, warmup, steps)
        kt = profile_steps(ctx)
        ctx.update_states(0)
        sub = ctx.chemistry(0.5 * ctx.get_dt(), record=True)
        ctx.close()
        hist = np.bincount(np.minimum(np.ceil(np.log2(np.maximum(sub, 1))).astype(int), 12), minlength=13)
        out["forms"]["%s/%s" % (prec, drv)] = {
            "value": d.ninner * steps / dt_s / 1e6, "ms_per_step": dt_s * 1e3 / steps,
            "chem_ms_per_launch": kt.get("chem", {}).get("ms_per_launch"),
            "passes": {"min": int(sub.min()), "median": float(np.median(sub)), "max": int(sub.max()),
                       "mean": float(sub.mean()), "log2_histogram": hist.tolist()}}
    base = out["forms"]["plain/per-cell"]["value"]
    for v in out["forms"].values():
        v["speedup_vs_plain_per_cell"] = v["value"] / base
    return out


# ---------------------------------------------------------------------------
# C5: 64-case ensemble
# ---------------------------------------------------------------------------
def ensemble_leg(ncases, sides, startup, steps, warmup, seed=7):
    """Independent reactive channel cases (H2/O2/Ar, Ns=9, NR=19), side per
    case drawn by seed, initial temperature and equivalence ratio perturbed
    per case.  After `startup` steps per case (the stiff start-up, plain
    form), forms: every case stepped alone one after another (plain, per
    cell), and the ensemble (one hipGraph per step, one work-queue chemistry
    launch for all cells) in the plain and the mixed form."""
    rng = np.random.default_rng(seed)
    picks = [int(v) for v in rng.choice(sides, size=ncases)]
    cases = []
    for k, n in enumerate(picks):
        s = xf.Setup(xf.case_path("reactive-channel"), nx=n, ny=n)
        U, T = s.initial("reactive-channel", seed=1000 + k)
        c = s.context()
        c.upload(U, T)
        c.initialize()
        c.steps(startup, graph=True)
        cases.append((s,) + c.download(0))
        c.close()
    progress("ensemble: %d cases past %d start-up steps" % (ncases, startup))
    cells = sum(s.desc.ninner for s, _, _ in cases)
    out = {"workload": "ensemble of %d reactive channel cases, sides from %s by seed %d" % (ncases, sides, seed),
           "cells": cells, "sides": picks, "startup_steps": startup, "steps": steps, "forms": {}}
    t_alone = 0.0
    alone = {}
    small = sorted(range(ncases), key=lambda k: picks[k])[:4]
    for k, (s, U, T) in enumerate(cases):
        ctx = s.context()
        ctx.upload(U, T)
        ctx.initialize()
        t_alone += timed_steps(ctx, warmup, steps)
        if k in small:
            alone[k] = ctx.download(0)
        ctx.close()
    out["forms"]["plain/one-by-one"] = {"value": cells * steps / t_alone / 1e6, "ms_per_step": t_alone * 1e3 / steps}
    for prec in ("plain", "mixed"):
        ctxs = []
        for s, U, T in cases:
            c = ctx_for(s, prec, "queue")
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
            s = cases[k][0]
            e = rel_errors(s, alone[k][0], alone[k][1], Ug, Tg)
            e["side"] = picks[k]
            e["bitwise"] = bool(np.array_equal(Ug.view(np.uint64), alone[k][0].view(np.uint64)))
            errs.append(e)
        res["error_vs_single_case_fp64"] = errs
        ens.close()
        for c in ctxs:
            c.close()
        out["forms"]["%s/ensemble" % prec] = res
    base = out["forms"]["plain/one-by-one"]["value"]
    for v in out["forms"].values():
        v["speedup_vs_one_by_one"] = v["value"] / base
    return out


def copy_probe(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (the library's streaming copy
    kernel): HBM copy bandwidth, and the scale for the FETCH_SIZE and
    WRITE_SIZE counters when run under rocprofv3 --pmc."""
    ms = xf.copy_probe(nbytes, reps)
    return {"bytes_read": nbytes, "bytes_written": nbytes, "reps": reps, "avg_ms": ms,
            "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps per form after the checkpoint")
    ap.add_argument("--startup", type=int, default=STARTUP, help="start-up steps before the checkpoint")
    ap.add_argument("--n", type=int, default=HEADLINE_N, help="headline grid n^3")
    ap.add_argument("--kernel-reps", type=int, default=60)
    ap.add_argument("--headline-only", action="store_true", help="only the mixing layer (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true", help="only the 1 GiB copy probe (counter scaling)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--c3-window-us", type=float, default=170.0)
    ap.add_argument("--c3-steps", type=int, default=200)
    ap.add_argument("--ensemble-cases", type=int, default=64)
    ap.add_argument("--ensemble-sides", default="128,256,512,1024")
    ap.add_argument("--c2-steps", type=int, default=500)
    ap.add_argument("--c3-only", action="store_true", help="only the shock tube window (its own JSON line)")
    ap.add_argument("--c2-only", action="store_true", help="only the shock-bubble leg (its own JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        print(json.dumps({"copy_probe": copy_probe()}))
        return
    if args.c3_only:
        print(json.dumps({"shock_tube": shock_tube(8192, args.c3_window_us * 1e-6, args.c3_steps, 10)}))
        return
    if args.c2_only:
        print(json.dumps({"shock_bubble": shock_bubble(args.c2_steps, 10, args.kernel_reps,
                                                       accuracy=not args.no_accuracy)}))
        return
    s, ml = mixing_layer(args.n, args.startup, args.warmup, args.steps)
    d = s.desc
    head = ml["forms"]["plain/per-cell"]
    fp64 = xf.fma_probe(10)
    roof = headline_roofline(d, head)
    roof["frac_measured_fma_rate"] = roof["achieved"] / fp64
    line = {
        "metric": "Mcell-updates/s per RK3 step, plain FP64 vs mixed/batched, at stated L1 err",
        "value": head["value"],
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ml["workload"], "grid": "%d^3" % args.n, "form": "plain FP64, per-cell chemistry",
                   "window": "steps %d-%d after %d start-up steps and %d warm-up steps" % (
                       args.startup + args.warmup + 1, args.startup + args.warmup + args.steps, args.startup,
                       args.warmup),
                   "finite": head["finite"]},
        "roofline": roof,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = mixing_layer_cpu(48, args.startup)
        progress("cpu baseline done")
    line["fp64_probe"] = {"tflops": fp64, "peak_datasheet": PEAK_FP64_TFLOPS}
    line["mixing_layer"] = ml
    if not args.headline_only:
        line["copy_probe"] = copy_probe()
        line["shock_bubble"] = shock_bubble(args.c2_steps, 10, args.kernel_reps, accuracy=not args.no_accuracy)
        line["shock_bubble"]["step_roofline"]["frac_measured_fma_rate"] = (
            line["shock_bubble"]["step_roofline"]["achieved_tflops"] / fp64)
        progress("shock-bubble done")
        line["vortex"] = vortex_leg()
        progress("vortex done")
        line["shock_tube"] = shock_tube(8192, args.c3_window_us * 1e-6, args.c3_steps, 10)
        progress("shock tube done")
        if args.ensemble_cases > 0:
            sides = [int(v) for v in args.ensemble_sides.split(",")]
            line["ensemble"] = ensemble_leg(args.ensemble_cases, sides, 5, 3, 1)
            progress("ensemble done")
    # the figures a reader looks for, last (they stay in a truncated tail)
    summ = {"mixing_layer_256": {k: {"value": round(f["value"], 2), "x": round(f["speedup_vs_plain_per_cell"], 3),
                                     "l1_rho": f.get("error_vs_plain_fp64", {}).get("l1_rel_rho"),
                                     "l1_T": f.get("error_vs_plain_fp64", {}).get("l1_rel_T")}
                                 for k, f in ml["forms"].items()},
            "roofline": "%s %.2f TFLOP/s = %.3f of %.1f" % (roof["kernel"], roof["achieved"], roof["frac"],
                                                             PEAK_FP64_TFLOPS)}
    if "shock_bubble" in line:
        sb = line["shock_bubble"]
        summ["shock_bubble_1024x512"] = {k: round(v["value"], 1) for k, v in sb["forms"].items()}
        summ["shock_bubble_roofline"] = "%s %.3f" % (sb["roofline"]["kernel"], sb["roofline"]["frac"])
    if "shock_tube" in line:
        summ["shock_tube_8192_window"] = {k: round(v["value"], 2) for k, v in line["shock_tube"]["forms"].items()}
    if "ensemble" in line:
        summ["ensemble_64"] = {k: round(v["value"], 2) for k, v in line["ensemble"]["forms"].items()}
    if "cpu_baseline" in line:
        summ["cpu_baseline"] = round(line["cpu_baseline"]["value"], 4)
    line["summary"] = summ
    print(json.dumps(line))


if __name__ == "__main__":
    main()
