"""Shared helpers for GPU-vs-oracle parity tests (test infrastructure)."""
import sys

import numpy as np

import xfluids_amd as xf
from oracle_lib import Oracle, config_from_setup

# Workloads at sizes the oracle finishes in seconds: (case file, workload,
# overrides, steps)
CASES = {
    "vortex": ("isentropic-vortex", "isentropic-vortex", dict(nx=64, ny=64), 10),
    "shock_bubble_z_janaf": ("shock-bubble", "shock-bubble", dict(nx=128, ny=64), 10),
    "shock_bubble_js_nasa": ("shock-bubble", "shock-bubble", dict(nx=96, ny=48, scheme=0, thermo=1), 6),
    "reactive_shock_tube": ("reactive-shock-tube", "reactive-shock-tube", dict(nx=512), 8),
    "reactive_channel": ("reactive-channel", "reactive-channel", dict(nx=24, ny=24), 3),
    "mixing_layer_3d": ("mixing-layer", "mixing-layer", dict(nx=16, ny=16, nz=16), 2),
}


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def make(case, seed=7, precision=xf.XF_PLAIN_FP64):
    cf, wl, over, nsteps = CASES[case]
    s = xf.Setup(xf.case_path(cf), **over)
    U, T = s.initial(wl, seed=seed)
    return s, U, T, nsteps


def oracle_for(s, U, T, use_libm=0):
    o = Oracle(config_from_setup(s, use_libm=use_libm))
    o.set_state(U, T)
    return o


def compare(a, b, E=None):
    """(bitwise-equal fraction, max relative difference per component)."""
    a = np.asarray(a)
    b = np.asarray(b)
    eq = np.mean(a.view(np.uint64) == b.view(np.uint64))
    if E:
        a2, b2 = a.reshape(-1, E), b.reshape(-1, E)
        scale = np.maximum(np.abs(a2).max(axis=0), 1e-300)
        rel = (np.abs(a2 - b2).max(axis=0) / scale)
    else:
        scale = max(np.abs(a).max(), 1e-300)
        rel = np.array([np.abs(a - b).max() / scale])
    return eq, rel
