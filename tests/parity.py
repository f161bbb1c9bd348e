"""Shared helpers for GPU-vs-oracle parity tests (test infrastructure)."""
import sys

import numpy as np

import xfluids_amd as xf
from oracle_lib import RU, Oracle, config_from_setup

# Workloads at sizes the oracle finishes in seconds: (case file, workload,
# overrides, steps)
CASES = {
    "vortex": ("isentropic-vortex", "isentropic-vortex", dict(nx=64, ny=64), 10),
    "shock_bubble_z_janaf": ("shock-bubble", "shock-bubble", dict(nx=128, ny=64), 10),
    "shock_bubble_js_nasa": ("shock-bubble", "shock-bubble", dict(nx=96, ny=48, scheme=0, thermo=1), 6),
    "shock_bubble_cu6": ("shock-bubble", "shock-bubble", dict(nx=96, ny=48, scheme=2), 6),
    "vortex_cu6": ("isentropic-vortex", "isentropic-vortex", dict(nx=48, ny=48, scheme=2), 8),
    # the ghost fill of the wall and symmetry codes: x viscous wall / slip wall,
    # y symmetry / no-slip wall
    "vortex_walls": ("isentropic-vortex-walls", "isentropic-vortex", dict(nx=48, ny=48), 8),
    "shock_bubble_pp": ("shock-bubble", "shock-bubble", dict(nx=96, ny=48, positivity=1), 6),
    "shock_bubble_ns2": ("shock-bubble-h2n2", "shock-bubble", dict(nx=96, ny=48), 6),
    "shock_bubble_3d": ("shock-bubble-3d", "shock-bubble", dict(nx=32, ny=16, nz=16), 4),
    "reactive_shock_tube": ("reactive-shock-tube", "reactive-shock-tube", dict(nx=512), 8),
    "reactive_shock_tube_pp": ("reactive-shock-tube", "reactive-shock-tube", dict(nx=512, positivity=1), 8),
    # Lie splitting: one full-dt Reaction after the RK stages (XFLUIDS.cpp:211-225)
    "reactive_shock_tube_lie": ("reactive-shock-tube-lie", "reactive-shock-tube", dict(nx=512), 8),
    "reactive_channel": ("reactive-channel", "reactive-channel", dict(nx=24, ny=24), 3),
    # NASA-7 (JANAF) fits in the chemistry's Gibbs energies and the states
    "reactive_channel_janaf": ("reactive-channel", "reactive-channel", dict(nx=24, ny=24, thermo=0), 3),
    "mixing_layer_3d": ("mixing-layer", "mixing-layer", dict(nx=16, ny=16, nz=16), 2),
    "shock_bubble_visc": ("shock-bubble", "shock-bubble", dict(nx=96, ny=48, viscous=(1, 1, 1)), 6),
    "mixing_layer_3d_inviscid": ("mixing-layer", "mixing-layer", dict(nx=16, ny=16, nz=16, viscous=(0, 0, 0)), 2),
    # combinations: WENO-CU6 in 3-D (three and nine species, the latter viscous
    # and reacting), and CU6 with the limiter and the viscous terms together
    "shock_bubble_3d_cu6": ("shock-bubble-3d", "shock-bubble", dict(nx=32, ny=16, nz=16, scheme=2), 4),
    "mixing_layer_3d_cu6": ("mixing-layer", "mixing-layer", dict(nx=16, ny=16, nz=16, scheme=2), 2),
    "shock_bubble_cu6_pp_visc": ("shock-bubble", "shock-bubble",
                                 dict(nx=96, ny=48, scheme=2, positivity=1, viscous=(1, 1, 1)), 6),
    # the reference's own shock-bubble mixtures, GhostSpecies (the last N2
    # carries no mass; Getrhoyi renormalises and writes back): Inert-SBI
    # (NS = 5, WENO-CU6, limiter, viscous, 3-D as cmake/init_sample.cmake
    # configures it), without fuel (NS = 4, WENO5-Z, 2-D), the reacting
    # RSBI-19REA (NS = 11), and the 21-reaction H2O-N2 mixing layer (NS = 10)
    "inert_sbi_3d": ("inert-sbi", "shock-bubble", dict(nx=32, ny=16, nz=16), 4),
    "inert_sbi_no_fuel": ("inert-sbi-without-fuel", "shock-bubble", dict(nx=64, ny=32), 6),
    "rsbi": ("rsbi", "shock-bubble", dict(nx=64, ny=32), 4),
    "mixing_layer_21": ("mixing-layer-21", "mixing-layer", dict(nx=16, ny=16, nz=16), 2),
}


def announce(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def make(case, seed=7, precision=xf.XF_PLAIN_FP64):
    cf, wl, over, nsteps = CASES[case]
    s = xf.Setup(xf.case_path(cf), **over)
    U, T = s.initial(wl, seed=seed)
    return s, U, T, nsteps


def oracle_for(s, U, T, use_libm=0, fused=0):
    o = Oracle(config_from_setup(s, use_libm=use_libm, fused=fused))
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


def field_errors(s, Uo, To, Ug, Tg):
    """Errors of a GPU state against the oracle's over interior cells: L1
    (sum |d| / sum |ref|) and Linf (max |d| / max |ref|) of rho, p, T, and
    of the mass fractions (absolute), with their ranges on both sides."""
    d = s.desc
    E = d.E
    shp = (d.Zmax, d.Ymax, d.Xmax, E)
    sl = (slice(d.Bwidth_Z, d.Bwidth_Z + d.Z_inner), slice(d.Bwidth_Y, d.Bwidth_Y + d.Y_inner),
          slice(d.Bwidth_X, d.Bwidth_X + d.X_inner))
    Uo = np.asarray(Uo).reshape(shp)[sl].reshape(-1, E)
    Ug = np.asarray(Ug).reshape(shp)[sl].reshape(-1, E)
    To = np.asarray(To).reshape(shp[:3])[sl].ravel()
    Tg = np.asarray(Tg).reshape(shp[:3])[sl].ravel()

    def prims(U, T):
        rho = U[:, 0]
        if d.cop:
            Y = np.concatenate([U[:, 5:] / rho[:, None], 1.0 - (U[:, 5:] / rho[:, None]).sum(axis=1)[:, None]],
                               axis=1)
            _, _, W = s.thermo_tables()
            p = rho * RU * (Y / np.asarray(W)[None, :]).sum(axis=1) * T
            return rho, p, T, Y
        ke = 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / rho
        p = (d.ncop_gamma - 1.0) * (U[:, 4] - ke)
        return rho, p, None, None

    ro, po, To_, Yo = prims(Uo, To)
    rg, pg, Tg_, Yg = prims(Ug, Tg)
    out = {}
    for name, a, b in (("rho", ro, rg), ("p", po, pg), ("T", To_, Tg_)):
        if a is None:
            continue
        out["l1_rel_" + name] = float(np.abs(b - a).sum() / np.abs(a).sum())
        out["linf_rel_" + name] = float(np.abs(b - a).max() / np.abs(a).max())
    if Yo is not None:
        out["l1_abs_Y"] = float(np.abs(Yg - Yo).mean())
        out["linf_abs_Y"] = float(np.abs(Yg - Yo).max())
        # per species k, in the case's species order
        out["l1_abs_Yk"] = [float(v) for v in np.abs(Yg - Yo).mean(axis=0)]
        out["linf_abs_Yk"] = [float(v) for v in np.abs(Yg - Yo).max(axis=0)]
        out["Y_range_oracle"] = [float(Yo.min()), float(Yo.max())]
        out["Y_range_gpu"] = [float(Yg.min()), float(Yg.max())]
    return out


# ---- whole-cell shifts along periodic axes -----------------------------------
def grid_shape(d, E=None):
    shp = (d.Zmax, d.Ymax, d.Xmax)
    return shp + (E,) if E else shp


def periodic_axes(d, which):
    """(array axis, ghost width, interior cells) of the named axes ("x", "y", "z")."""
    spec = {"x": (2, d.Bwidth_X, d.X_inner), "y": (1, d.Bwidth_Y, d.Y_inner), "z": (0, d.Bwidth_Z, d.Z_inner)}
    return [spec[w] for w in which]


def _axis_slice(ax, lo, hi, ndim):
    s = [slice(None)] * ndim
    s[ax] = slice(lo, hi)
    return tuple(s)


def periodic_fill(A, axes):
    """Ghost layers of a padded (Zmax, Ymax, Xmax[, E]) array along each
    periodic axis, copied from the opposite interior edge (in place)."""
    for ax, B, n in axes:
        A[_axis_slice(ax, 0, B, A.ndim)] = A[_axis_slice(ax, n, n + B, A.ndim)]
        A[_axis_slice(ax, B + n, 2 * B + n, A.ndim)] = A[_axis_slice(ax, B, 2 * B, A.ndim)]
    return A


def roll_interior(A, axes, shifts):
    """A copy with the interior rolled by whole cells along each periodic axis
    (all rows of the other axes, ghosts included), ghosts refilled."""
    A = A.copy()
    for (ax, B, n), s in zip(axes, shifts):
        sl = _axis_slice(ax, B, B + n, A.ndim)
        A[sl] = np.roll(A[sl], s, axis=ax)
    return periodic_fill(A, axes)


def interior(A, d):
    return A[(slice(d.Bwidth_Z, d.Bwidth_Z + d.Z_inner), slice(d.Bwidth_Y, d.Bwidth_Y + d.Y_inner),
              slice(d.Bwidth_X, d.Bwidth_X + d.X_inner))]


def shifted_pair(s, U0, T0, which, shifts):
    """The starting state with its ghosts filled periodically along `which`,
    and the same state rolled by `shifts` cells: (Ua, Ta), (Ub, Tb), flat."""
    d = s.desc
    axes = periodic_axes(d, which)
    Ua = periodic_fill(np.asarray(U0).reshape(grid_shape(d, d.E)).copy(), axes)
    Ta = periodic_fill(np.asarray(T0).reshape(grid_shape(d)).copy(), axes)
    Ub, Tb = roll_interior(Ua, axes, shifts), roll_interior(Ta, axes, shifts)
    return (Ua.ravel(), Ta.ravel()), (Ub.ravel(), Tb.ravel())


def unshift_compare(s, which, shifts, Ra, Rb, Rb0, E=1):
    """Interior of run a's result rolled like b's start against run b's
    result, for a field of E components (U: E = d.E; T: 1): (values that
    differ, max |diff| / max |value| per component -- the momenta against the
    largest momentum component, the partial densities against the density --,
    whether b's field moved away from its start Rb0)."""
    d = s.desc
    axes = periodic_axes(d, which)
    shp = grid_shape(d, E if E > 1 else None)
    A = interior(roll_interior(np.asarray(Ra).reshape(shp), axes, shifts), d).reshape(-1, E)
    B = interior(np.asarray(Rb).reshape(shp), d).reshape(-1, E)
    B0 = interior(np.asarray(Rb0).reshape(shp), d).reshape(-1, E)
    neq = int(np.count_nonzero(A.view(np.uint64) != B.view(np.uint64)))
    scale = np.abs(B).max(axis=0)
    if E >= 5:  # the momenta against the largest of them, partial densities against rho
        scale[1:4] = scale[1:4].max()
        scale[5:] = scale[0]
    rel = np.abs(A - B).max(axis=0) / np.maximum(scale, 1e-300)
    return neq, rel, not np.array_equal(B, B0)
