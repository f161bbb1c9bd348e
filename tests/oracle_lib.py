"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.  Builds the oracle configuration from an xfluids_amd.Setup (the
same thermo / mechanism tables the GPU context receives).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.environ.get("XF_ORACLE_SO", os.path.join(ROOT, "oracle", "_build", "liboracle.so"))
MAXNS, MAXNR = 16, 40
MAXE = MAXNS + 4


class OraConfig(C.Structure):
    _fields_ = [("Xmax", C.c_int), ("Ymax", C.c_int), ("Zmax", C.c_int),
                ("X_inner", C.c_int), ("Y_inner", C.c_int), ("Z_inner", C.c_int),
                ("Bwidth_X", C.c_int), ("Bwidth_Y", C.c_int), ("Bwidth_Z", C.c_int),
                ("DimX", C.c_int), ("DimY", C.c_int), ("DimZ", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("_dx", C.c_double), ("_dy", C.c_double), ("_dz", C.c_double), ("CFLnumber", C.c_double),
                ("BC", C.c_int * 6),
                ("NS", C.c_int), ("NR", C.c_int), ("cop", C.c_int), ("thermo", C.c_int), ("scheme", C.c_int),
                ("react", C.c_int), ("strang", C.c_int), ("ncop_gamma", C.c_double),
                ("Hia", C.c_double * (MAXNS * 21)), ("Hib", C.c_double * (MAXNS * 6)),
                ("Wi", C.c_double * MAXNS), ("_Wi", C.c_double * MAXNS), ("Ri", C.c_double * MAXNS),
                ("Nu_f", C.c_int * (MAXNR * MAXNS)), ("Nu_b", C.c_int * (MAXNR * MAXNS)),
                ("Nu_d", C.c_int * (MAXNR * MAXNS)), ("third_ind", C.c_int * MAXNR),
                ("React_ThirdCoef", C.c_double * (MAXNR * MAXNS)), ("Rargus", C.c_double * (MAXNR * 6)),
                ("use_libm", C.c_int), ("positivity", C.c_int),
                ("visc", C.c_int), ("visc_heat", C.c_int), ("visc_diffu", C.c_int),
                ("Yil_limiter", C.c_double), ("Dim_limiter", C.c_double),
                ("fcv", C.c_double * (MAXNS * 4)), ("fct", C.c_double * (MAXNS * 4)),
                ("Dkj", C.c_double * (MAXNS * MAXNS * 4)), ("fused", C.c_int), ("ghost", C.c_int)]


_lib = None
_DP = C.POINTER(C.c_double)
_IPT = C.POINTER(C.c_int)


def load(path):
    """ctypes handle of an oracle build (the plain or the counting library)."""
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    L = C.CDLL(path)
    v = C.c_void_p
    P = C.POINTER(OraConfig)
    sig = {
        "ora_create": (v, [P]), "ora_destroy": (None, [v]), "ora_num_eqn": (C.c_int, [v]),
        "ora_num_cells": (C.c_size_t, [v]), "ora_field": (_DP, [v, C.c_char_p]),
        "ora_boundary": (None, [v, C.c_int]), "ora_update_states": (C.c_int, [v, C.c_int]),
        "ora_compute_lu": (None, [v, C.c_int]), "ora_rk3": (None, [v, C.c_int, C.c_double]),
        "ora_get_dt": (C.c_double, [v]), "ora_chemistry": (None, [v, C.c_double]),
        "ora_reaction": (C.c_int, [v, C.c_double]), "ora_estimate_nan": (C.c_int, [v, C.c_int]),
        "ora_initialize": (C.c_int, [v]), "ora_steps": (C.c_int, [v, C.c_int, _DP]),
        "ora_exp": (C.c_double, [C.c_double]), "ora_log": (C.c_double, [C.c_double]),
        "ora_cp_i": (C.c_double, [P, C.c_double, C.c_int]), "ora_h_i": (C.c_double, [P, C.c_double, C.c_int]),
        "ora_s_i": (C.c_double, [P, C.c_double, C.c_int]), "ora_g_i": (C.c_double, [P, C.c_double, C.c_int]),
        "ora_e_i": (C.c_double, [P, C.c_double, C.c_int]),
        "ora_mix_cp": (C.c_double, [P, _DP, C.c_double]), "ora_mix_cv": (C.c_double, [P, _DP, C.c_double]),
        "ora_mix_h": (C.c_double, [P, _DP, C.c_double]), "ora_mix_e": (C.c_double, [P, _DP, C.c_double]),
        "ora_mix_R": (C.c_double, [P, _DP]), "ora_get_T": (C.c_double, [P, _DP, C.c_double, C.c_double]),
        "ora_weno5js": (C.c_double, [_DP, _DP]), "ora_weno5z": (C.c_double, [_DP, _DP]),
        "ora_weno5z_fused": (C.c_double, [_DP, _DP]),
        "ora_wenocu6": (C.c_double, [_DP, _DP, C.c_double]),
        "ora_eigen_lr": (None, [P, C.c_int, C.c_int, _DP, _DP, _DP, _DP, _DP] + [C.c_double] * 8),
        "ora_evalcv": (None, [P, _DP, _DP, _DP, C.c_double, C.c_double]),
        "ora_chemeq2": (C.c_int, [P, _DP, C.c_double, C.c_double, C.c_double]),
        "ora_face_c2": (C.c_double, [v, C.c_size_t, C.c_size_t, _DP, _DP]),
        "ora_set_time": (None, [v, C.c_double, C.c_int]), "ora_set_target_time": (None, [v, C.c_double]),
        "ora_get_time": (None, [v, _DP, _IPT, _IPT]),
        "ora_last_error": (C.c_int, [v, _IPT, _IPT, _IPT, _IPT, _IPT]),
    }
    for n, (r, a) in sig.items():
        f = getattr(L, n)
        f.restype, f.argtypes = r, a
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = load(ORACLE_SO)
    return _lib


RU = 6.02214076e26 * 1.380649e-23 * 1.0E-3


def config_from_setup(setup, use_libm=0, fused=0, **over):
    """OraConfig carrying the same geometry and tables as setup.context().
    fused = 1 selects the face-flux arithmetic of the fused FP64 device form
    (multiply-add pairs rounded once) instead of the reference's order."""
    d = setup.desc
    c = OraConfig()
    for name, _ in OraConfig._fields_:
        if hasattr(d, name) and name not in ("BC",):
            try:
                setattr(c, name, getattr(d, name))
            except TypeError:
                pass
    for q in range(6):
        c.BC[q] = d.BC[q]
    c.use_libm = use_libm
    c.fused = fused
    if d.cop:
        Hia, Hib, Wi = setup.thermo_tables()
        NS = d.NS
        for q in range(NS * 21):
            c.Hia[q] = Hia[q]
        for q in range(NS * 6):
            c.Hib[q] = Hib[q]
        for n in range(NS):
            c.Wi[n] = Wi[n]
            c._Wi[n] = 1.0 / Wi[n]
            c.Ri[n] = RU / Wi[n]
        if d.NR > 0:
            Nu_f, Nu_b, third, Rargus = setup.mechanism()
            for r in range(d.NR):
                s = 0.0
                for j in range(NS):
                    q = r * NS + j
                    c.Nu_f[q] = int(Nu_f[q])
                    c.Nu_b[q] = int(Nu_b[q])
                    c.Nu_d[q] = int(Nu_b[q]) - int(Nu_f[q])
                    c.React_ThirdCoef[q] = third[q]
                    s += third[q]
                c.third_ind[r] = 1 if s > 0.0 else 0
                for q in range(6):
                    c.Rargus[r * 6 + q] = Rargus[r * 6 + q]
        if d.visc:
            fcv, fct, Dkj = setup.transport_tables()
            for q in range(NS * 4):
                c.fcv[q] = fcv[q]
                c.fct[q] = fct[q]
            for q in range(NS * NS * 4):
                c.Dkj[q] = Dkj[q]
    for k, v in over.items():
        setattr(c, k, v)
    return c


class Oracle:
    """One case on the CPU oracle (AoS fields, reference layout)."""

    def __init__(self, cfg, L=None):
        self.cfg = cfg
        self.L = L or lib()
        self._h = self.L.ora_create(C.byref(cfg))
        self.E = self.L.ora_num_eqn(self._h)
        self.N = self.L.ora_num_cells(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.ora_destroy(self._h)
            self._h = None

    def field(self, name, n=None):
        sizes = {"U": self.E, "U1": self.E, "LU": self.E, "FluxF": self.E, "FluxG": self.E, "FluxH": self.E,
                 "FluxFw": self.E, "FluxGw": self.E, "FluxHw": self.E, "eigen_local_x": self.E,
                 "eigen_local_y": self.E, "eigen_local_z": self.E, "y": self.cfg.NS, "Vde": 9,
                 "Dkm": self.cfg.NS, "hi": self.cfg.NS}
        if name == "eigen_block":
            cnt = 3 * MAXE
        elif name == "uvw_c_max":
            cnt = 6
        else:
            cnt = self.N * sizes.get(name, 1)
        p = self.L.ora_field(self._h, name.encode())
        return np.ctypeslib.as_array(p, shape=(cnt,))

    def set_state(self, U, T):
        """Initial state; U1 = U as the reference's initial-condition kernels
        set it (e.g. src/solver_Ini/sample/2D-EulerVortex/ini_sample.hpp:84-88)."""
        self.field("U")[:] = U
        self.field("U1")[:] = U
        self.field("T")[:] = T

    def initialize(self):
        return self.L.ora_initialize(self._h)

    def steps(self, n):
        dts = np.zeros(max(n, 1))
        st = self.L.ora_steps(self._h, n, dts.ctypes.data_as(_DP))
        return st, dts[:n]

    def boundary(self, which):
        self.L.ora_boundary(self._h, which)

    def update_states(self, which):
        return self.L.ora_update_states(self._h, which)

    def compute_lu(self, which):
        self.L.ora_compute_lu(self._h, which)

    def rk3(self, stage, dt):
        self.L.ora_rk3(self._h, stage, dt)

    def get_dt(self):
        return self.L.ora_get_dt(self._h)

    def chemistry(self, dt):
        self.L.ora_chemistry(self._h, dt)

    def reaction(self, dt):
        return self.L.ora_reaction(self._h, dt)

    def set_time(self, time, step=0):
        self.L.ora_set_time(self._h, time, step)

    def set_target_time(self, target):
        self.L.ora_set_target_time(self._h, target)

    def time(self):
        t, s, r = C.c_double(), C.c_int(), C.c_int()
        self.L.ora_get_time(self._h, C.byref(t), C.byref(s), C.byref(r))
        return t.value, s.value, bool(r.value)

    def last_error(self):
        """(code, (i, j, k), step, stage) of the last physical error."""
        i, j, k, s, w = (C.c_int() for _ in range(5))
        code = self.L.ora_last_error(self._h, C.byref(i), C.byref(j), C.byref(k), C.byref(s), C.byref(w))
        return code, (i.value, j.value, k.value), s.value, w.value
