"""xfluids_amd: MI355X (gfx950) implementation of the XFluids per-step hot path.

Python binding over the C-ABI in include/xf_api.h and include/xf_host.h
(ctypes).  The compute path is the HIP library lib/libxfluids_amd.so; there is
no CPU fallback: loading fails loudly when the library is missing.

Typical use::

    import xfluids_amd as xf
    s = xf.Setup(xf.case_path("shock-bubble"))
    ctx = s.context()
    U, T = s.initial("shock-bubble", seed=1)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(10)
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libxfluids_amd.so")
RUNTIME_DAT = os.path.join(PKG, "data", "runtime.dat")
CASES = os.path.join(PKG, "cases")

XF_MAXNS = 16
XF_MAXNR = 40
XF_WENO5JS, XF_WENO5Z, XF_WENOCU6 = 0, 1, 2
XF_PLAIN_FP64, XF_MIXED, XF_FUSED_FP64 = 0, 1, 2
XF_CHEM_PER_CELL, XF_CHEM_QUEUE, XF_CHEM_WAVE = 0, 1, 2
XF_EINVAL, XF_EHIP, XF_ENOKERNEL = -1, -2, -3
ERRORS = {1: "XF_ERR_RHO_Y", 2: "XF_ERR_PRIM", 3: "XF_ERR_NAN", 5: "XF_ERR_CHEM",
          -1: "XF_EINVAL", -2: "XF_EHIP", -3: "XF_ENOKERNEL"}


class CaseDesc(C.Structure):
    """xf_case_desc (include/xf_api.h)."""
    _fields_ = [("Xmax", C.c_int), ("Ymax", C.c_int), ("Zmax", C.c_int),
                ("X_inner", C.c_int), ("Y_inner", C.c_int), ("Z_inner", C.c_int),
                ("Bwidth_X", C.c_int), ("Bwidth_Y", C.c_int), ("Bwidth_Z", C.c_int),
                ("DimX", C.c_int), ("DimY", C.c_int), ("DimZ", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("_dx", C.c_double), ("_dy", C.c_double), ("_dz", C.c_double),
                ("CFLnumber", C.c_double), ("BC", C.c_int * 6),
                ("NS", C.c_int), ("NR", C.c_int), ("cop", C.c_int), ("thermo", C.c_int),
                ("scheme", C.c_int), ("react", C.c_int), ("strang", C.c_int),
                ("ncop_gamma", C.c_double), ("precision", C.c_int), ("chem_driver", C.c_int),
                ("positivity", C.c_int), ("visc", C.c_int), ("visc_heat", C.c_int), ("visc_diffu", C.c_int),
                ("Yil_limiter", C.c_double), ("Dim_limiter", C.c_double), ("ghost", C.c_int)]

    @property
    def E(self):
        return self.NS + 4 if self.cop else 5

    @property
    def ncell(self):
        return self.Xmax * self.Ymax * self.Zmax

    @property
    def ninner(self):
        return self.X_inner * self.Y_inner * self.Z_inner


class XfError(RuntimeError):
    """A non-zero status.  For a physical error (status > 0): cell is the
    padded-grid (i, j, k) of the failure, step the Iteration and stage the RK
    stage (1..3) or 4 for a Reaction it happened in (0, 0 for a stage function
    called on its own)."""

    def __init__(self, status, where="", cell=None, step=0, stage=0):
        self.status = status
        self.cell = cell
        self.step = step
        self.stage = stage
        msg = "%s returned %d (%s)" % (where, status, ERRORS.get(status, "?"))
        if cell is not None:
            msg += " at cell (i,j,k)=%s" % (cell,)
        if step:
            msg += " in step %d, %s" % (step, "Reaction" if stage == 4 else "RK stage %d" % stage)
        super().__init__(msg)


_lib = None

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


_V = C.c_void_p
# C-ABI entry points: include/xf_api.h and include/xf_host.h
_SIGS = {
    "xf_create": (C.c_int, [C.POINTER(CaseDesc), C.POINTER(_V)]),
    "xf_destroy": (None, [_V]),
    "xf_num_eqn": (C.c_int, [_V]),
    "xf_get_desc": (C.c_int, [_V, C.POINTER(CaseDesc)]),
    "xf_num_cells": (C.c_size_t, [_V]),
    "xf_error_string": (C.c_char_p, [C.c_int]),
    "xf_set_thermo": (C.c_int, [_V, _DP, _DP, _DP]),
    "xf_set_mechanism": (C.c_int, [_V, _IP, _IP, _DP, _DP]),
    "xf_set_transport": (C.c_int, [_V, _DP, _DP, _DP]),
    "xf_mech_specialization_index": (C.c_int, [C.c_int, C.c_int, _IP, _IP, _DP, _DP]),
    "xf_set_chem_specialization": (C.c_int, [_V, C.c_int]),
    "xf_chem_specialization": (C.c_int, [_V]),
    "xf_set_chem_handover": (C.c_int, [_V, C.c_int]),
    "xf_upload_state": (C.c_int, [_V, C.c_int, _DP, _DP]),
    "xf_download_state": (C.c_int, [_V, C.c_int, _DP, _DP]),
    "xf_download_prims": (C.c_int, [_V] + [_DP] * 10),
    "xf_download_lu": (C.c_int, [_V, _DP]),
    "xf_download_maxima": (C.c_int, [_V, _DP, _DP]),
    "xf_boundary": (C.c_int, [_V, C.c_int]),
    "xf_update_states": (C.c_int, [_V, C.c_int]),
    "xf_compute_lu": (C.c_int, [_V, C.c_int]),
    "xf_rk3_stage": (C.c_int, [_V, C.c_int, C.c_double]),
    "xf_get_dt": (C.c_int, [_V, _DP]),
    "xf_get_dt_flow": (C.c_int, [C.POINTER(CaseDesc), _V, _V, _V, _V, _V, _V, _V, _V]),
    "xf_get_dt_flow_noalloc": (C.c_int, [C.POINTER(CaseDesc), _V, _V, _V, _V, _V, _V, _V, _V]),
    "xf_chemistry": (C.c_int, [_V, C.c_double, _IP]),
    "xf_reaction": (C.c_int, [_V, C.c_double]),
    "xf_estimate_nan": (C.c_int, [_V, C.c_int]),
    "xf_initialize": (C.c_int, [_V]),
    "xf_steps": (C.c_int, [_V, C.c_int, C.c_int, _DP]),
    "xf_steps_async": (C.c_int, [_V, C.c_int, C.c_int]),
    "xf_synchronize": (C.c_int, [_V]),
    "xf_stream": (_V, [_V]),
    "xf_last_error": (C.c_int, [_V, _IP, _IP, _IP, _IP]),
    "xf_last_error_step": (C.c_int, [_V, _IP, _IP]),
    "xf_set_time": (C.c_int, [_V, C.c_double, C.c_int]),
    "xf_set_target_time": (C.c_int, [_V, C.c_double]),
    "xf_get_time": (C.c_int, [_V, _DP, _IP, _IP]),
    "xf_set_profiling": (C.c_int, [_V, C.c_int]),
    "xf_set_profile_filter": (C.c_int, [_V, C.c_char_p]),
    "xf_time_kernel": (C.c_int, [_V, C.c_char_p, C.c_int, _DP]),
    "xf_kernel_times": (C.c_int, [_V, C.POINTER(C.c_char_p), _DP, _IP, C.c_int]),
    "xf_setup_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(_V), C.c_char_p, C.c_size_t]),
    "xf_setup_free": (None, [_V]),
    "xf_setup_override": (C.c_int, [_V, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]),
    "xf_setup_set_positivity": (C.c_int, [_V, C.c_int]),
    "xf_setup_set_reaction": (C.c_int, [_V, C.c_int]),
    "xf_setup_set_viscous": (C.c_int, [_V, C.c_int, C.c_int, C.c_int]),
    "xf_setup_transport": (C.c_int, [_V, _DP, _DP, _DP]),
    "xf_setup_transport_exact": (C.c_double, [_V, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
    "xf_setup_desc": (C.c_int, [_V, C.POINTER(CaseDesc)]),
    "xf_setup_thermo": (C.c_int, [_V, _DP, _DP, _DP]),
    "xf_setup_mechanism": (C.c_int, [_V, _IP, _IP, _DP, _DP]),
    "xf_setup_mass_fractions": (C.c_int, [_V, _DP, _DP]),
    "xf_setup_species": (C.c_char_p, [_V, C.c_int]),
    "xf_setup_domain": (C.c_int, [_V, _DP]),
    "xf_setup_initial": (C.c_int, [_V, C.c_char_p, C.c_ulonglong, _DP, _DP]),
    "xf_setup_context": (C.c_int, [_V, C.c_int, C.c_int, C.POINTER(_V)]),
    "xf_evolve_staged": (C.c_int, [_V, C.c_int, _DP]),
    "xf_checkpoint_write": (C.c_int, [C.c_char_p, _V, _DP, C.c_int, C.c_double, C.c_float]),
    "xf_checkpoint_read": (C.c_int, [C.c_char_p, _V, _DP, _IP, _DP, C.POINTER(C.c_float)]),
    "xf_save_checkpoint": (C.c_int, [_V, C.c_char_p, C.c_int, C.c_double, C.c_float]),
    "xf_load_checkpoint": (C.c_int, [_V, C.c_char_p, _IP, _DP, C.POINTER(C.c_float)]),
    "xf_debug_math": (C.c_int, [_DP, _DP, _DP, C.c_int]),
    "xf_copy_probe": (C.c_int, [C.c_size_t, C.c_int, _DP]),
    "xf_fma_probe": (C.c_int, [C.c_int, _DP]),
    "xf_set_tile_shape": (C.c_int, [_V, C.c_int, C.c_int]),
    "xf_set_visc_fusion": (C.c_int, [_V, C.c_int]),
    "xf_get_tile_shape": (C.c_int, [_V, _IP]),
    "xf_autotune": (C.c_int, [_V, C.c_int, _IP, _DP]),
    "xf_ensemble_create": (C.c_int, [C.POINTER(_V), C.c_int, C.POINTER(_V)]),
    "xf_ensemble_steps": (C.c_int, [_V, C.c_int, C.c_int]),
    "xf_ensemble_synchronize": (C.c_int, [_V]),
    "xf_ensemble_cells": (C.c_longlong, [_V]),
    "xf_ensemble_one_launch": (C.c_int, [_V]),
    "xf_ensemble_destroy": (None, [_V]),
}


def _sig(lib):
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("xfluids_amd: %s not built; run `python -m xfluids_amd.build`" % LIB_PATH)
        _lib = _sig(C.CDLL(LIB_PATH))
    return _lib


def exported_symbols():
    """Names of the C-ABI functions the library exports."""
    return list(_SIGS)


def _d(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_IP) if a is not None else None


def case_path(name):
    return os.path.join(CASES, name + ".json")


class Setup:
    """Case setup from a settings JSON (reference schema) + runtime.dat."""

    def __init__(self, settings, runtime_root=RUNTIME_DAT, nx=0, ny=0, nz=0, cfl=-1.0, scheme=-1, thermo=-1,
                 positivity=-1, viscous=None, reaction=None):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        st = L.xf_setup_load(settings.encode(), runtime_root.encode(), C.byref(h), err, 512)
        if st:
            raise XfError(st, "xf_setup_load: " + err.value.decode())
        self._h = h
        if nx or ny or nz or cfl > 0 or scheme >= 0 or thermo >= 0:
            st = L.xf_setup_override(h, nx, ny, nz, cfl, scheme, thermo)
            if st:
                raise XfError(st, "xf_setup_override")
        if positivity >= 0:
            st = L.xf_setup_set_positivity(h, int(positivity))
            if st:
                raise XfError(st, "xf_setup_set_positivity")
        if reaction is not None:
            # equations.Sources_React: False gives the inert step of a reacting case
            st = L.xf_setup_set_reaction(h, 1 if reaction else 0)
            if st:
                raise XfError(st, "xf_setup_set_reaction")
        if viscous is not None:
            # (visc, heat, diffu) switches, as equations.Visc / Visc_Heat / Visc_Diffu
            v, h_, df = (list(viscous) + [0, 0, 0])[:3]
            st = L.xf_setup_set_viscous(h, int(v), int(h_), int(df))
            if st:
                raise XfError(st, "xf_setup_set_viscous")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.xf_setup_free(self._h)
            self._h = None

    @property
    def desc(self):
        d = CaseDesc()
        lib().xf_setup_desc(self._h, C.byref(d))
        return d

    @property
    def species(self):
        d = self.desc
        return [lib().xf_setup_species(self._h, n).decode() for n in range(d.NS)]

    def thermo_tables(self):
        NS = self.desc.NS
        Hia, Hib, Wi = np.zeros(NS * 21), np.zeros(NS * 6), np.zeros(NS)
        lib().xf_setup_thermo(self._h, _d(Hia), _d(Hib), _d(Wi))
        return Hia, Hib, Wi

    def transport_tables(self):
        """Fits of ln(mu_k), ln(lambda_k), ln(p D_kj) in ln T: fcv[NS*4], fct[NS*4], Dkj[NS*NS*4]."""
        NS = self.desc.NS
        fcv, fct, Dkj = np.zeros(NS * 4), np.zeros(NS * 4), np.zeros(NS * NS * 4)
        st = lib().xf_setup_transport(self._h, _d(fcv), _d(fct), _d(Dkj))
        if st:
            raise XfError(st, "xf_setup_transport")
        return fcv, fct, Dkj

    def transport_exact(self, what, k, j, T, p=101325.0):
        """Kinetic-theory value behind a fit: what 0 mu_k [Pa s], 1 lambda_k [W/(m K)], 2 D_kj [cm^2/s]."""
        return lib().xf_setup_transport_exact(self._h, int(what), int(k), int(j), float(T), float(p))

    def mech_specialization_index(self):
        """Compiled mechanism whose records equal this case's tables, or -1
        (xf_mech_specialization_index; host only)."""
        d = self.desc
        if d.NR <= 0:
            return -1
        Nu_f, Nu_b, third, Rargus = self.mechanism()
        return int(lib().xf_mech_specialization_index(d.NS, d.NR, _i(Nu_f), _i(Nu_b), _d(third), _d(Rargus)))

    def mechanism(self):
        d = self.desc
        n = max(1, d.NR * d.NS)
        Nu_f, Nu_b = np.zeros(n, np.int32), np.zeros(n, np.int32)
        third, Rargus = np.zeros(n), np.zeros(max(1, d.NR * 6))
        lib().xf_setup_mechanism(self._h, _i(Nu_f), _i(Nu_b), _d(third), _d(Rargus))
        return Nu_f, Nu_b, third, Rargus

    def mass_fractions(self):
        NS = self.desc.NS
        a, b = np.zeros(NS), np.zeros(NS)
        lib().xf_setup_mass_fractions(self._h, _d(a), _d(b))
        return a, b

    def domain(self):
        v = np.zeros(6)
        lib().xf_setup_domain(self._h, _d(v))
        return v

    def initial(self, workload, seed=0):
        d = self.desc
        U = np.zeros(d.ncell * d.E)
        T = np.zeros(d.ncell)
        st = lib().xf_setup_initial(self._h, workload.encode(), seed, _d(U), _d(T))
        if st:
            raise XfError(st, "xf_setup_initial(%s)" % workload)
        return U, T

    def context(self, precision=XF_PLAIN_FP64, chem_driver=XF_CHEM_PER_CELL):
        h = C.c_void_p()
        st = lib().xf_setup_context(self._h, precision, chem_driver, C.byref(h))
        if st:
            raise XfError(st, "xf_setup_context")
        return Context(h)


class Context:
    """One case resident on the GPU (xf_ctx)."""

    def __init__(self, handle):
        self._h = handle
        d = CaseDesc()
        lib().xf_get_desc(handle, C.byref(d))
        self.desc = d

    @classmethod
    def create(cls, desc):
        h = C.c_void_p()
        st = lib().xf_create(C.byref(desc), C.byref(h))
        if st:
            raise XfError(st, "xf_create")
        return cls(h)

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.xf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, st, where):
        if st:
            cell, step, stage = None, 0, 0
            if st > 0:
                code, i, j, k = C.c_int(), C.c_int(), C.c_int(), C.c_int()
                lib().xf_last_error(self._h, C.byref(code), C.byref(i), C.byref(j), C.byref(k))
                cell = (i.value, j.value, k.value)
                s_, w_ = C.c_int(), C.c_int()
                lib().xf_last_error_step(self._h, C.byref(s_), C.byref(w_))
                step, stage = s_.value, w_.value
            raise XfError(st, where, cell, step, stage)
        return st

    # ---- clock -------------------------------------------------------------
    def set_time(self, time, step=0):
        self._chk(lib().xf_set_time(self._h, time, step), "xf_set_time")

    def set_target_time(self, target):
        self._chk(lib().xf_set_target_time(self._h, target), "xf_set_target_time")

    def time(self):
        """(physical time, steps counted, target reached) of the device clock."""
        t, s_, r = C.c_double(), C.c_int(), C.c_int()
        self._chk(lib().xf_get_time(self._h, C.byref(t), C.byref(s_), C.byref(r)), "xf_get_time")
        return t.value, s_.value, bool(r.value)

    # ---- data --------------------------------------------------------------
    def upload(self, U, T=None, which=0):
        U = np.ascontiguousarray(U, dtype=np.float64)
        Tn = None if T is None else np.ascontiguousarray(T, dtype=np.float64)
        self._chk(lib().xf_upload_state(self._h, which, _d(U), _d(Tn)), "xf_upload_state")

    def download(self, which=0):
        d = self.desc
        U = np.zeros(d.ncell * d.E)
        T = np.zeros(d.ncell)
        self._chk(lib().xf_download_state(self._h, which, _d(U), _d(T)), "xf_download_state")
        return U, T

    def prims(self):
        d = self.desc
        names = ["rho", "p", "u", "v", "w", "H", "c", "gamma", "e"]
        out = {n: np.zeros(d.ncell) for n in names}
        y = np.zeros(d.ncell * d.NS)
        self._chk(lib().xf_download_prims(self._h, *[_d(out[n]) for n in names], _d(y)), "xf_download_prims")
        out["y"] = y
        return out

    def lu(self):
        d = self.desc
        a = np.zeros(d.ncell * d.E)
        self._chk(lib().xf_download_lu(self._h, _d(a)), "xf_download_lu")
        return a

    def maxima(self):
        eb, uvw = np.zeros(3 * (XF_MAXNS + 4)), np.zeros(3)
        self._chk(lib().xf_download_maxima(self._h, _d(eb), _d(uvw)), "xf_download_maxima")
        return eb.reshape(3, XF_MAXNS + 4), uvw

    # ---- stage seam ----------------------------------------------------------
    def boundary(self, which=0):
        self._chk(lib().xf_boundary(self._h, which), "xf_boundary")

    def update_states(self, which=0):
        return self._chk(lib().xf_update_states(self._h, which), "xf_update_states")

    def compute_lu(self, which=0):
        self._chk(lib().xf_compute_lu(self._h, which), "xf_compute_lu")

    def rk3_stage(self, stage, dt):
        self._chk(lib().xf_rk3_stage(self._h, stage, dt), "xf_rk3_stage")

    def get_dt(self):
        dt = C.c_double()
        self._chk(lib().xf_get_dt(self._h, C.byref(dt)), "xf_get_dt")
        return dt.value

    def chemistry(self, dt_chem, record=False):
        sub = np.zeros(self.desc.ninner, np.int32) if record else None
        self._chk(lib().xf_chemistry(self._h, dt_chem, _i(sub)), "xf_chemistry")
        return sub

    def reaction(self, dt):
        self._chk(lib().xf_reaction(self._h, dt), "xf_reaction")

    def estimate_nan(self, flag):
        self._chk(lib().xf_estimate_nan(self._h, flag), "xf_estimate_nan")

    def initialize(self):
        self._chk(lib().xf_initialize(self._h), "xf_initialize")

    def steps(self, n, graph=False, record_dt=False):
        dts = np.zeros(n) if record_dt else None
        self._chk(lib().xf_steps(self._h, n, 1 if graph else 0, _d(dts)), "xf_steps")
        return dts

    def steps_async(self, n, graph=False):
        self._chk(lib().xf_steps_async(self._h, n, 1 if graph else 0), "xf_steps_async")

    def synchronize(self):
        self._chk(lib().xf_synchronize(self._h), "xf_synchronize")

    def evolve_staged(self, n):
        dts = np.zeros(n)
        self._chk(lib().xf_evolve_staged(self._h, n, _d(dts)), "xf_evolve_staged")
        return dts

    def stream(self):
        return lib().xf_stream(self._h)

    def set_profiling(self, on=True, only=None):
        """Per-launch HIP-event timing on the context stream; only = list of
        kernel classes to time (others run without events)."""
        self._chk(lib().xf_set_profiling(self._h, 1 if on else 0), "xf_set_profiling")
        csv = ",".join(only) if only else ""
        self._chk(lib().xf_set_profile_filter(self._h, csv.encode()), "xf_set_profile_filter")

    def time_kernel(self, name, reps=50):
        """Average device ms of one kernel class over reps back-to-back launches."""
        ms = C.c_double()
        self._chk(lib().xf_time_kernel(self._h, name.encode(), reps, C.byref(ms)), "xf_time_kernel")
        return ms.value

    def set_tile_shape(self, direction, shape):
        self._chk(lib().xf_set_tile_shape(self._h, direction, shape), "xf_set_tile_shape")

    def set_visc_fusion(self, on=True):
        """Viscous wall fluxes inside the sweeps (True) or as launches of
        their own before each sweep (False); the same bits either way."""
        self._chk(lib().xf_set_visc_fusion(self._h, 1 if on else 0), "xf_set_visc_fusion")

    def tile_shape(self):
        out = np.zeros(3, np.int32)
        self._chk(lib().xf_get_tile_shape(self._h, _i(out)), "xf_get_tile_shape")
        return out

    def set_chem_specialization(self, on=True):
        """Use the compiled rate evaluation of a shipped mechanism when the
        tables match one (default), or the generic table loop
        (xf_set_chem_specialization)."""
        self._chk(lib().xf_set_chem_specialization(self._h, 1 if on else 0), "xf_set_chem_specialization")

    def set_chem_handover(self, passes):
        """Queue chemistry: hand a cell to a whole wave after `passes` passes
        (0: never).  Switched off in this build: only 0 is accepted."""
        self._chk(lib().xf_set_chem_handover(self._h, passes), "xf_set_chem_handover")

    def chem_specialization(self):
        """Index of the compiled mechanism in use, or -1 (xf_chem_specialization)."""
        return int(lib().xf_chem_specialization(self._h))

    def autotune(self, reps=20):
        """Pick the fastest sweep tile shape per direction (xf_autotune);
        returns (shapes[3], ms[3, 3])."""
        shapes = np.zeros(3, np.int32)
        ms = np.zeros(9)
        self._chk(lib().xf_autotune(self._h, reps, _i(shapes), _d(ms)), "xf_autotune")
        return shapes, ms.reshape(3, 3)

    def save_checkpoint(self, path, step, time, time_consumption=0.0):
        """Write U in the reference's checkpoint format (xf_save_checkpoint)."""
        self._chk(lib().xf_save_checkpoint(self._h, str(path).encode(), step, time, time_consumption),
                  "xf_save_checkpoint")

    def load_checkpoint(self, path):
        """Read a checkpoint into U (T is kept); returns (step, time, time_consumption)."""
        st, tm, tc = C.c_int(), C.c_double(), C.c_float()
        self._chk(lib().xf_load_checkpoint(self._h, str(path).encode(), C.byref(st), C.byref(tm), C.byref(tc)),
                  "xf_load_checkpoint")
        return st.value, tm.value, tc.value

    def kernel_times(self):
        cap = 32
        names = (C.c_char_p * cap)()
        avg = np.zeros(cap)
        cnt = np.zeros(cap, np.int32)
        n = lib().xf_kernel_times(self._h, names, _d(avg), _i(cnt), cap)
        return {names[i].decode(): (float(avg[i]), int(cnt[i])) for i in range(min(n, cap))}


def debug_math(x):
    """Device exp(x), log(x) (the kernels' FP64 functions) for host values."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    e, l = np.zeros_like(x), np.zeros_like(x)
    st = lib().xf_debug_math(_d(x), _d(e), _d(l), x.size)
    if st:
        raise XfError(st, "xf_debug_math")
    return e, l


def copy_probe(nbytes=1 << 30, reps=5):
    """Average ms of a streaming device copy of nbytes (xf_copy_probe)."""
    ms = C.c_double(0.0)
    st = lib().xf_copy_probe(nbytes, reps, C.byref(ms))
    if st != 0:
        raise XfError(st, "xf_copy_probe")
    return ms.value


def fma_probe(reps=10):
    """Achieved FP64 fused multiply-add throughput in TFLOP/s (xf_fma_probe)."""
    tf = C.c_double(0.0)
    st = lib().xf_fma_probe(reps, C.byref(tf))
    if st != 0:
        raise XfError(st, "xf_fma_probe")
    return tf.value


class Ensemble:
    """Several contexts stepped together (xf_ensemble): each case's stages on
    its own stream, one chemistry launch for all cases per Reaction."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (_V * len(self.contexts))(*[c._h for c in self.contexts])
        h = _V()
        st = lib().xf_ensemble_create(arr, len(self.contexts), C.byref(h))
        if st:
            raise XfError(st, "xf_ensemble_create")
        self._h = h

    @property
    def one_launch(self):
        """True when each step is one launch per kernel class over all cases."""
        return lib().xf_ensemble_one_launch(self._h) == 1

    @property
    def cells(self):
        return int(lib().xf_ensemble_cells(self._h))

    def steps(self, n, graph=True):
        st = lib().xf_ensemble_steps(self._h, n, 1 if graph else 0)
        if st:
            raise XfError(st, "xf_ensemble_steps")

    def synchronize(self):
        st = lib().xf_ensemble_synchronize(self._h)
        if st:
            raise XfError(st, "xf_ensemble_synchronize")

    def close(self):
        if getattr(self, "_h", None):
            lib().xf_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def checkpoint_write(path, desc, U_aos, step, time, time_consumption=0.0):
    """Write an AoS state in the reference's checkpoint format (host only)."""
    U = np.ascontiguousarray(U_aos, dtype=np.float64)
    st = lib().xf_checkpoint_write(str(path).encode(), C.byref(desc), _d(U), step, time, time_consumption)
    if st:
        raise XfError(st, "xf_checkpoint_write")


def checkpoint_read(path, desc):
    """Read a checkpoint written for `desc`: (U_aos, step, time, time_consumption)."""
    E = desc.NS + 4 if desc.cop else 5
    U = np.zeros(E * desc.Xmax * desc.Ymax * desc.Zmax)
    st, tm, tc = C.c_int(), C.c_double(), C.c_float()
    r = lib().xf_checkpoint_read(str(path).encode(), C.byref(desc), _d(U), C.byref(st), C.byref(tm), C.byref(tc))
    if r:
        raise XfError(r, "xf_checkpoint_read")
    return U, st.value, tm.value, tc.value
