"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function include/*.h declares, the settings reader, and the runtime.dat
subset this package ships.

When the upstream tree is present (build container only), its settings files
and runtime.dat tables are read as data and compared; elsewhere those tests
skip.
"""
import glob
import json
import os
import re
import subprocess

import numpy as np
import pytest

import xfluids_amd as xf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
UPSTREAM = "/root/reference"


def header_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(xf_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_function():
    assert os.path.exists(xf.LIB_PATH), "build the HIP library first"
    out = subprocess.run(["nm", "-D", "--defined-only", xf.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    declared = header_functions()
    assert len(declared) >= 40
    assert declared <= exported, sorted(declared - exported)
    # the Python binding declares a signature for each of them
    assert declared <= set(xf.exported_symbols()), sorted(declared - set(xf.exported_symbols()))
    xf.lib()  # loads without a GPU; nothing is called


def test_error_strings_and_invalid_arguments():
    L = xf.lib()
    for code in (0, 1, 2, 3, 5, -1, -2, -3):
        assert L.xf_error_string(code)
    with pytest.raises(Exception):
        xf.Setup(os.path.join(ROOT, "no-such-file.json"))


def test_json_comments_cut_the_line_at_the_first_slash(tmp_path):
    # settings lines are cut at their first '/', so a path-valued string has
    # to be written as an array of path components
    p = tmp_path / "c.json"
    p.write_text('{\n "run": {"CFLnumber": 0.3}, // trailing comment\n'
                 ' "equations": {"Mixture": ["Reaction", "H2O_19_reaction"]},\n'
                 ' "mesh": {"DOMAIN_Size": [1.0, 0, 0], "Resolution": [64, 0, 0],\n'
                 '          "Boundarys": [1, 1, 2, 2, 2, 2]}  // another\n}\n')
    s = xf.Setup(str(p))
    d = s.desc
    assert abs(d.CFLnumber - 0.3) < 1e-15
    assert s.species == ["H2", "O2", "O", "H", "OH", "HO2", "H2O2", "H2O", "AR"]
    assert d.Xmax == 64 + 2 * d.Bwidth_X and d.Ymax == 1 and d.Zmax == 1
    assert d.cop == 1 and d.NS == 9 and d.NR == 19


def test_case_files_load():
    for f in sorted(glob.glob(os.path.join(xf.CASES, "*.json"))):
        s = xf.Setup(f)
        d = s.desc
        for dim, n, B, m in ((d.DimX, d.X_inner, d.Bwidth_X, d.Xmax), (d.DimY, d.Y_inner, d.Bwidth_Y, d.Ymax),
                             (d.DimZ, d.Z_inner, d.Bwidth_Z, d.Zmax)):
            assert m == (n + 2 * B if dim else 1)


@pytest.mark.skipif(not os.path.isdir(UPSTREAM), reason="upstream tree not present")
def test_upstream_settings_geometry():
    files = sorted(glob.glob(os.path.join(UPSTREAM, "settings", "*.json")))
    assert files
    for f in files:
        txt = "\n".join(line.split("/")[0] for line in open(f).read().splitlines())
        js = json.loads(re.sub(r",(\s*[}\]])", r"\1", txt))
        res = js["mesh"]["Resolution"]
        size = js["mesh"].get("DOMAIN_Size", [1.0, 1.0, 1.0])  # default of read_json.cpp:107
        s = xf.Setup(f)
        d = s.desc
        assert list(d.BC) == js["mesh"]["Boundarys"], f
        assert abs(d.CFLnumber - js["run"]["CFLnumber"]) < 1e-15
        for k, (dim, inner, mx, h) in enumerate(((d.DimX, d.X_inner, d.Xmax, d.dx), (d.DimY, d.Y_inner, d.Ymax, d.dy),
                                                 (d.DimZ, d.Z_inner, d.Zmax, d.dz))):
            if res[k]:
                assert dim and inner == res[k] and mx == res[k] + 8, f
                assert abs(h - size[k] / res[k]) <= 1e-15 * h, f
            else:
                assert not dim and mx == 1, f


def _blocks(path):
    toks = open(path).read().split()
    out = {}
    for i, t in enumerate(toks):
        if t.startswith("*") and t != "*END":
            vals = []
            for v in toks[i + 1:]:
                if v.startswith("*"):
                    break
                try:
                    vals.append(float(v))
                except ValueError:
                    pass
            out[t] = vals
    return out


@pytest.mark.skipif(not os.path.isdir(UPSTREAM), reason="upstream tree not present")
@pytest.mark.parametrize("fname", ["thermal_dynamics.dat", "thermal_dynamics_janaf.dat"])
def test_runtime_subset_matches_upstream_tables(fname):
    ours = _blocks(os.path.join(xf.RUNTIME_DAT, fname))
    theirs = _blocks(os.path.join(UPSTREAM, "runtime.dat", fname))
    assert len(ours) >= 9
    for k, v in ours.items():
        assert np.array_equal(np.array(v), np.array(theirs[k][:len(v)])), k


@pytest.mark.skipif(not os.path.isdir(UPSTREAM), reason="upstream tree not present")
@pytest.mark.parametrize("mech", ["H2O_19_reaction", "H2O-N2_19_reaction"])
def test_runtime_subset_mechanisms_match_upstream(mech):
    a = open(os.path.join(xf.RUNTIME_DAT, "Reaction", mech, "reaction_list.dat")).read()
    b = open(os.path.join(UPSTREAM, "runtime.dat", "Reaction", mech, "reaction_list.dat")).read()
    assert a == b


def test_checkpoint_format_round_trip(tmp_path):
    """The reference's checkpoint layout: int Step, double Time, float
    Time_consumption, then Emax*Xmax*Ymax*Zmax doubles (AoS, ghosts included)."""
    import struct
    for case, over in (("shock-bubble", dict(nx=24, ny=16)), ("isentropic-vortex", dict(nx=12, ny=10))):
        s = xf.Setup(xf.case_path(case), **over)
        d = s.desc
        E = d.NS + 4 if d.cop else 5
        n = E * d.Xmax * d.Ymax * d.Zmax
        U = np.random.default_rng(5).normal(size=n)
        p = tmp_path / ("ck_" + case)
        xf.checkpoint_write(p, d, U, 1234, 3.25e-4, 17.5)
        raw = p.read_bytes()
        assert len(raw) == 4 + 8 + 4 + 8 * n
        step, tm, tc = struct.unpack("<idf", raw[:16])
        assert (step, tm, tc) == (1234, 3.25e-4, 17.5)
        assert np.array_equal(np.frombuffer(raw[16:], dtype="<f8"), U)
        U2, step2, tm2, tc2 = xf.checkpoint_read(p, d)
        assert np.array_equal(U2.view(np.uint64), U.view(np.uint64)) and (step2, tm2, tc2) == (1234, 3.25e-4, 17.5)
        # a file for another geometry is refused
        p.write_bytes(raw + b"\0")
        with pytest.raises(xf.XfError):
            xf.checkpoint_read(p, d)
        p.write_bytes(raw[:-8])
        with pytest.raises(xf.XfError):
            xf.checkpoint_read(p, d)


def test_viscous_switches_need_a_mixture_and_wide_ghosts():
    """The viscous terms read species transport data and reach 4 cells into
    the ghost layers: a single-species case, or a ghost width below 4, is
    refused before any device allocation."""
    import ctypes as C
    with pytest.raises(xf.XfError):
        xf.Setup(xf.case_path("isentropic-vortex"), nx=16, ny=16, viscous=(1, 0, 0))
    s = xf.Setup(xf.case_path("shock-bubble"), nx=16, ny=8, viscous=(1, 1, 1))
    d = s.desc
    assert d.visc == 1 and d.visc_heat == 1 and d.visc_diffu == 1
    # BlSz.Yil_limiter = max(_dx, _dy, _dz) * clamp(1e10, 0, 1); Dim_limiter = clamp(2e-3, 0, 1)
    assert d.Yil_limiter == max(d._dx, d._dy, d._dz) and d.Dim_limiter == 2e-3
    bad = CaseDescCopy(d)
    bad.Bwidth_X, bad.Xmax = 3, bad.X_inner + 6
    h = C.c_void_p()
    assert xf.lib().xf_create(C.byref(bad), C.byref(h)) == -1
    nocop = CaseDescCopy(d)
    nocop.cop, nocop.NS = 0, 1
    assert xf.lib().xf_create(C.byref(nocop), C.byref(h)) == -1


def CaseDescCopy(d):
    import ctypes as C
    out = xf.CaseDesc()
    C.memmove(C.byref(out), C.byref(d), C.sizeof(d))
    return out


def test_shipped_mechanisms_have_compiled_rate_evaluations():
    """Every reactive case's tables (parsed by the host reader) equal the
    compile-time records of a shipped mechanism, and only exactly those: a
    changed Arrhenius factor, third-body efficiency or stoichiometric
    coefficient finds none (host-side matching, no GPU)."""
    found = set()
    for case in ("reactive-channel", "reactive-shock-tube", "mixing-layer"):
        s = xf.Setup(xf.case_path(case))
        d = s.desc
        k = s.mech_specialization_index()
        assert k >= 0, case
        found.add(k)
        Nu_f, Nu_b, third, Rargus = s.mechanism()
        L = xf.lib()
        idx = lambda *a: L.xf_mech_specialization_index(d.NS, d.NR, *a)
        i32 = lambda a: a.ctypes.data_as(xf.C.POINTER(xf.C.c_int))
        f64 = lambda a: a.ctypes.data_as(xf.C.POINTER(xf.C.c_double))
        assert idx(i32(Nu_f), i32(Nu_b), f64(third), f64(Rargus)) == k
        R2 = Rargus.copy()
        R2[6 * 3] = np.nextafter(R2[6 * 3], np.inf)        # A of reaction 3, one ulp
        assert idx(i32(Nu_f), i32(Nu_b), f64(third), f64(R2)) == -1
        t2 = third.copy()
        t2[np.flatnonzero(t2)[0]] *= 1.5
        assert idx(i32(Nu_f), i32(Nu_b), f64(t2), f64(Rargus)) == -1
        f2 = Nu_f.copy()
        f2[0] += 1
        assert idx(i32(f2), i32(Nu_b), f64(third), f64(Rargus)) == -1
        assert L.xf_mech_specialization_index(d.NS - 1, d.NR, i32(Nu_f), i32(Nu_b), f64(third), f64(Rargus)) == -1
    # the two shipped mechanisms (with argon and with nitrogen) are distinct
    assert len(found) == 2


def test_reaction_switch():
    """equations.Sources_React as a setup switch: off gives the inert step of
    a reacting case; on needs a mechanism."""
    s = xf.Setup(xf.case_path("reactive-channel"), reaction=False)
    assert s.desc.react == 0 and s.desc.NR == 19
    assert xf.Setup(xf.case_path("reactive-channel"), reaction=True).desc.react == 1
    with pytest.raises(xf.XfError):
        xf.Setup(xf.case_path("shock-bubble"), reaction=True)
