"""The operation-counting build of the oracle computes exactly what the
plain build computes, and the algorithmic flop figure bench.py reports is the
one that build counts."""
import os
import sys

import numpy as np

import xfluids_amd as xf
from oracle_lib import Oracle, config_from_setup, load

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle", "opcount"))
import count_flops  # noqa: E402


def test_counting_build_is_bitwise_the_oracle():
    Lc = load(count_flops.COUNT_SO) if os.path.exists(count_flops.COUNT_SO) else count_flops._load()[1]
    s = xf.Setup(xf.case_path("shock-bubble"), nx=48, ny=40)
    U, T = s.initial("shock-bubble", seed=5)
    a = Oracle(config_from_setup(s))
    b = Oracle(config_from_setup(s), Lc)
    for o in (a, b):
        o.set_state(U, T)
        assert o.initialize() == 0
        st, _ = o.steps(2)
        assert st == 0
    assert np.array_equal(a.field("U").view(np.uint64), b.field("U").view(np.uint64))
    assert np.array_equal(a.field("T").view(np.uint64), b.field("T").view(np.uint64))


def test_bench_face_flops_is_the_counted_figure():
    sys.path.insert(0, ROOT)
    import bench
    r = count_flops.measure("shock-bubble")
    assert r["solve_max_residual"] < 1e-6
    assert r["face_x_flops"] == r["face_y_flops"] == bench.FACE_FLOPS


def test_bench_step_constants_are_the_counted_figures():
    """The per-unit counts bench.py's whole-step roofline uses: exact for
    the fixed-cost stages, to 1 % for UpdateStates and GetDt (their counts
    depend on the Newton iterations and maxima of the seeded field)."""
    sys.path.insert(0, ROOT)
    import bench
    r = count_flops.measure("shock-bubble")
    assert r["lu_per_padded_cell_flops"] == bench.LU_PADDED_FLOPS
    assert r["lu_per_inner_cell_flops"] == bench.LU_INNER_FLOPS
    assert r["rk3_per_inner_cell_flops"] == bench.RK3_INNER_FLOPS
    assert abs(r["states_per_padded_cell_flops"] / bench.STATES_FLOPS - 1) < 0.01
    assert abs(r["dt_per_padded_cell_flops"] / bench.DT_PADDED_FLOPS - 1) < 0.01
    s = xf.Setup(xf.case_path(bench.CASE), nx=bench.NX, ny=bench.NY)
    sr = bench.step_roofline(s.desc, 800.0, 70.0)
    # about 24.2 k operations per cell-update: 3 stages x 2 directions x 3869
    # per face, plus states, LU, RK3 and dt
    assert 23500 < sr["flops_per_cell_update"] < 25000
    assert abs(sr["achieved_tflops"] - sr["flops_per_cell_update"] * 800e6 / 1e12) < 1e-9


def test_bench_headline_constants_are_the_counted_figures():
    """The 13-variable face flux (inviscid part, x = y = z) and the
    chemistry's per-cell and per-pass counts of the mixing-layer workload."""
    sys.path.insert(0, ROOT)
    import bench
    r = count_flops.measure("mixing-layer", nz=8, inviscid=True)
    assert r["solve_max_residual"] < 1e-6
    assert r["face_x_flops"] == r["face_y_flops"] == r["face_z_flops"] == bench.FACE_FLOPS_E13
    # the viscous wall flux the sweeps add per face: the viscous solve's face
    # counts less the inviscid ones (that solve is not exact, see bench.py)
    rv = count_flops.measure("mixing-layer", nz=8)
    for q in ("face_x_flops", "face_y_flops", "face_z_flops"):
        assert abs(rv[q] - r[q] - bench.FACE_VISC_FLOPS_E13) < 1.0, (q, rv[q] - r[q])
    c = count_flops.measure_chem("mixing-layer")
    assert c["fit_max_rel_residual_flops"] < 1e-3
    assert abs(c["per_pass_flops"] - bench.CHEM_PASS_FLOPS) < 1e-3
    assert abs(c["per_cell_flops"] - bench.CHEM_CELL_FLOPS) < 1e-3
    assert c["per_pass"][4:6] == [bench.CHEM_PASS_EXP, bench.CHEM_PASS_LOG]
