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
