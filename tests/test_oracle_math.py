"""Pins of the oracle's deterministic exp/log against libm (math module).

The plain FP64 GPU form and the oracle share this exp/log method; here it is
checked against correctly-rounded-ish libm results: exp within 1 ulp and log
within 2 ulp over wide ranges, special values exact.
"""
import math

import numpy as np

from oracle_lib import lib


def ulp_diff(a, b):
    ia = np.array(a, dtype=np.float64).view(np.int64)
    ib = np.array(b, dtype=np.float64).view(np.int64)
    return np.abs(ia - ib)


def test_exp_within_1ulp():
    L = lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-1, 1, 20000),
                         rng.uniform(-50, 50, 20000), rng.uniform(-745.13, -708.4, 5000),
                         rng.uniform(700.0, 709.78, 5000), [0.0, 1.0, -1.0, 0.5, 1e-300, -1e-300]])
    got = np.array([L.ora_exp(float(x)) for x in xs])
    ref = np.array([math.exp(float(x)) for x in xs])
    # the upper and lower bands reach the two-factor 2^k scaling and subnormal results
    assert ulp_diff(got, ref).max() <= 1


def test_log_within_2ulp():
    L = lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0.5, 2.0, 20000),
                         rng.uniform(100, 6000, 20000), [1.0, 2.0, 0.5, 1e-310, 5e-324, 1.7e308]])
    got = np.array([L.ora_log(float(x)) for x in xs])
    ref = np.array([math.log(float(x)) for x in xs])
    assert ulp_diff(got, ref).max() <= 2


def test_special_values():
    L = lib()
    assert L.ora_exp(0.0) == 1.0
    assert L.ora_log(1.0) == 0.0
    assert L.ora_exp(800.0) == math.inf
    assert L.ora_exp(-800.0) == 0.0
    assert math.isnan(L.ora_exp(math.nan))
    assert L.ora_log(0.0) == -math.inf
    assert math.isnan(L.ora_log(-1.0))
    assert L.ora_log(math.inf) == math.inf
