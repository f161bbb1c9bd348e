"""The oracle's two face-flux arithmetics (CPU).

The default oracle rounds every product and sum of the characteristic face
flux in the reference's expression order (Eigen_callback.h:198-229,
WENO5s_schemes.hpp:268-279).  cfg.fused = 1 rounds each multiply-add pair of
the projections, the GLF split, the back-projection and the WENO5-Z
indicators and candidate polynomials once: the arithmetic of the device's
fused FP64 form.  These tests pin the WENO5-Z rounding order against Python's
own IEEE double arithmetic (which never fuses), and measure how far the
fused variant drifts from the reference's rounding over whole steps.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import lib
from parity import compare, field_errors, make, oracle_for


def _weno5z_side_py(v1, v2, v3, v4, v5):
    # weno5Z_P body (WENO5s_schemes.hpp:268-282) evaluated by Python floats:
    # every operation is one IEEE double rounding, left to right as written
    eps = 1.0e-6
    s1 = 13.0 * (v1 - 2.0 * v2 + v3) * (v1 - 2.0 * v2 + v3) + 3.0 * (v1 - 4.0 * v2 + 3.0 * v3) * (v1 - 4.0 * v2 + 3.0 * v3)
    s2 = 13.0 * (v2 - 2.0 * v3 + v4) * (v2 - 2.0 * v3 + v4) + 3.0 * (v2 - v4) * (v2 - v4)
    s3 = 13.0 * (v3 - 2.0 * v4 + v5) * (v3 - 2.0 * v4 + v5) + 3.0 * (3.0 * v3 - 4.0 * v4 + v5) * (3.0 * v3 - 4.0 * v4 + v5)
    s5 = abs(s1 - s3)
    a1 = 0.1 * (1.0 + s5 / (s1 + eps))
    a2 = 0.6 * (1.0 + s5 / (s2 + eps))
    a3 = 0.3 * (1.0 + s5 / (s3 + eps))
    w1 = a1 / (a1 + a2 + a3)
    w2 = a2 / (a1 + a2 + a3)
    w3 = a3 / (a1 + a2 + a3)
    return w1 * (2.0 * v1 - 7.0 * v2 + 11.0 * v3) / 6.0 + w2 * (-v2 + 5.0 * v3 + 2.0 * v4) / 6.0 + \
        w3 * (2.0 * v3 + 5.0 * v4 - v5) / 6.0


def test_weno5z_rounds_as_the_reference_writes_it():
    L = lib()
    rng = np.random.default_rng(11)
    DP = C.POINTER(C.c_double)
    same_fused = 0
    n = 4000
    for t in range(n):
        scale = 10.0 ** rng.uniform(-6, 6)
        pp = rng.standard_normal(10) * scale
        mm = rng.standard_normal(10) * scale
        ref = _weno5z_side_py(*pp[1:6]) + _weno5z_side_py(mm[6], mm[5], mm[4], mm[3], mm[2])
        a = L.ora_weno5z(pp.ctypes.data_as(DP), mm.ctypes.data_as(DP))
        b = L.ora_weno5z_fused(pp.ctypes.data_as(DP), mm.ctypes.data_as(DP))
        assert a == ref, (t, a, ref)
        same_fused += (b == ref)
        assert abs(b - ref) <= 1e-12 * max(abs(ref), scale)
    # the fused variant rounds differently on a good share of inputs
    assert same_fused < 0.9 * n, same_fused


@pytest.mark.parametrize("case,nsteps", [("shock_bubble_js_nasa", 6), ("shock_bubble_z_janaf", 10), ("vortex", 10),
                                         ("mixing_layer_3d", 2)])
def test_fused_variant_drift(case, nsteps):
    """The fused arithmetic departs from the reference's rounding: report the
    drift after the parity cases' step counts.  Bounded well inside the mixed
    form's 1e-6 bar, and not zero (the switch reaches every face)."""
    s, U, T, _ = make(case)
    out = []
    for fused in (0, 1):
        o = oracle_for(s, U, T, fused=fused)
        assert o.initialize() == 0
        st, _ = o.steps(nsteps)
        assert st == 0
        out.append((o.field("U").copy(), o.field("T").copy()))
    eq, rel = compare(out[0][0], out[1][0], s.desc.E)
    err = field_errors(s, out[0][0], out[0][1], out[1][0], out[1][1])
    print("%s %d steps: bitwise fraction %.4f, max rel per component %s, %s" %
          (case, nsteps, eq, np.array2string(rel, precision=2),
           {k: "%.2e" % v for k, v in err.items() if k.startswith(("l1_rel", "linf_rel"))}))
    assert eq < 1.0
    for k in ("rho", "p", "T"):
        if "l1_rel_" + k in err:
            assert err["l1_rel_" + k] <= 1e-10, (k, err)
