"""Mixed-precision form (FP32 smoothness indicators, weights and
characteristic projections, and FP32 exponentials and Wilke factors in the
transport coefficients; FP64 everywhere else) against the FP64 oracle.

Bar, after the parity cases' step counts: L1 relative error of density,
pressure and temperature at most 1e-6, and the mass fractions no further
than 1e-6 outside the oracle's own range.
"""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import CASES, announce, field_errors, make, oracle_for

pytestmark = pytest.mark.gpu

TOL_L1 = 1e-6
TOL_Y = 1e-6


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_form_error_vs_oracle(case):
    announce("mixed_form_error_vs_oracle " + case)
    s, U, T, nsteps = make(case)
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    st, _ = o.steps(nsteps)
    assert st == 0
    ctx = s.context(precision=xf.XF_MIXED)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(nsteps, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    err = field_errors(s, o.field("U"), o.field("T"), Ug, Tg)
    print(err)
    for k in ("rho", "p", "T"):
        if "l1_rel_" + k in err:
            assert err["l1_rel_" + k] <= TOL_L1, (k, err)
    if "Y_range_gpu" in err:
        lo, hi = err["Y_range_oracle"]
        glo, ghi = err["Y_range_gpu"]
        assert glo >= lo - TOL_Y and ghi <= hi + TOL_Y, err
    assert np.all(np.isfinite(Ug))
