"""The reactive shock tube in the reference's output window (t = 170 us,
settings/1d-reactive-st.json): from the committed state the CPU oracle
reached there on 1024 cells (tests/golden/make_shock_tube_fixture.py), the
plain form with per-cell, queue and one-wave-per-cell chemistry must stay
bit-identical to the oracle, the queue must equal the per-cell driver pass for pass, and the
mixed form must keep L1 of rho, p and T within 1e-6."""
import os

import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce, field_errors, oracle_for

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "shock_tube_1024_170us.npz")
NSTEPS = 20


def _fixture():
    z = np.load(FIX)
    s = xf.Setup(xf.case_path("reactive-shock-tube"), nx=int(z["nx"]))
    return s, z["U"], z["T"], float(z["time"]), int(z["steps"])


_ref = {}


def _oracle():
    if "o" not in _ref:
        s, U, T, t, n = _fixture()
        o = oracle_for(s, U, T)
        assert o.initialize() == 0
        o.set_time(t, n)
        st, dts = o.steps(NSTEPS)
        assert st == 0
        _ref["o"] = (o.field("U").copy(), o.field("T").copy(), dts.copy(), o.time())
    return _ref["o"]


def _gpu(precision, driver):
    s, U, T, t, n = _fixture()
    ctx = s.context(precision=precision, chem_driver=driver)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.set_time(t, n)
    ctx.update_states(0)
    passes = ctx.chemistry(0.5 * ctx.get_dt(), record=True)
    ctx.upload(U, T)
    ctx.initialize()
    dts = ctx.steps(NSTEPS, graph=True, record_dt=True)
    Ug, Tg = ctx.download(0)
    clock = ctx.time()
    ctx.close()
    return s, Ug, Tg, dts, clock, passes


@pytest.mark.parametrize("driver", [xf.XF_CHEM_PER_CELL, xf.XF_CHEM_QUEUE, xf.XF_CHEM_WAVE])
def test_plain_bitwise_in_the_window(driver):
    announce("shock tube window, plain, driver %d" % driver)
    Uo, To, dto, clock_o = _oracle()
    s, Ug, Tg, dts, clock, passes = _gpu(xf.XF_PLAIN_FP64, driver)
    print("passes: median %g max %d (cells > 10: %d)" % (np.median(passes), passes.max(), (passes > 10).sum()))
    assert np.array_equal(Ug.view(np.uint64), Uo.view(np.uint64))
    assert np.array_equal(Tg.view(np.uint64), To.view(np.uint64))
    assert np.array_equal(dts, dto)
    assert clock[:2] == clock_o[:2]


def test_queue_passes_equal_per_cell():
    announce("shock tube window, queue vs per-cell passes")
    _, _, _, _, _, p1 = _gpu(xf.XF_PLAIN_FP64, xf.XF_CHEM_PER_CELL)
    _, _, _, _, _, p2 = _gpu(xf.XF_PLAIN_FP64, xf.XF_CHEM_QUEUE)
    assert np.array_equal(p1, p2)
    assert p1.max() > 1  # some cells do more than one pass here


def test_mixed_error_in_the_window():
    announce("shock tube window, mixed")
    Uo, To, _, _ = _oracle()
    s, Ug, Tg, _, _, _ = _gpu(xf.XF_MIXED, xf.XF_CHEM_QUEUE)
    err = field_errors(s, Uo, To, Ug, Tg)
    print({k: v for k, v in err.items() if not k.endswith("Yk")})
    for k in ("rho", "p", "T"):
        assert err["l1_rel_" + k] <= 1e-6, (k, err)
    lo, hi = err["Y_range_oracle"]
    glo, ghi = err["Y_range_gpu"]
    assert glo >= lo - 1e-6 and ghi <= hi + 1e-6
