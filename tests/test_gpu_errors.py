"""Error path and clock on the GPU, against the oracle.

A bad value planted in one interior cell must stop the step where the
reference's checks stop it (src/XFLUIDS.cpp:211-225, 285-287, 377-389,
441-525): same code, same cell, same step and stage, through the step
drivers (stream, hipGraph replays, per-step reads) and the staged driver.
Inside graph replays the failing pass halts the rest of the run on the
device, so the state equals the oracle's at the failure.  The clock clamps
dt so that the time lands on a target (src/XFLUIDS.cpp:196-198).
"""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce, make, oracle_for
from test_oracle_errors import PLANTED

pytestmark = pytest.mark.gpu

NSTEPS = 4


def _oracle(case, comp, val, cell, buf):
    s, U, T, _ = make(case)
    d = s.desc
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    idc = (cell[2] * d.Ymax + cell[1]) * d.Xmax + cell[0]
    o.field(buf)[idc * d.E + comp] = val
    st, _ = o.steps(NSTEPS)
    return s, U, T, idc, st, o.last_error(), o.field("U").copy()


def _gpu_ctx(s, U, T, idc, comp, val, buf, staged=False):
    """Context after initialisation with the value planted: in U, or in the
    state the stage-1 NaN check reads (the oracle's U1).  That state is the
    previous step's second-stage result, which the fused step keeps in U2 and
    the staged driver, like the reference, in U1; it keeps the unplanted
    values when U is planted, as the oracle's U1 does."""
    E = s.desc.E
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    Ug, _ = ctx.download(0)
    Up = Ug.copy()
    Up[idc * E + comp] = val
    stale = 1 if staged else 2
    if buf == "U":
        ctx.upload(Up, None, which=0)
        ctx.upload(Ug, None, which=1)
        ctx.upload(Ug, None, which=2)
    else:
        ctx.upload(Up, None, which=stale)
    return ctx


def same_bits(a, b):
    """Bitwise equal, NaN positions equal (NaN payloads may differ)."""
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64))


@pytest.mark.parametrize("driver", ["stream", "graph", "per_step", "staged"])
@pytest.mark.parametrize("case,comp,val,cell,buf,code,stage", PLANTED)
def test_planted_cell_matches_oracle(case, comp, val, cell, buf, code, stage, driver):
    announce("planted_cell %s comp %d %s %s -> code %d, driver %s" % (case, comp, val, buf, code, driver))
    s, U, T, idc, st_o, err_o, Uo = _oracle(case, comp, val, cell, buf)
    assert st_o == code
    ctx = _gpu_ctx(s, U, T, idc, comp, val, buf, staged=(driver == "staged"))
    with pytest.raises(xf.XfError) as ei:
        if driver == "staged":
            ctx.evolve_staged(NSTEPS)
        else:
            ctx.steps(NSTEPS, graph=(driver == "graph"), record_dt=(driver == "per_step"))
    e = ei.value
    print("gpu", e.status, e.cell, e.step, e.stage, "oracle", err_o)
    assert e.status == err_o[0] == code
    assert e.cell == err_o[1]
    # stopped where the oracle stopped: the conserved state is the oracle's
    # (stage-1 and Reaction failures come before any fused update writes U)
    Ug, _ = ctx.download(0)
    assert same_bits(Ug, Uo)
    if driver != "staged":
        assert (e.step, e.stage) == (err_o[2], err_o[3])
        assert ctx.time()[1] == err_o[2]
    ctx.close()


def test_nan_flux_divergence_cell():
    announce("nan_flux_divergence shock_bubble_z_janaf")
    s, U, T, idc, st_o, err_o, _ = _oracle("shock_bubble_z_janaf", 2, np.nan, (20, 12, 0), "U")
    assert err_o == (3, (20, 9, 0), 1, 1)
    ctx = _gpu_ctx(s, U, T, idc, 2, np.nan, "U")
    with pytest.raises(xf.XfError) as ei:
        ctx.steps(NSTEPS, graph=True)
    assert (ei.value.status, ei.value.cell, ei.value.step, ei.value.stage) == (3, (20, 9, 0), 1, 1)
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
def test_clock_lands_on_target(graph):
    announce("clock_lands_on_target graph=%s" % graph)
    s, U, T, _ = make("vortex")
    free = oracle_for(s, U, T)
    assert free.initialize() == 0
    _, dts = free.steps(4)
    target = dts[0] + dts[1] + 0.5 * dts[2]
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    o.set_target_time(target)
    o.steps(6)
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    ctx.set_target_time(target)
    got = ctx.steps(6, graph=graph, record_dt=True)
    t, step, reached = ctx.time()
    assert (step, reached) == (3, True)
    assert t == o.time()[0] == target
    np.testing.assert_array_equal(got[:2], dts[:2])
    assert got[2] == target - (dts[0] + dts[1])
    assert np.all(got[3:] == 0.0)
    Ug, _ = ctx.download(0)
    assert np.array_equal(Ug.view(np.uint64), o.field("U").view(np.uint64))
    # further steps do nothing until a new target is set
    ctx.steps(2, graph=graph)
    Ug2, _ = ctx.download(0)
    assert np.array_equal(Ug2.view(np.uint64), Ug.view(np.uint64))
    ctx.set_target_time(np.inf)
    ctx.steps(1, graph=graph)
    assert ctx.time()[1] == 4
    ctx.close()


def test_checkpoint_sets_clock(tmp_path):
    announce("checkpoint_sets_clock")
    s, U, T, _ = make("vortex")
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    path = str(tmp_path / "ubak.dat")
    ctx.save_checkpoint(path, 40, 1.25)
    ctx.load_checkpoint(path)
    assert ctx.time()[:2] == (1.25, 40)
    ctx.close()
