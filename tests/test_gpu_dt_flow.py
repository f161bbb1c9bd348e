"""GetDt on caller-owned FlowData arrays (xf_get_dt_flow) against the
context's GetDt (xf_get_dt) and the oracle: the same CFL maxima and dt, bit
for bit, from the primitives of the same UpdateStates pass
(src/solver_GetDt/GlobalDt_block.hpp:6-114)."""
import ctypes as C

import numpy as np
import pytest

import xfluids_amd as xf
from devmem import DeviceArray
from parity import announce, make, oracle_for

pytestmark = pytest.mark.gpu

CASES = ["vortex", "shock_bubble_z_janaf", "shock_bubble_3d", "reactive_shock_tube"]


def flow_arrays(ctx):
    pr = ctx.prims()
    return {k: DeviceArray(pr[k]) for k in ("rho", "p", "u", "v", "w")}


def call_dt_flow(fn, desc, arrs, dims):
    uvw = DeviceArray(n=3)
    dt = DeviceArray(n=1)
    ptr = lambda k, on=True: arrs[k].ptr if on else None
    st = fn(C.byref(desc), ptr("rho"), ptr("p"), ptr("u", dims[0]), ptr("v", dims[1]), ptr("w", dims[2]),
            uvw.ptr, dt.ptr, None)
    assert st == 0, xf.ERRORS.get(st, st)
    return uvw.get(), dt.get()[0]


@pytest.mark.parametrize("case", CASES)
def test_dt_flow_equals_context_dt(case):
    announce("dt_flow %s" % case)
    s, U, T, _ = make(case)
    d = s.desc
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    dt_ctx = ctx.get_dt()
    _, uvw_ctx = ctx.maxima()
    arrs = flow_arrays(ctx)
    uvw, dt = call_dt_flow(xf.lib().xf_get_dt_flow, d, arrs, (d.DimX, d.DimY, d.DimZ))
    ctx.close()
    assert np.array_equal(np.float64(dt).view(np.uint64), np.float64(dt_ctx).view(np.uint64))
    assert np.array_equal(uvw.view(np.uint64), np.asarray(uvw_ctx).view(np.uint64))
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    assert o.get_dt() == dt
