"""xf_get_dt_flow_noalloc (GetDt on caller arrays with its workspace in the
library's fixed device array) against xf_get_dt_flow (the same kernels with
an allocated workspace), bit for bit, on two calls in a row: a grid large
enough to use every reduction block of the fixed array, then a smaller one
(whose fold must not see the first call's leftover maxima)."""
import ctypes as C

import numpy as np
import pytest

import xfluids_amd as xf
from devmem import DeviceArray, hip
from parity import announce, make

pytestmark = pytest.mark.gpu


def _call(fn, desc, arrs, dims):
    uvw = DeviceArray(n=3)
    dt = DeviceArray(n=1)
    ptr = lambda k, on=True: arrs[k].ptr if on else None
    st = fn(C.byref(desc), ptr("rho"), ptr("p"), ptr("u", dims[0]), ptr("v", dims[1]), ptr("w", dims[2]),
            uvw.ptr, dt.ptr, None)
    assert st == 0, st
    assert hip().hipDeviceSynchronize() == 0
    return uvw.get(), dt.get()[0]


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


def test_noalloc_equals_allocating_dt_twice():
    L = xf.lib()
    # first call: a 3-D grid of 1.4 M padded cells (more than 4096 blocks of
    # 256 cells, so the reduction strides over the fixed array's every triple)
    big = xf.Setup(xf.case_path("mixing-layer"), nx=104, ny=104, nz=104).desc
    N = big.Xmax * big.Ymax * big.Zmax
    announce("noalloc dt: synthetic %dx%dx%d (%d cells)" % (big.Xmax, big.Ymax, big.Zmax, N))
    assert N > 4096 * 256
    rng = np.random.default_rng(5)
    host = {"rho": rng.uniform(0.1, 2.0, N), "p": rng.uniform(1e4, 2e5, N),
            "u": rng.normal(0, 300, N), "v": rng.normal(0, 300, N), "w": rng.normal(0, 300, N)}
    arrs = {k: DeviceArray(v) for k, v in host.items()}
    dims = (big.DimX, big.DimY, big.DimZ)
    uvw_a, dt_a = _call(L.xf_get_dt_flow, big, arrs, dims)
    uvw_n, dt_n = _call(L.xf_get_dt_flow_noalloc, big, arrs, dims)
    print("big: dt %r %r" % (dt_a, dt_n))
    assert _same(uvw_a, uvw_n) and _same(dt_a, dt_n)
    # the same maxima on the host
    c = np.sqrt(1.4 * host["p"] / host["rho"])
    ref = [np.max(np.abs(host[k]) + c) for k in ("u", "v", "w")]
    assert _same(uvw_n, ref)
    for a in arrs.values():
        a.free()
    # second call: a smaller 2-D case, from its own UpdateStates primitives,
    # also against the context's GetDt
    s, U, T, _ = make("vortex")
    d = s.desc
    announce("noalloc dt: vortex %dx%d" % (d.Xmax, d.Ymax))
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    dt_ctx = ctx.get_dt()
    pr = ctx.prims()
    ctx.close()
    arrs = {k: DeviceArray(pr[k]) for k in ("rho", "p", "u", "v", "w")}
    dims = (d.DimX, d.DimY, d.DimZ)
    uvw_n, dt_n = _call(L.xf_get_dt_flow_noalloc, d, arrs, dims)
    uvw_a, dt_a = _call(L.xf_get_dt_flow, d, arrs, dims)
    print("vortex: dt %r %r %r" % (dt_n, dt_a, dt_ctx))
    assert _same(uvw_a, uvw_n) and _same(dt_a, dt_n) and _same(dt_n, dt_ctx)


def test_noalloc_rejects_bad_arguments():
    L = xf.lib()
    d = xf.Setup(xf.case_path("isentropic-vortex"), nx=16, ny=16).desc
    a = DeviceArray(n=d.Xmax * d.Ymax)
    dt = DeviceArray(n=1)
    # a missing array of an active direction
    assert L.xf_get_dt_flow_noalloc(C.byref(d), a.ptr, a.ptr, None, a.ptr, None, None, dt.ptr, None) == xf.XF_EINVAL
    # more cells than the limit (2^31 - 1)
    huge = xf.CaseDesc.from_buffer_copy(d)
    huge.Xmax = huge.Ymax = 2048
    huge.Zmax = 1024
    assert L.xf_get_dt_flow_noalloc(C.byref(huge), a.ptr, a.ptr, a.ptr, a.ptr, a.ptr, None, dt.ptr,
                                    None) == xf.XF_EINVAL
