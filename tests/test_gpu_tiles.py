"""Sweep tile shapes and the autotuner: every shape computes the same bits,
and tuning on the live state leaves the following steps bit-identical to an
untuned run (the running GLF maxima are restored)."""
import itertools

import numpy as np
import pytest

from parity import announce, make

pytestmark = pytest.mark.gpu


def _run(s, U, T, nsteps, shapes=None, tune=False, precision=0):
    ctx = s.context(precision=precision)
    ctx.upload(U, T)
    ctx.initialize()
    if shapes is not None:
        for dr, sh in enumerate(shapes):
            ctx.set_tile_shape(dr, sh)
    chosen = None
    if tune:
        chosen, ms = ctx.autotune(reps=3)
        print("autotune", chosen, ms)
    ctx.steps(nsteps, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    return Ug, Tg, chosen


@pytest.mark.parametrize("case", ["shock_bubble_z_janaf", "mixing_layer_3d", "vortex"])
def test_every_tile_shape_is_bitwise_equal(case):
    announce("tile shapes " + case)
    s, U, T, nsteps = make(case)
    d = s.desc
    ref, Tref, _ = _run(s, U, T, nsteps)
    xs = range(2) if d.DimX else [0]
    ys = range(3) if d.DimY else [0]
    zs = range(3) if d.DimZ else [0]
    for shp in itertools.product(xs, ys, zs):
        Ug, Tg, _ = _run(s, U, T, nsteps, shapes=shp)
        assert np.array_equal(ref.view(np.uint64), Ug.view(np.uint64)), shp
        assert np.array_equal(Tref.view(np.uint64), Tg.view(np.uint64)), shp


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_results(precision):
    announce("autotune shock_bubble_z_janaf precision %d" % precision)
    s, U, T, nsteps = make("shock_bubble_z_janaf")
    ref, Tref, _ = _run(s, U, T, nsteps, precision=precision)
    Ug, Tg, chosen = _run(s, U, T, nsteps, tune=True, precision=precision)
    assert 0 <= chosen[0] < 2 and 0 <= chosen[1] < 3
    assert np.array_equal(ref.view(np.uint64), Ug.view(np.uint64))
