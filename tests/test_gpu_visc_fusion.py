"""Viscous wall fluxes computed inside the sweeps (the default) against
wall-flux launches of their own that write the flux planes the sweeps read:
both give the same bits, through the graph, per-step and staged drivers,
in the plain and the mixed form.  The parity tests check the default against
the oracle."""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce, make

pytestmark = pytest.mark.gpu

# in-sweep wall flux: the 9-, 10- and 11-species cases (3-D, 3-D, 2-D) and
# Inert-SBI (NS = 5, 3-D); the 3-species cases keep the separate launch in
# both settings
VISC_CASES = ["mixing_layer_3d", "mixing_layer_21", "rsbi", "inert_sbi_3d", "shock_bubble_visc",
              "shock_bubble_cu6_pp_visc"]


def _run(s, U, T, nsteps, fused, precision=xf.XF_PLAIN_FP64, mode="graph"):
    ctx = s.context(precision=precision)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.set_visc_fusion(fused)
    if mode == "staged":
        ctx.evolve_staged(nsteps)
    else:
        ctx.steps(nsteps, graph=(mode == "graph"))
    Ug, Tg = ctx.download(0)
    ctx.close()
    return Ug, Tg


@pytest.mark.parametrize("case", VISC_CASES)
@pytest.mark.parametrize("precision", [xf.XF_PLAIN_FP64, xf.XF_MIXED])
def test_fused_wall_flux_is_bitwise_equal(case, precision):
    announce("visc fusion %s precision %d" % (case, precision))
    s, U, T, nsteps = make(case)
    Ua, Ta = _run(s, U, T, nsteps, True, precision)
    Ub, Tb = _run(s, U, T, nsteps, False, precision)
    assert np.all(np.isfinite(Ua))
    assert np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64))
    assert np.array_equal(Ta.view(np.uint64), Tb.view(np.uint64))


@pytest.mark.parametrize("mode", ["step", "staged"])
def test_fused_wall_flux_in_other_drivers(mode):
    announce("visc fusion mixing_layer_3d driver %s" % mode)
    s, U, T, nsteps = make("mixing_layer_3d")
    Ua, Ta = _run(s, U, T, nsteps, True, mode=mode)
    Ub, Tb = _run(s, U, T, nsteps, False, mode=mode)
    assert np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64))
    assert np.array_equal(Ta.view(np.uint64), Tb.view(np.uint64))


def test_visc_fusion_switch_needs_viscous_terms():
    s, U, T, _ = make("vortex")
    ctx = s.context()
    with pytest.raises(xf.XfError):
        ctx.set_visc_fusion(False)
    ctx.close()


def test_switch_between_graph_replays():
    """Switching mid-run rebuilds the captured step: a fused step, then an
    unfused one, equals two unfused steps."""
    announce("visc fusion switch between graph replays")
    s, U, T, _ = make("mixing_layer_3d")
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(1, graph=True)
    ctx.set_visc_fusion(False)
    ctx.steps(1, graph=True)
    Ua, Ta = ctx.download(0)
    ctx.close()
    Ub, Tb = _run(s, U, T, 2, False)
    assert np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64))
    assert np.array_equal(Ta.view(np.uint64), Tb.view(np.uint64))
