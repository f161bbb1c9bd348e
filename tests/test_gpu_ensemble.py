"""Ensembles: independent cases of different sizes stepped together (one
graph launch per step, one chemistry launch for all cases; one launch per
kernel class over all cases when their configuration agrees) give every case
the bits it gets when stepped alone, in every precision form."""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce

pytestmark = pytest.mark.gpu

REACTIVE = [dict(nx=24, ny=24, seed=1), dict(nx=32, ny=16, seed=2), dict(nx=16, ny=40, seed=3), dict(nx=8, ny=8, seed=4)]
INERT = [dict(nx=64, ny=32, seed=1), dict(nx=48, ny=24, seed=2)]
VISCOUS = [dict(nx=64, ny=32, seed=1, viscous=(1, 1, 1)), dict(nx=48, ny=24, seed=2, viscous=(1, 1, 1)),
           dict(nx=40, ny=40, seed=3)]


POSITIVE = [dict(nx=64, ny=32, seed=1, positivity=1), dict(nx=40, ny=24, seed=2, positivity=1)]
THREE_D = [dict(nx=24, ny=16, nz=16, seed=1), dict(nx=16, ny=16, nz=8, seed=2), dict(nx=8, ny=8, nz=8, seed=3)]


def _cases(case, specs):
    out = []
    for sp in specs:
        kw = {k: sp[k] for k in ("nx", "ny", "nz", "positivity") if k in sp}
        s = xf.Setup(xf.case_path(case), viscous=sp.get("viscous"), **kw)
        U, T = s.initial(case if case != "shock-bubble-3d" else "shock-bubble", seed=sp["seed"])
        out.append((s, U, T))
    return out


def _alone(s, U, T, n, precision=xf.XF_PLAIN_FP64):
    ctx = s.context(precision=precision)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(n, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    return Ug, Tg


@pytest.mark.parametrize("case,specs,nsteps,one_launch", [
    ("reactive-channel", REACTIVE, 3, True), ("shock-bubble", INERT, 4, True),
    ("shock-bubble", VISCOUS, 3, False), ("shock-bubble", POSITIVE, 4, True), ("shock-bubble-3d", THREE_D, 3, True)])
@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("precision", [xf.XF_PLAIN_FP64, xf.XF_MIXED])
def test_ensemble_matches_single_cases(case, specs, nsteps, one_launch, graph, precision):
    announce("ensemble %s graph=%s precision=%d" % (case, graph, precision))
    cases = _cases(case, specs)
    ctxs = []
    for s, U, T in cases:
        ctx = s.context(precision=precision)
        ctx.upload(U, T)
        ctx.initialize()
        ctxs.append(ctx)
    ens = xf.Ensemble(ctxs)
    assert ens.cells == sum(s.desc.ninner for s, _, _ in cases)
    assert ens.one_launch == one_launch
    ens.steps(nsteps, graph=graph)
    ens.synchronize()
    got = [c.download(0) for c in ctxs]
    ens.close()
    for c in ctxs:
        c.close()
    for (s, U, T), (Ug, Tg) in zip(cases, got):
        Ua, Ta = _alone(s, U, T, nsteps, precision)
        assert np.array_equal(Ua.view(np.uint64), Ug.view(np.uint64))
        assert np.array_equal(Ta.view(np.uint64), Tg.view(np.uint64))


@pytest.mark.skip(reason="whole-wave chemistry hand-over switched off: its first GPU run did not finish and "
                         "the cause has not been shown")
@pytest.mark.parametrize("after", [1, 8])
def test_ensemble_handover_matches_single_cases(after):
    """The shared chemistry queue of an ensemble hands stiff cells of any
    case to whole waves: every case keeps its single-run bits."""
    announce("ensemble chemistry hand-over after %d passes" % after)
    cases = _cases("reactive-channel", REACTIVE)
    ctxs = []
    for s, U, T in cases:
        ctx = s.context()
        ctx.upload(U, T)
        ctx.initialize()
        ctxs.append(ctx)
    ctxs[0].set_chem_handover(after)
    ens = xf.Ensemble(ctxs)
    ens.steps(3, graph=True)
    ens.synchronize()
    got = [c.download(0) for c in ctxs]
    ens.close()
    for c in ctxs:
        c.close()
    for (s, U, T), (Ug, Tg) in zip(cases, got):
        Ua, Ta = _alone(s, U, T, 3)
        assert np.array_equal(Ua.view(np.uint64), Ug.view(np.uint64))
        assert np.array_equal(Ta.view(np.uint64), Tg.view(np.uint64))


def test_ensemble_rejects_mixed_species():
    a = xf.Setup(xf.case_path("reactive-channel"), nx=8, ny=8).context()
    b = xf.Setup(xf.case_path("shock-bubble"), nx=8, ny=8).context()
    with pytest.raises(xf.XfError):
        xf.Ensemble([a, b])
    with pytest.raises(xf.XfError):
        xf.Ensemble([a, a])
    a.close()
    b.close()
