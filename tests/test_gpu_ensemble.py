"""Ensembles: independent cases of different sizes stepped together (one
graph launch per step, one chemistry launch for all cases) give every case
the bits it gets when stepped alone."""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce

pytestmark = pytest.mark.gpu

REACTIVE = [dict(nx=24, ny=24, seed=1), dict(nx=32, ny=16, seed=2), dict(nx=16, ny=40, seed=3), dict(nx=8, ny=8, seed=4)]
INERT = [dict(nx=64, ny=32, seed=1), dict(nx=48, ny=24, seed=2)]
VISCOUS = [dict(nx=64, ny=32, seed=1, viscous=(1, 1, 1)), dict(nx=48, ny=24, seed=2, viscous=(1, 1, 1)),
           dict(nx=40, ny=40, seed=3)]


def _cases(case, specs):
    out = []
    for sp in specs:
        s = xf.Setup(xf.case_path(case), nx=sp["nx"], ny=sp["ny"], viscous=sp.get("viscous"))
        U, T = s.initial(case, seed=sp["seed"])
        out.append((s, U, T))
    return out


def _alone(s, U, T, n):
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(n, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    return Ug, Tg


@pytest.mark.parametrize("case,specs,nsteps", [("reactive-channel", REACTIVE, 3), ("shock-bubble", INERT, 4),
                                                ("shock-bubble", VISCOUS, 3)])
@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_matches_single_cases(case, specs, nsteps, graph):
    announce("ensemble %s graph=%s" % (case, graph))
    cases = _cases(case, specs)
    ctxs = []
    for s, U, T in cases:
        ctx = s.context()
        ctx.upload(U, T)
        ctx.initialize()
        ctxs.append(ctx)
    ens = xf.Ensemble(ctxs)
    assert ens.cells == sum(s.desc.ninner for s, _, _ in cases)
    ens.steps(nsteps, graph=graph)
    ens.synchronize()
    got = [c.download(0) for c in ctxs]
    ens.close()
    for c in ctxs:
        c.close()
    for (s, U, T), (Ug, Tg) in zip(cases, got):
        Ua, Ta = _alone(s, U, T, nsteps)
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
