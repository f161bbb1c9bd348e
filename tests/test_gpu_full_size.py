"""Size-independent properties at the workloads' full sizes, where the oracle
would take minutes per step.

* Shift invariance on periodic axes.  Every cell and face runs the same
  arithmetic wherever it sits, and the step's reductions are maxima and
  minima, which do not depend on order.  So a state rolled by whole cells
  along its periodic axes gives the rolled result, bit for bit.  This checks
  the tile, block and XCD index maps of every kernel at the real grid.
  - The isentropic vortex at 256 x 256 (periodic in x and y) runs its 200
    steps.
  - The mixing layer at 256^3 (periodic in x and z; nine species, viscous,
    inert step) runs three steps.
  Both forms.
* The reacting mixing layer at 256^3 is shift-invariant to rounding only.
  The chemistry writes T in interior cells alone, as the reference's does.
  So a periodic ghost cell keeps a Newton start that differs from its
  partner's in the last bits, and the faces next to a periodic boundary see
  it (reproduced on the oracle in test_oracle_shift.py).
* Conservation and the exact solution.  On the periodic vortex, the interior
  sums of rho, rho u, rho v and E stay put to rounding, and the density
  matches the advected exact vortex to the scheme's truncation error.
* Drivers.  The shock-bubble interaction at 1024 x 512 gives the same bits
  through the hipGraph step and the stage-by-stage driver over its 500
  steps.

The ghost layers of the starting state are filled on the host with the
periodic copies.  So a ghost cell's temperature (the Newton start of its own
state) is the one of the interior cell it copies, in the rolled run as in
the original one.
"""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce, shifted_pair, unshift_compare

pytestmark = pytest.mark.gpu

FORMS = [xf.XF_PLAIN_FP64, xf.XF_MIXED]


def _run(s, U, T, nsteps, precision):
    ctx = s.context(precision=precision)
    ctx.upload(U, T)
    ctx.initialize()
    ctx.steps(nsteps, graph=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    return Ug, Tg


def _shift_runs(s, U0, T0, which, shifts, nsteps, precision):
    (Ua, Ta), (Ub, Tb) = shifted_pair(s, U0, T0, which, shifts)
    Ra, RTa = _run(s, Ua, Ta, nsteps, precision)
    Rb, RTb = _run(s, Ub, Tb, nsteps, precision)
    assert np.all(np.isfinite(Rb))
    nU, relU, moved = unshift_compare(s, which, shifts, Ra, Rb, Ub, s.desc.E)
    nT, relT, _ = unshift_compare(s, which, shifts, RTa, RTb, Tb)
    print("differing values: U %d, T %d; max rel U %.3g, T %.3g" % (nU, nT, relU.max(), relT.max()))
    assert moved, "the state did not advance"
    return nU, relU, nT, relT


@pytest.mark.parametrize("precision", FORMS)
def test_vortex_256_shift_invariance(precision):
    announce("full size: vortex 256^2 shift invariance, precision %d" % precision)
    s = xf.Setup(xf.case_path("isentropic-vortex"), nx=256, ny=256)
    U0, T0 = s.initial("isentropic-vortex")
    nU, _, nT, _ = _shift_runs(s, U0, T0, "xy", (37, 101), 200, precision)
    assert nU == 0 and nT == 0


@pytest.mark.parametrize("precision", FORMS)
def test_mixing_layer_256_shift_invariance(precision):
    announce("full size: inert mixing layer 256^3 shift invariance, precision %d" % precision)
    s = xf.Setup(xf.case_path("mixing-layer"), nx=256, ny=256, nz=256, reaction=False)
    U0, T0 = s.initial("mixing-layer", seed=11)
    nU, _, nT, _ = _shift_runs(s, U0, T0, "xz", (77, 150), 3, precision)
    assert nU == 0 and nT == 0


def test_reacting_mixing_layer_256_shift_to_rounding():
    """Three reacting steps.  A misplaced index would move values by O(1).
    The ghost cells' own Newton starts move rho, the momenta, E and T by
    about 1e-13 (measured 2.5e-13).  The partial densities of the radicals
    move by up to 1e-8 of their own largest values (measured 9.8e-9): the
    stiff chemistry amplifies the difference, and a last-bit change of T can
    also flip one of CHEMEQ2's step-size decisions.  They are bounded here
    against rho, like mass fractions."""
    announce("full size: reacting mixing layer 256^3 shift invariance to rounding")
    s = xf.Setup(xf.case_path("mixing-layer"), nx=256, ny=256, nz=256)
    U0, T0 = s.initial("mixing-layer", seed=11)
    _, relU, _, relT = _shift_runs(s, U0, T0, "xz", (77, 150), 3, xf.XF_PLAIN_FP64)
    assert relU[:5].max() <= 1e-11 and relT.max() <= 1e-11, (relU, relT)
    assert relU[5:].max() <= 1e-9, relU


@pytest.mark.parametrize("precision", FORMS)
def test_vortex_256_conservation_and_exact_solution(precision):
    """200 steps: interior sums conserved to 1e-12 of their largest, and the
    L1 distance of rho from the exact vortex (the initial one advected by
    (t, t) modulo the 10 x 10 domain) at most 1e-6 (WENO5 at 256^2 reaches
    about 3e-8)."""
    announce("full size: vortex 256^2 conservation, precision %d" % precision)
    n = 256
    s = xf.Setup(xf.case_path("isentropic-vortex"), nx=n, ny=n)
    d = s.desc
    U0, T0 = s.initial("isentropic-vortex")
    shp = (d.Zmax, d.Ymax, d.Xmax, d.E)
    sl = (0, slice(d.Bwidth_Y, d.Bwidth_Y + d.Y_inner), slice(d.Bwidth_X, d.Bwidth_X + d.X_inner))
    inner = lambda A: np.asarray(A).reshape(shp)[sl]
    sums0 = inner(U0).sum(axis=(0, 1))
    ctx = s.context(precision=precision)
    ctx.upload(U0, T0)
    ctx.initialize()
    dts = ctx.steps(200, graph=True, record_dt=True)
    Ug, _ = ctx.download(0)
    ctx.close()
    Ui = inner(Ug)
    drift = np.abs(Ui.sum(axis=(0, 1)) - sums0) / np.abs(sums0).max()
    assert drift[[0, 1, 2, 4]].max() <= 1e-12, drift
    t = float(np.sum(dts))
    x = (np.arange(n) + 0.5) * (10.0 / n)
    X, Y = np.meshgrid(x, x)
    xv = np.mod(X - t, 10.0) - 5.0
    yv = np.mod(Y - t, 10.0) - 5.0
    rho_ex = (1.0 - 25.0 * np.exp(1.0 - (xv ** 2 + yv ** 2)) / (28.0 * np.pi ** 2)) ** 2.5
    l1 = float(np.abs(Ui[:, :, 0] - rho_ex).mean())
    assert l1 <= 1e-6, l1


def test_shock_bubble_1024_graph_equals_staged():
    announce("full size: shock-bubble 1024x512, 500 steps, graph vs staged")
    s = xf.Setup(xf.case_path("shock-bubble"), nx=1024, ny=512)
    U0, T0 = s.initial("shock-bubble")
    out = []
    for staged in (False, True):
        ctx = s.context()
        ctx.upload(U0, T0)
        ctx.initialize()
        if staged:
            ctx.evolve_staged(500)
        else:
            ctx.steps(500, graph=True)
        out.append(ctx.download(0))
        ctx.close()
    (Ua, Ta), (Ub, Tb) = out
    assert np.all(np.isfinite(Ua))
    assert np.array_equal(Ua.view(np.uint64), Ub.view(np.uint64))
    assert np.array_equal(Ta.view(np.uint64), Tb.view(np.uint64))
