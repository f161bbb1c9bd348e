"""The oracle's error path and clock (CPU).

The reference validates every UpdateStates (EstimateYiKernel: code 1 for a
negative / NaN / Inf density or a NaN / Inf mass fraction; then
EstimatePrimitiveVarKernel: code 2 for a negative / NaN / Inf rho, p or T;
Estimate_kernels.hpp:5-162) and runs EstimateNAN (code 3: U[0] < 0 or NaN /
Inf in U or LU; src/Fluids.cpp:47-87) before each RK update, on U1 at stages
1 and 2 and on U at stage 3 and after the Reaction (Fluids.cpp:963-990).  The
first failure stops the time loop (src/XFLUIDS.cpp:211-225, 285-287,
377-389).  A bad value planted in one interior cell must come back as that
cell, with the code, step and stage the reference's ordering gives.

The clock: Evolution clamps dt so that physicalTime lands on the output time
(src/XFLUIDS.cpp:196-198) and stops when it is reached (:172).
"""
import numpy as np
import pytest

from parity import make, oracle_for


def _plant(case, comp, val, cell, buf="U", nsteps=3):
    s, U, T, _ = make(case)
    d = s.desc
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    i, j, k = cell
    idc = (k * d.Ymax + j) * d.Xmax + i
    o.field(buf)[idc * d.E + comp] = val
    st, _ = o.steps(nsteps)
    return st, o.last_error()


# (case, component, value, cell, buffer) -> (code, stage)
PLANTED = [
    # a negative density: EstimateYiKernel in stage 1's UpdateStates
    ("vortex", 0, -1.0, (10, 12, 0), "U", 1, 1),
    # zero total energy: negative pressure (p = (gamma-1) rho e)
    ("vortex", 4, 0.0, (10, 12, 0), "U", 2, 1),
    # the stage-1 NaN check reads U1, the previous step's second-stage state
    ("vortex", 0, -1.0, (10, 12, 0), "U1", 3, 1),
    # a NaN partial density: NaN mass fraction
    ("shock_bubble_z_janaf", 5, np.nan, (20, 12, 0), "U", 1, 1),
    # a NaN species in the reactive tube: the Reaction's UpdateStates (Strang,
    # before the RK stages) sees it first
    ("reactive_shock_tube", 5, np.nan, (100, 0, 0), "U", 1, 4),
    # a negative partial density: the mass fraction stays finite, the
    # temperature inversion does not survive it
    ("reactive_shock_tube", 5, -1.0, (100, 0, 0), "U", 2, 4),
    # NaN transverse momentum in 1-D: the states stay finite (the damped
    # Newton's min/max drop the NaN), the chemistry runs, and the NaN check
    # after the Reaction (flag 4) finds it
    ("reactive_shock_tube", 2, np.nan, (100, 0, 0), "U", 3, 4),
]


@pytest.mark.parametrize("case,comp,val,cell,buf,code,stage", PLANTED)
def test_planted_cell_reported(case, comp, val, cell, buf, code, stage):
    st, (c, ijk, step, where) = _plant(case, comp, val, cell, buf)
    assert st == code and c == code
    assert ijk == cell
    assert step == 1 and where == stage


def test_nan_flux_divergence_reports_stencil_edge():
    # NaN y momentum in a 2-D mixture: the states pass, the face fluxes of
    # every face whose stencil holds the cell become NaN, and stage 1's
    # NaN check reports the lowest such cell (three cells below in y)
    st, (c, ijk, step, where) = _plant("shock_bubble_z_janaf", 2, np.nan, (20, 12, 0))
    assert (st, c, step, where) == (3, 3, 1, 1)
    assert ijk == (20, 9, 0)


def test_clock_lands_on_target():
    s, U, T, _ = make("vortex")
    free = oracle_for(s, U, T)
    assert free.initialize() == 0
    _, dts = free.steps(4)
    target = dts[0] + dts[1] + 0.5 * dts[2]
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    o.set_target_time(target)
    st, got = o.steps(6)
    assert st == 0
    t, step, reached = o.time()
    # two free steps, a third clamped to land on the target, then nothing
    assert step == 3 and reached
    np.testing.assert_array_equal(got[:2], dts[:2])
    assert got[2] == target - (dts[0] + dts[1])
    assert t == target
    assert np.all(got[3:] == 0.0)
    # a later target continues the loop from there
    o.set_target_time(np.inf)
    st, more = o.steps(1)
    assert st == 0 and o.time()[1] == 4 and more[0] > 0


def test_clock_restarts_from_checkpoint_header():
    s, U, T, _ = make("vortex")
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    o.set_time(1.25, 40)
    o.set_target_time(1.25)
    st, _ = o.steps(2)
    assert st == 0 and o.time() == (1.25, 40, True)


def test_shock_tube_fixture_is_post_reflection():
    """The committed 170 us state of the 1024-cell tube (made by the oracle,
    tests/golden/make_shock_tube_fixture.py): the reflected shock has heated
    the gas at the left wall to about 1200 K, and the oracle continues from
    it without error."""
    import os
    import xfluids_amd as xf
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shock_tube_1024_170us.npz"))
    s = xf.Setup(xf.case_path("reactive-shock-tube"), nx=int(z["nx"]))
    d = s.desc
    T = z["T"].reshape(-1)[d.Bwidth_X:d.Bwidth_X + d.X_inner]
    assert abs(float(z["time"]) - 170e-6) < 1e-12
    assert 1150.0 < T[:64].max() < 1300.0      # hot gas behind the reflected shock
    o = oracle_for(s, z["U"], z["T"])
    assert o.initialize() == 0
    o.set_time(float(z["time"]), int(z["steps"]))
    st, _ = o.steps(2)
    assert st == 0 and o.time()[1] == int(z["steps"]) + 2
