"""Plain-FP64 GPU path vs the CPU oracle on identical seeded inputs.

Bar: within 1e-10 relative (max |dU| / max |U| per conserved component, and
T) after several full time steps.  The plain form is built to be bit-identical
(the reference's operation order and rounding, no FMA contraction, the same
exp/log); the tests print the bit-exact fraction and require it for the
fluid-only cases.  The fused FP64 form is held to the same bar against the
oracle's fused variant (cfg.fused = 1), and its distance from the reference's
rounding is printed.
"""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import CASES, announce, compare, make, oracle_for

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _run_gpu(s, U, T, nsteps, mode, precision=xf.XF_PLAIN_FP64):
    ctx = s.context(precision=precision)
    ctx.upload(U, T)
    ctx.initialize()
    if mode == "staged":
        dts = ctx.evolve_staged(nsteps)
    else:
        dts = ctx.steps(nsteps, graph=(mode == "graph"), record_dt=True)
    Ug, Tg = ctx.download(0)
    ctx.close()
    return Ug, Tg, dts


def test_device_exp_log_bitwise():
    from oracle_lib import lib as olib
    announce("device exp/log vs oracle exp/log")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-700, 700, 20000), rng.uniform(-5, 5, 20000),
                        np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(100, 6000, 20000),
                        [1.0, 0.5, 2.0, 1e-300, 5e-324, 709.7, -745.0, 0.0]])
    eg, lg = xf.debug_math(x)
    L = olib()
    eo = np.array([L.ora_exp(v) for v in x])
    lo = np.array([L.ora_log(v) if v > 0 else np.nan for v in x])
    assert np.array_equal(eg.view(np.uint64), eo.view(np.uint64))
    pos = x > 0
    assert np.array_equal(lg[pos].view(np.uint64), lo[pos].view(np.uint64))


@pytest.mark.parametrize("case", list(CASES))
def test_steps_match_oracle(case):
    announce("steps_match_oracle " + case)
    s, U, T, nsteps = make(case)
    d = s.desc
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    st, dts_o = o.steps(nsteps)
    assert st == 0, "oracle reported error %d" % st
    Ug, Tg, dts_g = _run_gpu(s, U, T, nsteps, "fused")
    eqU, relU = compare(o.field("U"), Ug, d.E)
    eqT, relT = compare(o.field("T"), Tg)
    print("bitwise U %.6f T %.6f  max rel U %s T %.3e  dt rel %.3e" %
          (eqU, eqT, np.array2string(relU, precision=2), relT[0],
           np.abs(dts_o - dts_g).max() / dts_o.max()))
    assert np.all(relU <= TOL), relU
    if d.cop:
        assert relT[0] <= TOL
    assert np.abs(dts_o - dts_g).max() <= TOL * dts_o.max()
    if not d.react:
        assert eqU == 1.0, "fluid-only plain form should be bit-identical"


FUSED_CASES = ["vortex", "shock_bubble_z_janaf", "shock_bubble_js_nasa", "shock_bubble_3d", "shock_bubble_pp",
               "shock_bubble_visc", "mixing_layer_3d", "reactive_shock_tube"]


@pytest.mark.parametrize("case", FUSED_CASES)
def test_fused_form_matches_fused_oracle(case):
    announce("fused_form_matches_fused_oracle " + case)
    s, U, T, nsteps = make(case)
    d = s.desc
    o = oracle_for(s, U, T, fused=1)
    assert o.initialize() == 0
    st, dts_o = o.steps(nsteps)
    assert st == 0
    Ug, Tg, dts_g = _run_gpu(s, U, T, nsteps, "graph", precision=xf.XF_FUSED_FP64)
    eqU, relU = compare(o.field("U"), Ug, d.E)
    r = oracle_for(s, U, T)
    r.initialize()
    r.steps(nsteps)
    eqR, relR = compare(r.field("U"), Ug, d.E)
    print("fused form vs fused oracle: bitwise U %.6f max rel %s; vs reference rounding: bitwise %.4f max rel %s" %
          (eqU, np.array2string(relU, precision=2), eqR, np.array2string(relR, precision=2)))
    assert np.all(relU <= TOL), relU
    if not d.react:
        assert eqU == 1.0


@pytest.mark.parametrize("mode", ["graph", "staged"])
@pytest.mark.parametrize("case", ["vortex", "shock_bubble_z_janaf"])
def test_step_drivers_agree(case, mode):
    """hipGraph replay and the stage-by-stage seam give the fused result."""
    announce("step_drivers_agree %s %s" % (case, mode))
    s, U, T, nsteps = make(case)
    Uf, Tf, df = _run_gpu(s, U, T, nsteps, "fused")
    Um, Tm, dm = _run_gpu(s, U, T, nsteps, mode)
    assert np.array_equal(Uf.view(np.uint64), Um.view(np.uint64))
    assert np.array_equal(df, dm)


@pytest.mark.parametrize("mode", ["graph", "staged"])
@pytest.mark.parametrize("case", ["inert_sbi_no_fuel", "rsbi"])
def test_ghost_species_drivers_match_oracle(case, mode):
    """GhostSpecies mixtures through the hipGraph and the stage-by-stage
    drivers (the separate ghost fill): bitwise the oracle."""
    announce("ghost species %s %s" % (case, mode))
    s, U, T, nsteps = make(case)
    d = s.desc
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    st, dts_o = o.steps(nsteps)
    assert st == 0
    Ug, Tg, dts_g = _run_gpu(s, U, T, nsteps, mode)
    assert np.array_equal(o.field("U").view(np.uint64), Ug.view(np.uint64))
    assert np.array_equal(o.field("T").view(np.uint64), Tg.view(np.uint64))
    assert np.array_equal(dts_o, dts_g)


def test_ghost_species_seam():
    """One UpdateStates of a GhostSpecies mixture: the stored fractions (ghost
    = 0) and the written-back partial densities rho * Y_k equal the oracle's
    in every padded cell."""
    announce("ghost species seam inert_sbi_3d")
    s, U, T, _ = make("inert_sbi_3d")
    o = oracle_for(s, U, T)
    ctx = s.context()
    ctx.upload(U, T)
    o.boundary(0)
    ctx.boundary(0)
    assert o.update_states(0) == 0
    ctx.update_states(0)
    P = ctx.prims()
    assert np.array_equal(o.field("y").view(np.uint64), P["y"].view(np.uint64))
    assert np.all(P["y"].reshape(-1, s.desc.NS)[:, -1] == 0.0)
    Ug, _ = ctx.download(0)
    assert np.array_equal(o.field("U").view(np.uint64), Ug.view(np.uint64))
    assert not np.array_equal(Ug, U)  # the write-back changed the state
    ctx.close()


def test_stage_seam_matches_oracle():
    """One stage through the C-ABI seam: primitives, dt, LU, RK update."""
    announce("stage seam shock_bubble_js_nasa")
    s, U, T, _ = make("shock_bubble_js_nasa")
    d = s.desc
    o = oracle_for(s, U, T)
    o.initialize()
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    dt_o = o.get_dt()
    dt_g = ctx.get_dt()
    assert dt_o == dt_g
    o.boundary(0)
    ctx.boundary(0)
    assert o.update_states(0) == 0
    ctx.update_states(0)
    P = ctx.prims()
    for name in ("rho", "p", "u", "v", "H", "c", "gamma", "e"):
        eq, rel = compare(o.field(name), P[name])
        assert eq == 1.0, (name, eq, rel)
    eq, _ = compare(o.field("y"), P["y"])
    assert eq == 1.0
    o.compute_lu(0)
    ctx.compute_lu(0)
    LUo = o.field("LU").reshape(d.Zmax, d.Ymax, d.Xmax, d.E)
    LUg = ctx.lu().reshape(d.Zmax, d.Ymax, d.Xmax, d.E)
    sl = (slice(d.Bwidth_Z, d.Zmax - d.Bwidth_Z) if d.DimZ else slice(None),
          slice(d.Bwidth_Y, d.Ymax - d.Bwidth_Y), slice(d.Bwidth_X, d.Xmax - d.Bwidth_X))
    eq, rel = compare(LUo[sl].ravel(), LUg[sl].ravel(), d.E)
    assert eq == 1.0, rel
    eb_o = o.field("eigen_block").reshape(3, -1)
    eb_g, _ = ctx.maxima()
    assert np.array_equal(eb_o[:, :d.E], eb_g[:, :d.E])
    o.rk3(1, dt_o)
    ctx.rk3_stage(1, dt_g)
    ctx.close()


def test_chemistry_substeps_and_energy():
    """CHEMEQ2 on the GPU: same result and same predictor-pass counts."""
    announce("chemistry reactive_channel")
    s, U, T, _ = make("reactive_channel")
    o = oracle_for(s, U, T)
    o.initialize()
    ctx = s.context()
    ctx.upload(U, T)
    ctx.initialize()
    ctx.update_states(0)
    o.update_states(0)
    o.chemistry(2e-7)
    sub = ctx.chemistry(2e-7, record=True)
    Ug, Tg = ctx.download(0)
    d = s.desc
    eqU, relU = compare(o.field("U"), Ug, d.E)
    eqT, relT = compare(o.field("T"), Tg)
    print("chem bitwise U %.6f T %.6f  substeps min %d max %d mean %.1f" % (eqU, eqT, sub.min(), sub.max(), sub.mean()))
    assert np.all(relU <= TOL) and relT[0] <= TOL
    assert sub.min() >= 1
    ctx.close()
