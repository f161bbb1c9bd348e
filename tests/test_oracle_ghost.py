"""GhostSpecies mixtures in the oracle (CPU).

The reference's shock-bubble mixtures (runtime.dat/Inert-SBI,
Inert-SBI-without-fuel, Reaction/RSBI-19REA) list N2 twice and define
GhostSpecies: the last species is a ghost that carries no mass.  Getrhoyi
(src/solver_UpdateStates/Update_device.hpp:15-21) then sets its mass fraction
to 0, renormalises the other NS - 1 fractions to sum 1 and writes rho * Y_k
back into the state, and every later function reads those fractions.
"""
import numpy as np
import pytest

from parity import make, oracle_for


def _states(case):
    s, U, T, _ = make(case)
    d = s.desc
    o = oracle_for(s, U, T)
    o.boundary(0)
    U0 = o.field("U").copy().reshape(-1, d.E)
    assert o.update_states(0) == 0
    # copies: the fields are views of the oracle's buffers, freed with it
    return s, d, U0, o.field("U").copy().reshape(-1, d.E), o.field("y").copy().reshape(-1, d.NS)


@pytest.mark.parametrize("case", ["inert_sbi_3d", "inert_sbi_no_fuel", "rsbi", "mixing_layer_21"])
def test_getrhoyi_ghost_branch(case):
    s, d, U0, U1, y = _states(case)
    assert d.ghost == 1 and s.species[-1] == s.species[-2] == "N2"
    rho = U1[:, 0]
    assert np.array_equal(rho, U0[:, 0])
    # the ghost carries nothing
    assert np.all(y[:, -1] == 0.0)
    # the others sum to 1 within rounding
    assert np.abs(y[:, :-1].sum(axis=1) - 1.0).max() < 8 * np.finfo(float).eps
    # in the proportions of the state's partial densities
    r0 = U0[:, 5:] / U0[:, 5:].sum(axis=1)[:, None]
    assert np.abs(y[:, :-1] - r0).max() < 8 * np.finfo(float).eps
    # and written back as rho * Y_k, each one rounding
    assert np.array_equal(U1[:, 5:].view(np.uint64), (rho[:, None] * y[:, :-1]).view(np.uint64))
    # other components untouched
    assert np.array_equal(U1[:, :5], U0[:, :5])


@pytest.mark.parametrize("case", ["inert_sbi_no_fuel", "rsbi"])
def test_ghost_stays_empty_over_steps(case):
    """Over full steps (RK stages, limiter, chemistry where present) the
    ghost's fraction stays 0 after every UpdateStates and the partial
    densities of the others keep summing to the density."""
    s, U, T, nsteps = make(case)
    d = s.desc
    o = oracle_for(s, U, T)
    assert o.initialize() == 0
    st, _ = o.steps(nsteps)
    assert st == 0
    o.boundary(0)
    assert o.update_states(0) == 0
    Uo = o.field("U").reshape(-1, d.E)
    y = o.field("y").reshape(-1, d.NS)
    assert np.all(y[:, -1] == 0.0)
    assert np.abs(Uo[:, 5:].sum(axis=1) / Uo[:, 0] - 1.0).max() < 1e-14
    # WENO undershoots at the bubble edge stay small (the limiter bounds rho)
    assert y.min() > -1e-6 and y.max() < 1.0 + 1e-6
