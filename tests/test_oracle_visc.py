"""Pins of the viscous terms: the host's transport fits
(xfluids_amd/host/xf_transport.cpp) and the oracle's transport coefficients,
centre derivatives and wall fluxes (oracle/ora_visc.c).

* Kinetic theory behind the fits against published measurements: the
  viscosity and conductivity of N2, O2 and H2 at 300 K and 1000 K and the
  binary diffusion coefficients of H2-N2 and O2-N2 at 298 K, 1 atm (NIST
  Chemistry WebBook thermophysical data; Marrero & Mason, J. Phys. Chem.
  Ref. Data 1, 1972), within 2-5 %; the fits reproduce the kinetic-theory
  values at the fitting temperatures up to 2500 K within 2 %.
* Mixture rules: the oracle's cell viscosity is Wilke's rule (Bird, Stewart
  & Lightfoot, eq. 1.4-15/16) and its conductivity the same weights on the
  species conductivities, its D_km the mixture-averaged (1 - Y_k) /
  sum_j X_j / D_kj (Kee, Coltrin & Glarborg, eq. 12.178), evaluated here in
  numpy on the same fits; a pure gas gives the species' own values.
* Discrete operators: the fourth-order centre derivatives are exact on
  quartic velocity fields; for a velocity linear in one coordinate the wall
  flux is the Newtonian stress exactly: mu a for shear (directly, or through
  the centre derivatives for the cross term), 4/3 mu a for normal strain,
  and the stress times the face velocity in the energy.
* Navier-Stokes solutions: a periodic shear wave u = sin(ky) decays at
  nu k^2 within 0.5 %, and an isobaric temperature wave at kappa k^2 /
  (rho c_p) within 1 %.
* The single-process diffusion limits (Dim_max reduced into 0) clamp every
  species flux to zero: Visc_Diffu changes no result.
"""
import ctypes as C
import json

import numpy as np
import pytest

import xfluids_amd as xf
from oracle_lib import Oracle, config_from_setup, lib as olib

DP = C.POINTER(C.c_double)


def _fit(a, T):
    v = a[3]
    for i in (2, 1, 0):
        v = v * np.log(T) + a[i]
    return np.exp(v)


@pytest.fixture(scope="module")
def air():
    return xf.Setup(xf.case_path("shock-bubble"), nx=16, ny=8, viscous=(1, 1, 1))


@pytest.mark.parametrize("name,T,mu,kappa", [
    ("N2", 300.0, 17.89e-6, 0.02598), ("N2", 1000.0, 41.55e-6, 0.0666),
    ("O2", 300.0, 20.65e-6, 0.02633), ("H2", 300.0, 8.96e-6, 0.1868)])
def test_kinetic_theory_matches_measurements(air, name, T, mu, kappa):
    # measured: N2 1.1 %, 0.2 % / 1.4 %, 3.3 %; O2 0.0 % / 0.8 %; H2 0.4 % / 0.1 %
    k = air.species.index(name)
    assert abs(air.transport_exact(0, k, k, T) / mu - 1) < 0.02
    assert abs(air.transport_exact(1, k, k, T, 101325.0) / kappa - 1) < 0.05


@pytest.mark.parametrize("a,b,D", [("H2", "N2", 0.784), ("O2", "N2", 0.202)])
def test_binary_diffusion_matches_measurements(air, a, b, D):
    k, j = air.species.index(a), air.species.index(b)
    assert abs(air.transport_exact(2, k, j, 298.0, 101325.0) / D - 1) < 0.04


def test_fits_reproduce_kinetic_theory_at_the_nodes(air):
    # 3000 K is left out: there H2's reduced temperature 3000/38 lies in
    # (T*[35], T*[36]) = (75, 100), where the reference's interpolation reads
    # one row past its table (kept as it computes, +7 % in H2's viscosity at
    # that node; the cubic fit then sits 5 % below it)
    fcv, fct, Dkj = air.transport_tables()
    NS = air.desc.NS
    for T in (273.15, 500.0, 1000.0, 1500.0, 2000.0, 2500.0):
        for k in range(NS):
            assert abs(_fit(fcv[4 * k:4 * k + 4], T) / air.transport_exact(0, k, k, T) - 1) < 0.02
            assert abs(_fit(fct[4 * k:4 * k + 4], T) / air.transport_exact(1, k, k, T, 1.0) - 1) < 0.02
            for j in range(NS):
                q = 4 * (k * NS + j)
                assert abs(_fit(Dkj[q:q + 4], T) / air.transport_exact(2, k, j, T, 1.0) - 1) < 0.02
                assert np.array_equal(Dkj[q:q + 4], Dkj[4 * (j * NS + k):4 * (j * NS + k) + 4])


def _uniform_lu(s, Y, T0=300.0, p0=101325.0, vel=None):
    """Oracle after initialize + get_dt + compute_lu on a uniform state (inflow
    boundaries keep the ghost cells as given)."""
    d = s.desc
    cfg = config_from_setup(s)
    for q in range(6):
        cfg.BC[q] = 0
    o = Oracle(cfg)
    L = olib()
    yy = np.ascontiguousarray(Y, dtype=np.float64)
    R = L.ora_mix_R(C.byref(cfg), yy.ctypes.data_as(DP))
    e = L.ora_mix_e(C.byref(cfg), yy.ctypes.data_as(DP), T0)
    n = d.ncell
    rho = p0 / (R * T0)
    U = np.zeros((n, d.E))
    U[:, 0] = rho
    ke = np.zeros(n)
    if vel is not None:
        for c in range(3):
            U[:, 1 + c] = rho * vel[c]
            ke += 0.5 * rho * vel[c] ** 2
    U[:, 4] = rho * e + ke
    U[:, 5:] = rho * yy[None, :-1]
    o.set_state(U.ravel(), np.full(n, T0))
    assert o.initialize() == 0
    o.get_dt()
    o.compute_lu(0)
    return o, cfg, rho


@pytest.mark.parametrize("T", [300.0, 1500.0])
def test_mixture_rules_are_wilke_and_mixture_averaged(air, T):
    d = air.desc
    NS = d.NS
    _, _, W = air.thermo_tables()
    fcv, fct, Dkj = air.transport_tables()
    Y = np.array([0.02, 0.23, 0.75])
    o, cfg, rho = _uniform_lu(air, Y, T0=T)
    p = o.field("p")[0]
    X = (Y / W) / (Y / W).sum()
    mu = np.array([_fit(fcv[4 * k:4 * k + 4], T) for k in range(NS)])
    lam = np.array([_fit(fct[4 * k:4 * k + 4], T) for k in range(NS)])
    phi = np.array([[(1 + np.sqrt(mu[k] / mu[j]) * (W[j] / W[k]) ** 0.25) ** 2 / np.sqrt(8 * (1 + W[k] / W[j]))
                     for j in range(NS)] for k in range(NS)])
    den = phi @ X
    assert abs(o.field("va")[0] / np.sum(X * mu / den) - 1) < 1e-13
    assert abs(o.field("tca")[0] / np.sum(X * lam / den) - 1) < 1e-13
    D = np.array([[_fit(Dkj[4 * (k * NS + j):4 * (k * NS + j) + 4], T) / p for j in range(NS)] for k in range(NS)])
    n = d.ncell
    for k in range(NS):
        want = (1 - Y[k]) / sum(X[j] / D[k, j] for j in range(NS) if j != k) * 1e-4   # cm^2/s -> m^2/s
        assert abs(o.field("Dkm")[k * n] / want - 1) < 1e-12


def test_pure_gas_gives_species_values(air):
    fcv, fct, _ = air.transport_tables()
    k = air.species.index("N2")
    o, _, _ = _uniform_lu(air, np.array([0.0, 0.0, 1.0]), T0=800.0)
    assert abs(o.field("va")[0] / _fit(fcv[4 * k:4 * k + 4], 800.0) - 1) < 1e-14
    assert abs(o.field("tca")[0] / _fit(fct[4 * k:4 * k + 4], 800.0) - 1) < 1e-14


def _grid(s):
    d = s.desc
    x = (np.arange(d.Xmax) - d.Bwidth_X + 0.5) * d.dx
    y = (np.arange(d.Ymax) - d.Bwidth_Y + 0.5) * d.dy
    z = (np.arange(d.Zmax) - d.Bwidth_Z + 0.5) * d.dz if d.DimZ else np.zeros(1)
    Z, Yg, X = np.meshgrid(z, y, x, indexing="ij")
    return X.ravel(), Yg.ravel(), Z.ravel()


def test_centre_derivatives_exact_on_quartics():
    s = xf.Setup(xf.case_path("mixing-layer"), nx=10, ny=9, nz=8, viscous=(1, 0, 0))
    d = s.desc
    X, Yc, Z = _grid(s)
    sx, sy, sz = X / X.max(), Yc / Yc.max(), Z / Z.max()
    u = 3.0 * sx ** 4 - 2.0 * sy ** 3 + sx * sy * sz
    v = sy ** 4 + 0.5 * sz ** 2 - sx
    w = sz ** 3 * sx - sy
    Y = np.array(s.mass_fractions()[0])
    o, _, _ = _uniform_lu(s, Y, vel=(u, v, w))
    n = d.ncell
    Vde = o.field("Vde").reshape(9, n)
    Lx, Ly, Lz = X.max(), Yc.max(), Z.max()
    want = {0: 12 * sx ** 3 / Lx + sy * sz / Lx, 1: -1 / Lx, 2: sz ** 3 / Lx,
            3: -6 * sy ** 2 / Ly + sx * sz / Ly, 4: 4 * sy ** 3 / Ly, 5: -1 / Ly,
            6: sx * sy / Lz, 7: sz / Lz, 8: 3 * sz ** 2 * sx / Lz}
    inner = np.zeros((d.Zmax, d.Ymax, d.Xmax), bool)
    inner[d.Bwidth_Z:-d.Bwidth_Z, d.Bwidth_Y:-d.Bwidth_Y, d.Bwidth_X:-d.Bwidth_X] = True
    inner = inner.ravel()
    for q, wq in want.items():
        wq = np.broadcast_to(wq, (n,))
        scale = np.abs(wq[inner]).max()
        assert np.abs(Vde[q][inner] - wq[inner]).max() <= 1e-9 * scale, q


@pytest.mark.parametrize("kind", ["shear", "strain", "cross"])
@pytest.mark.parametrize("d_face", [0, 1, 2])
def test_wall_flux_of_linear_velocity_is_exact(d_face, kind):
    """Uniform rho and T, one velocity component linear in one coordinate
    (d the face normal, t = d + 1 mod 3):
    shear  u_t = 1 + a x_d: flux mu a in momentum t, mu a u_t(face) in energy;
    strain u_d = 1 + a x_d: (2 mu + lambda) a = 4/3 mu a in momentum d and
           4/3 mu a u_d(face) in energy;
    cross  u_d = 1 + a x_t: mu a in momentum t (through the centre
           derivatives interpolated to the face), nothing in energy.
    Every other component, the mass and the species carry no wall flux."""
    s = xf.Setup(xf.case_path("mixing-layer"), nx=8, ny=8, nz=8, viscous=(1, 0, 0))
    d = s.desc
    grid = _grid(s)
    t = (d_face + 1) % 3
    comp, along = {"shear": (t, d_face), "strain": (d_face, d_face), "cross": (d_face, t)}[kind]
    coord = grid[along]
    a = 100.0 / np.abs(coord).max()
    vel = [np.zeros(d.ncell) for _ in range(3)]
    vel[comp] = 1.0 + a * coord
    Y = np.array(s.mass_fractions()[0])
    fw = ("FluxFw", "FluxGw", "FluxHw")[d_face]
    on, cfg, rho = _uniform_lu(s, Y, vel=vel)
    s_off = xf.Setup(xf.case_path("mixing-layer"), nx=8, ny=8, nz=8, viscous=(0, 0, 0))
    off, _, _ = _uniform_lu(s_off, Y, vel=vel)
    E = d.E
    F = (off.field(fw) - on.field(fw)).reshape(-1, E)
    mu = on.field("va")[0]
    h = (d.dx, d.dy, d.dz)[d_face]
    inner = np.zeros((d.Zmax, d.Ymax, d.Xmax), bool)
    inner[d.Bwidth_Z:-d.Bwidth_Z, d.Bwidth_Y:-d.Bwidth_Y, d.Bwidth_X:-d.Bwidth_X] = True
    if kind == "cross":
        # the inflow boundary zeroes the ghost-cell centre derivatives the
        # two faces next to each boundary read
        sl = [slice(None)] * 3
        for cut in (slice(0, d.Bwidth_X + 2), slice(-(d.Bwidth_X + 2), None)):
            sl[2 - d_face] = cut
            inner[tuple(sl)] = False
    inner = inner.ravel()
    face = grid[along] + (0.5 * h if along == d_face else 0.0)
    u_face = 1.0 + a * face[inner]
    mom = {"shear": (1 + t, mu * a), "strain": (1 + d_face, 4.0 / 3.0 * mu * a), "cross": (1 + t, mu * a)}[kind]
    energy = {"shear": mu * a * u_face, "strain": 4.0 / 3.0 * mu * a * u_face, "cross": 0.0 * u_face}[kind]
    # the Newton temperatures differ from 300 K by ~1e-7, so mu is uniform to ~1e-9
    tol = 1e-7 * abs(mom[1])
    assert np.abs(F[inner, mom[0]] - mom[1]).max() <= tol
    assert np.abs(F[inner, 4] - energy).max() <= tol * np.abs(u_face).max()
    for q in (1, 2, 3):
        if q != mom[0]:
            assert np.abs(F[inner, q]).max() <= tol, q
    assert np.all(F[inner, 5:] == 0.0) and np.all(F[inner, 0] == 0.0)


def _wave(tmp_path, mode, n, L, nsteps, nx):
    js = {"run": {"CFLnumber": 0.4},
          "equations": {"Mixture": "H2-O2-N2", "Thermo": "JANAF", "Scheme": "WENO5Z", "Sources_React": False,
                        "Visc": True, "Visc_Heat": mode == "heat"},
          "mesh": {"DOMAIN_Size": [L * nx / n, L, 0.0], "Resolution": [nx, n, 0], "Boundarys": [3, 3, 3, 3, 3, 3]}}
    path = tmp_path / ("wave_%s.json" % mode)
    path.write_text(json.dumps(js))
    s = xf.Setup(str(path))
    d = s.desc
    cfg = config_from_setup(s)
    o = Oracle(cfg)
    B = d.Bwidth_Y
    y = np.repeat((np.arange(d.Ymax) - B + 0.5) * (L / n), d.Xmax)
    k = 2 * np.pi / L
    Y = np.ascontiguousarray(np.array(s.mass_fractions()[0]))
    Lb = olib()
    R = Lb.ora_mix_R(C.byref(cfg), Y.ctypes.data_as(DP))
    if mode == "shear":
        T, u = np.full(y.size, 300.0), np.sin(k * y)
    else:
        T, u = 300.0 + np.sin(k * y), np.zeros(y.size)
    rho = 101325.0 / (R * T)
    U = np.zeros((y.size, d.E))
    U[:, 0], U[:, 1] = rho, rho * u
    U[:, 4] = rho * np.array([Lb.ora_mix_e(C.byref(cfg), Y.ctypes.data_as(DP), t) for t in T]) + 0.5 * rho * u * u
    U[:, 5:] = rho[:, None] * Y[None, :-1]
    inner = np.zeros((d.Ymax, d.Xmax), bool)
    inner[B:B + n, d.Bwidth_X:d.Bwidth_X + d.X_inner] = True
    inner = inner.ravel()

    def amp():
        if mode == "shear":
            q = o.field("U").reshape(-1, d.E)
            val = q[inner, 1] / q[inner, 0]
        else:
            val = o.field("T")[inner] - 300.0
        return 2.0 / inner.sum() * np.sum(val * np.sin(k * y[inner]))

    o.set_state(U.ravel(), T.copy())
    assert o.initialize() == 0
    a0 = amp()
    o.get_dt()
    o.compute_lu(0)
    mu, kap = o.field("va")[inner].mean(), o.field("tca")[inner].mean()
    cp = Lb.ora_mix_cp(C.byref(cfg), Y.ctypes.data_as(DP), 300.0)
    o.set_state(U.ravel(), T.copy())
    assert o.initialize() == 0
    st, dts = o.steps(nsteps)
    assert st == 0
    rate = -np.log(amp() / a0) / dts.sum()
    diffusivity = mu / rho[inner].mean() if mode == "shear" else kap / (rho[inner].mean() * cp)
    return rate / (diffusivity * k * k)


def test_shear_wave_decays_at_viscous_rate(tmp_path):
    # 2-D (the x momentum's shear stress carries DimX_t): measured 0.9996
    assert abs(_wave(tmp_path, "shear", 64, 1e-4, 1000, 4) - 1) < 5e-3


def test_temperature_wave_decays_at_isobaric_diffusivity(tmp_path):
    # measured 0.9987; at L = 1e-4 the acoustic coupling adds ~1 %
    assert abs(_wave(tmp_path, "heat", 64, 4e-4, 4000, 0) - 1) < 1e-2


def test_diffusion_limits_clamp_species_fluxes():
    out = []
    for diffu in (0, 1):
        s = xf.Setup(xf.case_path("shock-bubble"), nx=48, ny=24, viscous=(1, 1, diffu))
        U, T = s.initial("shock-bubble", seed=3)
        o = Oracle(config_from_setup(s))
        o.set_state(U, T)
        assert o.initialize() == 0
        st, _ = o.steps(3)
        assert st == 0
        out.append(o.field("U").copy())
    assert np.array_equal(out[0], out[1])
