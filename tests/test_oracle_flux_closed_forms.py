"""Closed-form pins of the 3D two-point fluxes that no reference test covers
directly: flux_kennedy_gruber (src/equations/compressible_euler_energy_with_gravity_3d.jl:289-318),
flux_shima_etal (:248-276), flux_ranocha (:382-415) and the gravity terms
flux_nonconservative_souza_etal (:206-230) / waruszewski_etal (:135-159).

On states with one velocity v and one pressure p but different densities the
averages in these fluxes reduce to closed forms that tell the fluxes apart:
- Kennedy-Gruber averages the specific energy e = rho_e / rho, so its energy
  row carries the harmonic-type factor rho_avg (1/rho_l + 1/rho_r) / 2 =
  (rho_l + rho_r)^2 / (4 rho_l rho_r) on the internal energy; averaging rho_e
  instead (a plausible slip) drops that factor;
- Shima et al. carry the internal energy as p_avg v_n / (gamma - 1), so their
  energy flux is v_n (rho_avg |v|^2 / 2 + gamma p / (gamma - 1)) whatever the
  densities (pressure equilibrium);
- Ranocha's flux reduces to v_n (ln_mean(rho) |v|^2 / 2 + gamma p / (gamma - 1));
- both gravity terms have the energy entry v_avg . (momentum entries), with the
  arithmetic (Souza) or logarithmic (Waruszewski) mean density.
The discrete consequence is checked on rhs!: on a periodic mesh a state of
uniform velocity and pressure keeps both (dv/dt = dp/dt = 0 to round-off) with
the Shima or Ranocha volume flux and Ranocha surface flux, and does not with
the Kennedy-Gruber volume flux.
"""
import numpy as np
import pytest

import trixiatmo_hip as T
from oracle_lib import Oracle, params_for

P = Oracle.params(4, 1, c_p=1004.0, c_v=717.0)
GAMMA = 1004.0 / 717.0
KG, RANOCHA, SHIMA = 1, 0, 2
WARUSZEWSKI, SOUZA = 1, 2


def _pair_states(seed, k=25):
    """pairs (u_l, u_r, v, p, n) with one velocity and pressure, different
    densities (ratios from 0.3 to 3) and phi of either sign"""
    rng = np.random.default_rng(seed)
    for _ in range(k):
        v = rng.normal(size=3) * rng.uniform(0.1, 30)
        p = rng.uniform(0.5, 2.0) * 10 ** rng.uniform(-1, 5)
        rl = rng.uniform(0.2, 2.0)
        rr = rl * rng.uniform(0.3, 3.0)
        n = rng.normal(size=3)
        ul = np.array([rl, *(rl * v), p / (GAMMA - 1) + 0.5 * rl * v @ v, rng.uniform(-1e3, 1e3)])
        ur = np.array([rr, *(rr * v), p / (GAMMA - 1) + 0.5 * rr * v @ v, rng.uniform(-1e3, 1e3)])
        yield ul, ur, v, p, n


def _close(a, b, scale):
    return abs(a - b) <= 1e-13 * scale


def test_kennedy_gruber_uniform_velocity_and_pressure(oracle):
    for ul, ur, v, p, n in _pair_states(1):
        f = oracle.volume_flux(P, KG, ul, ur, n)
        rl, rr = ul[0], ur[0]
        ra, vn = 0.5 * (rl + rr), v @ n
        assert _close(f[0], ra * vn, abs(ra * vn) + 1e-300)
        for c in range(3):
            ref = ra * vn * v[c] + p * n[c]
            assert _close(f[1 + c], ref, abs(ra * vn * v[c]) + abs(p * n[c]))
        # e_avg = p / (gamma - 1) (1/rho_l + 1/rho_r) / 2 + |v|^2 / 2
        hm = (rl + rr) ** 2 / (4 * rl * rr)
        ref5 = vn * (p / (GAMMA - 1) * hm + 0.5 * ra * v @ v + p)
        scale = abs(vn) * (p / (GAMMA - 1) * hm + 0.5 * ra * v @ v + p)
        assert _close(f[4], ref5, scale), (f[4], ref5)
        # the average of rho_e (instead of e) gives a different energy flux
        wrong5 = vn * (p / (GAMMA - 1) + 0.5 * ra * v @ v + p)
        assert abs(f[4] - wrong5) > 1e-6 * scale * (hm - 1) or hm - 1 < 1e-3


def test_shima_uniform_velocity_and_pressure(oracle):
    for ul, ur, v, p, n in _pair_states(2):
        f = oracle.volume_flux(P, SHIMA, ul, ur, n)
        rl, rr = ul[0], ur[0]
        ra, vn = 0.5 * (rl + rr), v @ n
        assert _close(f[0], ra * vn, abs(ra * vn) + 1e-300)
        for c in range(3):
            assert _close(f[1 + c], ra * vn * v[c] + p * n[c], abs(ra * vn * v[c]) + abs(p * n[c]))
        ref5 = vn * (0.5 * ra * v @ v + GAMMA * p / (GAMMA - 1))
        assert _close(f[4], ref5, abs(ref5)), (f[4], ref5)


def test_shima_and_kennedy_gruber_differ_by_the_density_ratio(oracle):
    """f5(KG) - f5(Shima) = v_n p / (gamma - 1) (rho_l - rho_r)^2 / (4 rho_l rho_r)"""
    for ul, ur, v, p, n in _pair_states(3):
        d = oracle.volume_flux(P, KG, ul, ur, n)[4] - oracle.volume_flux(P, SHIMA, ul, ur, n)[4]
        rl, rr = ul[0], ur[0]
        ref = (v @ n) * p / (GAMMA - 1) * (rl - rr) ** 2 / (4 * rl * rr)
        scale = abs(v @ n) * (GAMMA * p / (GAMMA - 1) * (rl + rr) ** 2 / (4 * rl * rr) +
                              (rl + rr) * (v @ v))
        assert abs(d - ref) <= 1e-13 * scale, (d, ref)


def test_ranocha_uniform_velocity_and_pressure(oracle):
    for ul, ur, v, p, n in _pair_states(4):
        f = oracle.volume_flux(P, RANOCHA, ul, ur, n)
        rl, rr = ul[0], ur[0]
        rln = (rr - rl) / np.log(rr / rl) if abs(rr - rl) > 1e-12 * rl else rl
        vn = v @ n
        assert _close(f[0], rln * vn, abs(rln * vn) + 1e-300)
        ref5 = vn * (0.5 * rln * v @ v + GAMMA * p / (GAMMA - 1))
        assert abs(f[4] - ref5) <= 1e-12 * abs(ref5), (f[4], ref5)


@pytest.mark.parametrize("which", [SOUZA, WARUSZEWSKI])
def test_gravity_energy_entry_is_the_work_of_the_momentum_entry(oracle, which):
    """energy entry = v_avg . momentum entries, v_avg the arithmetic mean of
    the two velocities; with one velocity, rho_avg [phi] n . v"""
    rng = np.random.default_rng(5)
    for ul, ur, v, p, n in _pair_states(6):
        # different velocities on the two sides for the identity
        ur2 = ur.copy()
        w = v + rng.normal(size=3)
        ur2[1:4] = ur[0] * w
        g = oracle.noncons_flux(P, which, ul, ur2, n)
        vavg = 0.5 * (v + w)
        assert abs(g[4] - vavg @ g[1:4]) <= 1e-13 * (np.abs(vavg * g[1:4]).sum() + 1e-300)
        # uniform velocity: the closed form with the family's mean density
        g = oracle.noncons_flux(P, which, ul, ur, n)
        rl, rr = ul[0], ur[0]
        ra = 0.5 * (rl + rr) if which == SOUZA else (
            (rr - rl) / np.log(rr / rl) if abs(rr - rl) > 1e-12 * rl else rl)
        dphi = ur[5] - ul[5]
        ref = ra * dphi * (n @ v)
        assert abs(g[4] - ref) <= 1e-13 * abs(ra * dphi) * np.abs(n * v).sum(), (g[4], ref)
        assert g[0] == 0.0 and g[5] == 0.0


def pressure_equilibrium_state(mesh, eq, seed=0):
    """one velocity and one pressure everywhere, density varying by a factor
    of about 3 across the periodic cube, phi = 0"""
    rng = np.random.default_rng(seed)
    x = mesh.node_coordinates
    k = rng.integers(1, 3, size=(3, 3))
    ph = rng.uniform(0, 2 * np.pi, 3)
    rho = 1.0 + 0.5 * sum(np.sin(2 * np.pi * (k[m] @ x.reshape(3, -1)).reshape(x.shape[1:]) + ph[m])
                          for m in range(3)) / 3
    v = np.array([0.3, -0.2, 0.1])
    p = 1.0
    u = np.zeros((6,) + rho.shape)
    u[0] = rho
    u[1:4] = rho[None] * v[:, None, None, None, None]
    u[4] = p / (eq.gamma - 1) + 0.5 * rho * (v @ v)
    return u, v


def pressure_velocity_rates(du, u, gamma):
    du = np.asarray(du).reshape(u.shape, order="F")
    rho = u[0]
    v = u[1:4] / rho
    dv = (du[1:4] - v * du[0]) / rho
    dp = (gamma - 1) * (du[4] - (v * du[1:4]).sum(axis=0) + 0.5 * (v * v).sum(axis=0) * du[0])
    return np.abs(dv).max(), np.abs(dp).max(), np.abs(du[0]).max()


def equilibrium_case(volume_flux):
    mesh = T.P4estMesh((2, 2, 2), polydeg=3, coordinates_min=(0.0, 0.0, 0.0),
                       coordinates_max=(1.0, 1.0, 1.0), periodicity=(True, True, True))
    eq = T.equations.CompressibleEulerEnergyEquationsWithGravity3D(c_p=1004.0, c_v=717.0,
                                                                   gravity=9.81)
    solver = T.DGSEM(polydeg=3, surface_flux=("flux_ranocha", "flux_zero"),
                     volume_integral=T.solvers.VolumeIntegralFluxDifferencing(
                         (volume_flux, "flux_zero")))
    u, v = pressure_equilibrium_state(mesh, eq)
    return mesh, eq, solver, u


@pytest.mark.parametrize("volume_flux,keeps", [("flux_shima_etal", True), ("flux_ranocha", True),
                                               ("flux_kennedy_gruber", False)])
def test_rhs_pressure_equilibrium(oracle, volume_flux, keeps):
    """rhs! of the oracle on a periodic mesh: uniform velocity and pressure
    stay uniform (to 2e-12 of the density tendency, which is O(1) here: the
    rounding of sums of pressure terms of size |D| p ~ 10) with the Shima or
    Ranocha volume flux; with Kennedy-Gruber the pressure changes at 1e-3 of
    the density tendency or more"""
    mesh, eq, solver, u = equilibrium_case(volume_flux)
    du = oracle.rhs(params_for((mesh, eq, solver, 0.0)), mesh, u)
    dv, dp, drho = pressure_velocity_rates(du, u, eq.gamma)
    assert drho > 0.1
    if keeps:
        assert dv <= 2e-12 * drho and dp <= 2e-12 * drho, (dv, dp, drho)
    else:
        assert dp > 1e-3 * drho, (dp, drho)
