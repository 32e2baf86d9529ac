"""Pointwise fluxes of the oracle pinned by consistency with the textbook Euler
flux, symmetry, Tadmor's entropy-conservation condition and antisymmetry of
the gravity term (the properties the reference's own test_unit.jl style
checks)."""
import ctypes
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle_lib import Oracle

P = Oracle.params(4, 1, c_p=1004.0, c_v=717.0, lmars_c=340.0)
GAMMA = 1004.0 / 717.0


def prim2cons(rho, v, p, phi):
    v = np.asarray(v, float)
    return np.array([rho, *(rho * v), p / (GAMMA - 1) + 0.5 * rho * v @ v, phi])


def euler_flux(u, n):
    """textbook flux F(u).n of the conserved variables (rho, rho v, E)"""
    rho = u[0]
    v = u[1:4] / rho
    p = (GAMMA - 1) * (u[4] - 0.5 * rho * v @ v)
    vn = v @ n
    return np.array([rho * vn, *(rho * vn * v + p * np.asarray(n)), (u[4] + p) * vn, 0.0])


def states(seed, k=20):
    rng = np.random.default_rng(seed)
    for _ in range(k):
        yield (prim2cons(rng.uniform(0.5, 2), rng.normal(size=3), rng.uniform(0.5, 2),
                         rng.uniform(-1, 1)), rng.normal(size=3))


def test_ln_mean_against_high_precision(oracle):
    getcontext().prec = 50
    rng = np.random.default_rng(1)
    xs = list(rng.uniform(0.1, 10, 50))
    pairs = [(x, x * (1 + d)) for x in xs[:25] for d in (1e-9, 1e-5, 1e-3, 0.02, 0.5)]
    pairs += [(a, b) for a, b in zip(xs[25:], xs[:25])]
    for x, y in pairs:
        X, Y = Decimal(x), Decimal(y)
        exact = (Y - X) / (Y / X).ln() if x != y else X
        assert abs(oracle.ln_mean(x, y) / float(exact) - 1) < 4e-15
        assert abs(oracle.inv_ln_mean(x, y) * float(exact) - 1) < 4e-15


@pytest.mark.parametrize("which", [0, 1, 2])
def test_volume_flux_consistency_and_symmetry(oracle, which):
    for (u, n), (u2, _) in zip(states(2), states(3)):
        np.testing.assert_allclose(oracle.volume_flux(P, which, u, u, n), euler_flux(u, n),
                                   rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(oracle.volume_flux(P, which, u, u2, n),
                                   oracle.volume_flux(P, which, u2, u, n), rtol=1e-14, atol=1e-14)


def test_ranocha_tadmor_condition(oracle):
    """(w_r - w_l) . f# = (psi_r - psi_l) . n with psi = rho v for S = -rho s/(gamma-1)"""
    for (ul, n), (ur, _) in zip(states(4), states(5)):
        f = oracle.volume_flux(P, 0, ul, ur, n)
        dw = oracle.cons2entropy(P, ur) - oracle.cons2entropy(P, ul)
        dpsi = (ur[1:4] - ul[1:4]) @ n
        assert abs(dw @ f - dpsi) < 1e-12 * (1 + abs(dpsi) + np.abs(dw * f).sum())


def test_pointwise_flux_uses_phi_for_energy(oracle):
    """The reference's pointwise `flux` reads rho_e = last(u) = phi; with
    phi set to rho_e it is the textbook flux."""
    for u, n in states(6):
        v = u.copy()
        v[5] = u[4]
        np.testing.assert_allclose(oracle.flux(P, v, n)[:5], euler_flux(u, n)[:5], rtol=1e-13,
                                   atol=1e-13)
        w = u.copy()
        w[5] = 0.0
        assert abs(oracle.flux(P, w, n)[4] - (euler_flux(u, n)[4] - u[4] * (u[1:4] / u[0]) @ n)) < 1e-12


@pytest.mark.parametrize("which", [0, 1, 2, 3, 4, 5])
def test_surface_flux_consistency_and_conservation(oracle, which):
    for (ul, n), (ur, _) in zip(states(7), states(8)):
        f_lr = oracle.surface_flux(P, which, ul, ur, n)
        f_rl = oracle.surface_flux(P, which, ur, ul, -n)
        np.testing.assert_allclose(f_lr, -f_rl, rtol=1e-13, atol=1e-12)
        if which in (1, 2, 3, 4, 5):  # two-point fluxes reduce to the textbook flux
            np.testing.assert_allclose(oracle.surface_flux(P, which, ul, ul, n)[:5],
                                       euler_flux(ul, n)[:5], rtol=1e-13, atol=1e-12)
        else:  # LLF reduces to the (phi-reading) pointwise flux
            np.testing.assert_allclose(oracle.surface_flux(P, 0, ul, ul, n),
                                       oracle.flux(P, ul, n), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("which", [1, 2])
def test_gravity_term_antisymmetric(oracle, which):
    for (ul, n), (ur, _) in zip(states(9), states(10)):
        g1 = oracle.noncons_flux(P, which, ul, ur, n)
        g2 = oracle.noncons_flux(P, which, ur, ul, n)
        np.testing.assert_allclose(g1, -g2, rtol=1e-14, atol=1e-14)
        assert g1[0] == 0 and g1[5] == 0
        # rho_avg [phi] n in the momentum, rho_avg [phi] n.v_avg in the energy
        ra = oracle.ln_mean(ul[0], ur[0]) if which == 1 else 0.5 * (ul[0] + ur[0])
        np.testing.assert_allclose(g1[1:4], ra * (ur[5] - ul[5]) * n, rtol=1e-13)


@pytest.mark.parametrize("which,base", [(4, 2), (5, 1)])
def test_flux_plus_dissipation(oracle, which, base):
    """FluxPlusDissipation(f#, LLF) = f# - lambda_max/2 [u] with the
    max_abs_speed of the normal direction"""
    for (ul, n), (ur, _) in zip(states(11), states(12)):
        lam = oracle.L.ora_e3_max_abs_speed(ctypes.byref(P), ul.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            ur.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            np.asarray(n, float).ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        rho_l, rho_r = ul[0], ur[0]
        vl, vr = ul[1:4] / rho_l, ur[1:4] / rho_r
        pl = (GAMMA - 1) * (ul[4] - 0.5 * rho_l * vl @ vl)
        pr = (GAMMA - 1) * (ur[4] - 0.5 * rho_r * vr @ vr)
        nn = np.linalg.norm(n)
        lam_ref = max(abs(vl @ n) + np.sqrt(GAMMA * pl / rho_l) * nn,
                      abs(vr @ n) + np.sqrt(GAMMA * pr / rho_r) * nn)
        assert abs(lam / lam_ref - 1) < 1e-14
        np.testing.assert_allclose(oracle.surface_flux(P, which, ul, ur, n),
                                   oracle.surface_flux(P, base, ul, ur, n) - 0.5 * lam_ref * (ur - ul),
                                   rtol=1e-13, atol=1e-12)
