"""The oracle's SaveSolutionCallback output variables against closed forms.

cons2prim_and_vorticity (src/callbacks_step/save_solution_covariant.jl:65-117)
of the covariant form is pinned by the analytic relative vorticity of two
Williamson et al. (1992) states that the reference's elixirs start from:
case 2 (zonal flow v_0 cos(lat): zeta = 2 v_0 sin(lat) / a) and case 6 (the
Rossby-Haurwitz wave of stream function -a^2 w sin(lat) + a^2 K cos^R(lat)
sin(lat) cos(R lon): zeta = 2 w sin(lat) - K (R + 1)(R + 2) cos^R(lat)
sin(lat) cos(R lon), the second term a spherical harmonic of degree R + 1),
by spectral convergence of the collocation derivative with the polynomial
degree, and by the global velocity, which must equal the initial
condition's velocity at the nodes and be tangent to the sphere."""
import numpy as np
import pytest

from trixiatmo_hip import initial_conditions as ic
from trixiatmo_hip import workloads as W
from oracle_lib import SweOracle


def _output(oracle, name, cells, polydeg):
    case = W.covariant_elixir(name, cells=cells, polydeg=polydeg)
    aux = case.aux()
    so = SweOracle.from_parts(oracle, case.mesh, aux, case.equations, case.solver,
                              case.source_terms)
    u0 = case.u0()
    out = so.prim_vorticity(u0)
    x = case.mesh.node_coordinates.reshape(3, -1, order="F")
    return case, aux, u0, out, x


def _lon_lat(x):
    a = np.sqrt((x ** 2).sum(axis=0))
    return a, np.arctan2(x[1], x[0]), np.arcsin(x[2] / a)


def zonal_flow_error(oracle, polydeg, cells=4):
    _, _, _, out, x = _output(oracle, "geostrophic_balance", cells, polydeg)
    a, _, lat = _lon_lat(x)
    v_0 = 2.0 * np.pi * a / (12 * ic.SECONDS_PER_DAY)
    zeta = 2.0 * v_0 * np.sin(lat) / a
    return np.abs(out[5] - zeta).max() / np.abs(zeta).max(), out, x


def rossby_haurwitz_error(oracle, polydeg, cells=5):
    _, _, _, out, x = _output(oracle, "rossby_haurwitz", cells, polydeg)
    _, lon, lat = _lon_lat(x)
    w, K, R = ic._f32(7.848e-6), ic._f32(7.848e-6), 4.0
    zeta = 2 * w * np.sin(lat) - K * (R + 1) * (R + 2) * np.cos(lat) ** R * np.sin(lat) * \
        np.cos(R * lon)
    return np.abs(out[5] - zeta).max() / np.abs(zeta).max()


def test_vorticity_of_the_zonal_flow(oracle):
    err, out, x = zonal_flow_error(oracle, 5)
    err7, _, _ = zonal_flow_error(oracle, 7)
    err_h, _, _ = zonal_flow_error(oracle, 5, cells=8)
    print("case 2 vorticity rel err: polydeg 5 / 7 on 4x4 per face, polydeg 5 on 8x8:",
          err, err7, err_h)
    # measured 1.5e-4, 2.3e-6 and 5.0e-6: order-p convergence in p and in h
    assert err < 3e-4
    assert err7 < err / 30
    assert err_h < err / 16
    # the global velocity is the initial condition's: v_0 cos(lat) eastward, tangent
    a, lon, lat = _lon_lat(x)
    v_0 = 2.0 * np.pi * a / (12 * ic.SECONDS_PER_DAY)
    east = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)])
    v = out[1:4]
    np.testing.assert_allclose(v, v_0 * np.cos(lat) * east, atol=1e-12 * v_0.max())
    assert np.abs((v * x).sum(axis=0)).max() <= 1e-12 * v_0.max() * a.max()
    np.testing.assert_array_equal(out[4], 0.0)


def test_vorticity_of_the_rossby_haurwitz_wave(oracle):
    e3, e5 = rossby_haurwitz_error(oracle, 3), rossby_haurwitz_error(oracle, 5)
    print("case 6 vorticity rel err, polydeg 3 / 5:", e3, e5)
    assert e5 < 1e-3
    assert e5 < e3 / 10  # spectral convergence of the collocation derivative


def test_total_height_includes_the_topography(oracle):
    case, aux, u0, out, _ = _output(oracle, "isolated_mountain", 3, 3)
    hs = aux[19].ravel(order="F")
    assert np.abs(hs).max() > 100.0
    np.testing.assert_array_equal(out[4], hs)
    np.testing.assert_allclose(out[0], u0[0].ravel(order="F") + hs, rtol=1e-15)
