"""The rainy-Euler oracle (saturation Newton solve, fluxes, rhs!, stage
limiter order) against the reference's regression values and against closed
forms.

Pin: test/test_2d_rainy_euler.jl:7-31 (examples/euler/precipitation/tests/
convergence_test.jl: 6x6 periodic StructuredMesh on [0, 2]^2, polydeg 3, weak
form, flux_lax_friedrichs, manufactured source, NonlinearSolveDG with tol
1e-8 as the stage limiter of CarpenterKennedy2N54, StepsizeCallback(cfl = 1),
t = 2); test/test_2d_rainy_euler.jl:79-114 (the rainy bubble of
examples/euler/precipitation/buoyancy/elixir_hoeck_bubble.jl, 16 x 16 cells,
10 s: flux_ec_rain, LLF, source_terms_rainy, the simple slip wall and the
stage limiter together; with the reference's adaptive SSPRK43 restated, and
with this package's fixed-step integrator within the measured integrator gap);
test/test_unit.jl:526-544 (flux_ec_rain against flux on the reference's
state and normals).  Closed forms: saturated states of that test and unsaturated states
solve the saturation system exactly; the Jacobian and the derivative of the
saturation pressure match finite differences of the functions they
differentiate.  Invariants: conservation of dry air, of total water and of
momentum without gravity; consistency and symmetry of flux_ec_rain; entropy
conservation of flux_ec_rain.
"""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import initial_conditions as ic
from trixiatmo_hip.analysis import interpolation_matrix
from trixiatmo_hip.meshes import StructuredMesh2D
from oracle_lib import RainyOracle

EQ = T.equations.CompressibleRainyEulerEquations2D()

CONV_L2 = [2.39895785368954e-6, 3.892427386570826e-5, 1.5186141589998252e-5,
           0.007532307120716242, 0.02085654062623111, 51.988569260688045,
           1.7470730680079724e-8, 3.892376665583286e-5, 0.0002156633044466944]
CONV_LINF = [1.675032502916618e-5, 0.0002740479581908595, 8.691475272920579e-5,
             0.023414703432404593, 0.06248365429859781, 385.10196906188503,
             6.003313225070965e-8, 0.000274026861750043, 0.0007433951467703537]


def conv_mesh():
    return StructuredMesh2D((6, 6), (0.0, 0.0), (2.0, 2.0), polydeg=3)


def calc_error_norms_2d(u, mesh, exact, t):
    """Trixi's calc_error_norms for StructuredMesh{2} [ext]: interpolation to
    the LGL nodes of degree 2 polydeg, l2 weighted by w_i w_j |J|"""
    x_sol, _, _, _ = T.lgl_basis(mesh.polydeg)
    x_ana, w_ana, _, _ = T.lgl_basis(2 * mesh.polydeg)
    V = interpolation_matrix(x_sol, x_ana)
    Ui = np.einsum("ai,bj,vije->vabe", V, V, u)
    Xi = np.einsum("ai,bj,vije->vabe", V, V, mesh.node_coordinates)
    J = 0.25 * mesh.dx * mesh.dy
    W = (w_ana[:, None] * w_ana[None, :] * J)[:, :, None] * np.ones(mesh.n_elements)
    diff = exact(Xi, t) - Ui
    l2 = np.sqrt((diff ** 2 * W).sum(axis=(1, 2, 3)) / W.sum())
    return l2, np.abs(diff).max(axis=(1, 2, 3))


def test_convergence_test_regression(oracle):
    mesh = conv_mesh()
    ro = RainyOracle(oracle, mesh, volume_integral=0,
                     source="source_terms_convergence_test_rainy")
    exact = lambda x, t: ic.initial_condition_convergence_test_rainy(x, t, EQ)
    u0 = exact(mesh.node_coordinates, 0.0)
    u, steps = ro.solve(u0, 2.0, 1.0, tol=1e-8)
    l2, linf = calc_error_norms_2d(u.reshape(u0.shape, order="F"), mesh, exact, 2.0)
    np.testing.assert_allclose(l2, CONV_L2, rtol=1e-8)
    np.testing.assert_allclose(linf, CONV_LINF, rtol=1e-7)


def _conv_states():
    mesh = conv_mesh()
    u = ic.initial_condition_convergence_test_rainy(mesh.node_coordinates, 0.3, EQ)
    return u.reshape(9, -1, order="F").T


def test_newton_saturated_closed_form(oracle):
    """saturated states of convergence_test.jl:8-40: the solution is
    (e_s(T)/(R_v T), rho / c_l 3000, rho + 250) in closed form"""
    ro = RainyOracle(oracle, conv_mesh())
    rng = np.random.default_rng(0)
    for u in _conv_states()[::7]:
        guess = u[6:9] * (1 + 0.05 * rng.standard_normal(3))
        out, it = ro.newton_node(u, guess, tol=1e-9)
        assert 0 < it <= 20
        np.testing.assert_allclose(out[6:9], u[6:9], rtol=1e-11)
        assert np.abs(ro.residual(u, out[6:9])).max() < 1e-9


def test_newton_unsaturated_closed_form(oracle):
    """below saturation rho_vapour = rho_moist, rho_cloud = 0 and T solves the
    energy row, which is linear in T"""
    ro = RainyOracle(oracle, conv_mesh())
    rng = np.random.default_rng(1)
    for _ in range(20):
        Tt = rng.uniform(250, 300)
        rv = 0.5 * EQ.saturation_vapour_pressure(Tt) / (EQ.R_vapour * Tt)
        rd, rr = rng.uniform(0.5, 1.2), rng.uniform(0, 1e-4)
        v1, v2 = rng.uniform(-20, 20, 2)
        rho = rd + rv + rr
        E = (EQ.c_dry_air_const_volume * rd + EQ.c_vapour_const_volume * rv +
             EQ.c_liquid_water * rr) * Tt + rv * EQ.ref_latent_heat_vap_temp + \
            0.5 * rho * (v1 ** 2 + v2 ** 2)
        u = np.array([rd, rv, rr, rho * v1, rho * v2, E, rv * 0.9, 1e-5, Tt + 3])
        out, it = ro.newton_node(u, tol=1e-9)
        assert it > 0
        np.testing.assert_allclose(out[6:9], [rv, 0.0, Tt], rtol=1e-12, atol=1e-15)


def test_newton_reference_quirks(oracle):
    """a negative rain density is clamped in u but the residual uses the
    value read before the clamp; non-convergence keeps the last iterate and
    reports failure"""
    ro = RainyOracle(oracle, conv_mesh())
    u = _conv_states()[3].copy()
    u[2] = -1e-6
    out, it = ro.newton_node(u, tol=1e-9)
    assert out[2] == 0.0 and it >= 0
    u2 = u.copy()
    u2[2] = 0.0
    out2, _ = ro.newton_node(u2, tol=1e-9)
    assert abs(out[8] - out2[8]) > 1e-6   # the solve saw rho_rain = -1e-6
    u3 = _conv_states()[3].copy()
    out3, it3 = ro.newton_node(u3, u3[6:9] * np.array([1.0, 1.0, 1.3]), tol=1e-30, max_iter=3)
    assert it3 == -1


def test_saturation_pressure_derivative_and_jacobian(oracle):
    ro = RainyOracle(oracle, conv_mesh())
    for Tt in (230.0, 260.0, 273.15, 300.0):
        e, de = ro.svp(Tt)
        h = 1e-5 * Tt
        fd = (ro.svp(Tt + h)[0] - ro.svp(Tt - h)[0]) / (2 * h)
        np.testing.assert_allclose(de, fd, rtol=1e-7)
        np.testing.assert_allclose(e, EQ.saturation_vapour_pressure(Tt), rtol=1e-14)
    # 610.7 Pa at the reference temperature
    np.testing.assert_allclose(ro.svp(273.15)[0], 610.7, rtol=1e-15)
    for u in _conv_states()[::11]:
        for scale in ((1.0, 1.0, 1.0), (0.3, 1.0, 1.0)):  # saturated / unsaturated guess
            g = u[6:9] * np.array(scale) * np.array([1.0, 1.0, 1.001])
            J = ro.jacobian(u, g)
            fd = np.zeros((3, 3))
            for k in range(3):
                h = 1e-6 * abs(g[k])
                gp, gm = g.copy(), g.copy()
                gp[k] += h
                gm[k] -= h
                fd[:, k] = (ro.residual(u, gp) - ro.residual(u, gm)) / (2 * h)
            np.testing.assert_allclose(J, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def _random_state(rng, rain=True):
    Tt = rng.uniform(260, 300)
    rd = rng.uniform(0.8, 1.2)
    rv = 0.7 * EQ.saturation_vapour_pressure(Tt) / (EQ.R_vapour * Tt)
    rc = rng.uniform(0, 1e-4)
    rr = rng.uniform(1e-6, 1e-4) if rain else 0.0
    v = rng.uniform(-10, 10, 2)
    rho = rd + rv + rc + rr
    E = (EQ.c_dry_air_const_volume * rd + EQ.c_vapour_const_volume * rv +
         EQ.c_liquid_water * (rc + rr)) * Tt + rv * EQ.ref_latent_heat_vap_temp + \
        0.5 * rho * (v @ v)
    return np.array([rd, rv + rc, rr, rho * v[0], rho * v[1], E, rv, rc, Tt])


def test_flux_ec_rain_consistency_and_symmetry(oracle):
    ro = RainyOracle(oracle, conv_mesh())
    rng = np.random.default_rng(2)
    for _ in range(10):
        a, b = _random_state(rng), _random_state(rng)
        n = rng.standard_normal(2)
        np.testing.assert_allclose(ro.flux_ec(a, a, n), ro.flux(a, n), rtol=1e-12,
                                   atol=1e-12 * np.abs(ro.flux(a, n)).max())
        fab = ro.flux_ec(a, b, n)
        np.testing.assert_allclose(fab, ro.flux_ec(b, a, n), rtol=1e-14,
                                   atol=1e-14 * np.abs(fab).max())
        f_llf = ro.flux_llf(a, a, n)
        np.testing.assert_allclose(f_llf, ro.flux(a, n), rtol=1e-14,
                                   atol=1e-14 * np.abs(f_llf).max())


def test_flux_ec_rain_reference_consistency_pattern(oracle):
    """the reference's own known answer (test/test_unit.jl:526-544): on the
    state (1.0, 0.2, 0.1, 0.5, -0.4, 2.2, 0.1, 0.1, 300), whose energy and
    temperature do not agree, flux_ec_rain(u, u, n) equals flux(u, n) under
    Julia's isapprox (rtol sqrt(eps)) in every row except the energy row, for
    the four normals of that test"""
    ro = RainyOracle(oracle, conv_mesh())
    u = np.array([1.0, 0.2, 0.1, 0.5, -0.4, 2.2, 0.1, 0.1, 300.0])
    rtol = np.sqrt(np.finfo(float).eps)
    expected = [True, True, True, True, True, False, True, True, True]
    for n in ([1.0, 0.0], [0.0, 1.0], [0.5, -0.5], [-1.2, 0.3]):
        a, b = ro.flux_ec(u, u, np.array(n)), ro.flux(u, np.array(n))
        equal = np.abs(a - b) <= rtol * np.maximum(np.abs(a), np.abs(b))
        assert equal.tolist() == expected, (n, a, b)


def _bubble(nx=8, ny=6, periodic_y=False, seed=3):
    mesh = StructuredMesh2D((nx, ny), (0.0, 0.0), (4000.0, 3000.0), polydeg=3,
                            periodicity=(True, periodic_y))
    u, meta = ic.initial_condition_rainy_bubble(mesh, EQ, seed=seed, radius=800.0)
    return mesh, u


def _weights(mesh):
    _, w, _, _ = T.lgl_basis(mesh.polydeg)
    return (w[:, None] * w[None, :] * 0.25 * mesh.dx * mesh.dy)[:, :, None]


def test_conservation_with_rain(oracle):
    """periodic in both directions with phase changes and rain: dry air and
    total water (moist + rain) are conserved, momentum changes only by
    gravity"""
    mesh, u = _bubble(periodic_y=True)
    rng = np.random.default_rng(4)
    u = u.copy()
    u[2] += 1e-5 * rng.uniform(0, 1, u[2].shape)       # rain
    u[0] -= u[2]
    u[3] += rng.standard_normal(u[3].shape)              # motion
    ro = RainyOracle(oracle, mesh, volume_integral=1, source="source_terms_rainy")
    u, _, _ = ro.newton(u)
    u = u.reshape(9, 4, 4, -1, order="F")
    du = ro.rhs(u).reshape(u.shape, order="F")
    W = _weights(mesh)
    scale = np.abs(du).max(axis=(1, 2, 3)) * W.sum()
    assert abs((W * du[0]).sum()) <= 1e-12 * scale[0]
    assert abs((W * (du[1] + du[2])).sum()) <= 1e-12 * (scale[1] + scale[2])
    assert abs((W * du[3]).sum()) <= 1e-12 * scale[3]
    rho = u[0] + u[1] + u[2]
    np.testing.assert_allclose((W * du[4]).sum(), -(W * rho).sum() * EQ.gravity, rtol=1e-10)


def test_entropy_dissipation_of_llf(oracle):
    """flux_ec_rain in the volume; the local Lax-Friedrichs surface flux adds
    -lambda/2 [u], which raises dS/dt = int w(u) . du against the
    flux_ec_rain surface flux on the same state (cons2entropy is the gradient
    of the physical, concave entropy: dissipation produces it).  (flux_ec_rain itself
    leaves a Tadmor residual of ~1e-4 relative for the entropy of
    cons2entropy, so dS/dt is not at rounding level with it.)"""
    mesh, u = _bubble(periodic_y=True)
    rng = np.random.default_rng(5)
    u = u.copy()
    u[3] += 2 * rng.standard_normal(u[3].shape)
    u[4] += 2 * rng.standard_normal(u[4].shape)
    ro = RainyOracle(oracle, mesh, volume_integral=1, source=None, surface_flux=1)
    u, _, _ = ro.newton(u)
    dS_ec = ro.integrals(u, ro.rhs(u))[9]
    ro2 = RainyOracle(oracle, mesh, volume_integral=1, source=None, surface_flux=0)
    dS_llf = ro2.integrals(u, ro2.rhs(u))[9]
    assert dS_llf > dS_ec + 1e3 * abs(dS_ec)


def _entropy_vars(oracle, u):
    import ctypes
    from oracle_lib import _d
    for q in np.asarray(u).reshape(-1, 9):
        w = np.zeros(9)
        oracle.L.ora_rainy_cons2entropy(_d(np.ascontiguousarray(q)), _d(w))
        yield w


def test_lake_at_rest_without_gravity(oracle):
    """a uniform state at rest is a fixed point of the flux terms"""
    mesh = StructuredMesh2D((4, 3), (0.0, 0.0), (400.0, 300.0), polydeg=3,
                            periodicity=(True, False))
    rng = np.random.default_rng(6)
    s = _random_state(rng)
    s[3:5] = 0.0
    s[5] -= 0.0
    u = np.broadcast_to(s[:, None, None, None], (9, 4, 4, mesh.n_elements)).copy()
    for vi in (0, 1):
        ro = RainyOracle(oracle, mesh, volume_integral=vi, source=None)
        du = ro.rhs(u).reshape(u.shape, order="F")
        # rain falls: only the rain terms of the vertical flux survive, and
        # their divergence vanishes in the interior of a uniform column
        assert np.abs(du[:, :, :, 4:8][[0, 1, 3]]).max() < 1e-10


def test_mlp_oracle_gradient_finite_differences(oracle):
    """the FP64 loss gradient of the MLP oracle against central differences of
    its own loss (parameters perturbed by float32-representable steps)"""
    import ctypes
    from oracle_lib import _d
    rng = np.random.default_rng(0)
    p = (0.3 * rng.standard_normal(1250)).astype(np.float32)
    norm = np.array([1.0, 2.3e5, 1e-2, 0.1, 2e4, 5e-3, 260.0, 5e-3, 20.0, 5e-3], dtype=np.float32)
    x = np.stack([rng.uniform(0.8, 1.2, 64), rng.uniform(2e5, 2.6e5, 64),
                  rng.uniform(1e-3, 2e-2, 64)], axis=1).astype(np.float32)
    y = np.stack([rng.uniform(240, 290, 64), rng.uniform(1e-3, 1e-2, 64)], axis=1).astype(np.float32)
    fp = ctypes.POINTER(ctypes.c_float)

    def loss_grad(pp):
        L = ctypes.c_double()
        g = np.zeros(1250)
        oracle.L.ora_mlp_loss_grad(pp.ctypes.data_as(fp), norm.ctypes.data_as(fp),
                                   x.ctypes.data_as(fp), y.ctypes.data_as(fp), ctypes.c_long(64),
                                   ctypes.byref(L), _d(g))
        return L.value, g

    _, g = loss_grad(p)
    for idx in list(range(0, 96, 13)) + [96, 130, 500, 1100, 1160, 1190, 1224, 1248, 1249]:
        h = np.float32(1e-2) * max(abs(p[idx]), np.float32(0.05))
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (loss_grad(pp)[0] - loss_grad(pm)[0]) / (float(pp[idx]) - float(pm[idx]))
        assert abs(fd - g[idx]) <= 2e-3 * abs(g).max(), (idx, fd, g[idx])


def test_mlp_oracle_forward_matches_its_loss(oracle):
    """the oracle's forward outputs give back the oracle's loss (the forward
    that checks the GPU network is the one the gradient test pins)"""
    import ctypes
    from oracle_lib import _d
    rng = np.random.default_rng(1)
    p = (0.3 * rng.standard_normal(1250)).astype(np.float32)
    norm = np.array([1.0, 2.3e5, 1e-2, 0.1, 2e4, 5e-3, 260.0, 5e-3, 20.0, 5e-3], dtype=np.float32)
    x = np.stack([rng.uniform(0.8, 1.2, 200), rng.uniform(2e5, 2.6e5, 200),
                  rng.uniform(1e-3, 2e-2, 200)], axis=1).astype(np.float32)
    y = np.stack([rng.uniform(240, 290, 200), rng.uniform(1e-3, 1e-2, 200)], axis=1).astype(np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    L = ctypes.c_double()
    oracle.L.ora_mlp_loss_grad(p.ctypes.data_as(fp), norm.ctypes.data_as(fp), x.ctypes.data_as(fp),
                               y.ctypes.data_as(fp), ctypes.c_long(200), ctypes.byref(L),
                               _d(np.zeros(1250)))
    out = np.zeros((200, 2))
    oracle.L.ora_mlp_forward(p.ctypes.data_as(fp), norm.ctypes.data_as(fp), x.ctypes.data_as(fp),
                             ctypes.c_long(200), _d(out))
    scale = norm[8:].astype(np.float64)
    yt = (y.astype(np.float64) - norm[6:8]) / scale
    yo = (out - norm[6:8]) / scale
    np.testing.assert_allclose(((yo - yt) ** 2).sum(axis=1).mean(), L.value, rtol=1e-12)


# --- test/test_2d_rainy_euler.jl:79-114: the rainy bubble (elixir_hoeck_bubble.jl,
# 16 x 16 cells, 10 s): flux_ec_rain volume, flux_lax_friedrichs surface,
# source_terms_rainy, boundary_condition_simple_slip_wall, NonlinearSolveDG

# bars of the oracle run with the reference's integrator restated: three times
# the largest difference found (1.4e-6 l2, 3.5e-5 linf; rain 1e-4 / 6.5e-4)
BUBBLE_RTOL_L2, BUBBLE_RTOL_LINF = 3e-5, 1e-4
BUBBLE_RAIN_RTOL_L2, BUBBLE_RAIN_RTOL_LINF = 5e-4, 3e-3


def _bubble_oracle(oracle):
    import rainy_bubble_pin as B
    mesh = B.bubble_mesh()
    ro = RainyOracle(oracle, mesh, volume_integral=1, source="source_terms_rainy")
    u0 = B.exact(mesh.node_coordinates, 0.0)
    return B, mesh, ro, u0


def _bubble_bars(rain_l2, rain_linf, rest_l2, rest_linf):
    l2 = np.full(9, rest_l2)
    linf = np.full(9, rest_linf)
    l2[2], linf[2] = rain_l2, rain_linf
    return l2, linf


def test_bubble_initial_condition_solves_its_residuals():
    """every node of the bubble's initial state is a root of the reference's
    perturbation residual (saturated in the bubble core, relative humidity 0.2
    outside), and the hydrostatic layers are roots of theirs"""
    import rainy_bubble_pin as B
    L = B.layers()
    mesh = B.bubble_mesh()
    u = B.exact(mesh.node_coordinates, 0.0)
    T_ = u[8]
    es = EQ.saturation_vapour_pressure(T_)
    rvs = es / (EQ.R_vapour * T_)
    eps = EQ.R_dry_air / EQ.R_vapour
    x, z = mesh.node_coordinates
    r = np.sqrt((x - 1200) ** 2 + (z - 800) ** 2)
    p = (u[0] * EQ.R_dry_air + u[1] * EQ.R_vapour) * T_
    H = u[1] * (u[0] + rvs / eps) / (rvs * (u[0] + u[1] / eps))  # humidity residual solved for H
    assert np.allclose(H[r > 300], 0.2, rtol=1e-12)
    assert np.allclose(H[r < 200], 1.0, rtol=1e-12)
    assert np.all(np.diff(L.layer_data[:, 0]) < 0)  # pressure falls with height
    d = L.layer_data
    # hydrostatic residual of each layer on the one below (dz = 1)
    res1 = (d[1:, 0] - d[:-1, 0]) + (d[1:, 1] + d[1:, 2]) * EQ.gravity
    assert np.abs(res1).max() < 1e-9 * d[0, 0]
    assert np.abs(d[:, 0] - d[:, 3] * (d[:, 1] * EQ.R_dry_air + d[:, 2] * EQ.R_vapour)).max() \
        < 1e-10 * d[0, 0]
    assert np.all(p > 0)


def test_bubble_regression_with_the_reference_integrator(oracle):
    """the oracle integrated with the reference's adaptive SSPRK43 (default
    tolerances, stage limiter after every stage) reproduces the regression
    values of all nine variables"""
    B, mesh, ro, u0 = _bubble_oracle(oracle)
    u, steps = B.ssprk43_adaptive(lambda v, t: ro.rhs(v, t), lambda v: ro.newton(v)[0],
                                  ro._u(u0), B.T_END)
    u = u.reshape(u0.shape, order="F")
    l2, linf = calc_error_norms_2d(u, mesh, B.exact, B.T_END)
    print("SSPRK43 steps", steps, "\nl2", (l2 - B.BUBBLE_L2) / B.BUBBLE_L2,
          "\nlinf", (linf - B.BUBBLE_LINF) / B.BUBBLE_LINF)
    bl2, blinf = _bubble_bars(BUBBLE_RAIN_RTOL_L2, BUBBLE_RAIN_RTOL_LINF, BUBBLE_RTOL_L2,
                              BUBBLE_RTOL_LINF)
    assert np.all(np.abs(l2 - B.BUBBLE_L2) <= bl2 * B.BUBBLE_L2)
    assert np.all(np.abs(linf - B.BUBBLE_LINF) <= blinf * B.BUBBLE_LINF)


def bubble_ck2n54_bars(oracle):
    """the oracle's fixed-step CarpenterKennedy2N54 run (CFL 0.25, the
    integrator of this package) and the bars of its regression values: three
    times the measured gap between that run and the restated reference
    integrator, plus 1e-6 relative; returns (u, steps, l2 bar, linf bar)"""
    B, mesh, ro, u0 = _bubble_oracle(oracle)
    u_ck, steps = ro.solve(u0, B.T_END, B.CFL)
    u_ck = u_ck.reshape(u0.shape, order="F")
    u_ss, _ = B.ssprk43_adaptive(lambda v, t: ro.rhs(v, t), lambda v: ro.newton(v)[0],
                                 ro._u(u0), B.T_END)
    u_ss = u_ss.reshape(u0.shape, order="F")
    l2c, lic = calc_error_norms_2d(u_ck, mesh, B.exact, B.T_END)
    l2s, lis = calc_error_norms_2d(u_ss, mesh, B.exact, B.T_END)
    bar_l2 = 3 * np.abs(l2c - l2s) + 1e-6 * np.abs(l2s)
    bar_linf = 3 * np.abs(lic - lis) + 1e-6 * np.abs(lis)
    return u_ck, steps, bar_l2, bar_linf


def test_bubble_regression_fixed_steps(oracle):
    """the oracle with this package's integrator (fixed-step CK2N54 at CFL
    0.25): the regression values within the derived integrator gap"""
    import rainy_bubble_pin as B
    u, steps, bar_l2, bar_linf = bubble_ck2n54_bars(oracle)
    l2, linf = calc_error_norms_2d(u, B.bubble_mesh(), B.exact, B.T_END)
    print("CK2N54 steps", steps, "\nl2 diff/bar", np.abs(l2 - B.BUBBLE_L2) / bar_l2,
          "\nlinf diff/bar", np.abs(linf - B.BUBBLE_LINF) / bar_linf)
    assert np.all(np.abs(l2 - B.BUBBLE_L2) <= bar_l2)
    assert np.all(np.abs(linf - B.BUBBLE_LINF) <= bar_linf)
