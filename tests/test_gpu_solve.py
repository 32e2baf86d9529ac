"""GPU: the resident time loop (atmo_solve) against the host loop of
atmo_max_dt + atmo_step, the StepsizeCallback loop of the examples' solve
(examples/shallow_water/covariant/elixir_barotropic_instability.jl:67,79-80).

Both loops compute every step size from the same reduction with the same
operations, so the bar is exact: the same number of steps, the same final
time and the same state bit for bit, for the 3D Euler path (FP64 and mixed
precision), the covariant shallow-water path (single handles and ensembles,
whose members each take their own step sizes) and the rainy path with its
stage limiter and the time-dependent convergence-test source; also the
reference's convergence-test regression values reached by the resident loop."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import initial_conditions as ic
from trixiatmo_hip import workloads as W
from trixiatmo_hip.solvers import DGSEM, VolumeIntegralWeakForm
from test_oracle_rainy import CONV_L2, CONV_LINF, EQ, calc_error_norms_2d, conv_mesh

pytestmark = pytest.mark.gpu


def both_loops(make_semi, u0, t_end, cfl, check_every, timing=False):
    a = make_semi()
    u_host, n_host = T.solve(a, u0, (0.0, t_end), cfl)
    b = make_semi()
    if timing:
        b.enable_timing(True)  # launches one by one, outside a graph
    u_res, n_res = T.solve(b, u0, (0.0, t_end), cfl, resident=True, check_every=check_every)
    print(f"  host {n_host} steps, resident {n_res} steps", flush=True)
    return (u_host, n_host), (u_res, n_res), b


def euler3d(workload, **kw):
    mesh, eq, solver, bcs, u0 = workload(**kw)
    return (lambda: T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)), u0


@pytest.mark.parametrize("precision,check_every,timing",
                         [("fp64", 5, False), ("fp64", 64, True), ("fp32", 1, False)])
def test_resident_loop_3d_periodic_cube(precision, check_every, timing):
    print(f"case: periodic cube trees 3 polydeg 3 {precision} check_every {check_every} "
          f"timing {timing}", flush=True)
    make, u0 = euler3d(W.periodic_cube, trees=3, polydeg=3, seed=2, precision=precision)
    s = make()
    s.set_state(u0)
    t_end = 12.5 * s.max_dt(0.5)  # the last step is cut short
    (uh, nh), (ur, nr), semi = both_loops(make, u0, t_end, 0.5, check_every, timing)
    assert nh == nr and nh >= 13
    np.testing.assert_array_equal(ur, uh)
    steps, t = semi.solve_resident(t_end, 0.5)  # already there: no step
    assert steps[0] == 0
    np.testing.assert_array_equal(semi.get_state().reshape(semi.shape, order="F"), uh)


def test_resident_loop_3d_shell_with_gravity():
    print("case: shell 2x2 trees x 2 layers polydeg 3", flush=True)
    make, u0 = euler3d(W.shell, trees_per_face=2, layers=2, polydeg=3, seed=1)
    s = make()
    s.set_state(u0)
    t_end = 9.3 * s.max_dt(0.5)
    (uh, nh), (ur, nr), _ = both_loops(make, u0, t_end, 0.5, 4)
    assert nh == nr
    np.testing.assert_array_equal(ur, uh)


def test_resident_loop_covariant():
    print("case: zonal jet 3x3 trees polydeg 3", flush=True)
    case = W.zonal_jet(3, 3, seed=3)
    u0 = case.u0()
    s = case.semi()
    s.set_state(u0)
    t_end = 20.5 * s.max_dt(0.4)
    (uh, nh), (ur, nr), semi = both_loops(case.semi, u0, t_end, 0.4, 8)
    assert nh == nr
    np.testing.assert_array_equal(ur, uh)


def test_resident_ensemble_members_take_their_own_steps():
    """each member of a resident ensemble reproduces that member solved alone
    by the host loop: its own step count and its state bit for bit"""
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=4, polydeg=3, seed=6,
                                                    choices=(2, 3))
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    u0 = ens.compute_coefficients()
    ens.set_state(u0)
    t_end = 10.5 * ens.max_dt_members(0.4).max()
    print(f"case: ensemble of {ens.n_members} members, t_end {t_end}", flush=True)
    u_ens, steps = T.solve(ens, u0, (0.0, t_end), 0.4, resident=True, check_every=6)
    first = ens.member_first_element
    for m, (mesh, icm, _) in enumerate(members):
        single = T.SemidiscretizationHyperbolic(mesh, eq, solver, source_terms=src,
                                                initial_condition=icm)
        um0 = single.compute_coefficients()
        um, n = T.solve(single, um0, (0.0, t_end), 0.4)
        print(f"  member {m}: alone {n} steps, in the ensemble {steps[m]}", flush=True)
        assert steps[m] == n
        np.testing.assert_array_equal(u_ens[..., first[m]:first[m + 1]], um)


def test_swe_graph_follows_the_state_buffers():
    """an odd number of steps swaps the state buffers: the cached two-step
    graph is captured again instead of stepping the stale buffer"""
    case = W.zonal_jet(3, 3, seed=8)
    semi = case.semi()
    u0 = case.u0()
    semi.set_state(u0)
    dt = semi.max_dt(0.4)
    semi.step(dt, 2)
    semi.step(dt, 3)
    semi.step(dt, 2)
    a = semi.get_state()
    semi.set_state(u0)
    for _ in range(7):
        semi.step(dt, 1)
    np.testing.assert_array_equal(a, semi.get_state())


def rainy_make(cells=(16, 8)):
    mesh, eq, solver, src, lim, u0, _ = W.rainy_bubble(cells=cells, polydeg=3, seed=3)
    return (lambda: T.SemidiscretizationRainy(mesh, eq, solver, source_terms=src,
                                              stage_limiter=lim)), u0


@pytest.mark.parametrize("timing", [False, True])
def test_resident_loop_rainy_bubble(timing):
    print(f"case: rainy bubble 16x8 polydeg 3 timing {timing}", flush=True)
    make, u0 = rainy_make()
    s = make()
    s.set_state(u0)
    t_end = 7.5 * s.max_dt(0.5)
    (uh, nh), (ur, nr), _ = both_loops(make, u0, t_end, 0.5, 3, timing)
    assert nh == nr
    np.testing.assert_array_equal(ur, uh)


def test_resident_loop_rainy_convergence_regression():
    """the time-dependent source reads each stage's time from the device
    clock: the same bits as the host loop, and the reference's
    convergence_test.jl regression values"""
    mesh = conv_mesh()
    solver = DGSEM(polydeg=3, surface_flux="flux_lax_friedrichs",
                   volume_integral=VolumeIntegralWeakForm())

    def make():
        return T.SemidiscretizationRainy(mesh, EQ, solver,
                                         source_terms="source_terms_convergence_test_rainy",
                                         stage_limiter=T.NonlinearSolveDG(1e-8, 20))

    exact = lambda x, t: ic.initial_condition_convergence_test_rainy(x, t, EQ)
    u0 = exact(mesh.node_coordinates, 0.0)
    print("case: rainy convergence test to t = 2", flush=True)
    (uh, nh), (ur, nr), _ = both_loops(make, u0, 2.0, 1.0, 16)
    assert nh == nr
    np.testing.assert_array_equal(ur, uh)
    l2, linf = calc_error_norms_2d(ur, mesh, exact, 2.0)
    np.testing.assert_allclose(l2, CONV_L2, rtol=1e-7)
    np.testing.assert_allclose(linf, CONV_LINF, rtol=1e-6)


@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan"])
def test_euler3d_step_graph_matches_launches_one_by_one(precision):
    """atmo_step replays pairs of 3D steps from a captured graph whose stages
    read dt from device memory: the same bits as steps launched one by one,
    across odd step counts (which swap the state buffers) and changes of dt"""
    print(f"case: 3D shell 2x2 trees x 2 layers polydeg 3 {precision}", flush=True)
    make, u0 = euler3d(W.shell, trees_per_face=2, layers=2, polydeg=3, seed=4,
                       precision=precision)
    semi = make()
    semi.set_state(u0)
    dt = semi.max_dt(0.5)
    plan = [(dt, 2), (dt, 3), (0.7 * dt, 2), (dt, 4)]
    for d, k in plan:
        semi.step(d, k)
    a = semi.get_state()
    semi.set_state(u0)
    for d, k in plan:
        for _ in range(k):
            semi.step(d, 1)
    np.testing.assert_array_equal(a, semi.get_state())


@pytest.mark.parametrize("kind", ["euler3d", "covariant", "rainy"])
def test_a_nan_state_is_reported_not_stepped(kind):
    """a wave speed that is not a number makes max_dt and the resident loop
    fail with ATMO_ERR_NONFINITE (the reference's max_dt propagates NaN
    through Julia's max and the integrator stops)"""
    print(f"case: NaN in the state, {kind}", flush=True)
    if kind == "euler3d":
        make, u0 = euler3d(W.periodic_cube, trees=2, polydeg=3, seed=1)
    elif kind == "covariant":
        case = W.zonal_jet(3, 3, seed=1)
        make, u0 = case.semi, case.u0()
    else:
        make, u0 = rainy_make()
    flat = np.array(u0, copy=True).ravel(order="F")
    flat[u0.shape[0] * 7 + 1] = np.nan  # the second variable of node 7
    u = flat.reshape(u0.shape, order="F")
    semi = make()
    semi.set_state(u)
    with pytest.raises(T.AtmoError, match="NONFINITE"):
        semi.max_dt(0.5)
    semi.set_state(u)
    semi.set_time(0.0)
    with pytest.raises(T.AtmoError, match="NONFINITE"):
        semi.solve_resident(1.0, 0.5)
