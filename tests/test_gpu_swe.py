"""GPU parity of the covariant shallow-water path (SplitCovariantShallowWaterEquations2D
on the cubed sphere): the HIP rhs!, RK stages, max_dt and integrals against the
FP64 oracle on identical seeded states, the reference's regression values
integrated on the GPU, and ensembles of uneven members against the same
members run alone.

Tolerance: rhs <= 1e-12 relative per variable, the error divided by max over
nodes of sum_terms |term| (the oracle's magnitude mode), and <= 1e-12 relative
to max |du| where the tendency is resolved (the largest such error of these
cases is 2.9e-14); ensembles are bit-identical to the members run alone (the same
per-element arithmetic)."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import analysis
from trixiatmo_hip import workloads as W
from oracle_lib import SweOracle
from test_oracle_swe import DAY, PIN_RTOL, PINS

pytestmark = pytest.mark.gpu

RHS_RTOL = 1e-12


def perturbed(case, seed=5, amp=0.01):
    u0 = case.u0()
    rng = np.random.default_rng(seed)
    return u0 * (1.0 + amp * rng.standard_normal(u0.shape))


def check_rhs(oracle, case, u):
    semi = case.semi()
    print(f"case: E={semi.shape[-1]} n={semi.shape[1]} surface={case.solver.surface_flux} "
          f"volume={case.solver.volume_integral}", flush=True)
    so = SweOracle.for_semi(oracle, semi)
    du_gpu = semi.rhs(None, u)
    du_ref = so.rhs(u)
    mag = np.abs(so.rhs_magnitude(u).reshape(3, -1, order="F")).max(axis=1)
    diff = np.abs((du_gpu - du_ref).reshape(3, -1, order="F")).max(axis=1)
    err = diff / np.maximum(mag, 1e-300)
    plain = diff / np.maximum(np.abs(du_ref.reshape(3, -1, order="F")).max(axis=1), 1e-300)
    print("  rel err per variable (term scale):", err, " (|du| scale):", plain, flush=True)
    assert np.all(err <= RHS_RTOL), err
    # the |du| criterion applies where du is a resolved value, not rounding
    resolved = np.abs(du_ref.reshape(3, -1, order="F")).max(axis=1) > 1e-10 * mag
    assert np.all(plain[resolved] <= 1e-12), plain
    return semi, so, du_gpu, du_ref


@pytest.mark.parametrize("surface", [W.SWE_EC_LLF, "flux_lax_friedrichs", "flux_ec"])
@pytest.mark.parametrize("noncons", ["flux_nonconservative_surface_simplified",
                                     "flux_nonconservative_ec"])
def test_rhs_fluxes(oracle, surface, noncons):
    case = W.zonal_jet(4, 3, seed=1)
    case.solver = T.DGSEM(polydeg=3, surface_flux=(surface, noncons),
                          volume_integral=case.solver.volume_integral)
    check_rhs(oracle, case, perturbed(case))


@pytest.mark.parametrize("polydeg", [1, 2, 4, 5, 6, 7])
def test_rhs_polydeg(oracle, polydeg):
    case = W.zonal_jet(3, polydeg, seed=2)
    check_rhs(oracle, case, perturbed(case))


@pytest.mark.parametrize("polydeg", [2, 3, 5])
def test_rhs_standard_form_weak(oracle, polydeg):
    """CovariantShallowWaterEquations2D: weak form, its flux_lax_friedrichs and
    its geometric and Coriolis sources"""
    case = W.covariant_elixir("geostrophic_balance", cells=3, polydeg=polydeg)
    check_rhs(oracle, case, perturbed(case))


def test_rhs_central_volume_and_collocation_christoffel(oracle):
    case = W.covariant_elixir("rossby_haurwitz", cells=4,
                              volume_flux=("flux_central", "flux_nonconservative_ec"))
    check_rhs(oracle, case, perturbed(case))


def test_rhs_topography_lake_at_rest(oracle):
    case = W.covariant_elixir("well_balanced", cells=4)
    u0 = case.u0()
    semi, so, du_gpu, du_ref = check_rhs(oracle, case, u0)
    mag = so.rhs_magnitude(u0)
    assert np.abs(du_gpu).max() <= 1e-12 * mag.max()


def test_rhs_c2_single_member_size(oracle):
    """the single-member configuration: 6 x 32 x 32 elements, polydeg 3"""
    case = W.zonal_jet(32, 3, seed=0)
    check_rhs(oracle, case, perturbed(case))


def test_steps_max_dt_integrals(oracle):
    case = W.zonal_jet(6, 3, seed=3)
    semi = case.semi()
    so = SweOracle.for_semi(oracle, semi)
    u0 = perturbed(case, amp=0.002)
    semi.set_state(u0)
    dt_gpu = semi.max_dt(0.4)
    dt_ref = 0.4 * so.max_dt(u0)
    print("max_dt", dt_gpu, dt_ref, flush=True)
    assert abs(dt_gpu - dt_ref) <= 1e-14 * dt_ref
    semi.step(dt_ref, 10)
    u_gpu = semi.get_state()
    u_ref = so.ck2n54_steps(u0, dt_ref, 10)
    rel = np.abs((u_gpu - u_ref).reshape(3, -1, order="F")).max(axis=1) / \
        np.abs(u_ref.reshape(3, -1, order="F")).max(axis=1)
    print("10-step state rel err", rel, flush=True)
    assert np.all(rel <= 1e-13), rel
    ig = semi.integrals(fresh_rhs=True)
    ir = so.integrals(u_ref, so.rhs(u_ref))
    print("integrals", ig, ir, flush=True)
    np.testing.assert_allclose(ig[:2], ir[:2], rtol=1e-13)
    # dS/dt is a sum of cancelling terms: compare on the scale of the entropy rate terms
    assert abs(ig[2] - ir[2]) <= 1e-10 * abs(ir[1]) / 86400.0


@pytest.mark.parametrize("name,over,l2_ref,linf_ref", PINS, ids=[p[0] for p in PINS])
def test_covariant_elixir_regression_gpu(name, over, l2_ref, linf_ref):
    """the reference's one-day regression values, integrated by the HIP
    stage kernels with the StepsizeCallback loop"""
    case = W.covariant_elixir(name, **over)
    semi = case.semi()
    u0 = case.u0()
    u, steps = T.solve(semi, u0, (0.0, DAY), case.cfl)
    l2, linf = analysis.calc_error_norms_covariant(u, case.mesh, semi.aux_node_vars,
                                                   case.initial_condition, DAY)
    print(name, steps, "steps", l2, linf, flush=True)
    np.testing.assert_allclose(l2, l2_ref, rtol=PIN_RTOL)
    np.testing.assert_allclose(linf, linf_ref, rtol=PIN_RTOL)


def test_well_balanced_day_gpu():
    case = W.covariant_elixir("well_balanced")
    semi = case.semi()
    u0 = case.u0()
    u, _ = T.solve(semi, u0, (0.0, DAY), case.cfl)
    l2, linf = analysis.calc_error_norms_covariant(u, case.mesh, semi.aux_node_vars,
                                                   case.initial_condition, DAY)
    np.testing.assert_allclose(l2, 0.0, atol=1e-11)
    np.testing.assert_allclose(linf, 0.0, atol=1e-11)


def test_ensemble_matches_members():
    """uneven members in one handle: rhs, per-member max_dt and per-member
    time steps reproduce each member run alone, bit for bit"""
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=5, polydeg=3, seed=4,
                                                    choices=(2, 3, 5))
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    u0 = ens.compute_coefficients()
    print("ensemble elements", ens.shape[-1], "members", ens.n_members, flush=True)
    du = ens.rhs(None, u0).reshape(ens.shape, order="F")
    ens.set_state(u0)
    dts = ens.max_dt_members(0.4)
    ens.step_members(dts, 3)
    u_ens = ens.get_state().reshape(ens.shape, order="F")
    first = ens.member_first_element
    for m, (mesh, ic, _) in enumerate(members):
        single = T.SemidiscretizationHyperbolic(mesh, eq, solver, source_terms=src,
                                                initial_condition=ic)
        um = single.compute_coefficients()
        sl = slice(first[m], first[m + 1])
        np.testing.assert_array_equal(um, u0[..., sl])
        np.testing.assert_array_equal(single.rhs(None, um).reshape(single.shape, order="F"),
                                      du[..., sl])
        single.set_state(um)
        assert single.max_dt(0.4) == dts[m]
        single.step(dts[m], 3)
        np.testing.assert_array_equal(single.get_state().reshape(single.shape, order="F"),
                                      u_ens[..., sl])
    # the generic max_dt of an ensemble handle is the smallest member step
    assert ens.max_dt(0.4) == pytest.approx(ens.max_dt_members(0.4).min(), rel=0, abs=0)


def test_ensemble_conserves_mass_per_member():
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=4, polydeg=3, seed=7,
                                                    choices=(2, 4))
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    u0 = ens.compute_coefficients()
    ens.set_state(u0)
    ens.step_members(ens.max_dt_members(0.4), 20)
    u = ens.get_state().reshape(ens.shape, order="F")
    _, w, _, _ = T.lgl_basis(3)
    wJ = w[:, None, None] * w[None, :, None] * ens.aux_node_vars[12]
    first = ens.member_first_element
    for m in range(ens.n_members):
        sl = slice(first[m], first[m + 1])
        m0 = (wJ[..., sl] * u0[0][..., sl]).sum()
        m1 = (wJ[..., sl] * u[0][..., sl]).sum()
        assert abs(m1 - m0) <= 1e-13 * m0


def test_autotuner_shapes_and_graph_replay_are_bit_identical():
    """every workgroup shape gives the same bits, and steps replayed from the
    captured graph equal steps launched one by one"""
    case = W.zonal_jet(6, 3, seed=9)
    semi = case.semi()
    u0 = perturbed(case, amp=0.002)
    semi.set_state(u0)
    chosen, ms = semi.autotune()
    print("autotune", chosen, ms, flush=True)
    assert 0 <= chosen < 6 and np.all(ms > 0) and ms[chosen] == ms.min()
    outs = []
    for v in range(6):
        semi.set_kernel_variant(v)
        outs.append(semi.rhs(None, u0))
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    semi.set_kernel_variant(chosen)
    dt = semi.max_dt(0.4)
    semi.set_state(u0)
    semi.step(dt, 4)            # two graph replays
    a = semi.get_state()
    semi.set_state(u0)
    for _ in range(4):
        semi.step(dt, 1)        # launched one by one
    b = semi.get_state()
    np.testing.assert_array_equal(a, b)
    for v in (3, 5):            # persistent grids inside the graph as well
        semi.set_kernel_variant(v)
        semi.set_state(u0)
        semi.step(dt, 4)
        np.testing.assert_array_equal(semi.get_state(), a)


def test_inconsistent_metric_terms_are_refused(monkeypatch):
    """the stage recomputes J = sqrt(det G) and G^-1 from G: auxiliary
    variables given by the caller whose J or G^-1 disagree with their G are
    refused at creation"""
    from trixiatmo_hip import semidiscretization as S
    case = W.zonal_jet(2, 3, seed=0)
    good = S.covariant_aux

    def host_aux_semi():
        return S.SemidiscretizationHyperbolic(case.mesh, case.equations, case.solver,
                                              source_terms=case.source_terms,
                                              initial_condition=case.initial_condition,
                                              metric_terms=case.metric_terms,
                                              metric_terms_on_device=False)
    for row in (12, 16, 17, 18):
        def bad(mesh, metric_terms=None, auxiliary_field=None, row=row):
            aux = good(mesh, metric_terms, auxiliary_field)
            aux[row, 1, 2, 5] += 1e-6 * abs(aux[16 if row > 12 else 12, 1, 2, 5])
            return aux
        monkeypatch.setattr(S, "covariant_aux", bad)
        with pytest.raises(T.AtmoError, match="sqrt\\(det G\\)"):
            host_aux_semi()
    monkeypatch.setattr(S, "covariant_aux", good)
    host_aux_semi()
