"""The covariant form with FP32 geometry storage (precision "fp32_geometry",
ATMO_PREC_FP32_GEOMETRY): the stage kernel reads its metric terms, Christoffel
symbols, z / r and face maps as FP32 values (rounded once, to nearest, when
the handle is made) and does every operation in FP64; states, RK registers
and the analysis integrals stay FP64.

Its errors are stated against the FP64 oracle, not hidden in a tolerance:
- rhs: <= 1e-6 relative per variable on the scale of the terms (the oracle's
  magnitude mode), the rounding of the geometry entering linearly;
- state after 100 RK steps of the zonal jet: <= 1e-5 of the change of the
  state over the run;
- mass over the 100 steps: the quadrature with the Jacobian the scheme
  computes with (the area element rounded to FP32) drifts about 1e-13
  relative (<= 1e-12): the two sides of a face rotate the neighbour's momentum with
  independently rounded FP32 maps, so their mass fluxes agree to FP32
  rounding, not FP64; the analysis integral (FP64 Jacobian) drifts 1.5e-12
  (<= 1e-11).  The FP64 form conserves both to
  round-off.
The resident loop and ensembles keep their exact bar: the fused speeds read
the same FP32 values as the max-speed reduction, so the resident loop
reproduces the host loop bit for bit, and ensemble members reproduce the
members run alone."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import workloads as W
from oracle_lib import SweOracle

GEO = "fp32_geometry"


def perturbed(case, seed=5, amp=0.01):
    u0 = case.u0()
    rng = np.random.default_rng(seed)
    return u0 * (1.0 + amp * rng.standard_normal(u0.shape))


def test_covariant_precisions_accepted():
    """the covariant form takes fp64 and fp32_geometry; the flux-arithmetic
    forms of the 3D path are refused"""
    from trixiatmo_hip.semidiscretization import _swe_config
    case = W.zonal_jet(2, 3, precision=GEO)
    cfg = _swe_config(case.equations, case.solver, 24, case.source_terms)
    assert cfg.precision == 3
    for p in ("fp32_kahan", "fp32"):
        bad = W.zonal_jet(2, 3, precision=p)
        with pytest.raises(ValueError):
            _swe_config(bad.equations, bad.solver, 24, bad.source_terms)


@pytest.mark.gpu
def test_rhs_error_against_the_fp64_oracle(oracle):
    print("case: zonal jet 8x8 trees polydeg 3, fp32 geometry, perturbed", flush=True)
    case64 = W.zonal_jet(8, 3, seed=1)
    case32 = W.zonal_jet(8, 3, seed=1, precision=GEO)
    u = perturbed(case64)
    s64, s32 = case64.semi(), case32.semi()
    so = SweOracle.for_semi(oracle, s64)
    du_ref = so.rhs(u)
    mag = np.abs(so.rhs_magnitude(u).reshape(3, -1, order="F")).max(axis=1)
    du32 = s32.rhs(None, u)
    du64 = s64.rhs(None, u)
    err = np.abs((du32 - du_ref).reshape(3, -1, order="F")).max(axis=1) / mag
    plain = np.abs((du32 - du_ref).reshape(3, -1, order="F")).max(axis=1) / \
        np.abs(du_ref.reshape(3, -1, order="F")).max(axis=1)
    err64 = np.abs((du64 - du_ref).reshape(3, -1, order="F")).max(axis=1) / mag
    print("  fp32 geometry rel err (term scale):", err, " (|du| scale):", plain,
          " fp64 (term scale):", err64, flush=True)
    assert np.all(err <= 1e-6), err
    assert np.all(err64 <= 1e-12), err64
    # the rounding is visible: the two storage forms differ
    assert np.any(du32 != du64)


@pytest.mark.gpu
def test_state_after_100_steps_and_mass():
    print("case: zonal jet 6x6 trees polydeg 3, 100 steps, fp32 geometry vs fp64", flush=True)
    case64 = W.zonal_jet(6, 3, seed=2)
    case32 = W.zonal_jet(6, 3, seed=2, precision=GEO)
    u0 = case64.u0()
    s64, s32 = case64.semi(), case32.semi()
    s64.set_state(u0)
    dt = s64.max_dt(0.4)
    s32.set_state(u0)
    m0 = s32.integrals()[0]
    _, w, _, _ = T.lgl_basis(3)
    J32 = s32.aux_node_vars[12].astype(np.float32).astype(np.float64)
    wJ32 = w[:, None, None] * w[None, :, None] * J32
    own0 = (wJ32 * u0[0]).sum()
    s64.step(dt, 100)
    s32.step(dt, 100)
    a = s64.get_state().reshape(u0.shape, order="F")
    b = s32.get_state().reshape(u0.shape, order="F")
    change = np.abs((a - u0).reshape(3, -1, order="F")).max(axis=1)
    rel = np.abs((b - a).reshape(3, -1, order="F")).max(axis=1) / change
    drift = abs(s32.integrals()[0] - m0) / m0
    own = abs((wJ32 * b[0]).sum() - own0) / own0
    print("  100-step state err / state change:", rel, " mass drift (scheme's Jacobian):", own,
          " (FP64 Jacobian):", drift, flush=True)
    assert np.all(rel <= 1e-5), rel
    assert own <= 1e-12, own
    assert drift <= 1e-11, drift


@pytest.mark.gpu
def test_resident_loop_matches_the_host_loop():
    print("case: zonal jet 3x3 trees polydeg 3, fp32 geometry, resident vs host", flush=True)
    case = W.zonal_jet(3, 3, seed=3, precision=GEO)
    u0 = case.u0()
    s = case.semi()
    s.set_state(u0)
    t_end = 20.5 * s.max_dt(0.4)
    uh, nh = T.solve(case.semi(), u0, (0.0, t_end), 0.4)
    ur, nr = T.solve(case.semi(), u0, (0.0, t_end), 0.4, resident=True, check_every=8)
    print(f"  host {nh} steps, resident {nr} steps", flush=True)
    assert nh == nr
    np.testing.assert_array_equal(ur, uh)


@pytest.mark.gpu
def test_ensemble_matches_members():
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=4, polydeg=3, seed=4,
                                                    choices=(2, 3, 5), precision=GEO)
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    print("case: fp32-geometry ensemble elements", ens.shape[-1], flush=True)
    u0 = ens.compute_coefficients()
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
        np.testing.assert_array_equal(single.rhs(None, um).reshape(single.shape, order="F"),
                                      du[..., sl])
        single.set_state(um)
        assert single.max_dt(0.4) == dts[m]
        single.step(dts[m], 3)
        np.testing.assert_array_equal(single.get_state().reshape(single.shape, order="F"),
                                      u_ens[..., sl])


@pytest.mark.gpu
def test_resident_ensemble_members_take_their_own_steps():
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=3, polydeg=3, seed=8,
                                                    choices=(2, 3), precision=GEO)
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    u0 = ens.compute_coefficients()
    ens.set_state(u0)
    t_end = 8.5 * ens.max_dt_members(0.4).max()
    print(f"case: fp32-geometry resident ensemble of {ens.n_members}, t_end {t_end}", flush=True)
    u_ens, steps = T.solve(ens, u0, (0.0, t_end), 0.4, resident=True, check_every=4)
    first = ens.member_first_element
    for m, (mesh, icm, _) in enumerate(members):
        single = T.SemidiscretizationHyperbolic(mesh, eq, solver, source_terms=src,
                                                initial_condition=icm)
        um, n = T.solve(single, single.compute_coefficients(), (0.0, t_end), 0.4)
        assert steps[m] == n
        np.testing.assert_array_equal(u_ens[..., first[m]:first[m + 1]], um)


def _rhs_err(oracle, case, u):
    semi = case.semi()
    so = SweOracle.for_semi(oracle, semi)
    du_ref = so.rhs(u)
    mag = np.abs(so.rhs_magnitude(u).reshape(3, -1, order="F")).max(axis=1)
    du = semi.rhs(None, u)
    return np.abs((du - du_ref).reshape(3, -1, order="F")).max(axis=1) / mag


@pytest.mark.gpu
@pytest.mark.parametrize("polydeg", [1, 2, 4, 5, 6, 7])
def test_rhs_error_every_polydeg(oracle, polydeg):
    case = W.zonal_jet(3, polydeg, seed=2, precision=GEO)
    err = _rhs_err(oracle, case, perturbed(case))
    print(f"  polydeg {polydeg}: fp32 geometry rel err (term scale)", err, flush=True)
    assert np.all(err <= 1e-6), err


@pytest.mark.gpu
@pytest.mark.parametrize("name,over", [
    ("geostrophic_balance", {}),                        # standard form, weak form
    ("rossby_haurwitz", {"volume_flux": ("flux_central", "flux_nonconservative_ec")}),
    ("well_balanced", {}),                              # bottom topography
])
def test_rhs_error_other_forms(oracle, name, over):
    case = W.covariant_elixir(name, cells=3, polydeg=3, precision=GEO, **over)
    err = _rhs_err(oracle, case, perturbed(case))
    print(f"  {name}: fp32 geometry rel err (term scale)", err, flush=True)
    assert np.all(err <= 1e-6), err


@pytest.mark.gpu
def test_lake_at_rest_stays_balanced(oracle):
    """bottom topography with FP32 geometry storage: h_s stays FP64 (a row of
    its own), so h + h_s = const is balanced as in the FP64 form: the rhs of
    the lake at rest (the well_balanced elixir) vanishes to 1e-12 of the
    scale of its terms (with h_s rounded to FP32 it was 1.5e-9)"""
    print("case: well_balanced lake at rest, fp32 geometry", flush=True)
    case = W.covariant_elixir("well_balanced", cells=4, polydeg=3, precision=GEO)
    semi = case.semi()
    u0 = case.u0()
    du = semi.rhs(None, u0)
    so = SweOracle.for_semi(oracle, semi)
    mag = np.abs(so.rhs_magnitude(u0)).max()
    print(f"  max |du| / term scale {np.abs(du).max() / mag:.2e}", flush=True)
    assert np.abs(du).max() <= 1e-12 * mag
