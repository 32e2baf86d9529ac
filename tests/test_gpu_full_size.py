"""GPU parity at the full sizes of the benchmarked configurations: one HIP
rhs! against the FP64 oracle on the same state for the 3D shell 6 x 16^2 x 8
at polydeg 4 (1,536,000 nodes), the precision-stress shell 6 x 8^2 x 4 at
polydeg 7 (786,432 nodes), the rainy bubble 256 x 128 at polydeg 3 (524,288
nodes), the covariant single member 6 x 32^2 and the 32-member covariant
ensemble (1.69 M nodes, every member against the oracle run on that member
alone); and the semi-discrete entropy balance of the entropy-conservative
precision-stress configuration in its three forms.

Errors are reported two ways per conserved variable: against the scale of the
summed terms (max over nodes of sum_terms |term|, the oracle's magnitude
mode: a near-hydrostatic tendency is a small residual of large cancelling
terms, and rounding scales with those terms) and against max |du| itself.
Bars (FP64): <= 1e-12 on the term scale; on the plain |du| scale <= 1e-10 for
the 3D and covariant paths and <= 1e-11 for the rainy path, about ten times
the largest error these comparisons show (printed), so that a regression of a
few ulps in a cancelling sum shows up."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import workloads as W
from oracle_lib import Oracle, RainyOracle, SweOracle, params_for

pytestmark = pytest.mark.gpu

TERM_RTOL = 1e-12
PLAIN_RTOL_3D = 1e-10
PLAIN_RTOL_SWE = 1e-10
PLAIN_RTOL_RAINY = 1e-11


def errors(du, ref, mag, nv):
    du, ref = du.reshape(nv, -1, order="F"), ref.reshape(nv, -1, order="F")
    diff = np.abs(du - ref).max(axis=1)
    term = diff / np.maximum(np.abs(mag.reshape(nv, -1, order="F")).max(axis=1), 1e-300)
    plain = diff / np.maximum(np.abs(ref).max(axis=1), 1e-300)
    return term, plain


def evolved(semi, u0, steps=20, cfl=0.5):
    """the state after `steps` CK2N54 steps on the GPU: the workloads start at
    rest, where the mass and energy tendencies are rounding noise (every
    volume and gravity energy flux vanishes with the velocity); after a few
    steps the temperature bumps have set the air in motion"""
    semi.set_state(u0)
    semi.step(semi.max_dt(cfl), steps)
    return semi.get_state().reshape(u0.shape, order="F")


def check_3d(oracle, mesh, eq, solver, bcs, u0, label):
    print(f"case: {label}, E={mesh.n_elements} n={mesh.nnodes}", flush=True)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    u0 = evolved(semi, u0)
    du = semi.rhs(None, u0)
    p = params_for((mesh, eq, solver, 0.0))
    ref = oracle.rhs(p, mesh, u0)
    mag = oracle.rhs_magnitude(p, mesh, u0)
    term, plain = errors(du, ref, mag, 6)
    print(f"  {label} rel err (term scale): {term[:5]}\n  {label} rel err (|du| scale): "
          f"{plain[:5]}", flush=True)
    assert np.all(term[:5] <= TERM_RTOL), term
    assert np.all(plain[:5] <= PLAIN_RTOL_3D), plain
    # phi has no flux, but the LLF dissipation acts on its rounding-level jumps
    # across faces (as in the reference); that row is compared absolutely
    d5 = np.abs(du.reshape(6, -1, order="F")[5] - ref.reshape(6, -1, order="F")[5]).max()
    assert d5 <= 1e-12 * max(1.0, np.abs(ref).max()), d5
    return semi, du, ref, mag


def test_c3_shell_full_size(oracle):
    """C3: 6 x 16^2 x 8, polydeg 4, KG + Waruszewski volume, KG + LLF +
    Waruszewski surface, slip walls (the benchmarked FP64 workload)"""
    check_3d(oracle, *W.shell(16, 8, polydeg=4, seed=0), "C3 shell 6x16^2x8 p4")


# errors of the reduced-precision 3D forms on C3's evolved state against the
# FP64 oracle, per conserved variable: on the scale of the terms, and on the
# scale of max |du|; bars about three times the measured values (printed):
# mixed 7.9e-6 / 1.5e-10 .. 2.5e-10 / 8.0e-6 of the term scale and 3.8e-4 /
# 1.1e-5 .. 1.7e-5 / 3.9e-4 of max |du|; FP32 geometry 1.4e-8 .. 1.9e-8 and
# 6.8e-7 / 1.3e-3 .. 1.4e-3 / 6.8e-7
REDUCED_BARS = {
    "fp32_kahan": {"term": [5e-7, 1e-9, 1e-9, 1e-9, 5e-7],
                   "plain": [3e-5, 5e-5, 5e-5, 5e-5, 3e-5]},
    "fp32_geometry": {"term": [5e-8] * 5, "plain": [2e-6, 4e-3, 4e-3, 4e-3, 2e-6]},
}


@pytest.mark.parametrize("precision", ["fp32_kahan", "fp32_geometry"])
def test_c3_reduced_precision_on_the_evolved_state(oracle, precision):
    """The mixed form (FP32 fluxes, Kahan sums) and the FP32-geometry form on
    the benchmarked shell after 20 FP64 steps.  Near hydrostatic balance the
    tendencies are small residuals of much larger terms, so the FP32 rounding
    of those terms (or of the metric identities) is a visible fraction of
    them.  The mixed form takes the local Lax-Friedrichs jumps from the FP64
    states (rounded after the subtraction): its mass and energy rows carry
    about 1e-7 of the term scale (8e-6 when the jumps were differences of
    FP32-rounded states), its momentum rows (background split in FP64)
    2e-10, both near 1e-5 of max |du|.  The FP32-geometry form is at the FP32
    rounding of the metric terms, 1.4e-8 to 1.9e-8 of the term scale, in
    every row, which is 1e-3 of the momentum rows' max |du|."""
    mesh, eq, solver, bcs, u0 = W.shell(16, 8, polydeg=4, seed=0)
    print(f"case: C3 {precision} on the evolved state", flush=True)
    u = evolved(T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs), u0)
    p = params_for((mesh, eq, solver, 0.0))
    ref = oracle.rhs(p, mesh, u)
    mag = oracle.rhs_magnitude(p, mesh, u)
    mesh, eq, solver, bcs, _ = W.shell(16, 8, polydeg=4, seed=0, precision=precision)
    du = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs).rhs(None, u)
    term, plain = errors(du, ref, mag, 6)
    print(f"  {precision} rel err (term scale): {np.array2string(term[:5], max_line_width=200)}\n"
          f"  {precision} rel err (|du| scale): {np.array2string(plain[:5], max_line_width=200)}",
          flush=True)
    bars = REDUCED_BARS[precision]
    assert np.all(term[:5] <= np.asarray(bars["term"])), term
    assert np.all(plain[:5] <= np.asarray(bars["plain"])), plain


def test_c4_precision_stress_full_size(oracle):
    """C4: 6 x 8^2 x 4, polydeg 7, flux_ranocha + Waruszewski in the volume
    and on the surface"""
    check_3d(oracle, *W.precision_stress("fp64"), "C4 shell 6x8^2x4 p7")


def test_c5_rainy_full_size(oracle):
    """C5: 256 x 128 cells, polydeg 3, flux_ec_rain volume, LLF surface,
    source_terms_rainy, slip walls in y; the state after the saturation
    solve (the stage limiter's output, as every rhs! of a run sees it)"""
    mesh, eq, solver, src, lim, u0, info = W.rainy_bubble(seed=0)
    print(f"case: C5 rainy 256x128 p3, RH {info['relative_humidity']}", flush=True)
    semi = T.SemidiscretizationRainy(mesh, eq, solver, source_terms=src, stage_limiter=lim)
    ro = RainyOracle(oracle, mesh, volume_integral=1, source=src)
    u, _, _ = ro.newton(u0)
    u = u.reshape(u0.shape, order="F")
    # winds so that the tendencies are not only a hydrostatic residual
    rng = np.random.default_rng(11)
    u = u.copy()
    u[3] += 2.0 * rng.standard_normal(u[3].shape) * u[0]
    u[4] += 2.0 * rng.standard_normal(u[4].shape) * u[0]
    du = semi.rhs(None, u)
    ref = ro.rhs(u)
    d = np.abs((du - ref).reshape(9, -1, order="F")).max(axis=1)
    plain = d / np.maximum(np.abs(ref.reshape(9, -1, order="F")).max(axis=1), 1e-300)
    print(f"  C5 rel err (|du| scale): {plain}", flush=True)
    # rows 7-9 (rho_v, rho_c, T) change only by the LLF dissipation of the
    # surface flux (the saturation solve resets them in every stage)
    assert np.all(plain <= PLAIN_RTOL_RAINY), plain


def _swe_check(oracle, semi, aux, mesh, eq, solver, src, u, du):
    so = SweOracle.from_parts(oracle, mesh, aux, eq, solver, src)
    ref = so.rhs(u)
    mag = so.rhs_magnitude(u)
    return errors(du, ref, mag, 3)


def test_c2_single_member_full_size(oracle):
    case = W.zonal_jet(32, 3, seed=0)
    print("case: C2 single member 6x32^2 p3", flush=True)
    semi = case.semi()
    u0 = case.u0()
    rng = np.random.default_rng(5)
    u = u0 * (1.0 + 0.01 * rng.standard_normal(u0.shape))
    du = semi.rhs(None, u)
    term, plain = _swe_check(oracle, semi, semi.aux_node_vars, case.mesh, case.equations,
                             case.solver, case.source_terms, u, du)
    print(f"  C2 single rel err (term scale): {term}\n  C2 single rel err (|du| scale): "
          f"{plain}", flush=True)
    assert np.all(term <= TERM_RTOL) and np.all(plain <= PLAIN_RTOL_SWE), (term, plain)


def test_c2_ensemble_full_size(oracle):
    """the 32-member ensemble in one handle; every member's slice of one
    ensemble rhs! against the oracle on that member alone"""
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=32, polydeg=3, seed=0)
    ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src)
    u0 = ens.compute_coefficients()
    rng = np.random.default_rng(6)
    u = u0 * (1.0 + 0.01 * rng.standard_normal(u0.shape))
    print(f"case: C2 ensemble, {ens.n_members} members, {ens.shape[-1]} elements", flush=True)
    du = ens.rhs(None, u).reshape(ens.shape, order="F")
    first = ens.member_first_element
    worst_t, worst_p = np.zeros(3), np.zeros(3)
    for m, (mesh, ic, _) in enumerate(members):
        sl = slice(first[m], first[m + 1])
        term, plain = _swe_check(oracle, ens, ens.aux_node_vars[..., sl], mesh, eq, solver, src,
                                 u[..., sl], du[..., sl])
        worst_t, worst_p = np.maximum(worst_t, term), np.maximum(worst_p, plain)
    print(f"  C2 ensemble worst member rel err (term scale): {worst_t}\n"
          f"  C2 ensemble worst member rel err (|du| scale): {worst_p}", flush=True)
    assert np.all(worst_t <= TERM_RTOL) and np.all(worst_p <= PLAIN_RTOL_SWE), (worst_t, worst_p)


# semi-discrete entropy balance of the entropy-conservative C4 configuration:
# |dS/dt| against the scale of its terms, sum_nodes w J sum_v |w_v| T_v with
# T_v the sum of the magnitudes of the terms of du_v (the oracle's magnitude
# mode): on a near-balanced state du is a small residual of large terms and
# its rounding, summed over 786,432 nodes, scales with those terms
EC_BOUND = {"fp64": 1e-12, "fp32_kahan": 1e-9, "fp32": 1e-9}


def _isothermal_noisy(u0, eq, seed=5):
    """u0's density times 1 + 1e-3 N(0, 1), a random 1 m/s flow, 250 K everywhere"""
    rng = np.random.default_rng(seed)
    uf = u0.reshape(6, -1, order="F").copy()
    n = uf.shape[1]
    rho = uf[0] * (1.0 + 1e-3 * rng.standard_normal(n))
    v = rng.standard_normal((3, n))
    uf[0] = rho
    uf[1:4] = rho * v
    uf[4] = rho * (eq.c_p - eq.c_v) * 250.0 / (eq.gamma - 1.0) + 0.5 * rho * (v ** 2).sum(0)
    return uf.reshape(u0.shape, order="F")


def _entropy_rate(oracle, precision, u):
    """(GPU dS/dt from the integrals, host sum of w . du over the GPU du,
    oracle dS/dt, term scale) on the state u of the precision-stress shell"""
    mesh, eq, solver, bcs, _ = W.precision_stress("fp64")
    p = params_for((mesh, eq, solver, 0.0))
    mag = oracle.rhs_magnitude(p, mesh, u).reshape(6, -1, order="F")
    ref = oracle.integrals(p, mesh, u, oracle.rhs(p, mesh, u))[2]
    mesh, eq, solver, bcs, _ = W.precision_stress(precision)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    semi.set_state(u)
    ints = semi.integrals(fresh_rhs=True)
    du = semi.rhs(None, u).reshape(6, -1, order="F")
    uf = u.reshape(6, -1, order="F")
    w = _cons2entropy(uf, eq)
    _, wq, _, _ = T.lgl_basis(mesh.polydeg)
    W3 = (wq[:, None, None] * wq[None, :, None] * wq[None, None, :])[..., None] / \
        mesh.inverse_jacobian
    wJ = W3.ravel(order="F")
    term_scale = (wJ * (np.abs(w[:5]) * np.abs(mag[:5])).sum(axis=0)).sum()
    rate = (wJ * (w[:5] * du[:5]).sum(axis=0)).sum()
    print(f"  dS/dt {ints[2]:.3e} (host sum {rate:.3e}, oracle {ref:.3e}); term scale "
          f"{term_scale:.3e}: ratios {abs(ints[2]) / term_scale:.2e}, oracle "
          f"{abs(ref) / term_scale:.2e}", flush=True)
    assert term_scale > 0.0
    return ints[2], rate, ref, term_scale


@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan", "fp32"])
def test_c4_entropy_conservation_isothermal(oracle, precision):
    """flux_ranocha + Waruszewski in the volume and on the surface, no
    dissipation, slip walls: the semi-discrete entropy rate
    dS/dt = int w(u) . du (analysis_covariant.jl:88-105 pattern) vanishes up
    to rounding on any state of one temperature (the gravity pair term
    cancels in w . du only between nodes of equal rho / p; the oracle's
    test_entropy_balance_of_the_gravity_term_needs_equal_temperatures shows
    both sides of that).  The state: noisy density, a random flow, 250 K.
    FP64 to 1e-12 of the scale of its terms; the FP32 flux forms to 1e-9
    (their FP32 rounding enters only the perturbation from the FP64
    hydrostatic background)."""
    mesh, eq, solver, bcs, u0 = W.precision_stress("fp64")
    print(f"case: C4 entropy rate, isothermal state, {precision}", flush=True)
    gpu, host, ref, scale = _entropy_rate(oracle, precision, _isothermal_noisy(u0, eq))
    assert abs(ref) <= 1e-13 * scale
    assert abs(gpu) <= EC_BOUND[precision] * scale
    assert abs(host) <= EC_BOUND[precision] * scale


@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan", "fp32"])
def test_c4_entropy_rate_evolved_matches_oracle(oracle, precision):
    """on the FP64 state after 20 steps the temperature varies, and the
    scheme's own entropy rate is no longer rounding (the oracle's is about
    4.5e-11 of the term scale: the gravity term's imbalance between nodes of
    unequal temperature); the GPU's rate equals the oracle's to 1e-12 of the
    term scale in FP64 and to 1e-9 in the FP32 forms"""
    mesh, eq, solver, bcs, u0 = W.precision_stress("fp64")
    print(f"case: C4 entropy rate, evolved state, {precision}", flush=True)
    u = evolved(T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs), u0)
    gpu, host, ref, scale = _entropy_rate(oracle, precision, u)
    assert abs(gpu - ref) <= EC_BOUND[precision] * scale
    assert abs(host - ref) <= EC_BOUND[precision] * scale


def _cons2entropy(u, eq):
    """entropy variables of compressible_euler_energy_with_gravity_3d.jl:581-601
    (w6 = 0), vectorised; checked against the oracle's cons2entropy on a
    sample in test_cons2entropy_vectorised"""
    g = eq.gamma
    rho = u[0]
    v = u[1:4] / rho
    vsq = (v * v).sum(axis=0)
    p = (g - 1) * (u[4] - 0.5 * rho * vsq)
    s = np.log(p) - g * np.log(rho)
    rp = rho / p
    return np.stack([(g - s) / (g - 1) - 0.5 * rp * vsq, rp * v[0], rp * v[1], rp * v[2], -rp,
                     np.zeros_like(rho)])


def test_cons2entropy_vectorised(oracle):
    mesh, eq, solver, bcs, u0 = W.shell(2, 2, polydeg=3, seed=1)
    u0 = u0.copy()
    u0[1:4] += 10.0 * np.sin(mesh.node_coordinates / 1e6) * u0[0]
    p = params_for((mesh, eq, solver, 0.0))
    uf = u0.reshape(6, -1, order="F")
    w = _cons2entropy(uf, eq)
    for i in range(0, uf.shape[1], 97):
        np.testing.assert_allclose(w[:, i], oracle.cons2entropy(p, np.ascontiguousarray(uf[:, i])),
                                   rtol=1e-13,
                                   atol=1e-13 * np.abs(w[:, i]).max())
