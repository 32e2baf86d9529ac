"""The 3D form with FP32 geometry storage (precision "fp32_geometry",
ATMO_PREC_FP32_GEOMETRY): the stage kernel reads the contravariant vectors
from an FP32 copy (rounded once, to nearest, when the handle is made) and
widens each value where it is staged; states, fluxes, sums, the RK registers,
1/J and every reduction stay FP64.

Its errors are stated against the FP64 oracle, not hidden in a tolerance:
- rhs: <= 2e-7 relative per conserved variable on the scale of the terms
  (the oracle's magnitude mode): the rounding of the vectors (6e-8) enters
  the fluxes linearly, times the few terms that cancel;
- mass over 50 RK steps: conserved to FP64 round-off (<= 1e-13): an
  interface uses the primary side's FP32 normal on both sides, and a volume
  pair the average of the two nodes' FP32 vectors, which is symmetric;
- entropy rate of the entropy-conservative shell on a state of one
  temperature: the metric identities hold to the FP32 rounding only, so the
  rate is that rounding times the terms (<= 1e-6 of the term scale);
- the resident loop and both workgroup shapes reproduce the host loop bit
  for bit (the step size reads the FP64 vectors in both)."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import workloads as W
from oracle_lib import params_for

pytestmark = pytest.mark.gpu

GEO = "fp32_geometry"


def _case(name, precision):
    if name == "cube":
        return W.periodic_cube(trees=3, polydeg=3, seed=4, precision=precision)
    if name == "shell7_ec":
        out = W.shell(2, 2, polydeg=7, seed=6, precision=precision, volume_flux="flux_ranocha",
                      surface_flux="flux_ranocha")
    else:
        out = W.shell(3, 2, polydeg=4, seed=5, precision=precision)
    mesh, eq, solver, bcs, u0 = out
    u0 = u0.copy()
    u0[1:4] += 10.0 * np.sin(mesh.node_coordinates / 1e6) * u0[0]
    return mesh, eq, solver, bcs, u0


@pytest.mark.parametrize("name", ["cube", "shell_flow", "shell7_ec"])
def test_rhs_error_against_the_fp64_oracle(oracle, name):
    mesh, eq, solver, bcs, u0 = _case(name, GEO)
    print(f"case: {name} fp32 geometry, E={mesh.n_elements} n={mesh.nnodes}", flush=True)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    du = semi.rhs(None, u0).reshape(6, -1, order="F")
    p = params_for((mesh, eq, solver, 0.0))
    ref = oracle.rhs(p, mesh, u0).reshape(6, -1, order="F")
    mag = np.abs(oracle.rhs_magnitude(p, mesh, u0).reshape(6, -1, order="F")).max(axis=1)
    diff = np.abs(du - ref).max(axis=1)
    term = diff / np.maximum(mag, 1e-300)
    plain = diff / np.maximum(np.abs(ref).max(axis=1), 1e-300)
    print("  rel err (term scale):", term[:5], "\n  rel err (|du| scale):", plain[:5], flush=True)
    assert np.all(term[:5] <= 2e-7), term
    # phi: only the surface dissipation of rounding-level jumps between the
    # two sides' node coordinates
    assert np.abs(du[5] - ref[5]).max() <= 1e-7 * max(1.0, np.abs(ref).max())


def test_mass_conservation_over_steps():
    mesh, eq, solver, bcs, u0 = _case("shell_flow", GEO)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    semi.set_state(u0)
    i0 = semi.integrals(fresh_rhs=True)
    semi.step(semi.max_dt(0.5), 50)
    i1 = semi.integrals(fresh_rhs=True)
    drift = (i1[0] - i0[0]) / i0[0]
    print("mass drift over 50 steps:", drift, flush=True)
    assert abs(drift) <= 1e-13


def test_entropy_rate_of_the_ec_shell():
    """flux_ranocha + Waruszewski, no dissipation, slip walls, one
    temperature (where the FP64 form's rate is rounding, 1e-18 of the terms)"""
    from oracle_lib import Oracle
    from test_gpu_full_size import _cons2entropy, _isothermal_noisy
    mesh, eq, solver, bcs, u0 = W.shell(2, 2, polydeg=7, seed=6, precision=GEO,
                                        volume_flux="flux_ranocha", surface_flux="flux_ranocha")
    u = _isothermal_noisy(u0, eq)
    o = Oracle()
    p = params_for((mesh, eq, solver, 0.0))
    mag = o.rhs_magnitude(p, mesh, u).reshape(6, -1, order="F")
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    semi.set_state(u)
    rate = semi.integrals(fresh_rhs=True)[2]
    w = _cons2entropy(u.reshape(6, -1, order="F"), eq)
    _, wq, _, _ = T.lgl_basis(mesh.polydeg)
    wJ = ((wq[:, None, None] * wq[None, :, None] * wq[None, None, :])[..., None] /
          mesh.inverse_jacobian).ravel(order="F")
    scale = (wJ * (np.abs(w[:5]) * np.abs(mag[:5])).sum(axis=0)).sum()
    print(f"dS/dt {rate:.3e}, term scale {scale:.3e}, ratio {abs(rate) / scale:.2e}", flush=True)
    assert abs(rate) <= 1e-6 * scale


@pytest.mark.parametrize("name", ["shell_p4", "cube_p3"])
def test_shapes_and_resident_loop_are_bit_identical(name):
    if name == "shell_p4":
        mesh, eq, solver, bcs, u0 = _case("shell_flow", GEO)
    else:
        mesh, eq, solver, bcs, u0 = W.periodic_cube(trees=3, polydeg=3, seed=2, precision=GEO)
    print(f"case: {name} fp32 geometry, E={mesh.n_elements}", flush=True)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    semi.set_state(u0)
    chosen, ms = semi.autotune()
    print("autotune", chosen, ms[:2], flush=True)
    t_end = 6.5 * semi.max_dt(0.5)  # the last step is cut short
    out = []
    for v in (0, 1):
        for resident in (False, True):
            s = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
            s.set_kernel_variant(v)
            u, n = T.solve(s, u0, (0.0, t_end), 0.5, resident=resident, check_every=4)
            out.append((u, n))
    print("  steps", [n for _, n in out], flush=True)
    assert len({n for _, n in out}) == 1
    for u, _ in out[1:]:
        np.testing.assert_array_equal(u, out[0][0])
