"""GPU: SaveSolutionCallback output variables and Trixi-compatible files.

The output variables are computed on the device from the resident state
(atmo_solution_variables): the covariant cons2prim_and_vorticity against the
oracle (itself pinned by the closed-form vorticity of Williamson cases 2 and
6) within 1e-13 of each variable's scale; the 3D and rainy cons2prim against
the primitive values a state was built from.  Solution and restart files are
written from the device output and read back; a run resumed from a restart
file continues bit for bit like the run that wrote it."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import solution_files as SF
from trixiatmo_hip import workloads as W
from oracle_lib import SweOracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["barotropic_instability", "isolated_mountain"])
def test_cons2prim_and_vorticity_matches_oracle(oracle, name):
    print(f"case: covariant {name} 4x4 per face polydeg 3, perturbed", flush=True)
    case = W.covariant_elixir(name, cells=4, polydeg=3)
    semi = case.semi()
    rng = np.random.default_rng(1)
    u = case.u0()
    u = u * (1.0 + 0.01 * rng.standard_normal(u.shape))
    semi.set_state(u)
    out, names = SF.solution_variables(semi, "cons2prim_and_vorticity")
    assert names == ("H", "v1", "v2", "v3", "b", "vorticity")
    ref = SweOracle.for_semi(oracle, semi).prim_vorticity(u)
    scale = np.maximum(np.abs(ref).max(axis=1), 1e-300)
    err = np.abs(out - ref).max(axis=1) / scale
    print("  rel err per variable", err, flush=True)
    assert np.all(err <= 1e-13), err
    cons, cnames = SF.solution_variables(semi, "cons2cons")
    assert cnames == ("h", "h_vcon1", "h_vcon2")
    np.testing.assert_array_equal(cons, u.reshape(3, -1, order="F"))


def test_euler3d_cons2prim_recovers_the_primitive_state():
    print("case: 3D shell 3x3 trees x 2 layers polydeg 3", flush=True)
    mesh, eq, solver, bcs, _ = W.shell(3, 2, polydeg=3, seed=0)
    semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs)
    n_nodes = semi.ndofs
    rng = np.random.default_rng(4)
    rho = rng.uniform(0.5, 1.5, n_nodes)
    v = rng.uniform(-30, 30, (3, n_nodes))
    p = rng.uniform(2e4, 1e5, n_nodes)
    phi = rng.uniform(0, 3e5, n_nodes)
    gamma = eq.c_p / eq.c_v
    # conservative state of the energy form: rho e = p / (gamma - 1) + rho |v|^2 / 2
    u = np.stack([rho, rho * v[0], rho * v[1], rho * v[2],
                  p / (gamma - 1) + 0.5 * rho * (v ** 2).sum(axis=0), phi])
    semi.set_state(np.ascontiguousarray(u.T).ravel())
    out, names = SF.solution_variables(semi, "cons2prim")
    assert names == ("rho", "v1", "v2", "v3", "p", "phi")
    prim = np.stack([rho, v[0], v[1], v[2], p, phi])
    np.testing.assert_array_equal(out[0], rho)
    np.testing.assert_array_equal(out[5], phi)
    np.testing.assert_allclose(out[1:4], prim[1:4], rtol=0, atol=1e-13 * 30)
    np.testing.assert_allclose(out[4], p, rtol=1e-11)  # p from rho e - |v|^2 / 2: cancellation


def test_rainy_cons2prim_velocities():
    mesh, eq, solver, src, lim, u0, _ = W.rainy_bubble(cells=(8, 4), polydeg=3, seed=2)
    semi = T.SemidiscretizationRainy(mesh, eq, solver, source_terms=src, stage_limiter=lim)
    u = np.array(u0.reshape(9, -1, order="F"))
    rng = np.random.default_rng(5)
    vel = rng.uniform(-20, 20, (2, u.shape[1]))
    rho = u[0] + u[1] + u[2]
    u[3], u[4] = rho * vel[0], rho * vel[1]
    semi.set_state(np.ascontiguousarray(u.T).ravel())
    out, names = SF.solution_variables(semi, "cons2prim")
    assert names[3:5] == ("v1", "v2")
    np.testing.assert_allclose(out[3:5], vel, rtol=1e-15, atol=1e-14)
    for v in (0, 1, 2, 5, 6, 7, 8):
        np.testing.assert_array_equal(out[v], u[v])


def test_solution_file_and_restart_resume(tmp_path):
    print("case: covariant zonal jet 3x3 polydeg 3 with files", flush=True)
    case = W.zonal_jet(3, 3, seed=6)
    semi = case.semi()
    u0 = case.u0()
    semi.set_state(u0)
    semi.set_time(0.0)
    semi.solve_resident(3000.0, 0.4)
    fn = SF.save_solution_file(semi, 3000.0, 100.0, 7, "cons2prim_and_vorticity",
                               output_directory=str(tmp_path))
    assert fn.endswith("solution_000000007.h5")
    attrs, data, names = SF.load_solution_file(fn, semi.ndofs)
    out, _ = SF.solution_variables(semi, "cons2prim_and_vorticity")
    np.testing.assert_array_equal(data, out)
    assert names == ["H", "v1", "v2", "v3", "b", "vorticity"]
    assert attrs["equations"] == "SplitCovariantShallowWaterEquations2D"
    assert attrs["ndims"] == 2 and attrs["n_vars"] == 6 and attrs["n_elements"] == 54
    rf = SF.save_restart_file(semi, 3000.0, 100.0, 7, output_directory=str(tmp_path))
    # continue the original, and a fresh handle resumed from the restart file
    semi.solve_resident(6000.0, 0.4)
    fresh = case.semi()
    a = SF.load_restart_file(fresh, rf)
    assert a["time"] == 3000.0 and a["timestep"] == 7
    fresh.solve_resident(6000.0, 0.4)
    np.testing.assert_array_equal(fresh.get_state(), semi.get_state())
