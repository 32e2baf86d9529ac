"""Face connectivity built on the device (atmo_mesh3d_connectivity_device,
atmo_mesh2d_connectivity_device) against the host builders' lists: the
interface element ids and node-index codes and, in 3D, the boundary
elements, codes and kinds must be equal entry for entry (same primary
sides, same order, same orientations) on Cartesian bricks with every
periodicity pattern, on cubed-sphere shells at several polydeg up to the
benchmarked 6 x 16^2 x 8, and on the 2D cubed sphere up to 6 x 32^2."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip.meshes import device_connectivity

pytestmark = pytest.mark.gpu

pytest.importorskip("torch")

KEYS_3D = ("interface_neighbor_ids", "interface_node_indices", "boundary_element",
           "boundary_node_indices", "boundary_kind")


def _check(mesh, keys):
    dev = device_connectivity(mesh)
    for k in keys:
        host = getattr(mesh, k)
        print(f"  {k}: host {host.shape}, device {dev[k].shape}", flush=True)
        np.testing.assert_array_equal(dev[k], host, err_msg=k)


@pytest.mark.parametrize("trees,periodic,polydeg", [
    ((2, 3, 2), (True, False, True), 2),
    ((3, 3, 3), (True, True, True), 3),
    ((2, 2, 2), (False, False, False), 1),
    ((4, 1, 2), (False, True, False), 4),
    ((3, 2, 5), (True, True, False), 7),
])
def test_cartesian_brick(trees, periodic, polydeg):
    print(f"case: brick {trees} periodic {periodic} polydeg {polydeg}", flush=True)
    mesh = T.P4estMesh(trees, polydeg=polydeg, coordinates_min=(-1.0, 0.5, 2.0),
                       coordinates_max=(1.0, 3.5, 4.0), periodicity=periodic)
    _check(mesh, KEYS_3D)


@pytest.mark.parametrize("trees,layers,polydeg", [(2, 2, 1), (3, 2, 4), (4, 3, 7), (16, 8, 4)])
def test_cubed_sphere_shell(trees, layers, polydeg):
    print(f"case: shell 6x{trees}^2x{layers} polydeg {polydeg}", flush=True)
    mesh = T.P4estMeshCubedSphereTopography(trees, layers, 6.371e6, 3.0e4, polydeg=polydeg)
    _check(mesh, KEYS_3D)


@pytest.mark.parametrize("trees,polydeg", [(2, 1), (3, 3), (5, 7), (8, 2), (32, 3)])
def test_cubed_sphere_2d(trees, polydeg):
    print(f"case: cubed sphere 6x{trees}^2 polydeg {polydeg}", flush=True)
    mesh = T.P4estMeshCubedSphere2D(trees, 6.371e6, polydeg=polydeg)
    _check(mesh, ("interface_neighbor_ids", "interface_node_indices"))


def _pair(make):
    """rhs! and 3 RK steps of a handle given the host lists and of one that
    builds them on the device (connectivity_on_device)"""
    out = []
    for on_device in (False, True):
        semi, u0 = make(on_device)
        du = semi.rhs(None, u0)
        semi.set_state(u0)
        semi.step(semi.max_dt(0.4), 3)
        out.append((du, semi.get_state()))
    return out


@pytest.mark.parametrize("case", ["shell", "brick"])
def test_3d_handle_from_node_coordinates_alone(case):
    """metric terms and connectivity both built on the device: the same bits
    as a handle given the host's lists"""
    print(f"case: 3D handle from coordinates, {case}", flush=True)
    if case == "shell":
        mesh, eq, solver, bcs, u0 = T.workloads.shell(3, 2, polydeg=4, seed=3)
    else:
        mesh, eq, solver, bcs, u0 = T.workloads.periodic_cube(trees=3, polydeg=3, seed=3)

    def make(on_device):
        return T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs,
                                              connectivity_on_device=on_device), u0
    (d0, s0), (d1, s1) = _pair(make)
    np.testing.assert_array_equal(d1, d0)
    np.testing.assert_array_equal(s1, s0)


def test_covariant_handles_from_node_coordinates_alone():
    """auxiliary variables and connectivity built on the device, for one mesh
    and for an ensemble of uneven members"""
    print("case: covariant handles from coordinates", flush=True)
    case = T.workloads.zonal_jet(6, 3, seed=2)
    u0 = case.u0()

    def make(on_device):
        return T.SemidiscretizationHyperbolic(
            case.mesh, case.equations, case.solver, source_terms=case.source_terms,
            initial_condition=case.initial_condition, metric_terms=case.metric_terms,
            connectivity_on_device=on_device), u0
    (d0, s0), (d1, s1) = _pair(make)
    np.testing.assert_array_equal(d1, d0)
    np.testing.assert_array_equal(s1, s0)
    members, eq, solver, src = T.workloads.zonal_jet_ensemble(n_members=4, polydeg=3, seed=1,
                                                              choices=(3, 5, 6))
    out = []
    for on_device in (False, True):
        ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src,
                                           connectivity_on_device=on_device)
        u = ens.compute_coefficients()
        out.append(ens.rhs(None, u))
    np.testing.assert_array_equal(out[1], out[0])
