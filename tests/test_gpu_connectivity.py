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
