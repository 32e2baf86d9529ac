"""GPU: P4estMesh{3} metric terms computed on the device.

atmo_metric_terms3d_device (and atmo_create given node coordinates only)
computes the Jacobian matrix, the invariant-curl contravariant vectors and
1/J with the host mesh builder's operation order and no fused multiply-adds,
so the bar is exact: the same bits as the host's for the periodic cube and
the cubed-sphere shell at every polydeg, and the same rhs bits from a handle
built either way (FP64 and the mixed path, whose FP32 metric copy is narrowed
on the device)."""
import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import workloads as W
from trixiatmo_hip._lib import check, lib
from trixiatmo_hip.meshes import P4estMesh, P4estMeshCubedSphereTopography

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def device_metric_terms(mesh):
    n, E = mesh.nnodes, mesh.n_elements
    xyz = torch.from_numpy(np.ascontiguousarray(mesh.node_coordinates.ravel(order="F"))).cuda()
    cv = torch.full((9 * n ** 3 * E,), float("nan"), dtype=torch.float64, device="cuda")
    ij = torch.full((n ** 3 * E,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    check(lib().atmo_metric_terms3d_device(mesh.polydeg, E, xyz.data_ptr(), cv.data_ptr(),
                                           ij.data_ptr()))
    return (cv.cpu().numpy().reshape((3, 3, n, n, n, E), order="F"),
            ij.cpu().numpy().reshape((n, n, n, E), order="F"))


@pytest.mark.parametrize("polydeg", [1, 2, 3, 4, 5, 6, 7])
def test_metric_terms_match_host_shell(polydeg):
    print(f"case: shell 3x3 trees x 2 layers polydeg {polydeg}", flush=True)
    mesh = P4estMeshCubedSphereTopography(3, 2, 6.371e6, 30000.0, polydeg=polydeg)
    cv, ij = device_metric_terms(mesh)
    np.testing.assert_array_equal(cv, mesh.contravariant_vectors)
    np.testing.assert_array_equal(ij, mesh.inverse_jacobian)


@pytest.mark.parametrize("polydeg", [1, 4, 7])
def test_metric_terms_match_host_cube(polydeg):
    print(f"case: cube 2x3x2 trees polydeg {polydeg}", flush=True)
    mesh = P4estMesh((2, 3, 2), polydeg=polydeg, coordinates_min=(0.0, -1.0, 0.5),
                     coordinates_max=(1.0, 2.0, 1.5), periodicity=(True, False, True))
    cv, ij = device_metric_terms(mesh)
    np.testing.assert_array_equal(cv, mesh.contravariant_vectors)
    np.testing.assert_array_equal(ij, mesh.inverse_jacobian)


@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan"])
def test_handle_from_node_coordinates_gives_the_same_rhs(precision):
    print(f"case: shell 4x4 trees x 2 layers polydeg 4 {precision}", flush=True)
    mesh, eq, solver, bcs, u0 = W.shell(4, 2, polydeg=4, seed=3, precision=precision)
    dev = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs, metric_terms_on_device=True)
    host = T.SemidiscretizationHyperbolic(mesh, eq, solver, bcs, metric_terms_on_device=False)
    np.testing.assert_array_equal(dev.rhs(None, u0), host.rhs(None, u0))
    for s in (dev, host):
        s.set_state(u0)
    assert dev.max_dt(0.5) == host.max_dt(0.5)
