"""The covariant auxiliary variables built on the device
(atmo_covariant_aux_device, init_auxiliary_node_variables! of
src/solvers/dgsem_p4est/containers_2d_manifold_in_3d_covariant.jl:232-489)
against the host builder of atmo_mesh2d_cubed_sphere, and handles made from
node coordinates alone against handles given the host's auxiliary variables.

Bars: every row but the exact Christoffel symbols bit for bit (the device
evaluates the same expressions in the same order with contraction off); the
collocation Christoffel symbols bit for bit; the exact ones (forward-mode
dual numbers on the device, complex step on the host) to 1e-14 of the
largest magnitude of the six symbols.  The rhs of a handle built from coordinates equals
the host-aux handle's to 1e-13 of max |du| (the Christoffel symbols enter the
geometric source linearly)."""
import ctypes

import numpy as np
import pytest

import trixiatmo_hip as T
from trixiatmo_hip import workloads as W
from trixiatmo_hip._lib import check, lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def device_aux(mesh, collocation, topography=None):
    n, E = mesh.nnodes, mesh.n_elements
    X = torch.from_numpy(np.ascontiguousarray(mesh.node_coordinates.ravel(order="F"))).cuda()
    hs = None
    if topography is not None:
        hs = torch.from_numpy(np.ascontiguousarray(topography.ravel(order="F"))).cuda()
    aux = torch.full((26 * n * n * E,), np.nan, dtype=torch.float64, device="cuda")
    check(lib().atmo_covariant_aux_device(mesh.polydeg, E, ctypes.c_void_p(X.data_ptr()),
                                          1 if collocation else 0,
                                          ctypes.c_void_p(hs.data_ptr()) if hs is not None
                                          else None, ctypes.c_void_p(aux.data_ptr())))
    return aux.cpu().numpy().reshape(26, -1, order="F")


@pytest.mark.parametrize("collocation", [False, True])
@pytest.mark.parametrize("polydeg", [1, 2, 3, 4, 5, 6, 7])
def test_device_aux_matches_the_host_builder(polydeg, collocation):
    mesh = T.P4estMeshCubedSphere2D(3, W.SWE_RADIUS, polydeg=polydeg)
    print(f"case: aux polydeg {polydeg} collocation {collocation}", flush=True)
    host = mesh.data(collocation).aux_node_vars.reshape(26, -1, order="F")
    rng = np.random.default_rng(polydeg)
    hs = rng.uniform(0.0, 2000.0, (mesh.nnodes, mesh.nnodes, mesh.n_elements))
    dev = device_aux(mesh, collocation, hs)
    np.testing.assert_array_equal(dev[:19], host[:19])
    np.testing.assert_array_equal(dev[19], hs.ravel(order="F"))
    if collocation:
        np.testing.assert_array_equal(dev[20:], host[20:])
    else:
        # some symbols vanish analytically and both builders leave rounding
        # noise there: the error is measured on the scale of all six rows
        err = np.abs(dev[20:] - host[20:]).max(axis=1) / np.abs(host[20:]).max()
        print("  exact Christoffel err per row / max |Gamma|:", err, flush=True)
        assert np.all(err <= 1e-14), err
    assert np.all(device_aux(mesh, collocation)[19] == 0.0)


def _rhs_pair(case, metric_terms=None, field=None):
    mesh, eq, solver = case.mesh, case.equations, case.solver
    out = []
    for on_device in (True, False):
        semi = T.SemidiscretizationHyperbolic(mesh, eq, solver, source_terms=case.source_terms,
                                              initial_condition=case.initial_condition,
                                              metric_terms=metric_terms, auxiliary_field=field,
                                              metric_terms_on_device=on_device)
        u = semi.compute_coefficients()
        out.append(semi.rhs(None, u).reshape(3, -1, order="F"))
    return out


@pytest.mark.parametrize("christoffel", ["ChristoffelSymbolsAutodiff()",
                                         "ChristoffelSymbolsCollocationDerivative()"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_geometry"])
def test_handle_from_coordinates_matches_host_aux(christoffel, precision):
    case = W.zonal_jet(8, 3, precision=precision)
    print(f"case: zonal jet 6x8^2 p3, {christoffel}, {precision}", flush=True)
    mt = T.MetricTermsCovariantSphere(christoffel)
    dev, host = _rhs_pair(case, mt)
    err = np.abs(dev - host).max(axis=1) / np.abs(host).max(axis=1)
    print("  rhs rel err (device aux vs host aux):", err, flush=True)
    if mt.collocation:
        np.testing.assert_array_equal(dev, host)
    else:
        assert np.all(err <= 1e-13), err


def test_handle_from_coordinates_with_topography():
    """the isolated-mountain case: bottom topography row from the host array"""
    case = W.covariant_elixir("isolated_mountain", cells=6, polydeg=3)
    print("case: isolated mountain 6x6^2 p3", flush=True)
    dev, host = _rhs_pair(case, case.metric_terms, case.auxiliary_field)
    err = np.abs(dev - host).max(axis=1) / np.abs(host).max(axis=1)
    print("  rhs rel err (device aux vs host aux):", err, flush=True)
    assert np.all(err <= 1e-13), err


def test_ensemble_from_coordinates_matches_host_aux():
    """members of uneven size: each element's radius is read from the first
    node of its own member"""
    members, eq, solver, src = W.zonal_jet_ensemble(n_members=5, polydeg=3, seed=2,
                                                    choices=(4, 6, 8))
    print("case: ensemble of 5 members, device aux vs host aux", flush=True)
    out = []
    for on_device in (True, False):
        ens = T.SemidiscretizationEnsemble(members, eq, solver, source_terms=src,
                                           metric_terms_on_device=on_device)
        u = ens.compute_coefficients()
        out.append(ens.rhs(None, u).reshape(3, -1, order="F"))
    err = np.abs(out[0] - out[1]).max(axis=1) / np.abs(out[1]).max(axis=1)
    print("  rhs rel err:", err, flush=True)
    assert np.all(err <= 1e-13), err
