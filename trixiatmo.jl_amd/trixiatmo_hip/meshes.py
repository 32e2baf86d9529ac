"""Mesh constructors mirroring TrixiAtmo/Trixi names (P4estMesh{3} family).

Each mesh carries the arrays Trixi keeps in its element, interface and
boundary containers, as numpy arrays in Julia column-major order (first index
fastest), so they can be handed to the semidiscretization unchanged.
"""
import ctypes

import numpy as np

from ._lib import check, c_double_p, c_int32_p, lib

# Trixi node-index symbols and their integer codes in the C ABI
NODE_INDEX_CODES = {"begin": 0, "end": 1, "i_forward": 2, "i_backward": 3,
                    "j_forward": 4, "j_backward": 5}


class P4estMesh3D:
    """A level-0 P4estMesh{3}: one element per tree, geometry and connectivity.

    Attributes (numpy, Julia layout): node_coordinates [3,n,n,n,E],
    contravariant_vectors [3,3,n,n,n,E], inverse_jacobian [n,n,n,E],
    interface_neighbor_ids [2,I], interface_node_indices [3,2,I],
    boundary_element [B], boundary_node_indices [3,B], boundary_kind [B]
    (0 = :inside, 1 = :outside, 2 = other).
    """

    def __init__(self, handle, polydeg):
        L = lib()
        self.polydeg = polydeg
        n = polydeg + 1
        ne, ni, nb = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(L.atmo_mesh3d_sizes(handle, ctypes.byref(ne), ctypes.byref(ni), ctypes.byref(nb)))
        self.n_elements, self.n_interfaces, self.n_boundaries = ne.value, ni.value, nb.value
        xyz = c_double_p()
        ptrs_i = [c_int32_p() for _ in range(5)]
        # the metric terms are not requested here: they are computed on first
        # use (host) or by the semidiscretization on the device
        check(L.atmo_mesh3d_arrays(handle, ctypes.byref(xyz), None, None,
                                   *[ctypes.byref(p) for p in ptrs_i]))
        E, I, B = self.n_elements, self.n_interfaces, self.n_boundaries

        def dcopy(p, shape):
            cnt = int(np.prod(shape))
            if cnt == 0:
                return np.zeros(shape, dtype=np.float64, order="F")
            return np.ctypeslib.as_array(p, shape=(cnt,)).copy().reshape(shape, order="F")

        def icopy(p, shape):
            cnt = int(np.prod(shape))
            if cnt == 0:
                return np.zeros(shape, dtype=np.int32, order="F")
            return np.ctypeslib.as_array(p, shape=(cnt,)).copy().reshape(shape, order="F")

        self.node_coordinates = dcopy(xyz, (3, n, n, n, E))
        self._metric = None
        self.interface_neighbor_ids = icopy(ptrs_i[0], (2, I))
        self.interface_node_indices = icopy(ptrs_i[1], (3, 2, I))
        self.boundary_element = icopy(ptrs_i[2], (B,))
        self.boundary_node_indices = icopy(ptrs_i[3], (3, B))
        self.boundary_kind = icopy(ptrs_i[4], (B,))
        L.atmo_mesh3d_destroy(handle)

    @property
    def nnodes(self):
        return self.polydeg + 1

    def _metric_terms(self):
        if self._metric is None:
            n, E = self.nnodes, self.n_elements
            xyz = np.ascontiguousarray(self.node_coordinates.ravel(order="F"))
            cv = np.empty(9 * n ** 3 * E)
            ij = np.empty(n ** 3 * E)
            check(lib().atmo_metric_terms3d(self.polydeg, E, xyz.ctypes.data_as(c_double_p),
                                            cv.ctypes.data_as(c_double_p),
                                            ij.ctypes.data_as(c_double_p)))
            self._metric = (cv.reshape((3, 3, n, n, n, E), order="F"),
                            ij.reshape((n, n, n, E), order="F"))
        return self._metric

    @property
    def contravariant_vectors(self):
        """[3,3,n,n,n,E] in Kopriva's invariant curl form (host, computed on first use)"""
        return self._metric_terms()[0]

    @property
    def inverse_jacobian(self):
        """[n,n,n,E] (host, computed on first use)"""
        return self._metric_terms()[1]


def P4estMesh(trees_per_dimension, *, polydeg, coordinates_min, coordinates_max,
              periodicity=(True, True, True)):
    """Trixi `P4estMesh(trees_per_dimension; polydeg, coordinates_min,
    coordinates_max, periodicity)` in 3D with initial_refinement_level = 0."""
    t = np.asarray(trees_per_dimension, dtype=np.int32)
    lo = np.asarray(coordinates_min, dtype=np.float64)
    hi = np.asarray(coordinates_max, dtype=np.float64)
    per = np.asarray([1 if p else 0 for p in periodicity], dtype=np.int32)
    h = ctypes.c_void_p()
    check(lib().atmo_mesh3d_cartesian(polydeg, t.ctypes.data_as(c_int32_p),
                                      lo.ctypes.data_as(c_double_p), hi.ctypes.data_as(c_double_p),
                                      per.ctypes.data_as(c_int32_p), ctypes.byref(h)))
    mesh = P4estMesh3D(h, polydeg)
    mesh.period = np.where(per != 0, hi - lo, 0.0)
    return mesh


def P4estMeshCubedSphereTopography(trees_per_face_dimension, layers, inner_radius, thickness,
                                   *, polydeg):
    """src/meshes/p4est_cubed_sphere_topography.jl:114-163 with the perfect
    sphere topography and initial_refinement_level = 0."""
    h = ctypes.c_void_p()
    check(lib().atmo_mesh3d_cubed_sphere_shell(polydeg, trees_per_face_dimension, layers,
                                               float(inner_radius), float(thickness),
                                               ctypes.byref(h)))
    return P4estMesh3D(h, polydeg)


def device_connectivity(mesh):
    """The face connectivity of `mesh` (P4estMesh3D or P4estMeshCubedSphere2D)
    built on the device from its node coordinates alone
    (atmo_mesh3d_connectivity_device / atmo_mesh2d_connectivity_device): a
    dict of the interface (and, in 3D, boundary) arrays in the layout of the
    mesh's own attributes."""
    import torch
    E, n = mesh.n_elements, mesh.nnodes
    X = torch.from_numpy(np.ascontiguousarray(mesh.node_coordinates.ravel(order="F"))).cuda()
    ni, nb = ctypes.c_int32(0), ctypes.c_int32(0)
    i32 = dict(dtype=torch.int32, device="cuda")
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    if isinstance(mesh, P4estMesh3D):
        ids, codes = torch.full((2 * 3 * E,), -1, **i32), torch.full((6 * 3 * E,), -1, **i32)
        be, bc, bk = (torch.full((6 * E,), -1, **i32), torch.full((3 * 6 * E,), -1, **i32),
                      torch.full((6 * E,), -1, **i32))
        period = np.ascontiguousarray(getattr(mesh, "period", np.zeros(3)), dtype=np.float64)
        check(lib().atmo_mesh3d_connectivity_device(
            mesh.polydeg, E, ptr(X), period.ctypes.data_as(c_double_p), ptr(ids), ptr(codes),
            ptr(be), ptr(bc), ptr(bk), ctypes.byref(ni), ctypes.byref(nb)))
        I, B = ni.value, nb.value
        return {"interface_neighbor_ids": ids.cpu().numpy()[:2 * I].reshape(2, I, order="F"),
                "interface_node_indices": codes.cpu().numpy()[:6 * I].reshape(3, 2, I, order="F"),
                "boundary_element": be.cpu().numpy()[:B],
                "boundary_node_indices": bc.cpu().numpy()[:3 * B].reshape(3, B, order="F"),
                "boundary_kind": bk.cpu().numpy()[:B]}
    ids, codes = torch.full((2 * 2 * E,), -1, **i32), torch.full((4 * 2 * E,), -1, **i32)
    check(lib().atmo_mesh2d_connectivity_device(mesh.polydeg, E, ptr(X), ptr(ids), ptr(codes),
                                                ctypes.byref(ni)))
    I = ni.value
    return {"interface_neighbor_ids": ids.cpu().numpy()[:2 * I].reshape(2, I, order="F"),
            "interface_node_indices": codes.cpu().numpy()[:4 * I].reshape(2, 2, I, order="F")}


class P4estMeshCubedSphere2DData:
    """A level-0 P4estMesh{2} of the sphere (one element per tree) as
    P4estMeshCubedSphere2D builds it, with the 26 auxiliary variables per node
    of the covariant form.

    Attributes (numpy, Julia layout): node_coordinates [3,n,n,E],
    aux_node_vars [26,n,n,E] (bottom topography row 19 zero),
    interface_neighbor_ids [2,I], interface_node_indices [2,2,I] (codes
    begin/end/i_forward/i_backward).
    """

    def __init__(self, trees_per_face_dimension, radius, polydeg, christoffel_collocation):
        L = lib()
        h = ctypes.c_void_p()
        check(L.atmo_mesh2d_cubed_sphere(polydeg, trees_per_face_dimension, float(radius),
                                         1 if christoffel_collocation else 0, ctypes.byref(h)))
        try:
            ne, ni = ctypes.c_int32(), ctypes.c_int32()
            check(L.atmo_mesh2d_sizes(h, ctypes.byref(ne), ctypes.byref(ni)))
            pd = [c_double_p() for _ in range(2)]
            pi = [c_int32_p() for _ in range(2)]
            check(L.atmo_mesh2d_arrays(h, *[ctypes.byref(p) for p in pd],
                                       *[ctypes.byref(p) for p in pi]))
            n, E, I = polydeg + 1, ne.value, ni.value

            def copy(p, shape):
                cnt = int(np.prod(shape))
                return np.ctypeslib.as_array(p, shape=(cnt,)).copy().reshape(shape, order="F")

            self.node_coordinates = copy(pd[0], (3, n, n, E))
            self.aux_node_vars = copy(pd[1], (26, n, n, E))
            self.interface_neighbor_ids = copy(pi[0], (2, I))
            self.interface_node_indices = copy(pi[1], (2, 2, I))
        finally:
            L.atmo_mesh2d_destroy(h)
        self.polydeg = polydeg
        self.trees_per_face_dimension = trees_per_face_dimension
        self.radius = float(radius)
        self.christoffel_collocation = bool(christoffel_collocation)
        self.n_elements, self.n_interfaces = E, I

    @property
    def nnodes(self):
        return self.polydeg + 1


class P4estMeshCubedSphere2D:
    """`P4estMeshCubedSphere2D(trees_per_face_dimension, radius; polydeg,
    element_local_mapping = true)` (src/meshes/p4est_cubed_sphere.jl:47-79)
    with initial_refinement_level = 0.

    The auxiliary variables depend on the metric terms chosen at the
    semidiscretization (`MetricTermsCovariantSphere(christoffel_symbols =
    ...)`); `data(christoffel_collocation)` builds them for either choice.
    """

    def __init__(self, trees_per_face_dimension, radius, *, polydeg,
                 element_local_mapping=True):
        if not element_local_mapping:
            raise ValueError("the covariant form uses element_local_mapping = true")
        self.trees_per_face_dimension = int(trees_per_face_dimension)
        self.radius = float(radius)
        self.polydeg = int(polydeg)
        self._data = {}
        d = self.data(False)
        self.node_coordinates = d.node_coordinates
        self.interface_neighbor_ids = d.interface_neighbor_ids
        self.interface_node_indices = d.interface_node_indices
        self.n_elements, self.n_interfaces = d.n_elements, d.n_interfaces

    def data(self, christoffel_collocation=False):
        key = bool(christoffel_collocation)
        if key not in self._data:
            self._data[key] = P4estMeshCubedSphere2DData(self.trees_per_face_dimension,
                                                         self.radius, self.polydeg, key)
        return self._data[key]

    @property
    def nnodes(self):
        return self.polydeg + 1


def lgl_basis(polydeg):
    """LobattoLegendreBasis(polydeg): (nodes, weights, D, Dsplit)."""
    n = polydeg + 1
    x = np.zeros(n)
    w = np.zeros(n)
    D = np.zeros(n * n)
    Ds = np.zeros(n * n)
    check(lib().atmo_lgl_basis(n, x.ctypes.data_as(c_double_p), w.ctypes.data_as(c_double_p),
                               D.ctypes.data_as(c_double_p), Ds.ctypes.data_as(c_double_p)))
    return x, w, D.reshape(n, n), Ds.reshape(n, n)


class StructuredMesh2D:
    """`StructuredMesh(cells_per_dimension, coordinates_min, coordinates_max;
    periodicity)` [ext] for a uniform Cartesian rectangle: elements numbered x
    fastest, node (i, j) of element (ex, ey) at
    x = x_min + dx (ex + (xi_i + 1) / 2), y likewise.  `periodicity[1]` false
    means boundary_condition_simple_slip_wall at y_min and y_max."""

    def __init__(self, cells_per_dimension, coordinates_min, coordinates_max, *, polydeg,
                 periodicity=(True, True)):
        self.nx, self.ny = (int(c) for c in cells_per_dimension)
        self.x0, self.y0 = (float(c) for c in coordinates_min)
        self.x1, self.y1 = (float(c) for c in coordinates_max)
        if not periodicity[0]:
            raise ValueError("the rainy path is periodic in x")
        self.periodic_y = bool(periodicity[1])
        self.polydeg = int(polydeg)
        self.n_elements = self.nx * self.ny
        n = self.polydeg + 1
        xi, _, _, _ = lgl_basis(self.polydeg)
        dx = (self.x1 - self.x0) / self.nx
        dy = (self.y1 - self.y0) / self.ny
        ex = np.arange(self.nx)
        ey = np.arange(self.ny)
        X = self.x0 + dx * (ex[None, None, :, None] + 0.5 * (xi[:, None, None, None] + 1.0))
        Y = self.y0 + dy * (ey[None, None, None, :] + 0.5 * (xi[None, :, None, None] + 1.0))
        X, Y = np.broadcast_arrays(X, Y)
        # [2, i, j, e] with e = ex + nx * ey
        self.node_coordinates = np.stack([X.reshape(n, n, -1, order="F"),
                                          Y.reshape(n, n, -1, order="F")])
        self.dx, self.dy = dx, dy

    @property
    def nnodes(self):
        return self.polydeg + 1
