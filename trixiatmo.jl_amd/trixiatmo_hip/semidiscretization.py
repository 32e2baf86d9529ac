"""SemidiscretizationHyperbolic over libatmo_hip (3D P4est meshes).

Mirrors Trixi's `semi = SemidiscretizationHyperbolic(mesh, equations,
initial_condition, solver; boundary_conditions)` and `Trixi.rhs!(du, u, semi,
t)`; the work runs in the HIP kernels of libatmo_hip.so.
"""
import ctypes

import numpy as np

from ._lib import AtmoConfig, AtmoGeometry3D, check, dptr, iptr, lib
from .equations import EARTH_ROTATION_RATE

ATMO_EQ_EULER_ENERGY_GRAVITY_3D = 1
ATMO_BC_SLIP_WALL = 0
BOUNDARY_CONDITIONS = {"boundary_condition_slip_wall": ATMO_BC_SLIP_WALL}
# source_terms name -> (ATMO_SRC_*, rotation rate)
SOURCE_TERMS = {None: (0, 0.0), "source_terms_coriolis": (1, EARTH_ROTATION_RATE)}


class SemidiscretizationHyperbolic:
    def __init__(self, mesh, equations, solver, boundary_conditions=None, source_terms=None):
        self.mesh, self.equations, self.solver = mesh, equations, solver
        if solver.polydeg != mesh.polydeg:
            raise ValueError("solver and mesh polydeg differ")
        codes = solver.codes()
        src, omega = SOURCE_TERMS[source_terms]
        cfg = AtmoConfig(equations=ATMO_EQ_EULER_ENERGY_GRAVITY_3D, polydeg=solver.polydeg,
                         n_elements=mesh.n_elements, c_p=equations.c_p, c_v=equations.c_v,
                         gravity=equations.gravity, source_terms=src, rotation_rate=omega,
                         **codes)
        bc = boundary_conditions or {}
        kinds = {0: "inside", 1: "outside", 2: "other"}
        bcodes = np.zeros(mesh.n_boundaries, dtype=np.int32)
        for b in range(mesh.n_boundaries):
            name = bc.get(kinds[int(mesh.boundary_kind[b])], "boundary_condition_slip_wall")
            bcodes[b] = BOUNDARY_CONDITIONS[name]
        self._keep = [np.ascontiguousarray(mesh.contravariant_vectors.ravel(order="F")),
                      np.ascontiguousarray(mesh.inverse_jacobian.ravel(order="F")),
                      np.ascontiguousarray(mesh.interface_neighbor_ids.ravel(order="F")),
                      np.ascontiguousarray(mesh.interface_node_indices.ravel(order="F")),
                      np.ascontiguousarray(mesh.boundary_element),
                      np.ascontiguousarray(mesh.boundary_node_indices.ravel(order="F")),
                      bcodes]
        k = self._keep
        geo = AtmoGeometry3D(contravariant_vectors=dptr(k[0]), inverse_jacobian=dptr(k[1]),
                             n_interfaces=mesh.n_interfaces, interface_neighbor_ids=iptr(k[2]),
                             interface_node_indices=iptr(k[3]), n_boundaries=mesh.n_boundaries,
                             boundary_element=iptr(k[4]), boundary_node_indices=iptr(k[5]),
                             boundary_condition=iptr(k[6]))
        h = ctypes.c_void_p()
        check(lib().atmo_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)))
        self._h = h
        n = mesh.nnodes
        self.shape = (equations.nvariables, n, n, n, mesh.n_elements)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().atmo_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def ndofs(self):
        """Trixi's DOF count: nodes (E n^3), not nodes x variables."""
        return self.mesh.n_elements * self.mesh.nnodes ** 3

    def _flat(self, u):
        a = np.asarray(u, dtype=np.float64)
        if a.shape == self.shape:
            a = a.ravel(order="F")
        if a.size != int(np.prod(self.shape)):
            raise ValueError("state has the wrong size")
        return np.ascontiguousarray(a)

    def rhs(self, du, u, t=0.0):
        """Trixi.rhs!(du, u, semi, t): fills du (flat or [v,i,j,k,e] F-ordered)."""
        uf = self._flat(u)
        out = np.empty_like(uf)
        check(lib().atmo_rhs(self._h, dptr(out), dptr(uf), float(t)), self._h)
        if du is not None:
            if du.shape == self.shape:
                du[...] = out.reshape(self.shape, order="F")
            else:
                du[...] = out
        return out

    def set_state(self, u):
        uf = self._flat(u)
        check(lib().atmo_set_state(self._h, dptr(uf)), self._h)

    def get_state(self):
        out = np.empty(int(np.prod(self.shape)))
        check(lib().atmo_get_state(self._h, dptr(out)), self._h)
        return out

    def step(self, dt, n_steps=1):
        """n_steps CarpenterKennedy2N54 steps of size dt on the resident state."""
        check(lib().atmo_step(self._h, float(dt), int(n_steps)), self._h)

    def max_dt(self, cfl):
        dt = ctypes.c_double()
        check(lib().atmo_max_dt(self._h, float(cfl), ctypes.byref(dt)), self._h)
        return dt.value

    def integrals(self, fresh_rhs=True):
        """(mass, entropy, dS/dt, integral of each conserved variable...)"""
        out = np.zeros(9)
        check(lib().atmo_integrals(self._h, 1 if fresh_rhs else 0, dptr(out)), self._h)
        return out

    def synchronize(self):
        check(lib().atmo_synchronize(self._h), self._h)

    def enable_timing(self, on=True):
        check(lib().atmo_enable_timing(self._h, 1 if on else 0), self._h)

    def kernel_times(self, kernel_class=0):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        check(lib().atmo_kernel_times(self._h, kernel_class, ctypes.byref(n), ctypes.byref(ms)),
              self._h)
        return n.value, ms.value
