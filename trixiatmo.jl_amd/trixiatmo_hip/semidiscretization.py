"""SemidiscretizationHyperbolic over libatmo_hip (3D P4est meshes).

Mirrors Trixi's `semi = SemidiscretizationHyperbolic(mesh, equations,
initial_condition, solver; boundary_conditions)` and `Trixi.rhs!(du, u, semi,
t)`; the work runs in the HIP kernels of libatmo_hip.so.
"""
import ctypes

import numpy as np

from ._lib import (AtmoConfig, AtmoGeometry2D, AtmoGeometry3D, AtmoSweConfig, check, dptr, iptr,
                   lib)
from .equations import (EARTH_ROTATION_RATE, SWE_SURFACE_FLUXES, SWE_SURFACE_LLF_STANDARD,
                        SWE_SURFACE_NONCONS, SWE_VOLUME_FLUXES, SWE_VOLUME_NONCONS,
                        SWE_VOLUME_WEAK_FORM, CovariantShallowWaterEquations2D,
                        SplitCovariantShallowWaterEquations2D)
from .solvers import PRECISIONS, VolumeIntegralWeakForm

ATMO_EQ_EULER_ENERGY_GRAVITY_3D = 1
ATMO_BC_SLIP_WALL = 0
BOUNDARY_CONDITIONS = {"boundary_condition_slip_wall": ATMO_BC_SLIP_WALL}
# source_terms name -> (ATMO_SRC_*, rotation rate)
SOURCE_TERMS = {None: (0, 0.0), "source_terms_coriolis": (1, EARTH_ROTATION_RATE)}


# covariant source terms -> ATMO_SRC_* of atmo_swe_config
COVARIANT_SOURCE_TERMS = {None: 0, "source_terms_geometric_coriolis": 1}
# source_terms_geometric_coriolis of the standard form (covariant_shallow_water.jl:215-238)
COVARIANT_STANDARD_SOURCE_TERMS = {None: 0, "source_terms_geometric_coriolis": 2}
COVARIANT_EQUATIONS = (SplitCovariantShallowWaterEquations2D, CovariantShallowWaterEquations2D)


def swe_codes(equations, solver, source_terms):
    """(volume_flux, surface_flux, surface_noncons, source_terms) codes of a
    covariant semidiscretization (atmo_swe_config)"""
    if isinstance(equations, CovariantShallowWaterEquations2D):
        sf = solver.surface_flux if isinstance(solver.surface_flux, str) else solver.surface_flux[0]
        if not isinstance(solver.volume_integral, VolumeIntegralWeakForm) or \
                sf != "flux_lax_friedrichs":
            raise ValueError("CovariantShallowWaterEquations2D runs with VolumeIntegralWeakForm() "
                             "and flux_lax_friedrichs")
        return (SWE_VOLUME_WEAK_FORM, SWE_SURFACE_LLF_STANDARD, 0,
                COVARIANT_STANDARD_SOURCE_TERMS[source_terms])
    vol, vnc = solver.volume_integral.volume_flux
    sf, snc = solver.surface_flux
    if vnc not in SWE_VOLUME_NONCONS:
        raise ValueError(f"unsupported volume nonconservative flux {vnc}")
    return (SWE_VOLUME_FLUXES[vol], SWE_SURFACE_FLUXES[sf], SWE_SURFACE_NONCONS[snc],
            COVARIANT_SOURCE_TERMS[source_terms])


class MetricTermsCovariantSphere:
    """`MetricTermsCovariantSphere(christoffel_symbols = ChristoffelSymbolsAutodiff())`
    (src/semidiscretization/metric_terms.jl:28-73): the Christoffel symbols
    come from exact derivatives of the analytic metric ("autodiff") or from
    the collocation derivative of the nodal metric ("collocation")."""

    def __init__(self, christoffel_symbols="ChristoffelSymbolsAutodiff()"):
        if christoffel_symbols not in ("ChristoffelSymbolsAutodiff()",
                                       "ChristoffelSymbolsCollocationDerivative()"):
            raise ValueError(f"unknown christoffel_symbols {christoffel_symbols}")
        self.christoffel_symbols = christoffel_symbols

    @property
    def collocation(self):
        return self.christoffel_symbols == "ChristoffelSymbolsCollocationDerivative()"


def _swe_config(equations, solver, n_elements, source_terms):
    if solver.precision not in ("fp64", "fp32_geometry"):
        raise ValueError("the covariant form runs in FP64 (precision fp64 or fp32_geometry)")
    vol, sf, snc, src = swe_codes(equations, solver, source_terms)
    return AtmoSweConfig(polydeg=solver.polydeg, n_elements=n_elements, volume_flux=vol,
                         surface_flux=sf, surface_noncons=snc, source_terms=src,
                         precision=PRECISIONS[solver.precision],
                         gravity=equations.gravity, rotation_rate=equations.rotation_rate)


def covariant_aux(mesh, metric_terms=None, auxiliary_field=None):
    """aux_node_vars [26,n,n,E] of init_auxiliary_node_variables!
    (src/solvers/dgsem_p4est/containers_2d_manifold_in_3d_covariant.jl:232-305),
    the bottom topography row filled from `auxiliary_field(x)` at the nodes."""
    mt = metric_terms or MetricTermsCovariantSphere()
    aux = mesh.data(mt.collocation).aux_node_vars.copy(order="F")
    if auxiliary_field is not None:
        aux[19] = auxiliary_field(mesh.node_coordinates)
    return aux


class SemidiscretizationHyperbolic:
    """`SemidiscretizationHyperbolic(mesh, equations, initial_condition, solver;
    source_terms, boundary_conditions)`.  For P4estMesh{3} the initial state
    is built by the caller; for the covariant form on P4estMeshCubedSphere2D
    (src/semidiscretization/semidiscretization_hyperbolic_2d_manifold_in_3d.jl:78-111)
    `initial_condition` is a transformed initial condition (x, t, aux) -> u,
    and `metric_terms` / `auxiliary_field` are the reference's keywords."""

    def __init__(self, mesh, equations, solver, boundary_conditions=None, source_terms=None,
                 initial_condition=None, metric_terms=None, auxiliary_field=None,
                 metric_terms_on_device=True, connectivity_on_device=False):
        self.mesh, self.equations, self.solver = mesh, equations, solver
        # the interface (and boundary) lists: the mesh's host arrays, or built
        # by the handle on the device from the node coordinates
        self.connectivity_on_device = connectivity_on_device
        # covariant form: the auxiliary variables of the handle are computed on
        # the device (metric_terms_on_device) or uploaded from the host
        self.aux_on_device = metric_terms_on_device
        self.initial_condition = initial_condition
        self.source_terms = source_terms
        if isinstance(equations, COVARIANT_EQUATIONS):
            self._init_covariant(source_terms, metric_terms, auxiliary_field)
            return
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
        by_kind = [BOUNDARY_CONDITIONS[bc.get(kinds[q], "boundary_condition_slip_wall")]
                   for q in range(3)]
        # P4estMesh{3} metric terms: computed on the device from the node
        # coordinates (the same bits as the host's), or uploaded from the host
        if metric_terms_on_device or connectivity_on_device:
            xyz = np.ascontiguousarray(mesh.node_coordinates.ravel(order="F"))
        else:
            xyz = None
        if metric_terms_on_device:
            geo_arrays = [None, None]
        else:
            geo_arrays = [np.ascontiguousarray(mesh.contravariant_vectors.ravel(order="F")),
                          np.ascontiguousarray(mesh.inverse_jacobian.ravel(order="F"))]
        self._keep = [*geo_arrays,
                      np.ascontiguousarray(mesh.interface_neighbor_ids.ravel(order="F")),
                      np.ascontiguousarray(mesh.interface_node_indices.ravel(order="F")),
                      np.ascontiguousarray(mesh.boundary_element),
                      np.ascontiguousarray(mesh.boundary_node_indices.ravel(order="F")),
                      bcodes,
                      np.ascontiguousarray(getattr(mesh, "period", np.zeros(3)), dtype=np.float64)]
        k = self._keep
        geo = AtmoGeometry3D(contravariant_vectors=dptr(k[0]) if k[0] is not None else None,
                             inverse_jacobian=dptr(k[1]) if k[1] is not None else None,
                             n_interfaces=-1 if connectivity_on_device else mesh.n_interfaces,
                             interface_neighbor_ids=iptr(k[2]),
                             interface_node_indices=iptr(k[3]), n_boundaries=mesh.n_boundaries,
                             boundary_element=iptr(k[4]), boundary_node_indices=iptr(k[5]),
                             boundary_condition=iptr(k[6]),
                             node_coordinates=dptr(xyz) if xyz is not None else None,
                             period=dptr(k[7]),
                             boundary_condition_by_kind=(ctypes.c_int32 * 3)(*by_kind))
        h = ctypes.c_void_p()
        check(lib().atmo_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)))
        del xyz
        self._h = h
        n = mesh.nnodes
        self.shape = (equations.nvariables, n, n, n, mesh.n_elements)

    def _init_covariant(self, source_terms, metric_terms, auxiliary_field):
        mesh, solver = self.mesh, self.solver
        if solver.polydeg != mesh.polydeg:
            raise ValueError("solver and mesh polydeg differ")
        # the host copy of the auxiliary variables serves the analysis and the
        # initial conditions; the handle builds its own on the device from the
        # node coordinates (aux_on_device) or takes this copy
        self.aux_node_vars = covariant_aux(mesh, metric_terms, auxiliary_field)
        cfg = _swe_config(self.equations, solver, mesh.n_elements, source_terms)
        mt = metric_terms or MetricTermsCovariantSphere()
        k = [np.ascontiguousarray(self.aux_node_vars.ravel(order="F")),
             np.ascontiguousarray(mesh.node_coordinates.ravel(order="F")),
             np.ascontiguousarray(mesh.interface_neighbor_ids.ravel(order="F"), dtype=np.int32),
             np.ascontiguousarray(mesh.interface_node_indices.ravel(order="F"), dtype=np.int32),
             np.ascontiguousarray(self.aux_node_vars[19].ravel(order="F"))]
        geo = AtmoGeometry2D(aux_node_vars=None if self.aux_on_device else dptr(k[0]),
                             node_coordinates=dptr(k[1]),
                             n_interfaces=-1 if self.connectivity_on_device else mesh.n_interfaces,
                             interface_neighbor_ids=iptr(k[2]),
                             interface_node_indices=iptr(k[3]), n_members=1,
                             member_first_element=None,
                             christoffel_collocation=1 if mt.collocation else 0,
                             bottom_topography=dptr(k[4]) if auxiliary_field is not None else None)
        h = ctypes.c_void_p()
        check(lib().atmo_swe_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)))
        self._h = h
        n = mesh.nnodes
        self.shape = (3, n, n, mesh.n_elements)
        self.n_members = 1

    def compute_coefficients(self, t=0.0):
        """Trixi.compute_coefficients(initial_condition, t, semi) for the
        covariant form: the conserved state at the nodes, [3,n,n,E]."""
        if self.initial_condition is None:
            raise ValueError("no initial condition")
        return self.initial_condition(self.mesh.node_coordinates, t, self.aux_node_vars)

    @property
    def is_covariant(self):
        return isinstance(self.equations, COVARIANT_EQUATIONS)

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
        """Trixi's DOF count: nodes (E n^d), not nodes x variables."""
        return int(np.prod(self.shape[1:]))

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

    def set_time(self, t):
        """time of the resident state (the start of solve_resident)"""
        check(lib().atmo_set_time(self._h, float(t)), self._h)

    def solve_resident(self, t_end, cfl, check_every=64):
        """StepsizeCallback(cfl) time loop from the resident time to t_end on
        the device (see atmo_solve); returns (steps per member, time reached)."""
        return solve_resident(self._h, getattr(self, "n_members", 1), t_end, cfl, check_every)

    def integrals(self, fresh_rhs=True):
        """(mass, entropy, dS/dt, integral of each conserved variable...)"""
        out = np.zeros(3 if self.is_covariant else 9)
        check(lib().atmo_integrals(self._h, 1 if fresh_rhs else 0, dptr(out)), self._h)
        return out

    def synchronize(self):
        check(lib().atmo_synchronize(self._h), self._h)

    def autotune(self):
        """time every workgroup shape of the stage kernel on the resident
        state, keep the fastest; returns (chosen shape, ms per shape)"""
        v = ctypes.c_int32()
        ms = np.zeros(6)
        check(lib().atmo_autotune(self._h, ctypes.byref(v), dptr(ms)), self._h)
        return v.value, ms

    def set_kernel_variant(self, variant):
        """pin the workgroup shape of the stage kernel"""
        check(lib().atmo_set_kernel_variant(self._h, int(variant)), self._h)

    def enable_timing(self, on=True):
        check(lib().atmo_enable_timing(self._h, 1 if on else 0), self._h)

    def kernel_times(self, kernel_class=0):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        check(lib().atmo_kernel_times(self._h, kernel_class, ctypes.byref(n), ctypes.byref(ms)),
              self._h)
        return n.value, ms.value


class SemidiscretizationEnsemble(SemidiscretizationHyperbolic):
    """Independent covariant shallow-water members on their own cubed-sphere
    meshes (uneven element counts) in one handle: every RK stage is one
    launch over all members' elements, each member advancing with its own
    time step.  `members` is a list of (mesh, transformed initial condition,
    auxiliary_field or None)."""

    def __init__(self, members, equations, solver, source_terms=None, metric_terms=None,
                 metric_terms_on_device=True, connectivity_on_device=False):
        self.equations, self.solver = equations, solver
        self.connectivity_on_device = connectivity_on_device
        self.source_terms = source_terms
        self.mesh = None
        self.initial_condition = None
        auxs, xyzs, nbrs, idxs, first = [], [], [], [], [0]
        self.members = []
        for mesh, ic, field in members:
            if mesh.polydeg != solver.polydeg:
                raise ValueError("solver and mesh polydeg differ")
            aux = covariant_aux(mesh, metric_terms, field)
            auxs.append(aux)
            xyzs.append(mesh.node_coordinates)
            nbrs.append(mesh.interface_neighbor_ids + first[-1])
            idxs.append(mesh.interface_node_indices)
            self.members.append((mesh, ic, aux, first[-1]))
            first.append(first[-1] + mesh.n_elements)
        self.aux_node_vars = np.concatenate(auxs, axis=-1)
        self.node_coordinates = np.concatenate(xyzs, axis=-1)
        nbr = np.concatenate(nbrs, axis=-1)
        idx = np.concatenate(idxs, axis=-1)
        E = first[-1]
        self.member_first_element = np.asarray(first, dtype=np.int32)
        self.n_members = len(members)
        cfg = _swe_config(equations, solver, E, source_terms)
        mt = metric_terms or MetricTermsCovariantSphere()
        k = [np.ascontiguousarray(self.aux_node_vars.ravel(order="F")),
             np.ascontiguousarray(self.node_coordinates.ravel(order="F")),
             np.ascontiguousarray(nbr.ravel(order="F"), dtype=np.int32),
             np.ascontiguousarray(idx.ravel(order="F"), dtype=np.int32),
             self.member_first_element,
             np.ascontiguousarray(self.aux_node_vars[19].ravel(order="F"))]
        has_topography = any(field is not None for _, _, field in members)
        geo = AtmoGeometry2D(aux_node_vars=None if metric_terms_on_device else dptr(k[0]),
                             node_coordinates=dptr(k[1]),
                             n_interfaces=-1 if connectivity_on_device else nbr.shape[1],
                             interface_neighbor_ids=iptr(k[2]),
                             interface_node_indices=iptr(k[3]), n_members=self.n_members,
                             member_first_element=iptr(k[4]),
                             christoffel_collocation=1 if mt.collocation else 0,
                             bottom_topography=dptr(k[5]) if has_topography else None)
        h = ctypes.c_void_p()
        check(lib().atmo_swe_create(ctypes.byref(cfg), ctypes.byref(geo), ctypes.byref(h)))
        self._h = h
        n = solver.polydeg + 1
        self.shape = (3, n, n, E)

    @property
    def is_covariant(self):
        return True

    def compute_coefficients(self, t=0.0):
        parts = [ic(mesh.node_coordinates, t, aux) for mesh, ic, aux, _ in self.members]
        return np.concatenate(parts, axis=-1)

    def max_dt_members(self, cfl):
        """StepsizeCallback max_dt of every member, [n_members]."""
        dt = np.zeros(self.n_members)
        check(lib().atmo_max_dt_members(self._h, float(cfl), dptr(dt)), self._h)
        return dt

    def step_members(self, dt, n_steps=1):
        """n_steps CarpenterKennedy2N54 steps, member m with time step dt[m]."""
        d = np.ascontiguousarray(dt, dtype=np.float64)
        if d.shape != (self.n_members,):
            raise ValueError("one time step per member")
        check(lib().atmo_step_members(self._h, dptr(d), int(n_steps)), self._h)


def solve_resident(handle, n_members, t_end, cfl, check_every=64):
    """atmo_solve on a handle: (steps taken per member, time reached)"""
    steps = np.zeros(n_members, dtype=np.int64)
    t = ctypes.c_double()
    check(lib().atmo_solve(handle, float(t_end), float(cfl), int(check_every),
                           steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           ctypes.byref(t)), handle)
    return steps, t.value


def solve(semi, u0, tspan, cfl, resident=False, check_every=64):
    """`solve(ode, CarpenterKennedy2N54(williamson_condition = false);
    callback = StepsizeCallback(cfl))` on the resident state: before every
    step dt = cfl * max_dt(u), and the last step is cut to end at tspan[1]
    as OrdinaryDiffEq's tstop handling does
    (examples/shallow_water/covariant/elixir_barotropic_instability.jl:67,79-80).
    Host loop by default; `resident=True` keeps the loop on the device
    (same steps and state, the host waiting only every `check_every` steps;
    ensemble members then each take their own step sizes).  Returns
    (final state, steps), steps per member for a resident ensemble."""
    t, t_end = float(tspan[0]), float(tspan[1])
    semi.set_state(u0)
    semi.set_time(t)
    if resident:
        steps, _ = semi.solve_resident(t_end, cfl, check_every)
        if getattr(semi, "n_members", 1) == 1:
            steps = int(steps[0])
    else:
        steps = 0
        while t < t_end:
            dt = min(semi.max_dt(cfl), t_end - t)
            semi.step(dt, 1)
            t += dt
            steps += 1
    return semi.get_state().reshape(semi.shape, order="F"), steps
