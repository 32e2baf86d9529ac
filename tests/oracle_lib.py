"""ctypes access to the CPU oracle (oracle/build/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it, never the
other way round, and nothing in the product imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "build", "liboracle.so")

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int)


class E3Params(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("n_elements", ctypes.c_int), ("c_p", ctypes.c_double),
                ("c_v", ctypes.c_double), ("gravity", ctypes.c_double),
                ("volume_flux", ctypes.c_int), ("volume_noncons", ctypes.c_int),
                ("surface_flux", ctypes.c_int), ("surface_noncons", ctypes.c_int),
                ("lmars_c", ctypes.c_double), ("coriolis_rate", ctypes.c_double)]


class Connectivity(ctypes.Structure):
    _fields_ = [("n_interfaces", ctypes.c_int), ("neighbor_ids", ip), ("node_indices", ip),
                ("n_boundaries", ctypes.c_int), ("bnd_element", ip), ("bnd_node_indices", ip),
                ("bnd_kind", ip)]


class SweParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("n_elements", ctypes.c_int), ("gravity", ctypes.c_double),
                ("rotation_rate", ctypes.c_double), ("volume_flux", ctypes.c_int),
                ("surface_flux", ctypes.c_int), ("surface_noncons", ctypes.c_int),
                ("source_terms", ctypes.c_int)]


class Connectivity2D(ctypes.Structure):
    _fields_ = [("n_interfaces", ctypes.c_int), ("neighbor_ids", ip), ("node_indices", ip)]


def _stale():
    t = os.path.getmtime(ORACLE_LIB)
    return any(os.path.getmtime(os.path.join(ORACLE_DIR, f)) > t
               for f in os.listdir(ORACLE_DIR) if f.endswith((".cpp", ".h")))


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_LIB) or _stale():
            subprocess.run(["make", "-C", ORACLE_DIR, "-j8"], check=True, capture_output=True)
        L = ctypes.CDLL(ORACLE_LIB)
        L.ora_ln_mean.restype = ctypes.c_double
        L.ora_inv_ln_mean.restype = ctypes.c_double
        L.ora_ln_mean.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ora_inv_ln_mean.argtypes = [ctypes.c_double, ctypes.c_double]
        L.ora_e3_max_abs_speed.restype = ctypes.c_double
        L.ora_e3_entropy.restype = ctypes.c_double
        L.ora_e3_max_dt.restype = ctypes.c_double
        L.ora_swe_max_dt.restype = ctypes.c_double
        self.L = L

    # ---- basis
    def lgl(self, n):
        x = np.zeros(n)
        w = np.zeros(n)
        self.L.ora_lgl_nodes_weights(n, _d(x), _d(w))
        return x, w

    def dmat(self, n):
        x, w = self.lgl(n)
        D = np.zeros(n * n)
        self.L.ora_derivative_matrix(n, _d(x), _d(D))
        return D.reshape(n, n)

    def dsplit(self, n):
        x, w = self.lgl(n)
        D = np.zeros(n * n)
        self.L.ora_derivative_split(n, _d(x), _d(w), _d(D))
        return D.reshape(n, n)

    def ln_mean(self, x, y):
        return self.L.ora_ln_mean(x, y)

    def inv_ln_mean(self, x, y):
        return self.L.ora_inv_ln_mean(x, y)

    def ck2n54(self):
        A, B, c = np.zeros(5), np.zeros(5), np.zeros(5)
        self.L.ora_ck2n54_coefficients(_d(A), _d(B), _d(c))
        return A, B, c

    # ---- 3D Euler
    @staticmethod
    def params(n, n_elements, c_p=1004.0, c_v=717.0, gravity=9.81, volume_flux=0,
               volume_noncons=0, surface_flux=0, surface_noncons=0, lmars_c=0.0,
               coriolis_rate=0.0):
        return E3Params(n, n_elements, c_p, c_v, gravity, volume_flux, volume_noncons,
                        surface_flux, surface_noncons, lmars_c, coriolis_rate)

    def _vec(self, fn, p, *args):
        out = np.zeros(6)
        getattr(self.L, fn)(ctypes.byref(p), *args, _d(out))
        return out

    def flux(self, p, u, n):
        return self._vec("ora_e3_flux", p, _d(np.ascontiguousarray(u, float)), _d(np.ascontiguousarray(n, float)))

    def volume_flux(self, p, which, ul, ur, n):
        return self._vec("ora_e3_volume_flux", p, which, _d(np.ascontiguousarray(ul, float)),
                         _d(np.ascontiguousarray(ur, float)), _d(np.ascontiguousarray(n, float)))

    def noncons_flux(self, p, which, ul, ur, n):
        return self._vec("ora_e3_noncons_flux", p, which, _d(np.ascontiguousarray(ul, float)),
                         _d(np.ascontiguousarray(ur, float)), _d(np.ascontiguousarray(n, float)))

    def surface_flux(self, p, which, ul, ur, n):
        return self._vec("ora_e3_surface_flux", p, which, _d(np.ascontiguousarray(ul, float)),
                         _d(np.ascontiguousarray(ur, float)), _d(np.ascontiguousarray(n, float)))

    def cons2entropy(self, p, u):
        return self._vec("ora_e3_cons2entropy", p, _d(np.ascontiguousarray(u, float)))

    def entropy(self, p, u):
        return self.L.ora_e3_entropy(ctypes.byref(p), _d(np.ascontiguousarray(u, float)))

    def _conn(self, mesh, bc_kinds=None):
        keep = [np.ascontiguousarray(mesh.interface_neighbor_ids.ravel(order="F"), dtype=np.int32),
                np.ascontiguousarray(mesh.interface_node_indices.ravel(order="F"), dtype=np.int32),
                np.ascontiguousarray(mesh.boundary_element, dtype=np.int32),
                np.ascontiguousarray(mesh.boundary_node_indices.ravel(order="F"), dtype=np.int32),
                np.zeros(mesh.n_boundaries, dtype=np.int32)]
        c = Connectivity(mesh.n_interfaces, _i(keep[0]), _i(keep[1]), mesh.n_boundaries,
                         _i(keep[2]), _i(keep[3]), _i(keep[4]))
        return c, keep

    def _geo(self, mesh):
        cv = np.ascontiguousarray(mesh.contravariant_vectors.ravel(order="F"))
        ij = np.ascontiguousarray(mesh.inverse_jacobian.ravel(order="F"))
        return cv, ij

    def rhs(self, p, mesh, u):
        c, keep = self._conn(mesh)
        cv, ij = self._geo(mesh)
        uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)
        du = np.zeros_like(uf)
        rc = self.L.ora_e3_rhs(ctypes.byref(p), ctypes.byref(c), _d(cv), _d(ij), _d(uf), _d(du))
        assert rc == 0
        return du

    def rhs_magnitude(self, p, mesh, u):
        c, keep = self._conn(mesh)
        cv, ij = self._geo(mesh)
        uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)
        du = np.zeros_like(uf)
        assert self.L.ora_e3_rhs_magnitude(ctypes.byref(p), ctypes.byref(c), _d(cv), _d(ij),
                                           _d(uf), _d(du)) == 0
        return du

    def ck2n54_steps(self, p, mesh, u, dt, n_steps):
        c, keep = self._conn(mesh)
        cv, ij = self._geo(mesh)
        uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float).copy()
        rc = self.L.ora_e3_ck2n54(ctypes.byref(p), ctypes.byref(c), _d(cv), _d(ij), _d(uf),
                                  ctypes.c_double(dt), n_steps)
        assert rc == 0
        return uf

    def max_dt(self, p, mesh, u):
        cv, ij = self._geo(mesh)
        uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)
        return self.L.ora_e3_max_dt(ctypes.byref(p), _d(cv), _d(ij), _d(uf))

    def integrals(self, p, mesh, u, du=None):
        _, ij = self._geo(mesh)
        uf = np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)
        out = np.zeros(9)
        dptr = _d(np.ascontiguousarray(du)) if du is not None else None
        self.L.ora_e3_integrals(ctypes.byref(p), _d(ij), _d(uf), dptr, _d(out))
        return out


class SweOracle:
    """The covariant shallow-water oracle bound to one mesh / aux set:
    nbr [2,I], idx [2,2,I], aux [26,n,n,E], xyz [3,n,n,E] (F-ordered)."""

    def __init__(self, oracle, params, aux, xyz, nbr, idx):
        self.L, self.p = oracle.L, params
        f = lambda a, t: np.ascontiguousarray(np.asarray(a).ravel(order="F"), dtype=t)
        self._keep = [f(aux, float), f(xyz, float), f(nbr, np.int32), f(idx, np.int32)]
        self.c = Connectivity2D(int(np.asarray(nbr).shape[1]), _i(self._keep[2]),
                                _i(self._keep[3]))

    @classmethod
    def for_semi(cls, oracle, semi):
        return cls.from_parts(oracle, semi.mesh, semi.aux_node_vars, semi.equations, semi.solver,
                              semi.source_terms)

    @classmethod
    def from_parts(cls, oracle, mesh, aux, equations, solver, source_terms):
        codes = swe_codes(equations, solver, source_terms)
        p = SweParams(mesh.nnodes, mesh.n_elements, equations.gravity, equations.rotation_rate,
                      *codes)
        return cls(oracle, p, aux, mesh.node_coordinates, mesh.interface_neighbor_ids,
                   mesh.interface_node_indices)

    def _u(self, u):
        return np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)

    def rhs(self, u):
        uf = self._u(u)
        du = np.zeros_like(uf)
        assert self.L.ora_swe_rhs(ctypes.byref(self.p), ctypes.byref(self.c), _d(self._keep[0]),
                                  _d(self._keep[1]), _d(uf), _d(du)) == 0
        return du

    def rhs_magnitude(self, u):
        uf = self._u(u)
        du = np.zeros_like(uf)
        assert self.L.ora_swe_rhs_magnitude(ctypes.byref(self.p), ctypes.byref(self.c),
                                            _d(self._keep[0]), _d(self._keep[1]), _d(uf),
                                            _d(du)) == 0
        return du

    def max_dt(self, u):
        return self.L.ora_swe_max_dt(ctypes.byref(self.p), _d(self._keep[0]), _d(self._u(u)))

    def ck2n54_steps(self, u, dt, n_steps):
        uf = self._u(u).copy()
        assert self.L.ora_swe_ck2n54(ctypes.byref(self.p), ctypes.byref(self.c),
                                     _d(self._keep[0]), _d(self._keep[1]), _d(uf),
                                     ctypes.c_double(dt), n_steps) == 0
        return uf

    def integrals(self, u, du=None):
        out = np.zeros(3)
        dptr = _d(np.ascontiguousarray(du)) if du is not None else None
        self.L.ora_swe_integrals(ctypes.byref(self.p), _d(self._keep[0]), _d(self._u(u)), dptr,
                                 _d(out))
        return out

    def prim_vorticity(self, u):
        """cons2prim_and_vorticity: [6, nodes] (H, v1, v2, v3, b, vorticity)"""
        uf = self._u(u)
        out = np.zeros(2 * uf.size)
        self.L.ora_swe_prim_vorticity(ctypes.byref(self.p), _d(self._keep[0]), _d(uf), _d(out))
        return out.reshape(6, -1)

    def solve(self, u0, t_end, cfl):
        """StepsizeCallback(cfl) + CarpenterKennedy2N54 to t_end (the last
        step cut to end at t_end)."""
        u = self._u(u0).copy()
        t, steps = 0.0, 0
        while t < t_end:
            dt = min(cfl * self.max_dt(u), t_end - t)
            u = self.ck2n54_steps(u, dt, 1)
            t += dt
            steps += 1
        return u, steps


def swe_codes(equations, solver, source_terms):
    """(volume_flux, surface_flux, surface_noncons, source_terms) codes of a
    covariant semidiscretization (the oracle takes the library's codes)"""
    from trixiatmo_hip.semidiscretization import swe_codes as codes
    return codes(equations, solver, source_terms)


def params_for(semi_args):
    """oracle parameters for (mesh, equations, solver, source omega)"""
    mesh, eq, solver, omega = semi_args
    codes = solver.codes()
    return Oracle.params(mesh.nnodes, mesh.n_elements, eq.c_p, eq.c_v, eq.gravity,
                         codes["volume_flux"], codes["volume_noncons"], codes["surface_flux"],
                         codes["surface_noncons"], codes["lmars_c"], omega)


class RainyParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
                ("x0", ctypes.c_double), ("y0", ctypes.c_double), ("x1", ctypes.c_double),
                ("y1", ctypes.c_double), ("periodic_y", ctypes.c_int),
                ("volume_integral", ctypes.c_int), ("source", ctypes.c_int),
                ("surface_flux", ctypes.c_int)]


RAINY_SOURCES = {None: 0, "source_terms_rainy": 1, "source_terms_convergence_test_rainy": 2,
                 "source_terms_no_phase_change": 3}


class RainyOracle:
    """The rainy-Euler oracle bound to one StructuredMesh2D and solver choice:
    volume_integral 0 weak form / 1 flux_ec_rain, surface_flux 0 LLF / 1
    flux_ec_rain."""

    def __init__(self, oracle, mesh, volume_integral=1, source=None, surface_flux=0):
        self.L = oracle.L
        for fn in ("ora_rainy_max_dt", "ora_rainy_max_abs_speed_naive"):
            getattr(self.L, fn).restype = ctypes.c_double
        self.p = RainyParams(mesh.nnodes, mesh.nx, mesh.ny, mesh.x0, mesh.y0, mesh.x1, mesh.y1,
                             1 if mesh.periodic_y else 0, volume_integral,
                             RAINY_SOURCES[source], surface_flux)
        self.shape = (9, mesh.nnodes, mesh.nnodes, mesh.n_elements)

    def _u(self, u):
        return np.ascontiguousarray(np.asarray(u).ravel(order="F"), dtype=float)

    def rhs(self, u, t=0.0):
        uf = self._u(u)
        du = np.zeros_like(uf)
        assert self.L.ora_rainy_rhs(ctypes.byref(self.p), _d(uf), ctypes.c_double(t), _d(du)) == 0
        return du

    def max_dt(self, u):
        return self.L.ora_rainy_max_dt(ctypes.byref(self.p), _d(self._u(u)))

    def newton(self, u, tol=1e-9, max_iter=20):
        """NonlinearSolveDG on a copy of u: (u_out, iterations per node, failures)"""
        uf = self._u(u).copy()
        its = np.zeros(uf.size // 9, dtype=np.int32)
        fails = self.L.ora_rainy_newton(ctypes.byref(self.p), _d(uf), ctypes.c_double(tol),
                                        max_iter, its.ctypes.data_as(ip))
        return uf, its, fails

    def ck2n54(self, u, t, dt, n_steps=1, tol=1e-9, max_iter=20):
        uf = self._u(u).copy()
        assert self.L.ora_rainy_ck2n54(ctypes.byref(self.p), _d(uf), ctypes.c_double(t),
                                       ctypes.c_double(dt), n_steps, ctypes.c_double(tol),
                                       max_iter) == 0
        return uf

    def integrals(self, u, du=None):
        out = np.zeros(10)
        dptr = _d(np.ascontiguousarray(du)) if du is not None else None
        self.L.ora_rainy_integrals(ctypes.byref(self.p), _d(self._u(u)), dptr, _d(out))
        return out

    def solve(self, u0, t_end, cfl, tol=1e-9, max_iter=20):
        u = self._u(u0).copy()
        t, steps = 0.0, 0
        while t < t_end:
            dt = min(cfl * self.max_dt(u), t_end - t)
            u = self.ck2n54(u, t, dt, 1, tol, max_iter)
            t += dt
            steps += 1
        return u, steps

    # pointwise functions
    def residual(self, u, guess):
        r = np.zeros(3)
        self.L.ora_rainy_residual(_d(np.ascontiguousarray(u, float)),
                                  _d(np.ascontiguousarray(guess, float)), _d(r))
        return r

    def jacobian(self, u, guess):
        J = np.zeros(9)
        self.L.ora_rainy_residual_jacobian(_d(np.ascontiguousarray(u, float)),
                                           _d(np.ascontiguousarray(guess, float)), _d(J))
        return J.reshape(3, 3, order="F")

    def svp(self, T):
        a, b = ctypes.c_double(), ctypes.c_double()
        self.L.ora_rainy_saturation_vapour_pressure(ctypes.c_double(T), ctypes.byref(a),
                                                    ctypes.byref(b))
        return a.value, b.value

    def newton_node(self, u, guess=None, tol=1e-9, max_iter=20):
        uu = np.ascontiguousarray(u, float).copy()
        g = _d(np.ascontiguousarray(guess, float)) if guess is not None else None
        it = self.L.ora_rainy_newton_node(_d(uu), g, ctypes.c_double(tol), max_iter)
        return uu, it

    def _flux2(self, fn, ul, ur, n):
        f = np.zeros(9)
        getattr(self.L, fn)(_d(np.ascontiguousarray(ul, float)), _d(np.ascontiguousarray(ur, float)),
                            _d(np.ascontiguousarray(n, float)), _d(f))
        return f

    def flux_ec(self, ul, ur, n):
        return self._flux2("ora_rainy_flux_ec", ul, ur, n)

    def flux_llf(self, ul, ur, n):
        return self._flux2("ora_rainy_flux_llf", ul, ur, n)

    def flux(self, u, n):
        f = np.zeros(9)
        self.L.ora_rainy_flux(_d(np.ascontiguousarray(u, float)), _d(np.ascontiguousarray(n, float)),
                              _d(f))
        return f

    def source(self, u):
        s = np.zeros(9)
        self.L.ora_rainy_source(_d(np.ascontiguousarray(u, float)), _d(s))
        return s
