"""Rainy Euler in 2D over libatmo_hip: the semidiscretization on a uniform
Cartesian StructuredMesh with NonlinearSolveDG as stage limiter, the
standalone saturation solve, and the MLP that seeds it.

Mirrors `SemidiscretizationHyperbolic(mesh, CompressibleRainyEulerEquations2D(),
initial_condition, DGSEM(...); source_terms, boundary_conditions)` and
`NonlinearSolveDG(saturation_residual, saturation_residual_jacobian,
SVector(7, 8, 9); tolerance, max_iterations)`
(src/callbacks_stage/nonlinear_solve_dg.jl:19-45).  Device arrays (training
samples, iteration counts) are torch tensors on the GPU.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import (AtmoNewtonStats, AtmoRainyConfig, c_float_p, check, dptr, lib)
from .solvers import VolumeIntegralFluxDifferencing, VolumeIntegralWeakForm

RAINY_SURFACE_FLUXES = {"flux_lax_friedrichs": 0, "flux_ec_rain": 1}
RAINY_SOURCES = {None: 0, "source_terms_rainy": 1, "source_terms_convergence_test_rainy": 2,
                 "source_terms_no_phase_change": 3}


@dataclass(frozen=True)
class NonlinearSolveDG:
    """the stage limiter solving the saturation system at every node"""
    tolerance: float = 1e-9
    max_iterations: int = 20


def _ptr(t):
    """device pointer of a torch tensor for the library, which runs on its own
    stream and reads the memory at once: work that torch has queued on the
    tensor's stream is finished first (the library's entry points synchronise
    their own stream before they return, so results are complete for torch)"""
    if t is None:
        return None
    if t.is_cuda:
        import torch
        torch.cuda.current_stream(t.device).synchronize()
    return ctypes.c_void_p(t.data_ptr())


class SemidiscretizationRainy:
    def __init__(self, mesh, equations, solver, source_terms=None, stage_limiter=None):
        self.mesh, self.equations, self.solver = mesh, equations, solver
        self.source_terms = source_terms
        self.stage_limiter = stage_limiter
        vi = solver.volume_integral
        if isinstance(vi, VolumeIntegralWeakForm):
            vol = 0
        elif isinstance(vi, VolumeIntegralFluxDifferencing) and vi.volume_flux in (
                "flux_ec_rain", ("flux_ec_rain",)):
            vol = 1
        else:
            raise ValueError(f"unsupported volume integral {vi}")
        sf = solver.surface_flux if isinstance(solver.surface_flux, str) else solver.surface_flux[0]
        lim = stage_limiter or NonlinearSolveDG(1e-9, 0)
        cfg = AtmoRainyConfig(polydeg=mesh.polydeg, cells_x=mesh.nx, cells_y=mesh.ny,
                              x_min=mesh.x0, y_min=mesh.y0, x_max=mesh.x1, y_max=mesh.y1,
                              periodic_y=1 if mesh.periodic_y else 0, volume_integral=vol,
                              surface_flux=RAINY_SURFACE_FLUXES[sf],
                              source_terms=RAINY_SOURCES[source_terms],
                              newton_tolerance=float(lim.tolerance),
                              newton_max_iterations=int(lim.max_iterations) if stage_limiter else 0)
        h = ctypes.c_void_p()
        check(lib().atmo_rainy_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        n = mesh.nnodes
        self.shape = (9, n, n, mesh.n_elements)
        self._mlp = None

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
        return int(np.prod(self.shape[1:]))

    def _flat(self, u):
        a = np.asarray(u, dtype=np.float64)
        if a.shape == self.shape:
            a = a.ravel(order="F")
        if a.size != int(np.prod(self.shape)):
            raise ValueError("state has the wrong size")
        return np.ascontiguousarray(a)

    def rhs(self, du, u, t=0.0):
        uf = self._flat(u)
        out = np.empty_like(uf)
        check(lib().atmo_rhs(self._h, dptr(out), dptr(uf), float(t)), self._h)
        if du is not None:
            du[...] = out.reshape(du.shape, order="F") if du.shape == self.shape else out
        return out

    def set_state(self, u, t=0.0):
        uf = self._flat(u)
        check(lib().atmo_set_state(self._h, dptr(uf)), self._h)
        check(lib().atmo_set_time(self._h, float(t)), self._h)

    def get_state(self):
        out = np.empty(int(np.prod(self.shape)))
        check(lib().atmo_get_state(self._h, dptr(out)), self._h)
        return out

    def step(self, dt, n_steps=1):
        check(lib().atmo_step(self._h, float(dt), int(n_steps)), self._h)

    def set_time(self, t):
        check(lib().atmo_set_time(self._h, float(t)), self._h)

    def solve_resident(self, t_end, cfl, check_every=64):
        """StepsizeCallback(cfl) time loop from the resident time to t_end on
        the device, the stage limiter in every stage; returns (steps, time
        reached)."""
        from .semidiscretization import solve_resident
        return solve_resident(self._h, 1, t_end, cfl, check_every)

    def max_dt(self, cfl):
        dt = ctypes.c_double()
        check(lib().atmo_max_dt(self._h, float(cfl), ctypes.byref(dt)), self._h)
        return dt.value

    def integrals(self):
        out = np.zeros(9)
        check(lib().atmo_integrals(self._h, 0, dptr(out)), self._h)
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

    # saturation solve
    def newton(self, seeded=False, iterations=None):
        """NonlinearSolveDG on the resident state; `iterations` an int32 device
        tensor of one entry per node (optional).  `seeded`: False (the
        previous u[7:9]), True or "network" (the attached MLP), or
        "closed_form" (all vapour with T from the energy row, one saturation
        adjustment where that is supersaturated).  Returns the statistics."""
        st = AtmoNewtonStats()
        check(lib().atmo_saturation_newton(self._h, _seed_mode(seeded), _ptr(iterations),
                                           ctypes.byref(st)), self._h)
        return _stats(st)

    def enable_newton_stats(self, on=True):
        check(lib().atmo_enable_newton_stats(self._h, 1 if on else 0), self._h)

    def newton_stats(self):
        st = AtmoNewtonStats()
        check(lib().atmo_newton_stats_read(self._h, ctypes.byref(st)), self._h)
        return _stats(st)

    def features(self, x, y):
        """fill device tensors x [nodes, 3], y [nodes, 2] (float32) from the
        resident state"""
        check(lib().atmo_rainy_features(self._h, _ptr(x), _ptr(y)), self._h)

    def attach_mlp(self, mlp):
        self._mlp = mlp
        check(lib().atmo_rainy_attach_mlp(self._h, mlp.handle if mlp else None), self._h)

    def seed_stage_limiter(self, on=True):
        """seed the Newton solve of every stage: True or "network" with the
        attached network, "closed_form" with the closed-form guess, False not"""
        check(lib().atmo_rainy_seed_stage_limiter(self._h, _seed_mode(on)), self._h)


def _seed_mode(seeded):
    modes = {False: 0, True: 1, "network": 1, "closed_form": 2, None: 0}
    if seeded not in modes:
        raise ValueError(f"unknown seed {seeded!r}")
    return modes[seeded]


def _stats(st):
    return {"nodes": st.nodes, "updates": st.updates, "max_updates": st.max_updates,
            "failures": st.failures, "fallbacks": st.fallbacks,
            "updates_per_node": st.updates / max(st.nodes, 1)}


class SaturationMLP:
    """3 -> 32 -> 32 -> 2 tanh network in FP32 that maps (rho, internal energy
    density, total water) to (T, rho_vapour / rho), trained and evaluated by
    the library's HIP kernels."""

    N_PARAMS = 1250

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        check(lib().atmo_mlp_create(int(seed), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().atmo_mlp_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_normalization(self, in_mean, in_scale, out_mean, out_scale):
        norm = np.ascontiguousarray(np.concatenate([in_mean, in_scale, out_mean, out_scale]),
                                    dtype=np.float32)
        check(lib().atmo_mlp_set_normalization(self._h, norm.ctypes.data_as(c_float_p)))
        self.norm = norm

    def params(self):
        p = np.zeros(self.N_PARAMS, dtype=np.float32)
        check(lib().atmo_mlp_get_params(self._h, p.ctypes.data_as(c_float_p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().atmo_mlp_set_params(self._h, p.ctypes.data_as(c_float_p)))

    def loss_grad(self, x, y):
        loss = ctypes.c_float()
        g = np.zeros(self.N_PARAMS, dtype=np.float32)
        check(lib().atmo_mlp_loss_grad(self._h, _ptr(x), _ptr(y), int(x.shape[0]),
                                       ctypes.byref(loss), g.ctypes.data_as(c_float_p)))
        return loss.value, g

    def train(self, x, y, perm, batch, n_batches, lr):
        loss = ctypes.c_float()
        check(lib().atmo_mlp_train(self._h, _ptr(x), _ptr(y), _ptr(perm), int(x.shape[0]),
                                   int(batch), int(n_batches), float(lr), ctypes.byref(loss)))
        return loss.value

    def forward(self, x, y):
        check(lib().atmo_mlp_forward(self._h, _ptr(x), _ptr(y), int(x.shape[0])))


def harvest_features(semi, states, steps=0, dt=None):
    """training samples from rainy states: for each state, `steps` stage-limited
    RK steps on the device, then (rho, internal energy, total water) ->
    (T, q_v) of every node; returns float32 device tensors x [N, 3], y [N, 2]"""
    import torch
    xs, ys = [], []
    n = semi.ndofs
    for u in states:
        semi.set_state(u)
        if steps:
            semi.step(dt if dt is not None else semi.max_dt(0.5), steps)
        x = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        y = torch.empty((n, 2), dtype=torch.float32, device="cuda")
        semi.features(x, y)
        xs.append(x)
        ys.append(y)
    return torch.cat(xs), torch.cat(ys)


def train_saturation_mlp(x, y, epochs=20, batch=4096, lr=3e-3, seed=0, log=None):
    """train a SaturationMLP on device samples with the library's kernels:
    inputs and outputs standardised by their means and standard deviations,
    Adam on mini-batches of a fresh permutation every epoch, the learning
    rate halved over the last third.  Returns (mlp, loss history)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    mlp = SaturationMLP(seed)
    xm, xs = x.double().mean(0), x.double().std(0)
    ym, ys = y.double().mean(0), y.double().std(0)
    mlp.set_normalization(xm.cpu().numpy(), xs.cpu().numpy() + 1e-30, ym.cpu().numpy(),
                          ys.cpu().numpy() + 1e-30)
    n = x.shape[0]
    n_batches = n // batch
    history = []
    for ep in range(epochs):
        perm = torch.randperm(n, device="cuda", generator=g).to(torch.int32)
        rate = lr if ep < 2 * epochs // 3 else 0.5 * lr
        loss = mlp.train(x, y, perm, batch, n_batches, rate)
        history.append(loss)
        if log:
            log(f"mlp epoch {ep}: last batch loss {loss:.3e}")
    return mlp, history
