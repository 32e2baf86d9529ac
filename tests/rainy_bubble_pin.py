"""Shared set-up of the reference's rainy-bubble regression
(test/test_2d_rainy_euler.jl:79-114: examples/euler/precipitation/buoyancy/elixir_hoeck_bubble.jl
at cells_per_dimension = (16, 16), tspan = (0, 10)) and the time integrator
the reference ran it with.

The elixir integrates with OrdinaryDiffEq's adaptive SSPRK43 at its default
tolerances (abstol 1e-6, reltol 1e-3) with NonlinearSolveDG as the stage
limiter, and no step-size callback.  `ssprk43_adaptive` restates that
integrator on the oracle: Kraaijevanger's four-stage third-order SSP method
u1 = u + dt/2 L(u), u2 = u1 + dt/2 L(u1), u3 = 2/3 u + 1/3 (u2 + dt/2 L(u2)),
u_new = u3 + dt/2 L(u3), the stage limiter after each stage, the embedded
second-order solution 1/3 u + 2/3 (u2 + dt/2 L(u2)) for the error estimate,
the RMS norm of the error scaled by abstol + reltol max(|u|, |u_new|), a PI
step-size controller (beta1 = 7/30, beta2 = 2/15, safety 0.9, step ratio in
[1/10, 5]) and Hairer's initial step.  These controller details are
OrdinaryDiffEq defaults [ext]; with an I controller instead the regression
values move by less than 1e-5 relative (rain 1e-4), well inside the bars."""
import numpy as np

import trixiatmo_hip as T
from trixiatmo_hip import initial_conditions as ic
from trixiatmo_hip.meshes import StructuredMesh2D

EQ = T.equations.CompressibleRainyEulerEquations2D()

# test/test_2d_rainy_euler.jl:82-103
BUBBLE_L2 = np.array([7.959523735914366e-5, 4.8983178144498016e-5, 1.2695889115092138e-9,
                      0.019951374355214414, 0.033830586959875855, 126.25361431033272,
                      4.899300381610471e-5, 3.4512985541873764e-7, 0.0016455875546178209])
BUBBLE_LINF = np.array([0.0011817049726382534, 0.0007153644712404369, 2.0653953346431095e-8,
                        0.12229462313083396, 0.1760334548347, 1831.4029281113471,
                        0.0007153644712404369, 6.113680304613196e-6, 0.008505948056836132])
T_END = 10.0
CFL = 0.25  # the fixed-step CarpenterKennedy2N54 runs of the oracle and the GPU

_layers = None


def layers():
    global _layers
    if _layers is None:
        # AtmosphereLayersRainyBubble(equations; total_height = 2400 + 1)
        _layers = ic.AtmosphereLayersRainyBubble(EQ, total_height=2401.0)
    return _layers


def bubble_mesh():
    """StructuredMesh((16, 16), (0, 0), (2400, 2400), periodicity = (true, false)), polydeg 3"""
    return StructuredMesh2D((16, 16), (0.0, 0.0), (2400.0, 2400.0), polydeg=3,
                            periodicity=(True, False))


def exact(x, t):
    return ic.initial_condition_bubble_rainy(x, t, EQ, layers())


def ssprk43_adaptive(rhs, limiter, u0, t_end, abstol=1e-6, reltol=1e-3):
    """the adaptive SSPRK43 loop described in the module docstring on flat
    FP64 states; rhs(u, t) -> du, limiter(u) -> u; returns (u, steps)"""
    u = np.array(u0, dtype=float)
    t = 0.0
    f0 = rhs(u, t)
    sc = abstol + np.abs(u) * reltol
    d0, d1 = np.sqrt(np.mean((u / sc) ** 2)), np.sqrt(np.mean((f0 / sc) ** 2))
    dt0 = 0.01 * d0 / d1 if min(d0, d1) >= 1e-5 else 1e-6
    f1 = rhs(u + dt0 * f0, t + dt0)
    d2 = np.sqrt(np.mean(((f1 - f0) / sc) ** 2)) / dt0
    dt1 = (0.01 / max(d1, d2)) ** (1 / 3) if max(d1, d2) > 1e-15 else max(1e-6, dt0 * 1e-3)
    dt = min(100 * dt0, dt1)
    beta1, beta2, gamma, qmin, qmax = 7 / 30, 2 / 15, 0.9, 0.2, 10.0
    errold, steps, fs = 1e-4, 0, f0
    while t < t_end:
        dt = min(dt, t_end - t)
        h = 0.5 * dt
        a = limiter(u + h * fs)
        b = limiter(a + h * rhs(a, t + h))
        c3 = b + h * rhs(b, t + dt)
        embedded = (1 / 3) * u + (2 / 3) * c3
        c = limiter((2 / 3) * u + (1 / 3) * c3)
        un = limiter(c + h * rhs(c, t + h))
        sc = abstol + np.maximum(np.abs(u), np.abs(un)) * reltol
        err = np.sqrt(np.mean(((embedded - un) / sc) ** 2))
        if err <= 1.0:
            q = err ** beta1 / errold ** beta2 if err > 0 else 0.0
            q = min(1 / qmin, max(1 / qmax, q / gamma))
            u, t, steps = un, t + dt, steps + 1
            errold = max(err, 1e-4)
            fs = rhs(u, t)
        else:
            q = min(1 / qmin, max(1.0, err ** beta1 / gamma))
        dt = dt / q
    return u, steps
