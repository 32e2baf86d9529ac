"""Seeded synthetic initial states for the workloads this library is measured on.

Generators use numpy's PCG64 (`np.random.default_rng(seed)`); the state is
evaluated on the host at the mesh nodes and handed over in Trixi layout.
"""
import numpy as np

from .equations import EARTH_RADIUS


def _modes(rng, n_modes=4):
    k = rng.integers(1, 4, size=(n_modes, 3))
    phase = rng.uniform(0.0, 2.0 * np.pi, size=n_modes)
    return k, phase


def _mode_sum(x, k, phase):
    # x: [3, ...]; f = 1/4 sum_m sin(2 pi k_m . x + phi_m)
    out = np.zeros(x.shape[1:])
    for km, pm in zip(k, phase):
        out += np.sin(2.0 * np.pi * (km[0] * x[0] + km[1] * x[1] + km[2] * x[2]) + pm)
    return 0.25 * out


def initial_condition_periodic_modes(mesh, equations, seed=0):
    """rho = 1 + 0.05 f, (v1, v2, v3) = 0.1 (g1, g2, g3), p = 1 + 0.05 h, phi = 0,
    each field a sum of 4 seeded sine modes with integer wavenumbers 1..3 per
    axis and uniform phases (periodic unit cube)."""
    rng = np.random.default_rng(seed)
    x = mesh.node_coordinates
    fields = [_mode_sum(x, *_modes(rng)) for _ in range(5)]
    rho = 1.0 + 0.05 * fields[0]
    v1, v2, v3 = (0.1 * fields[1], 0.1 * fields[2], 0.1 * fields[3])
    p = 1.0 + 0.05 * fields[4]
    phi = np.zeros_like(rho)
    cons = equations.prim2cons(rho, v1, v2, v3, p, phi)
    return np.stack(cons, axis=0)  # [6, n, n, n, E]


def initial_condition_isothermal_perturbed(mesh, equations, seed=0, T0=250.0, n_bumps=10,
                                           amplitude=1.0, radius=500e3,
                                           inner_radius=EARTH_RADIUS):
    """Isothermal hydrostatic shell at rest (the form of
    src/equations/initial_condition_compressible_euler_3d.jl:171-196 with
    T0 = 250 K): phi = g (a - a^2/r), p = p0 exp(-phi / (R T0)), plus seeded
    Gaussian temperature bumps (great-circle e-folding radius `radius`,
    vertically uniform); rho = p / (R T)."""
    rng = np.random.default_rng(seed)
    x = mesh.node_coordinates
    r = np.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2)
    g = equations.gravity
    a = inner_radius
    phi = g * (a - a * a / r)
    p = equations.p_0 * np.exp(-phi / (equations.R * T0))
    unit = x / r
    T = np.full_like(r, T0)
    for _ in range(n_bumps):
        z = rng.uniform(-1.0, 1.0)
        lon = rng.uniform(0.0, 2.0 * np.pi)
        c = np.array([np.sqrt(1 - z * z) * np.cos(lon), np.sqrt(1 - z * z) * np.sin(lon), z])
        cosang = np.clip(unit[0] * c[0] + unit[1] * c[1] + unit[2] * c[2], -1.0, 1.0)
        dist = a * np.arccos(cosang)
        T = T + amplitude * np.exp(-(dist / radius) ** 2)
    rho = p / (equations.R * T)
    zero = np.zeros_like(rho)
    cons = equations.prim2cons(rho, zero, zero, zero, p, phi)
    return np.stack(cons, axis=0)


def _cartesian_to_sphere(x):
    """src/auxiliary/auxiliary.jl:25-34: (lambda in [0, 2 pi), latitude, r)."""
    r = np.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2)
    lam = np.arctan2(x[1], x[0])
    lam = np.where(lam < 0, lam + 2 * np.pi, lam)
    phi = np.arcsin(x[2] / r)
    return lam, phi, r


def _baroclinic_basic_state(lon, lat, z):
    """src/equations/initial_condition_compressible_euler_3d.jl:1-61
    (Ullrich et al. 2013, Table 1 and Appendix A)."""
    a = EARTH_RADIUS
    b = 2.0
    g = 9.80616
    k = 3.0
    p0 = 1.0e5
    R = 287.0
    TE, TP = 310.0, 240.0
    lapse = 0.005
    omega = 7.292e-5
    r = z + a
    T0 = 0.5 * (TE + TP)
    A = 1.0 / lapse
    B = (T0 - TP) / (T0 * TP)
    C = 0.5 * (k + 2) * (TE - TP) / (TE * TP)
    H = R * T0 / g
    sz = z / (b * H)
    t1 = np.exp(lapse / T0 * z)
    t2 = np.exp(-sz ** 2)
    tau1 = A * lapse / T0 * t1 + B * (1 - 2 * sz ** 2) * t2
    tau2 = C * (1 - 2 * sz ** 2) * t2
    it1 = A * (t1 - 1) + B * z * t2
    it2 = C * z * t2
    t3 = r / a * np.cos(lat)
    t4 = t3 ** k - k / (k + 2) * t3 ** (k + 2)
    T = 1.0 / ((r / a) ** 2 * (tau1 - tau2 * t4))
    big_u = g / a * k * T * it2 * (t3 ** (k - 1) - t3 ** (k + 1))
    t5 = a * np.cos(lat)
    u = -omega * t5 + np.sqrt(omega ** 2 * t5 ** 2 + t5 * big_u)
    p = p0 * np.exp(-g / R * (it1 - it2 * t4))
    rho = p / (R * T)
    return rho, u, p


def _baroclinic_perturbation(lon, lat, z):
    """:63-106 (stream-function perturbation, analytical derivative)."""
    d0 = 1.0 / 6.0
    vp = 1.0
    plon = np.pi / 9.0
    plat = 2.0 * np.pi / 9.0
    pertz = 15000.0
    gc = np.arccos(np.clip(np.sin(plat) * np.sin(lat) +
                           np.cos(plat) * np.cos(lat) * np.cos(lon - plon), -1.0, 1.0))
    inside = (z <= pertz) & (gc <= d0)
    taper = 1 - 3 * z ** 2 / pertz ** 2 + 2 * z ** 3 / pertz ** 3
    s_ = np.sin(0.5 * np.pi * gc / d0)
    c_ = np.cos(0.5 * np.pi * gc / d0)
    factor = 16 / (3 * np.sqrt(3)) * vp * taper * c_ ** 3 * s_
    with np.errstate(divide="ignore", invalid="ignore"):
        up = -factor * (-np.sin(plat) * np.cos(lat) +
                        np.cos(plat) * np.sin(lat) * np.cos(lon - plon)) / np.sin(gc)
        vpp = factor * np.cos(plat) * np.sin(lon - plon) / np.sin(gc)
    up = np.where(inside, up, 0.0)
    vpp = np.where(inside, vpp, 0.0)
    return up, vpp


def initial_condition_baroclinic_instability(x, equations):
    """src/equations/initial_condition_compressible_euler_3d.jl:132-158 at
    Cartesian points x [3, ...]; returns conserved variables [6, ...]."""
    lon, lat, r = _cartesian_to_sphere(x)
    z = np.maximum(r - EARTH_RADIUS, 0.0)
    rho, u, p = _baroclinic_basic_state(lon, lat, z)
    up, vp = _baroclinic_perturbation(lon, lat, z)
    u = u + up
    v = vp
    v1 = -np.sin(lon) * u - np.sin(lat) * np.cos(lon) * v
    v2 = np.cos(lon) * u - np.sin(lat) * np.sin(lon) * v
    v3 = np.cos(lat) * v
    phi = -EARTH_RADIUS ** 2 * 9.80616 / r
    return np.stack(equations.prim2cons(rho, v1, v2, v3, p, phi), axis=0)
