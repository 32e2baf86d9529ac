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


# ---------------------------------------------------------------------------
# Covariant shallow water on the sphere (src/equations/reference_data.jl).
# Every function takes node positions x [3, ...] and the time t and returns the
# tuple (H, v_x, v_y, v_z, h_s) of total height, Cartesian velocity and bottom
# topography that the reference's initial conditions return for
# GlobalCartesianCoordinates; `transform_initial_condition` turns it into the
# conserved covariant variables with the auxiliary variables of each node.

EARTH_GRAVITATIONAL_ACCELERATION = 9.80616
EARTH_ROTATION_RATE = 7.292e-5
SECONDS_PER_DAY = 8.64e4


def _f32(v):
    """A Julia Float32 literal such as 7.848f-6, promoted to Float64."""
    return float(np.float32(v))


def _lon_lat(x):
    a = np.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2)
    return a, np.arctan2(x[1], x[0]), np.arcsin(x[2] / a)


def spherical2cartesian(vlon, vlat, vrad, x):
    """src/equations/equations.jl:294-307"""
    _, lon, lat = _lon_lat(x)
    sinlon, coslon = np.sin(lon), np.cos(lon)
    sinlat, coslat = np.sin(lat), np.cos(lat)
    vx = -sinlon * vlon - sinlat * coslon * vlat + coslat * coslon * vrad
    vy = coslon * vlon - sinlat * sinlon * vlat + coslat * sinlon * vrad
    vz = coslat * vlat + sinlat * vrad
    return vx, vy, vz


def _spherical_state(h, vlon, vlat, b, x):
    vx, vy, vz = spherical2cartesian(vlon, vlat, np.zeros_like(h), x)
    return h, vx, vy, vz, b


def bottom_topography_unsteady_solid_body_rotation(x):
    """reference_data.jl:386-388"""
    return 0.5 * (EARTH_ROTATION_RATE * x[2]) ** 2 / EARTH_GRAVITATIONAL_ACCELERATION


def initial_condition_unsteady_solid_body_rotation(x, t):
    """reference_data.jl:348-384 (Lauter et al. 2005)."""
    a = np.sqrt(x[0] ** 2 + x[1] ** 2 + x[2] ** 2)
    v_0 = 2.0 * np.pi * a / (12 * SECONDS_PER_DAY)
    k1 = 133681.0
    alpha = np.pi / 4
    wt = EARTH_ROTATION_RATE * t
    b1 = np.array([np.cos(wt), np.sin(wt), 0.0])
    b2 = np.array([-np.sin(wt), np.cos(wt), 0.0])
    b3 = np.array([0.0, 0.0, 1.0])
    c = np.array([-np.sin(alpha), 0.0, np.cos(alpha)])
    nrm = x / a
    omega_dot_x = EARTH_ROTATION_RATE * x[2]
    phi_t = np.array([c @ b1, c @ b2, c @ b3])
    v = v_0 * np.stack([phi_t[1] * nrm[2] - phi_t[2] * nrm[1],
                        phi_t[2] * nrm[0] - phi_t[0] * nrm[2],
                        phi_t[0] * nrm[1] - phi_t[1] * nrm[0]])
    pn = phi_t[0] * nrm[0] + phi_t[1] * nrm[1] + phi_t[2] * nrm[2]
    H = (-0.5 * (v_0 * pn + omega_dot_x) ** 2 + 0.5 * omega_dot_x ** 2 + k1) / \
        EARTH_GRAVITATIONAL_ACCELERATION
    return H, v[0], v[1], v[2], bottom_topography_unsteady_solid_body_rotation(x)


def initial_condition_geostrophic_balance(x, t):
    """reference_data.jl:145-162 (Williamson et al. 1992, case 2): zonal flow
    v_0 cos(lat), v_0 = 2 pi a / 12 days, in geostrophic balance with
    h = (2.94e4 - (a Omega v_0 + v_0^2 / 2) sin^2(lat)) / g; the 2.94f4 and
    0.5f0 literals are exact in Float32."""
    a, _, lat = _lon_lat(x)
    v_0 = 2.0 * np.pi * a / (12 * SECONDS_PER_DAY)
    vlon = v_0 * np.cos(lat)
    h = 1 / EARTH_GRAVITATIONAL_ACCELERATION * \
        (2.94e4 - (a * EARTH_ROTATION_RATE * v_0 + 0.5 * v_0 ** 2) * np.sin(lat) ** 2)
    return _spherical_state(h, vlon, np.zeros_like(h), np.zeros_like(h), x)


def bottom_topography_isolated_mountain(x):
    """reference_data.jl:286-294 (Williamson et al. 1992, case 5)."""
    a, lon, lat = _lon_lat(x)
    lon_0, lat_0 = -np.pi / 2, np.pi / 6
    b_0 = 2000.0
    R = np.pi / 9
    return b_0 * (1.0 - np.sqrt(np.minimum(R ** 2, (lon - lon_0) ** 2 + (lat - lat_0) ** 2)) / R)


def initial_condition_isolated_mountain(x, t):
    """reference_data.jl:262-275"""
    a, _, lat = _lon_lat(x)
    h_0, v_0 = 5960.0, 20.0
    vlon = v_0 * np.cos(lat)
    h = h_0 - 1 / EARTH_GRAVITATIONAL_ACCELERATION * \
        (a * EARTH_ROTATION_RATE * v_0 + 0.5 * v_0 ** 2) * np.sin(lat) ** 2
    return _spherical_state(h, vlon, np.zeros_like(h), bottom_topography_isolated_mountain(x), x)


def initial_condition_well_balanced(x, t):
    """examples/shallow_water/covariant/elixir_well_balanced.jl:13: H = 5960 m
    at rest (with the isolated-mountain topography as auxiliary field)."""
    H = np.full(x.shape[1:], 5960.0)
    z = np.zeros_like(H)
    return H, z, z, z, z


def initial_condition_rossby_haurwitz(x, t):
    """reference_data.jl:202-235 (Williamson et al. 1992, case 6); omega and
    K are Float32 literals in the reference and are promoted the same way."""
    a, lon, lat = _lon_lat(x)
    h_0 = 8.0e3
    omega = _f32(7.848e-6)
    K = _f32(7.848e-6)
    R = 4.0
    cl = np.cos(lat)
    A = 0.5 * omega * (2 * EARTH_ROTATION_RATE + omega) * cl ** 2 + \
        0.25 * K ** 2 * cl ** (2 * R) * ((R + 1) * cl ** 2 + (2 * R ** 2 - R - 2) -
                                         2 * R ** 2 / cl ** 2)
    B = 2 * (EARTH_ROTATION_RATE + omega) * K / ((R + 1) * (R + 2)) * cl ** R * \
        ((R ** 2 + 2 * R + 2) - (R + 1) ** 2 * cl ** 2)
    C = 0.25 * K ** 2 * cl ** (2 * R) * ((R + 1) * cl ** 2 - (R + 2))
    h = h_0 + (1 / EARTH_GRAVITATIONAL_ACCELERATION) * \
        (a ** 2 * A + a ** 2 * B * np.cos(R * lon) + a ** 2 * C * np.cos(2 * R * lon))
    vlon = a * omega * cl + a * K * cl ** (R - 1) * (R * np.sin(lat) ** 2 - cl ** 2) * \
        np.cos(R * lon)
    vlat = -a * K * R * cl ** (R - 1) * np.sin(lat) * np.sin(R * lon)
    return _spherical_state(h, vlon, vlat, np.zeros_like(h), x)


def galewsky_velocity(lat, u_0, lat_0, lat_1):
    """reference_data.jl:457-465"""
    lat = np.asarray(lat, dtype=float)
    inside = (lat_0 < lat) & (lat < lat_1)
    ls = np.where(inside, lat, 0.5 * (lat_0 + lat_1))
    u = u_0 / np.exp(-4 / (lat_1 - lat_0) ** 2) * np.exp(1 / (ls - lat_0) * 1 / (ls - lat_1))
    return np.where(inside, u, 0.0)


def _galewsky_integral(lat, u_0, lat_0, lat_1, a):
    """integral from pi/2 to lat of u (2 Omega sin + u tan / a) (the quadgk
    call of reference_data.jl:436-437); the integrand vanishes outside
    (lat_0, lat_1), so adaptive Gauss-Kronrod runs on that interval only."""
    from scipy.integrate import quad

    def f(s):
        u = float(galewsky_velocity(s, u_0, lat_0, lat_1))
        return u * (2 * EARTH_ROTATION_RATE * np.sin(s) + u * np.tan(s) / a)

    flat = np.ravel(lat)
    out = np.empty_like(flat)
    for q, l in enumerate(flat):
        lo = min(max(l, lat_0), lat_1)
        val, _ = quad(f, lo, lat_1, epsabs=0.0, epsrel=1e-13, limit=200)
        out[q] = -val
    return out.reshape(np.shape(lat))


def initial_condition_barotropic_instability(x, t):
    """reference_data.jl:425-455 (Galewsky et al. 2004)."""
    a, lon, lat = _lon_lat(x)
    u_0 = 80.0
    lat_0 = np.pi / 7
    lat_1 = np.pi / 2 - lat_0
    vlon = galewsky_velocity(lat, u_0, lat_0, lat_1)
    radius = float(np.ravel(a)[0])
    h = 10158.0 - a / EARTH_GRAVITATIONAL_ACCELERATION * \
        _galewsky_integral(lat, u_0, lat_0, lat_1, radius)
    alpha, beta = 1 / 3, 1 / 15
    lat_2 = np.pi / 4
    bump = 120.0 * np.cos(lat) * np.exp(-((lon / alpha) ** 2)) * \
        np.exp(-((lat_2 - lat) / beta) ** 2)
    h = np.where((-np.pi < lon) & (lon < np.pi), h + bump, h)
    return _spherical_state(h, vlon, np.zeros_like(h), np.zeros_like(h), x)


def initial_condition_zonal_jet_bumps(seed, radius=6.371e6, gravity=9.81,
                                      rotation_rate=EARTH_ROTATION_RATE):
    """Seeded balanced zonal jet plus three Gaussian height bumps: solid-body
    zonal flow u_0 cos(lat) with u_0 ~ U[20, 40] m/s in geostrophic balance
    with h_0 = 5960 m (Williamson et al. 1992, case 2 form), plus three bumps
    A exp(-(d / w)^2) of amplitude A ~ U[10, 120] m, width w ~ U[300, 1500] km
    and uniformly distributed centres (d = great-circle distance)."""
    rng = np.random.default_rng(seed)
    u_0 = rng.uniform(20.0, 40.0)
    amp = rng.uniform(10.0, 120.0, size=3)
    width = rng.uniform(3.0e5, 1.5e6, size=3)
    zc = rng.uniform(-1.0, 1.0, size=3)
    lc = rng.uniform(-np.pi, np.pi, size=3)
    centres = np.stack([np.sqrt(1 - zc ** 2) * np.cos(lc), np.sqrt(1 - zc ** 2) * np.sin(lc), zc])

    def ic(x, t):
        a, _, lat = _lon_lat(x)
        h = 5960.0 - (a * rotation_rate * u_0 + 0.5 * u_0 ** 2) * np.sin(lat) ** 2 / gravity
        nrm = x / a
        for m in range(3):
            cosd = np.clip(centres[0, m] * nrm[0] + centres[1, m] * nrm[1] +
                           centres[2, m] * nrm[2], -1.0, 1.0)
            d = a * np.arccos(cosd)
            h = h + amp[m] * np.exp(-(d / width[m]) ** 2)
        vlon = u_0 * np.cos(lat)
        return _spherical_state(h, vlon, np.zeros_like(h), np.zeros_like(h), x)

    return ic


def transform_initial_condition(initial_condition):
    """src/equations/equations.jl:94-101: (x, t, aux) -> conserved variables
    (h, h v^1, h v^2) with v^i = a^i . v and h = H - h_s, h_s from the
    auxiliary variables (row 19), aux [26, ...], x [3, ...]."""

    def transformed(x, t, aux):
        H, vx, vy, vz, _ = initial_condition(x, t)
        # basis_contravariant: 2 x 3, column-major in aux rows 6..11
        vcon1 = aux[6] * vx + aux[8] * vy + aux[10] * vz
        vcon2 = aux[7] * vx + aux[9] * vy + aux[11] * vz
        h = H - aux[19]
        return np.stack([h, h * vcon1, h * vcon2])

    return transformed


# ---------------------------------------------------------------------------
# Rainy Euler (CompressibleRainyEulerEquations2D)

def initial_condition_convergence_test_rainy(x, t, equations):
    """examples/euler/precipitation/tests/convergence_test.jl:8-40: a
    saturated travelling wave with rho = 2 + 0.1 sin(pi (x + y - t)),
    T = rho + 250, rho_vapour = e_s(T) / (R_v T), v = (1, 1); x [2, ...]."""
    eq = equations
    c = 2.0
    A = 0.1
    L = 2.0
    f = 1 / L
    w = 2 * np.pi * f
    rho = c + A * np.sin(w * (x[0] + x[1] - t))
    T = rho + 250.0
    rv = eq.saturation_vapour_pressure(T) / (eq.R_vapour * T)
    rc = rho / eq.c_liquid_water * 3000
    rr = rho / eq.c_liquid_water * 1000
    rm = rv + rc
    rd = rho - rm - rr
    E = (eq.c_dry_air_const_volume * rd + eq.c_vapour_const_volume * rv +
         eq.c_liquid_water * (rc + rr)) * T
    E = E + rv * eq.ref_latent_heat_vap_temp + rho
    return np.stack([rd, rm, rr, rho, rho, E, rv, rc, T])


def moist_hydrostatic_column(z, equations, theta=300.0, rel_humidity=0.5, p_surface=1.0e5,
                             dz_max=5.0):
    """pressure, temperature and vapour density at heights z of a hydrostatic
    atmosphere with constant dry potential temperature theta and constant
    relative humidity: dp/dz = -(rho_d + rho_v) g with
    T = theta (p / p_ref)^(R_d / c_pd), p_v = RH e_s(T), integrated by RK4
    with steps of at most dz_max from the surface."""
    eq = equations
    R_d, c_pd = eq.R_dry_air, eq.c_dry_air_const_pressure

    def state(p):
        T = theta * (p / eq.ref_pressure) ** (R_d / c_pd)
        pv = rel_humidity * eq.saturation_vapour_pressure(T)
        rv = pv / (eq.R_vapour * T)
        rd = (p - pv) / (R_d * T)
        return T, rv, rd

    def rhs(p):
        T, rv, rd = state(p)
        return -(rd + rv) * eq.gravity

    zs = np.unique(np.asarray(z, dtype=float).ravel())
    ps = np.empty_like(zs)
    p, zc = p_surface, 0.0
    for k, zt in enumerate(zs):
        nsub = max(1, int(np.ceil((zt - zc) / dz_max)))
        h = (zt - zc) / nsub
        for _ in range(nsub):
            k1 = rhs(p)
            k2 = rhs(p + 0.5 * h * k1)
            k3 = rhs(p + 0.5 * h * k2)
            k4 = rhs(p + h * k3)
            p = p + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        zc = zt
        ps[k] = p
    pz = np.interp(np.asarray(z, dtype=float), zs, ps)
    return pz


def initial_condition_rainy_bubble(mesh, equations, seed=0, theta=300.0, radius=2000.0,
                                   amplitude=2.0, relative_humidity=None):
    """Seeded moist warm bubble: a hydrostatic background with dry potential
    temperature 300 K and a relative humidity drawn from U[0.2, 0.95] (or the
    given one), plus a temperature bump amplitude cos^2(pi r / (2 R)) of
    R = 2 km at a random centre, at the background pressure and relative
    humidity; at rest, no cloud or rain, the saturation variables
    (rho_vapour, 0, T) as given (a relative humidity above 1 leaves the
    state supersaturated: the saturation solve then condenses cloud)."""
    eq = equations
    rng = np.random.default_rng(seed)
    rh = rng.uniform(0.2, 0.95)
    if relative_humidity is not None:
        rh = float(relative_humidity)
    xc = rng.uniform(mesh.x0 + radius, mesh.x1 - radius)
    zc = rng.uniform(mesh.y0 + radius, mesh.y0 + 0.5 * (mesh.y1 - mesh.y0))
    x, z = mesh.node_coordinates
    p = moist_hydrostatic_column(z - mesh.y0, eq, theta, rh)
    T = theta * (p / eq.ref_pressure) ** (eq.R_dry_air / eq.c_dry_air_const_pressure)
    r = np.sqrt((x - xc) ** 2 + (z - zc) ** 2)
    T = T + np.where(r < radius, amplitude * np.cos(0.5 * np.pi * r / radius) ** 2, 0.0)
    pv = rh * eq.saturation_vapour_pressure(T)
    rv = pv / (eq.R_vapour * T)
    rd = (p - pv) / (eq.R_dry_air * T)
    zero = np.zeros_like(rd)
    E = (eq.c_dry_air_const_volume * rd + eq.c_vapour_const_volume * rv) * T + \
        rv * eq.ref_latent_heat_vap_temp
    u = np.stack([rd, rv, zero, zero, zero, E, rv, zero, T])
    return u, {"relative_humidity": rh, "centre": (xc, zc)}


# ---------------------------------------------------------------------------
# rainy bubble of examples/euler/precipitation/buoyancy/elixir_hoeck_bubble.jl

def rainy_theta_d(z, equations):
    """hydrostatic dry potential temperature of the rainy bubble
    (src/equations/compressible_rainy_euler_2d.jl:1066-1083): 283 K at the
    surface pressure 85000 Pa, stratification 1.3e-5 / m"""
    eq = equations
    theta0 = 283.0 * (eq.ref_pressure / 85000.0) ** (eq.R_dry_air / eq.c_dry_air_const_pressure)
    return theta0 * np.exp(1.3e-5 * np.asarray(z, dtype=float))


def _svp_and_derivative(T, eq):
    """saturation vapour pressure and its temperature derivative
    (compressible_rainy_euler_2d.jl:397-446)"""
    k1 = (eq.c_vapour_const_pressure - eq.c_liquid_water) / eq.R_vapour
    k2 = (eq.ref_latent_heat_vap_temp -
          (eq.c_vapour_const_pressure - eq.c_liquid_water) * eq.ref_temperature) / eq.R_vapour
    es = eq.ref_saturation_pressure * (T / eq.ref_temperature) ** k1 * \
        np.exp(k2 * (1.0 / eq.ref_temperature - 1.0 / T))
    return es, es * (k1 / T + k2 / (T * T))


class AtmosphereLayersRainyBubble:
    """`AtmosphereLayersRainyBubble(equations; total_height, precision = 1)`
    (src/equations/compressible_rainy_euler_2d.jl:1086-1131): hydrostatic
    layers of (pressure, rho_dry, rho_vapour, T) at relative humidity 0.2,
    one every `precision` metres, each the root of the four residuals of
    generate_hydrostatic_residual (:997-1031) given the layer below: the
    first at z = 0.01 above the surface pressure 85000 Pa, then up in steps
    of `precision`.  The reference finds the roots with NLsolve to a residual
    tolerance (1e-8, then 1e-10); here Newton's method with the analytic
    Jacobian goes to the root (the two agree to ~1e-15 relative in the
    pressure).  Rows keep the reference's indexing: row n (0-based n - 1 in
    its 1-based table) is used as the layer at height (n - 1) precision."""

    def __init__(self, equations, total_height, precision=1.0):
        eq = equations
        self.total_height = float(total_height)
        self.precision = float(precision)
        n = int(round(self.total_height / self.precision))
        data = np.zeros((n + 1, 4))
        data[0] = self._solve(85000.0, 0.01, 0.01, np.array([85000.0, 1.4, 0.04, 300.0]), eq)
        z, dz = 0.01, self.precision
        for i in range(n):
            z += dz
            data[i + 1] = self._solve(data[i, 0], z, dz, data[i].copy(), eq)
        self.layer_data = data

    @staticmethod
    def _solve(p_lower, z, dz, x, eq):
        R_d, R_v, eps = eq.R_dry_air, eq.R_vapour, eq.R_dry_air / eq.R_vapour
        kappa = R_d / eq.c_dry_air_const_pressure
        H, g = 0.2, eq.gravity
        th = rainy_theta_d(z, eq)
        for _ in range(60):
            p, rd, rv, T = x
            es, des = _svp_and_derivative(T, eq)
            rvs = es / (R_v * T)
            drvs = (des * T - es) / (R_v * T * T)
            pk = (eq.ref_pressure / p) ** kappa
            r = np.array([(p - p_lower) / dz + (rd + rv) * g,
                          p - T * (rd * R_d + rv * R_v),
                          th - T * pk,
                          1000.0 * (rvs * (rd + rv / eps) * H - rv * (rd + rvs / eps))])
            J = np.array([[1.0 / dz, g, g, 0.0],
                          [1.0, -T * R_d, -T * R_v, -(rd * R_d + rv * R_v)],
                          [T * kappa * pk / p, 0.0, 0.0, -pk],
                          [0.0, 1000.0 * (rvs * H - rv), 1000.0 * (rvs * H / eps - rd - rvs / eps),
                           1000.0 * (drvs * (rd + rv / eps) * H - rv * drvs / eps)]])
            step = np.linalg.solve(J, r)
            x = x - step
            if np.all(np.abs(step) <= 4e-16 * np.abs(x)):
                break
        return x


def initial_condition_bubble_rainy(x, t, equations, atmosphere_layers):
    """`initial_condition_bubble_rainy` (src/equations/compressible_rainy_euler_2d.jl:1133-1206)
    at points x [2, ...]: relative humidity 0.2 outside a bubble centred at
    (1200, 800) m, 1 inside radius 200 m with a cos^2 transition to 300 m;
    the layer pressure interpolated linearly in height; (rho_dry,
    rho_vapour, T) the root of generate_perturbation_residual (:1033-1064)
    from the layer's values (Newton's method to the root; the reference's
    NLsolve stops at a residual of 1e-9); at rest, no cloud or rain."""
    eq = equations
    R_d, R_v, eps = eq.R_dry_air, eq.R_vapour, eq.R_dry_air / eq.R_vapour
    kappa = R_d / eq.c_dry_air_const_pressure
    xs = np.asarray(x[0], dtype=float)
    z = np.asarray(x[1], dtype=float)
    r = np.sqrt((xs - 1200.0) ** 2 + (z - 800.0) ** 2)
    H = np.where(r > 300.0, 0.2,
                 np.where(r > 200.0,
                          0.2 + 0.8 * np.cos(np.pi * (r - 200.0) / 200.0) ** 2, 1.0))
    L = atmosphere_layers
    dz = L.precision
    n = np.floor((z + np.finfo(float).eps) / dz).astype(np.int64) + 1  # 1-based, as the reference
    z_lower = (n - 1) * dz
    z_upper = n * dz
    top = z_lower == L.total_height
    z_upper = np.where(top, z_lower + dz, z_upper)
    n = np.where(top, n - 1, n)
    n = np.where(n == 0, 1, n)
    if np.any((z > L.total_height) & ~np.isclose(z, L.total_height)):
        raise ValueError("The atmosphere does not match the simulation domain")
    data = L.layer_data
    p_lo, p_up = data[n - 1, 0], data[n, 0]
    ph = (p_up * (z - z_lower) + p_lo * (z_upper - z)) / dz
    # the root: T from the potential-temperature residual, then Newton on
    # (rho_dry, rho_vapour) for the humidity and pressure residuals
    T = rainy_theta_d(z, eq) / (eq.ref_pressure / ph) ** kappa
    es, _ = _svp_and_derivative(T, eq)
    rvs = es / (R_v * T)
    rd, rv = data[n - 1, 1].copy(), data[n - 1, 2].copy()
    for _ in range(60):
        f1 = 30.0 * (rvs * (rd + rv / eps) * H - rv * (rd + rvs / eps))
        f2 = ph - (rd * R_d + rv * R_v) * T
        a11 = 30.0 * (rvs * H - rv)
        a12 = 30.0 * (rvs * H / eps - rd - rvs / eps)
        a21, a22 = -R_d * T, -R_v * T
        det = a11 * a22 - a12 * a21
        d_rd = (f1 * a22 - a12 * f2) / det
        d_rv = (a11 * f2 - a21 * f1) / det
        rd, rv = rd - d_rd, rv - d_rv
        if np.all(np.abs(d_rd) <= 4e-16 * np.abs(rd)) and np.all(np.abs(d_rv) <= 4e-16 * np.abs(rv)):
            break
    E = (eq.c_dry_air_const_volume * rd + eq.c_vapour_const_volume * rv) * T + \
        rv * eq.ref_latent_heat_vap_temp
    zero = np.zeros_like(rd)
    return np.stack([rd, rv, zero, zero, zero, E, rv, zero, T])
