"""Equation systems and flux names (the Julia function names of TrixiAtmo/Trixi)."""
from dataclasses import dataclass

# src/equations/reference_data.jl:11-61
EARTH_RADIUS = 6.37122e6
EARTH_GRAVITATIONAL_ACCELERATION = 9.80616
EARTH_ROTATION_RATE = 7.292e-5

# two-point volume fluxes -> ATMO_VOL_*
VOLUME_FLUXES = {"flux_ranocha": 0, "flux_kennedy_gruber": 1, "flux_shima_etal": 2}
# nonconservative gravity terms -> ATMO_NC_*
NONCONSERVATIVE_FLUXES = {"flux_zero": 0, "flux_nonconservative_waruszewski_etal": 1,
                          "flux_nonconservative_souza_etal": 2}
# surface fluxes -> ATMO_SURF_*
SURFACE_FLUXES = {"flux_lax_friedrichs": 0, "flux_ranocha": 1, "flux_kennedy_gruber": 2,
                  "FluxLMARS": 3,
                  "FluxPlusDissipation(flux_kennedy_gruber, DissipationLocalLaxFriedrichs())": 4,
                  "FluxPlusDissipation(flux_ranocha, DissipationLocalLaxFriedrichs())": 5}


# covariant shallow water: volume fluxes -> ATMO_SWE_VOL_*, surface fluxes ->
# ATMO_SWE_SURF_*, surface nonconservative terms -> ATMO_SWE_NC_*
SWE_VOLUME_FLUXES = {"flux_ec": 0, "flux_central": 1}
# the standard (non-split) covariant form: VolumeIntegralWeakForm() and its
# flux_lax_friedrichs (ATMO_SWE_VOL_WEAK_FORM, ATMO_SWE_SURF_LAX_FRIEDRICHS_STANDARD)
SWE_VOLUME_WEAK_FORM = 2
SWE_SURFACE_LLF_STANDARD = 3
SWE_VOLUME_NONCONS = {"flux_nonconservative_ec"}
SWE_SURFACE_FLUXES = {"FluxPlusDissipation(flux_ec, DissipationLocalLaxFriedrichs())": 0,
                      "flux_lax_friedrichs": 1, "flux_ec": 2}
SWE_SURFACE_NONCONS = {"flux_nonconservative_surface_simplified": 0,
                       "flux_nonconservative_ec": 1}


def FluxPlusDissipation(numerical_flux, dissipation="DissipationLocalLaxFriedrichs()"):
    """`FluxPlusDissipation(flux, DissipationLocalLaxFriedrichs())`: a two-point
    flux plus local Lax-Friedrichs dissipation -lambda_max/2 [u] (the
    construction of src/equations/equations.jl:256-290)."""
    name = f"FluxPlusDissipation({numerical_flux}, {dissipation})"
    if name not in SURFACE_FLUXES and name not in SWE_SURFACE_FLUXES:
        raise ValueError(f"unsupported surface flux {name}")
    return name


@dataclass(frozen=True)
class CompressibleEulerEnergyEquationsWithGravity3D:
    """src/equations/compressible_euler_energy_with_gravity_3d.jl:45-68:
    variables (rho, rho_v1, rho_v2, rho_v3, rho_e, phi)."""
    c_p: float
    c_v: float
    gravity: float
    p_0: float = 100_000.0

    @property
    def gamma(self):
        return self.c_p / self.c_v

    @property
    def R(self):
        return self.c_p - self.c_v

    nvariables = 6
    varnames = ("rho", "rho_v1", "rho_v2", "rho_v3", "rho_e", "phi")

    def prim2cons(self, rho, v1, v2, v3, p, phi):
        """:626-636"""
        e = p / (self.gamma - 1.0) + 0.5 * rho * (v1 * v1 + v2 * v2 + v3 * v3)
        return rho, rho * v1, rho * v2, rho * v3, e, phi


@dataclass(frozen=True)
class FluxLMARS:
    """FluxLMARS(c): low-Mach approximate Riemann solver with sound speed c."""
    speed_of_sound: float


@dataclass(frozen=True)
class SplitCovariantShallowWaterEquations2D:
    """src/equations/covariant_shallow_water_split.jl:1-64: shallow water on a
    2D manifold in 3D, variables (h, h v^1, h v^2) with contravariant
    momentum, split form with entropy-conservative two-point fluxes.  The
    global coordinate system is GlobalCartesianCoordinates()."""
    gravity: float
    rotation_rate: float
    global_coordinate_system: str = "GlobalCartesianCoordinates()"

    nvariables = 3
    varnames = ("h", "h_vcon1", "h_vcon2")

    def __post_init__(self):
        if self.global_coordinate_system != "GlobalCartesianCoordinates()":
            raise ValueError("the covariant path uses GlobalCartesianCoordinates()")


@dataclass(frozen=True)
class CovariantShallowWaterEquations2D:
    """src/equations/covariant_shallow_water.jl:1-100: the standard covariant
    shallow-water form, variables (h, h v^1, h v^2), conservative flux with the
    pressure term G^{ij} g h^2 / 2 (no nonconservative terms, no bottom
    topography), for VolumeIntegralWeakForm().  Global coordinates are
    Cartesian: the reference's GlobalSphericalCoordinates() give the same
    metric and the same map between neighbouring contravariant bases up to
    rounding (the state is contravariant in both)."""
    gravity: float
    rotation_rate: float
    global_coordinate_system: str = "GlobalCartesianCoordinates()"

    nvariables = 3
    varnames = ("h", "h_vcon1", "h_vcon2")

    def __post_init__(self):
        if self.global_coordinate_system != "GlobalCartesianCoordinates()":
            raise ValueError("the covariant path uses GlobalCartesianCoordinates()")


@dataclass(frozen=True)
class CompressibleRainyEulerEquations2D:
    """src/equations/compressible_rainy_euler_2d.jl:47-113: variables (rho_dry,
    rho_moist, rho_rain, rho_v1, rho_v2, energy_density, rho_vapour, rho_cloud,
    temperature); the last three are the solution of the saturation system."""
    c_liquid_water: float = 4186.0
    c_dry_air_const_pressure: float = 1004.0
    c_dry_air_const_volume: float = 717.0
    c_vapour_const_pressure: float = 1885.0
    c_vapour_const_volume: float = 1424.0
    ref_saturation_pressure: float = 610.7
    ref_temperature: float = 273.15
    ref_latent_heat_vap_temp: float = 2500000.0
    ref_pressure: float = 100000.0
    gravity: float = 9.80616
    rain_water_distr: float = 8000000.0
    v_mean_rain: float = 130.0

    nvariables = 9
    varnames = ("rho_dry", "rho_moist", "rho_rain", "rho_v1", "rho_v2", "energy_density",
                "rho_vapour", "rho_cloud", "temperature")

    @property
    def R_dry_air(self):
        return self.c_dry_air_const_pressure - self.c_dry_air_const_volume

    @property
    def R_vapour(self):
        return self.c_vapour_const_pressure - self.c_vapour_const_volume

    def saturation_vapour_pressure(self, T):
        """:353-373 (Clausius-Clapeyron)"""
        import numpy as np
        k1 = (self.c_vapour_const_pressure - self.c_liquid_water) / self.R_vapour
        k2 = (self.ref_latent_heat_vap_temp -
              (self.c_vapour_const_pressure - self.c_liquid_water) * self.ref_temperature) / self.R_vapour
        return self.ref_saturation_pressure * (T / self.ref_temperature) ** k1 * \
            np.exp(k2 * (1 / self.ref_temperature - 1 / T))
