"""The named workloads (mesh + equations + solver + seeded state)."""
from .equations import CompressibleEulerEnergyEquationsWithGravity3D, EARTH_RADIUS
from .initial_conditions import (initial_condition_isothermal_perturbed,
                                 initial_condition_periodic_modes)
from .meshes import P4estMesh, P4estMeshCubedSphereTopography
from .solvers import DGSEM, VolumeIntegralFluxDifferencing

SHELL_RADIUS = 6.371e6


def periodic_cube(trees=4, polydeg=3, seed=0, precision="fp64"):
    """Dry Euler (energy form) on the periodic unit cube, trees^3 elements,
    flux_ranocha volume, flux_lax_friedrichs surface, no gravity (phi = 0)."""
    eq = CompressibleEulerEnergyEquationsWithGravity3D(c_p=1004.0, c_v=717.0, gravity=9.80616)
    mesh = P4estMesh((trees,) * 3, polydeg=polydeg, coordinates_min=(0.0, 0.0, 0.0),
                     coordinates_max=(1.0, 1.0, 1.0), periodicity=(True, True, True))
    # mixed paths split off the mean pressure p = 1 (a constant background:
    # phi = 0); rho_bg = p0 / (R T0) = 1 is unused without gravity
    solver = DGSEM(polydeg=polydeg, surface_flux=("flux_lax_friedrichs", "flux_zero"),
                   volume_integral=VolumeIntegralFluxDifferencing(("flux_ranocha", "flux_zero")),
                   precision=precision,
                   hydrostatic_background=None if precision == "fp64" else (1.0, 1.0 / eq.R))
    u0 = initial_condition_periodic_modes(mesh, eq, seed)
    return mesh, eq, solver, {}, u0


SHELL_SURFACE_FLUX = "FluxPlusDissipation(flux_kennedy_gruber, DissipationLocalLaxFriedrichs())"


def shell(trees_per_face=16, layers=8, polydeg=4, seed=0, precision="fp64",
          volume_flux="flux_kennedy_gruber", surface_flux=SHELL_SURFACE_FLUX,
          surface_noncons="flux_nonconservative_waruszewski_etal"):
    """Dry Euler (energy form) with gravity on the cubed-sphere shell of radius
    6.371e6 m and 30 km height, isothermal 250 K background plus 10 seeded
    1 K bumps of 500 km radius, slip walls inside and outside.

    The local Lax-Friedrichs surface flux is built as
    FluxPlusDissipation(flux_kennedy_gruber, DissipationLocalLaxFriedrichs()):
    the reference's `flux_lax_friedrichs` averages the pointwise `flux`, whose
    energy entry reads phi instead of rho_e
    (src/equations/compressible_euler_energy_with_gravity_3d.jl:80), and with
    phi ~ 3e5 J/kg at the model top that run diverges within 10 steps."""
    eq = CompressibleEulerEnergyEquationsWithGravity3D(c_p=1004.0, c_v=717.0, gravity=9.81)
    mesh = P4estMeshCubedSphereTopography(trees_per_face, layers, SHELL_RADIUS, 30000.0,
                                          polydeg=polydeg)
    solver = DGSEM(polydeg=polydeg, surface_flux=(surface_flux, surface_noncons),
                   volume_integral=VolumeIntegralFluxDifferencing(
                       (volume_flux, "flux_nonconservative_waruszewski_etal")),
                   precision=precision,
                   hydrostatic_background=None if precision == "fp64" else (eq.p_0, 250.0))
    u0 = initial_condition_isothermal_perturbed(mesh, eq, seed, inner_radius=SHELL_RADIUS)
    bcs = {"inside": "boundary_condition_slip_wall", "outside": "boundary_condition_slip_wall"}
    return mesh, eq, solver, bcs, u0


def baroclinic_instability(trees_per_face=8, layers=4, polydeg=5, precision="fp64"):
    """examples/euler/dry_air/global_circulation/elixir_energy_baroclinic_instability.jl:14-37:
    flux_ranocha + waruszewski volume, FluxLMARS(340) + flux_zero surface,
    slip walls, source_terms_coriolis, shell of EARTH_RADIUS and 30 km."""
    from .equations import EARTH_GRAVITATIONAL_ACCELERATION, FluxLMARS
    from .initial_conditions import initial_condition_baroclinic_instability
    eq = CompressibleEulerEnergyEquationsWithGravity3D(
        c_p=1004.0, c_v=717.0, gravity=EARTH_GRAVITATIONAL_ACCELERATION)
    mesh = P4estMeshCubedSphereTopography(trees_per_face, layers, EARTH_RADIUS, 30000.0,
                                          polydeg=polydeg)
    solver = DGSEM(polydeg=polydeg, surface_flux=(FluxLMARS(340.0), "flux_zero"),
                   volume_integral=VolumeIntegralFluxDifferencing(
                       ("flux_ranocha", "flux_nonconservative_waruszewski_etal")),
                   precision=precision)
    u0 = initial_condition_baroclinic_instability(mesh.node_coordinates, eq)
    bcs = {"inside": "boundary_condition_slip_wall", "outside": "boundary_condition_slip_wall"}
    return mesh, eq, solver, bcs, u0
