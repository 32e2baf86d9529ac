"""The named workloads (mesh + equations + solver + seeded state)."""
from .equations import CompressibleEulerEnergyEquationsWithGravity3D, EARTH_RADIUS
from .initial_conditions import (initial_condition_isothermal_perturbed,
                                 initial_condition_periodic_modes)
from .meshes import P4estMesh, P4estMeshCubedSphereTopography
from .solvers import DGSEM, VolumeIntegralFluxDifferencing, VolumeIntegralWeakForm

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
                   hydrostatic_background=None if precision in ("fp64", "fp32_geometry")
                   else (1.0, 1.0 / eq.R))
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
                   hydrostatic_background=None if precision in ("fp64", "fp32_geometry")
                   else (eq.p_0, 250.0))
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


# ---------------------------------------------------------------------------
# covariant shallow water on the cubed sphere

SWE_EC_LLF = "FluxPlusDissipation(flux_ec, DissipationLocalLaxFriedrichs())"


class CovariantCase:
    """mesh, equations, solver and the semidiscretization keywords of one
    covariant shallow-water setup; `semi()` creates the HIP semidiscretization,
    `aux()` the host auxiliary variables (no device needed)."""

    def __init__(self, mesh, equations, solver, initial_condition, source_terms,
                 auxiliary_field=None, christoffel="ChristoffelSymbolsAutodiff()", cfl=0.4):
        from .semidiscretization import MetricTermsCovariantSphere
        self.mesh, self.equations, self.solver = mesh, equations, solver
        self.initial_condition = initial_condition
        self.source_terms = source_terms
        self.auxiliary_field = auxiliary_field
        self.metric_terms = MetricTermsCovariantSphere(christoffel)
        self.cfl = cfl

    def aux(self):
        from .semidiscretization import covariant_aux
        return covariant_aux(self.mesh, self.metric_terms, self.auxiliary_field)

    def u0(self, t=0.0):
        return self.initial_condition(self.mesh.node_coordinates, t, self.aux())

    def semi(self):
        from .semidiscretization import SemidiscretizationHyperbolic
        return SemidiscretizationHyperbolic(self.mesh, self.equations, self.solver,
                                            source_terms=self.source_terms,
                                            initial_condition=self.initial_condition,
                                            metric_terms=self.metric_terms,
                                            auxiliary_field=self.auxiliary_field)


def covariant_elixir(name, cells=5, polydeg=3, **override):
    """The covariant elixirs of examples/shallow_water/covariant with the
    Cartesian global coordinate system, at `cells` trees per face dimension:
    "geostrophic_balance" (the standard form, weak form), and the split-form
    "unsteady_solid_body_rotation_EC", "barotropic_instability",
    "isolated_mountain", "rossby_haurwitz", "well_balanced".  `override`
    replaces volume_flux / surface_flux / christoffel as the reference's tests
    do through @test_trixi_include."""
    from . import initial_conditions as ic
    from .equations import (EARTH_GRAVITATIONAL_ACCELERATION, EARTH_ROTATION_RATE,
                            CovariantShallowWaterEquations2D,
                            SplitCovariantShallowWaterEquations2D)
    from .meshes import P4estMeshCubedSphere2D
    mesh = P4estMeshCubedSphere2D(cells, EARTH_RADIUS, polydeg=polydeg)
    spec = {
        # elixir_geostrophic_balance.jl:9-37: the standard form with the weak
        # form and flux_lax_friedrichs (the elixir's GlobalSphericalCoordinates()
        # gives the same metric and basis maps up to rounding)
        "geostrophic_balance": dict(
            ic=ic.initial_condition_geostrophic_balance, field=None, standard=True,
            surface_flux="flux_lax_friedrichs", cfl=0.4),
        # examples/shallow_water/covariant/elixir_unsteady_solid_body_rotation_EC.jl:20-46
        "unsteady_solid_body_rotation_EC": dict(
            ic=ic.initial_condition_unsteady_solid_body_rotation,
            field=ic.bottom_topography_unsteady_solid_body_rotation,
            volume_flux=("flux_ec", "flux_nonconservative_ec"),
            surface_flux=("flux_ec", "flux_nonconservative_surface_simplified"), cfl=0.4),
        # elixir_barotropic_instability.jl:20-67
        "barotropic_instability": dict(
            ic=ic.initial_condition_barotropic_instability, field=None,
            volume_flux=("flux_ec", "flux_nonconservative_ec"),
            surface_flux=(SWE_EC_LLF, "flux_nonconservative_surface_simplified"), cfl=0.4),
        # elixir_isolated_mountain.jl:21-71
        "isolated_mountain": dict(
            ic=ic.initial_condition_isolated_mountain,
            field=ic.bottom_topography_isolated_mountain,
            volume_flux=("flux_ec", "flux_nonconservative_ec"),
            surface_flux=(SWE_EC_LLF, "flux_nonconservative_surface_simplified"), cfl=0.4),
        # elixir_rossby_haurwitz.jl:20-70
        "rossby_haurwitz": dict(
            ic=ic.initial_condition_rossby_haurwitz, field=None,
            volume_flux=("flux_ec", "flux_nonconservative_ec"),
            surface_flux=(SWE_EC_LLF, "flux_nonconservative_surface_simplified"), cfl=0.4),
        # elixir_well_balanced.jl:13-73
        "well_balanced": dict(
            ic=ic.initial_condition_well_balanced,
            field=ic.bottom_topography_isolated_mountain,
            volume_flux=("flux_ec", "flux_nonconservative_ec"),
            surface_flux=(SWE_EC_LLF, "flux_nonconservative_ec"), cfl=0.1),
    }[name]
    spec.update(override)
    if spec.get("standard"):
        eq = CovariantShallowWaterEquations2D(EARTH_GRAVITATIONAL_ACCELERATION,
                                              EARTH_ROTATION_RATE)
        solver = DGSEM(polydeg=polydeg, surface_flux=spec["surface_flux"],
                       volume_integral=VolumeIntegralWeakForm(),
                       precision=spec.get("precision", "fp64"))
    else:
        eq = SplitCovariantShallowWaterEquations2D(EARTH_GRAVITATIONAL_ACCELERATION,
                                                   EARTH_ROTATION_RATE)
        solver = DGSEM(polydeg=polydeg, surface_flux=tuple(spec["surface_flux"]),
                       volume_integral=VolumeIntegralFluxDifferencing(tuple(spec["volume_flux"])),
                       precision=spec.get("precision", "fp64"))
    return CovariantCase(mesh, eq, solver, ic.transform_initial_condition(spec["ic"]),
                         "source_terms_geometric_coriolis", spec["field"],
                         spec.get("christoffel", "ChristoffelSymbolsAutodiff()"), spec["cfl"])


SWE_RADIUS = 6.371e6
SWE_GRAVITY = 9.81


def zonal_jet(trees_per_face=32, polydeg=3, seed=0, precision="fp64"):
    """Covariant shallow water on the cubed sphere of radius 6.371e6 m with
    g = 9.81: a seeded balanced zonal jet (u_0 in [20, 40] m/s, mean depth
    5960 m) plus three Gaussian height bumps; flux_ec +
    flux_nonconservative_ec volume, FluxPlusDissipation(flux_ec,
    DissipationLocalLaxFriedrichs()) + flux_nonconservative_surface_simplified
    surface, source_terms_geometric_coriolis; FP64 arithmetic with FP64 or
    (precision "fp32_geometry") FP32 metric storage."""
    from . import initial_conditions as ic
    from .equations import EARTH_ROTATION_RATE, SplitCovariantShallowWaterEquations2D
    from .meshes import P4estMeshCubedSphere2D
    eq = SplitCovariantShallowWaterEquations2D(SWE_GRAVITY, EARTH_ROTATION_RATE)
    mesh = P4estMeshCubedSphere2D(trees_per_face, SWE_RADIUS, polydeg=polydeg)
    solver = DGSEM(polydeg=polydeg,
                   surface_flux=(SWE_EC_LLF, "flux_nonconservative_surface_simplified"),
                   volume_integral=VolumeIntegralFluxDifferencing(
                       ("flux_ec", "flux_nonconservative_ec")), precision=precision)
    jet = ic.initial_condition_zonal_jet_bumps(seed, radius=SWE_RADIUS, gravity=SWE_GRAVITY)
    return CovariantCase(mesh, eq, solver, ic.transform_initial_condition(jet),
                         "source_terms_geometric_coriolis")


def ensemble_sizes(n_members=32, seed=0, choices=(8, 16, 24, 32)):
    """trees per face dimension of each member, drawn uniformly from `choices`"""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [int(k) for k in rng.choice(choices, size=n_members)]


def zonal_jet_ensemble(n_members=32, polydeg=3, seed=0, choices=(8, 16, 24, 32),
                       precision="fp64"):
    """Ensemble of independent zonal-jet members with uneven element counts
    6 k^2, k drawn from `choices`; member m draws its jet and bumps from seed
    1000 * seed + m + 1.  Returns (members [(mesh, ic, None)], equations,
    solver, source_terms) for SemidiscretizationEnsemble."""
    from . import initial_conditions as ic
    sizes = ensemble_sizes(n_members, seed, choices)
    meshes = {}
    members = []
    for m, k in enumerate(sizes):
        if k not in meshes:
            meshes[k] = zonal_jet(k, polydeg).mesh
        jet = ic.initial_condition_zonal_jet_bumps(1000 * seed + m + 1, radius=SWE_RADIUS,
                                                   gravity=SWE_GRAVITY)
        members.append((meshes[k], ic.transform_initial_condition(jet), None))
    base = zonal_jet(1, polydeg, precision=precision)
    return members, base.equations, base.solver, base.source_terms


# ---------------------------------------------------------------------------
# rainy Euler in 2D

def rainy_bubble(cells=(256, 128), polydeg=3, seed=0, domain=(20000.0, 10000.0),
                 newton_tolerance=1e-9, newton_max_iterations=20, periodic_y=False,
                 relative_humidity=None):
    """Moist warm bubble on a 20 km x 10 km StructuredMesh, 256 x 128 cells,
    polydeg 3: CompressibleRainyEulerEquations2D, flux_ec_rain volume,
    flux_lax_friedrichs surface, source_terms_rainy, NonlinearSolveDG stage
    limiter; periodic in x and boundary_condition_simple_slip_wall in y (a
    hydrostatic column is not periodic in height); a hydrostatic background of
    dry potential temperature 300 K and seeded relative humidity in
    [0.2, 0.95] (or the given one) with a 2 K bubble of 2 km radius at a
    seeded position.  Returns (mesh, equations, solver, source_terms, stage_limiter, u0, info)."""
    from .equations import CompressibleRainyEulerEquations2D
    from .initial_conditions import initial_condition_rainy_bubble
    from .meshes import StructuredMesh2D
    from .rainy import NonlinearSolveDG
    eq = CompressibleRainyEulerEquations2D()
    mesh = StructuredMesh2D(cells, (0.0, 0.0), domain, polydeg=polydeg,
                            periodicity=(True, periodic_y))
    solver = DGSEM(polydeg=polydeg, surface_flux="flux_lax_friedrichs",
                   volume_integral=VolumeIntegralFluxDifferencing("flux_ec_rain"))
    u0, info = initial_condition_rainy_bubble(mesh, eq, seed=seed,
                                              relative_humidity=relative_humidity)
    lim = NonlinearSolveDG(newton_tolerance, newton_max_iterations)
    return mesh, eq, solver, "source_terms_rainy", lim, u0, info


def precision_stress(precision="fp64", seed=0):
    """The shell at 6 x 8^2 trees x 4 layers, polydeg 7 (786,432 nodes), with
    the isothermal 250 K background and 10 seeded 1 K bumps, entropy
    conservative: (flux_ranocha, flux_nonconservative_waruszewski_etal) in
    the volume and on the surface, no dissipation, slip walls."""
    return shell(trees_per_face=8, layers=4, polydeg=7, seed=seed, precision=precision,
                 volume_flux="flux_ranocha", surface_flux="flux_ranocha",
                 surface_noncons="flux_nonconservative_waruszewski_etal")
