"""trixiatmo_hip: host-side mirror of TrixiAtmo.jl's DGSEM path over libatmo_hip.

The right-hand side, the Runge-Kutta stages and the reductions run in the
hand-written gfx950 HIP kernels of libatmo_hip.so (C ABI: include/atmo_hip.h).
"""
from ._lib import AtmoError, lib
from .equations import (EARTH_GRAVITATIONAL_ACCELERATION, EARTH_RADIUS, EARTH_ROTATION_RATE,
                        CompressibleEulerEnergyEquationsWithGravity3D, FluxLMARS,
                        CovariantShallowWaterEquations2D, FluxPlusDissipation,
                        SplitCovariantShallowWaterEquations2D)
from .initial_conditions import (initial_condition_baroclinic_instability,
                                 initial_condition_isothermal_perturbed,
                                 initial_condition_periodic_modes)
from .meshes import P4estMesh, P4estMeshCubedSphere2D, P4estMeshCubedSphereTopography, lgl_basis
from .semidiscretization import (MetricTermsCovariantSphere, SemidiscretizationEnsemble,
                                 SemidiscretizationHyperbolic, solve)
from .solvers import DGSEM, VolumeIntegralFluxDifferencing
from .rainy import NonlinearSolveDG, SaturationMLP, SemidiscretizationRainy
from . import analysis, solution_files, workloads

__all__ = [
    "AtmoError", "lib", "EARTH_GRAVITATIONAL_ACCELERATION", "EARTH_RADIUS", "EARTH_ROTATION_RATE",
    "CompressibleEulerEnergyEquationsWithGravity3D", "FluxLMARS", "FluxPlusDissipation",
    "initial_condition_baroclinic_instability", "initial_condition_isothermal_perturbed", "initial_condition_periodic_modes",
    "P4estMesh", "P4estMeshCubedSphereTopography", "lgl_basis", "SemidiscretizationHyperbolic",
    "DGSEM", "VolumeIntegralFluxDifferencing", "analysis", "workloads",
    "SplitCovariantShallowWaterEquations2D", "CovariantShallowWaterEquations2D", "P4estMeshCubedSphere2D", "MetricTermsCovariantSphere",
    "SemidiscretizationEnsemble", "solve", "NonlinearSolveDG", "SaturationMLP",
    "SemidiscretizationRainy", "solution_files",
]
