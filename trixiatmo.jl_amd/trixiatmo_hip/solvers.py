"""DGSEM solver description (Trixi `DGSEM(polydeg, surface_flux, volume_integral)`)."""
from dataclasses import dataclass
from typing import Optional, Tuple, Union

from .equations import (FluxLMARS, NONCONSERVATIVE_FLUXES, SURFACE_FLUXES, VOLUME_FLUXES)

PRECISIONS = {"fp64": 0, "fp32_kahan": 1, "fp32": 2, "fp32_geometry": 3}


@dataclass(frozen=True)
class VolumeIntegralFluxDifferencing:
    """`VolumeIntegralFluxDifferencing((symmetric_flux, nonconservative_flux))`"""
    volume_flux: Tuple[str, str]


@dataclass(frozen=True)
class VolumeIntegralWeakForm:
    """`VolumeIntegralWeakForm()`"""


@dataclass(frozen=True)
class DGSEM:
    """`DGSEM(polydeg = p, surface_flux = (flux, noncons), volume_integral = ...)`

    `precision` selects the arithmetic of the flux evaluations and volume sums
    ("fp64", "fp32_kahan" or "fp32"); RK registers stay FP64 in every form.
    "fp32_geometry" keeps every operation in FP64 and stores the metric terms
    in FP32: the 3D contravariant vectors, or the covariant form's metric
    terms, Christoffel symbols and face maps. The covariant form takes "fp64"
    or "fp32_geometry".
    """
    polydeg: int
    surface_flux: Union[Tuple[Union[str, FluxLMARS], str], str]
    volume_integral: Union[VolumeIntegralFluxDifferencing, VolumeIntegralWeakForm]
    precision: str = "fp64"
    # mixed paths: (p0, T0) of an isothermal hydrostatic background whose
    # pressure and density are split off in FP64 (None: no split)
    hydrostatic_background: Optional[Tuple[float, float]] = None

    def codes(self):
        vol, vnc = self.volume_integral.volume_flux
        sf, snc = self.surface_flux
        lmars_c = 0.0
        if isinstance(sf, FluxLMARS):
            lmars_c = float(sf.speed_of_sound)
            sf = "FluxLMARS"
        p0, T0 = self.hydrostatic_background or (0.0, 0.0)
        return dict(volume_flux=VOLUME_FLUXES[vol], volume_noncons=NONCONSERVATIVE_FLUXES[vnc],
                    surface_flux=SURFACE_FLUXES[sf], surface_noncons=NONCONSERVATIVE_FLUXES[snc],
                    precision=PRECISIONS[self.precision], lmars_c=lmars_c,
                    background_p0=float(p0), background_T0=float(T0))
