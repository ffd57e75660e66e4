"""SPEEDY parameters: mirror of jcm/physics/speedy/params.py (Parameters and its
seven groups, same names and defaults), packed for the C ABI in declaration
order.  Two multipliers are added: ``forcing.co2_scale`` (scales the CO2
absorptivity ``ablco2``) and ``surface_flux.drag_scale`` (scales ``cdl`` and
``cds``); at 1.0 every expression equals the reference's."""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

from . import _native


@dataclass
class ConvectionParameters:          # params.py:9-30
    psmin: float = 0.8
    trcnv: float = 6.0
    rhil: float = 0.7
    rhbl: float = 0.9
    entmax: float = 0.5
    smf: float = 0.8


@dataclass
class ForcingParameters:             # params.py:32-47
    increase_co2: bool = False
    co2_year_ref: int = 1950
    co2_scale: float = 1.0


@dataclass
class CondensationParameters:        # params.py:49-66
    trlsc: float = 4.0
    rhlsc: float = 0.9
    drhlsc: float = 0.1
    rhblsc: float = 0.95


@dataclass
class ShortwaveRadiationParameters:  # params.py:68-129
    albcl: float = 0.43
    albcls: float = 0.50
    absdry: float = 0.033
    absaer: float = 0.033
    abswv1: float = 0.022
    abswv2: float = 15.000
    abscl1: float = 0.015
    abscl2: float = 0.15
    ablwin: float = 0.3
    ablwv1: float = 0.7
    ablwv2: float = 50.0
    ablcl1: float = 12.0
    ablcl2: float = 0.6
    rhcl1: float = 0.30
    rhcl2: float = 1.00
    qacl: float = 0.20
    wpcl: float = 0.2
    pmaxcl: float = 10.0
    clsmax: float = 0.60
    clsminl: float = 0.15
    gse_s0: float = 0.25
    gse_s1: float = 0.40


@dataclass
class ModRadConParameters:           # params.py:131-153
    albsea: float = 0.07
    albice: float = 0.60
    albsn: float = 0.60
    epslw: float = 0.05
    emisfc: float = 0.98


@dataclass
class SurfaceFluxParameters:         # params.py:155-209
    fwind0: float = 0.95
    ftemp0: float = 1.0
    fhum0: float = 0.0
    cdl: float = 2.4e-3
    cds: float = 1.0e-3
    chl: float = 1.2e-3
    chs: float = 0.9e-3
    vgust: float = 5.0
    ctday: float = 1.0e-2
    dtheta: float = 3.0
    fstab: float = 0.67
    clambda: float = 7.0
    clambsn: float = 7.0
    lscasym: bool = True
    lskineb: bool = True
    hdrag: float = 2000.0
    drag_scale: float = 1.0


@dataclass
class VerticalDiffusionParameters:   # params.py:211-232
    trshc: float = 6.0
    trvdi: float = 24.0
    trvds: float = 6.0
    redshc: float = 0.5
    rhgrad: float = 0.5
    segrad: float = 0.1


# The parameter groups in the order of the C struct (speedy_abi.h SpeedyParams).
_GROUPS = ("convection", "condensation", "shortwave_radiation", "mod_radcon", "surface_flux",
           "vertical_diffusion", "forcing")

# The five scalars the fused VJP differentiates, in the order of g_params.
VJP_PARAMS = (("convection", "rhbl"), ("convection", "trcnv"), ("condensation", "rhlsc"),
              ("forcing", "co2_scale"), ("surface_flux", "drag_scale"))


@dataclass
class Parameters:                    # params.py:234-294
    convection: ConvectionParameters = field(default_factory=ConvectionParameters)
    condensation: CondensationParameters = field(default_factory=CondensationParameters)
    shortwave_radiation: ShortwaveRadiationParameters = field(default_factory=ShortwaveRadiationParameters)
    mod_radcon: ModRadConParameters = field(default_factory=ModRadConParameters)
    surface_flux: SurfaceFluxParameters = field(default_factory=SurfaceFluxParameters)
    vertical_diffusion: VerticalDiffusionParameters = field(default_factory=VerticalDiffusionParameters)
    forcing: ForcingParameters = field(default_factory=ForcingParameters)

    @classmethod
    def default(cls) -> "Parameters":
        return cls()

    def replace(self, **groups) -> "Parameters":
        """Parameters with some fields changed: replace(convection=dict(rhbl=0.8))."""
        new = dataclasses.replace(self)
        for g, changes in groups.items():
            setattr(new, g, dataclasses.replace(getattr(self, g), **changes))
        return new

    def abi_key(self) -> tuple:
        """Every value to_abi packs, group by group: equal keys give equal ABI structs
        (a cheap test for in-place changes between calls)."""
        return tuple(tuple(vars(getattr(self, grp)).values()) for grp in _GROUPS)

    def to_abi(self) -> _native.SpeedyParams:
        p = _native.SpeedyParams()
        for grp in _GROUPS:
            for f in dataclasses.fields(getattr(self, grp)):
                v = getattr(getattr(self, grp), f.name)
                setattr(p, f.name, int(v) if isinstance(v, bool) or f.name == "co2_year_ref" else float(v))
        return p
