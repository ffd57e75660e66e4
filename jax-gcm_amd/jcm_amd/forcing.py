"""ForcingData: mirror of jcm/forcing.py:8-45 (fields and lfluxland flag).

Fields are (ix, il) arrays, or (ix, il, ndays) climatologies that are sliced
at the model day (speedy_physics.py:119-121, utils.py:104-108).
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Any

import numpy as np

DAYS_YEAR = 365.2425          # date.py:7
FORCING_FIELDS = ("alb0", "sice_am", "snowc_am", "soilw_am", "stl_am", "sea_surface_temperature")


@dataclass
class ForcingData:
    alb0: Any
    sice_am: Any
    snowc_am: Any
    soilw_am: Any
    stl_am: Any
    sea_surface_temperature: Any
    lfluxland: bool = False

    @classmethod
    def zeros(cls, nodal_shape, lfluxland=False, **kw):
        d = {k: np.zeros(nodal_shape, np.float32) for k in FORCING_FIELDS}
        d.update({k: np.asarray(v, np.float32) for k, v in kw.items()})
        return cls(**d, lfluxland=lfluxland)

    @classmethod
    def ones(cls, nodal_shape, lfluxland=True, **kw):
        d = {k: np.ones(nodal_shape, np.float32) for k in FORCING_FIELDS}
        d.update({k: np.asarray(v, np.float32) for k, v in kw.items()})
        return cls(**d, lfluxland=lfluxland)

    def copy(self, **kw) -> "ForcingData":
        return replace(self, **kw)


def model_day(tyear: float) -> int:
    """DateData.model_day (date.py:40-41): round(tyear * 365.2425) in float32."""
    return int(np.round(np.float32(tyear) * np.float32(DAYS_YEAR)))


def day_slice(arr, day: int):
    """arr[:, :, day] for a climatology (clamped like a JAX gather), else arr."""
    if arr.ndim > 2:
        return arr[:, :, min(max(day, 0), arr.shape[2] - 1)]
    return arr
