"""DateData: mirror of jcm/date.py:9-48."""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from . import _native
from .forcing import model_day as _model_day


@dataclass
class DateData:
    tyear: float = 0.0          # fraction of the year
    model_year: int = 1950
    model_step: int = 0
    dt_seconds: float = 1800.0

    @classmethod
    def zeros(cls, **kw):
        return cls(**kw)

    def model_day(self) -> int:
        return _model_day(self.tyear)

    def copy(self, **kw) -> "DateData":
        return replace(self, **kw)

    def to_abi(self) -> _native.SpeedyDate:
        d = _native.SpeedyDate()
        d.tyear = float(np.float32(self.tyear))
        d.model_year = int(self.model_year)
        d.model_step = int(self.model_step)
        d.dt_seconds = float(self.dt_seconds)
        return d
