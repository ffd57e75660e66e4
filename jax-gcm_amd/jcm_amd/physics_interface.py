"""Physics plugin API: mirror of jcm/physics_interface.py:22-194 and :305-342.

State and tendency fields are level-first arrays of shape (kx, ix, il) (or
(nmember, kx, ix, il) for a stacked ensemble); surface pressure is (ix, il).
Fields may be torch tensors on the GPU (kept device-resident and used
without copies when float32 and contiguous) or NumPy arrays (uploaded).
"""
from __future__ import annotations

from dataclasses import dataclass, fields, replace
from typing import Any, Tuple


@dataclass
class PhysicsState:                     # physics_interface.py:22-67
    u_wind: Any
    v_wind: Any
    temperature: Any
    specific_humidity: Any
    geopotential: Any
    normalized_surface_pressure: Any

    def copy(self, **kw) -> "PhysicsState":
        return replace(self, **kw)


@dataclass
class PhysicsTendency:                  # physics_interface.py:89-120
    u_wind: Any
    v_wind: Any
    temperature: Any
    specific_humidity: Any

    def copy(self, **kw) -> "PhysicsTendency":
        return replace(self, **kw)

    def __add__(self, other: "PhysicsTendency") -> "PhysicsTendency":
        return PhysicsTendency(*(getattr(self, f.name) + getattr(other, f.name) for f in fields(self)))


class Physics:                          # physics_interface.py:137-194
    def compute_tendencies(self, state: PhysicsState, forcing, geometry, date) -> Tuple[PhysicsTendency, Any]:
        raise NotImplementedError("Physics compute_tendencies method not implemented.")

    def get_empty_data(self, geometry) -> Any:
        return None

    def data_struct_to_dict(self, struct: Any, geometry=None, sep: str = ".") -> dict:
        """Flatten a physics data struct to {"group.field[.channel]": array}
        (physics_interface.py:157-194).  Accepts this library's PhysicsData."""
        if struct is None:
            return {}
        if hasattr(struct, "to_dict"):
            d = struct.to_dict()
            return d if sep == "." else {k.replace(".", sep): v for k, v in d.items()}
        raise ValueError(f"Unsupported type {type(struct)}")


def verify_state(state: PhysicsState) -> PhysicsState:
    """physics_interface.py:305-319: negative specific humidity becomes 0."""
    q = state.specific_humidity
    return state.copy(specific_humidity=q.clamp_min(0.0) if hasattr(q, "clamp_min") else q * (q >= 0))


def verify_tendencies(state: PhysicsState, tendencies: PhysicsTendency, time_step) -> PhysicsTendency:
    """physics_interface.py:321-342: limit the q tendency so q stays >= 0."""
    q, dq = state.specific_humidity, tendencies.specific_humidity
    ok = q + time_step * dq >= 0
    if hasattr(dq, "where"):
        new = dq.where(ok, -q / time_step)
    else:
        import numpy as np
        new = np.where(ok, dq, -q / time_step)
    return tendencies.copy(specific_humidity=new)


def dynamics_state_to_physics_state(state, dynamics) -> PhysicsState:
    """physics_interface.py:196-234, on the GPU (jcm_amd.transforms)."""
    from .transforms import dynamics_state_to_physics_state as f
    return f(state, dynamics)


def physics_state_to_dynamics_state(physics_state: PhysicsState, dynamics):
    """physics_interface.py:237-273, on the GPU (jcm_amd.transforms)."""
    from .transforms import physics_state_to_dynamics_state as f
    return f(physics_state, dynamics)


def physics_tendency_to_dynamics_tendency(physics_tendency: PhysicsTendency, dynamics):
    """physics_interface.py:275-303, on the GPU (jcm_amd.transforms)."""
    from .transforms import physics_tendency_to_dynamics_tendency as f
    return f(physics_tendency, dynamics)
