"""jcm_amd: the SPEEDY column-physics path of JCM (jax-gcm), on MI355X.

Public API mirrors the reference modules it replaces:
  SpeedyPhysics            jcm/physics/speedy/speedy_physics.py
  DiagnosticsCollector     jcm/model.py (the diagnostics' running mean, on the device)
  Parameters               jcm/physics/speedy/params.py
  PhysicsState, PhysicsTendency, Physics, verify_state, verify_tendencies
                           jcm/physics_interface.py
  Geometry                 jcm/geometry.py
  ForcingData              jcm/forcing.py
  DateData                 jcm/date.py
"""
from .date import DateData
from .forcing import ForcingData
from .geometry import Geometry
from .params import Parameters
from .physics_interface import Physics, PhysicsState, PhysicsTendency, verify_state, verify_tendencies

__all__ = ["DateData", "ForcingData", "Geometry", "Parameters", "Physics", "PhysicsState", "PhysicsTendency",
           "verify_state", "verify_tendencies", "SpeedyPhysics", "DiagnosticsCollector"]


def __getattr__(name):
    if name in ("SpeedyPhysics", "DiagnosticsCollector"):
        from . import speedy_physics
        return getattr(speedy_physics, name)
    raise AttributeError(name)
