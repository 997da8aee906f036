"""MI355X-native FP64 FCT tracer advection (drop-in for fesom_jax.tracer_adv)."""

from .config import DT_DEFAULT, R_EARTH, TracerConfig
from .mesh import Mesh, load_mesh, save_mesh, level_masks, check_cw_orientation

__all__ = ["DT_DEFAULT", "R_EARTH", "TracerConfig", "Mesh", "load_mesh", "save_mesh",
           "level_masks", "check_cw_orientation"]
