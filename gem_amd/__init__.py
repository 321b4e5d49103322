"""gem_amd -- MI355X (gfx950) implementation of GEM's semi-Lagrangian advection step
(gemdyn adz_* layer): departure-point trajectories, index/weight step, tricubic Lagrange
interpolation and the Bermejo-Conde mass fixer, behind a C ABI (include/adz_hip.h,
libadz_hip.so) with a Fortran ISO_C_BINDING layer and this Python host mirror."""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
