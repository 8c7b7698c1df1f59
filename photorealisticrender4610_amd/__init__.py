"""MI355X (gfx950) path-tracing hot path of the PhotoRealisticRender4610
renderer: a HIP library (libprr.so, C ABI in include/prr_abi.h) plus the
host-side scene description used to drive it from Python."""
from . import abi, scenes  # noqa: F401

__all__ = ["abi", "scenes"]
