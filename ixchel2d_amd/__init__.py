"""MI355X-native (gfx950) SIMPLE iteration of the Ixchel2D staggered-mesh
Navier-Stokes + energy solver, behind a C-ABI (include/ixchel2d.h)."""
from . import case  # noqa: F401

__all__ = ["case", "simple"]


def __getattr__(name):
    # the HIP binding is loaded on first use so that case tooling works on
    # machines without the built library
    if name == "simple":
        from . import simple
        return simple
    raise AttributeError(name)
