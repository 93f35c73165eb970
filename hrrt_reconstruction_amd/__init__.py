"""MI355X (gfx950) implementation of the hrrt_osem3d OSEM hot path."""
from ._lib import HrrtError, LIB_PATH  # noqa: F401

__all__ = ["HrrtError", "LIB_PATH", "Layout", "Osem3D", "make_params"]


def __getattr__(name):
    if name in ("Layout", "Osem3D", "make_params"):
        from . import osem
        return getattr(osem, name)
    raise AttributeError(name)
