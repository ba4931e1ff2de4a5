"""polyred_amd: polyred's rasteriser and deferred shader on AMD MI355X (gfx950).

The product is libprh (polyred_amd/_lib/libprh.so, C ABI in include/prh.h);
this package holds its Python binding (abi), the host-side mirror of the
render option API (render) and the seeded workload scenes (scenes).
"""
from . import abi  # noqa: F401
from . import scenes  # noqa: F401

__all__ = ["abi", "scenes", "render"]
