"""isoext_amd -- MI355X-native iso-surface extraction with isoext's Python API.

Drop-in names (GuangyanCai/isoext src/isoext/__init__.py:6-26):
UniformGrid, SparseGrid, Intersection, marching_cubes, dual_contouring,
get_intersection, write_obj, make_grid, gaussian_smooth.

The compute path is libisoext_amd.so (hand-written HIP for gfx950, C ABI in
include/isoext_amd.h); torch provides device memory and streams.
"""

from __future__ import annotations

import importlib.util

if importlib.util.find_spec("torch") is None:
    raise ImportError("PyTorch is required but not installed. Please install PyTorch with ROCm support.\n")

from .grid import Grid, UniformGrid
from .mc import marching_cubes
from .utils import gaussian_smooth, make_grid, write_obj

__all__ = [
    "Grid",
    "UniformGrid",
    "marching_cubes",
    "gaussian_smooth",
    "make_grid",
    "write_obj",
]

from .dc import Intersection, dual_contouring, get_intersection  # noqa: E402
from .sparse import SparseGrid  # noqa: E402

__all__ += ["Intersection", "SparseGrid", "dual_contouring", "get_intersection"]
