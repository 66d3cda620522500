"""Input/output helpers of the drop-in API (GuangyanCai/isoext src/isoext/utils.py).

``write_obj`` is the reference's OBJ format: one ``v x y z`` line per vertex
with each float32 widened to a Python float and printed with repr, then
``f a b c`` lines with 1-based indices; a None or empty mesh writes an empty
file (src/isoext/utils.py:42-63).  The text is produced by the library's
multi-threaded C++ writer (``isoext_write_obj``).  ``make_grid`` and ``gaussian_smooth`` are
the reference's input-side conveniences restated in torch.
"""

from __future__ import annotations

import torch
import torch.nn.functional as F


def gaussian_smooth(field: torch.Tensor, sigma: float = 1.0, kernel_size: int | None = None) -> torch.Tensor:
    """Separable-equivalent 3D Gaussian blur with replicate padding
    (src/isoext/utils.py:5-39).  Kernel size defaults to int(6*sigma) | 1."""
    if kernel_size is None:
        kernel_size = int(6 * sigma) | 1
    half = kernel_size // 2
    x = torch.arange(kernel_size, device=field.device, dtype=field.dtype) - half
    g = torch.exp(-0.5 * (x / sigma) ** 2)
    g = g / g.sum()
    k3 = (g[:, None, None] * g[None, :, None] * g[None, None, :]).view(1, 1, kernel_size, kernel_size, kernel_size)
    f5 = F.pad(field.view(1, 1, *field.shape), [half] * 6, mode="replicate")
    return F.conv3d(f5, k3).reshape(field.shape)


def write_obj(obj_path: str, v: torch.Tensor, f: torch.Tensor, threads: int = 8) -> None:
    """Write a mesh as Wavefront OBJ text (vertices, then 1-based faces) with
    the reference's exact bytes (src/isoext/utils.py:42-63): coordinates as
    Python float reprs.  Formatting runs in the library's C++ writer on
    `threads` host threads; an empty mesh leaves an empty file."""
    import ctypes

    from . import _lib
    L = _lib.lib()
    # opening in Python first raises exactly what the reference's open() raises
    with open(obj_path, "w"):
        pass
    if v is None or f is None or v.numel() == 0 or f.numel() == 0:
        return
    path = str(obj_path).encode()
    vh = v.detach().to("cpu", torch.float64).contiguous()
    fh = f.detach().to("cpu", torch.int64).contiguous()
    _lib.check(L.isoext_write_obj(path, ctypes.c_void_p(vh.data_ptr()), vh.shape[0],
                                  ctypes.c_void_p(fh.data_ptr()), fh.shape[0], int(threads)), "write_obj")


def make_grid(aabb: list[float], res: int | list[int], device: str = "cuda") -> torch.Tensor:
    """(x_res, y_res, z_res, 3) lattice over aabb = [xmin, ymin, zmin, xmax, ymax, zmax]
    (src/isoext/utils.py:66-83)."""
    if isinstance(res, int):
        res = [res] * 3
    axes = [torch.linspace(aabb[i], aabb[i + 3], res[i], device=device) for i in range(3)]
    return torch.stack(torch.meshgrid(axes, indexing="ij"), dim=-1)
