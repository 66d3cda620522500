"""Input/output helpers of the drop-in API (GuangyanCai/isoext src/isoext/utils.py).

``write_obj`` is the reference's OBJ format: one ``v x y z`` line per vertex
with each float32 widened to a Python float and printed with repr, then
``f a b c`` lines with 1-based indices; a None or empty mesh writes an empty
file (src/isoext/utils.py:42-63).  ``make_grid`` and ``gaussian_smooth`` are
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


def write_obj(obj_path: str, v: torch.Tensor, f: torch.Tensor) -> None:
    """Write a mesh as Wavefront OBJ text (vertices, then 1-based faces)."""
    with open(obj_path, "w") as out:
        if v is None or f is None or v.numel() == 0 or f.numel() == 0:
            return
        verts = v.tolist()
        faces = (f + 1).tolist()
        out.writelines([f"v {a} {b} {c}\n" for a, b, c in verts])
        out.writelines([f"f {a} {b} {c}\n" for a, b, c in faces])


def make_grid(aabb: list[float], res: int | list[int], device: str = "cuda") -> torch.Tensor:
    """(x_res, y_res, z_res, 3) lattice over aabb = [xmin, ymin, zmin, xmax, ymax, zmax]
    (src/isoext/utils.py:66-83)."""
    if isinstance(res, int):
        res = [res] * 3
    axes = [torch.linspace(aabb[i], aabb[i + 3], res[i], device=device) for i in range(3)]
    return torch.stack(torch.meshgrid(axes, indexing="ij"), dim=-1)
