"""Signed-distance primitives and CSG operators in torch (input generation).

Same callables and formulas as the reference's isoext.sdf
(src/isoext/sdf.py:23-221): sphere, torus (xy-plane), axis-aligned cuboid,
union (min), smooth union, intersection (max), negation, translation and
axis-angle rotation, plus autograd gradients/normals.  These produce grid
values; they are not on the extraction path.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Protocol

import torch


class SDFProtocol(Protocol):
    def __call__(self, p: torch.Tensor) -> torch.Tensor: ...


def get_sdf_grad(sdf: SDFProtocol, p: torch.Tensor) -> torch.Tensor:
    """Gradient of sdf at points p (..., 3) by autograd."""
    p = p.requires_grad_()
    d = sdf(p)
    return torch.autograd.grad(d, p, grad_outputs=torch.ones_like(d))[0]


def get_sdf_normal(sdf: SDFProtocol, p: torch.Tensor) -> torch.Tensor:
    """Unit-length gradient of sdf at points p."""
    return torch.nn.functional.normalize(get_sdf_grad(sdf, p), dim=-1)


class SDF(ABC):
    @abstractmethod
    def __call__(self, p: torch.Tensor) -> torch.Tensor: ...


@dataclass
class SphereSDF(SDF):
    radius: float

    def __call__(self, p):
        return p.norm(dim=-1) - self.radius


@dataclass
class TorusSDF(SDF):
    """Torus around the z axis: major radius R, tube radius r."""

    R: float
    r: float

    def __call__(self, p):
        q = p[..., [0, 1]].norm(dim=-1) - self.R
        return torch.stack([q, p[..., 2]], dim=-1).norm(dim=-1) - self.r


@dataclass
class CuboidSDF(SDF):
    """Axis-aligned box centred at the origin; `size` holds full edge lengths."""

    size: list[float]

    def __call__(self, p):
        half = torch.tensor(self.size, device=p.device, dtype=p.dtype) / 2
        q = torch.abs(p) - half
        qmax = q.max(dim=-1).values
        outside = torch.norm(torch.maximum(q, torch.zeros_like(q)), dim=-1)
        return outside + torch.minimum(qmax, torch.zeros_like(qmax))


@dataclass
class CylinderSDF(SDF):
    """Infinite cylinder around the z axis, |p_xy| - radius (not a class of
    the reference; the field of the dual-contouring benchmark carves one)."""

    radius: float

    def __call__(self, p):
        return p[..., :2].norm(dim=-1) - self.radius


@dataclass
class UnionOp(SDF):
    sdf_list: list

    def __call__(self, p):
        return torch.stack([s(p) for s in self.sdf_list], dim=-1).min(dim=-1).values


@dataclass
class SmoothUnionOp(SDF):
    """Log-sum-exp blend of the children with sharpness k."""

    sdf_list: list
    k: float

    def __call__(self, p):
        ds = [s(p) for s in self.sdf_list]
        d = ds[0]
        for d2 in ds[1:]:
            d = -self.k * torch.log(torch.exp(-d / self.k) + torch.exp(-d2 / self.k))
        return d


@dataclass
class IntersectionOp(SDF):
    sdf_list: list

    def __call__(self, p):
        return torch.stack([s(p) for s in self.sdf_list], dim=-1).max(dim=-1).values


@dataclass
class NegationOp(SDF):
    sdf: SDF

    def __call__(self, p):
        return -self.sdf(p)


@dataclass
class TranslationOp(SDF):
    sdf: SDF
    offset: list[float]

    def __call__(self, p):
        return self.sdf(p - torch.tensor(self.offset).to(p))


@dataclass
class RotationOp(SDF):
    """Rotate the child about `axis` by `angle` (degrees unless use_degree=False)."""

    sdf: SDF
    axis: list[float]
    angle: float
    use_degree: bool = True

    def __post_init__(self):
        a = torch.nn.functional.normalize(torch.tensor(self.axis).float(), dim=0).reshape(3, 1)
        ang = torch.tensor(self.angle).float()
        if self.use_degree:
            ang = torch.deg2rad(ang)
        s, c = torch.sin(ang), torch.cos(ang)
        k = torch.zeros((3, 3))
        k[0, 1], k[0, 2] = -a[2], a[1]
        k[1, 0], k[1, 2] = a[2], -a[0]
        k[2, 0], k[2, 1] = -a[1], a[0]
        self.R = c * torch.eye(3) + s * k + (1 - c) * (a @ a.T)

    def __call__(self, p):
        return self.sdf(p @ self.R.to(p))
