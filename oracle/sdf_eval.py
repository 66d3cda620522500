"""CPU restatement of the reference's SDF primitives and CSG operators --
TEST INFRASTRUCTURE ONLY (tests/ may import it; the product never does).

``eval_tree`` evaluates an isoext.sdf tree in numpy float32 with the
formulas of src/isoext/sdf.py (SphereSDF :70-77, TorusSDF :80-94, CuboidSDF
:97-115, UnionOp :118-126, SmoothUnionOp :129-149, IntersectionOp :152-160,
NegationOp :163-170, TranslationOp :173-181, RotationOp :184-221), every
norm summed as (x*x + y*y) + z*z.  torch's own norm reduction rounds its
sum differently in about one point in ten (by an ulp), so a torch
evaluation agrees with this one to a few ulps, not bit for bit.
``eval_tree_grad`` gives the same tree's value and gradient in float64 with
torch autograd's conventions at kinks.

``eval_program`` interprets the flat program format of the HIP evaluator
(isoext_sdf_upload: op | dst << 8 | src << 12 | acc << 16 plus a parameter
offset) in numpy float32 with the same op order.

Parity pins (tests/test_csg.py): closed forms of every primitive and
operator (distances to spheres, tori, boxes and cylinders at points whose
distance is known exactly; a 90-degree rotation mapping axes onto axes; min
/ max / negation of known values; the log-sum-exp blend of two equal
values, d - k log 2); ``eval_tree`` against a torch evaluation of the same
tree within 4 ulps; ``eval_tree_grad`` against float64 central differences;
``eval_program`` against ``eval_tree`` bit for bit on random trees.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


def _norm(*c):
    s = c[0] * c[0]
    for x in c[1:]:
        s = s + x * x
    return np.sqrt(s)


def _nanmin(a, b):
    return np.where(np.isnan(b) | (b < a), b, a)


def _nanmax(a, b):
    return np.where(np.isnan(b) | (b > a), b, a)


def _cls(node) -> str:
    return type(node).__name__


def _rotation(node) -> np.ndarray:
    return np.asarray(node.R, dtype=np.float32).reshape(3, 3)


def eval_tree(node, p: np.ndarray) -> np.ndarray:
    """Values of an isoext.sdf tree at points p (..., 3), float32."""
    p = np.asarray(p, dtype=np.float32)
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    k = _cls(node)
    if k == "SphereSDF":
        return _norm(x, y, z) - f32(node.radius)
    if k == "TorusSDF":
        t = _norm(x, y) - f32(node.R)
        return _norm(t, z) - f32(node.r)
    if k == "CuboidSDF":
        h = np.asarray(node.size, dtype=np.float32) / f32(2)
        u = [np.abs(c) - h[i] for i, c in enumerate((x, y, z))]
        m = _nanmax(_nanmax(u[0], u[1]), u[2])
        o = [np.where(np.isnan(c), c, np.maximum(c, f32(0))) for c in u]
        inn = np.where(np.isnan(m), m, np.minimum(m, f32(0)))
        return _norm(*o) + inn
    if k == "CylinderSDF":
        return _norm(x, y) - f32(node.radius)
    if k in ("UnionOp", "IntersectionOp"):
        vals = [eval_tree(c, p) for c in node.sdf_list]
        r = vals[0]
        for v in vals[1:]:
            r = _nanmin(r, v) if k == "UnionOp" else _nanmax(r, v)
        return r
    if k == "SmoothUnionOp":
        kk = f32(node.k)
        vals = [eval_tree(c, p) for c in node.sdf_list]
        d = vals[0]
        with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
            for d2 in vals[1:]:
                d = (-kk) * np.log(np.exp(-d / kk) + np.exp(-d2 / kk))
        return d.astype(np.float32)
    if k == "NegationOp":
        return -eval_tree(node.sdf, p)
    if k == "TranslationOp":
        return eval_tree(node.sdf, p - np.asarray(node.offset, dtype=np.float32))
    if k == "RotationOp":
        R = _rotation(node)
        q = np.stack([(x * R[0, j] + y * R[1, j]) + z * R[2, j] for j in range(3)], -1)
        return eval_tree(node.sdf, q)
    raise TypeError(f"unknown SDF node {k}")


def eval_tree_grad(node, p: np.ndarray):
    """(value, gradient) of a tree at points p (..., 3) in float64, with
    autograd's conventions: |x|' = sign(x), zero-length norms and the
    untaken operand of min / max get gradient 0, ties of max(q, 0) and
    min(m, 0) split evenly."""
    p = np.asarray(p, dtype=np.float64)
    k = _cls(node)
    shp = p.shape[:-1]

    def unit(v):
        n = np.sqrt((v * v).sum(-1, keepdims=True))
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(n > 0, v / np.where(n > 0, n, 1), 0.0), n[..., 0]

    if k == "SphereSDF":
        u, n = unit(p)
        return n - node.radius, u
    if k == "CylinderSDF":
        u, n = unit(p[..., :2])
        return n - node.radius, np.concatenate([u, np.zeros(shp + (1,))], -1)
    if k == "TorusSDF":
        uxy, l = unit(p[..., :2])
        t = l - node.R
        w, m = unit(np.stack([t, p[..., 2]], -1))
        g = np.concatenate([w[..., :1] * uxy, w[..., 1:]], -1)
        return m - node.r, g
    if k == "CuboidSDF":
        h = np.asarray(node.size, dtype=np.float64) / 2
        u = np.abs(p) - h
        am = np.argmax(u, -1)
        m = np.take_along_axis(u, am[..., None], -1)[..., 0]
        o = np.maximum(u, 0)
        uo, l = unit(o)
        dmax = np.where(u > 0, 1.0, np.where(u == 0, 0.5, 0.0))
        du = uo * dmax
        din = np.where(m < 0, 1.0, np.where(m == 0, 0.5, 0.0))
        du = du + np.eye(3)[am] * din[..., None]
        return l + np.minimum(m, 0), du * np.sign(p)
    if k in ("UnionOp", "IntersectionOp"):
        vg = [eval_tree_grad(c, p) for c in node.sdf_list]
        v, g = vg[0]
        for v2, g2 in vg[1:]:
            take = (v2 < v) if k == "UnionOp" else (v2 > v)
            v = np.where(take, v2, v)
            g = np.where(take[..., None], g2, g)
        return v, g
    if k == "SmoothUnionOp":
        vg = [eval_tree_grad(c, p) for c in node.sdf_list]
        v, g = vg[0]
        for v2, g2 in vg[1:]:
            e1, e2 = np.exp(-v / node.k), np.exp(-v2 / node.k)
            s = e1 + e2
            v = -node.k * np.log(s)
            g = (e1 / s)[..., None] * g + (e2 / s)[..., None] * g2
        return v, g
    if k == "NegationOp":
        v, g = eval_tree_grad(node.sdf, p)
        return -v, -g
    if k == "TranslationOp":
        return eval_tree_grad(node.sdf, p - np.asarray(node.offset, dtype=np.float64))
    if k == "RotationOp":
        R = _rotation(node).astype(np.float64)
        v, g = eval_tree_grad(node.sdf, p @ R)
        return v, g @ R.T
    raise TypeError(f"unknown SDF node {k}")


NPARAM = {1: 1, 2: 4, 3: 2, 4: 3, 5: 1, 6: 0, 7: 1, 8: 0, 9: 3, 10: 9}


def eval_program(ops: np.ndarray, params: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Interpret a flat SDF program at points p (..., 3) in float32."""
    p = np.asarray(p, dtype=np.float32)
    shp = p.shape[:-1]
    prm = np.asarray(params, dtype=np.float32)
    v = [np.zeros(shp, np.float32) for _ in range(8)]
    q = [np.zeros(shp + (3,), np.float32) for _ in range(4)]
    q[0] = p.copy()
    for w, o in np.asarray(ops, dtype=np.int64).reshape(-1, 2):
        op, dst, src, acc = int(w) & 0xFF, (int(w) >> 8) & 7, (int(w) >> 12) & 7, (int(w) >> 16) & 3
        a = prm[int(o):int(o) + NPARAM[op]]
        x, y, z = (q[src & 3][..., i] for i in range(3))
        if op == 9:
            q[dst & 3] = np.stack([x - a[0], y - a[1], z - a[2]], -1)
            continue
        if op == 10:
            q[dst & 3] = np.stack([(x * a[j] + y * a[3 + j]) + z * a[6 + j] for j in range(3)], -1)
            continue
        cur = v[dst]
        if op == 1:
            r = _norm(x, y, z) - a[0]
        elif op == 2:
            r = _norm(x - a[0], y - a[1], z - a[2]) - a[3]
        elif op == 3:
            r = _norm(_norm(x, y) - a[0], z) - a[1]
        elif op == 4:
            u = [np.abs(c) - a[i] for i, c in enumerate((x, y, z))]
            m = _nanmax(_nanmax(u[0], u[1]), u[2])
            oo = [np.where(np.isnan(c), c, np.maximum(c, f32(0))) for c in u]
            r = _norm(*oo) + np.where(np.isnan(m), m, np.minimum(m, f32(0)))
        elif op == 5:
            r = _norm(x, y) - a[0]
        elif op == 6:
            r = v[src]
        elif op == 7:
            with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
                r = ((-a[0]) * np.log(np.exp(-cur / a[0]) + np.exp(-v[src] / a[0]))).astype(np.float32)
        elif op == 8:
            r = -cur
        else:
            raise ValueError(f"unknown op {op}")
        if acc == 1:
            r = _nanmin(cur, r)
        elif acc == 2:
            r = _nanmax(cur, r)
        v[dst] = np.asarray(r, dtype=np.float32)
    return v[0]
