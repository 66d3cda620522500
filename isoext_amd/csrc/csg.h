// Signed-distance programs: the reference's SDF primitives and CSG operators
// (src/isoext/sdf.py:70-221) flattened into a short register program that a
// kernel evaluates at any point, so a field can be produced where the
// extraction needs it instead of being stored as a grid.
//
// A program is a list of ops over 8 value slots and 4 point slots; point
// slot 0 holds the evaluation point, the result is value slot 0.  Each op
// is two int32: a code word (op | dst << 8 | src << 12 | acc << 16) and the
// offset of its parameters.  Primitives read point slot `src` and write
// value slot `dst`, combined with the slot's previous value by `acc` (store,
// min or max); transforms write point slot `dst` from point slot `src`.
// Every formula keeps the op order of the reference's torch code (norms as
// sqrt((x*x + y*y) + z*z), uncontracted), so a program evaluated here and
// the stored grid of the same program written by isoext_sdf_grid agree bit
// for bit.
//
//   op              params             value
//   SPHERE          r                  |q| - r                    SphereSDF
//   SPHERE_AT       cx, cy, cz, r      |q - c| - r                TranslationOp(SphereSDF)
//   TORUS           R, r               |(|q_xy| - R, q_z)| - r    TorusSDF
//   CUBOID          hx, hy, hz         |max(|q| - h, 0)| + min(max_i(|q_i| - h_i), 0)   CuboidSDF (h = size / 2)
//   CYLINDER_Z      r                  |q_xy| - r                 (infinite z-cylinder; no reference class)
//   VALUE           -                  v[src]                     UnionOp / IntersectionOp with acc
//   SMOOTH_MIN      k                  -k log(exp(-v[dst]/k) + exp(-v[src]/k))   SmoothUnionOp
//   NEG             -                  -v[dst]                    NegationOp
//   TRANSLATE       ox, oy, oz         q[dst] = q[src] - o        TranslationOp
//   ROTATE          R (3x3 row-major)  q[dst] = q[src] @ R        RotationOp
//
// min / max propagate NaN like torch.min / torch.max; on ties the earlier
// operand wins (the gradient then follows it).
#pragma once

#include "common.h"

namespace iso {

constexpr int kCsgMaxOps = 256;
constexpr int kCsgMaxParams = 1024;
constexpr int kCsgVals = 8;
constexpr int kCsgPts = 4;

enum : int {
    CSG_SPHERE = 1,
    CSG_SPHERE_AT = 2,
    CSG_TORUS = 3,
    CSG_CUBOID = 4,
    CSG_CYLINDER_Z = 5,
    CSG_VALUE = 6,
    CSG_SMOOTH_MIN = 7,
    CSG_NEG = 8,
    CSG_TRANSLATE = 9,
    CSG_ROTATE = 10,
};
enum : int { CSG_STORE = 0, CSG_ACC_MIN = 1, CSG_ACC_MAX = 2 };

// Parameters read by each op (0 for unknown codes).
ISO_HD int csg_param_count(int op) {
    switch (op) {
    case CSG_SPHERE: return 1;
    case CSG_SPHERE_AT: return 4;
    case CSG_TORUS: return 2;
    case CSG_CUBOID: return 3;
    case CSG_CYLINDER_Z: return 1;
    case CSG_SMOOTH_MIN: return 1;
    case CSG_TRANSLATE: return 3;
    case CSG_ROTATE: return 9;
    default: return 0;
    }
}

// Validated device form of a program (built on the host by csg_build).
struct CsgProgram {
    int2 ops[kCsgMaxOps];
    float prm[kCsgMaxParams];
    int n_ops;
    int lipschitz;   // 1: no smooth min, so the field is 1-Lipschitz and sign culling is exact
    float pscale;    // sum of |parameters| (part of the culling margin)
    int pad;
};

// ---- evaluation ----------------------------------------------------------------------

__device__ __forceinline__ float csg_norm2(float x, float y) { return sqrtf(x * x + y * y); }
__device__ __forceinline__ float csg_norm3(float x, float y, float z) {
    float s = x * x + y * y;
    s = s + z * z;
    return sqrtf(s);
}
// b replaces a in a min (max): b is NaN, or strictly smaller (larger) while a is not NaN
__device__ __forceinline__ bool csg_take_min(float a, float b) { return b != b || b < a; }
__device__ __forceinline__ bool csg_take_max(float a, float b) { return b != b || b > a; }

// Register slots addressed by a wave-uniform (or per-lane) index: short
// vectors read and written through select chains over constant components,
// so the slots stay in VGPRs (an element address taken with a variable
// index would move the slots to private or LDS memory).
typedef float CsgV4 __attribute__((ext_vector_type(4)));
typedef float CsgV8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ float csg_sel(CsgV4 v, int s) {
    float r = v.x;
    r = s == 1 ? v.y : r;
    r = s == 2 ? v.z : r;
    r = s == 3 ? v.w : r;
    return r;
}
__device__ __forceinline__ float csg_sel(CsgV8 v, int s) {
    float r = v.s0;
    r = s == 1 ? v.s1 : r;
    r = s == 2 ? v.s2 : r;
    r = s == 3 ? v.s3 : r;
    r = s == 4 ? v.s4 : r;
    r = s == 5 ? v.s5 : r;
    r = s == 6 ? v.s6 : r;
    r = s == 7 ? v.s7 : r;
    return r;
}
struct CsgSlots4 {
    CsgV4 a;
    __device__ __forceinline__ void fill(float x) { a = (CsgV4) x; }
    __device__ __forceinline__ float get(int s) const { return csg_sel(a, s); }
    __device__ __forceinline__ void set(int s, float x) {
        CsgV4 t = a;
        t.x = s == 0 ? x : t.x;
        t.y = s == 1 ? x : t.y;
        t.z = s == 2 ? x : t.z;
        t.w = s == 3 ? x : t.w;
        a = t;
    }
};
struct CsgSlots8 {
    CsgV8 a;
    __device__ __forceinline__ void fill(float x) { a = (CsgV8) x; }
    __device__ __forceinline__ float get(int s) const { return csg_sel(a, s); }
    __device__ __forceinline__ void set(int s, float x) {
        CsgV8 t = a;
        t.s0 = s == 0 ? x : t.s0;
        t.s1 = s == 1 ? x : t.s1;
        t.s2 = s == 2 ? x : t.s2;
        t.s3 = s == 3 ? x : t.s3;
        t.s4 = s == 4 ? x : t.s4;
        t.s5 = s == 5 ? x : t.s5;
        t.s6 = s == 6 ? x : t.s6;
        t.s7 = s == 7 ? x : t.s7;
        a = t;
    }
};

struct CsgOpWord {
    int op, dst, src, acc, prm;
};
template <bool UNI>
__device__ __forceinline__ CsgOpWord csg_decode(const CsgProgram &P, int i) {
    const int2 w = P.ops[i];
    int a = w.x, b = w.y;
    if (UNI) {
        a = __builtin_amdgcn_readfirstlane(a);
        b = __builtin_amdgcn_readfirstlane(b);
    }
    CsgOpWord o;
    o.op = a & 0xff;
    o.dst = (a >> 8) & 7;
    o.src = (a >> 12) & 7;
    o.acc = (a >> 16) & 3;
    o.prm = b;
    return o;
}

// Value of primitive `op` (SPHERE .. CYLINDER_Z) at (x, y, z).
__device__ __forceinline__ float csg_prim(int op, const float *a, float x, float y, float z) {
    switch (op) {
    case CSG_SPHERE: return csg_norm3(x, y, z) - a[0];
    case CSG_SPHERE_AT: return csg_norm3(x - a[0], y - a[1], z - a[2]) - a[3];
    case CSG_TORUS: {
        const float t = csg_norm2(x, y) - a[0];
        return csg_norm2(t, z) - a[1];
    }
    case CSG_CUBOID: {
        const float ux = fabsf(x) - a[0], uy = fabsf(y) - a[1], uz = fabsf(z) - a[2];
        float m = ux;
        if (csg_take_max(m, uy)) m = uy;
        if (csg_take_max(m, uz)) m = uz;
        const float ox = ux != ux ? ux : fmaxf(ux, 0.0f), oy = uy != uy ? uy : fmaxf(uy, 0.0f),
                    oz = uz != uz ? uz : fmaxf(uz, 0.0f);
        const float in = m != m ? m : fminf(m, 0.0f);
        return csg_norm3(ox, oy, oz) + in;
    }
    default: return csg_norm2(x, y) - a[0];   // CYLINDER_Z
    }
}

__device__ __forceinline__ float csg_combine(int acc, float cur, float r) {
    if (acc == CSG_ACC_MIN) return csg_take_min(cur, r) ? r : cur;
    if (acc == CSG_ACC_MAX) return csg_take_max(cur, r) ? r : cur;
    return r;
}

// Field value at (px, py, pz).  UNI: every lane of the wave runs the same
// program (ops decoded into scalar registers).  Primitives on point slot 0
// into value slot 0 (most ops of a union of shapes) skip the slot selects.
template <bool UNI>
__device__ float csg_value(const CsgProgram &P, float px, float py, float pz) {
    CsgSlots8 v;
    CsgSlots4 qx, qy, qz;
    v.fill(0.0f);
    qx.fill(0.0f);
    qy.fill(0.0f);
    qz.fill(0.0f);
    qx.a.x = px;
    qy.a.x = py;
    qz.a.x = pz;
    int n = P.n_ops;
    if (UNI) n = __builtin_amdgcn_readfirstlane(n);
    for (int i = 0; i < n; ++i) {
        const CsgOpWord o = csg_decode<UNI>(P, i);
        const float *a = P.prm + o.prm;
        if (o.op >= CSG_SPHERE && o.op <= CSG_CYLINDER_Z && o.dst == 0 && o.src == 0) {
            v.a.s0 = csg_combine(o.acc, v.a.s0, csg_prim(o.op, a, qx.a.x, qy.a.x, qz.a.x));
            continue;
        }
        const int ps = o.src & (kCsgPts - 1);
        const float x = qx.get(ps), y = qy.get(ps), z = qz.get(ps);
        if (o.op == CSG_TRANSLATE || o.op == CSG_ROTATE) {
            float nx, ny, nz;
            if (o.op == CSG_TRANSLATE) {
                nx = x - a[0];
                ny = y - a[1];
                nz = z - a[2];
            } else {   // row vector times R: ((x R0j + y R1j) + z R2j)
                nx = x * a[0] + y * a[3];
                nx = nx + z * a[6];
                ny = x * a[1] + y * a[4];
                ny = ny + z * a[7];
                nz = x * a[2] + y * a[5];
                nz = nz + z * a[8];
            }
            const int pd = o.dst & (kCsgPts - 1);
            qx.set(pd, nx);
            qy.set(pd, ny);
            qz.set(pd, nz);
            continue;
        }
        const float cur = v.get(o.dst);
        float r;
        if (o.op >= CSG_SPHERE && o.op <= CSG_CYLINDER_Z) {
            r = csg_prim(o.op, a, x, y, z);
        } else if (o.op == CSG_VALUE) {
            r = v.get(o.src);
        } else if (o.op == CSG_SMOOTH_MIN) {
            const float k = a[0], b = v.get(o.src);
            r = (-k) * logf(expf(-cur / k) + expf(-b / k));
        } else if (o.op == CSG_NEG) {
            r = -cur;
        } else {
            r = cur;
        }
        v.set(o.dst, csg_combine(o.acc, cur, r));
    }
    return v.a.s0;
}

// Value and gradient d value / d p by forward differentiation, with torch
// autograd's conventions at kinks: |x|' = sign(x) (0 at 0), a zero-length
// norm has gradient 0, max(q, 0) and min(m, 0) split ties evenly, and
// min / max pass the gradient of the operand they take.
template <bool UNI>
__device__ float csg_value_grad(const CsgProgram &P, float px, float py, float pz, float g[3]) {
    CsgSlots8 v, gx, gy, gz;
    CsgSlots4 qx, qy, qz;
    CsgSlots4 J00, J01, J02, J10, J11, J12, J20, J21, J22;   // Jij = d q_s[j] / d p[i] per slot s
    v.fill(0.0f);
    gx.fill(0.0f);
    gy.fill(0.0f);
    gz.fill(0.0f);
    qx.fill(0.0f);
    qy.fill(0.0f);
    qz.fill(0.0f);
    J00.fill(0.0f);
    J01.fill(0.0f);
    J02.fill(0.0f);
    J10.fill(0.0f);
    J11.fill(0.0f);
    J12.fill(0.0f);
    J20.fill(0.0f);
    J21.fill(0.0f);
    J22.fill(0.0f);
    J00.a.x = J11.a.x = J22.a.x = 1.0f;
    qx.a.x = px;
    qy.a.x = py;
    qz.a.x = pz;
    int n = P.n_ops;
    if (UNI) n = __builtin_amdgcn_readfirstlane(n);
    for (int i = 0; i < n; ++i) {
        const CsgOpWord o = csg_decode<UNI>(P, i);
        const float *a = P.prm + o.prm;
        const int ps = o.src & (kCsgPts - 1);
        const float x[3] = {qx.get(ps), qy.get(ps), qz.get(ps)};
        const float Jp[3][3] = {{J00.get(ps), J01.get(ps), J02.get(ps)},
                                {J10.get(ps), J11.get(ps), J12.get(ps)},
                                {J20.get(ps), J21.get(ps), J22.get(ps)}};
        if (o.op == CSG_TRANSLATE || o.op == CSG_ROTATE) {
            float nq[3], nJ[3][3];
            if (o.op == CSG_TRANSLATE) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    nq[c] = x[c] - a[c];
#pragma unroll
                    for (int j = 0; j < 3; ++j) nJ[c][j] = Jp[c][j];
                }
            } else {
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const float t = x[0] * a[j] + x[1] * a[3 + j];
                    nq[j] = t + x[2] * a[6 + j];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const float u = Jp[c][0] * a[j] + Jp[c][1] * a[3 + j];
                        nJ[c][j] = u + Jp[c][2] * a[6 + j];
                    }
                }
            }
            const int pd = o.dst & (kCsgPts - 1);
            qx.set(pd, nq[0]);
            qy.set(pd, nq[1]);
            qz.set(pd, nq[2]);
            J00.set(pd, nJ[0][0]);
            J01.set(pd, nJ[0][1]);
            J02.set(pd, nJ[0][2]);
            J10.set(pd, nJ[1][0]);
            J11.set(pd, nJ[1][1]);
            J12.set(pd, nJ[1][2]);
            J20.set(pd, nJ[2][0]);
            J21.set(pd, nJ[2][1]);
            J22.set(pd, nJ[2][2]);
            continue;
        }
        const float cur = v.get(o.dst);
        const float cg[3] = {gx.get(o.dst), gy.get(o.dst), gz.get(o.dst)};
        float r, dq[3] = {0.0f, 0.0f, 0.0f}, rg[3] = {0.0f, 0.0f, 0.0f};
        bool local = true;   // rg = Jp dq (a primitive of the point slot) or given directly
        switch (o.op) {
        case CSG_SPHERE:
        case CSG_SPHERE_AT: {
            const bool at = o.op == CSG_SPHERE_AT;
            const float d0 = at ? x[0] - a[0] : x[0], d1 = at ? x[1] - a[1] : x[1], d2 = at ? x[2] - a[2] : x[2];
            const float l = csg_norm3(d0, d1, d2);
            r = l - (at ? a[3] : a[0]);
            if (l != 0.0f) {
                dq[0] = d0 / l;
                dq[1] = d1 / l;
                dq[2] = d2 / l;
            }
            break;
        }
        case CSG_TORUS: {
            const float l = csg_norm2(x[0], x[1]), t = l - a[0], m = csg_norm2(t, x[2]);
            r = m - a[1];
            const float dt = m != 0.0f ? t / m : 0.0f;
            if (l != 0.0f) {
                dq[0] = dt * (x[0] / l);
                dq[1] = dt * (x[1] / l);
            }
            dq[2] = m != 0.0f ? x[2] / m : 0.0f;
            break;
        }
        case CSG_CUBOID: {
            float u[3], o3[3], sg[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                u[c] = fabsf(x[c]) - a[c];
                o3[c] = u[c] != u[c] ? u[c] : fmaxf(u[c], 0.0f);
                sg[c] = x[c] > 0.0f ? 1.0f : (x[c] < 0.0f ? -1.0f : 0.0f);
            }
            int am = 0;
            float m = u[0];
            if (csg_take_max(m, u[1])) {
                m = u[1];
                am = 1;
            }
            if (csg_take_max(m, u[2])) {
                m = u[2];
                am = 2;
            }
            const float in = m != m ? m : fminf(m, 0.0f);
            const float l = csg_norm3(o3[0], o3[1], o3[2]);
            r = l + in;
            const float din = m < 0.0f ? 1.0f : (m == 0.0f ? 0.5f : 0.0f);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float dmax = u[c] > 0.0f ? 1.0f : (u[c] == 0.0f ? 0.5f : 0.0f);
                float du = l != 0.0f ? (o3[c] / l) * dmax : 0.0f;
                if (c == am) du = du + din;
                dq[c] = du * sg[c];
            }
            break;
        }
        case CSG_CYLINDER_Z: {
            const float l = csg_norm2(x[0], x[1]);
            r = l - a[0];
            if (l != 0.0f) {
                dq[0] = x[0] / l;
                dq[1] = x[1] / l;
            }
            break;
        }
        case CSG_VALUE:
            r = v.get(o.src);
            rg[0] = gx.get(o.src);
            rg[1] = gy.get(o.src);
            rg[2] = gz.get(o.src);
            local = false;
            break;
        case CSG_SMOOTH_MIN: {
            const float k = a[0], b = v.get(o.src);
            const float e1 = expf(-cur / k), e2 = expf(-b / k), s = e1 + e2;
            r = (-k) * logf(s);
            const float w1 = e1 / s, w2 = e2 / s;
            rg[0] = w1 * cg[0] + w2 * gx.get(o.src);
            rg[1] = w1 * cg[1] + w2 * gy.get(o.src);
            rg[2] = w1 * cg[2] + w2 * gz.get(o.src);
            local = false;
            break;
        }
        case CSG_NEG:
            r = -cur;
            rg[0] = -cg[0];
            rg[1] = -cg[1];
            rg[2] = -cg[2];
            local = false;
            break;
        default:
            r = cur;
            rg[0] = cg[0];
            rg[1] = cg[1];
            rg[2] = cg[2];
            local = false;
            break;
        }
        if (local) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float t = Jp[c][0] * dq[0] + Jp[c][1] * dq[1];
                rg[c] = t + Jp[c][2] * dq[2];
            }
        }
        bool take = true;
        if (o.acc == CSG_ACC_MIN) take = csg_take_min(cur, r);
        else if (o.acc == CSG_ACC_MAX) take = csg_take_max(cur, r);
        v.set(o.dst, take ? r : cur);
        gx.set(o.dst, take ? rg[0] : cg[0]);
        gy.set(o.dst, take ? rg[1] : cg[1]);
        gz.set(o.dst, take ? rg[2] : cg[2]);
    }
    g[0] = gx.a.s0;
    g[1] = gy.a.s0;
    g[2] = gz.a.s0;
    return v.a.s0;
}

// Unit normal from a gradient as torch.nn.functional.normalize (eps 1e-12).
__device__ __forceinline__ void csg_normalize(const float g[3], float n[3]) {
    const float l = fmaxf(csg_norm3(g[0], g[1], g[2]), 1e-12f);
    n[0] = g[0] / l;
    n[1] = g[1] / l;
    n[2] = g[2] / l;
}

// Sign of the whole box of radius R (largest distance from the evaluated
// centre to any grid point of the box) from the value vc at its centre:
// +1 every grid point is outside (v - level >= 0), -1 every point is inside,
// 0 undecided.  The field is 1-Lipschitz (rotations are orthonormal to FP32
// rounding, covered by the 1e-4 relative slack) and the margin of
// 1e-4 * (sum |params| + coordinate scale + R + |vc|) is far above the FP32
// evaluation error, so a decided box holds no point whose evaluated value
// has the other sign.
__device__ __forceinline__ int csg_box_sign(const CsgProgram &P, float vc, double R, double coord_scale,
                                            float level) {
    if (!P.lipschitz) return 0;
    const double m = 1e-4 * ((double) P.pscale + coord_scale + R + fabs((double) vc));
    const double d = (double) vc - (double) level, lim = R * 1.0001 + m;
    if (!(fabs(d) <= 1e300) || !(lim <= 1e300)) return 0;   // NaN or infinite: decide per point
    if (d > lim) return 1;
    if (-d > lim) return -1;
    return 0;
}

// Stage a program into LDS (all threads of the block call it).
__device__ __forceinline__ void csg_stage(CsgProgram &dst, const CsgProgram *src) {
    const int n = src->n_ops;
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst.ops[i] = src->ops[i];
    for (int i = threadIdx.x; i < kCsgMaxParams; i += blockDim.x) dst.prm[i] = src->prm[i];
    if (threadIdx.x == 0) {
        dst.n_ops = n;
        dst.lipschitz = src->lipschitz;
        dst.pscale = src->pscale;
    }
}

}   // namespace iso

namespace iso {
// Host: validate a program (ops: n_ops x 2 int32, params: n_params floats,
// both in host memory) into its device form; returns an error message or
// nullptr.
const char *csg_build(const int32_t *ops, int n_ops, const float *params, int n_params, CsgProgram &out);
// Program values at the grid points of (shape, aabb) / at n points (values
// and unit gradients, either may be null).
hipError_t sdf_grid_launch(const CsgProgram *prog, const int64_t shape[3], const float amin[3], const float amax[3],
                           float *out, hipStream_t st);
hipError_t sdf_points_launch(const CsgProgram *prog, const float *pts, int64_t n, float *values, float *normals,
                             hipStream_t st);
}   // namespace iso
