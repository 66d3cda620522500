// Shared device/host helpers for the iso-surface kernels (gfx950).
//
// Float semantics follow the reference exactly, with every operation
// uncontracted (the whole library is built with -ffp-contract=off):
//   grid coordinate  get_vtx_pos_op   include/utils.cuh:62-80
//   case bits        get_case_num_op  include/utils.cuh:91-101
//   edge parameter   process_cube_op  src/mc/nagae.cu:55-57
//   lerp                              include/math.cuh:77-81
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define ISO_HD __host__ __device__ __forceinline__

namespace iso {

// Corner k of a cell sits at (k >> 2, (k >> 1) & 1, k & 1); edge e joins
// corners edge_a(e) < edge_b(e) (include/shared_luts.cuh:34-38).
// Edge axis: 0 = x, 1 = y, 2 = z.  Owner corner = the lower endpoint.
// Packed in immediates (4-bit / 2-bit fields) so that a lookup with a
// per-lane edge index is ALU work rather than a memory load.
constexpr uint64_t kEdgeAPacked = 0x231046540210ull;   // 0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2
constexpr uint64_t kEdgeBPacked = 0x675467752331ull;   // 1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6
constexpr uint32_t kEdgeAxisPacked = 0x6666u;          // 2, 1, 2, 1, 2, 1, 2, 1, 0, 0, 0, 0
ISO_HD constexpr int edge_a(int e) { return (int) ((kEdgeAPacked >> (4 * e)) & 15); }
ISO_HD constexpr int edge_b(int e) { return (int) ((kEdgeBPacked >> (4 * e)) & 15); }
ISO_HD constexpr int edge_axis(int e) { return (int) ((kEdgeAxisPacked >> (2 * e)) & 3); }

// Inclusive scan of `v` over the 64 lanes of a wave (lane i gets
// op(v_0, ..., v_i)) with DPP lane moves instead of LDS permutes: four
// row_shr steps scan each 16-lane row, then row_bcast:15 carries a row's
// last lane into the next row (used by rows 1 and 3) and row_bcast:31
// carries lane 31 into rows 2 and 3.  All 64 lanes must be active.
template <class Op>
__device__ __forceinline__ int wave_incl_scan(int v, Op op) {
    const int lane = (int) __lane_id(), r = lane & 15;
    int t = __builtin_amdgcn_mov_dpp(v, 0x111, 0xf, 0xf, false);
    if (r >= 1) v = op(v, t);
    t = __builtin_amdgcn_mov_dpp(v, 0x112, 0xf, 0xf, false);
    if (r >= 2) v = op(v, t);
    t = __builtin_amdgcn_mov_dpp(v, 0x114, 0xf, 0xf, false);
    if (r >= 4) v = op(v, t);
    t = __builtin_amdgcn_mov_dpp(v, 0x118, 0xf, 0xf, false);
    if (r >= 8) v = op(v, t);
    t = __builtin_amdgcn_mov_dpp(v, 0x142, 0xf, 0xf, false);
    if (lane & 16) v = op(v, t);
    t = __builtin_amdgcn_mov_dpp(v, 0x143, 0xf, 0xf, false);
    if (lane >= 32) v = op(v, t);
    return v;
}
__device__ __forceinline__ int wave_incl_sum(int v) {
    return wave_incl_scan(v, [](int a, int b) { return a + b; });
}
__device__ __forceinline__ int wave_incl_max(int v) {
    return wave_incl_scan(v, [](int a, int b) { return a > b ? a : b; });
}

ISO_HD int32_t f2i_bits(float f) {
    union {
        float f;
        int32_t i;
    } u;
    u.f = f;
    return u.i;
}

ISO_HD float grid_coord(int64_t i, int64_t res, float amin, float amax) {
    float size = amax - amin;
    float p = (float) i / (float) (res - 1);
    float prod = p * size;
    return amin + prod;
}

// a / b for 0 <= a < 2^52 and 0 < b: 32-bit division when both fit, else a
// correctly rounded FP64 quotient (within 1 of the integer one for such a)
// corrected by one step.  Replaces the much longer 64-bit integer division.
ISO_HD int64_t div_fast(int64_t a, int64_t b) {
    if (((uint64_t) a | (uint64_t) b) >> 32 == 0) return (int64_t) ((uint32_t) a / (uint32_t) b);
    int64_t q = (int64_t) ((double) a / (double) b);
    if (q * b > a) --q;
    else if ((q + 1) * b <= a) ++q;
    return q;
}

// Unsigned division of any 32-bit dividend by an invariant divisor d in
// [1, 2^31) with a multiply-high, an add and two shifts (Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994, section 4): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// t = mulhi(m, a), a / d = (t + ((a - t) >> min(l, 1))) >> max(l - 1, 0).
struct UDiv32 {
    uint32_t m = 0;
    int s1 = 0, s2 = 0;
    UDiv32() = default;
    ISO_HD explicit UDiv32(uint32_t d) {
        const int l = d > 1 ? 32 - __builtin_clz(d - 1) : 0;
        m = (uint32_t) (div_fast((int64_t) (((1ull << l) - d) << 32), (int64_t) d) + 1);
        s1 = l > 0 ? 1 : 0;
        s2 = l > 0 ? l - 1 : 0;
    }
    ISO_HD uint32_t div(uint32_t a) const {
        const uint32_t t = (uint32_t) (((uint64_t) m * a) >> 32);
        return (t + ((a - t) >> s1)) >> s2;
    }
};

ISO_HD float lerp1(float t, float a, float b) {
    float s = 1.0f - t;
    float u = s * a;
    float v = t * b;
    return u + v;
}

// Interpolation parameter of a crossing edge.  NaN (only from NaN/inf
// inputs) is mapped to 0 so a vertex never carries NaN.
ISO_HD float edge_t(float v0, float v1, float level) {
    float denom = v1 - v0;
    float t = (denom != 0.0f) ? (level - v0) / denom : 0.0f;
    if (t != t) t = 0.0f;
    return t;
}

ISO_HD bool inside(float v, float level) { return (v - level) < 0.0f; }

// Order-preserving map of a float to an integer (ulp steps between two
// floats = difference of their keys; -0 and +0 map to the same key).
ISO_HD int32_t ordered_key(float f) {
    int32_t b = f2i_bits(f);
    return b >= 0 ? b : (int32_t) (0x80000000u - (uint32_t) b);
}

// A vertex on an edge may coincide bitwise with a vertex of another edge only
// when its along-axis coordinate r is within a few ulps of an endpoint
// coordinate (the off-axis coordinates of a lerp between equal values stay
// within 3 ulps of the grid line).  16 ulp steps is a safe margin.
ISO_HD bool near_endpoint(float r, float a, float b) {
    int64_t kr = ordered_key(r), ka = ordered_key(a), kb = ordered_key(b);
    int64_t da = kr - ka, db = kr - kb;
    da = da < 0 ? -da : da;
    db = db < 0 ? -db : db;
    return da <= 16 || db <= 16;
}

ISO_HD uint32_t f3_hash(float x, float y, float z) {
    // -0 and +0 compare equal, so hash their shared ordered key.
    uint32_t h = 2166136261u;
    uint32_t k[3] = {(uint32_t) ordered_key(x), (uint32_t) ordered_key(y),
                     (uint32_t) ordered_key(z)};
    for (int i = 0; i < 3; ++i) {
        h ^= k[i];
        h *= 16777619u;
        h ^= h >> 15;
    }
    return h;
}

ISO_HD uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// trilinear_interp (src/its.cu:170-184), restated op for op
__device__ __forceinline__ float trilinear(float tx, float ty, float tz, const float *c) {
    float a = 1.0f - tz;
    float c00 = c[0] * a + c[1] * tz;
    float c01 = c[2] * a + c[3] * tz;
    float c10 = c[4] * a + c[5] * tz;
    float c11 = c[6] * a + c[7] * tz;
    float b = 1.0f - ty;
    float c0 = c00 * b + c01 * ty;
    float c1 = c10 * b + c11 * ty;
    float d = 1.0f - tx;
    return c0 * d + c1 * tx;
}

// compute_normals_op (src/its.cu:186-268) for one point in a cell.
__device__ inline void cell_normal(const float lo[3], const float hi[3], const float *cv, float px, float py,
                            float pz, float &nx, float &ny, float &nz) {
    float sx = hi[0] - lo[0], sy = hi[1] - lo[1], sz = hi[2] - lo[2];
    float tx = (px - lo[0]) / sx, ty = (py - lo[1]) / sy, tz = (pz - lo[2]) / sz;
    tx = fmaxf(0.01f, fminf(0.99f, tx));
    ty = fmaxf(0.01f, fminf(0.99f, ty));
    tz = fmaxf(0.01f, fminf(0.99f, tz));
    const float eps = 0.02f;
    float pxx = fminf(tx + eps, 0.99f), mxx = fmaxf(tx - eps, 0.01f);
    float pyy = fminf(ty + eps, 0.99f), myy = fmaxf(ty - eps, 0.01f);
    float pzz = fminf(tz + eps, 0.99f), mzz = fmaxf(tz - eps, 0.01f);
    float dx = (trilinear(pxx, ty, tz, cv) - trilinear(mxx, ty, tz, cv)) / ((pxx - mxx) * sx);
    float dy = (trilinear(tx, pyy, tz, cv) - trilinear(tx, myy, tz, cv)) / ((pyy - myy) * sy);
    float dz = (trilinear(tx, ty, pzz, cv) - trilinear(tx, ty, mzz, cv)) / ((pzz - mzz) * sz);
    float d2 = dx * dx + dy * dy;
    d2 = d2 + dz * dz;
    float len = sqrtf(d2);
    if (len > 1e-8f) {
        nx = dx / len;
        ny = dy / len;
        nz = dz / len;
    } else {
        nx = 0.0f;
        ny = 0.0f;
        nz = 1.0f;
    }
}

}   // namespace iso
