// Backward of marching cubes: vertex-position gradients -> grid values.
//
// A vertex on the edge from point a (value va, position Pa) to point b is
// v = (1 - t) Pa + t Pb with t = (L - va) / (vb - va) (src/mc/nagae.cu:55-59).
// Hence dv/dva = (Pb - Pa) (L - vb) / (vb - va)^2 and
//       dv/dvb = -(Pb - Pa) (L - va) / (vb - va)^2.
// A welded vertex takes its position from one representative edge, so its
// gradient goes to that edge's two values.  Accumulation uses float atomics
// (a grid value is shared by up to six edges).

#include "backward.h"
#include "common.h"

namespace iso {

namespace {

constexpr int kB = 256;

__device__ __forceinline__ void edge_grad(const float *g, float va, float vb, float level, const float pa[3],
                                          const float pb[3], float &ga, float &gb) {
    float d = vb - va;
    ga = gb = 0.0f;
    if (!(d != 0.0f)) return;
    float s = g[0] * (pb[0] - pa[0]) + g[1] * (pb[1] - pa[1]) + g[2] * (pb[2] - pa[2]);
    float inv2 = 1.0f / (d * d);
    ga = s * (level - vb) * inv2;
    gb = -s * (level - va) * inv2;
    if (!isfinite(ga)) ga = 0.0f;
    if (!isfinite(gb)) gb = 0.0f;
}

__global__ void k_bwd_dense(const float *grad_v, const int64_t *keys, int64_t nv, const float *values,
                            DenseGeom g, float level, float *grad_values) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (i >= nv) return;
    int64_t key = keys[i];
    const int64_t p = div_fast(key, 3);
    const int axis = (int) (key - 3 * p);
    const int64_t r = div_fast(p, g.Z), z = p - r * g.Z;
    const int64_t x = div_fast(r, g.Y), y = r - x * g.Y;
    int64_t stride = axis == 0 ? g.Y * g.Z : (axis == 1 ? g.Z : 1);
    int64_t q = p + stride;
    float pa[3] = {grid_coord(x, g.X, g.amin[0], g.amax[0]), grid_coord(y, g.Y, g.amin[1], g.amax[1]),
                   grid_coord(z, g.Z, g.amin[2], g.amax[2])};
    float pb[3] = {pa[0], pa[1], pa[2]};
    if (axis == 0) pb[0] = grid_coord(x + 1, g.X, g.amin[0], g.amax[0]);
    if (axis == 1) pb[1] = grid_coord(y + 1, g.Y, g.amin[1], g.amax[1]);
    if (axis == 2) pb[2] = grid_coord(z + 1, g.Z, g.amin[2], g.amax[2]);
    float ga, gb;
    edge_grad(grad_v + 3 * i, values[p], values[q], level, pa, pb, ga, gb);
    if (ga != 0.0f) atomicAdd(&grad_values[p], ga);
    if (gb != 0.0f) atomicAdd(&grad_values[q], gb);
}


__global__ void k_bwd_sparse(const float *grad_v, const int64_t *vsrc, int64_t nv, const float *values,
                             const int64_t *ids, DenseGeom g, float level, float *grad_values) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (i >= nv) return;
    int64_t src = vsrc[i];
    const int64_t slot = div_fast(src, 12);
    const int e = (int) (src - 12 * slot);
    const int64_t id = ids[slot];
    const int64_t cz = g.Z - 1, cy = g.Y - 1;
    const int64_t r = div_fast(id, cz), z = id - r * cz;
    const int64_t x = div_fast(r, cy), y = r - x * cy;
    int ca = edge_a(e), cb = edge_b(e);
    float pa[3] = {grid_coord(x + (ca >> 2), g.X, g.amin[0], g.amax[0]),
                   grid_coord(y + ((ca >> 1) & 1), g.Y, g.amin[1], g.amax[1]),
                   grid_coord(z + (ca & 1), g.Z, g.amin[2], g.amax[2])};
    float pb[3] = {grid_coord(x + (cb >> 2), g.X, g.amin[0], g.amax[0]),
                   grid_coord(y + ((cb >> 1) & 1), g.Y, g.amin[1], g.amax[1]),
                   grid_coord(z + (cb & 1), g.Z, g.amin[2], g.amax[2])};
    float ga, gb;
    edge_grad(grad_v + 3 * i, values[8 * slot + ca], values[8 * slot + cb], level, pa, pb, ga, gb);
    if (ga != 0.0f) atomicAdd(&grad_values[8 * slot + ca], ga);
    if (gb != 0.0f) atomicAdd(&grad_values[8 * slot + cb], gb);
}

}   // namespace

hipError_t mc_backward_dense(const float *grad_v, const int64_t *keys, int64_t nv, const float *values,
                             const DenseGeom &g, float level, float *grad_values, hipStream_t st) {
    if (nv == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bwd_dense, dim3((unsigned) ((nv + kB - 1) / kB)), dim3(kB), 0, st, grad_v, keys, nv, values,
                       g, level, grad_values);
    return hipGetLastError();
}

hipError_t mc_backward_sparse(const float *grad_v, const int64_t *vsrc, int64_t nv, const float *values,
                              const int64_t *ids, const DenseGeom &g, float level, float *grad_values,
                              hipStream_t st) {
    if (nv == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bwd_sparse, dim3((unsigned) ((nv + kB - 1) / kB)), dim3(kB), 0, st, grad_v, vsrc, nv,
                       values, ids, g, level, grad_values);
    return hipGetLastError();
}

}   // namespace iso
