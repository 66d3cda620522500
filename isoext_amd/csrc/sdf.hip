// On-device SDF generation for the benchmark workloads: the union (min) of K
// spheres evaluated directly into a dense (X, Y, Z) grid in one pass.  Each
// distance is sqrt((dx*dx + dy*dy) + dz*dz) - r in fp32 with the grid
// coordinates of get_vtx_pos_op (include/utils.cuh:62-80).

#include "common.h"

namespace iso {

namespace {

constexpr int kMaxSpheres = 256;

__global__ void k_sdf_spheres(float *out, int64_t X, int64_t Y, int64_t Z, float ax, float ay, float az, float bx,
                              float by, float bz, const float4 *spheres, int k) {
    __shared__ float4 sp[kMaxSpheres];
    for (int i = threadIdx.x; i < k; i += blockDim.x) sp[i] = spheres[i];
    __syncthreads();
    const int64_t n = X * Y * Z;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        int64_t z = i % Z, y = (i / Z) % Y, x = i / (Y * Z);
        float px = grid_coord(x, X, ax, bx), py = grid_coord(y, Y, ay, by), pz = grid_coord(z, Z, az, bz);
        float best = INFINITY;
        for (int s = 0; s < k; ++s) {
            float dx = px - sp[s].x, dy = py - sp[s].y, dz = pz - sp[s].z;
            float d2 = dx * dx + dy * dy;
            d2 = d2 + dz * dz;
            float d = sqrtf(d2) - sp[s].w;
            best = fminf(best, d);
        }
        out[i] = best;
    }
}

}   // namespace

}   // namespace iso

extern "C" int isoext_sdf_spheres(const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                                  const float *spheres, int k, float *out, void *stream) {
    if (k < 1 || k > iso::kMaxSpheres) return -1;
    int64_t n = shape[0] * shape[1] * shape[2];
    if (n == 0) return 0;
    unsigned grid = (unsigned) std::min<int64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(iso::k_sdf_spheres, dim3(grid), dim3(256), 0, (hipStream_t) stream, out, shape[0], shape[1],
                       shape[2], aabb_min[0], aabb_min[1], aabb_min[2], aabb_max[0], aabb_max[1], aabb_max[2],
                       (const float4 *) spheres, k);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
