// On-device SDF generation: the union (min) of K spheres, and programs of
// the reference's SDF primitives and CSG operators (csg.h), evaluated into a
// dense (X, Y, Z) grid at the grid coordinates of get_vtx_pos_op
// (include/utils.cuh:62-80), or at given points with analytic normals.

#include <cmath>
#include <cstring>

#include "common.h"
#include "csg.h"

namespace iso {

namespace {

constexpr int kMaxSpheres = 256;

// Each distance is sqrt((dx*dx + dy*dy) + dz*dz) - r in fp32.
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

// Program values at every grid point (z fastest, coalesced stores).
__global__ void __launch_bounds__(256) k_sdf_grid(const CsgProgram *prog, float *out, int64_t X, int64_t Y, int64_t Z,
                                                  float ax, float ay, float az, float bx, float by, float bz) {
    __shared__ CsgProgram sp;
    csg_stage(sp, prog);
    __syncthreads();
    const int64_t n = X * Y * Z;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const int64_t r = i / Z, z = i - r * Z, x = r / Y, y = r - x * Y;
        out[i] = csg_value<true>(sp, grid_coord(x, X, ax, bx), grid_coord(y, Y, ay, by), grid_coord(z, Z, az, bz));
    }
}

// Program values (optional) and unit gradients (optional) at n points.
__global__ void __launch_bounds__(256) k_sdf_points(const CsgProgram *prog, const float *pts, int64_t n,
                                                    float *values, float *normals) {
    __shared__ CsgProgram sp;
    csg_stage(sp, prog);
    __syncthreads();
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        float v;
        if (normals) {
            float g[3], nv[3];
            v = csg_value_grad<true>(sp, px, py, pz, g);
            csg_normalize(g, nv);
            normals[3 * i] = nv[0];
            normals[3 * i + 1] = nv[1];
            normals[3 * i + 2] = nv[2];
        } else {
            v = csg_value<true>(sp, px, py, pz);
        }
        if (values) values[i] = v;
    }
}

unsigned stride_blocks(int64_t n) { return (unsigned) std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

// Largest singular value of a 3x3 matrix (power iteration on M^T M, FP64).
double spectral_norm3(const float *m) {
    double a[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += (double) m[3 * k + i] * (double) m[3 * k + j];
            a[i][j] = s;
        }
    double v[3] = {1.0, 0.7, 0.3}, lam = 0.0;
    for (int it = 0; it < 200; ++it) {
        double w[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) w[i] += a[i][j] * v[j];
        const double l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (!(l > 0.0)) return 0.0;
        for (int i = 0; i < 3; ++i) v[i] = w[i] / l;
        lam = l;
    }
    return std::sqrt(lam);
}

}   // namespace

const char *csg_build(const int32_t *ops, int n_ops, const float *params, int n_params, CsgProgram &out) {
    if (!ops || n_ops < 1 || n_ops > kCsgMaxOps) return "an SDF program holds 1 to 256 ops";
    if (n_params < 0 || n_params > kCsgMaxParams || (n_params > 0 && !params))
        return "an SDF program holds at most 1024 parameters";
    std::memset(&out, 0, sizeof out);
    out.n_ops = n_ops;
    out.lipschitz = 1;
    double ps = 0.0;
    for (int i = 0; i < n_params; ++i) {
        out.prm[i] = params[i];
        if (!std::isfinite(params[i])) out.lipschitz = 0;
        else ps += std::fabs((double) params[i]);
    }
    for (int i = 0; i < n_ops; ++i) {
        const int w = ops[2 * i], o = ops[2 * i + 1];
        const int op = w & 0xff, dst = (w >> 8) & 15, src = (w >> 12) & 15, acc = (w >> 16) & 0xffff;
        if (op < CSG_SPHERE || op > CSG_ROTATE) return "unknown SDF program op";
        if (acc > CSG_ACC_MAX) return "SDF program op with an unknown combine mode";
        const bool point_op = op == CSG_TRANSLATE || op == CSG_ROTATE;
        const bool prim = op >= CSG_SPHERE && op <= CSG_CYLINDER_Z;
        if (point_op && (dst >= kCsgPts || src >= kCsgPts || acc != 0)) return "SDF program transform slots out of range";
        if (prim && (dst >= kCsgVals || src >= kCsgPts)) return "SDF program primitive slots out of range";
        if (!point_op && !prim) {
            if (dst >= kCsgVals || src >= kCsgVals) return "SDF program value slots out of range";
            if ((op == CSG_SMOOTH_MIN || op == CSG_NEG) && acc != 0) return "SDF program op cannot combine";
        }
        const int np = csg_param_count(op);
        if (np > 0 && (o < 0 || o > n_params - np)) return "SDF program parameter offset out of range";
        out.ops[i].x = w;
        out.ops[i].y = np > 0 ? o : 0;
        if (op == CSG_SMOOTH_MIN) out.lipschitz = 0;
        if (op == CSG_ROTATE && !(spectral_norm3(params + o) <= 1.0 + 1e-5)) out.lipschitz = 0;
    }
    out.pscale = (float) std::min(ps, 1e30);
    return nullptr;
}

hipError_t sdf_grid_launch(const CsgProgram *prog, const int64_t shape[3], const float amin[3], const float amax[3],
                           float *out, hipStream_t st) {
    const int64_t n = shape[0] * shape[1] * shape[2];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sdf_grid, dim3(stride_blocks(n)), dim3(256), 0, st, prog, out, shape[0], shape[1], shape[2],
                       amin[0], amin[1], amin[2], amax[0], amax[1], amax[2]);
    return hipGetLastError();
}

hipError_t sdf_points_launch(const CsgProgram *prog, const float *pts, int64_t n, float *values, float *normals,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sdf_points, dim3(stride_blocks(n)), dim3(256), 0, st, prog, pts, n, values, normals);
    return hipGetLastError();
}

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
