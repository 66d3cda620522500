#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iso {

struct DenseGeom {
    int64_t X, Y, Z;   // points per axis
    float amin[3], amax[3];
};

hipError_t mc_backward_dense(const float *grad_v, const int64_t *keys, int64_t nv, const float *values,
                             const DenseGeom &g, float level, float *grad_values, hipStream_t st);
hipError_t mc_backward_sparse(const float *grad_v, const int64_t *vsrc, int64_t nv, const float *values,
                              const int64_t *ids, const DenseGeom &g, float level, float *grad_values,
                              hipStream_t st);

}   // namespace iso
