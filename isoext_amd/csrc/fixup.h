#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_dense.h"

namespace iso {

struct FixupGrid {
    int64_t X, Y, Z;   // points per axis
};

hipError_t weld_fixup(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                      const FixupGrid &g, hipStream_t st, int64_t *keys = nullptr);

}   // namespace iso
