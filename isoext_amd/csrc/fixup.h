#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mc_dense.h"

namespace iso {

struct FixupGrid {
    int64_t X, Y, Z;   // points per axis
    // Optional, dense point-major numbering: exclusive vertex / face prefix of
    // (point row, z-segment) entry e = blk[e / kMcScanBlock] + cnt[e], S
    // segments per row.  Lets the face remap skip the faces that cannot
    // reference a relocated tail vertex.
    const int64_t *blk_v = nullptr, *blk_f = nullptr;
    const uint32_t *cnt_v = nullptr, *cnt_f = nullptr;
    int64_t S = 0;
};

hipError_t weld_fixup(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                      const FixupGrid &g, hipStream_t st, int64_t *keys = nullptr);
// Stream-ordered variant: no host synchronisation (except to grow the
// scratch); final_dev[0], [1] receive the final vertex / face counts and
// final_dev[2] is non-zero when the lists were inconsistent.  nv / nf are the
// provisional counts.
hipError_t weld_fixup_async(const SpecialLists &sp, float *v, int32_t *f, int64_t nv, int64_t nf,
                            const FixupGrid &g, hipStream_t st, int64_t *keys, int64_t *final_dev);

}   // namespace iso
