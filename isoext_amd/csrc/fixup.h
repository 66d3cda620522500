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

// Scratch a caller supplies so that a fix-up allocates nothing: device bytes
// (at least the *_bytes functions' figures) and host staging (small: item
// records and counts).  Without it the fix-up keeps per-thread caches.
struct FixupScratch {
    char *dev;
    size_t dev_bytes;
    char *host;
    size_t host_bytes;
};

hipError_t weld_fixup(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                      const FixupGrid &g, hipStream_t st, int64_t *keys = nullptr,
                      const FixupScratch *ext = nullptr);
// Device scratch bytes weld_fixup needs for these lists and provisional counts.
size_t weld_fixup_bytes(const SpecialLists &sp, int64_t nv, int64_t nf);
// One weld fix-up for one workgroup: every phase of weld_fixup in a single
// kernel (hash tables in global scratch, removed / dropped lists sorted in
// LDS), for lists small enough (fixup_small_ok).  Items are independent: a
// batch launches one workgroup per item.
struct FixSmallItem {
    SpecialLists sp;
    float *v;
    int32_t *f;
    int64_t *vkeys;
    int64_t nv, nf;
    FixupGrid g;
    int64_t *keys;   // nslots
    void *info;      // nslots vertex records
    int32_t *pslot;  // npslots
    void *groups;    // npslots position groups
    int64_t nslots, npslots;
    int64_t *final_dev;   // final vertex / face counts, inconsistency flag
};
constexpr int64_t kFixSmallMax = 2048;     // removed vertices and dropped faces the small path holds
constexpr int64_t kFixSmallSlots = 2048;   // candidate + special-edge entries the small path takes (one workgroup is slower beyond)
// final[2] of an item whose removed vertices did not fit the small path: its
// outputs are untouched and the multi-kernel fix-up must run instead
constexpr int64_t kFixSmallOverflow = 2;
constexpr int kFixMaxDevices = 64;        // per-thread scratch caches are kept per device up to this id
bool fixup_small_ok(const SpecialLists &sp);
// Scratch bytes of one small fix-up item and the item's table sizes.
size_t fixup_small_scratch(const SpecialLists &sp, int64_t &nslots, int64_t &npslots);
hipError_t fixup_small_launch(const FixSmallItem *items_dev, int n, hipStream_t st);
// Small fix-ups of n items in one launch and one synchronisation; final[3i],
// final[3i+1] = final vertex / face counts of item i, final[3i+2] != 0 when
// its lists were inconsistent.
hipError_t weld_fixup_small_batch(int n, const SpecialLists *sp, float *const *v, int32_t *const *f,
                                  int64_t *const *vkeys, const int64_t *nv, const int64_t *nf, const FixupGrid *g,
                                  hipStream_t st, int64_t *final_host, const FixupScratch *ext = nullptr);
// Device bytes (and host staging bytes) of weld_fixup_small_batch.
size_t weld_fixup_small_batch_bytes(int n, const SpecialLists *sp, size_t *host_bytes);

// Stream-ordered variant: no host synchronisation (except to grow the
// scratch); final_dev[0], [1] receive the final vertex / face counts and
// final_dev[2] is non-zero when the lists were inconsistent.  nv / nf are the
// provisional counts.
hipError_t weld_fixup_async(const SpecialLists &sp, float *v, int32_t *f, int64_t nv, int64_t nf,
                            const FixupGrid &g, hipStream_t st, int64_t *keys, int64_t *final_dev);

}   // namespace iso
