#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace iso {

struct SparseGeom {
    int64_t X, Y, Z;   // points per axis
    float amin[3], amax[3];
};

struct SparseMcIn {
    const float *values;   // (n, 8)
    const int64_t *ids;    // (n) dense cell ids
    int64_t n;
    SparseGeom g;
    float level;
    int method;
};

struct SparseMcWork {
    uint32_t *cst;        // per cell: crossing-edge mask (bits 0-11), case index (bits 16-23)
    int64_t *cell_inst;   // first instance of each active cell
    int32_t *inst_cell;   // cell slot of each instance
    float *ipos;          // instance positions
    int64_t *isrc;        // slot * 12 + edge of each instance
    int64_t *rkey;        // instance index of its representative (first instance at its position)
    uint8_t *used;        // the representative is a corner of a kept face
    int32_t *vid;         // vertex id of the instance's representative
    int *err;
};

struct SparseMcOut {
    float *v;
    int64_t cap_v;
    int32_t *f;
    int64_t cap_f;
    int64_t *vsrc;   // optional: slot * 12 + edge of each output vertex
};

hipError_t sparse_corner_points(const int64_t *ids, int64_t n, const SparseGeom &g, float *out, hipStream_t st);
// Sorted-key set operations of the population path: union of sorted unique a
// with sorted b (duplicates allowed), or a minus b.  out (capacity na + nb /
// na) receives the sorted result, slot[k] the index in a of out[k] (-1 for a
// key new from b), *n_out (device) its length.  No host synchronisation.
int64_t sparse_setop_scratch_bytes(int64_t na, int64_t nb);
hipError_t sparse_setop(const int64_t *a, int64_t na, const int64_t *b, int64_t nb, bool difference, void *scratch,
                        int64_t *out, int64_t *slot, int64_t *n_out, hipStream_t st);
// out row k = old_values row slot[k], or dflt where slot[k] < 0 (8 floats a row).
hipError_t sparse_gather_rows(const float *old_values, const int64_t *slot, int64_t n, float dflt, float *out,
                              hipStream_t st);
hipError_t sparse_crossing_flags(const float *values, int64_t n, float level, uint8_t *flags, hipStream_t st);
size_t sparse_mc_work_bytes(int64_t n, int64_t ninst);
hipError_t sparse_mc_count(const SparseMcIn &in, int64_t *scratch, int64_t &ninst, hipStream_t st);
hipError_t sparse_mc(const SparseMcIn &in, int64_t ninst, void *work, size_t work_bytes, SparseMcOut &o,
                     int64_t &nv, int64_t &nf, hipStream_t st);

}   // namespace iso
