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
    int64_t *cell_inst;   // first instance of each active cell
    float *ipos;          // instance positions
    int64_t *isrc;        // slot * 12 + edge of each instance
    int64_t *islot;       // hash slot of each instance
    unsigned long long *hkey, *hmin;
    uint32_t *hused;
    int64_t hmask;
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
hipError_t sparse_crossing_flags(const float *values, int64_t n, float level, uint8_t *flags, hipStream_t st);
size_t sparse_mc_work_bytes(int64_t n, int64_t ninst);
hipError_t sparse_mc_count(const SparseMcIn &in, int64_t *scratch, int64_t &ninst, hipStream_t st);
hipError_t sparse_mc(const SparseMcIn &in, int64_t ninst, void *work, size_t work_bytes, SparseMcOut &o,
                     int64_t &nv, int64_t &nf, hipStream_t st);

}   // namespace iso
