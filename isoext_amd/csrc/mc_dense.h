// Device-side parameter blocks shared by the marching-cubes kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace iso {

enum : int {
    ISO_ERR_CAPACITY = 1,    // an output buffer was too small (counts are still exact)
    ISO_ERR_SPIN = 2,        // a bounded inter-workgroup wait gave up
    ISO_ERR_TOO_LARGE = 4,   // more than 2^31-1 vertices or faces
    ISO_ERR_SPECIAL = 8,     // a special-vertex list overflowed (counts are still exact)
};

// A vertex whose position may coincide with another edge's vertex.
struct SpecialVertex {
    int64_t vid;    // provisional vertex id
    int64_t edge;   // owner point linear id * 3 + axis
    float x, y, z, pad;
};

// One crossing edge of a cell that holds a special vertex.
struct SpecialEdge {
    int64_t vid;
    int64_t edge;
};

struct SpecialLists {
    unsigned long long n_cand, n_scell, n_kept, n_hole;
    int64_t cap_cand, cap_scell, cap_kept, cap_hole;
    SpecialVertex *cand;   // candidate vertices (emitted once, by their owner)
    SpecialEdge *scell;    // crossing edges of special cells (one entry per cell)
    int64_t *kept;         // face slots kept in special cells
    int64_t *hole;         // face slots dropped as degenerate
};

struct McParams {
    const float *values;
    int64_t X, Y, Z;   // points per axis
    float amin[3], amax[3];
    float level;
    int method;        // 0 = nagae, 1 = lorensen
    int all_special;   // treat every vertex as a weld candidate
    int vec4;          // Z % 4 == 0 and values 16-byte aligned: rows load as float4
    int NI, NJ, NK;    // tiles per axis
    int64_t ntiles;
    float *out_v;
    int64_t cap_v;
    int64_t *out_key;   // optional: owner point linear id * 3 + axis per vertex
    int32_t *out_f;
    int64_t cap_f;
    uint64_t *status;   // per ticket: look-back state
    uint32_t *ticket;
    uint32_t *epoch_ptr;
    uint64_t *pub;      // per tile: published boundary vertex ids
    int *err;
    int64_t *totals;    // [0] vertices, [1] faces (provisional)
    SpecialLists *sp;
};

int mc_config_count();
bool mc_config_tiles(int cfg, const int64_t shape[3], int &ni, int &nj, int &nk, int64_t &pub);
hipError_t mc_dense_launch(int cfg, McParams &p, const SpecialLists &sp_init, hipStream_t stream,
                           hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

}   // namespace iso
