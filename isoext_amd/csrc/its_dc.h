#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace iso {

struct CsgProgram;

struct GridDesc {
    const float *values;         // dense (X, Y, Z) or sparse (n, 8)
    const int64_t *sparse_ids;   // sorted dense cell ids of the sparse slots, or nullptr
    int64_t n_sparse;
    int64_t X, Y, Z;             // points per axis
    float amin[3], amax[3];
    // dense grids in batched dual contouring: the field given by an SDF
    // program (device form, csg.h) instead of values; csg_normals = 1 takes
    // the program's analytic unit gradients as normals (else trilinear)
    const CsgProgram *csg = nullptr;
    int csg_normals = 0;
};

struct ItsOut {
    int64_t *cell_indices;   // dense: cell id; sparse: slot
    int64_t *cell_dense;     // dense cell id
    int64_t *cell_offsets;   // ncells (+1 written by the host)
    float *points;
    float *normals;          // may be null
    uint8_t *edge_local;
    uint8_t *is_out;
    int32_t *entry_cell;
    // batches: the arrays above start at the grid's first cell / entry of
    // batch-wide arrays, whose indices (cell_offsets, entry_cell values) are
    // offset by these bases
    int64_t a_base = 0, e_base = 0;
    // optional, dense grids: per (cell row, 64-cell word) -- index
    // (x * (Y-1) + y) * W + z / 64, W = ceil(Z / 64) -- the word's active-cell
    // mask and the (batch-wide) active index of its first active cell
    uint64_t *amask_tab = nullptr;
    int32_t *apre_tab = nullptr;
};

struct ItsIn {
    int64_t ncells, npts;
    const int64_t *cell_indices;
    const int64_t *cell_dense;
    const int64_t *cell_offsets;
    const float *points;
    const float *normals;
    const uint8_t *edge_local;
    const uint8_t *is_out;
    const int32_t *entry_cell;
    // optional (dense grids): the lookup tables of ItsOut::amask_tab / apre_tab
    const uint64_t *amask_tab = nullptr;
    const int32_t *apre_tab = nullptr;
};

struct DcScratch {
    float *dual_v;
    float *cond;
    unsigned long long *n_fp64;
    uint8_t *used;
    uint8_t *bnd;         // vertex on its cell's box (may coincide with a neighbour's)
    uint8_t *qn;          // dense sets: quads emitted per cell
    uint8_t *qmask;       // dense grids without a set: canonical quad edges and orientations per cell
    int64_t *bsum;
    int64_t *part;        // chunk sums of the block-sum scans (dc_scan_part_pairs(nb) pairs)
    int64_t *tot;
    int32_t *tri_cells;
    int64_t cap_tri;
    int32_t *hslots, *slot_min;
    int64_t hmask;
    int32_t *rep_slot, *final_id;
    int64_t *akey;      // active-cell hash: dense cell id -> active index
    int32_t *aval;
    int64_t amask;
    float *out_v;
    int32_t *out_f;
    int32_t *out_vcell;   // optional: active cell of each output vertex
    int64_t nquads;
};

int64_t its_scratch_pairs(const GridDesc &g);
hipError_t its_count(const GridDesc &g, float level, int64_t *bsum, int64_t *tot, hipStream_t st);
hipError_t its_emit(const GridDesc &g, float level, const int64_t *bsum_pre, const ItsOut &o, hipStream_t st);
hipError_t its_normals(const GridDesc &g, const ItsIn &in, float *normals, hipStream_t st);
// p[0, n) = v on the stream (no host synchronisation).
hipError_t fill_i64(int64_t *p, int64_t n, int64_t v, hipStream_t st);
// Runs dual contouring and synchronises once; n_fp64 returns the number of
// FP64 re-solves.
hipError_t dc_run(const GridDesc &g, const ItsIn &in, float reg, float svd_tol, float cond_fp64, DcScratch &w,
                  int64_t &nv, int64_t &nf, int64_t &n_fp64, hipStream_t st);
// Batched dual contouring of dense grids g[0, n) (n <= 256) whose
// intersection sets are concatenated in `in` (grid i's active cells start at
// cell0[i]; cell offsets and entry cells index the concatenated arrays).
// Output vertices come out grouped by grid in grid order (grid i's first
// vertex is the sum of nv[0..i)), face corners are grid-local vertex ids, and
// quads are grouped likewise.  One host synchronisation; nv / nq per grid.
struct DcBatchDev {
    void *items;              // dc_batch_item_bytes() * n
    int32_t *cell_item;       // per active cell
    int64_t *vbase, *qbase;   // n + 1 each
};
constexpr int kMaxDcBatchGrids = 256;
// Pairs of scan scratch for nb block-sum pairs.
inline int64_t dc_scan_part_pairs(int64_t nb) { return nb / 1024 + 2; }
size_t dc_batch_item_bytes();
// amask / apre: per grid, the lookup tables written by the intersection pass
// (ItsOut::amask_tab / apre_tab; the active-cell hash is not built).
// levels != nullptr: no intersection set (in holds only ncells and
// cell_dense); each cell's edges, points and normals come from the grid's
// values at levels[i], with identical results.
hipError_t dc_run_batch(int n, const GridDesc *g, const int64_t *cell0, const uint64_t *const *amask,
                        const int32_t *const *apre, const float *levels, const ItsIn &in, float reg, float svd_tol,
                        float cond_fp64, DcScratch &w, const DcBatchDev &d, int64_t *nv, int64_t *nq,
                        int64_t &n_fp64, hipStream_t st);
// Backward of dual contouring: dL/d(output vertex) -> dL/d(grid values)
// (through the QEF solve and the edge interpolation of the intersection
// points) and, when grad_normals != nullptr, dL/d(normals).
hipError_t dc_backward(const GridDesc &g, const ItsIn &in, float reg, float svd_tol, float level, int64_t nv,
                       const int32_t *vcell, const float *grad_v, float *grad_values, float *grad_normals,
                       hipStream_t st);

}   // namespace iso
