// Device-side parameter blocks shared by the marching-cubes kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace iso {

struct CsgProgram;

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
    int64_t need;          // special capacity a call needs (segment overflow included)
    int64_t dup;           // two candidates share a position: 1, none: 0, not checked: -1
};

// The emit kernel appends special entries through nseg (<= kSpSeg)
// independent counters per list (workgroup b uses segment b % nseg, each
// counter on its own 256-byte line), so the atomics of different workgroups
// do not serialise on one address; a gather pass concatenates the segments
// into the contiguous SpecialLists.
constexpr int kSpSeg = 256;
static_assert((kSpSeg & (kSpSeg - 1)) == 0, "segment counts are powers of two");
constexpr int kSpStride = 32;
struct SpecialSegs {
    unsigned long long *ctr;   // [4][nseg * kSpStride]: cand, scell, kept, hole
    int64_t cap[4];            // entries per segment
    int nseg;
    SpecialVertex *cand;       // [nseg * cap[0]]
    SpecialEdge *scell;        // [nseg * cap[1]]
    int64_t *kept, *hole;      // [nseg * cap[2]], [nseg * cap[3]]
};
// Entries of each list per unit of special capacity: cand 1, scell 12,
// kept 5, hole 5.
inline int64_t sp_seg_cap(int64_t special_cap, int k, int nseg) {
    const int64_t mult = k == 1 ? 12 : (k >= 2 ? 5 : 1);
    return (special_cap * mult + nseg - 1) / nseg;
}

// Entries of (point row, z-segment) counted and scanned per workgroup.
constexpr int kMcScanBlock = 1024;

// Flat classify: 64-point chunks per wave task (plus one read past the end).
constexpr int kFlatChunks = 32;
constexpr int64_t kFlatPts = 64 * kFlatChunks;

struct McParams {
    const float *values;
    int64_t X, Y, Z;   // points per axis
    float amin[3], amax[3];
    float level;
    int method;        // 0 = nagae, 1 = lorensen
    int all_special;   // treat every vertex as a weld candidate
    float cand_eps;    // weld-candidate prefilter: t within cand_eps of 0 or 1
    int vec4;          // Z % 4 == 0 and values 16-byte aligned: rows load as float4
    int its;           // count pass: 0 marching cubes (vertices, faces), 1 intersections
                       // (active cells, crossing edges)
    int64_t W;         // 64-point chunks per point row, ceil(Z / 64)
    int nzc;           // chunks per z-segment (= emit tile depth)
    int64_t S;         // z-segments per point row, ceil(W / nzc)
    int64_t NI, NJ, NK;   // emit tiles per axis
    uint64_t *bits;    // (X*Y, W) inside masks: bit z%64 of word z/64
    uint32_t *cnt_v;   // (X*Y*S) vertex counts -> block-local exclusive offsets
    uint32_t *cnt_f;   // (X*Y*S) face counts   -> block-local exclusive offsets
    uint32_t *bsum_v, *bsum_f;   // per scan block totals
    int64_t *blk_v, *blk_f;      // per scan block global exclusive bases
    int64_t nblk;
    float *out_v;
    int64_t cap_v;
    int64_t *out_key;   // optional: owner point linear id * 3 + axis per vertex
    int32_t *out_f;
    int64_t cap_f;
    int *err;
    int64_t *totals;    // [0] vertices, [1] faces (provisional)
    SpecialLists *sp;   // contiguous lists (written by the gather pass)
    SpecialSegs seg;    // segmented lists (written by emit)
    // fused field (values unused when nsph > 0): v(p) = min_k |p - c_k| - r_k
    // over nsph <= kMaxFusedSpheres spheres (cx, cy, cz, r), evaluated where
    // needed with the sphere-union generator's formula
    const float4 *spheres;
    int nsph;
    // fused field given by an SDF program (device form, csg.h); values
    // unused when set
    const CsgProgram *csg;
    int csg_list;      // the classify's undecided-block list fits in cnt_v (see csg_list_fits)
    // nblk <= 256 (single grids): every emit workgroup scans the block totals
    // itself, and workgroup 0 writes blk_v / blk_f and the mesh size (no
    // separate scan launch)
    int scan_in_emit;
    // chunk index -> (row, chunk) -> (x, y) and chunk -> z-segment by
    // multiply-high, valid when narrow (X * Y * W < 2^32); set by
    // mc_params_divisors once W, Y and nzc are final
    int narrow;
    UDiv32 div_w, div_y, div_zc;
    // optional active-tile list (single-grid marching cubes): the count pass
    // appends each emit tile (TX x TY cell rows, segment tk) that has a face
    // once (tile_flag, zeroed by the reset kernel, marks it), and emit block b
    // < *tile_count takes tile tile_list[b]; the other blocks exit
    int *tile_flag;
    int *tile_list;
    int *tile_count;
    int tx, ty;
    int tile_base;   // id of the grid's first tile in the list's id space (a batch's tile ids)
    UDiv32 div_s, div_tx, div_ty;
    // read-once pipeline (k_stream, k_stream_scan*, k_stream_flush, k_faces):
    // regions of kStrRX x-slabs x kStrRY point rows x one 64-point word,
    // ordered (x block, y block, word) with the word fastest; a region's
    // four waves own 8 rows each
    int64_t NRX, NRY, NR;
    uint32_t *vofs;      // region-major entries: vertices of the wave before (point row, word)
    uint32_t *fofs;      // region-major entries: faces of the wave before (cell row, word)
    uint32_t *wcnt_v;    // (NR, 4): vertices of each wave
    uint32_t *wcnt_f;    // (NR, 4): faces of each wave
    int64_t *wb_v;       // (NR, 4): first vertex id of each wave (chunk-relative until the flush)
    int64_t *wb_f;       // (NR, 4): first face id of each wave (chunk-relative until the flush)
    // optional (single grids): open-addressing table of candidate indices by
    // position, dmask + 1 slots, cleared by the reset kernel
    int32_t *dslots;
    int64_t dmask;
    int64_t *alist;      // waves with vertices (k_stream_scan1), *acount of them
    int *acount;
    int64_t *csum_v, *csum_f;   // per chunk of kStrScanChunk waves: vertices / faces
    int64_t *cofs_v, *cofs_f;   // per chunk: first vertex / face id
    // vertex arena: k_stream writes each slab's vertices where a segment
    // counter puts them, k_stream_flush copies them to their final ids
    void *arena;         // kStrArenaSeg segments of arena_seg_cap entries (the face buffer, not yet used)
    int64_t arena_seg_cap;
    unsigned long long *arena_ctr;   // one counter per segment, on its own 256-byte line
    int32_t *arena_off;  // (NR, 4, kStrRX): arena entry of each slab's first vertex
    uint8_t *wover;      // (NR, 4): the wave's vertices did not fit its arena segment
};

// Read-once pipeline geometry.
constexpr int kStrRX = 16;        // x-slabs per region
constexpr int kStrRows = 8;       // point rows per wave
constexpr int kStrWaves = 4;
constexpr int kStrRY = kStrRows * kStrWaves;
constexpr int kStrArenaSeg = 256;  // arena segments (spread the reservation atomics)
constexpr int kStrScanChunk = 256; // waves per first-level scan block
// The read-once pipeline serves single grids whose emit tiles are 16 x 16
// cell rows x one word (tile configuration 6).
constexpr int kStrTileConfig = 6;
constexpr int kMaxFusedSpheres = 256;
// Classify blocks of a program field: 4 x 4 point rows x one 64-point word.
__host__ __device__ inline int64_t csg_classify_tasks(int64_t X, int64_t Y, int64_t W) {
    return ((X + 3) / 4) * ((Y + 3) / 4) * W;
}
// The undecided-block list fits: uint32 block ids in cnt_v, the count and
// the blocks' quarter states in cnt_f (X*Y*S entries each).
inline bool csg_list_fits(int64_t X, int64_t Y, int64_t W, int64_t S) {
    const int64_t nb = csg_classify_tasks(X, Y, W);
    return nb + 1 <= X * Y * S && nb <= (int64_t) UINT32_MAX;
}
constexpr int kCsgFillGroups = 2048;   // phase-2 workgroups
constexpr int kCsgFillWaves = 4;       // waves per phase-2 workgroup (they share a block's evaluations)
constexpr int kCsgFillBatchGroups = 4096;   // phase-2 workgroups of a batch (one wave each)
// Emit tile configuration of SDF-program fields, (8, 8, 1): small tiles, so
// that a tile's distinct corner points fit its LDS cache.
constexpr int kProgramTileConfig = 5;

// Fill McParams' invariant-divisor fields from X, Y, W and nzc.
void mc_params_divisors(McParams &p);

// Stage timings (ms) of the last timed launch: classify, count, scan, emit,
// and the whole extraction (first to last kernel).
struct McTimes {
    float classify, count, scan, emit, total;
};

// One grid of a ragged batch: its parameters (pointers into the batch
// workspace) and where its work starts in the batch-wide index spaces.
struct McBatchItem {
    McParams p;
    SpecialLists sp_init;   // contiguous special lists (counters zeroed by reset)
    int64_t ctask0;         // first flat classify task
    int64_t blk0;           // first count block
    int64_t tile0;          // first emit tile
    int64_t gtask0;         // first program classify task (items with an SDF program)
    int mode;               // classify loads: 0 Z%64 == 0 float4, 1 float4, 2 scalar
};

struct McBatchHdr {
    int64_t n, nctask, nblk, ntiles;   // items, flat classify tasks, count blocks, emit tiles
    int64_t ngtask = 0;                // program classify tasks (items with an SDF program)
    int n_prog = 0;                    // items with an SDF program
    // active-tile list of the batch (null: every tile runs): one flag per
    // batch tile and one list length per tile configuration
    int *tile_flag = nullptr;
    int *tile_count = nullptr;
};
constexpr int kMaxTileConfigs = 8;

// Per-item result of a batch phase, gathered for one read-back.
struct McBatchSummary {
    int64_t err, nv, nf;
    SpecialLists sp;   // the item's contiguous special lists (counts, need)
};

struct ItsOut;
// Intersections of a dense grid through the mask pipeline (p.its = 1):
// classify + count + scan (totals: active cells, intersection points), then
// the emit pass writing the reference's entry order.
hipError_t its_dense_count_launch(McParams &p, hipStream_t stream);
hipError_t its_dense_emit_launch(int cfg, McParams &p, const ItsOut &o, hipStream_t stream);

constexpr int kMaxBatch = 256;        // items per batch
// Tile configuration of batched intersection passes (16 x 16 cell rows x 2
// words): in a batch every grid's tiles share one launch, so the
// single-grid rule (enough tiles to fill the GPU) does not apply, and this
// configuration measured fastest over a mix of 32^3 to 512^3 grids.
constexpr int kBatchTileConfig = 2;
constexpr int kMaxSpSegBatch = 16;    // special-list segments per batch item
static_assert((kMaxSpSegBatch & (kMaxSpSegBatch - 1)) == 0, "segment counts are powers of two");

int mc_config_count();
int64_t mc_config_tiles(int cfg, const int64_t shape[3]);
int mc_auto_config(const int64_t shape[3]);
bool mc_config_geom(int cfg, int &tx, int &ty, int &nzc);
// Uploads the triangle tables once per process (a synchronous copy: call it
// before capturing launches into a graph).
hipError_t mc_upload_luts();
hipError_t mc_dense_launch(int cfg, McParams &p, const SpecialLists &sp_init, hipStream_t stream,
                           hipEvent_t *ev = nullptr);
// The read-once pipeline: reset, k_stream (masks, counts, vertices into the
// arena), k_stream_scan (wave bases), k_stream_flush (vertices to their
// ids), k_faces over the listed tiles, special gather.  Same outputs as
// mc_dense_launch up to the order of vertices and faces.
hipError_t mc_stream_launch(McParams &p, const SpecialLists &sp_init, hipStream_t stream, hipEvent_t *ev = nullptr);
// Batch phases: classify + count + scan of every item, then emit + gather.
// Both end with a kernel that writes one McBatchSummary per item.
hipError_t mc_batch_count_launch(const McBatchItem *items, const McBatchHdr *hdr, const McBatchHdr &h,
                                 McBatchSummary *summary, hipStream_t stream);
// Intersections of a batch of dense grids (items counted with p.its = 1):
// one launch per tile configuration, outs[i] the grid's output arrays.
hipError_t its_batch_emit_launch(const int64_t *tile_begin, const int64_t *tile_end, const McBatchItem *items,
                                 const int32_t *order, const ItsOut *outs, const McBatchHdr *hdr, hipStream_t stream,
                                 const int *tile_list = nullptr, const int *tile_count = nullptr);

// Items may use different tile configurations: configuration c owns the
// emit tiles [tile_begin[c], tile_end[c]) (one launch each); order lists the
// items by first tile.
// tile_list / tile_count: the batch's active-tile list (configuration c's
// entries start at tile_begin[c], its length is tile_count[c]); null: every
// tile runs.
hipError_t mc_batch_emit_launch(const int64_t *tile_begin, const int64_t *tile_end, const McBatchItem *items,
                                const int32_t *order, const McBatchHdr *hdr, const McBatchHdr &h,
                                McBatchSummary *summary, hipStream_t stream, const int *tile_list = nullptr,
                                const int *tile_count = nullptr);

}   // namespace iso
