// Dense marching cubes: classify -> count -> scan -> emit.
//
// Reference path replaced: mc::marching_cubes on a UniformGrid
// (src/mc/mc.cu:17-68 -> get_case_num_op include/utils.cuh:82-102,
// compaction src/mc/mc.cu:33-42, Nagae/Lorensen::process_cube_op
// src/mc/nagae.cu:34-80 / src/mc/lorensen.cu:34-80, vertex_welding
// src/utils.cu:32-59).
//
// Design (gfx950, wave64):
//   * classify streams the value grid once with dwordx4 loads.  16 lanes
//     turn 64 consecutive z-values of a point row into one 64-bit inside mask
//     (4-bit nibbles merged with three in-group shuffles); only the masks,
//     1 bit per point, are written.
//   * Vertices are numbered in point-major order: a vertex belongs to the
//     lower endpoint of its grid edge (its owner point) and vertices are
//     ordered by (owner x, y, z, axis x < y < z).  Faces are ordered by
//     (cell x, y, z, LUT slot).  Both numberings are a prefix sum over
//     (point row, z-segment) entries plus a popcount inside the segment, so
//     every vertex id is a pure function of the masks: welding without
//     hashing, sorting or any wait between workgroups.
//   * count derives per-entry vertex and face counts from the masks and
//     scans them per workgroup; scan_top turns the block totals into global
//     bases and the mesh size.
//   * emit tiles the cells (TX x TY cell rows x one z-segment of NZC 64-cell
//     chunks), stages the mask patch in LDS, rebuilds per-chunk vertex and
//     face bases, compacts the active cells through an LDS scan and gives
//     each one lane: it gathers the 8 corner values, interpolates with the
//     reference's float op order and writes the vertices its cell owns and
//     its faces.  Output is deterministic.
//   * Vertices whose position could coincide bitwise with another edge's
//     vertex (within 16 ulps of a grid corner) are recorded, together with the
//     faces of their cells, for the exact position-based weld of the fix-up
//     pass (the reference welds by position equality and drops degenerate
//     triangles, src/utils.cu:49-55, src/mc/nagae.cu:71).

#include <atomic>
#include <mutex>
#include <type_traits>
#include <utility>

#include "common.h"
#include "its_dc.h"
#include "mc_dense.h"
#include "csg.h"
#include "mc_luts.h"

namespace iso {

static __constant__ uint64_t cTris[2][256];
static __constant__ uint8_t cNtri[2][256];   // triangles per case (the low 4 bits of cTris)


// Append to special list k (0 cand, 1 scell, 2 kept, 3 hole) in segment
// (seg + thread) mod nseg: a workgroup's entries spread over all segments,
// so a grid with few workgroups needs no more capacity than its entries.
// Returns the entry index, or -1 when the segment is full (the counter
// still counts so the host can size a retry).
__device__ __forceinline__ int64_t seg_push(const McParams &p, int k, int seg) {
    seg = (seg + (int) threadIdx.x) & (p.seg.nseg - 1);   // nseg: a power of two
    const unsigned long long i = atomicAdd(&p.seg.ctr[(k * p.seg.nseg + seg) * kSpStride], 1ull);
    return (int64_t) i < p.seg.cap[k] ? (int64_t) seg * p.seg.cap[k] + (int64_t) i : -1;
}

__device__ __forceinline__ void set_err(const McParams &p, int code) { atomicOr(p.err, code); }

// Clears what an extraction's later kernels count into (special-list
// counters, tile flags, the duplicate-check slots, the header); thread t of
// a grid of `stride` threads.
__device__ __forceinline__ void reset_state(const McParams &p, SpecialLists init, int64_t t, int64_t stride) {
    for (int64_t i = t; i < 4 * (int64_t) p.seg.nseg; i += stride) p.seg.ctr[i * kSpStride] = 0;
    if (p.tile_flag) {
        const int64_t ntiles = p.NI * p.NJ * p.NK;
        for (int64_t i = t; i < ntiles; i += stride) p.tile_flag[i] = 0;
    }
    if (p.dslots)
        for (int64_t i = t; i <= p.dmask; i += stride) p.dslots[i] = -1;
    if (t == 0) {
        if (p.tile_count) *p.tile_count = 0;
        if (p.csg) p.cnt_f[0] = 0;   // the program classify's block list count
        *p.err = 0;
        p.totals[0] = p.totals[1] = 0;
        init.n_cand = init.n_scell = init.n_kept = init.n_hole = 0;
        init.need = 0;
        init.dup = p.dslots ? 0 : -1;
        if (p.sp) *p.sp = init;   // the intersection pipeline keeps no special lists
    }
}

// Bits [0, n), n clamped to [0, 64].
__device__ __forceinline__ uint64_t low_bits(int64_t n) {
    return n >= 64 ? ~0ull : (n <= 0 ? 0ull : ((1ull << n) - 1));
}

// Position of the r-th set bit (r < popcount(m)).
__device__ __forceinline__ int nth_set_bit(uint64_t m, int r) {
    int pos = 0;
    uint32_t x = (uint32_t) m;
    int c = __popc(x);
    if (r >= c) {
        r -= c;
        x = (uint32_t) (m >> 32);
        pos = 32;
    }
#pragma unroll
    for (int half = 16; half >= 1; half >>= 1) {
        c = __popc(x & ((1u << half) - 1));
        if (r >= c) {
            r -= c;
            x >>= half;
            pos += half;
        }
    }
    return pos;
}

// Vertices owned by the points of one 64-point chunk of a point row: b the
// chunk's mask, bn0 bit 0 of the next chunk, bx / by the masks of the rows at
// x+1 / y+1 (0 when absent), zlo the chunk's first z.
__device__ __forceinline__ int chunk_vertices(uint64_t b, uint64_t bn0, uint64_t bx, uint64_t by, bool hasx,
                                              bool hasy, int64_t zlo, int64_t Z) {
    const uint64_t pv = low_bits(Z - zlo), ev = low_bits(Z - 1 - zlo);
    const uint64_t zs = (b >> 1) | (bn0 << 63);
    int n = __popcll((b ^ zs) & ev);
    if (hasx) n += __popcll((b ^ bx) & pv);
    if (hasy) n += __popcll((b ^ by) & pv);
    return n;
}

// Corner masks of the 64 cells of a chunk of a cell row: m[k] bit j = corner
// k of cell z = zlo + j is inside.  Corner k sits at (k>>2, (k>>1)&1, k&1)
// (include/utils.cuh:32-60).
struct CellMasks {
    uint64_t m[8];
};

__device__ __forceinline__ CellMasks cell_masks(uint64_t a, uint64_t an, uint64_t b, uint64_t bn, uint64_t d0,
                                                uint64_t d0n, uint64_t d1, uint64_t d1n) {
    CellMasks c;
    c.m[0] = a;
    c.m[1] = (a >> 1) | ((an & 1) << 63);
    c.m[2] = b;
    c.m[3] = (b >> 1) | ((bn & 1) << 63);
    c.m[4] = d0;
    c.m[5] = (d0 >> 1) | ((d0n & 1) << 63);
    c.m[6] = d1;
    c.m[7] = (d1 >> 1) | ((d1n & 1) << 63);
    return c;
}

__device__ __forceinline__ uint64_t cell_active(const CellMasks &c) {
    uint64_t o = 0, a = ~0ull;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        o |= c.m[k];
        a &= c.m[k];
    }
    return o & ~a;
}

__device__ __forceinline__ int cell_case(const CellMasks &c, int j) {
    int cs = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) cs |= (int) ((c.m[k] >> j) & 1) << k;
    return cs;
}

// Per-cell interpolation context: corner values and the two grid
// coordinates of the cell along each axis.  Only ever indexed with
// compile-time constants (dynamic indexing would put it in scratch).
struct CellCtx {
    float v[8];
    float c[3][2];
    float level;
};

// Calls f(std::integral_constant<int, E>) for E = 0..11 (unrolled).
template <typename F, int... E>
__device__ __forceinline__ void for_edges_impl(F &&f, std::integer_sequence<int, E...>) {
    (f(std::integral_constant<int, E>{}), ...);
}
template <typename F>
__device__ __forceinline__ void for_edges(F &&f) {
    for_edges_impl(f, std::make_integer_sequence<int, 12>{});
}

__device__ __forceinline__ int crossing_status(int cs) {
    int st = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (((cs >> edge_a(e)) & 1) != ((cs >> edge_b(e)) & 1)) st |= 1 << e;
    return st;
}

// Vertex position on crossing edge E, with the reference's op order.
template <int E>
__device__ __forceinline__ void edge_pos_c(const CellCtx &cc, float &x, float &y, float &z) {
    constexpr int ca = edge_a(E), cb = edge_b(E);
    const float tt = edge_t(cc.v[ca], cc.v[cb], cc.level);
    x = lerp1(tt, cc.c[0][ca >> 2], cc.c[0][cb >> 2]);
    y = lerp1(tt, cc.c[1][(ca >> 1) & 1], cc.c[1][(cb >> 1) & 1]);
    z = lerp1(tt, cc.c[2][ca & 1], cc.c[2][cb & 1]);
}

// Same, for a per-lane edge index (only used in the rare special cells).
__device__ __forceinline__ void edge_pos_dyn(const CellCtx &cc, int e, float &x, float &y, float &z) {
    switch (e) {
    case 0: edge_pos_c<0>(cc, x, y, z); break;
    case 1: edge_pos_c<1>(cc, x, y, z); break;
    case 2: edge_pos_c<2>(cc, x, y, z); break;
    case 3: edge_pos_c<3>(cc, x, y, z); break;
    case 4: edge_pos_c<4>(cc, x, y, z); break;
    case 5: edge_pos_c<5>(cc, x, y, z); break;
    case 6: edge_pos_c<6>(cc, x, y, z); break;
    case 7: edge_pos_c<7>(cc, x, y, z); break;
    case 8: edge_pos_c<8>(cc, x, y, z); break;
    case 9: edge_pos_c<9>(cc, x, y, z); break;
    case 10: edge_pos_c<10>(cc, x, y, z); break;
    default: edge_pos_c<11>(cc, x, y, z); break;
    }
}

template <int E>
__device__ __forceinline__ bool edge_is_candidate(const CellCtx &cc) {
    constexpr int ca = edge_a(E), cb = edge_b(E), axis = edge_axis(E);
    constexpr int ia = axis == 0 ? (ca >> 2) : (axis == 1 ? ((ca >> 1) & 1) : (ca & 1));
    constexpr int ib = axis == 0 ? (cb >> 2) : (axis == 1 ? ((cb >> 1) & 1) : (cb & 1));
    const float tt = edge_t(cc.v[ca], cc.v[cb], cc.level);
    const float a = cc.c[axis][ia], b = cc.c[axis][ib];
    return near_endpoint(lerp1(tt, a, b), a, b);
}

// Global key of edge e of cell (gx, gy, gz): owner point linear id * 3 + axis.
__device__ __forceinline__ int64_t edge_key(const McParams &p, int64_t gx, int64_t gy, int64_t gz, int e) {
    int ca = edge_a(e);
    int64_t ox = gx + (ca >> 2), oy = gy + ((ca >> 1) & 1), oz = gz + (ca & 1);
    return ((ox * p.Y + oy) * p.Z + oz) * 3 + edge_axis(e);
}

// ---- classify ------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

// k_classify_v4 (Z % 64 == 0, float4-aligned rows): word w covers values
// [64w, 64w + 64); 16 lanes load it with dwordx4 loads and merge their
// nibbles with three in-group shuffles.
// RESET: the grid first clears the extraction's state (reset_state), which
// nothing in classify reads, so the stored-grid pipeline needs no separate
// reset launch.
template <bool RESET>
__global__ void __launch_bounds__(256) k_classify_v4(const float *__restrict__ values, int64_t nwords, float level,
                                                     uint64_t *__restrict__ bits, McParams rp, SpecialLists init) {
    if constexpr (RESET)
        reset_state(rp, init, (int64_t) blockIdx.x * blockDim.x + threadIdx.x, (int64_t) gridDim.x * blockDim.x);
    constexpr int U = 4;   // words in flight per 16-lane group
    const int lane = threadIdx.x & 63;
    const int grp = lane >> 4, q = lane & 15;
    const int64_t wave = (int64_t) blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nwave = (int64_t) gridDim.x * 4;
    for (int64_t base = wave * (4 * U); base < nwords; base += nwave * (4 * U)) {
        float v[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t w = base + 4 * u + grp;
            v[u][0] = v[u][1] = v[u][2] = v[u][3] = 0.0f;
            if (w < nwords) {
                // streamed once: nontemporal loads keep the values out of L2's working set
                const f32x4 f = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(values + 64 * w + 4 * q));
                v[u][0] = f.x;
                v[u][1] = f.y;
                v[u][2] = f.z;
                v[u][3] = f.w;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint32_t x = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) x |= (uint32_t) inside(v[u][j], level) << j;
            x |= __shfl_down(x, 1, 16) << 4;
            x |= __shfl_down(x, 2, 16) << 8;
            x |= __shfl_down(x, 4, 16) << 16;
            const uint32_t hi = __shfl_down(x, 8, 16);
            const int64_t w = base + 4 * u + grp;
            if (q == 0 && w < nwords) bits[w] = (uint64_t) x | ((uint64_t) hi << 32);
        }
    }
}


// Flat classify (any row length): task t covers the grid's flat points
// [t * kFlatPts, (t + 1) * kFlatPts).  A wave loads kFlatChunks + 1
// consecutive 64-point chunks (coalesced 256-byte loads, no idle lanes
// whatever Z is), their compare ballots go to LDS, and each lane assembles
// one (row, word) mask of a word that starts inside the task: 64 bits of the
// flat bit stream at the word's start (a funnel shift of two ballots), cut
// at the row's end.  The extra chunk covers words that run past the task.
__device__ __forceinline__ int64_t words_before(int64_t f, int64_t Z, int64_t W) {
    const int64_t r = div_fast(f, Z);
    return r * W + min<int64_t>(W, (f - r * Z + 63) >> 6);
}

__device__ __forceinline__ void classify_flat_task(const float *__restrict__ values, int64_t N, int64_t Z,
                                                   int64_t W, float level, int64_t t, uint64_t *__restrict__ bits,
                                                   uint64_t *ballots) {
    constexpr int K = kFlatChunks + 1;
    const int lane = threadIdx.x & 63;
    const int64_t f0 = t * kFlatPts;
    float v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int64_t f = f0 + 64 * k + lane;
        v[k] = __builtin_nontemporal_load(values + (f < N ? f : 0));   // past the end: a valid dummy
    }
    uint64_t mine = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint64_t m = __ballot(f0 + 64 * k + lane < N && inside(v[k], level));
        if (lane == k) mine = m;
    }
    if (lane < K) ballots[lane] = mine;
    __builtin_amdgcn_wave_barrier();
    const int64_t g0 = words_before(f0, Z, W), g1 = words_before(min(f0 + kFlatPts, N), Z, W);
    for (int64_t g = g0 + lane; g < g1; g += 64) {
        const int64_t row = div_fast(g, W), c = g - row * W;
        const int o = (int) (row * Z + 64 * c - f0), k = o >> 6, sh = o & 63;
        uint64_t w = ballots[k] >> sh;
        if (sh) w |= ballots[k + 1] << (64 - sh);
        const int64_t len = Z - 64 * c;
        if (len < 64) w &= (1ull << len) - 1ull;
        bits[g] = w;
    }
    __builtin_amdgcn_wave_barrier();
}

template <bool RESET>
__global__ void __launch_bounds__(256) k_classify(const float *__restrict__ values, int64_t N, int64_t Z, int64_t W,
                                                  float level, uint64_t *__restrict__ bits, McParams rp,
                                                  SpecialLists init) {
    if constexpr (RESET)
        reset_state(rp, init, (int64_t) blockIdx.x * blockDim.x + threadIdx.x, (int64_t) gridDim.x * blockDim.x);
    __shared__ uint64_t ballots[4][kFlatChunks + 1];
    const int64_t wave = __builtin_amdgcn_readfirstlane((int) (blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int64_t nwave = (int64_t) gridDim.x * 4;
    const int64_t ntask = (N + kFlatPts - 1) / kFlatPts;
    for (int64_t t = wave; t < ntask; t += nwave)
        classify_flat_task(values, N, Z, W, level, t, bits, ballots[threadIdx.x >> 6]);
}

// ---- count + block scan ----------------------------------------------------------
// Entry i = (point row r, z-segment s), r = x * Y + y.  Vertex count: edges
// owned by the row's points in the segment; face count: triangles of the
// cell row (x, y) in the segment (0 when x = X-1 or y = Y-1).
template <int NT>
__device__ __forceinline__ void block_exscan2(uint32_t &a, uint32_t &b, uint32_t &tot_a, uint32_t &tot_b,
                                              uint32_t *sa, uint32_t *sb) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t ia = (uint32_t) wave_incl_sum((int) a), ib = (uint32_t) wave_incl_sum((int) b);
    if (lane == 63) {
        sa[wave] = ia;
        sb[wave] = ib;
    }
    __syncthreads();
    uint32_t pa = 0, pb = 0;
    tot_a = tot_b = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wave) {
            pa += sa[w];
            pb += sb[w];
        }
        tot_a += sa[w];
        tot_b += sb[w];
    }
    a = pa + ia - a;
    b = pb + ib - b;
}

// First chunk (point row r, chunk c) of entry i, as the global chunk index r * W + c.
__device__ __forceinline__ int64_t entry_chunk(const McParams &p, int64_t i) {
    const int64_t r = div_fast(i, p.S), s = i - r * p.S;
    return r * p.W + s * p.nzc;
}

// One workgroup per kMcScanBlock entries.  Lanes walk the block's chunks in
// memory order (coalesced mask loads), add their counts into the chunk's
// entry in LDS, then the entries are scanned.  NT = 1024 for grids with few
// blocks (more lanes in flight per block), 256 otherwise.
template <int NT>
__device__ __forceinline__ void count_block(const McParams &p, int64_t blk) {
    constexpr int PER = kMcScanBlock / NT;
    __shared__ uint32_t ent_v[kMcScanBlock], ent_f[kMcScanBlock];
    __shared__ uint32_t sa[NT / 64], sb[NT / 64];
    __shared__ uint8_t ntri[256];
    const int tid = threadIdx.x;
    for (int c = tid; c < 256; c += NT)
        ntri[c] = p.its ? (uint8_t) __popc(crossing_status(c)) : (uint8_t) (cTris[p.method][c] & 15);
    const int64_t n = p.X * p.Y * p.S;
    const int64_t i0 = blk * kMcScanBlock;
    const int64_t i1 = min(i0 + (int64_t) kMcScanBlock, n);
    for (int k = tid; k < kMcScanBlock; k += NT) ent_v[k] = ent_f[k] = 0;
    // last tile seen per slot (tile id mod 256) by this workgroup: repeats of
    // a tile skip the global flag
    __shared__ int tseen[256];
    if (p.tile_list)
        for (int k = tid; k < 256; k += NT) tseen[k] = -1;
    __syncthreads();
    const int64_t g0 = entry_chunk(p, i0);
    const int64_t g1 = i1 < n ? entry_chunk(p, i1) : p.X * p.Y * p.W;
    const int64_t YW = p.Y * p.W;
    const int lane = tid & 63, wave = tid >> 6;
    __shared__ int32_t mark[NT / 64][64];
    const bool narrow = p.narrow;
    for (int64_t gb = g0; gb < g1; gb += NT) {   // every lane runs every round (wave-wide steps below)
        const int64_t g = gb + tid;
        const bool valid = g < g1;
        const int64_t gg = valid ? g : g0;
        int64_t r, c, x, y, seg;
        if (narrow) {
            const uint32_t g32 = (uint32_t) gg, r32 = p.div_w.div(g32), c32 = g32 - r32 * (uint32_t) p.W,
                           x32 = p.div_y.div(r32);
            r = r32;
            c = c32;
            x = x32;
            y = r32 - x32 * (uint32_t) p.Y;
            seg = p.div_zc.div(c32);
        } else {
            r = div_fast(gg, p.W);
            c = gg - r * p.W;
            x = div_fast(r, p.Y);
            y = r - x * p.Y;
            seg = div_fast(c, p.nzc);
        }
        const bool hasx = x + 1 < p.X, hasy = y + 1 < p.Y, hn = c + 1 < p.W;
        const int64_t zlo = 64 * c;
        // all eight words are loaded before any is used: one memory round
        // trip per round instead of two
        const bool hxy = hasx && hasy, hxyn = hxy && hn;
        const uint64_t *q = p.bits + gg;
        const uint64_t b = q[0], bn = hn ? q[1] : 0ull;
        const uint64_t bx = hasx ? q[YW] : 0ull, by = hasy ? q[p.W] : 0ull;
        const uint64_t bxy = hxy ? q[YW + p.W] : 0ull;
        const uint64_t byn = hxyn ? q[p.W + 1] : 0ull, bxn = hxyn ? q[YW + 1] : 0ull,
                       bxyn = hxyn ? q[YW + p.W + 1] : 0ull;
        uint32_t cv = p.its ? 0u : (uint32_t) chunk_vertices(b, bn & 1, bx, by, hasx, hasy, zlo, p.Z);
        uint32_t cf = 0;
        const CellMasks m = cell_masks(b, bn, by, byn, bx, bxn, bxy, bxyn);
        uint64_t act = hxy ? cell_active(m) & low_bits(p.Z - 1 - zlo) : 0ull;
        // bit 0 of the next word of rows (0,0), (0,1), (1,0), (1,1)
        const uint32_t nxt = (uint32_t) (bn & 1) | (uint32_t) (byn & 1) << 1 | (uint32_t) (bxn & 1) << 2 |
                             (uint32_t) (bxyn & 1) << 3;
        if (hxy) {
            if (p.its) {
                // crossing edges of the active cells, bit-parallel over the 64
                // cells: edge e crosses where its two corner masks differ
                cv = (uint32_t) __popcll(act);
                for_edges([&](auto E) { cf += (uint32_t) __popcll((m.m[edge_a(E)] ^ m.m[edge_b(E)]) & act); });
            }
        }
        if (!valid) act = 0;
        const int e = (int) (r * p.S + seg - i0);
        if (valid && cv) atomicAdd(&ent_v[e], cv);
        if (valid && cf) atomicAdd(&ent_f[e], cf);
        if (p.its) continue;
        // triangles: the wave's active cells are dealt out one per lane per
        // round (owner lane by a running maximum over marked first slots),
        // so a word with many active cells does not serialise its wave
        const int cnt = __popcll(act);
        const int incl = wave_incl_sum(cnt);
        const int total = __shfl(incl, 63), excl = incl - cnt;
        for (int base = 0; base < total; base += 64) {
            mark[wave][lane] = -1;
            __builtin_amdgcn_wave_barrier();
            if (cnt > 0 && excl < base + 64 && incl > base) mark[wave][max(excl, base) - base] = lane;
            __builtin_amdgcn_wave_barrier();
            const int own = max(wave_incl_max(mark[wave][lane]), 0);
            const int k = base + lane;
            const uint64_t oact = __shfl(act, own);
            const uint64_t w00 = __shfl(b, own), w01 = __shfl(by, own), w10 = __shfl(bx, own), w11 = __shfl(bxy, own);
            const uint32_t on = (uint32_t) __shfl((int) nxt, own);
            const int oexcl = __shfl(excl, own), oe = __shfl(e, own);
            if (k < total) {
                const int j = nth_set_bit(oact, k - oexcl);
                uint32_t up;   // bit j + 1 of the four rows
                if (j < 63)
                    up = (uint32_t) ((w00 >> (j + 1)) & 1) | (uint32_t) ((w01 >> (j + 1)) & 1) << 1 |
                         (uint32_t) ((w10 >> (j + 1)) & 1) << 2 | (uint32_t) ((w11 >> (j + 1)) & 1) << 3;
                else
                    up = on;
                const int cs = (int) (((w00 >> j) & 1) | ((up & 1) << 1) | (((w01 >> j) & 1) << 2) |
                                      (((up >> 1) & 1) << 3) | (((w10 >> j) & 1) << 4) | (((up >> 2) & 1) << 5) |
                                      (((w11 >> j) & 1) << 6) | (((up >> 3) & 1) << 7));
                atomicAdd(&ent_f[oe], (uint32_t) ntri[cs]);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    uint32_t cv[PER], cf[PER], tv = 0, tf = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        cv[e] = ent_v[tid * PER + e];
        cf[e] = ent_f[tid * PER + e];
        tv += cv[e];
        tf += cf[e];
    }
    uint32_t bv, bf;
    block_exscan2<NT>(tv, tf, bv, bf, sa, sb);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int64_t i = i0 + tid * PER + e;
        if (i < n) {
            p.cnt_v[i] = tv;
            p.cnt_f[i] = tf;
            // tiles with a face (intersection mode: with an active cell); the
            // list exists only with narrow indices (decided on the host)
            if (p.tile_list && (p.its ? cv[e] : cf[e])) {
                const uint32_t r = p.div_s.div((uint32_t) i), sg = (uint32_t) i - r * (uint32_t) p.S,
                               x = p.div_y.div(r), y = r - x * (uint32_t) p.Y;
                const int t = p.tile_base +
                              (int) ((p.div_tx.div(x) * (uint32_t) p.NJ + p.div_ty.div(y)) * (uint32_t) p.NK + sg);
                if (atomicExch(&tseen[t & 255], t) != t && atomicCAS(&p.tile_flag[t], 0, 1) == 0)
                    p.tile_list[atomicAdd(p.tile_count, 1)] = t;
            }
        }
        tv += cv[e];
        tf += cf[e];
    }
    if (tid == 0) {
        p.bsum_v[blk] = bv;
        p.bsum_f[blk] = bf;
    }
}

template <int NT>
__global__ void __launch_bounds__(NT) k_count(McParams p) { count_block<NT>(p, blockIdx.x); }

// Blocks below which the count pass uses 1024-lane workgroups.
constexpr int64_t kCountWideBelow = 2048;

// Exclusive scan of the block totals by one workgroup: thread t sums a
// contiguous run of ceil(nblk / 1024) totals, one block-wide scan of the run
// sums, then each thread writes its run's prefixes.  Writes the mesh size and
// flags meshes beyond int32 indices.
__device__ __forceinline__ void scan_top(const McParams &p) {
    constexpr int NT = 1024, NW = NT / 64;
    __shared__ int64_t sv[NW], sf[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t per = (p.nblk + NT - 1) / NT;
    const int64_t lo = min<int64_t>((int64_t) tid * per, p.nblk), hi = min<int64_t>(lo + per, p.nblk);
    int64_t v = 0, f = 0;
#pragma unroll 4
    for (int64_t i = lo; i < hi; ++i) {
        v += (int64_t) p.bsum_v[i];
        f += (int64_t) p.bsum_f[i];
    }
    int64_t iv = v, iff = f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t yv = __shfl_up(iv, o), yf = __shfl_up(iff, o);
        if (lane >= o) {
            iv += yv;
            iff += yf;
        }
    }
    if (lane == 63) {
        sv[wave] = iv;
        sf[wave] = iff;
    }
    __syncthreads();
    int64_t rv = iv - v, rf = iff - f, tv = 0, tf = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wave) {
            rv += sv[w];
            rf += sf[w];
        }
        tv += sv[w];
        tf += sf[w];
    }
#pragma unroll 4
    for (int64_t i = lo; i < hi; ++i) {
        const int64_t a = (int64_t) p.bsum_v[i], b = (int64_t) p.bsum_f[i];
        p.blk_v[i] = rv;
        p.blk_f[i] = rf;
        rv += a;
        rf += b;
    }
    if (tid == 0) {
        p.totals[0] = tv;
        p.totals[1] = tf;
        if (tv > 0x7fffffffll || tf > 0x7fffffffll) set_err(p, ISO_ERR_TOO_LARGE);
    }
    __syncthreads();   // sv / sf are reused by the next call in the same workgroup
}

__global__ void __launch_bounds__(1024) k_scan_top(McParams p) { scan_top(p); }

// ---- fused sphere-union field ------------------------------------------------------
// v(p) = min_k (sqrt((dx*dx + dy*dy) + dz*dz) - r_k): the sphere-union
// generator's formula and op order (sdf.hip), evaluated where the pipeline
// needs it instead of read from a grid.  Classify needs only sign bits
// (v - level < 0 iff some sphere has d_k - level < 0, since inside() is
// monotone in v), emit the 8 corners of its active cells.  Spheres that
// cannot matter inside a box of grid points are culled with bounds computed
// in FP64 from the box's grid coordinates and a margin (1e-5 of the
// distances involved) far above the FP32 evaluation error, so every value
// evaluated is bit-identical to the generated grid's (fminf is exact and
// order-independent).

__device__ __forceinline__ float sphere_d(float px, float py, float pz, float4 sp) {
    const float dx = px - sp.x, dy = py - sp.y, dz = pz - sp.z;
    float d2 = dx * dx + dy * dy;
    d2 = d2 + dz * dz;
    return sqrtf(d2) - sp.w;
}

// World box of grid points [i0, i1] per axis (grid_coord is monotone).
struct Box3 {
    double lo[3], hi[3];
};
__device__ __forceinline__ Box3 point_box(const McParams &p, int64_t x0, int64_t x1, int64_t y0, int64_t y1,
                                          int64_t z0, int64_t z1) {
    const int64_t a0[3] = {x0, y0, z0}, a1[3] = {x1, y1, z1}, n[3] = {p.X, p.Y, p.Z};
    Box3 b;
    for (int a = 0; a < 3; ++a) {
        const double u = grid_coord(a0[a], n[a], p.amin[a], p.amax[a]);
        const double v = grid_coord(a1[a], n[a], p.amin[a], p.amax[a]);
        b.lo[a] = fmin(u, v);
        b.hi[a] = fmax(u, v);
    }
    return b;
}

// Lower / upper bound of |p - c| - r over the box, and the culling margin.
__device__ __forceinline__ void sphere_bounds(const Box3 &b, float4 sp, double &lb, double &ub, double &m) {
    const double c[3] = {sp.x, sp.y, sp.z};
    double dn = 0.0, dx = 0.0;
    for (int a = 0; a < 3; ++a) {
        const double g = fmax(fmax(b.lo[a] - c[a], c[a] - b.hi[a]), 0.0);
        const double f = fmax(fabs(c[a] - b.lo[a]), fabs(c[a] - b.hi[a]));
        dn += g * g;
        dx += f * f;
    }
    dx = sqrt(dx);
    lb = sqrt(dn) - (double) sp.w;
    ub = dx - (double) sp.w;
    m = 1e-5 * (dx + fabs((double) sp.w)) + 1e-30;
}

// Classify-side bounds in FP32 without square roots: whether sphere sp can
// reach below `level` somewhere in the box (keep) and whether it is below
// it on the whole box (all_in).  Distances are compared squared; the margin
// is 1e-5 of an L1 bound of the largest distance (>= the Euclidean one),
// far above the rounding of the squared sums (a few ulps).
__device__ __forceinline__ void sphere_sign_bounds(const float lo[3], const float hi[3], float4 sp, float level,
                                                   bool &keep, bool &all_in) {
    const float c[3] = {sp.x, sp.y, sp.z};
    float dn = 0.0f, dx = 0.0f, l1 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float g = fmaxf(fmaxf(lo[a] - c[a], c[a] - hi[a]), 0.0f);
        const float f = fmaxf(fabsf(c[a] - lo[a]), fabsf(c[a] - hi[a]));
        dn += g * g;
        dx += f * f;
        l1 += f;
    }
    const float m = 1e-5f * (l1 + fabsf(sp.w)) + 1e-30f;
    const float tk = level + sp.w + m;   // keep iff min distance < tk
    keep = tk > 0.0f && dn < tk * tk;
    const float ta = level + sp.w - m;   // all inside iff max distance < ta
    all_in = ta > 0.0f && dx < ta * ta;
}

// Inside masks of the fused field: a wave task is a 4 x 4 block of point
// rows times one 64-point word.  Once per task (one sphere per lane,
// ballots): if some sphere is below the level on the whole box, every point
// is inside; if none reaches below it anywhere, none is; otherwise each
// row's word is the ballot of "inside one of the spheres that reach below".
__global__ void __launch_bounds__(256) k_classify_spheres(McParams p) {
    __shared__ float4 sph[kMaxFusedSpheres];
    for (int i = threadIdx.x; i < p.nsph; i += 256) sph[i] = p.spheres[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t wave = __builtin_amdgcn_readfirstlane((int) (blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int64_t nwave = (int64_t) gridDim.x * 4;
    const int64_t nbx = (p.X + 3) / 4, nby = (p.Y + 3) / 4, ntask = nbx * nby * p.W;
    constexpr int NCH = kMaxFusedSpheres / 64;
    for (int64_t t = wave; t < ntask; t += nwave) {
        const int64_t c = t % p.W, r = t / p.W, yb = r % nby, xb = r / nby;
        const int64_t xa = 4 * xb, ya = 4 * yb, za = 64 * c;
        const int nxr = (int) min<int64_t>(4, p.X - xa), nyr = (int) min<int64_t>(4, p.Y - ya);
        const int nzp = (int) min<int64_t>(64, p.Z - za);
        // grid coordinates: lanes 0-3 the block's x rows, 4-7 its y rows, every lane its z
        const float cxy = lane < 4 ? grid_coord(xa + min(lane, nxr - 1), p.X, p.amin[0], p.amax[0])
                                   : grid_coord(ya + min(max(lane - 4, 0), nyr - 1), p.Y, p.amin[1], p.amax[1]);
        const int64_t z = za + lane;
        const bool zok = z < p.Z;
        const float pz = grid_coord(zok ? z : p.Z - 1, p.Z, p.amin[2], p.amax[2]);
        float blo[3], bhi[3];
        {
            const float x0c = __shfl(cxy, 0), x1c = __shfl(cxy, nxr - 1);
            const float y0c = __shfl(cxy, 4), y1c = __shfl(cxy, 4 + nyr - 1);
            const float z0c = __shfl(pz, 0), z1c = __shfl(pz, nzp - 1);
            blo[0] = fminf(x0c, x1c);
            bhi[0] = fmaxf(x0c, x1c);
            blo[1] = fminf(y0c, y1c);
            bhi[1] = fmaxf(y0c, y1c);
            blo[2] = fminf(z0c, z1c);
            bhi[2] = fmaxf(z0c, z1c);
        }
        uint64_t rel[NCH], any_rel = 0, full = 0;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
            const int k = 64 * ch + lane;
            bool keep = false, all_in = false;
            if (k < p.nsph) sphere_sign_bounds(blo, bhi, sph[k], p.level, keep, all_in);
            rel[ch] = __ballot(keep);
            any_rel |= rel[ch];
            full |= __ballot(all_in);
        }
        if (full || !any_rel) {   // the whole block is inside / outside
            if (lane < 16) {
                const int64_t x = xa + (lane >> 2), y = ya + (lane & 3);
                if (x < p.X && y < p.Y) p.bits[(x * p.Y + y) * p.W + c] = full ? low_bits(p.Z - za) : 0ull;
            }
            continue;
        }
        uint64_t mine = 0;
        for (int i = 0; i < 16; ++i) {
            const int64_t x = xa + (i >> 2), y = ya + (i & 3);
            const float px = __shfl(cxy, i >> 2), py = __shfl(cxy, 4 + (i & 3));
            if (x >= p.X || y >= p.Y) continue;
            bool in = false;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                uint64_t m = rel[ch];
                while (m && !in) {
                    const int k = 64 * ch + __ffsll((unsigned long long) m) - 1;
                    m &= m - 1;
                    in = inside(sphere_d(px, py, pz, sph[k]), p.level);
                }
            }
            const uint64_t word = __ballot(zok && in);
            if (lane == i) mine = word;
        }
        if (lane < 16) {
            const int64_t x = xa + (lane >> 2), y = ya + (lane & 3);
            if (x < p.X && y < p.Y) p.bits[(x * p.Y + y) * p.W + c] = mine;
        }
    }
}

// ---- fused SDF program field ---------------------------------------------------------
// Inside masks of v(p) = program(p) (csg.h) with exact sign culling: the
// field is 1-Lipschitz (programs without smooth min), so a box of grid
// points whose centre value is farther from the level than the box radius
// (plus a margin far above the FP32 evaluation error) has one sign.
// Phase 1 (one lane per block of 4 x 4 point rows x one 64-point word)
// tests each block at its centre, then an undecided block at its four
// 16-point quarters, writes the words of decided blocks and lists the
// others with their quarters' states; phase 2 (four waves per listed block)
// tests the block's 64 columns of 4 x 4 points (one z per lane) and
// evaluates the points of the columns still undecided, 64 at a time,
// assembling the rows' words.  Every evaluated value is the stored grid's
// (same program, same grid coordinates), and decided points have the sign
// their value would give.  The list lives in cnt_v (block ids) and cnt_f
// (its count in cnt_f[0], the quarter states after it), free until the
// count pass; when it does not fit (thin grids with deep emit segments, see
// csg_list_fits), phase 2 walks every block and repeats the centre test
// (p.csg_list = 0).

struct CsgBlock {
    int64_t xa, ya, c, za;
    int nx, ny, nz;
};
__device__ __forceinline__ CsgBlock csg_block(const McParams &p, int64_t b) {
    const int64_t nby = (p.Y + 3) / 4;
    CsgBlock k;
    k.c = b % p.W;
    const int64_t r = b / p.W, by = r % nby, bx = r / nby;
    k.xa = 4 * bx;
    k.ya = 4 * by;
    k.za = 64 * k.c;
    k.nx = (int) min<int64_t>(4, p.X - k.xa);
    k.ny = (int) min<int64_t>(4, p.Y - k.ya);
    k.nz = (int) min<int64_t>(64, p.Z - k.za);
    return k;
}

__device__ __forceinline__ double csg_coord_scale(const McParams &p) {
    double m = 0.0;
    for (int a = 0; a < 3; ++a) m = fmax(m, fmax(fabs((double) p.amin[a]), fabs((double) p.amax[a])));
    return m;
}

// Sign of the box of grid points [xa, xa + nx) x [ya, ya + ny) x [za, za + nz)
// from its centre value (every lane of the wave evaluates its own box, or
// all the same one).
__device__ __forceinline__ int csg_points_sign(const McParams &p, const CsgProgram &P, int64_t xa, int nx, int64_t ya,
                                               int ny, int64_t za, int nz, double cs) {
    const float x0 = grid_coord(xa, p.X, p.amin[0], p.amax[0]), x1 = grid_coord(xa + nx - 1, p.X, p.amin[0], p.amax[0]);
    const float y0 = grid_coord(ya, p.Y, p.amin[1], p.amax[1]), y1 = grid_coord(ya + ny - 1, p.Y, p.amin[1], p.amax[1]);
    const float z0 = grid_coord(za, p.Z, p.amin[2], p.amax[2]), z1 = grid_coord(za + nz - 1, p.Z, p.amin[2], p.amax[2]);
    const float cx = 0.5f * (x0 + x1), cy = 0.5f * (y0 + y1), cz = 0.5f * (z0 + z1);
    const double ex = fmax(fabs((double) cx - x0), fabs((double) x1 - cx));
    const double ey = fmax(fabs((double) cy - y0), fabs((double) y1 - cy));
    const double ez = fmax(fabs((double) cz - z0), fabs((double) z1 - cz));
    const float vc = csg_value<true>(P, cx, cy, cz);
    return csg_box_sign(P, vc, sqrt(ex * ex + ey * ey + ez * ez), cs, p.level);
}

__device__ __forceinline__ int csg_block_sign(const McParams &p, const CsgProgram &P, const CsgBlock &k, double cs) {
    return csg_points_sign(p, P, k.xa, k.nx, k.ya, k.ny, k.za, k.nz, cs);
}

__device__ __forceinline__ void csg_write_block(const McParams &p, const CsgBlock &k, int sign) {
    const uint64_t w = sign < 0 ? low_bits(k.nz) : 0ull;
    for (int r = 0; r < 16; ++r)
        if ((r >> 2) < k.nx && (r & 3) < k.ny) p.bits[((k.xa + (r >> 2)) * p.Y + (k.ya + (r & 3))) * p.W + k.c] = w;
}

// Phase 1 for block b (valid: b is a block of the grid); all lanes call it.
// A block undecided at its centre tests its four 16-point quarters (radius
// about a quarter of the block's); if all four are decided its words are
// written, else it is listed with the quarters' states (2 bits each: 0
// undecided, 1 inside, 2 outside or empty) in cnt_f[1 + list index].
__device__ void csg_test_block(const McParams &p, const CsgProgram &P, int64_t b, bool valid, double cs) {
    const CsgBlock k = csg_block(p, valid ? b : 0);
    const int sign = csg_block_sign(p, P, k, cs);
    if (valid && sign != 0) csg_write_block(p, k, sign);
    bool und = valid && sign == 0;
    if (!__ballot(und)) return;
    uint32_t qs = 0;
    if (und) {
        for (int q = 0; q < 4; ++q) {
            const int nzq = min(16, k.nz - 16 * q);
            const int sq = nzq > 0 ? csg_points_sign(p, P, k.xa, k.nx, k.ya, k.ny, k.za + 16 * q, nzq, cs) : 1;
            qs |= (uint32_t) (sq < 0 ? 1 : (sq > 0 ? 2 : 0)) << (2 * q);
        }
        bool all = true;
        uint64_t w = 0;
        for (int q = 0; q < 4; ++q) {
            const uint32_t st = (qs >> (2 * q)) & 3u;
            all = all && st != 0;
            if (st == 1) w |= 0xffffull << (16 * q);
        }
        if (all) {
            w &= low_bits(k.nz);
            for (int r = 0; r < 16; ++r)
                if ((r >> 2) < k.nx && (r & 3) < k.ny) p.bits[((k.xa + (r >> 2)) * p.Y + (k.ya + (r & 3))) * p.W + k.c] = w;
            und = false;
        }
    }
    const uint64_t lst = __ballot(und);
    if (!p.csg_list || !lst) return;
    const int lane = threadIdx.x & 63;
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(p.cnt_f, (unsigned) __popcll(lst));
    base = __shfl(base, 0);
    if (und) {
        const unsigned idx = base + (unsigned) __popcll(lst & ((1ull << lane) - 1ull));
        p.cnt_v[idx] = (uint32_t) b;
        p.cnt_f[1 + idx] = qs;
    }
}

// Phase 2 for an undecided block b by the NW waves of a workgroup (all of
// them call it); qs: its quarters' states from phase 1 (0: test every
// column).  Wave 0 tests the columns and shares the masks through LDS
// (wacc[16], wacc[17]); the rounds of point evaluations are dealt to the
// waves, whose partial words meet in LDS (wacc[0, 16)).
template <int NW>
__device__ void csg_fill_block(const McParams &p, const CsgProgram &P, int64_t b, uint32_t qs, double cs,
                               uint64_t *wacc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const CsgBlock k = csg_block(p, b);
    // columns: 4 x 4 rows at one z per lane
    const float x0 = grid_coord(k.xa, p.X, p.amin[0], p.amax[0]), x1 = grid_coord(k.xa + k.nx - 1, p.X, p.amin[0], p.amax[0]);
    const float y0 = grid_coord(k.ya, p.Y, p.amin[1], p.amax[1]), y1 = grid_coord(k.ya + k.ny - 1, p.Y, p.amin[1], p.amax[1]);
    const float cx = 0.5f * (x0 + x1), cy = 0.5f * (y0 + y1);
    const double ex = fmax(fabs((double) cx - x0), fabs((double) x1 - cx));
    const double ey = fmax(fabs((double) cy - y0), fabs((double) y1 - cy));
    if (wave == 0) {
        const bool zok = lane < k.nz;
        const uint32_t qst = (qs >> (2 * (lane >> 4))) & 3u;   // this lane's quarter, decided in phase 1
        int s2 = 1;
        if (zok && qst != 0) {
            s2 = qst == 1 ? -1 : 1;
        } else if (zok) {
            const float pz = grid_coord(k.za + lane, p.Z, p.amin[2], p.amax[2]);
            s2 = csg_box_sign(P, csg_value<true>(P, cx, cy, pz), sqrt(ex * ex + ey * ey), cs, p.level);
        }
        const uint64_t kin = __ballot(zok && s2 < 0), nd = __ballot(zok && s2 == 0);
        if (lane < 16) wacc[lane] = 0ull;
        if (lane == 0) {
            wacc[16] = kin;
            wacc[17] = nd;
        }
    }
    __syncthreads();
    const uint64_t known_in = wacc[16], need = wacc[17];
    const int nn = __popcll(need);
    uint64_t word = 0;   // lane r < 16: this wave's bits of row r
    for (int j0 = 64 * wave; j0 < 16 * nn; j0 += 64 * NW) {
        const int j = j0 + lane, r = lane & 15;
        const bool act = j < 16 * nn && (r >> 2) < k.nx && (r & 3) < k.ny;
        const int L = j < 16 * nn ? nth_set_bit(need, j >> 4) : 0;
        const float px = grid_coord(k.xa + (act ? (r >> 2) : 0), p.X, p.amin[0], p.amax[0]);
        const float py = grid_coord(k.ya + (act ? (r & 3) : 0), p.Y, p.amin[1], p.amax[1]);
        const float qz = grid_coord(k.za + L, p.Z, p.amin[2], p.amax[2]);
        const int in = act && inside(csg_value<true>(P, px, py, qz), p.level) ? 1 : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int bq = __shfl(in, r + 16 * q), lq = __shfl(L, r + 16 * q);
            if (bq) word |= 1ull << lq;
        }
    }
    if (lane < 16 && word) atomicOr((unsigned long long *) &wacc[lane], (unsigned long long) word);
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 16 && (t >> 2) < k.nx && (t & 3) < k.ny)
        p.bits[((k.xa + (t >> 2)) * p.Y + (k.ya + (t & 3))) * p.W + k.c] = (known_in | wacc[t]) & low_bits(k.nz);
    __syncthreads();   // wacc is reused by the next block
}

// Phase 2 over the list, or over every block (repeating the test); one
// block per workgroup at a time.
template <int NW>
__device__ void csg_fill(const McParams &p, const CsgProgram &P, double cs, int64_t first, int64_t stride,
                         uint64_t *wacc) {
    if (p.csg_list) {
        const int64_t n = __builtin_amdgcn_readfirstlane((int) *(volatile uint32_t *) p.cnt_f);
        for (int64_t i = first; i < n; i += stride)
            csg_fill_block<NW>(p, P, (int64_t) p.cnt_v[i], p.cnt_f[1 + i], cs, wacc);
    } else {
        const int64_t nb = csg_classify_tasks(p.X, p.Y, p.W);
        for (int64_t b = first; b < nb; b += stride)
            if (csg_block_sign(p, P, csg_block(p, b), cs) == 0) csg_fill_block<NW>(p, P, b, 0u, cs, wacc);
    }
}

__global__ void __launch_bounds__(256) k_csg_blocks(McParams p) {
    __shared__ CsgProgram P;
    csg_stage(P, p.csg);
    __syncthreads();
    const int64_t nb = csg_classify_tasks(p.X, p.Y, p.W);
    const double cs = csg_coord_scale(p);
    for (int64_t base = (int64_t) blockIdx.x * 256; base < nb; base += (int64_t) gridDim.x * 256)
        csg_test_block(p, P, base + threadIdx.x, base + threadIdx.x < nb, cs);
}

__global__ void __launch_bounds__(64 * kCsgFillWaves) k_csg_fill(McParams p) {
    __shared__ CsgProgram P;
    __shared__ uint64_t wacc[18];
    csg_stage(P, p.csg);
    __syncthreads();
    csg_fill<kCsgFillWaves>(p, P, csg_coord_scale(p), blockIdx.x, gridDim.x, wacc);
}

// LDS of the fused emit: per 64-point z chunk of the tile, the spheres that
// can be the minimum somewhere in the chunk's box (about 2 KB in all).
// Field sources of the emit kernels: a stored grid, the fused sphere union,
// a fused SDF program.
enum : int { kSrcGrid = 0, kSrcSpheres = 1, kSrcProgram = 2 };

template <int SRC, int NZC>
struct FusedLds {};

// Program fields: the distinct corner points of an emit tile's active cells,
// evaluated once each.  need[w]: bit b = point (px, py, 64 c + b) of point
// word w = (px * (TY + 1) + py) * (NZC + 1) + c is a corner of an active
// cell; pre: exclusive offsets of the words' points into val.
template <int TX, int TY, int NZC, bool ON>
struct CsgTileCache {};
template <int TX, int TY, int NZC>
struct CsgTileCache<TX, TY, NZC, true> {
    static constexpr int NPW = (TX + 1) * (TY + 1) * (NZC + 1);
    static constexpr int CAP = 4096;   // more distinct corners: each cell evaluates its own 8
    uint64_t need[NPW];
    uint32_t pre[NPW + 1];
    float val[CAP];
};
template <int NZC>
struct FusedLds<kSrcProgram, NZC> {
    CsgProgram prog;
};
template <int NZC>
struct FusedLds<kSrcSpheres, NZC> {
    static constexpr int CAP = 128 / NZC > 8 ? 128 / NZC : 8;   // spheres per chunk
    float4 sph[NZC][CAP];
    int n[NZC];   // culled spheres; > CAP: evaluate all of p.spheres instead
    double red[16];
};

// Keeps sphere k unless lb_k - m_k > min_j (ub_j + m_j) (then it is larger
// than another sphere's value everywhere in the box).  Block-wide; all
// threads call it.
template <int NT, int CAP>
__device__ void cull_spheres(const McParams &p, const Box3 &box, float4 *out, int &n_out, double *red) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) n_out = 0;
    double best = INFINITY;
    for (int k = tid; k < p.nsph; k += NT) {
        double lb, ub, m;
        sphere_bounds(box, p.spheres[k], lb, ub, m);
        best = fmin(best, ub + m);
    }
    for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o));
    if (lane == 0) red[wave] = best;
    __syncthreads();
    double tmin = INFINITY;
    for (int w = 0; w < NT / 64; ++w) tmin = fmin(tmin, red[w]);
    for (int k = tid; k < p.nsph; k += NT) {
        double lb, ub, m;
        const float4 sp = p.spheres[k];
        sphere_bounds(box, sp, lb, ub, m);
        if (lb - m <= tmin) {
            const int slot = atomicAdd(&n_out, 1);
            if (slot < CAP) out[slot] = sp;
        }
    }
    __syncthreads();
}

// ---- emit ------------------------------------------------------------------------
template <int TX, int TY, int NZC, int NT>
struct EmitLds {
    static constexpr int PX = TX + 2, PY = TY + 2, PC = NZC + 1;   // mask patch
    static constexpr int QY = TY + 1, QC = NZC + 1;                // vertex bases
    static constexpr int NUNITS = TX * TY * NZC;                   // (cell row, chunk), chunk fastest
    // active-cell offsets fit 16 bits when the tile has fewer than 2^16
    // cells; slot counts (at most 5 faces per slot) and unit indices always
    // do.  The narrow types keep the (16, 16, 2) tile under 32 KB, five
    // workgroups per CU.
    using NaT = std::conditional_t<(TX * TY * 64 * NZC < 65536), uint16_t, uint32_t>;
    static_assert(NUNITS < 32768 && 5 * NT < 65536, "narrow slot types");
    uint64_t w[PX * PY * PC];
    uint64_t amask[NUNITS];           // active cells of each unit
    int32_t vb[(TX + 1) * QY * QC];   // first vertex id per (point row, chunk)
    int32_t fb[TX * TY];              // first face id per cell row (this z-segment)
    NaT na[NUNITS + 1];               // active cells per unit -> exclusive offsets
    int16_t unit[NT];                 // batch: unit of each slot
    uint16_t fsc[NT + 1];             // batch: faces per slot -> exclusive offsets
    uint32_t tsum[NT / 64];
    int32_t carry[2];                 // faces of the straddling row's cells before a batch
    int32_t ids[12][NT];              // batch: vertex id of each edge of each slot's cell
    float cx[TX + 1], cy[TY + 1], cz[64 * NZC + 1];   // grid coordinates of the tile's points
};

template <int TX, int TY, int NZC, int NT>
__device__ __forceinline__ uint64_t PW(const EmitLds<TX, TY, NZC, NT> &s, int px, int py, int j) {
    using L = EmitLds<TX, TY, NZC, NT>;
    return s.w[(px * L::PY + py) * L::PC + j];
}

template <int TX, int TY, int NZC, int NT>
__device__ __forceinline__ CellMasks unit_cells(const EmitLds<TX, TY, NZC, NT> &s, int cx, int cy, int j) {
    return cell_masks(PW(s, cx, cy, j), PW(s, cx, cy, j + 1), PW(s, cx, cy + 1, j), PW(s, cx, cy + 1, j + 1),
                      PW(s, cx + 1, cy, j), PW(s, cx + 1, cy, j + 1), PW(s, cx + 1, cy + 1, j),
                      PW(s, cx + 1, cy + 1, j + 1));
}

// Global id of the vertex on edge (owner point (opx, opy, oz) tile-local,
// axis): the (point row, chunk) base plus the rank among the row's crossing
// edges in (z, axis) order.  hx / hy: tile-local point rows below which the
// row at x+1 / y+1 exists.
template <int TX, int TY, int NZC, int NT>
__device__ __forceinline__ int vertex_id(const EmitLds<TX, TY, NZC, NT> &s, int hx, int hy, int opx, int opy, int oz,
                                         int axis) {
    using L = EmitLds<TX, TY, NZC, NT>;
    const int j = oz >> 6, k = oz & 63;
    const uint64_t b = PW(s, opx, opy, j);
    const uint64_t xm = opx < hx ? (b ^ PW(s, opx + 1, opy, j)) : 0ull;
    const uint64_t ym = opy < hy ? (b ^ PW(s, opx, opy + 1, j)) : 0ull;
    const uint64_t zm = b ^ (b >> 1);   // only bits below k are used
    const uint64_t low = (1ull << k) - 1;
    int r = __popcll(xm & low) + __popcll(ym & low) + __popcll(zm & low);
    if (axis >= 1) r += (int) ((xm >> k) & 1);
    if (axis == 2) r += (int) ((ym >> k) & 1);
    return s.vb[(opx * L::QY + opy) * L::QC + j] + r;
}

template <int N, int NT, typename T>
__device__ __forceinline__ void block_exscan_lds(T *a, uint32_t *tsum) {
    constexpr int NW = NT / 64, PER = (N + NT - 1) / NT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lo = tid * PER, hi = min(lo + PER, N);
    uint32_t sum = 0;
    for (int i = lo; i < hi; ++i) sum += a[i];
    const uint32_t inc = (uint32_t) wave_incl_sum((int) sum);
    if (lane == 63) tsum[wave] = inc;
    __syncthreads();
    uint32_t run = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        if (w < wave) run += tsum[w];
        tot += tsum[w];
    }
    run += inc - sum;
    for (int i = lo; i < hi; ++i) {
        const uint32_t v = a[i];
        a[i] = (T) run;
        run += v;
    }
    if (tid == 0) a[N] = (T) tot;
    __syncthreads();
}

// Inclusive running maximum over a[0, NT) (one element per thread).
template <int NT, typename T>
__device__ __forceinline__ void block_maxscan_lds(T *a, uint32_t *tsum) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int32_t m = wave_incl_max((int32_t) a[tid]);
    if (lane == 63) tsum[wave] = (uint32_t) m;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < NW; ++w)
        if (w < wave) m = max(m, (int32_t) tsum[w]);
    a[tid] = (T) m;
    __syncthreads();
}


// Weld-candidate test with a division-free prefilter: an edge's vertex can
// lie within 16 ulps of an endpoint coordinate only when its parameter t is
// within cand_eps of 0 or 1, i.e. |level - va| or |vb - level| is at most
// cand_eps * |vb - va| (cand_eps holds a 2x margin over the rounding of
// both sides); the exact test runs for the edges that pass.
template <int E>
__device__ __forceinline__ bool edge_maybe_candidate(const CellCtx &cc, float eps) {
    constexpr int ca = edge_a(E), cb = edge_b(E);
    const float d = fabsf(cc.v[cb] - cc.v[ca]) * eps;
    // written so that NaN or infinite operands pass to the exact test
    return !(fabsf(cc.level - cc.v[ca]) > d && fabsf(cc.v[cb] - cc.level) > d);
}

template <int TX, int TY, int NZC, int NT, int SRC = kSrcGrid>
__device__ __forceinline__ void emit_tile(const McParams &p, int64_t tile, int spseg) {
    using L = EmitLds<TX, TY, NZC, NT>;
    __shared__ L s;
    __shared__ FusedLds<SRC, NZC> fl;
    __shared__ CsgTileCache<TX, TY, NZC, SRC == kSrcProgram> cache;
    const int tid = threadIdx.x;
    if (*(volatile int *) p.err & ISO_ERR_TOO_LARGE) return;

    // small grids (nblk <= NT): the block totals are scanned here, into the
    // LDS of the id table (used only by the batches below, after the bases
    // are read); workgroup 0 also publishes the bases and the mesh size
    const int64_t *bv_tab = p.blk_v, *bf_tab = p.blk_f;
    if (p.scan_in_emit) {
        static_assert(12 * NT * sizeof(int32_t) >= 2 * NT * sizeof(int64_t), "scan table in the id table");
        __shared__ int64_t wsa[NT / 64], wsb[NT / 64];
        int64_t *tab = reinterpret_cast<int64_t *>(&s.ids[0][0]);
        const int lane = tid & 63, wave = tid >> 6;
        const int64_t a = tid < p.nblk ? (int64_t) p.bsum_v[tid] : 0, b = tid < p.nblk ? (int64_t) p.bsum_f[tid] : 0;
        int64_t ia = a, ib = b;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t ya = __shfl_up(ia, o), yb = __shfl_up(ib, o);
            if (lane >= o) {
                ia += ya;
                ib += yb;
            }
        }
        if (lane == 63) {
            wsa[wave] = ia;
            wsb[wave] = ib;
        }
        __syncthreads();
        int64_t ra = ia - a, rb = ib - b, ta = 0, tb = 0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) {
            if (w < wave) {
                ra += wsa[w];
                rb += wsb[w];
            }
            ta += wsa[w];
            tb += wsb[w];
        }
        tab[tid] = ra;
        tab[NT + tid] = rb;
        if (blockIdx.x == 0) {
            if (tid < p.nblk) {
                p.blk_v[tid] = ra;
                p.blk_f[tid] = rb;
            }
            if (tid == 0) {
                p.totals[0] = ta;
                p.totals[1] = tb;
                if (ta > 0x7fffffffll || tb > 0x7fffffffll) set_err(p, ISO_ERR_TOO_LARGE);
            }
        }
        __syncthreads();
        bv_tab = tab;
        bf_tab = tab + NT;
    }

    const int64_t tk = tile % p.NK, tj = (tile / p.NK) % p.NJ, ti = tile / (p.NK * p.NJ);
    const int64_t x0 = ti * TX, y0 = tj * TY, c0 = tk * NZC, z0 = 64 * c0;
    const int nx = (int) min<int64_t>(TX, p.X - 1 - x0);
    const int ny = (int) min<int64_t>(TY, p.Y - 1 - y0);
    const int hx = (int) min<int64_t>(TX + 2, p.X - 1 - x0);   // point rows px < hx have a row at px + 1
    const int hy = (int) min<int64_t>(TY + 2, p.Y - 1 - y0);

    // ---- mask patch: point rows [x0, x0+TX+2) x [y0, y0+TY+2), chunks [c0, c0+NZC];
    // all loads are issued before the first use, together with the global
    // bases the tile needs if it has active cells
    constexpr int NP = L::PX * L::PY * L::PC, IT = (NP + NT - 1) / NT;
    constexpr int CR = (TX * TY + NT - 1) / NT, PR = ((TX + 1) * (TY + 1) + NT - 1) / NT;
    uint64_t wv[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = tid + k * NT;
        const int j = i % L::PC, r = i / L::PC, py = r % L::PY, px = r / L::PY;
        const int64_t gx = x0 + px, gy = y0 + py, c = c0 + j;
        wv[k] = (i < NP && gx < p.X && gy < p.Y && c < p.W) ? p.bits[(gx * p.Y + gy) * p.W + c] : 0ull;
    }
    int64_t fbase_pf[CR], vbase_pf[PR], vnext_pf[PR];
#pragma unroll
    for (int k = 0; k < CR; ++k) {
        const int i = tid + k * NT, cx = i / TY, cy = i % TY;
        fbase_pf[k] = 0;
        if (i < TX * TY && cx < nx && cy < ny) {
            const int64_t e = ((x0 + cx) * p.Y + (y0 + cy)) * p.S + tk;
            fbase_pf[k] = bf_tab[e / kMcScanBlock] + (int64_t) p.cnt_f[e];
        }
    }
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int i = tid + k * NT, px = i / (ny + 1), py = i % (ny + 1);
        vbase_pf[k] = vnext_pf[k] = 0;
        if (i < (nx + 1) * (ny + 1)) {
            const int64_t e = ((x0 + px) * p.Y + (y0 + py)) * p.S + tk;
            vbase_pf[k] = bv_tab[e / kMcScanBlock] + (int64_t) p.cnt_v[e];
            if (tk + 1 < p.S) vnext_pf[k] = bv_tab[(e + 1) / kMcScanBlock] + (int64_t) p.cnt_v[e + 1];
        }
    }
    for (int c = tid; c < TX + TY + 64 * NZC + 3; c += NT) {
        if (c <= TX) {
            s.cx[c] = grid_coord(x0 + c, p.X, p.amin[0], p.amax[0]);
        } else if (c <= TX + TY + 1) {
            s.cy[c - TX - 1] = grid_coord(y0 + (c - TX - 1), p.Y, p.amin[1], p.amax[1]);
        } else {
            s.cz[c - TX - TY - 2] = grid_coord(z0 + (c - TX - TY - 2), p.Z, p.amin[2], p.amax[2]);
        }
    }
    if constexpr (SRC == kSrcProgram) csg_stage(fl.prog, p.csg);
    if constexpr (SRC == kSrcSpheres) {
        using F = FusedLds<kSrcSpheres, NZC>;
        for (int j = 0; j < NZC; ++j) {
            const int64_t za = min(z0 + 64 * j, p.Z - 1);
            cull_spheres<NT, F::CAP>(p, point_box(p, x0, min(x0 + TX, p.X - 1), y0, min(y0 + TY, p.Y - 1), za,
                                                  min(za + 64, p.Z - 1)),
                                     fl.sph[j], fl.n[j], fl.red);
        }
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = tid + k * NT;
        if (i < NP) s.w[i] = wv[k];
    }
    __syncthreads();

    // ---- active cells per unit (cell row, chunk) --------------------------------
    for (int u = tid; u < L::NUNITS; u += NT) {
        const int j = u % NZC, r = u / NZC, cy = r % TY, cx = r / TY;
        const int64_t zlo = z0 + 64 * j;
        uint64_t act = 0;
        if (cx < nx && cy < ny && zlo < p.Z - 1)
            act = cell_active(unit_cells(s, cx, cy, j)) & low_bits(p.Z - 1 - zlo);
        s.amask[u] = act;
        s.na[u] = (typename L::NaT) __popcll(act);
    }
    __syncthreads();
    block_exscan_lds<L::NUNITS, NT>(s.na, s.tsum);
    const int nactive = (int) s.na[L::NUNITS];
    if (nactive == 0) return;

    // ---- program fields: every corner point of the active cells evaluated once
    bool cached = false;
    if constexpr (SRC == kSrcProgram) {
        using C = CsgTileCache<TX, TY, NZC, true>;
        for (int pw = tid; pw < C::NPW; pw += NT) {
            const int c = pw % (NZC + 1), r = pw / (NZC + 1), py = r % (TY + 1), px = r / (TY + 1);
            uint64_t m = 0;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int cx = px - (d >> 1), cy = py - (d & 1);
                if (cx < 0 || cy < 0 || cx >= TX || cy >= TY) continue;
                const int ub = (cx * TY + cy) * NZC;
                if (c < NZC) {
                    const uint64_t a = s.amask[ub + c];
                    m |= a | (a << 1);
                }
                if (c > 0) m |= s.amask[ub + c - 1] >> 63;
            }
            cache.need[pw] = m;
            cache.pre[pw] = (uint32_t) __popcll(m);
        }
        __syncthreads();
        block_exscan_lds<C::NPW, NT>(cache.pre, s.tsum);
        const int U = (int) cache.pre[C::NPW];
        cached = U <= C::CAP;
        if (cached) {
            for (int j = tid; j < U; j += NT) {
                int lo = 0, hi = C::NPW - 1;   // the word holding point j: first w with pre[w + 1] > j
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (cache.pre[mid + 1] > (uint32_t) j) hi = mid;
                    else lo = mid + 1;
                }
                const int c = lo % (NZC + 1), r = lo / (NZC + 1), py = r % (TY + 1), px = r / (TY + 1);
                const int zl = 64 * c + nth_set_bit(cache.need[lo], j - (int) cache.pre[lo]);
                cache.val[j] = csg_value<true>(fl.prog, s.cx[px], s.cy[py], s.cz[zl]);
            }
            __syncthreads();
        }
    }

    // ---- global bases: faces per cell row, vertices per (point row, chunk) ------
#pragma unroll
    for (int k = 0; k < CR; ++k) {
        const int i = tid + k * NT;
        if (i < TX * TY) s.fb[i] = (int32_t) fbase_pf[k];
    }
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int i = tid + k * NT;
        if (i >= (nx + 1) * (ny + 1)) continue;
        const int px = i / (ny + 1), py = i % (ny + 1);
        const int64_t gx = x0 + px, gy = y0 + py;
        const bool hasx = gx + 1 < p.X, hasy = gy + 1 < p.Y;
        int64_t base = vbase_pf[k];
        int32_t *vb = &s.vb[(px * L::QY + py) * L::QC];
        for (int j = 0; j < NZC; ++j) {
            vb[j] = (int32_t) base;
            if (c0 + j < p.W)
                base += chunk_vertices(PW(s, px, py, j), PW(s, px, py, j + 1) & 1, PW(s, px + 1, py, j),
                                       PW(s, px, py + 1, j), hasx, hasy, 64 * (c0 + j), p.Z);
        }
        vb[NZC] = (tk + 1 < p.S) ? (int32_t) vnext_pf[k] : (int32_t) base;
    }
    if (tid == 0) s.carry[0] = 0;

    // ---- emit, NT active cells per batch ----------------------------------------
    const float lvl = p.level;
    for (int j0 = 0, bt = 0; j0 < nactive; j0 += NT, bt ^= 1) {
        // slot -> unit: each unit overlapping the batch marks its first slot,
        // a running maximum carries it over the unit's other slots
        s.unit[tid] = -1;
        __syncthreads();
        for (int u = tid; u < L::NUNITS; u += NT) {
            const int a0 = (int) s.na[u], a1 = (int) s.na[u + 1];
            if (a1 > a0 && a0 < j0 + NT && a1 > j0) s.unit[max(a0, j0) - j0] = u;
        }
        __syncthreads();
        block_maxscan_lds<NT>(s.unit, s.tsum);

        const int jj = j0 + tid;
        const bool live = jj < nactive;
        int u = 0, cs = 0, k = 0;
        float pre[8];   // corner values, loaded before the face scan so that the scan hides their latency
        if (live) {
            u = s.unit[tid];
            k = nth_set_bit(s.amask[u], jj - (int) s.na[u]);
            const int jc = u % NZC, r = u / NZC, cyy = r % TY, cxx = r / TY;
            if constexpr (SRC == kSrcGrid) {
                const float *r00 = p.values + ((x0 + cxx) * p.Y + (y0 + cyy)) * p.Z + (z0 + 64 * jc + k);
                const float *r01 = r00 + p.Z;
                const float *r10 = r00 + p.Y * p.Z;
                const float *r11 = r10 + p.Z;
                pre[0] = r00[0];
                pre[1] = r00[1];
                pre[2] = r01[0];
                pre[3] = r01[1];
                pre[4] = r10[0];
                pre[5] = r10[1];
                pre[6] = r11[0];
                pre[7] = r11[1];
            }
            // corner k of the cell sits at (k>>2, (k>>1)&1, k&1)
            const uint64_t w00 = PW(s, cxx, cyy, jc), w01 = PW(s, cxx, cyy + 1, jc);
            const uint64_t w10 = PW(s, cxx + 1, cyy, jc), w11 = PW(s, cxx + 1, cyy + 1, jc);
            uint32_t up;   // bit z+1 of the four rows (bit 0 of the next chunk when z is the chunk's last)
            if (k < 63) {
                up = (uint32_t) ((w00 >> (k + 1)) & 1) | (uint32_t) ((w01 >> (k + 1)) & 1) << 1 |
                     (uint32_t) ((w10 >> (k + 1)) & 1) << 2 | (uint32_t) ((w11 >> (k + 1)) & 1) << 3;
            } else {
                up = (uint32_t) (PW(s, cxx, cyy, jc + 1) & 1) | (uint32_t) (PW(s, cxx, cyy + 1, jc + 1) & 1) << 1 |
                     (uint32_t) (PW(s, cxx + 1, cyy, jc + 1) & 1) << 2 |
                     (uint32_t) (PW(s, cxx + 1, cyy + 1, jc + 1) & 1) << 3;
            }
            cs = (int) (((w00 >> k) & 1) | ((up & 1) << 1) | (((w01 >> k) & 1) << 2) | (((up >> 1) & 1) << 3) |
                        (((w10 >> k) & 1) << 4) | (((up >> 2) & 1) << 5) | (((w11 >> k) & 1) << 6) |
                        (((up >> 3) & 1) << 7));
        }
        s.fsc[tid] = live ? (uint16_t) cNtri[p.method][cs] : (uint16_t) 0;
        __syncthreads();
        block_exscan_lds<NT, NT>(s.fsc, s.tsum);
        // faces of the straddling row's cells before the next batch
        if (tid == NT - 1 && j0 + NT < nactive) {
            const int rl = s.unit[NT - 1] / NZC;
            const int rs = (int) s.na[rl * NZC], re = (int) s.na[(rl + 1) * NZC];
            s.carry[bt ^ 1] = re > j0 + NT ? (rs < j0 ? s.carry[bt] : 0) + (int) (s.fsc[NT] - s.fsc[max(rs - j0, 0)])
                                           : 0;
        }
        if (live) {
            const int jc = u % NZC, r = u / NZC, cy = r % TY, cx = r / TY;
            const int rs = (int) s.na[r * NZC];   // the row's first active cell
            const int fo = (int) (s.fsc[tid] - s.fsc[max(rs - j0, 0)]) + (rs < j0 ? s.carry[bt] : 0);
            const int zl = 64 * jc + k;
            const int64_t gx = x0 + cx, gy = y0 + cy, gz = z0 + zl;
            CellCtx cc;
            if constexpr (SRC == kSrcProgram) {
                if (cached) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int px = cx + (q >> 2), py = cy + ((q >> 1) & 1), zq = zl + (q & 1);
                        const int w = (px * (TY + 1) + py) * (NZC + 1) + (zq >> 6);
                        const uint64_t below = (1ull << (zq & 63)) - 1ull;
                        cc.v[q] = cache.val[cache.pre[w] + (uint32_t) __popcll(cache.need[w] & below)];
                    }
                } else {
                    const float xs[2] = {s.cx[cx], s.cx[cx + 1]}, ys[2] = {s.cy[cy], s.cy[cy + 1]};
                    const float zs[2] = {s.cz[zl], s.cz[zl + 1]};
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        cc.v[q] = csg_value<true>(fl.prog, xs[q >> 2], ys[(q >> 1) & 1], zs[q & 1]);
                }
            } else if constexpr (SRC == kSrcSpheres) {
                const float xs[2] = {s.cx[cx], s.cx[cx + 1]}, ys[2] = {s.cy[cy], s.cy[cy + 1]};
                const float zs[2] = {s.cz[zl], s.cz[zl + 1]};
                using F = FusedLds<kSrcSpheres, NZC>;
                const int ns = fl.n[jc];
                float best[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) best[q] = INFINITY;
                if (ns <= F::CAP) {   // the chunk's culled spheres (LDS)
                    for (int j = 0; j < ns; ++j) {
                        const float4 sp = fl.sph[jc][j];
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            best[q] = fminf(best[q], sphere_d(xs[q >> 2], ys[(q >> 1) & 1], zs[q & 1], sp));
                    }
                } else {   // too many kept: all spheres (the same minimum)
                    for (int j = 0; j < p.nsph; ++j) {
                        const float4 sp = p.spheres[j];
#pragma unroll
                        for (int q = 0; q < 8; ++q)
                            best[q] = fminf(best[q], sphere_d(xs[q >> 2], ys[(q >> 1) & 1], zs[q & 1], sp));
                    }
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) cc.v[q] = best[q];
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) cc.v[q] = pre[q];
            }
            cc.c[0][0] = s.cx[cx];
            cc.c[0][1] = s.cx[cx + 1];
            cc.c[1][0] = s.cy[cy];
            cc.c[1][1] = s.cy[cy + 1];
            cc.c[2][0] = s.cz[zl];
            cc.c[2][1] = s.cz[zl + 1];
            cc.level = lvl;
            const uint32_t status = (uint32_t) crossing_status(cs);
            // weld candidates: a vertex within a few ulps of a grid corner
            uint32_t candm = 0;
            if (p.all_special) {
                candm = status;
            } else {
                // a cell whose every corner lies further from the level than
                // cand_eps * (max - min) of its corners has no edge that passes
                // the prefilter (|vb - va| <= max - min, and rounding is
                // monotone); NaN or infinite corners fall through to it
                float vlo = cc.v[0], vhi = cc.v[0];
#pragma unroll
                for (int q = 1; q < 8; ++q) {
                    vlo = fminf(vlo, cc.v[q]);
                    vhi = fmaxf(vhi, cc.v[q]);
                }
                const float dcell = (vhi - vlo) * p.cand_eps;
                bool near = false;
#pragma unroll
                for (int q = 0; q < 8; ++q) near |= !(fabsf(cc.level - cc.v[q]) > dcell);
                uint32_t maybe = 0;
                if (near)
                    for_edges([&](auto E) {
                        if (((status >> E) & 1) && edge_maybe_candidate<E>(cc, p.cand_eps)) maybe |= 1u << E;
                    });
                if (maybe) {
                    for_edges([&](auto E) {
                        if (((maybe >> E) & 1) && edge_is_candidate<E>(cc)) candm |= 1u << E;
                    });
                }
            }
            const bool special = candm != 0;
            // vertex ids of the cell's 12 edges (meaningful where crossing):
            // ranks in the owner rows' words from the masks of the cell's
            // four point rows and their x+1 / y+1 neighbours
            int32_t id[12];
            {
                const uint64_t b00 = PW(s, cx, cy, jc), b01 = PW(s, cx, cy + 1, jc);
                const uint64_t b10 = PW(s, cx + 1, cy, jc), b11 = PW(s, cx + 1, cy + 1, jc);
                const uint64_t b20 = PW(s, cx + 2, cy, jc), b21 = PW(s, cx + 2, cy + 1, jc);
                const uint64_t b02 = PW(s, cx, cy + 2, jc), b12 = PW(s, cx + 1, cy + 2, jc);
                const bool hx0 = cx < hx, hx1 = cx + 1 < hx, hy0 = cy < hy, hy1 = cy + 1 < hy;
                const uint64_t xm00 = hx0 ? b00 ^ b10 : 0ull, xm01 = hx0 ? b01 ^ b11 : 0ull;
                const uint64_t xm10 = hx1 ? b10 ^ b20 : 0ull, xm11 = hx1 ? b11 ^ b21 : 0ull;
                const uint64_t ym00 = hy0 ? b00 ^ b01 : 0ull, ym10 = hy0 ? b10 ^ b11 : 0ull;
                const uint64_t ym01 = hy1 ? b01 ^ b02 : 0ull, ym11 = hy1 ? b11 ^ b12 : 0ull;
                const uint64_t zm00 = b00 ^ (b00 >> 1), zm01 = b01 ^ (b01 >> 1), zm10 = b10 ^ (b10 >> 1);
                const uint64_t zm11 = b11 ^ (b11 >> 1);
                const uint64_t low = (1ull << k) - 1;
                auto rank = [&](uint64_t xm, uint64_t ym, uint64_t zm) {
                    return __popcll(xm & low) + __popcll(ym & low) + __popcll(zm & low);
                };
                auto bit = [&](uint64_t m, int q) { return (int) ((m >> q) & 1); };
                constexpr int QC = L::QC, QY = L::QY;
                const int v00 = s.vb[(cx * QY + cy) * QC + jc], v01 = s.vb[(cx * QY + cy + 1) * QC + jc];
                const int v10 = s.vb[((cx + 1) * QY + cy) * QC + jc];
                const int v11 = s.vb[((cx + 1) * QY + cy + 1) * QC + jc];
                const int r00 = v00 + rank(xm00, ym00, zm00), r01 = v01 + rank(xm01, ym01, zm01);
                const int r10 = v10 + rank(xm10, ym10, zm10), r11 = v11 + rank(xm11, ym11, zm11);
                // point (row, k): x-edge, y-edge, z-edge in that order; the
                // point at k + 1 follows the three edges of k
                const int x00 = bit(xm00, k), y00 = bit(ym00, k), z00 = bit(zm00, k);
                const int x01 = bit(xm01, k), y01 = bit(ym01, k), z01 = bit(zm01, k);
                const int x10 = bit(xm10, k), y10 = bit(ym10, k), z10 = bit(zm10, k);
                id[8] = r00;                                  // corner 0 x
                id[3] = r00 + x00;                            // corner 0 y
                id[0] = r00 + x00 + y00;                      // corner 0 z
                id[11] = r01;                                 // corner 2 x
                id[2] = r01 + x01 + y01;                      // corner 2 z
                id[7] = r10 + x10;                            // corner 4 y
                id[4] = r10 + x10 + y10;                      // corner 4 z
                id[6] = r11 + bit(xm11, k) + bit(ym11, k);    // corner 6 z
                if (k < 63) {
                    const int s00 = r00 + x00 + y00 + z00, s01 = r01 + x01 + y01 + z01;
                    const int s10 = r10 + x10 + y10 + z10;
                    id[9] = s00;                              // corner 1 x
                    id[1] = s00 + bit(xm00, k + 1);           // corner 1 y
                    id[10] = s01;                             // corner 3 x
                    id[5] = s10 + bit(xm10, k + 1);           // corner 5 y
                } else {   // corners at z + 1 sit in the next word
                    id[9] = vertex_id(s, hx, hy, cx, cy, zl + 1, 0);
                    id[1] = vertex_id(s, hx, hy, cx, cy, zl + 1, 1);
                    id[10] = vertex_id(s, hx, hy, cx, cy + 1, zl + 1, 0);
                    id[5] = vertex_id(s, hx, hy, cx + 1, cy, zl + 1, 1);
                }
#pragma unroll
                for (int e = 0; e < 12; ++e) s.ids[e][tid] = id[e];
            }
            // vertices owned by this cell: the edges of its lowest corner, plus
            // those of upper corners on the grid's upper faces (no cell above)
            {
                const bool lastx = gx == p.X - 2, lasty = gy == p.Y - 2, lastz = gz == p.Z - 2;
                uint32_t emit = (1u << 0) | (1u << 3) | (1u << 8);   // corner 0: z, y, x
                if (lastx) emit |= (1u << 7) | (1u << 4);             // corner 4: y, z
                if (lasty) emit |= (1u << 11) | (1u << 2);            // corner 2: x, z
                if (lastz) emit |= (1u << 9) | (1u << 1);             // corner 1: x, y
                if (lastx && lasty) emit |= 1u << 6;                  // corner 6: z
                if (lastx && lastz) emit |= 1u << 5;                  // corner 5: y
                if (lasty && lastz) emit |= 1u << 10;                 // corner 3: x
                const uint32_t m = emit & status;
                for_edges([&](auto E) {
                    if (!((m >> E) & 1)) return;
                    const int64_t v = id[E];
                    float x, y, z;
                    edge_pos_c<E>(cc, x, y, z);
                    if (v >= 0 && v < p.cap_v) {
                        *reinterpret_cast<float3 *>(p.out_v + 3 * v) = make_float3(x, y, z);
                        if (p.out_key) p.out_key[v] = edge_key(p, gx, gy, gz, E);
                    }   // else: the host sees totals beyond the capacity and retries
                    if ((candm >> E) & 1) {
                        const int64_t i = seg_push(p, 0, spseg);
                        if (i >= 0) {
                            SpecialVertex sv;
                            sv.vid = v;
                            sv.edge = edge_key(p, gx, gy, gz, E);
                            sv.x = x;
                            sv.y = y;
                            sv.z = z;
                            sv.pad = 0.0f;
                            p.seg.cand[i] = sv;
                        }
                    }
                });
            }
            // faces
            const uint64_t wl = cTris[p.method][cs];
            const int ntri = (int) (wl & 15);
            int64_t slot = (int64_t) s.fb[r] + fo;
            bool dropped = false;   // a face of this cell is degenerate
            for (int q = 0; q < ntri; ++q, ++slot) {
                const int e0 = (int) ((wl >> (4 + 12 * q)) & 15);
                const int e1 = (int) ((wl >> (8 + 12 * q)) & 15);
                const int e2 = (int) ((wl >> (12 + 12 * q)) & 15);
                bool keep = true;
                if (special) {
                    float xa, ya, za, xb, yb, zb, xc, yc, zc;
                    edge_pos_dyn(cc, e0, xa, ya, za);
                    edge_pos_dyn(cc, e1, xb, yb, zb);
                    edge_pos_dyn(cc, e2, xc, yc, zc);
                    const bool d01 = xa == xb && ya == yb && za == zb;
                    const bool d02 = xa == xc && ya == yc && za == zc;
                    const bool d12 = xb == xc && yb == yc && zb == zc;
                    keep = !(d01 || d02 || d12);
                    dropped |= !keep;
                    const int64_t i = seg_push(p, keep ? 2 : 3, spseg);
                    if (i >= 0) (keep ? p.seg.kept : p.seg.hole)[i] = slot;
                }
                if (slot >= 0 && slot < p.cap_f) {
                    int32_t *dst = p.out_f + 3 * slot;
                    if (keep) {
                        *reinterpret_cast<int3 *>(dst) = make_int3(s.ids[e0][tid], s.ids[e1][tid], s.ids[e2][tid]);
                    } else {
                        *reinterpret_cast<int3 *>(dst) = make_int3(-1, -1, -1);
                    }
                }
            }
            // the crossing edges of a cell that dropped a face: the fix-up
            // needs them to tell whether a vertex still has a face (a vertex
            // next to a cell that dropped none is referenced by its faces)
            if (special && dropped) {
                uint32_t m = (uint32_t) status;
                while (m) {
                    const int e = __ffs(m) - 1;
                    m &= m - 1;
                    const int64_t i = seg_push(p, 1, spseg);
                    if (i >= 0) {
                        SpecialEdge se;
                        se.vid = s.ids[e][tid];
                        se.edge = edge_key(p, gx, gy, gz, e);
                        p.seg.scell[i] = se;
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Tile of emit block b: the b-th listed active tile when the count pass
// built the list (blocks past the list's length have nothing to do), else b.
__device__ __forceinline__ int64_t emit_block_tile(const McParams &p) {
    if (!p.tile_list) return blockIdx.x;
    return (int) blockIdx.x < *p.tile_count ? (int64_t) p.tile_list[blockIdx.x] : -1;
}

// Five waves per SIMD: 96 VGPRs, and the narrow LDS slot types keep the
// (16, 16, 2) and smaller tiles under 32 KB (five workgroups per CU).
template <int TX, int TY, int NZC, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5, 8))) k_emit(McParams p) {
    const int64_t tile = emit_block_tile(p);
    if (tile >= 0) emit_tile<TX, TY, NZC, NT>(p, tile, (int) (blockIdx.x % p.seg.nseg));
}

template <int TX, int TY, int NZC, int NT>
__global__ void __launch_bounds__(NT) k_emit_spheres(McParams p) {
    const int64_t tile = emit_block_tile(p);
    if (tile >= 0) emit_tile<TX, TY, NZC, NT, kSrcSpheres>(p, tile, (int) (blockIdx.x % p.seg.nseg));
}

template <int TX, int TY, int NZC, int NT>
__global__ void __launch_bounds__(NT) k_emit_program(McParams p) {
    const int64_t tile = emit_block_tile(p);
    if (tile >= 0) emit_tile<TX, TY, NZC, NT, kSrcProgram>(p, tile, (int) (blockIdx.x % p.seg.nseg));
}

// Reset of the per-call counters (and the active-tile flags when listed);
// grid-stride over max(4 * nseg, tiles) threads.
__global__ void mc_dense_reset(McParams p, SpecialLists init) {
    reset_state(p, init, (int64_t) blockIdx.x * blockDim.x + threadIdx.x, (int64_t) gridDim.x * blockDim.x);
}

// Concatenate the special-list segments (in segment order) into the
// contiguous lists and record the counts and the capacity a call needs.
// Workgroup b copies segment b of every list; its offset is the sum of the
// (clamped) counts of the segments before it.
static_assert(kSpSeg == 256, "sp_gather_seg maps one thread to each segment");
__device__ __forceinline__ void sp_gather_seg(const McParams &p, int b) {
    // 256 threads, nseg <= 256: thread t reads segment t's four counters in
    // one round of loads; the offsets (clamped counts of the segments before
    // b), block 0's totals (unclamped) and the capacity a call needs come
    // from the same wave reductions, and the four lists are copied in one
    // interleaved round per 256 entries
    constexpr int NW = 4;
    __shared__ int64_t part[4][NW], own[4];
    __shared__ unsigned long long tot[4][NW], needw[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nseg = p.seg.nseg;
    unsigned long long c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = tid < nseg ? p.seg.ctr[(k * nseg + tid) * kSpStride] : 0ull;
    unsigned long long need = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t cl = min((int64_t) c[k], p.seg.cap[k]);
        if (tid == b) own[k] = cl;
        int64_t v = tid < b ? cl : 0;
        unsigned long long t = c[k];
        for (int o = 32; o > 0; o >>= 1) {
            v += __shfl_xor(v, o);
            t += __shfl_xor(t, o);
        }
        if (lane == 0) {
            part[k][wave] = v;
            tot[k][wave] = t;
        }
        const unsigned long long mult = k == 1 ? 12 : (k >= 2 ? 5 : 1);
        if (c[k] > 0) need = max(need, ((c[k] + mult - 1) / mult) * (unsigned long long) nseg);
    }
    for (int o = 32; o > 0; o >>= 1) need = max(need, (unsigned long long) __shfl_xor(need, o));
    if (lane == 0) needw[wave] = need;
    __syncthreads();
    int64_t off[4], n[4], nmax = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        off[k] = part[k][0] + part[k][1] + part[k][2] + part[k][3];
        n[k] = own[k];
        nmax = max(nmax, n[k]);
    }
    const int64_t b64 = b;
    for (int64_t i = tid; i < nmax; i += 256) {
        SpecialVertex cv{};
        SpecialEdge se{};
        int64_t kv = 0, hv = 0;
        if (i < n[0]) cv = p.seg.cand[b64 * p.seg.cap[0] + i];
        if (i < n[1]) se = p.seg.scell[b64 * p.seg.cap[1] + i];
        if (i < n[2]) kv = p.seg.kept[b64 * p.seg.cap[2] + i];
        if (i < n[3]) hv = p.seg.hole[b64 * p.seg.cap[3] + i];
        if (i < n[0]) p.sp->cand[off[0] + i] = cv;
        if (i < n[1]) p.sp->scell[off[1] + i] = se;
        if (i < n[2]) p.sp->kept[off[2] + i] = kv;
        if (i < n[3]) p.sp->hole[off[3] + i] = hv;
    }
    if (b == 0 && tid == 0) {
        unsigned long long raw[4], nd = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) raw[k] = tot[k][0] + tot[k][1] + tot[k][2] + tot[k][3];
#pragma unroll
        for (int w = 0; w < NW; ++w) nd = max(nd, needw[w]);
        p.sp->n_cand = raw[0];
        p.sp->n_scell = raw[1];
        p.sp->n_kept = raw[2];
        p.sp->n_hole = raw[3];
        p.sp->need = (int64_t) nd;
    }
}

__global__ void __launch_bounds__(256) k_sp_gather(McParams p) { sp_gather_seg(p, blockIdx.x); }

// Whether two gathered weld candidates share a position (as the fix-up
// groups them: equal coordinates): candidate indices go into an
// open-addressing table keyed by position.  With no shared position and no
// dropped face the fix-up has nothing to do.
// Duplicate check of the weld candidates: every candidate's position goes
// into the dslots hash; two candidates at one position set sp->dup (the weld
// fix-up is skipped when none do).  Candidates are read from their segments
// (index s * cap[0] + j), so the check does not wait for the gather: both
// run in one launch.
__device__ __forceinline__ void sp_dup_insert(const McParams &p, int32_t i) {
    const SpecialVertex *cv = p.seg.cand;
    const SpecialVertex v = cv[i];
    const uint64_t mask = (uint64_t) p.dmask;
    const uint64_t h = f3_hash(v.x, v.y, v.z) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const int64_t slot = (int64_t) ((h + probe) & mask);
        int32_t cur = p.dslots[slot];
        if (cur < 0) {
            const int32_t prev = atomicCAS(&p.dslots[slot], -1, i);
            if (prev < 0) break;
            cur = prev;
        }
        const SpecialVertex o = cv[cur];
        if (o.x == v.x && o.y == v.y && o.z == v.z) {
            atomicExch((unsigned long long *) &p.sp->dup, 1ull);
            break;
        }
    }
}

// Workgroups [0, nseg) gather segment b of the special lists; the others
// check the candidates for duplicate positions, segment by segment.
__global__ void __launch_bounds__(256) k_sp_gather_dup(McParams p) {
    const int nseg = p.seg.nseg;
    if ((int) blockIdx.x < nseg) {
        sp_gather_seg(p, blockIdx.x);
        return;
    }
    const int nb = (int) gridDim.x - nseg, b = (int) blockIdx.x - nseg;
    const int64_t cap0 = p.seg.cap[0];
    for (int sg = b; sg < nseg; sg += nb) {
        const int64_t n = min((int64_t) p.seg.ctr[(int64_t) sg * kSpStride], cap0);
        for (int64_t j = threadIdx.x; j < n; j += 256) sp_dup_insert(p, (int32_t) ((int64_t) sg * cap0 + j));
    }
}

// ---- ragged batches ----------------------------------------------------------------
// Every kernel walks the batch-wide index space of its stage (classify words,
// count blocks, emit tiles) and finds the owning item by binary search over
// the items' start offsets, staged in LDS; the per-item work is the
// single-grid code.

// Last item whose start is <= x (starts staged in LDS; at most kMaxBatch items).
__device__ __forceinline__ int find_start(const int64_t *st, int n, int64_t x) {
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (st[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Clears a batch's state (tile flags and counts, every item's special-list
// counters, error flag, totals, program block-list count and list
// descriptor); thread t0 of a grid of `stride` threads.  Nothing the batch
// classify does reads it, so the classify kernel runs this first.
__device__ __forceinline__ void reset_batch_state(const McBatchItem *items, const McBatchHdr *hdr, int64_t t0,
                                                  int64_t stride) {
    const int n = (int) hdr->n;
    if (hdr->tile_flag) {
        for (int64_t t = t0; t < hdr->ntiles; t += stride) hdr->tile_flag[t] = 0;
        if (t0 < kMaxTileConfigs) hdr->tile_count[t0] = 0;
    }
    for (int64_t t = t0; t < (int64_t) n * 4 * kSpSeg; t += stride) {
        const int i = (int) (t / (4 * kSpSeg)), c = (int) (t % (4 * kSpSeg));
        const McParams &p = items[i].p;
        if (c < 4 * p.seg.nseg) p.seg.ctr[c * kSpStride] = 0;
        if (c == 0) {
            *p.err = 0;
            p.totals[0] = p.totals[1] = 0;
            if (p.csg && p.X >= 2 && p.Y >= 2 && p.Z >= 2) p.cnt_f[0] = 0;   // the program classify's list count
            SpecialLists init = items[i].sp_init;
            init.n_cand = init.n_scell = init.n_kept = init.n_hole = 0;
            init.need = 0;
            init.dup = -1;
            *p.sp = init;
        }
    }
}

// Flat classify tasks over the batch (each grid owns a range of whole
// tasks), after the batch's state is cleared.
__global__ void __launch_bounds__(256) k_classify_batch(const McBatchItem *__restrict__ items,
                                                        const McBatchHdr *__restrict__ hdr) {
    __shared__ int64_t starts[kMaxBatch];
    __shared__ uint64_t ballots[4][kFlatChunks + 1];
    reset_batch_state(items, hdr, (int64_t) blockIdx.x * 256 + threadIdx.x, (int64_t) gridDim.x * 256);
    const int n = (int) hdr->n;
    const int64_t ntask = hdr->nctask;
    for (int i = threadIdx.x; i < n; i += 256) starts[i] = items[i].ctask0;
    __syncthreads();
    const int64_t wave = __builtin_amdgcn_readfirstlane((int) (blockIdx.x * 4 + (threadIdx.x >> 6)));
    const int64_t nwave = (int64_t) gridDim.x * 4;
    for (int64_t t = wave; t < ntask; t += nwave) {
        const int i = __builtin_amdgcn_readfirstlane(find_start(starts, n, t));
        const McParams &p = items[i].p;
        classify_flat_task(p.values, p.X * p.Y * p.Z, p.Z, p.W, p.level, t - items[i].ctask0, p.bits,
                           ballots[threadIdx.x >> 6]);
    }
}

// Program items of a batch, phase 1: the items' block ranges are
// consecutive and padded to 256 blocks (gtask0 non-decreasing in item
// order; items without a program own none), so a workgroup's 256 blocks
// belong to one item; the program is staged again when the item changes.
__global__ void __launch_bounds__(256) k_csg_blocks_batch(const McBatchItem *__restrict__ items,
                                                         const McBatchHdr *__restrict__ hdr) {
    __shared__ CsgProgram P;
    __shared__ int64_t starts[kMaxBatch];
    const int n = (int) hdr->n;
    for (int i = threadIdx.x; i < n; i += 256) starts[i] = items[i].gtask0;
    __syncthreads();
    int cur = -1;
    double cs = 0.0;
    for (int64_t base = (int64_t) blockIdx.x * 256; base < hdr->ngtask; base += (int64_t) gridDim.x * 256) {
        const int i = __builtin_amdgcn_readfirstlane(find_start(starts, n, base));
        const McParams &p = items[i].p;
        if (i != cur) {
            __syncthreads();
            csg_stage(P, p.csg);
            __syncthreads();
            cur = i;
            cs = csg_coord_scale(p);
        }
        const int64_t b = base - items[i].gtask0 + threadIdx.x;
        csg_test_block(p, P, b, b < csg_classify_tasks(p.X, p.Y, p.W), cs);
    }
}

// Phase 2 of every program item of a batch by NW-wave workgroups.  The
// items' phase-2 blocks (listed undecided blocks, or every block of an item
// without a list) form one concatenated range that the workgroups walk with
// a grid stride (neighbouring blocks, often of similar cost, go to different
// workgroups); the item only grows along a workgroup's walk, so it stages
// each program at most once.  A batch with one program item takes
// four-wave groups (fused DC at 256^3: 0.78 vs 0.89 ms with one wave); a
// batch of several takes one-wave groups, which measured faster under the
// server's concurrent batches (program requests: 363 vs 538 ms per pass).
template <int NW>
__global__ void __launch_bounds__(64 * NW) k_csg_fill_batch(const McBatchItem *__restrict__ items,
                                                           const McBatchHdr *__restrict__ hdr) {
    constexpr int NT = 64 * NW;
    __shared__ CsgProgram P;
    __shared__ uint64_t wacc[18];
    __shared__ uint32_t work[kMaxBatch + 1];
    __shared__ int64_t wstart[kMaxBatch + 1];
    __shared__ uint32_t tsum[NT / 64];
    const int n = (int) hdr->n;
    for (int i = threadIdx.x; i < kMaxBatch; i += NT) {
        uint32_t w = 0;
        if (i < n) {
            const McParams &p = items[i].p;
            if (p.csg && p.X >= 2 && p.Y >= 2 && p.Z >= 2)
                w = p.csg_list ? *(volatile uint32_t *) p.cnt_f : (uint32_t) csg_classify_tasks(p.X, p.Y, p.W);
        }
        work[i] = w;
    }
    __syncthreads();
    block_exscan_lds<kMaxBatch, NT>(work, tsum);
    for (int i = threadIdx.x; i <= n; i += NT) wstart[i] = (int64_t) work[i];
    __syncthreads();
    const int64_t total = wstart[n];
    int cur = -1;
    double cs = 0.0;
    for (int64_t g = blockIdx.x; g < total; g += gridDim.x) {
        const int i = __builtin_amdgcn_readfirstlane(find_start(wstart, n, g));
        const McParams &p = items[i].p;
        if (i != cur) {
            __syncthreads();
            csg_stage(P, p.csg);
            __syncthreads();
            cur = i;
            cs = csg_coord_scale(p);
        }
        const int64_t k = g - wstart[i];
        if (p.csg_list)
            csg_fill_block<NW>(p, P, (int64_t) p.cnt_v[k], p.cnt_f[1 + k], cs, wacc);
        else if (csg_block_sign(p, P, csg_block(p, k), cs) == 0)
            csg_fill_block<NW>(p, P, k, 0u, cs, wacc);
    }
}

// One workgroup per count block (grid = the batch's block count); the
// item comes from a search over the item starts staged in LDS.
template <int NT>
__global__ void __launch_bounds__(NT) k_count_batch(const McBatchItem *__restrict__ items,
                                                    const McBatchHdr *__restrict__ hdr) {
    __shared__ int64_t starts[kMaxBatch];
    const int n = (int) hdr->n;
    const int64_t gb = blockIdx.x;
    if (gb >= hdr->nblk) return;
    for (int i = threadIdx.x; i < n; i += NT) starts[i] = items[i].blk0;
    __syncthreads();
    const int i = __builtin_amdgcn_readfirstlane(find_start(starts, n, gb));
    count_block<NT>(items[i].p, gb - items[i].blk0);
}

__global__ void __launch_bounds__(1024) k_scan_top_batch(const McBatchItem *__restrict__ items,
                                                         const McBatchHdr *__restrict__ hdr) {
    for (int64_t i = blockIdx.x; i < hdr->n; i += gridDim.x) {
        const McParams p = items[i].p;
        __syncthreads();
        scan_top(p);
    }
}

// One workgroup per emit tile of the items using this configuration (tiles
// [t_begin, t_begin + grid)); items are found in tile order.
// SRC kSrcGrid takes the stored-grid items, kSrcProgram the program items;
// each skips the other kind's tiles.
template <int TX, int TY, int NZC, int NT, int SRC>
__global__ void __launch_bounds__(NT) k_emit_batch(const McBatchItem *__restrict__ items,
                                                   const int32_t *__restrict__ order,
                                                   const McBatchHdr *__restrict__ hdr, int64_t t_begin,
                                                   const int *__restrict__ tlist, const int *__restrict__ tcount) {
    __shared__ int64_t starts[kMaxBatch];
    const int n = (int) hdr->n;
    int64_t gt = t_begin + blockIdx.x;
    if (tlist) {   // this configuration's active tiles only
        if ((int) blockIdx.x >= *tcount) return;
        gt = tlist[blockIdx.x];
    }
    if (gt >= hdr->ntiles) return;
    for (int k = threadIdx.x; k < n; k += NT) starts[k] = items[order[k]].tile0;
    __syncthreads();
    const int i = order[__builtin_amdgcn_readfirstlane(find_start(starts, n, gt))];
    if ((items[i].p.csg != nullptr) != (SRC == kSrcProgram)) return;
    const int64_t t = gt - items[i].tile0;
    emit_tile<TX, TY, NZC, NT, SRC>(items[i].p, t, (int) (t % items[i].p.seg.nseg));
}

// grid (segments, items)
__global__ void __launch_bounds__(256) k_sp_gather_batch(const McBatchItem *items, const McBatchHdr *hdr) {
    const int64_t i = blockIdx.y;
    if (i >= hdr->n || (int) blockIdx.x >= items[i].p.seg.nseg) return;
    sp_gather_seg(items[i].p, blockIdx.x);
}

__global__ void k_batch_summary(const McBatchItem *items, const McBatchHdr *hdr, McBatchSummary *out) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < hdr->n; i += (int64_t) gridDim.x * blockDim.x) {
        const McParams &p = items[i].p;
        McBatchSummary s;
        s.err = *p.err;
        s.nv = p.totals[0];
        s.nf = p.totals[1];
        s.sp = *p.sp;
        out[i] = s;
    }
}

// ---- intersections of a dense grid -------------------------------------------------
// Entry order of get_intersection (src/its.cu:93-159): active cells by
// ascending cell id, each cell's crossing edges by edge index.  Cell ids run
// over cell rows (x, y) then z, so the (point row, z-segment) entries of the
// count pass (p.its = 1: active cells, crossing edges) give every active
// cell's index and first entry by a prefix sum plus in-segment popcounts.

template <int TX, int TY, int NZC, int NT>
struct ItsLds {
    static constexpr int PX = TX + 1, PY = TY + 1, PC = NZC + 1;
    static constexpr int NUNITS = TX * TY * NZC;   // (cell row, chunk), chunk fastest
    uint64_t w[PX * PY * PC];
    uint64_t amask[NUNITS];   // active cells of each unit
    int64_t abase[TX * TY];   // active index of the row's first active cell (this z-segment)
    int64_t ebase[TX * TY];   // first entry of the row's cells
    uint32_t na[NUNITS + 1];  // active cells per unit -> exclusive offsets
    int32_t unit[NT];         // batch: unit of each slot
    uint32_t esc[NT + 1];     // batch: entries per slot -> exclusive offsets
    uint32_t tsum[NT / 64];
    int32_t carry[2];         // entries of the straddling row's cells before a batch
    uint8_t ncross[256];
    float cx[TX + 1], cy[TY + 1], cz[64 * NZC + 1];
};

template <int TX, int TY, int NZC, int NT>
__device__ __forceinline__ CellMasks its_unit_cells(const ItsLds<TX, TY, NZC, NT> &s, int cx, int cy, int j) {
    using L = ItsLds<TX, TY, NZC, NT>;
    auto W = [&](int px, int py, int jj) { return s.w[(px * L::PY + py) * L::PC + jj]; };
    return cell_masks(W(cx, cy, j), W(cx, cy, j + 1), W(cx, cy + 1, j), W(cx, cy + 1, j + 1), W(cx + 1, cy, j),
                      W(cx + 1, cy, j + 1), W(cx + 1, cy + 1, j), W(cx + 1, cy + 1, j + 1));
}

// Same tiling and slot machinery as the marching-cubes emit: active index of
// a cell = its row's base + its rank in the row, first entry = the row's base
// + a block scan of crossing counts (with a carry for the row that straddles
// two batches).
template <int TX, int TY, int NZC, int NT>
__device__ __forceinline__ void its_tile(const McParams &p, const ItsOut &o, int64_t tile) {
    using L = ItsLds<TX, TY, NZC, NT>;
    __shared__ L s;
    const int tid = threadIdx.x;
    const int64_t tk = tile % p.NK, tj = (tile / p.NK) % p.NJ, ti = tile / (p.NK * p.NJ);
    const int64_t x0 = ti * TX, y0 = tj * TY, c0 = tk * NZC, z0 = 64 * c0;
    const int nx = (int) min<int64_t>(TX, p.X - 1 - x0);
    const int ny = (int) min<int64_t>(TY, p.Y - 1 - y0);

    for (int c = tid; c < 256; c += NT) s.ncross[c] = (uint8_t) __popc(crossing_status(c));
    for (int i = tid; i < L::PX * L::PY * L::PC; i += NT) {
        const int j = i % L::PC, r = i / L::PC, py = r % L::PY, px = r / L::PY;
        const int64_t gx = x0 + px, gy = y0 + py, c = c0 + j;
        s.w[i] = (gx < p.X && gy < p.Y && c < p.W) ? p.bits[(gx * p.Y + gy) * p.W + c] : 0ull;
    }
    for (int c = tid; c < TX + TY + 64 * NZC + 3; c += NT) {
        if (c <= TX) {
            s.cx[c] = grid_coord(x0 + c, p.X, p.amin[0], p.amax[0]);
        } else if (c <= TX + TY + 1) {
            s.cy[c - TX - 1] = grid_coord(y0 + (c - TX - 1), p.Y, p.amin[1], p.amax[1]);
        } else {
            s.cz[c - TX - TY - 2] = grid_coord(z0 + (c - TX - TY - 2), p.Z, p.amin[2], p.amax[2]);
        }
    }
    for (int i = tid; i < TX * TY; i += NT) {
        const int cx = i / TY, cy = i % TY;
        s.abase[i] = s.ebase[i] = 0;
        if (cx >= nx || cy >= ny) continue;
        const int64_t e = ((x0 + cx) * p.Y + (y0 + cy)) * p.S + tk;
        s.abase[i] = p.blk_v[e / kMcScanBlock] + (int64_t) p.cnt_v[e];
        s.ebase[i] = p.blk_f[e / kMcScanBlock] + (int64_t) p.cnt_f[e];
    }
    __syncthreads();
    for (int u = tid; u < L::NUNITS; u += NT) {
        const int j = u % NZC, r = u / NZC, cy = r % TY, cx = r / TY;
        const int64_t zlo = z0 + 64 * j;
        uint64_t act = 0;
        if (cx < nx && cy < ny && zlo < p.Z - 1)
            act = cell_active(its_unit_cells(s, cx, cy, j)) & low_bits(p.Z - 1 - zlo);
        s.amask[u] = act;
        s.na[u] = (uint32_t) __popcll(act);
    }
    __syncthreads();
    block_exscan_lds<L::NUNITS, NT>(s.na, s.tsum);
    if (o.amask_tab) {
        // per (cell row, word): active cells and the active index of the
        // first of them, for the neighbour lookups of the quads
        for (int u = tid; u < L::NUNITS; u += NT) {
            const int j = u % NZC, r = u / NZC, cy = r % TY, cx = r / TY;
            const int64_t c = c0 + j;
            if (cx >= nx || cy >= ny || 64 * c >= p.Z - 1) continue;
            const int64_t t = ((x0 + cx) * (p.Y - 1) + (y0 + cy)) * p.W + c;
            o.amask_tab[t] = s.amask[u];
            o.apre_tab[t] = (int32_t) (s.abase[r] + (int64_t) (s.na[u] - s.na[r * NZC]) + o.a_base);
        }
    }
    const int nactive = (int) s.na[L::NUNITS];
    if (tid == 0) s.carry[0] = 0;
    for (int j0 = 0, bt = 0; j0 < nactive; j0 += NT, bt ^= 1) {
        s.unit[tid] = -1;
        __syncthreads();
        for (int u = tid; u < L::NUNITS; u += NT) {
            const int a0 = (int) s.na[u], a1 = (int) s.na[u + 1];
            if (a1 > a0 && a0 < j0 + NT && a1 > j0) s.unit[max(a0, j0) - j0] = u;
        }
        __syncthreads();
        block_maxscan_lds<NT>(s.unit, s.tsum);
        const int jj = j0 + tid;
        const bool live = jj < nactive;
        int u = 0, k = 0, cs = 0;
        float cv[8];   // corner values, loaded before the entry scan so that the scan hides their latency
        if (live) {
            u = s.unit[tid];
            k = nth_set_bit(s.amask[u], jj - (int) s.na[u]);
            const int jc = u % NZC, r = u / NZC;
            cs = cell_case(its_unit_cells(s, r / TY, r % TY, jc), k);
        }
        if (live && o.points) {
            const int jc = u % NZC, r = u / NZC;
            const float *r00 = p.values + ((x0 + r / TY) * p.Y + (y0 + r % TY)) * p.Z + (z0 + 64 * jc + k);
            const float *r01 = r00 + p.Z;
            const float *r10 = r00 + p.Y * p.Z;
            const float *r11 = r10 + p.Z;
            cv[0] = r00[0];
            cv[1] = r00[1];
            cv[2] = r01[0];
            cv[3] = r01[1];
            cv[4] = r10[0];
            cv[5] = r10[1];
            cv[6] = r11[0];
            cv[7] = r11[1];
        }
        s.esc[tid] = live ? (uint32_t) s.ncross[cs] : 0u;
        __syncthreads();
        block_exscan_lds<NT, NT>(s.esc, s.tsum);
        if (tid == NT - 1 && j0 + NT < nactive) {
            const int rl = s.unit[NT - 1] / NZC;
            const int rs = (int) s.na[rl * NZC], re = (int) s.na[(rl + 1) * NZC];
            s.carry[bt ^ 1] = re > j0 + NT ? (rs < j0 ? s.carry[bt] : 0) + (int) (s.esc[NT] - s.esc[max(rs - j0, 0)])
                                           : 0;
        }
        if (live) {
            const int jc = u % NZC, r = u / NZC, cy = r % TY, cx = r / TY;
            const int rs = (int) s.na[r * NZC];
            const int64_t a = s.abase[r] + (jj - rs);
            int64_t e = s.ebase[r] + (int64_t) (s.esc[tid] - s.esc[max(rs - j0, 0)]) + (rs < j0 ? s.carry[bt] : 0);
            const int zl = 64 * jc + k;
            const int64_t gx = x0 + cx, gy = y0 + cy, gz = z0 + zl;
            const int64_t id = (gx * (p.Y - 1) + gy) * (p.Z - 1) + gz;
            if (a < p.cap_v) {   // capacities from the count pass (defensive)
                if (o.cell_indices) o.cell_indices[a] = id;
                o.cell_dense[a] = id;
                if (o.cell_offsets) o.cell_offsets[a] = e + o.e_base;
            }
            if (a < p.cap_v && o.points) {
                const float lo[3] = {s.cx[cx], s.cy[cy], s.cz[zl]};
                const float hi[3] = {s.cx[cx + 1], s.cy[cy + 1], s.cz[zl + 1]};
                const int st = crossing_status(cs);
                for_edges([&](auto E) {
                    if (!((st >> E) & 1) || e >= p.cap_f) return;
                    constexpr int ca = edge_a(E), cb = edge_b(E);
                    const float t = edge_t(cv[ca], cv[cb], p.level);
                    const float px = lerp1(t, (ca >> 2) ? hi[0] : lo[0], (cb >> 2) ? hi[0] : lo[0]);
                    const float py = lerp1(t, ((ca >> 1) & 1) ? hi[1] : lo[1], ((cb >> 1) & 1) ? hi[1] : lo[1]);
                    const float pz = lerp1(t, (ca & 1) ? hi[2] : lo[2], (cb & 1) ? hi[2] : lo[2]);
                    *reinterpret_cast<float3 *>(o.points + 3 * e) = make_float3(px, py, pz);
                    o.edge_local[e] = (uint8_t) E;
                    o.is_out[e] = cv[ca] <= cv[cb] ? 1 : 0;
                    o.entry_cell[e] = (int32_t) (a + o.a_base);
                    if (o.normals) {
                        float nxv, nyv, nzv;
                        cell_normal(lo, hi, cv, px, py, pz, nxv, nyv, nzv);
                        *reinterpret_cast<float3 *>(o.normals + 3 * e) = make_float3(nxv, nyv, nzv);
                    }
                    ++e;
                });
            }
        }
        __syncthreads();
    }
}

template <int TX, int TY, int NZC, int NT>
__global__ void __launch_bounds__(NT) k_its_tiles(McParams p, ItsOut o) {
    const int64_t tile = emit_block_tile(p);
    if (tile >= 0) its_tile<TX, TY, NZC, NT>(p, o, tile);
}

// Batch: one workgroup per intersection tile of the grids using this
// configuration, grids found in tile order (as the marching-cubes emit).
template <int TX, int TY, int NZC, int NT>
__global__ void __launch_bounds__(NT) k_its_tiles_batch(const McBatchItem *__restrict__ items,
                                                        const int32_t *__restrict__ order,
                                                        const ItsOut *__restrict__ outs,
                                                        const McBatchHdr *__restrict__ hdr, int64_t t_begin,
                                                        const int *__restrict__ tlist, const int *__restrict__ tcount) {
    __shared__ int64_t starts[kMaxBatch];
    const int n = (int) hdr->n;
    int64_t gt = t_begin + blockIdx.x;
    if (tlist) {   // this configuration's tiles with an active cell only
        if ((int) blockIdx.x >= *tcount) return;
        gt = tlist[blockIdx.x];
    }
    if (gt >= hdr->ntiles) return;
    for (int k = threadIdx.x; k < n; k += NT) starts[k] = items[order[k]].tile0;
    __syncthreads();
    const int i = order[__builtin_amdgcn_readfirstlane(find_start(starts, n, gt))];
    its_tile<TX, TY, NZC, NT>(items[i].p, outs[i], gt - items[i].tile0);
}

__global__ void k_its_reset(McParams p) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        *p.err = 0;
        p.totals[0] = p.totals[1] = 0;
        if (p.tile_count) *p.tile_count = 0;
    }
    if (p.tile_flag) {
        const int64_t ntiles = p.NI * p.NJ * p.NK;
        for (int64_t i = t; i < ntiles; i += (int64_t) gridDim.x * blockDim.x) p.tile_flag[i] = 0;
    }
}

// ---- host side ---------------------------------------------------------------

// Emit tile configurations the autotuner may choose from: (cell rows along
// x, along y, 64-cell chunks along z, threads).
#define ISO_MC_CONFIGS(X)   \
    X(0, 16, 16, 4, 256)    \
    X(1, 8, 8, 2, 256)      \
    X(2, 16, 16, 2, 256)    \
    X(3, 16, 16, 1, 256)    \
    X(4, 16, 16, 8, 256)    \
    X(5, 8, 8, 1, 256)      \
    X(6, 16, 16, 1, 256)

int mc_config_count() {
    int n = 0;
#define X(id, tx, ty, nzc, nt) ++n;
    ISO_MC_CONFIGS(X)
#undef X
    return n;
}

bool mc_config_geom(int cfg, int &tx, int &ty, int &nzc) {
    switch (cfg) {
#define X(id, tx_, ty_, nzc_, nt) \
    case id:                      \
        tx = tx_;                 \
        ty = ty_;                 \
        nzc = nzc_;               \
        return true;
        ISO_MC_CONFIGS(X)
#undef X
    default:
        return false;
    }
}

int64_t mc_config_tiles(int cfg, const int64_t shape[3]) {
    int tx, ty, nzc;
    if (!mc_config_geom(cfg, tx, ty, nzc) || shape[0] < 2 || shape[1] < 2 || shape[2] < 2) return 0;
    const int64_t cell_chunks = (shape[2] - 1 + 63) / 64;
    return ((shape[0] - 1 + tx - 1) / tx) * ((shape[1] - 1 + ty - 1) / ty) * ((cell_chunks + nzc - 1) / nzc);
}

// First tile in the order that still gives every CU several workgroups
// (16x16x64 cells from about 384^3 up, 8x8x128 near 256^3); when none
// does, the one with the most tiles.  With empty tiles skipped through the
// active-tile list, 16x16x64 extracted fastest at 384^3..1024^3 on MI355X
// (0.5-8% ahead of 16x16x128, which never has more tiles and so is not in
// the order); 16x16x256 never wins.
int mc_auto_config(const int64_t shape[3]) {
    static const int order[] = {3, 1, 5};
    constexpr int64_t kWant = 2048;
    int best = 0;
    int64_t best_tiles = -1;
    for (int cfg : order) {
        const int64_t t = mc_config_tiles(cfg, shape);
        if (t >= kWant) return cfg;
        if (t > best_tiles) {
            best_tiles = t;
            best = cfg;
        }
    }
    return best;
}

hipError_t mc_upload_luts() {
    // once per device (each GPU holds its own copy of the constant tables);
    // host threads that call first together wait for one upload
    constexpr int kMaxDevices = 64;
    static std::atomic<bool> done[kMaxDevices];
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (done[dev].load(std::memory_order_acquire)) return hipSuccess;
    std::lock_guard<std::mutex> lock(mu);
    if (done[dev].load(std::memory_order_relaxed)) return hipSuccess;
    uint64_t h[2][256];
    uint8_t n[2][256];
    for (int i = 0; i < 256; ++i) {
        h[0][i] = kNagaeTris[i];
        h[1][i] = kLorensenTris[i];
        n[0][i] = (uint8_t) (h[0][i] & 15);
        n[1][i] = (uint8_t) (h[1][i] & 15);
    }
    e = hipMemcpyToSymbol(HIP_SYMBOL(cTris), h, sizeof(h));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cNtri), n, sizeof(n));
    if (e == hipSuccess) done[dev].store(true, std::memory_order_release);
    return e;
}

// Inside masks of a whole grid: dwordx4 loads when rows are whole aligned
// words, the flat kernel otherwise.
// reset: the stored-grid classify clears the extraction's state itself
// (returns true); program and sphere fields keep the separate reset kernel.
static bool launch_classify(const McParams &p, hipStream_t stream, const SpecialLists *reset = nullptr) {
    const int64_t nwords = p.X * p.Y * p.W;
    if (p.csg) {
        const int64_t nb = csg_classify_tasks(p.X, p.Y, p.W);
        const unsigned blocks = (unsigned) std::max<int64_t>(1, std::min<int64_t>((nb + 255) / 256, 8192));
        hipLaunchKernelGGL(k_csg_blocks, dim3(blocks), dim3(256), 0, stream, p);
        hipLaunchKernelGGL(k_csg_fill, dim3(kCsgFillGroups), dim3(64 * kCsgFillWaves), 0, stream, p);
    } else if (p.nsph > 0) {
        const int64_t ntask = ((p.X + 3) / 4) * ((p.Y + 3) / 4) * p.W;
        const unsigned blocks = (unsigned) std::max<int64_t>(1, std::min<int64_t>((ntask + 3) / 4, 8192));
        hipLaunchKernelGGL(k_classify_spheres, dim3(blocks), dim3(256), 0, stream, p);
    } else if (p.vec4 && p.Z % 64 == 0) {
        const unsigned blocks = (unsigned) std::max<int64_t>(1, std::min<int64_t>((nwords + 63) / 64, 8192));
        if (reset)
            hipLaunchKernelGGL(k_classify_v4<true>, dim3(blocks), dim3(256), 0, stream, p.values, nwords, p.level,
                               p.bits, p, *reset);
        else
            hipLaunchKernelGGL(k_classify_v4<false>, dim3(blocks), dim3(256), 0, stream, p.values, nwords, p.level,
                               p.bits, p, SpecialLists{});
        return reset != nullptr;
    } else {
        const int64_t N = p.X * p.Y * p.Z, ntask = (N + kFlatPts - 1) / kFlatPts;
        const unsigned blocks = (unsigned) std::max<int64_t>(1, std::min<int64_t>((ntask + 3) / 4, 8192));
        if (reset)
            hipLaunchKernelGGL(k_classify<true>, dim3(blocks), dim3(256), 0, stream, p.values, N, p.Z, p.W, p.level,
                               p.bits, p, *reset);
        else
            hipLaunchKernelGGL(k_classify<false>, dim3(blocks), dim3(256), 0, stream, p.values, N, p.Z, p.W, p.level,
                               p.bits, p, SpecialLists{});
        return reset != nullptr;
    }
    return false;
}

void mc_params_divisors(McParams &p) {
    p.narrow = p.X > 0 && p.Y > 0 && p.W > 0 && p.nzc > 0 && p.X * p.Y * p.W <= (int64_t) UINT32_MAX;
    if (p.narrow) {
        p.div_w = UDiv32((uint32_t) p.W);
        p.div_y = UDiv32((uint32_t) p.Y);
        p.div_zc = UDiv32((uint32_t) p.nzc);
        p.div_s = UDiv32((uint32_t) p.S);
    }
    // the active-tile list needs narrow indices and tile ids below 2^31
    if (p.tile_list && p.tx > 0 && p.ty > 0 && p.narrow && p.tile_base >= 0 &&
        p.tile_base + p.NI * p.NJ * p.NK <= (int64_t) INT32_MAX) {
        p.div_tx = UDiv32((uint32_t) p.tx);
        p.div_ty = UDiv32((uint32_t) p.ty);
    } else {
        p.tile_flag = p.tile_list = p.tile_count = nullptr;
    }
}

hipError_t mc_dense_launch(int cfg, McParams &p, const SpecialLists &sp_init, hipStream_t stream,
                           hipEvent_t *ev) {
    mc_params_divisors(p);
    hipError_t e = mc_upload_luts();
    if (e != hipSuccess) return e;
    auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], stream) : hipSuccess; };
    if ((e = mark(0)) != hipSuccess) return e;
    // a stored grid's classify kernel clears the state itself; program and
    // sphere fields, whose classify counts into it, take the reset kernel
    const bool grid_field = !p.csg && p.nsph <= 0;
    if (!grid_field) {
        const int64_t ntiles = p.NI * p.NJ * p.NK;
        const unsigned rblocks = (unsigned) std::max<int64_t>(
            (4 * kSpSeg + 255) / 256, p.tile_flag ? std::min<int64_t>((ntiles + 255) / 256, 1024) : 0);
        hipLaunchKernelGGL(mc_dense_reset, dim3(rblocks), dim3(256), 0, stream, p, sp_init);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }

    launch_classify(p, stream, grid_field ? &sp_init : nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(1)) != hipSuccess) return e;

    if (p.nblk < kCountWideBelow)
        hipLaunchKernelGGL(k_count<1024>, dim3((unsigned) p.nblk), dim3(1024), 0, stream, p);
    else
        hipLaunchKernelGGL(k_count<256>, dim3((unsigned) p.nblk), dim3(256), 0, stream, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(2)) != hipSuccess) return e;
    // up to 256 count blocks the emit workgroups scan the block totals
    // themselves: one launch fewer on small grids
    p.scan_in_emit = p.nblk <= 256 ? 1 : 0;
    if (!p.scan_in_emit) {
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, p);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = mark(3)) != hipSuccess) return e;

    const unsigned tiles = (unsigned) (p.NI * p.NJ * p.NK);
    switch (cfg) {
#define X(id, tx, ty, nzc, nt)                                                                     \
    case id:                                                                                       \
        if (p.csg)                                                                                 \
            hipLaunchKernelGGL((k_emit_program<tx, ty, nzc, nt>), dim3(tiles), dim3(nt), 0, stream, p); \
        else if (p.nsph > 0)                                                                       \
            hipLaunchKernelGGL((k_emit_spheres<tx, ty, nzc, nt>), dim3(tiles), dim3(nt), 0, stream, p); \
        else                                                                                       \
            hipLaunchKernelGGL((k_emit<tx, ty, nzc, nt>), dim3(tiles), dim3(nt), 0, stream, p);    \
        break;
        ISO_MC_CONFIGS(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(4)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sp_gather_dup, dim3((unsigned) p.seg.nseg + (p.dslots ? 64u : 0u)), dim3(256), 0, stream,
                       p);
    return hipGetLastError();
}

hipError_t its_dense_count_launch(McParams &p, hipStream_t stream) {
    mc_params_divisors(p);
    hipError_t e = mc_upload_luts();
    if (e != hipSuccess) return e;
    // a stored grid's classify clears the state itself (reset_state covers
    // what k_its_reset clears)
    const bool grid_field = !p.csg && p.nsph <= 0;
    if (!grid_field) {
        const unsigned rblocks = p.tile_flag ? (unsigned) std::max<int64_t>(
                                                   1, std::min<int64_t>((p.NI * p.NJ * p.NK + 255) / 256, 1024))
                                             : 1;
        hipLaunchKernelGGL(k_its_reset, dim3(rblocks), dim3(256), 0, stream, p);
    }
    const SpecialLists none{};
    launch_classify(p, stream, grid_field ? &none : nullptr);
    if (p.nblk < kCountWideBelow)
        hipLaunchKernelGGL(k_count<1024>, dim3((unsigned) p.nblk), dim3(1024), 0, stream, p);
    else
        hipLaunchKernelGGL(k_count<256>, dim3((unsigned) p.nblk), dim3(256), 0, stream, p);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, stream, p);
    return hipGetLastError();
}

hipError_t its_dense_emit_launch(int cfg, McParams &p, const ItsOut &o, hipStream_t stream) {
    mc_params_divisors(p);
    const unsigned tiles = (unsigned) (p.NI * p.NJ * p.NK);
    switch (cfg) {
#define X(id, tx, ty, nzc, nt)                                                                        \
    case id:                                                                                          \
        hipLaunchKernelGGL((k_its_tiles<tx, ty, nzc, nt>), dim3(tiles), dim3(nt), 0, stream, p, o);   \
        break;
        ISO_MC_CONFIGS(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t mc_batch_count_launch(const McBatchItem *items, const McBatchHdr *hdr, const McBatchHdr &h,
                                 McBatchSummary *summary, hipStream_t stream) {
    hipError_t e = mc_upload_luts();
    if (e != hipSuccess) return e;
    // the classify kernel clears the batch's state first (no reset launch)
    const unsigned cb = (unsigned) std::max<int64_t>(1, std::min<int64_t>((h.nctask + 3) / 4, 8192));
    hipLaunchKernelGGL(k_classify_batch, dim3(cb), dim3(256), 0, stream, items, hdr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (h.ngtask > 0) {
        const unsigned gb = (unsigned) std::min<int64_t>((h.ngtask + 255) / 256, 8192);
        hipLaunchKernelGGL(k_csg_blocks_batch, dim3(gb), dim3(256), 0, stream, items, hdr);
        if (h.n_prog == 1)
            hipLaunchKernelGGL(k_csg_fill_batch<kCsgFillWaves>, dim3(kCsgFillGroups), dim3(64 * kCsgFillWaves), 0,
                               stream, items, hdr);
        else
            hipLaunchKernelGGL(k_csg_fill_batch<1>, dim3(kCsgFillBatchGroups), dim3(64), 0, stream, items, hdr);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const unsigned nb = (unsigned) std::max<int64_t>(1, h.nblk);
    if (h.nblk < kCountWideBelow)
        hipLaunchKernelGGL(k_count_batch<1024>, dim3(nb), dim3(1024), 0, stream, items, hdr);
    else
        hipLaunchKernelGGL(k_count_batch<256>, dim3(nb), dim3(256), 0, stream, items, hdr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const unsigned sb = (unsigned) std::max<int64_t>(1, std::min<int64_t>(h.n, 256));
    hipLaunchKernelGGL(k_scan_top_batch, dim3(sb), dim3(1024), 0, stream, items, hdr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_summary, dim3((unsigned) ((h.n + 255) / 256)), dim3(256), 0, stream, items, hdr,
                       summary);
    return hipGetLastError();
}

hipError_t its_batch_emit_launch(const int64_t *tile_begin, const int64_t *tile_end, const McBatchItem *items,
                                 const int32_t *order, const ItsOut *outs, const McBatchHdr *hdr, hipStream_t stream,
                                 const int *tile_list, const int *tile_count) {
    for (int c = 0; c < mc_config_count(); ++c) {
        const int64_t nt = tile_end[c] - tile_begin[c];
        if (nt <= 0) continue;
        switch (c) {
#define X(id, tx, ty, nzc, nth)                                                                                  \
    case id:                                                                                                     \
        hipLaunchKernelGGL((k_its_tiles_batch<tx, ty, nzc, nth>), dim3((unsigned) nt), dim3(nth), 0, stream, items, \
                           order, outs, hdr, tile_begin[c], tile_list ? tile_list + tile_begin[c] : nullptr,    \
                           tile_list ? tile_count + c : nullptr);                                            \
        break;
            ISO_MC_CONFIGS(X)
#undef X
        default:
            return hipErrorInvalidValue;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t mc_batch_emit_launch(const int64_t *tile_begin, const int64_t *tile_end, const McBatchItem *items,
                                const int32_t *order, const McBatchHdr *hdr, const McBatchHdr &h,
                                McBatchSummary *summary, hipStream_t stream, const int *tile_list,
                                const int *tile_count) {
    hipError_t e = hipSuccess;
    for (int c = 0; c < mc_config_count(); ++c) {
        const int64_t nt = tile_end[c] - tile_begin[c];
        if (nt <= 0) continue;
        switch (c) {
#define X(id, tx, ty, nzc, nth)                                                                              \
    case id:                                                                                                 \
        if (h.n_prog < h.n)                                                                                  \
            hipLaunchKernelGGL((k_emit_batch<tx, ty, nzc, nth, kSrcGrid>), dim3((unsigned) nt), dim3(nth), 0,  \
                               stream, items, order, hdr, tile_begin[c],                                     \
                               tile_list ? tile_list + tile_begin[c] : nullptr,                               \
                               tile_list ? tile_count + c : nullptr);                                        \
        if (h.n_prog > 0)                                                                                    \
            hipLaunchKernelGGL((k_emit_batch<tx, ty, nzc, nth, kSrcProgram>), dim3((unsigned) nt), dim3(nth), 0, \
                               stream, items, order, hdr, tile_begin[c],                                     \
                               tile_list ? tile_list + tile_begin[c] : nullptr,                               \
                               tile_list ? tile_count + c : nullptr);                                        \
        break;
            ISO_MC_CONFIGS(X)
#undef X
        default:
            return hipErrorInvalidValue;
        }
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_sp_gather_batch, dim3(kMaxSpSegBatch, (unsigned) h.n), dim3(256), 0, stream, items, hdr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_summary, dim3((unsigned) ((h.n + 255) / 256)), dim3(256), 0, stream, items, hdr,
                       summary);
    return hipGetLastError();
}

// ==== read-once pipeline ============================================================
// k_stream reads every value once.  Workgroup = region of kStrRX x-slabs x
// kStrRY point rows x one 64-point word; the regions are spread so that
// each XCD sweeps one contiguous part of the grid (the neighbouring words
// and rows whose halo values a region reads are then in flight on the same
// L2).  Wave k owns point rows [8k, 8k + 8) of the region and slides along
// x with the rows of the next slabs in flight.  Per slab it ballots the
// inside masks of its 9 rows (the 8 it owns and the one above) into a
// per-wave LDS table and checks whether the slab is all inside or all
// outside; only a slab that is not uniform with its successor does more.
// There one lane per row derives the row's crossing-edge and active-cell
// masks, and the slab's vertices and active cells are each handled
// compacted, one per lane: the vertex positions go to a segment of the
// arena (one atomic per slab), the cells' LUT triangle counts give each
// row's first face.  Vertex order in a wave: slab, row, axis (x < y < z),
// z; face order: slab, row, z, LUT slot.  At the end the wave writes its
// masks and per-row offsets region-major (str_entry) as whole lines.
// k_stream_scan1/2 turn the waves' counts into first ids, k_stream_flush
// moves the vertices from the arena to their ids (and flags the rows that
// own weld candidates), and k_faces emits the faces of the listed tiles
// from the masks and offsets alone.

// Region-major index of point row (x, y), word w.
__device__ __forceinline__ int64_t str_entry(const McParams &p, int64_t x, int64_t y, int64_t w) {
    const int64_t rx = x / kStrRX, s = x - rx * kStrRX;
    const int64_t ry = y / kStrRY, yy = y - ry * kStrRY;
    const int64_t r = (rx * p.NRY + ry) * p.W + w;
    return ((r * kStrWaves + (yy >> 3)) * kStrRX + s) * kStrRows + (yy & 7);
}
// (region, wave) of point row (x, y), word w: index into wb_v / wb_f.
__device__ __forceinline__ int64_t str_wave(const McParams &p, int64_t x, int64_t y, int64_t w) {
    const int64_t r = ((x / kStrRX) * p.NRY + y / kStrRY) * p.W + w;
    return r * kStrWaves + ((y % kStrRY) >> 3);
}
// Region of k_stream workgroup b: XCD b % 8 takes regions [k * per, (k + 1) * per).
__device__ __forceinline__ int64_t str_region(int64_t b, int64_t NR) {
    const int64_t per = (NR + 7) / 8;
    return (b & 7) * per + (b >> 3);
}

// Bit 31 of a vofs entry: the point row owns a weld-candidate vertex in the word.
constexpr uint32_t kStrCandFlag = 0x80000000u;

// An arena entry: a vertex position and where it came from.
struct StrStage {
    float x, y, z;
    uint32_t info;   // axis (bits 0-1), z (2-7), row (8-10), slab (11-15), candidate (16)
};

// One slab of the wave's rows: rows 0..8 at this lane's z, and in lanes
// 0..8 the z-halo value of row `lane` (the next word's first point).
struct StrSlab {
    float v[kStrRows + 1];
    float h;
};

// Per-wave LDS of k_stream.
struct StrLds {
    float v[2][kStrRows + 1][64];            // values of slabs st (slot st & 1) and st + 1
    float h[2][kStrRows + 1];                // their z-halo values
    uint64_t mt[kStrRX + 1][kStrRows + 1];   // inside masks: slabs 0..16 x rows 0..8
    uint32_t ht[kStrRX + 1];                 // z-halo inside bits of each slab (bit j: row j)
    uint32_t vt[kStrRX * kStrRows];          // first vertex of each (slab, row), wave-relative
    uint32_t ft[kStrRX * kStrRows];          // first face of each (slab, row), wave-relative
    uint64_t em[kStrRows][4];                // surface slab: x / y / z crossing edges, active cells per row
};
// Grid coordinates of a region's points.
struct StrCoords {
    float cx[kStrRX + 1];
    float cy[kStrRY + 1];
    float cz[65];
};

// Region geometry of the calling wave.
struct StrWave {
    int64_t r, w, rx, ry, x0, y0, z0;
    int nx, ny, nrow;
    bool zin, hz;
    uint64_t pv, ev;
    float cz0, cz1;
    const float *wbase;  // point (0, y0, z0)
    int64_t slab;        // points between x-slabs (Y Z)
    uint32_t rowbytes;   // bytes between point rows (4 Z)
    uint32_t loff;       // this lane's byte offset in a row (clamped to the grid)
    uint32_t hoff;       // this lane's halo value: row min(lane, nrow - 1), point 64
};
__device__ __forceinline__ StrWave str_wave_geom(const McParams &p, int64_t r, int wave, int lane) {
    StrWave g;
    g.r = r;
    g.w = r % p.W;
    const int64_t rt = r / p.W;
    g.ry = rt % p.NRY;
    g.rx = rt / p.NRY;
    g.x0 = g.rx * kStrRX;
    g.y0 = g.ry * kStrRY + wave * kStrRows;
    g.z0 = 64 * g.w;
    g.nx = (int) min<int64_t>(kStrRX, p.X - g.x0);
    g.ny = (int) max<int64_t>(0, min<int64_t>(kStrRows, p.Y - g.y0));
    g.nrow = (int) max<int64_t>(0, min<int64_t>(kStrRows + 1, p.Y - g.y0));   // rows incl. the one above
    g.zin = g.z0 + lane < p.Z;
    g.hz = g.z0 + 64 < p.Z;
    g.pv = low_bits(p.Z - g.z0);
    g.ev = low_bits(p.Z - 1 - g.z0);
    g.cz0 = grid_coord(g.z0 + lane, p.Z, p.amin[2], p.amax[2]);
    g.cz1 = grid_coord(g.z0 + lane + 1, p.Z, p.amin[2], p.amax[2]);
    g.wbase = p.values ? p.values + g.y0 * p.Z + g.z0 : nullptr;
    g.slab = p.Y * p.Z;
    g.rowbytes = (uint32_t) (4 * p.Z);
    g.loff = 4u * (uint32_t) min<int64_t>(lane, p.Z - 1 - g.z0);
    g.hoff = (uint32_t) max(0, min(lane, g.nrow - 1)) * g.rowbytes + (g.hz ? 256u : g.loff);
    return g;
}

// Rows 0..8 of slab x at this lane's z, and the z-halo values.  Branch
// free: rows past the grid re-read its last row, lanes past its end the
// row's last point, and without a next word the halo lanes their own row
// (all masked out of every ballot); each address is a uniform row pointer
// plus the lane's 32-bit offset.
// FULL: the wave has all 9 rows, a whole word and a next word (no clamps;
// raw buffer loads from the slab's row 0, per-lane 32-bit offsets).
template <bool FULL>
__device__ __forceinline__ void str_load(const McParams &p, int64_t x, const StrWave &g, StrSlab &o) {
    const float *basep = g.wbase + x * g.slab;
    if constexpr (FULL) {
        const __amdgpu_buffer_rsrc_t r =
            __builtin_amdgcn_make_buffer_rsrc((void *) basep, (short) 0, 0x7fffffff, 0x00020000);
#pragma unroll
        for (int j = 0; j <= kStrRows; ++j)
            o.v[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                   r, (int) (g.loff + (uint32_t) j * g.rowbytes), 0, 0));
        o.h = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int) g.hoff, 0, 0));
    } else {
        const char *base = reinterpret_cast<const char *>(basep);
        const char *row = base;
#pragma unroll
        for (int j = 0; j <= kStrRows; ++j) {
            o.v[j] = *reinterpret_cast<const float *>(row + g.loff);
            if (j + 1 < g.nrow) row += g.rowbytes;
        }
        o.h = *reinterpret_cast<const float *>(base + g.hoff);
    }
}

// Vertex on a crossing edge from the value v0 at its owner point to v1;
// cand: the along-axis coordinate is within 16 ulps of an endpoint's.
__device__ __forceinline__ float3 str_vertex(float v0, float v1, float level, int axis, float cx0, float cx1,
                                             float cy0, float cy1, float cz0, float cz1, bool all_special,
                                             bool &cand) {
    const float tt = edge_t(v0, v1, level);
    const float px = lerp1(tt, cx0, axis == 0 ? cx1 : cx0);
    const float py = lerp1(tt, cy0, axis == 1 ? cy1 : cy0);
    const float pz = lerp1(tt, cz0, axis == 2 ? cz1 : cz0);
    const float ra = axis == 0 ? px : (axis == 1 ? py : pz);
    const float a = axis == 0 ? cx0 : (axis == 1 ? cy0 : cz0);
    const float b = axis == 0 ? cx1 : (axis == 1 ? cy1 : cz1);
    cand = all_special || near_endpoint(ra, a, b);
    return make_float3(px, py, pz);
}

// Inclusive sum over lanes 0..15 (one DPP row) of each lane's v.
__device__ __forceinline__ uint32_t str_row16_incl(uint32_t v) {
    v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x112, 0xf, 0xf, true);   // row_shr:2
    v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x114, 0xf, 0xf, true);   // row_shr:4
    v += (uint32_t) __builtin_amdgcn_update_dpp(0, (int) v, 0x118, 0xf, 0xf, true);   // row_shr:8
    return v;
}

// Commit slab s: inside masks to mt[s] (lane j: row j), halo bits to
// ht[s], values to ring slot s & 1; uin / uout: every point of the slab
// (its 9 rows and their halo values) is inside / outside.
template <bool FULL>
__device__ __forceinline__ void str_commit(const StrSlab &sl, const StrWave &g, float level, int lane, StrLds &w,
                                           int s, bool &uin, bool &uout) {
    const int slot = s & 1;
    uint64_t m[kStrRows + 1];
    uint64_t all = ~0ull, any = 0ull;
#pragma unroll
    for (int j = 0; j <= kStrRows; ++j) {
        if (FULL) {
            m[j] = __ballot(inside(sl.v[j], level));
        } else {
            const uint64_t b = __ballot(inside(sl.v[j], level)) & g.pv;
            m[j] = j < g.nrow ? b : 0ull;
        }
        w.v[slot][j][lane] = sl.v[j];
        if (FULL || j < g.nrow) {
            all &= m[j];
            any |= m[j];
        }
    }
    const uint32_t rows = FULL ? 0x1ffu : (uint32_t) low_bits(g.nrow);
    const uint32_t hb = (FULL || g.hz) ? (uint32_t) __ballot(inside(sl.h, level)) & rows : 0u;
    if (lane <= kStrRows) w.h[slot][lane] = sl.h;
    const uint32_t hall = (FULL || g.hz) ? rows : 0u;
    uin = (FULL ? all == ~0ull : (all & g.pv) == g.pv) && hb == hall;
    uout = any == 0ull && hb == 0u;
    uint64_t mine;
    if (uin || uout) {
        mine = (uin && (FULL || lane < g.nrow)) ? g.pv : 0ull;
    } else {
        mine = 0ull;
#pragma unroll
        for (int j = 0; j <= kStrRows; ++j) mine = lane == j ? m[j] : mine;
    }
    if (lane <= kStrRows) w.mt[s][lane] = mine;
    if (lane == 0) w.ht[s] = hb;
}

// Largest q in [0, 8) with P[q] <= i (P[0] = 0, P nondecreasing), and P[q].
__device__ __forceinline__ int str_row_of(uint32_t i, const uint32_t (&P)[kStrRows + 1], uint32_t &base) {
    int r = 0;
    base = 0;
#pragma unroll
    for (int q = 1; q < kStrRows; ++q) {
        const bool ge = i >= P[q];
        r += ge ? 1 : 0;
        base = ge ? P[q] : base;
    }
    return r;
}

// The work of a slab that is not uniform with its successor: vertices into
// the arena, per-row first vertex / face (lanes 0..7: vrow, frow).
template <bool KEYED>
__device__ __forceinline__ void str_surface(const McParams &p, const StrWave &g, StrLds &w, const StrCoords &sc,
                                            const uint8_t *ntri, int lane, int st, bool hasx, int seg,
                                            uint32_t &vrow, uint32_t &frow, uint32_t &vcount, uint32_t &fcount,
                                            bool &over, int32_t &my_off, bool &any_faces) {
    const int cs = st & 1, ns = cs ^ 1;
    const int j = lane & 7;
    const uint32_t hc_all = w.ht[st], hn_all = hasx ? w.ht[st + 1] : 0u;
    // one lane per row: crossing edges and active cells
    uint64_t xm = 0, ym = 0, zm = 0, act = 0;
    if (lane < g.ny) {
        const uint64_t m = w.mt[st][j], my = w.mt[st][j + 1];
        const uint64_t mx = hasx ? w.mt[st + 1][j] : 0ull, mxy = hasx ? w.mt[st + 1][j + 1] : 0ull;
        const bool yok = j + 1 < g.nrow;
        const uint32_t hc = (hc_all >> j) & 1u, hcy = (hc_all >> (j + 1)) & 1u;
        const uint32_t hn = (hn_all >> j) & 1u, hny = (hn_all >> (j + 1)) & 1u;
        zm = (m ^ ((m >> 1) | ((uint64_t) hc << 63))) & g.ev;
        if (yok) ym = (m ^ my) & g.pv;
        if (hasx) xm = (m ^ mx) & g.pv;
        if (hasx && yok) {
            const uint64_t o = m | my | mx | mxy, n = m & my & mx & mxy;
            const uint64_t o1 = (o >> 1) | ((uint64_t) ((hc | hcy | hn | hny) != 0u) << 63);
            const uint64_t n1 = (n >> 1) | ((uint64_t) ((hc & hcy & hn & hny) != 0u) << 63);
            act = (o | o1) & ~(n & n1) & g.ev;
        }
    }
    if (lane < kStrRows) {
        w.em[j][0] = xm;
        w.em[j][1] = ym;
        w.em[j][2] = zm;
        w.em[j][3] = act;
    }
    const uint32_t nv = (uint32_t) (__popcll(xm) + __popcll(ym) + __popcll(zm)), na = (uint32_t) __popcll(act);
    const uint32_t vin = str_row16_incl(nv), ain = str_row16_incl(na);
    uint32_t VP[kStrRows + 1], AP[kStrRows + 1];
    VP[0] = AP[0] = 0;
#pragma unroll
    for (int q = 0; q < kStrRows; ++q) {
        VP[q + 1] = (uint32_t) __builtin_amdgcn_readlane((int) vin, q);
        AP[q + 1] = (uint32_t) __builtin_amdgcn_readlane((int) ain, q);
    }
    const uint32_t nslab = VP[kStrRows], nact = AP[kStrRows];
    vrow = vcount + vin - nv;
    __builtin_amdgcn_wave_barrier();
    // vertices, one per lane, into the arena
    if (nslab && !over) {
        // arena segment: a hash of (wave, slab), so neighbouring surface
        // waves spread over all segments; a full one sends the reservation on
        // to up to three others before the wave takes the flush's slow path
        const uint32_t hsh = ((uint32_t) seg * (uint32_t) kStrRX + (uint32_t) st) * 0x9E3779B1u;
        int64_t a0 = -1;
        for (int t = 0; t < 4 && a0 < 0; ++t) {
            const int sg = (int) ((hsh >> 24) + (uint32_t) t * 67u) & (kStrArenaSeg - 1);
            unsigned long long off = 0;
            if (lane == 0) off = atomicAdd(&p.arena_ctr[sg * kSpStride], (unsigned long long) nslab);
            off = __shfl(off, 0);
            if ((int64_t) off + nslab <= p.arena_seg_cap) a0 = (int64_t) sg * p.arena_seg_cap + (int64_t) off;
        }
        if (a0 < 0) {
            over = true;
        } else {
            if (lane == st) my_off = (int32_t) a0;
            StrStage *arena = reinterpret_cast<StrStage *>(p.arena) + a0;
            const float cx0 = sc.cx[st], cx1 = sc.cx[st + 1];
            const int yl0 = (int) (g.y0 - g.ry * kStrRY);
            for (uint32_t b = 0; b < nslab; b += 64) {
                const uint32_t i = b + (uint32_t) lane;
                if (i < nslab) {
                    uint32_t base;
                    const int rr = str_row_of(i, VP, base);
                    const uint64_t exm = w.em[rr][0], eym = w.em[rr][1], ezm = w.em[rr][2];
                    const uint32_t nxm = (uint32_t) __popcll(exm), nym = (uint32_t) __popcll(eym);
                    uint32_t k = i - base;
                    int axis = 0;
                    uint64_t mm = exm;
                    if (k >= nxm) {
                        k -= nxm;
                        axis = 1;
                        mm = eym;
                        if (k >= nym) {
                            k -= nym;
                            axis = 2;
                            mm = ezm;
                        }
                    }
                    const int z = nth_set_bit(mm, (int) k);
                    const float v0 = w.v[cs][rr][z];
                    const float v1 = axis == 0 ? w.v[ns][rr][z]
                                               : (axis == 1 ? w.v[cs][rr + 1][z] : (z < 63 ? w.v[cs][rr][z + 1] : w.h[cs][rr]));
                    bool cand;
                    const float3 pos = str_vertex(v0, v1, p.level, axis, cx0, cx1, sc.cy[yl0 + rr], sc.cy[yl0 + rr + 1],
                                                  sc.cz[z], sc.cz[z + 1], p.all_special, cand);
                    StrStage e;
                    e.x = pos.x;
                    e.y = pos.y;
                    e.z = pos.z;
                    e.info = (uint32_t) axis | ((uint32_t) z << 2) | ((uint32_t) rr << 8) | ((uint32_t) st << 11) |
                             ((uint32_t) cand << 16);
                    arena[i] = e;
                }
            }
        }
    }
    // active cells, one per lane: LUT triangle counts -> each row's first face
    if (nact) {
        any_faces = true;
        const uint32_t aex = ain - na;   // lanes 0..7: the row's first active cell
        uint32_t F0 = 0, fr = 0;
        for (uint32_t b = 0; b < nact; b += 64) {
            const uint32_t i = b + (uint32_t) lane;
            int nt = 0;
            if (i < nact) {
                uint32_t base;
                const int rr = str_row_of(i, AP, base);
                const int z = nth_set_bit(w.em[rr][3], (int) (i - base));
                const uint64_t a = w.mt[st][rr], ay = w.mt[st][rr + 1];
                const uint64_t ax = w.mt[st + 1][rr], axy = w.mt[st + 1][rr + 1];
                auto bit = [&](uint64_t mm) { return (int) ((mm >> z) & 1ull); };
                auto up = [&](uint64_t mm, uint32_t h) {
                    return z < 63 ? (int) ((mm >> (z + 1)) & 1ull) : (int) (h & 1u);
                };
                const int c = bit(a) | up(a, hc_all >> rr) << 1 | bit(ay) << 2 | up(ay, hc_all >> (rr + 1)) << 3 |
                              bit(ax) << 4 | up(ax, hn_all >> rr) << 5 | bit(axy) << 6 |
                              up(axy, hn_all >> (rr + 1)) << 7;
                nt = (int) ntri[c];
            }
            const int incl = wave_incl_sum(nt), excl = incl - nt;
            const uint32_t tot = (uint32_t) __shfl(incl, 63);
            const bool here = aex >= b && aex < b + 64;
            const int ex = __shfl(excl, here ? (int) (aex - b) : lane);
            if (lane < kStrRows && here) fr = F0 + (uint32_t) ex;
            F0 += tot;
        }
        if (aex >= nact) fr = F0;
        frow = fcount + fr;
        fcount += F0;
    }
    vcount += nslab;
}

// The wave's slab sweep (FULL: all 9 rows, a whole word and a next word).
template <bool FULL, bool KEYED>
__device__ __forceinline__ void str_sweep(const McParams &p, const StrWave &g, StrLds &w, const StrCoords &sc,
                                          const uint8_t *ntri, int lane, int seg, uint32_t &vcount, uint32_t &fcount,
                                          bool &over, int32_t &my_off, bool &any_faces) {
    const float level = p.level;
    // slab q is in buffer q % 3: step st loads slab st + 3 into the
    // buffer of slab st (committed a step earlier), commits slab st + 1
    // every step loads and commits: slabs past the region's halo slab or
    // the grid re-read the last one in reach (an L2 hit; unused)
    const int64_t xlast = min<int64_t>(g.x0 + g.nx, p.X - 1);
    StrSlab B0, B1, B2;
    bool uin0, uout0, uin1 = false, uout1 = false;
    str_load<FULL>(p, g.x0, g, B0);
    str_load<FULL>(p, min<int64_t>(g.x0 + 1, xlast), g, B1);
    str_load<FULL>(p, min<int64_t>(g.x0 + 2, xlast), g, B2);
    str_commit<FULL>(B0, g, level, lane, w, 0, uin0, uout0);
    auto step = [&](int st, StrSlab &ld, const StrSlab &cm) {
        const int64_t x = g.x0 + st;
        const bool hasx = x + 1 < p.X;
        str_load<FULL>(p, min<int64_t>(x + 3, xlast), g, ld);
        str_commit<FULL>(cm, g, level, lane, w, st + 1, uin1, uout1);
        __builtin_amdgcn_wave_barrier();
        const bool trivial = hasx ? ((uin0 && uin1) || (uout0 && uout1)) : (uin0 || uout0);
        uint32_t vrow = vcount, frow = fcount;
        if (!trivial)
            str_surface<KEYED>(p, g, w, sc, ntri, lane, st, hasx, seg, vrow, frow, vcount, fcount, over, my_off,
                               any_faces);
        if (lane < kStrRows) {
            w.vt[st * kStrRows + lane] = vrow;
            w.ft[st * kStrRows + lane] = frow;
        }
        uin0 = uin1;
        uout0 = uout1;
        __builtin_amdgcn_wave_barrier();
    };
    for (int st = 0; st < g.nx; st += 3) {
        step(st, B0, B1);
        if (st + 1 < g.nx) step(st + 1, B1, B2);
        if (st + 2 < g.nx) step(st + 2, B2, B0);
    }
}

template <bool KEYED>
__global__ void __launch_bounds__(256) k_stream(McParams p) {
    __shared__ StrLds sw[kStrWaves];
    __shared__ StrCoords sc;
    __shared__ uint8_t ntri[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform per wave
    const int64_t r = str_region(blockIdx.x, p.NR);
    if (r >= p.NR) return;
    const StrWave g = str_wave_geom(p, r, wave, lane);
    ntri[tid] = (uint8_t) (cTris[p.method][tid] & 15);
    {
        const int64_t xr = g.rx * kStrRX, yr = g.ry * kStrRY;
        if (tid <= kStrRX) {
            sc.cx[tid] = grid_coord(xr + tid, p.X, p.amin[0], p.amax[0]);
        } else if (tid <= kStrRX + 1 + kStrRY) {
            const int k = tid - kStrRX - 1;
            sc.cy[k] = grid_coord(yr + k, p.Y, p.amin[1], p.amax[1]);
        } else if (tid <= kStrRX + kStrRY + 2 + 64) {
            const int k = tid - kStrRX - kStrRY - 2;
            sc.cz[k] = grid_coord(g.z0 + k, p.Z, p.amin[2], p.amax[2]);
        }
    }
    __syncthreads();
    const int64_t wid = r * kStrWaves + wave;
    StrLds &w = sw[wave];
    const int seg = (int) (uint32_t) wid;   // arena segment hash input
    uint32_t vcount = 0, fcount = 0;
    bool any_faces = false, over = false;
    int32_t my_off = 0;   // lane s: arena entry of slab s's first vertex
    if (g.ny > 0) {
        if (g.nrow == kStrRows + 1 && g.hz)
            str_sweep<true, KEYED>(p, g, w, sc, ntri, lane, seg, vcount, fcount, over, my_off, any_faces);
        else
            str_sweep<false, KEYED>(p, g, w, sc, ntri, lane, seg, vcount, fcount, over, my_off, any_faces);
    }
    // region-major entries of this wave: whole lines
    const int64_t e0 = wid * (int64_t) (kStrRX * kStrRows);
#pragma unroll
    for (int q = 0; q < kStrRX * kStrRows / 64; ++q) {
        const int e = q * 64 + lane, s = e / kStrRows, j = e % kStrRows;
        const bool ok = g.ny > 0 && s < g.nx;
        p.bits[e0 + e] = ok ? w.mt[s][j] : 0ull;
        p.vofs[e0 + e] = ok ? w.vt[e] : 0u;
        p.fofs[e0 + e] = ok ? w.ft[e] : 0u;
    }
    if (lane < kStrRX) p.arena_off[wid * kStrRX + lane] = my_off;
    if (lane == 0) {
        p.wcnt_v[wid] = vcount;
        p.wcnt_f[wid] = fcount;
        p.wover[wid] = over ? 1 : 0;
    }
    // the wave's emit tile (16 x 16 cell rows, one word) has a face
    if (any_faces && lane == 0 && p.tile_list) {
        const int64_t tj = g.y0 / 16;
        const int t = (int) ((g.rx * p.NJ + tj) * p.NK + g.w);
        if (atomicCAS(&p.tile_flag[t], 0, 1) == 0) p.tile_list[atomicAdd(p.tile_count, 1)] = t;
    }
}

// First level of the wave-count scan: each block of kStrScanChunk waves
// writes the waves' first ids relative to the chunk and the chunk's sums.
__global__ void __launch_bounds__(kStrScanChunk) k_stream_scan1(McParams p) {
    constexpr int NW = kStrScanChunk / 64;
    __shared__ int sv[NW], sf[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t nw = p.NR * kStrWaves;
    const int64_t i = (int64_t) blockIdx.x * kStrScanChunk + tid;
    const int v = i < nw ? (int) p.wcnt_v[i] : 0, f = i < nw ? (int) p.wcnt_f[i] : 0;
    const int iv = wave_incl_sum(v), iff = wave_incl_sum(f);
    if (lane == 63) {
        sv[wave] = iv;
        sf[wave] = iff;
    }
    __syncthreads();
    int64_t bv = 0, bf = 0, tv = 0, tf = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        if (q < wave) {
            bv += sv[q];
            bf += sf[q];
        }
        tv += sv[q];
        tf += sf[q];
    }
    if (i < nw) {
        p.wb_v[i] = bv + iv - v;
        p.wb_f[i] = bf + iff - f;
    }
    if (tid == 0) {
        p.csum_v[blockIdx.x] = tv;
        p.csum_f[blockIdx.x] = tf;
    }
    // the waves with vertices, for the flush
    const uint64_t act = __ballot(v > 0);
    int at = 0;
    if (lane == 0 && act) at = atomicAdd(p.acount, __popcll(act));
    at = __shfl(at, 0);
    if (v > 0) p.alist[at + __popcll(act & ((1ull << lane) - 1ull))] = i;
}

// Chunk bases into the waves' first vertex / face ids.
__global__ void __launch_bounds__(kStrScanChunk) k_stream_fin(McParams p) {
    const int64_t nw = p.NR * kStrWaves;
    const int64_t i = (int64_t) blockIdx.x * kStrScanChunk + threadIdx.x;
    if (i < nw) {
        p.wb_v[i] += p.cofs_v[blockIdx.x];
        p.wb_f[i] += p.cofs_f[blockIdx.x];
    }
}

// Second level (one workgroup; each thread a contiguous run of chunks):
// the chunks' first ids, the mesh size; meshes beyond int32 indices are
// flagged.
__global__ void __launch_bounds__(1024) k_stream_scan2(McParams p) {
    constexpr int NT = 1024, NW = NT / 64;
    __shared__ int64_t sv[NW], sf[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t n = (p.NR * kStrWaves + kStrScanChunk - 1) / kStrScanChunk;
    const int64_t per = (n + NT - 1) / NT;
    const int64_t lo = min<int64_t>((int64_t) tid * per, n), hi = min<int64_t>(lo + per, n);
    int64_t v = 0, f = 0;
    for (int64_t i = lo; i < hi; ++i) {
        v += p.csum_v[i];
        f += p.csum_f[i];
    }
    int64_t iv = v, iff = f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t yv = __shfl_up(iv, o), yf = __shfl_up(iff, o);
        if (lane >= o) {
            iv += yv;
            iff += yf;
        }
    }
    if (lane == 63) {
        sv[wave] = iv;
        sf[wave] = iff;
    }
    __syncthreads();
    int64_t rv = iv - v, rf = iff - f, tv = 0, tf = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        if (q < wave) {
            rv += sv[q];
            rf += sf[q];
        }
        tv += sv[q];
        tf += sf[q];
    }
    for (int64_t i = lo; i < hi; ++i) {
        p.cofs_v[i] = rv;
        p.cofs_f[i] = rf;
        rv += p.csum_v[i];
        rf += p.csum_f[i];
    }
    if (tid == 0) {
        p.totals[0] = tv;
        p.totals[1] = tf;
        if (tv > 0x7fffffffll || tf > 0x7fffffffll) set_err(p, ISO_ERR_TOO_LARGE);
    }
}

// Each wave finalises its first vertex / face id, then moves its vertices
// from the arena to their ids, one per lane (keys, weld-candidate records
// and the owner rows' candidate flags on the way); a wave whose vertices
// did not fit the arena reads its rows again and writes them directly.
// One wave's flush (wid: a wave with vertices).
template <bool KEYED>
__device__ __forceinline__ void str_flush_wave(const McParams &p, int64_t wid, int lane) {
    const int64_t r = wid / kStrWaves;
    const int wave = (int) (wid - r * kStrWaves);
    const int64_t base = p.wb_v[wid];
    const uint32_t cnt = p.wcnt_v[wid];
    const StrWave g = str_wave_geom(p, r, wave, lane);
    const int spseg = (int) (r % p.seg.nseg);
    const int64_t e0 = wid * (int64_t) (kStrRX * kStrRows);
    auto emit = [&](int64_t vid, float3 pos, int st, int j, int z, int axis, bool cand) {
        if (vid >= p.cap_v) return;   // the host sees totals beyond the capacity and retries
        *reinterpret_cast<float3 *>(p.out_v + 3 * vid) = pos;
        const int64_t owner = ((g.x0 + st) * p.Y + (g.y0 + j)) * p.Z + g.z0 + z;
        if (KEYED) p.out_key[vid] = owner * 3 + axis;
        if (cand) {
            atomicOr(&p.vofs[e0 + st * kStrRows + j], kStrCandFlag);
            const int64_t q = seg_push(p, 0, spseg);
            if (q >= 0) {
                SpecialVertex sv;
                sv.vid = vid;
                sv.edge = owner * 3 + axis;
                sv.x = pos.x;
                sv.y = pos.y;
                sv.z = pos.z;
                sv.pad = 0.0f;
                p.seg.cand[q] = sv;
            }
        }
    };
    if (!p.wover[wid]) {
        // slab st's vertices: [S[st], S[st + 1]) of the wave, at arena_off[st]
        const uint32_t s0l = lane < g.nx ? (p.vofs[e0 + lane * kStrRows] & ~kStrCandFlag) : 0xffffffffu;
        const int a0l = lane < kStrRX ? p.arena_off[wid * kStrRX + lane] : 0;
        uint32_t S[kStrRX];
#pragma unroll
        for (int q = 0; q < kStrRX; ++q) S[q] = (uint32_t) __builtin_amdgcn_readlane((int) s0l, q);
        const StrStage *arena = reinterpret_cast<const StrStage *>(p.arena);
        // four batches of 64 vertices: all loads first, then the stores
        constexpr int NB = 4;
        for (uint32_t b = 0; b < cnt; b += 64 * NB) {
            StrStage e[NB];
            int sts[NB];
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const uint32_t i = b + 64 * k + (uint32_t) lane;
                int st = 0;
#pragma unroll
                for (int q = 1; q < kStrRX; ++q) st += i >= S[q] ? 1 : 0;
                const uint32_t s0 = (uint32_t) __shfl((int) s0l, st);
                const int a0 = __shfl(a0l, st);
                sts[k] = st;
                if (i < cnt) e[k] = arena[(int64_t) a0 + (i - s0)];
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const uint32_t i = b + 64 * k + (uint32_t) lane;
                if (i < cnt)
                    emit(base + i, make_float3(e[k].x, e[k].y, e[k].z), sts[k], (int) ((e[k].info >> 8) & 7),
                         (int) ((e[k].info >> 2) & 63), (int) (e[k].info & 3), (e[k].info >> 16) & 1u);
            }
        }
        return;
    }
    // slow path: the wave's rows again, vertices written at their ids
    for (int st = 0; st < g.nx; ++st) {
        const int64_t x = g.x0 + st;
        const bool hasx = x + 1 < p.X;
        StrSlab A, B;
        str_load<false>(p, x, g, A);
        if (hasx) str_load<false>(p, x + 1, g, B);
        uint64_t ma[kStrRows + 1], mb[kStrRows + 1];
#pragma unroll
        for (int j = 0; j <= kStrRows; ++j) {
            ma[j] = j < g.nrow ? __ballot(inside(A.v[j], p.level)) & g.pv : 0ull;
            mb[j] = (hasx && j < g.nrow) ? __ballot(inside(B.v[j], p.level)) & g.pv : 0ull;
        }
        const uint32_t ha = g.hz ? (uint32_t) __ballot(inside(A.h, p.level)) & (uint32_t) low_bits(g.nrow) : 0u;
        const float cx0 = grid_coord(x, p.X, p.amin[0], p.amax[0]);
        const float cx1 = grid_coord(x + 1, p.X, p.amin[0], p.amax[0]);
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (int j = 0; j < kStrRows; ++j) {
            if (j >= g.ny) break;
            const int64_t y = g.y0 + j;
            const uint64_t m = ma[j];
            const uint64_t zm = (m ^ ((m >> 1) | ((uint64_t) ((ha >> j) & 1u) << 63))) & g.ev;
            const uint64_t ym = (j + 1 < g.nrow) ? (m ^ ma[j + 1]) & g.pv : 0ull;
            const uint64_t xm = hasx ? (m ^ mb[j]) & g.pv : 0ull;
            if (!(xm | ym | zm)) continue;
            const int64_t rb = base + (int64_t) (p.vofs[e0 + st * kStrRows + j] & ~kStrCandFlag);
            const float cy0 = grid_coord(y, p.Y, p.amin[1], p.amax[1]);
            const float cy1 = grid_coord(y + 1, p.Y, p.amin[1], p.amax[1]);
            const float zn = __shfl(A.v[j], (lane + 1) & 63), hv = __shfl(A.h, j);
            uint32_t off = 0;
#pragma unroll
            for (int axis = 0; axis < 3; ++axis) {
                const uint64_t mm = axis == 0 ? xm : (axis == 1 ? ym : zm);
                if ((mm >> lane) & 1ull) {
                    const float v1 = axis == 0 ? B.v[j] : (axis == 1 ? A.v[j + 1] : (lane < 63 ? zn : hv));
                    bool cand;
                    const float3 pos = str_vertex(A.v[j], v1, p.level, axis, cx0, cx1, cy0, cy1, g.cz0, g.cz1,
                                                  p.all_special, cand);
                    emit(rb + off + (uint32_t) __popcll(mm & below), pos, st, j, lane, axis, cand);
                }
                off += (uint32_t) __popcll(mm);
            }
        }
    }
}

// Moves each listed wave's vertices from the arena to their ids (keys,
// weld-candidate records and the owner rows' candidate flags on the way); a
// wave whose vertices did not fit the arena reads its rows again and writes
// them directly.  Persistent: the waves of the grid step through the list.
template <bool KEYED>
__global__ void __launch_bounds__(256) k_stream_flush(McParams p) {
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (*(volatile int *) p.err & ISO_ERR_TOO_LARGE) return;
    const int64_t n = *p.acount;
    for (int64_t k = (int64_t) blockIdx.x * kStrWaves + wave; k < n; k += (int64_t) gridDim.x * kStrWaves) {
        const int64_t wid = p.alist[k];
        const int lo = __builtin_amdgcn_readfirstlane((int) (uint32_t) wid);
        const int hi = __builtin_amdgcn_readfirstlane((int) (uint32_t) (wid >> 32));
        str_flush_wave<KEYED>(p, (int64_t) (((uint64_t) (uint32_t) hi << 32) | (uint32_t) lo), lane);
    }
}

// k_faces: one workgroup per listed tile of TX x TY cell rows x one word.
// The tile stages its mask patch (rows + 2, words + 1, from the region-major
// masks) and its vertex / face bases (wave base + row offset), compacts the
// active cells and writes their faces; vertex ids are the owner row's base
// plus the rank of the edge in (axis, z) order.  A cell is special when one
// of its vertices is a weld candidate (the owner row's vofs flag): its
// faces are tested for bitwise-equal corners (positions read back from the
// vertex output) and recorded for the fix-up, with its crossing edges.

// Id of the vertex on the x- (axis 0) or y-edge (axis 1) of tile-local
// point row (opx, opy) at z = 0 of the patch's second word.
template <int TX, int TY, int NZC, int NT>
__device__ __forceinline__ int str_vertex_id_next(const EmitLds<TX, TY, NZC, NT> &s, int hx, int hy, int opx,
                                                  int opy, int axis) {
    using L = EmitLds<TX, TY, NZC, NT>;
    const uint64_t b = PW(s, opx, opy, 1);
    const uint64_t xm = opx < hx ? (b ^ PW(s, opx + 1, opy, 1)) : 0ull;
    return s.vb[(opx * L::QY + opy) * L::QC + 1] + (axis == 1 ? __popcll(xm) : 0);
}

template <int TX, int TY, int NT>
__device__ __forceinline__ void faces_tile(const McParams &p, int64_t tile, int spseg) {
    constexpr int NZC = 1;
    using L = EmitLds<TX, TY, NZC, NT>;
    __shared__ L s;
    const int tid = threadIdx.x;
    if (*(volatile int *) p.err & ISO_ERR_TOO_LARGE) return;
    const int64_t tk = tile % p.NK, tj = (tile / p.NK) % p.NJ, ti = tile / (p.NK * p.NJ);
    const int64_t x0 = ti * TX, y0 = tj * TY, c0 = tk;
    const int nx = (int) min<int64_t>(TX, p.X - 1 - x0);
    const int ny = (int) min<int64_t>(TY, p.Y - 1 - y0);
    const int hx = (int) min<int64_t>(TX + 2, p.X - 1 - x0);
    const int hy = (int) min<int64_t>(TY + 2, p.Y - 1 - y0);
    constexpr int NP = L::PX * L::PY * L::PC, IT = (NP + NT - 1) / NT;
    constexpr int CR = (TX * TY + NT - 1) / NT, PR = ((TX + 1) * (TY + 1) * L::QC + NT - 1) / NT;
    uint64_t wv[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = tid + k * NT;
        const int j = i % L::PC, rr = i / L::PC, py = rr % L::PY, px = rr / L::PY;
        const int64_t gx = x0 + px, gy = y0 + py, c = c0 + j;
        wv[k] = (i < NP && gx < p.X && gy < p.Y && c < p.W) ? p.bits[str_entry(p, gx, gy, c)] : 0ull;
    }
    int64_t fbase_pf[CR], vbase_pf[PR];
#pragma unroll
    for (int k = 0; k < CR; ++k) {
        const int i = tid + k * NT, cx = i / TY, cy = i % TY;
        fbase_pf[k] = 0;
        if (i < TX * TY && cx < nx && cy < ny) {
            const int64_t gx = x0 + cx, gy = y0 + cy;
            fbase_pf[k] = p.wb_f[str_wave(p, gx, gy, c0)] + (int64_t) p.fofs[str_entry(p, gx, gy, c0)];
        }
    }
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int i = tid + k * NT;
        const int j = i % L::QC, rr = i / L::QC, py = rr % (TY + 1), px = rr / (TY + 1);
        const int64_t gx = x0 + px, gy = y0 + py, c = c0 + j;
        vbase_pf[k] = 0;
        if (i < (TX + 1) * (TY + 1) * L::QC && gx < p.X && gy < p.Y && c < p.W) {
            const uint32_t o = p.vofs[str_entry(p, gx, gy, c)];
            // bit 62 of the prefetched base carries the row's candidate flag
            vbase_pf[k] = (p.wb_v[str_wave(p, gx, gy, c)] + (int64_t) (o & ~kStrCandFlag)) |
                          ((int64_t) (o >> 31) << 62);
        }
    }
    for (int c = tid; c < TX + TY + 64 + 3; c += NT) {
        if (c <= TX) {
            s.cx[c] = grid_coord(x0 + c, p.X, p.amin[0], p.amax[0]);
        } else if (c <= TX + TY + 1) {
            s.cy[c - TX - 1] = grid_coord(y0 + (c - TX - 1), p.Y, p.amin[1], p.amax[1]);
        } else {
            s.cz[c - TX - TY - 2] = grid_coord(64 * c0 + (c - TX - TY - 2), p.Z, p.amin[2], p.amax[2]);
        }
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = tid + k * NT;
        if (i < NP) s.w[i] = wv[k];
    }
    __syncthreads();
    for (int u = tid; u < L::NUNITS; u += NT) {
        const int cy = u % TY, cx = u / TY;
        const int64_t zlo = 64 * c0;
        uint64_t act = 0;
        if (cx < nx && cy < ny && zlo < p.Z - 1) act = cell_active(unit_cells(s, cx, cy, 0)) & low_bits(p.Z - 1 - zlo);
        s.amask[u] = act;
        s.na[u] = (typename L::NaT) __popcll(act);
    }
    __syncthreads();
    block_exscan_lds<L::NUNITS, NT>(s.na, s.tsum);
    const int nactive = (int) s.na[L::NUNITS];
    if (nactive == 0) return;
#pragma unroll
    for (int k = 0; k < CR; ++k) {
        const int i = tid + k * NT;
        if (i < TX * TY) s.fb[i] = (int32_t) fbase_pf[k];
    }
    // the candidate flags of the patch's (point row, word) entries
    __shared__ uint8_t vflag[(TX + 1) * (TY + 1) * L::QC];
#pragma unroll
    for (int k = 0; k < PR; ++k) {
        const int i = tid + k * NT;
        if (i < (TX + 1) * (TY + 1) * L::QC) {
            s.vb[i] = (int32_t) (vbase_pf[k] & ((1ll << 62) - 1));
            vflag[i] = (uint8_t) ((uint64_t) vbase_pf[k] >> 62);
        }
    }
    if (tid == 0) s.carry[0] = 0;
    for (int j0 = 0, bt = 0; j0 < nactive; j0 += NT, bt ^= 1) {
        s.unit[tid] = -1;
        __syncthreads();
        for (int u = tid; u < L::NUNITS; u += NT) {
            const int a0 = (int) s.na[u], a1 = (int) s.na[u + 1];
            if (a1 > a0 && a0 < j0 + NT && a1 > j0) s.unit[max(a0, j0) - j0] = u;
        }
        __syncthreads();
        block_maxscan_lds<NT>(s.unit, s.tsum);
        const int jj = j0 + tid;
        const bool live = jj < nactive;
        int u = 0, cs = 0, k = 0;
        if (live) {
            u = s.unit[tid];
            k = nth_set_bit(s.amask[u], jj - (int) s.na[u]);
            const int cyy = u % TY, cxx = u / TY;
            const uint64_t w00 = PW(s, cxx, cyy, 0), w01 = PW(s, cxx, cyy + 1, 0);
            const uint64_t w10 = PW(s, cxx + 1, cyy, 0), w11 = PW(s, cxx + 1, cyy + 1, 0);
            uint32_t up;
            if (k < 63) {
                up = (uint32_t) ((w00 >> (k + 1)) & 1) | (uint32_t) ((w01 >> (k + 1)) & 1) << 1 |
                     (uint32_t) ((w10 >> (k + 1)) & 1) << 2 | (uint32_t) ((w11 >> (k + 1)) & 1) << 3;
            } else {
                up = (uint32_t) (PW(s, cxx, cyy, 1) & 1) | (uint32_t) (PW(s, cxx, cyy + 1, 1) & 1) << 1 |
                     (uint32_t) (PW(s, cxx + 1, cyy, 1) & 1) << 2 | (uint32_t) (PW(s, cxx + 1, cyy + 1, 1) & 1) << 3;
            }
            cs = (int) (((w00 >> k) & 1) | ((up & 1) << 1) | (((w01 >> k) & 1) << 2) | (((up >> 1) & 1) << 3) |
                        (((w10 >> k) & 1) << 4) | (((up >> 2) & 1) << 5) | (((w11 >> k) & 1) << 6) |
                        (((up >> 3) & 1) << 7));
        }
        s.fsc[tid] = live ? (uint16_t) cNtri[p.method][cs] : (uint16_t) 0;
        __syncthreads();
        block_exscan_lds<NT, NT>(s.fsc, s.tsum);
        if (tid == NT - 1 && j0 + NT < nactive) {
            const int rl = s.unit[NT - 1];
            const int rs = (int) s.na[rl], re = (int) s.na[rl + 1];
            s.carry[bt ^ 1] = re > j0 + NT ? (rs < j0 ? s.carry[bt] : 0) + (int) (s.fsc[NT] - s.fsc[max(rs - j0, 0)])
                                           : 0;
        }
        if (live) {
            const int rrow = u, cy = u % TY, cx = u / TY;
            const int rs = (int) s.na[rrow];
            const int fo = (int) (s.fsc[tid] - s.fsc[max(rs - j0, 0)]) + (rs < j0 ? s.carry[bt] : 0);
            const int zl = k;
            const int64_t gx = x0 + cx, gy = y0 + cy, gz = 64 * c0 + zl;
            const uint32_t status = (uint32_t) crossing_status(cs);
            int32_t id[12];
            {
                const uint64_t b00 = PW(s, cx, cy, 0), b01 = PW(s, cx, cy + 1, 0);
                const uint64_t b10 = PW(s, cx + 1, cy, 0), b11 = PW(s, cx + 1, cy + 1, 0);
                const uint64_t b20 = PW(s, cx + 2, cy, 0), b21 = PW(s, cx + 2, cy + 1, 0);
                const uint64_t b02 = PW(s, cx, cy + 2, 0), b12 = PW(s, cx + 1, cy + 2, 0);
                const bool hx0 = cx < hx, hx1 = cx + 1 < hx, hy0 = cy < hy, hy1 = cy + 1 < hy;
                const uint64_t xm00 = hx0 ? b00 ^ b10 : 0ull, xm01 = hx0 ? b01 ^ b11 : 0ull;
                const uint64_t xm10 = hx1 ? b10 ^ b20 : 0ull, xm11 = hx1 ? b11 ^ b21 : 0ull;
                const uint64_t ym00 = hy0 ? b00 ^ b01 : 0ull, ym10 = hy0 ? b10 ^ b11 : 0ull;
                const uint64_t ym01 = hy1 ? b01 ^ b02 : 0ull, ym11 = hy1 ? b11 ^ b12 : 0ull;
                const uint64_t zm00 = b00 ^ (b00 >> 1), zm01 = b01 ^ (b01 >> 1), zm10 = b10 ^ (b10 >> 1);
                const uint64_t zm11 = b11 ^ (b11 >> 1);
                // ids in the row order of k_stream: x-edges by z, then
                // y-edges, then z-edges
                const uint64_t low = (1ull << k) - 1;
                auto pc = [](uint64_t m) { return __popcll(m); };
                constexpr int QC = L::QC, QY = L::QY;
                const int v00 = s.vb[(cx * QY + cy) * QC], v01 = s.vb[(cx * QY + cy + 1) * QC];
                const int v10 = s.vb[((cx + 1) * QY + cy) * QC];
                const int v11 = s.vb[((cx + 1) * QY + cy + 1) * QC];
                const int X00 = pc(xm00), X01 = pc(xm01), X10 = pc(xm10), X11 = pc(xm11);
                const int XY00 = X00 + pc(ym00), XY01 = X01 + pc(ym01), XY10 = X10 + pc(ym10);
                const int XY11 = X11 + pc(ym11);
                id[8] = v00 + pc(xm00 & low);
                id[3] = v00 + X00 + pc(ym00 & low);
                id[0] = v00 + XY00 + pc(zm00 & low);
                id[11] = v01 + pc(xm01 & low);
                id[2] = v01 + XY01 + pc(zm01 & low);
                id[7] = v10 + X10 + pc(ym10 & low);
                id[4] = v10 + XY10 + pc(zm10 & low);
                id[6] = v11 + XY11 + pc(zm11 & low);
                if (k < 63) {
                    const uint64_t low1 = (low << 1) | 1ull;
                    id[9] = v00 + pc(xm00 & low1);
                    id[1] = v00 + X00 + pc(ym00 & low1);
                    id[10] = v01 + pc(xm01 & low1);
                    id[5] = v10 + X10 + pc(ym10 & low1);
                } else {
                    id[9] = str_vertex_id_next(s, hx, hy, cx, cy, 0);
                    id[1] = str_vertex_id_next(s, hx, hy, cx, cy, 1);
                    id[10] = str_vertex_id_next(s, hx, hy, cx, cy + 1, 0);
                    id[5] = str_vertex_id_next(s, hx, hy, cx + 1, cy, 1);
                }
#pragma unroll
                for (int e = 0; e < 12; ++e) s.ids[e][tid] = id[e];
            }
            // a weld candidate on one of the cell's edges makes it special:
            // edges whose owner row is flagged are tested exactly, on the
            // vertex's along-axis coordinate (the criterion k_stream used)
            bool special = p.all_special && status != 0;
            if (!special) {
                constexpr int QC = L::QC, QY = L::QY;
                uint32_t m = status;
                while (m) {
                    const int e = __ffs(m) - 1;
                    m &= m - 1;
                    const int ca = edge_a(e), cb = edge_b(e), axis = edge_axis(e);
                    const int ox = cx + (ca >> 2), oy = cy + ((ca >> 1) & 1), oz = zl + (ca & 1);
                    if (!vflag[(ox * QY + oy) * QC + (oz >> 6)]) continue;
                    const int64_t v = id[e];
                    if (v < 0 || v >= p.cap_v) continue;
                    const float *pv = p.out_v + 3 * v;
                    const float ra = pv[axis];
                    float a, b;
                    if (axis == 0) {
                        a = s.cx[cx + (ca >> 2)];
                        b = s.cx[cx + (cb >> 2)];
                    } else if (axis == 1) {
                        a = s.cy[cy + ((ca >> 1) & 1)];
                        b = s.cy[cy + ((cb >> 1) & 1)];
                    } else {
                        a = s.cz[zl + (ca & 1)];
                        b = s.cz[zl + (cb & 1)];
                    }
                    if (near_endpoint(ra, a, b)) special = true;
                }
            }
            const uint64_t wl = cTris[p.method][cs];
            const int ntri = (int) (wl & 15);
            int64_t slot = (int64_t) s.fb[rrow] + fo;
            bool dropped = false;   // a face of this cell is degenerate
            for (int q = 0; q < ntri; ++q, ++slot) {
                const int e0 = (int) ((wl >> (4 + 12 * q)) & 15);
                const int e1 = (int) ((wl >> (8 + 12 * q)) & 15);
                const int e2 = (int) ((wl >> (12 + 12 * q)) & 15);
                const int32_t i0 = s.ids[e0][tid], i1 = s.ids[e1][tid], i2 = s.ids[e2][tid];
                bool keep = true;
                if (special) {
                    auto at = [&](int32_t v) {
                        return v >= 0 && v < p.cap_v ? *reinterpret_cast<const float3 *>(p.out_v + 3 * (int64_t) v)
                                                     : make_float3(0.0f, 0.0f, 0.0f);
                    };
                    const float3 a = at(i0), b = at(i1), c = at(i2);
                    const bool d01 = a.x == b.x && a.y == b.y && a.z == b.z;
                    const bool d02 = a.x == c.x && a.y == c.y && a.z == c.z;
                    const bool d12 = b.x == c.x && b.y == c.y && b.z == c.z;
                    keep = !(d01 || d02 || d12);
                    dropped |= !keep;
                    const int64_t i = seg_push(p, keep ? 2 : 3, spseg);
                    if (i >= 0) (keep ? p.seg.kept : p.seg.hole)[i] = slot;
                }
                if (slot >= 0 && slot < p.cap_f) {
                    int32_t *dst = p.out_f + 3 * slot;
                    *reinterpret_cast<int3 *>(dst) = keep ? make_int3(i0, i1, i2) : make_int3(-1, -1, -1);
                }
            }
            // crossing edges of a cell that dropped a face (see emit_tile)
            if (special && dropped) {
                uint32_t m = status;
                while (m) {
                    const int e = __ffs(m) - 1;
                    m &= m - 1;
                    const int64_t i = seg_push(p, 1, spseg);
                    if (i >= 0) {
                        SpecialEdge se;
                        se.vid = s.ids[e][tid];
                        se.edge = edge_key(p, gx, gy, gz, e);
                        p.seg.scell[i] = se;
                    }
                }
            }
        }
        __syncthreads();
    }
}

template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) k_faces(McParams p) {
    const int64_t tile = emit_block_tile(p);
    if (tile >= 0) faces_tile<TX, TY, NT>(p, tile, (int) (blockIdx.x % p.seg.nseg));
}

// Reset of the read-once pipeline's per-call state.
__global__ void k_stream_reset(McParams p, SpecialLists init) {
    const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x, nt = (int64_t) gridDim.x * blockDim.x;
    if (t < 4 * p.seg.nseg) p.seg.ctr[t * kSpStride] = 0;
    if (t < kStrArenaSeg) p.arena_ctr[t * kSpStride] = 0;
    if (p.tile_flag) {
        const int64_t ntiles = p.NI * p.NJ * p.NK;
        for (int64_t i = t; i < ntiles; i += nt) p.tile_flag[i] = 0;
    }
    if (p.dslots)
        for (int64_t i = t; i <= p.dmask; i += nt) p.dslots[i] = -1;
    if (t == 0) {
        if (p.tile_count) *p.tile_count = 0;
        *p.acount = 0;
        *p.err = 0;
        p.totals[0] = p.totals[1] = 0;
        init.n_cand = init.n_scell = init.n_kept = init.n_hole = 0;
        init.need = 0;
        init.dup = p.dslots ? 0 : -1;
        *p.sp = init;
    }
}

hipError_t mc_stream_launch(McParams &p, const SpecialLists &sp_init, hipStream_t stream, hipEvent_t *ev) {
    mc_params_divisors(p);
    hipError_t e = mc_upload_luts();
    if (e != hipSuccess) return e;
    auto mark = [&](int i) { return ev ? hipEventRecord(ev[i], stream) : hipSuccess; };
    if ((e = mark(0)) != hipSuccess) return e;
    const int64_t ntiles = p.NI * p.NJ * p.NK;
    const int64_t rwork = std::max<int64_t>(ntiles, 1024);
    const unsigned rblocks = (unsigned) std::min<int64_t>((rwork + 255) / 256, 2048);
    hipLaunchKernelGGL(k_stream_reset, dim3(rblocks), dim3(256), 0, stream, p, sp_init);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // stage times: reset, k_stream, scan + flush, k_faces, total
    if ((e = mark(1)) != hipSuccess) return e;
    const unsigned sblocks = (unsigned) (8 * ((p.NR + 7) / 8));
    if (p.out_key)
        hipLaunchKernelGGL(k_stream<true>, dim3(sblocks), dim3(256), 0, stream, p);
    else
        hipLaunchKernelGGL(k_stream<false>, dim3(sblocks), dim3(256), 0, stream, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(2)) != hipSuccess) return e;
    const int64_t nchunk = (p.NR * kStrWaves + kStrScanChunk - 1) / kStrScanChunk;
    hipLaunchKernelGGL(k_stream_scan1, dim3((unsigned) nchunk), dim3(kStrScanChunk), 0, stream, p);
    hipLaunchKernelGGL(k_stream_scan2, dim3(1), dim3(1024), 0, stream, p);
    hipLaunchKernelGGL(k_stream_fin, dim3((unsigned) nchunk), dim3(kStrScanChunk), 0, stream, p);
    const unsigned fblocks = (unsigned) std::min<int64_t>(2048, std::max<int64_t>(1, (p.NR + 7) / 8));
    if (p.out_key)
        hipLaunchKernelGGL(k_stream_flush<true>, dim3(fblocks), dim3(256), 0, stream, p);
    else
        hipLaunchKernelGGL(k_stream_flush<false>, dim3(fblocks), dim3(256), 0, stream, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(3)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_faces<16, 16, 256>), dim3((unsigned) ntiles), dim3(256), 0, stream, p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = mark(4)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_sp_gather_dup, dim3((unsigned) p.seg.nseg + (p.dslots ? 64u : 0u)), dim3(256), 0, stream,
                       p);
    return hipGetLastError();
}

}   // namespace iso
