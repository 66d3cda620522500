// Dense marching cubes: one fused, read-once pass over the value grid.
//
// Reference path replaced: mc::marching_cubes on a UniformGrid
// (src/mc/mc.cu:17-68 -> get_case_num_op include/utils.cuh:82-102,
// compaction src/mc/mc.cu:33-42, Nagae/Lorensen::process_cube_op
// src/mc/nagae.cu:34-80 / src/mc/lorensen.cu:34-80, vertex_welding
// src/utils.cu:32-59).
//
// Design (gfx950, wave64):
//   * The grid of cells is cut into tiles of TX x TY x (64*NZC) cells.  A
//     workgroup takes a ticket (processing order = ticket order, highest
//     linear tile first) and streams its tile's values plus a one-point
//     upper halo from HBM exactly once, turning each 64-point row chunk into a
//     64-bit inside/outside mask with one ballot.  The masks live in LDS.
//   * Every vertex belongs to one grid edge, every edge to its lower endpoint
//     ("owner point"), every owner point to one tile.  Vertices are numbered
//     per tile in (x, y, z-chunk) unit order, then (z, axis) inside a unit, so
//     a vertex's index is a popcount over the LDS masks.  No hashing and no
//     sort: welding is by construction.
//   * Per-unit face and vertex counts are scanned in LDS; tile bases come from
//     a single-pass decoupled look-back over the tickets.  A tile then
//     publishes the ids of the vertices on its lower boundary planes, which
//     its lower neighbours (processed later) reference.
//   * Active cells gather their 8 corner values (L2/MALL hits), interpolate
//     with the reference's float op order and write faces and the vertices
//     they own.  Output is deterministic: faces in tile order, then cell
//     order, then LUT slot.
//   * Vertices whose position could coincide bitwise with another edge's
//     vertex (within 16 ulps of a grid corner) are recorded, together with the
//     faces of their cells, for the exact position-based weld of the fix-up
//     pass (the reference welds by position equality and drops degenerate
//     triangles, src/utils.cu:49-55, src/mc/nagae.cu:71).

#include "common.h"
#include "mc_dense.h"
#include "mc_luts.h"

namespace iso {

static __constant__ uint64_t cTris[2][256];

// Number of triangles of each case, per method (low nibble of the packed LUT).
__device__ __forceinline__ int tri_count(int method, int c) { return (int) (cTris[method][c] & 15); }

__device__ __forceinline__ uint64_t pack_status(uint64_t state, uint64_t f, uint64_t v) {
    return (state << 62) | ((f & 0x7fffffffull) << 31) | (v & 0x7fffffffull);
}

__device__ __forceinline__ int64_t wave_sum64(int64_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Append to a device list; returns the slot, or -1 when full (the counter
// still counts so the host can size a retry).
__device__ __forceinline__ int64_t list_push(unsigned long long *counter, int64_t cap) {
    unsigned long long i = atomicAdd(counter, 1ull);
    return (int64_t) i < cap ? (int64_t) i : -1;
}

template <int TX, int TY, int NZC>
struct TileGeom {
    static constexpr int PX = TX + 1;            // point planes incl. halo
    static constexpr int PY = TY + 1;            // point rows incl. halo
    static constexpr int PC = NZC + 1;           // 64-point chunks incl. halo
    static constexpr int TZ = 64 * NZC;          // cells along z
    static constexpr int NWORDS = PX * PY * PC;  // LDS mask words
    static constexpr int NUNITS = PX * PY * PC;  // count units (point rows x chunks)
    // publish buffer layout per tile (entries of 64 bits)
    static constexpr int PUB_X = PY * (TZ + 1) * 2;   // plane px=0: (py, z) x {y, z}
    static constexpr int PUB_Y = PX * (TZ + 1) * 2;   // plane py=0: (px, z) x {x, z}
    static constexpr int PUB_Z = PX * PY * 2;         // plane z=0: (px, py) x {x, y}
    static constexpr int PUB = PUB_X + PUB_Y + PUB_Z;
};

struct TileCtx {
    int64_t x0, y0, z0;    // global origin (cells == owner points)
    int nx, ny, nz;        // cells in this tile
    int nxo, nyo, nzo;     // owned points (cells, +1 on the grid's upper face)
    int nchunk;            // loaded 64-point chunks per row (covers nz+1)
    int64_t tile;          // linear tile index
    int ti, tj, tk;
};

template <int TX, int TY, int NZC>
struct Lds {
    using G = TileGeom<TX, TY, NZC>;
    uint64_t w[G::NWORDS];
    uint32_t fo[G::NUNITS + 1];   // face count, then exclusive offset
    uint32_t vo[G::NUNITS + 1];   // vertex count, then exclusive offset
    uint32_t na[G::NUNITS + 1];   // active-cell count, then exclusive offset
    uint32_t tsum_f[16], tsum_v[16], tsum_a[16];
    int64_t fbase, vbase;
    int64_t ticket;
};

template <int TX, int TY, int NZC>
__device__ __forceinline__ uint64_t W(const Lds<TX, TY, NZC> &s, int px, int py, int c) {
    using G = TileGeom<TX, TY, NZC>;
    return s.w[(px * G::PY + py) * G::PC + c];
}

// Crossing masks of the three edges owned by the points of unit (px, py, c):
// bit b = point z = 64c + b (tile-local).  Only owned points with existing
// edges are set.
template <int TX, int TY, int NZC>
__device__ __forceinline__ void unit_masks(const Lds<TX, TY, NZC> &s, const McParams &p,
                                           const TileCtx &t, int px, int py, int c, uint64_t &xc,
                                           uint64_t &yc, uint64_t &zc) {
    uint64_t a = W(s, px, py, c);
    uint64_t an = (c + 1 < t.nchunk) ? W(s, px, py, c + 1) : 0ull;
    uint64_t as = (a >> 1) | (an << 63);
    int zlo = 64 * c;
    int nown = t.nzo - zlo;   // owned points in this chunk
    uint64_t own = nown >= 64 ? ~0ull : (nown <= 0 ? 0ull : ((1ull << nown) - 1));
    // z-edge exists for global z < Z-1  <=>  local z < nz (cells along z)
    int nze = t.nz - zlo;
    uint64_t zexist = nze >= 64 ? ~0ull : (nze <= 0 ? 0ull : ((1ull << nze) - 1));
    zc = (a ^ as) & own & zexist;
    yc = 0;
    xc = 0;
    if (t.y0 + py < p.Y - 1) yc = (a ^ W(s, px, py + 1, c)) & own;
    if (t.x0 + px < p.X - 1) xc = (a ^ W(s, px + 1, py, c)) & own;
}

template <int TX, int TY, int NZC>
__device__ __forceinline__ int unit_index(int px, int py, int c) {
    using G = TileGeom<TX, TY, NZC>;
    return (px * G::PY + py) * G::PC + c;
}

// Tile-local vertex id of the edge (point (px,py,z), axis) owned by this tile.
template <int TX, int TY, int NZC>
__device__ __forceinline__ int64_t local_vid(const Lds<TX, TY, NZC> &s, const McParams &p,
                                             const TileCtx &t, int px, int py, int z, int axis) {
    int c = z >> 6, b = z & 63;
    uint64_t xc, yc, zc;
    unit_masks(s, p, t, px, py, c, xc, yc, zc);
    uint64_t below = (b == 0) ? 0ull : (~0ull >> (64 - b));
    int64_t r = __popcll(xc & below) + __popcll(yc & below) + __popcll(zc & below);
    if (axis >= 1) r += (xc >> b) & 1;
    if (axis >= 2) r += (yc >> b) & 1;
    return (int64_t) s.vo[unit_index<TX, TY, NZC>(px, py, c)] + r;
}

// Publish-buffer slot of an in-plane edge on a tile's lower boundary.
// plane 0: px == 0 (edges y, z); 1: py == 0 (x, z); 2: z == 0 (x, y).
template <int TX, int TY, int NZC>
__device__ __forceinline__ int64_t pub_slot(int plane, int lx, int ly, int lz, int axis) {
    using G = TileGeom<TX, TY, NZC>;
    if (plane == 0) return (int64_t) (ly * (G::TZ + 1) + lz) * 2 + (axis == 1 ? 0 : 1);
    if (plane == 1) return G::PUB_X + (int64_t) (lx * (G::TZ + 1) + lz) * 2 + (axis == 0 ? 0 : 1);
    return G::PUB_X + G::PUB_Y + (int64_t) (lx * G::PY + ly) * 2 + (axis == 0 ? 0 : 1);
}

__device__ __forceinline__ void set_err(const McParams &p, int code) {
    atomicOr(p.err, code);
}

// Per-cell interpolation context: corner values and the two grid
// coordinates of the cell along each axis.
struct CellCtx {
    float cv[8];
    float c[3][2];
    float level;
};

__device__ __forceinline__ int crossing_status(int cs) {
    int st = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (((cs >> kEdgeA[e]) & 1) != ((cs >> kEdgeB[e]) & 1)) st |= 1 << e;
    return st;
}

// Vertex position on crossing edge e, with the reference's op order.
__device__ __forceinline__ void edge_pos(const CellCtx &cc, int e, float &x, float &y, float &z) {
    int ca = kEdgeA[e], cb = kEdgeB[e];
    float tt = edge_t(cc.cv[ca], cc.cv[cb], cc.level);
    x = lerp1(tt, cc.c[0][ca >> 2], cc.c[0][cb >> 2]);
    y = lerp1(tt, cc.c[1][(ca >> 1) & 1], cc.c[1][(cb >> 1) & 1]);
    z = lerp1(tt, cc.c[2][ca & 1], cc.c[2][cb & 1]);
}

__device__ __forceinline__ bool edge_is_candidate(const CellCtx &cc, int e) {
    int ca = kEdgeA[e], cb = kEdgeB[e], axis = kEdgeAxis[e];
    float tt = edge_t(cc.cv[ca], cc.cv[cb], cc.level);
    int ia = axis == 0 ? (ca >> 2) : (axis == 1 ? ((ca >> 1) & 1) : (ca & 1));
    int ib = axis == 0 ? (cb >> 2) : (axis == 1 ? ((cb >> 1) & 1) : (cb & 1));
    float a = cc.c[axis][ia], b = cc.c[axis][ib];
    return near_endpoint(lerp1(tt, a, b), a, b);
}

// Global key of edge e of cell (gx, gy, gz): owner point linear id * 3 + axis.
__device__ __forceinline__ int64_t edge_key(const McParams &p, int64_t gx, int64_t gy, int64_t gz,
                                            int e) {
    int ca = kEdgeA[e];
    int64_t ox = gx + (ca >> 2), oy = gy + ((ca >> 1) & 1), oz = gz + (ca & 1);
    return ((ox * p.Y + oy) * p.Z + oz) * 3 + kEdgeAxis[e];
}

// Global vertex id of edge e of tile-local cell (px, py, zl): from the LDS
// ranks when this tile owns the edge, else from the owner's published entry.
template <int TX, int TY, int NZC>
__device__ __noinline__ int64_t edge_vid(const Lds<TX, TY, NZC> &s, const McParams &p,
                                         const TileCtx &t, int px, int py, int zl, int e,
                                         int64_t vbase, uint64_t epoch) {
    using G = TileGeom<TX, TY, NZC>;
    int ca = kEdgeA[e], axis = kEdgeAxis[e];
    int ox = px + (ca >> 2), oy = py + ((ca >> 1) & 1), oz = zl + (ca & 1);
    if (ox < t.nxo && oy < t.nyo && oz < t.nzo) return vbase + local_vid(s, p, t, ox, oy, oz, axis);
    int64_t gox = t.x0 + ox, goy = t.y0 + oy, goz = t.z0 + oz;
    int oi = (int) min<int64_t>(gox / TX, p.NI - 1);
    int oj = (int) min<int64_t>(goy / TY, p.NJ - 1);
    int ok = (int) min<int64_t>(goz / G::TZ, p.NK - 1);
    int lx = (int) (gox - (int64_t) oi * TX), ly = (int) (goy - (int64_t) oj * TY),
        lz = (int) (goz - (int64_t) ok * G::TZ);
    // the owner lies above this tile along an axis other than the edge's;
    // that axis names the boundary plane it published
    int plane = (oi != t.ti) ? 0 : ((oj != t.tj) ? 1 : 2);
    int64_t otile = ((int64_t) oi * p.NJ + oj) * p.NK + ok;
    const uint64_t *src = p.pub + otile * (int64_t) G::PUB + pub_slot<TX, TY, NZC>(plane, lx, ly, lz, axis);
    uint64_t wv = ld_agent(src);
    int spins = 0;
    while ((wv & 0xffffffff00000000ull) != epoch) {
        if (++spins > kSpinLimit || ((spins & 1023) == 0 && (ld_agent32((const uint32_t *) p.err) & ISO_ERR_SPIN))) {
            set_err(p, ISO_ERR_SPIN);
            return -1;
        }
        __builtin_amdgcn_s_sleep(1);
        wv = ld_agent(src);
    }
    return (int64_t) (wv & 0xffffffffull);
}

template <int TX, int TY, int NZC, int NT>
__global__ void __launch_bounds__(NT) mc_dense_kernel(McParams p) {
    using G = TileGeom<TX, TY, NZC>;
    __shared__ Lds<TX, TY, NZC> s;
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    // ---- ticket -> tile (descending linear order: upper neighbours first) ---
    if (tid == 0) s.ticket = (int64_t) atomicAdd(p.ticket, 1u);
    __syncthreads();
    const int64_t ticket = s.ticket;
    TileCtx t;
    t.tile = p.ntiles - 1 - ticket;
    t.tk = (int) (t.tile % p.NK);
    t.tj = (int) ((t.tile / p.NK) % p.NJ);
    t.ti = (int) (t.tile / ((int64_t) p.NK * p.NJ));
    t.x0 = (int64_t) t.ti * TX;
    t.y0 = (int64_t) t.tj * TY;
    t.z0 = (int64_t) t.tk * G::TZ;
    t.nx = (int) min<int64_t>(TX, p.X - 1 - t.x0);
    t.ny = (int) min<int64_t>(TY, p.Y - 1 - t.y0);
    t.nz = (int) min<int64_t>(G::TZ, p.Z - 1 - t.z0);
    t.nxo = t.nx + (t.ti == p.NI - 1 ? 1 : 0);
    t.nyo = t.ny + (t.tj == p.NJ - 1 ? 1 : 0);
    t.nzo = t.nz + (t.tk == p.NK - 1 ? 1 : 0);
    t.nchunk = (t.nz + 1 + 63) >> 6;

    // ---- sweep: stream values once --------------------------------------------
    // The first 64*NZC points of each point row are the row's body: 16*NZC
    // lanes each load 4 consecutive values (one dwordx4), turn them into a
    // 4-bit inside mask and the 16 lanes of each 64-point chunk pack their
    // nibbles into the chunk's 64-bit word.  The one-point z halo is loaded
    // separately.
    {
        constexpr int LPR = 16 * NZC;        // lanes per row body
        constexpr int RPW = 64 / LPR;        // rows per wave instruction
        constexpr int U = 8;                 // row loads in flight per wave
        const int nrow = (t.nx + 1) * (t.ny + 1);
        const int body = min(t.nz + 1, G::TZ);
        const bool vec4 = p.vec4 != 0;
        const int lrow = lane / LPR, lq = lane % LPR;
        const int zq = 4 * lq;
        for (int base = wave * RPW; base < nrow; base += NW * RPW * U) {
            float v[U][4];
            int nvalid[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int row = base + u * NW * RPW + lrow;
                nvalid[u] = 0;
                v[u][0] = v[u][1] = v[u][2] = v[u][3] = 0.0f;
                if (row < nrow) {
                    const int py = row % (t.ny + 1), px = row / (t.ny + 1);
                    const float *src = p.values + ((t.x0 + px) * p.Y + (t.y0 + py)) * p.Z + t.z0 + zq;
                    const int nv4 = min(4, body - zq);
                    nvalid[u] = nv4 < 0 ? 0 : nv4;
                    if (nv4 == 4 && vec4) {
                        const float4 q = *reinterpret_cast<const float4 *>(src);
                        v[u][0] = q.x;
                        v[u][1] = q.y;
                        v[u][2] = q.z;
                        v[u][3] = q.w;
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            if (j < nv4) v[u][j] = src[j];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int row = base + u * NW * RPW + lrow;
                uint32_t x = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) x |= (uint32_t) (j < nvalid[u] && inside(v[u][j], p.level)) << j;
                x |= __shfl_down(x, 1, 16) << 4;
                x |= __shfl_down(x, 2, 16) << 8;
                x |= __shfl_down(x, 4, 16) << 16;
                const uint32_t hi = __shfl_down(x, 8, 16);
                if (row < nrow && (lq & 15) == 0) {
                    const int py = row % (t.ny + 1), px = row / (t.ny + 1);
                    s.w[(px * G::PY + py) * G::PC + (lq >> 4)] = (uint64_t) x | ((uint64_t) hi << 32);
                }
            }
        }
        // halo point z = 64*NZC (present when the tile has a full z extent)
        if (t.nz + 1 > G::TZ) {
            for (int row = tid; row < nrow; row += NT) {
                const int py = row % (t.ny + 1), px = row / (t.ny + 1);
                const float hv = p.values[((t.x0 + px) * p.Y + (t.y0 + py)) * p.Z + t.z0 + G::TZ];
                s.w[(px * G::PY + py) * G::PC + NZC] = inside(hv, p.level) ? 1ull : 0ull;
            }
        }
    }
    __syncthreads();

    // ---- per-unit face / vertex counts (one lane per unit) ------------------
    const int ncz = (t.nzo + 63) >> 6;
    for (int u = tid; u < G::NUNITS; u += NT) {
        int c = u % G::PC, r = u / G::PC, py = r % G::PY, px = r / G::PY;
        uint32_t fc = 0, vc = 0, ac = 0;
        if (px < t.nxo && py < t.nyo && c < ncz) {
            uint64_t xc, yc, zc;
            unit_masks(s, p, t, px, py, c, xc, yc, zc);
            vc = __popcll(xc) + __popcll(yc) + __popcll(zc);
            if (px < t.nx && py < t.ny) {
                int ncell = t.nz - 64 * c;
                uint64_t cm = ncell >= 64 ? ~0ull : (ncell <= 0 ? 0ull : ((1ull << ncell) - 1));
                uint64_t a = W(s, px, py, c), b = W(s, px, py + 1, c);
                uint64_t d0 = W(s, px + 1, py, c), d1 = W(s, px + 1, py + 1, c);
                bool hn = c + 1 < t.nchunk;
                uint64_t as = (a >> 1) | (hn ? W(s, px, py, c + 1) << 63 : 0ull);
                uint64_t bs = (b >> 1) | (hn ? W(s, px, py + 1, c + 1) << 63 : 0ull);
                uint64_t d0s = (d0 >> 1) | (hn ? W(s, px + 1, py, c + 1) << 63 : 0ull);
                uint64_t d1s = (d1 >> 1) | (hn ? W(s, px + 1, py + 1, c + 1) << 63 : 0ull);
                uint64_t any = a | as | b | bs | d0 | d0s | d1 | d1s;
                uint64_t all = a & as & b & bs & d0 & d0s & d1 & d1s;
                uint64_t act = any & ~all & cm;
                ac = (uint32_t) __popcll(act);
                while (act) {
                    int k = __ffsll((unsigned long long) act) - 1;
                    act &= act - 1;
                    int cs = (int) ((a >> k) & 1) | (int) ((as >> k) & 1) << 1 |
                             (int) ((b >> k) & 1) << 2 | (int) ((bs >> k) & 1) << 3 |
                             (int) ((d0 >> k) & 1) << 4 | (int) ((d0s >> k) & 1) << 5 |
                             (int) ((d1 >> k) & 1) << 6 | (int) ((d1s >> k) & 1) << 7;
                    fc += tri_count(p.method, cs);
                }
            }
        }
        s.fo[u] = fc;
        s.vo[u] = vc;
        s.na[u] = ac;
    }
    __syncthreads();

    // ---- block exclusive scan of unit counts ---------------------------------
    {
        constexpr int PER = (G::NUNITS + NT - 1) / NT;
        const int lo = tid * PER, hi = min(lo + PER, G::NUNITS);
        uint32_t sf = 0, sv = 0, sa = 0;
        for (int u = lo; u < hi; ++u) {
            sf += s.fo[u];
            sv += s.vo[u];
            sa += s.na[u];
        }
        // inclusive scan of thread sums across the block
        uint32_t incf = sf, incv = sv, inca = sa;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t yf = __shfl_up(incf, o), yv = __shfl_up(incv, o), ya = __shfl_up(inca, o);
            if (lane >= o) {
                incf += yf;
                incv += yv;
                inca += ya;
            }
        }
        if (lane == 63) {
            s.tsum_f[wave] = incf;
            s.tsum_v[wave] = incv;
            s.tsum_a[wave] = inca;
        }
        __syncthreads();
        uint32_t wf = 0, wv = 0, wa = 0, tot_f = 0, tot_v = 0, tot_a = 0;
        for (int w = 0; w < NW; ++w) {
            if (w < wave) {
                wf += s.tsum_f[w];
                wv += s.tsum_v[w];
                wa += s.tsum_a[w];
            }
            tot_f += s.tsum_f[w];
            tot_v += s.tsum_v[w];
            tot_a += s.tsum_a[w];
        }
        uint32_t ef = wf + incf - sf, ev = wv + incv - sv, ea = wa + inca - sa;
        for (int u = lo; u < hi; ++u) {
            uint32_t f = s.fo[u], v = s.vo[u], a = s.na[u];
            s.fo[u] = ef;
            s.vo[u] = ev;
            s.na[u] = ea;
            ef += f;
            ev += v;
            ea += a;
        }
        __syncthreads();
        if (tid == 0) {
            s.fo[G::NUNITS] = tot_f;
            s.vo[G::NUNITS] = tot_v;
            s.na[G::NUNITS] = tot_a;
        }
    }
    __syncthreads();

    // ---- decoupled look-back over tickets for the tile's global bases --------
    if (wave == 0) {
        const uint64_t tf = s.fo[G::NUNITS], tv = s.vo[G::NUNITS];
        int64_t exf = 0, exv = 0;
        if (ticket == 0) {
            if (lane == 0) st_agent(&p.status[0], pack_status(2, tf, tv));
        } else {
            if (lane == 0) st_agent(&p.status[ticket], pack_status(1, tf, tv));
            int64_t look = ticket - 1;
            bool done = false;
            while (!done) {
                int64_t idx = look - lane;
                uint64_t st = 0;
                if (idx >= 0) {
                    int spins = 0;
                    st = ld_agent(&p.status[idx]);
                    while ((st >> 62) == 0) {
                        if (++spins > kSpinLimit ||
                            ((spins & 1023) == 0 && (ld_agent32((const uint32_t *) p.err) & ISO_ERR_SPIN))) {
                            set_err(p, ISO_ERR_SPIN);
                            st = pack_status(2, 0, 0);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                        st = ld_agent(&p.status[idx]);
                    }
                }
                uint64_t incl = __ballot(idx >= 0 && (st >> 62) == 2);
                int kstop = incl ? (__ffsll((unsigned long long) incl) - 1) : 63;
                bool take = idx >= 0 && lane <= kstop;
                int64_t f = take ? (int64_t) ((st >> 31) & 0x7fffffffull) : 0;
                int64_t v = take ? (int64_t) (st & 0x7fffffffull) : 0;
                exf += wave_sum64(f);
                exv += wave_sum64(v);
                if (incl || look - 63 <= 0) done = true;
                look -= 64;
            }
            if (exf + (int64_t) tf > 0x7fffffffll || exv + (int64_t) tv > 0x7fffffffll) {
                if (lane == 0) set_err(p, ISO_ERR_TOO_LARGE);
            }
            if (lane == 0) st_agent(&p.status[ticket], pack_status(2, exf + tf, exv + tv));
        }
        if (lane == 0) {
            s.fbase = exf;
            s.vbase = exv;
            if (ticket == p.ntiles - 1) {
                p.totals[0] = exv + (int64_t) tv;
                p.totals[1] = exf + (int64_t) tf;
            }
        }
    }
    __syncthreads();
    const int64_t fbase = s.fbase, vbase = s.vbase;
    const uint64_t epoch = (uint64_t) (*p.epoch_ptr) << 32;
    uint64_t *pub = p.pub + t.tile * (int64_t) G::PUB;

    // ---- publish ids of crossing in-plane edges on the lower boundary planes -
    // plane px == 0 (read by tile ti-1): edges y, z of points (0, py, z)
    if (t.ti > 0) {
        for (int u = tid; u < t.nyo * ncz; u += NT) {
            int py = u / ncz, c = u % ncz;
            uint64_t xc, yc, zc;
            unit_masks(s, p, t, 0, py, c, xc, yc, zc);
            uint64_t m = yc | zc;
            while (m) {
                int b = __ffsll((unsigned long long) m) - 1;
                m &= m - 1;
                int z = 64 * c + b;
                if ((yc >> b) & 1)
                    st_agent(&pub[pub_slot<TX, TY, NZC>(0, 0, py, z, 1)],
                             epoch | (uint64_t) (vbase + local_vid(s, p, t, 0, py, z, 1)));
                if ((zc >> b) & 1)
                    st_agent(&pub[pub_slot<TX, TY, NZC>(0, 0, py, z, 2)],
                             epoch | (uint64_t) (vbase + local_vid(s, p, t, 0, py, z, 2)));
            }
        }
    }
    // plane py == 0 (read by tile tj-1): edges x, z of points (px, 0, z)
    if (t.tj > 0) {
        for (int u = tid; u < t.nxo * ncz; u += NT) {
            int px = u / ncz, c = u % ncz;
            uint64_t xc, yc, zc;
            unit_masks(s, p, t, px, 0, c, xc, yc, zc);
            uint64_t m = xc | zc;
            while (m) {
                int b = __ffsll((unsigned long long) m) - 1;
                m &= m - 1;
                int z = 64 * c + b;
                if ((xc >> b) & 1)
                    st_agent(&pub[pub_slot<TX, TY, NZC>(1, px, 0, z, 0)],
                             epoch | (uint64_t) (vbase + local_vid(s, p, t, px, 0, z, 0)));
                if ((zc >> b) & 1)
                    st_agent(&pub[pub_slot<TX, TY, NZC>(1, px, 0, z, 2)],
                             epoch | (uint64_t) (vbase + local_vid(s, p, t, px, 0, z, 2)));
            }
        }
    }
    // plane z == 0 (read by tile tk-1): edges x, y of points (px, py, 0)
    if (t.tk > 0) {
        for (int u = tid; u < t.nxo * t.nyo; u += NT) {
            int px = u / t.nyo, py = u % t.nyo;
            uint64_t xc, yc, zc;
            unit_masks(s, p, t, px, py, 0, xc, yc, zc);
            if (xc & 1)
                st_agent(&pub[pub_slot<TX, TY, NZC>(2, px, py, 0, 0)],
                         epoch | (uint64_t) (vbase + local_vid(s, p, t, px, py, 0, 0)));
            if (yc & 1)
                st_agent(&pub[pub_slot<TX, TY, NZC>(2, px, py, 0, 1)],
                         epoch | (uint64_t) (vbase + local_vid(s, p, t, px, py, 0, 1)));
        }
    }

    // ---- emit: one lane per active cell (cells compacted through the scan) ---
    const float lvl = p.level;
    const int nactive = (int) s.na[G::NUNITS];
    for (int j0 = 0; j0 < nactive; j0 += NT) {
        const int j = j0 + tid;
        if (j >= nactive) break;
        // unit holding the j-th active cell: last u with na[u] <= j
        int lo = 0, hi = G::NUNITS;
        while (hi - lo > 1) {
            int mid = (lo + hi) >> 1;
            if (s.na[mid] <= (uint32_t) j) lo = mid;
            else hi = mid;
        }
        const int u = lo;
        int rank = j - (int) s.na[u];
        const int c = u % G::PC, r = u / G::PC, py = r % G::PY, px = r / G::PY;
        const int ncell = t.nz - 64 * c;
        const uint64_t cm = ncell >= 64 ? ~0ull : ((1ull << ncell) - 1);
        const uint64_t a = W(s, px, py, c), b = W(s, px, py + 1, c);
        const uint64_t d0 = W(s, px + 1, py, c), d1 = W(s, px + 1, py + 1, c);
        const bool hn = c + 1 < t.nchunk;
        const uint64_t as = (a >> 1) | (hn ? W(s, px, py, c + 1) << 63 : 0ull);
        const uint64_t bs = (b >> 1) | (hn ? W(s, px, py + 1, c + 1) << 63 : 0ull);
        const uint64_t d0s = (d0 >> 1) | (hn ? W(s, px + 1, py, c + 1) << 63 : 0ull);
        const uint64_t d1s = (d1 >> 1) | (hn ? W(s, px + 1, py + 1, c + 1) << 63 : 0ull);
        uint64_t act = (a | as | b | bs | d0 | d0s | d1 | d1s) & ~(a & as & b & bs & d0 & d0s & d1 & d1s) & cm;
        // skip the active cells before this one, adding up their faces
        int64_t fslot = fbase + s.fo[u];
        for (; rank > 0; --rank) {
            int k0 = __ffsll((unsigned long long) act) - 1;
            act &= act - 1;
            int cs0 = (int) ((a >> k0) & 1) | (int) ((as >> k0) & 1) << 1 | (int) ((b >> k0) & 1) << 2 |
                      (int) ((bs >> k0) & 1) << 3 | (int) ((d0 >> k0) & 1) << 4 |
                      (int) ((d0s >> k0) & 1) << 5 | (int) ((d1 >> k0) & 1) << 6 |
                      (int) ((d1s >> k0) & 1) << 7;
            fslot += tri_count(p.method, cs0);
        }
        const int64_t gx = t.x0 + px, gy = t.y0 + py;
        const float cx0 = grid_coord(gx, p.X, p.amin[0], p.amax[0]);
        const float cx1 = grid_coord(gx + 1, p.X, p.amin[0], p.amax[0]);
        const float cy0 = grid_coord(gy, p.Y, p.amin[1], p.amax[1]);
        const float cy1 = grid_coord(gy + 1, p.Y, p.amin[1], p.amax[1]);
        {
            const int k = __ffsll((unsigned long long) act) - 1;
            const int zl = 64 * c + k;
            const int64_t gz = t.z0 + zl;
            int cs = (int) ((a >> k) & 1) | (int) ((as >> k) & 1) << 1 | (int) ((b >> k) & 1) << 2 |
                     (int) ((bs >> k) & 1) << 3 | (int) ((d0 >> k) & 1) << 4 |
                     (int) ((d0s >> k) & 1) << 5 | (int) ((d1 >> k) & 1) << 6 |
                     (int) ((d1s >> k) & 1) << 7;
            // corner values (gathered; the tile was just streamed)
            CellCtx cc;
            {
                const float *r00 = p.values + (gx * p.Y + gy) * p.Z + gz;
                const float *r01 = r00 + p.Z;
                const float *r10 = r00 + p.Y * p.Z;
                const float *r11 = r10 + p.Z;
                cc.cv[0] = r00[0];
                cc.cv[1] = r00[1];
                cc.cv[2] = r01[0];
                cc.cv[3] = r01[1];
                cc.cv[4] = r10[0];
                cc.cv[5] = r10[1];
                cc.cv[6] = r11[0];
                cc.cv[7] = r11[1];
            }
            cc.c[0][0] = cx0;
            cc.c[0][1] = cx1;
            cc.c[1][0] = cy0;
            cc.c[1][1] = cy1;
            cc.c[2][0] = grid_coord(gz, p.Z, p.amin[2], p.amax[2]);
            cc.c[2][1] = grid_coord(gz + 1, p.Z, p.amin[2], p.amax[2]);
            cc.level = lvl;
            const int status = crossing_status(cs);
            // weld candidates: a vertex within a few ulps of a grid corner
            uint32_t candm = 0;
            {
                uint32_t m = (uint32_t) status;
                while (m) {
                    int e = __ffs(m) - 1;
                    m &= m - 1;
                    if (p.all_special || edge_is_candidate(cc, e)) candm |= 1u << e;
                }
            }
            const bool special = candm != 0;
            // vertices emitted by this cell: edges of owned corners that have
            // no lower-indexed cell in the tile to emit them
            {
                bool lastx = (px == t.nx - 1) && (t.nxo == t.nx + 1);
                bool lasty = (py == t.ny - 1) && (t.nyo == t.ny + 1);
                bool lastz = (zl == t.nz - 1) && (t.nzo == t.nz + 1);
                uint32_t emit = (1u << 0) | (1u << 3) | (1u << 8);          // v0: z, y, x
                if (lastx) emit |= (1u << 7) | (1u << 4);                    // v4: y, z
                if (lasty) emit |= (1u << 11) | (1u << 2);                   // v2: x, z
                if (lastz) emit |= (1u << 9) | (1u << 1);                    // v1: x, y
                if (lastx && lasty) emit |= 1u << 6;                         // v6: z
                if (lastx && lastz) emit |= 1u << 5;                         // v5: y
                if (lasty && lastz) emit |= 1u << 10;                        // v3: x
                uint32_t m = emit & (uint32_t) status;
                while (m) {
                    int e = __ffs(m) - 1;
                    m &= m - 1;
                    // emitted edges are owned by this tile
                    int ca = kEdgeA[e];
                    int64_t v = vbase + local_vid(s, p, t, px + (ca >> 2), py + ((ca >> 1) & 1),
                                                  zl + (ca & 1), kEdgeAxis[e]);
                    float x, y, z;
                    edge_pos(cc, e, x, y, z);
                    if (v >= 0 && v < p.cap_v) {
                        p.out_v[3 * v + 0] = x;
                        p.out_v[3 * v + 1] = y;
                        p.out_v[3 * v + 2] = z;
                        if (p.out_key) p.out_key[v] = edge_key(p, gx, gy, gz, e);
                    } else {
                        set_err(p, ISO_ERR_CAPACITY);
                    }
                    if ((candm >> e) & 1) {
                        int64_t i = list_push(&p.sp->n_cand, p.sp->cap_cand);
                        if (i >= 0) {
                            SpecialVertex sv;
                            sv.vid = v;
                            sv.edge = edge_key(p, gx, gy, gz, e);
                            sv.x = x;
                            sv.y = y;
                            sv.z = z;
                            sv.pad = 0.0f;
                            p.sp->cand[i] = sv;
                        }
                    }
                }
            }
            // faces
            const uint64_t w = cTris[p.method][cs];
            const int ntri = (int) (w & 15);
            for (int q = 0; q < ntri; ++q) {
                int e0 = (int) ((w >> (4 + 12 * q)) & 15);
                int e1 = (int) ((w >> (8 + 12 * q)) & 15);
                int e2 = (int) ((w >> (12 + 12 * q)) & 15);
                int64_t slot = fslot++;
                bool keep = true;
                if (special) {
                    float x0, y0, z0, x1, y1, z1, x2, y2, z2;
                    edge_pos(cc, e0, x0, y0, z0);
                    edge_pos(cc, e1, x1, y1, z1);
                    edge_pos(cc, e2, x2, y2, z2);
                    bool d01 = x0 == x1 && y0 == y1 && z0 == z1;
                    bool d02 = x0 == x2 && y0 == y2 && z0 == z2;
                    bool d12 = x1 == x2 && y1 == y2 && z1 == z2;
                    keep = !(d01 || d02 || d12);
                    int64_t i = list_push(keep ? &p.sp->n_kept : &p.sp->n_hole,
                                          keep ? p.sp->cap_kept : p.sp->cap_hole);
                    if (i >= 0) (keep ? p.sp->kept : p.sp->hole)[i] = slot;
                }
                if (slot < p.cap_f) {
                    int32_t *dst = p.out_f + 3 * slot;
                    if (keep) {
                        dst[0] = (int32_t) edge_vid<TX, TY, NZC>(s, p, t, px, py, zl, e0, vbase, epoch);
                        dst[1] = (int32_t) edge_vid<TX, TY, NZC>(s, p, t, px, py, zl, e1, vbase, epoch);
                        dst[2] = (int32_t) edge_vid<TX, TY, NZC>(s, p, t, px, py, zl, e2, vbase, epoch);
                    } else {
                        dst[0] = dst[1] = dst[2] = -1;
                    }
                } else {
                    set_err(p, ISO_ERR_CAPACITY);
                }
            }
            if (special) {
                uint32_t m = (uint32_t) status;
                while (m) {
                    int e = __ffs(m) - 1;
                    m &= m - 1;
                    int64_t i = list_push(&p.sp->n_scell, p.sp->cap_scell);
                    if (i >= 0) {
                        SpecialEdge se;
                        se.vid = edge_vid<TX, TY, NZC>(s, p, t, px, py, zl, e, vbase, epoch);
                        se.edge = edge_key(p, gx, gy, gz, e);
                        p.sp->scell[i] = se;
                    }
                }
            }
        }
    }
}

__global__ void mc_dense_reset(uint64_t *status, int64_t n, uint32_t *ticket, uint32_t *epoch,
                               int *err, int64_t *totals, SpecialLists *sp, SpecialLists init) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (int64_t) gridDim.x * blockDim.x) status[i] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *ticket = 0;
        *epoch = *epoch + 1;
        *err = 0;
        totals[0] = totals[1] = 0;
        init.n_cand = init.n_scell = init.n_kept = init.n_hole = 0;
        *sp = init;
    }
}

// ---- host side ---------------------------------------------------------------

template <int TX, int TY, int NZC, int NT>
struct Cfg {
    using G = TileGeom<TX, TY, NZC>;
    static void tiles(const int64_t shape[3], int &ni, int &nj, int &nk) {
        ni = (int) ((shape[0] - 1 + TX - 1) / TX);
        nj = (int) ((shape[1] - 1 + TY - 1) / TY);
        nk = (int) ((shape[2] - 1 + G::TZ - 1) / G::TZ);
    }
    static hipError_t launch(McParams &p, hipStream_t stream) {
        hipLaunchKernelGGL((mc_dense_kernel<TX, TY, NZC, NT>), dim3((unsigned) p.ntiles), dim3(NT), 0,
                           stream, p);
        return hipGetLastError();
    }
};

// Tile configurations the autotuner may choose from.
#define ISO_MC_CONFIGS(X)       \
    X(0, 16, 16, 4, 256)        \
    X(1, 32, 16, 4, 256)        \
    X(2, 16, 16, 2, 256)        \
    X(3, 8, 8, 4, 256)          \
    X(4, 32, 32, 4, 512)        \
    X(5, 16, 8, 4, 256)

int mc_config_count() {
    int n = 0;
#define X(id, tx, ty, nzc, nt) ++n;
    ISO_MC_CONFIGS(X)
#undef X
    return n;
}

bool mc_config_tiles(int cfg, const int64_t shape[3], int &ni, int &nj, int &nk, int64_t &pub) {
    switch (cfg) {
#define X(id, tx, ty, nzc, nt)                          \
    case id:                                            \
        Cfg<tx, ty, nzc, nt>::tiles(shape, ni, nj, nk); \
        pub = TileGeom<tx, ty, nzc>::PUB;               \
        return true;
        ISO_MC_CONFIGS(X)
#undef X
    default:
        return false;
    }
}

hipError_t mc_upload_luts() {
    static bool done = false;
    if (done) return hipSuccess;
    uint64_t h[2][256];
    for (int i = 0; i < 256; ++i) {
        h[0][i] = kNagaeTris[i];
        h[1][i] = kLorensenTris[i];
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cTris), h, sizeof(h));
    if (e == hipSuccess) done = true;
    return e;
}

hipError_t mc_dense_launch(int cfg, McParams &p, const SpecialLists &sp_init, hipStream_t stream,
                           hipEvent_t ev0, hipEvent_t ev1) {
    hipError_t e = mc_upload_luts();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mc_dense_reset, dim3((unsigned) std::min<int64_t>(1024, (p.ntiles + 255) / 256 + 1)),
                       dim3(256), 0, stream, p.status, p.ntiles, p.ticket, p.epoch_ptr, p.err,
                       p.totals, p.sp, sp_init);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the kernel reads the epoch from device memory after the reset bumps it
    if (ev0) {
        e = hipEventRecord(ev0, stream);
        if (e != hipSuccess) return e;
    }
    switch (cfg) {
#define X(id, tx, ty, nzc, nt)                     \
    case id:                                       \
        e = Cfg<tx, ty, nzc, nt>::launch(p, stream); \
        break;
        ISO_MC_CONFIGS(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
    if (e == hipSuccess && ev1) e = hipEventRecord(ev1, stream);
    return e;
}

}   // namespace iso
