// SparseGrid kernels: corner points, crossing-cell filter, and marching cubes
// over per-cell corner values with an exact position weld.
//
// Reference path replaced:
//   SparseGrid::get_points_from_cell_indices  src/grid/sparse.cu:21-36
//   SparseGrid::filter_cell_indices           src/grid/sparse.cu:150-179
//   mc::marching_cubes on a SparseGrid        src/mc/mc.cu:17-68 with the
//     identity get_cells (src/grid/sparse.cu:61-69) and vertex_welding
//     (src/utils.cu:32-59)
//
// Each sparse cell carries its own 8 corner values, so two cells sharing an
// edge may disagree about its vertex; the reference welds by position only.
// Sparse MC therefore welds instances with equal positions (equality on
// floats), with the representative of a position = the smallest (slot * 12
// + edge) that produces it; vertex ids follow representative order and
// faces follow (slot, LUT slot) order, so output is deterministic.  The
// instances that can share a position are found on the sorted ids: a column
// table gives the slots of every (x, y) cell column, and a neighbour is a
// binary search inside its column (see k_smc_inst).
// Degenerate triangles (two equal corners) are dropped as in
// src/mc/nagae.cu:71; a position is emitted only if a kept triangle uses it.
// The per-vertex representative is exported for the backward pass.

#include <atomic>
#include <mutex>
#include "common.h"
#include "mc_luts.h"
#include "sparse.h"

namespace iso {

namespace {

constexpr int kB = 256;

static __constant__ uint64_t cSTris[2][256];

__device__ __forceinline__ void cell_xyz(int64_t id, int64_t Y, int64_t Z, int64_t &x, int64_t &y, int64_t &z) {
    const int64_t cz = Z - 1, cy = Y - 1;
    const int64_t r = div_fast(id, cz);
    z = id - r * cz;
    x = div_fast(r, cy);
    y = r - x * cy;
}

__global__ void k_corner_points(const int64_t *ids, int64_t n, SparseGeom g, float *out) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    int64_t x, y, z;
    cell_xyz(ids[i], g.Y, g.Z, x, y, z);
    float c[3][2] = {{grid_coord(x, g.X, g.amin[0], g.amax[0]), grid_coord(x + 1, g.X, g.amin[0], g.amax[0])},
                     {grid_coord(y, g.Y, g.amin[1], g.amax[1]), grid_coord(y + 1, g.Y, g.amin[1], g.amax[1])},
                     {grid_coord(z, g.Z, g.amin[2], g.amax[2]), grid_coord(z + 1, g.Z, g.amin[2], g.amax[2])}};
    float *o = out + 24 * i;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        o[3 * k + 0] = c[0][k >> 2];
        o[3 * k + 1] = c[1][(k >> 1) & 1];
        o[3 * k + 2] = c[2][k & 1];
    }
}

__global__ void k_crossing_flags(const float *values, int64_t n, float level, uint8_t *flags) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) c |= inside(values[8 * i + k], level) << k;
    flags[i] = (c != 0 && c != 255) ? 1 : 0;
}

struct SCell {
    float cv[8];
    float c[3][2];
    int cs;
};

__device__ __forceinline__ bool load_cell(const SparseMcIn &in, int64_t i, SCell &s) {
    int cs = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        s.cv[k] = in.values[8 * i + k];
        cs |= inside(s.cv[k], in.level) << k;
    }
    s.cs = cs;
    if (cs == 0 || cs == 255) return false;
    int64_t x, y, z;
    cell_xyz(in.ids[i], in.g.Y, in.g.Z, x, y, z);
    s.c[0][0] = grid_coord(x, in.g.X, in.g.amin[0], in.g.amax[0]);
    s.c[0][1] = grid_coord(x + 1, in.g.X, in.g.amin[0], in.g.amax[0]);
    s.c[1][0] = grid_coord(y, in.g.Y, in.g.amin[1], in.g.amax[1]);
    s.c[1][1] = grid_coord(y + 1, in.g.Y, in.g.amin[1], in.g.amax[1]);
    s.c[2][0] = grid_coord(z, in.g.Z, in.g.amin[2], in.g.amax[2]);
    s.c[2][1] = grid_coord(z + 1, in.g.Z, in.g.amin[2], in.g.amax[2]);
    return true;
}

__device__ __forceinline__ void spos(const SCell &s, float level, int e, float &x, float &y, float &z) {
    int ca = edge_a(e), cb = edge_b(e);
    float t = edge_t(s.cv[ca], s.cv[cb], level);
    x = lerp1(t, s.c[0][ca >> 2], s.c[0][cb >> 2]);
    y = lerp1(t, s.c[1][(ca >> 1) & 1], s.c[1][(cb >> 1) & 1]);
    z = lerp1(t, s.c[2][ca & 1], s.c[2][cb & 1]);
}

__device__ __forceinline__ int scross(int cs) {
    int st = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (((cs >> edge_a(e)) & 1) != ((cs >> edge_b(e)) & 1)) st |= 1 << e;
    return st;
}

// Pass 1: crossing-edge instances per cell -> block sums; optionally each
// cell's crossing-edge mask.
__global__ void k_smc_count(SparseMcIn in, int64_t *bsum, uint32_t *cst) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    int64_t k = 0;
    if (i < in.n) {
        int cs = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) cs |= inside(in.values[8 * i + j], in.level) << j;
        const int st = (cs != 0 && cs != 255) ? scross(cs) : 0;
        k = __popc(st);
        if (cst) cst[i] = (uint32_t) st | ((uint32_t) cs << 16);
    }
    __shared__ int64_t sh[kB];
    sh[threadIdx.x] = k;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned) o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

__device__ __forceinline__ int64_t block_excl(int64_t v, int64_t *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kB; o <<= 1) {
        int64_t y = threadIdx.x >= (unsigned) o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    int64_t r = sh[threadIdx.x] - v;
    __syncthreads();
    return r;
}

// ---- neighbour lookup on the sorted ids ------------------------------------------
// The cells of column (x, y) (ids (x (Y-1) + y) (Z-1) + z, sorted) are the
// slots [cf[c], cf[c + 1]) of that column c = x (Y-1) + y, cf[c] the first
// slot whose column is >= c (cf[ncol] = n); a neighbour is found by a binary
// search inside its column.
__global__ void k_smc_cols(const int64_t *ids, int64_t n, int64_t cz, int32_t *cf) {
    const int64_t s = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (s >= n) return;
    const int64_t col = div_fast(ids[s], cz);
    if (s == 0 || div_fast(ids[s - 1], cz) != col) cf[col] = (int32_t) s;
}

// Suffix minimum over cf (empty columns take the next column's start):
// per 1024-entry chunk, then over the chunk minima, then applied.
__global__ void k_cols_chunk(int32_t *cf, int64_t m, int32_t *cmin) {
    __shared__ int32_t sh[1024];
    const int64_t i = (int64_t) blockIdx.x * 1024 + (1023 - threadIdx.x);   // reversed within the chunk
    sh[threadIdx.x] = i < m ? cf[i] : INT32_MAX;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int32_t y = threadIdx.x >= (unsigned) o ? sh[threadIdx.x - o] : INT32_MAX;
        __syncthreads();
        sh[threadIdx.x] = min(sh[threadIdx.x], y);
        __syncthreads();
    }
    if (i < m) cf[i] = sh[threadIdx.x];
    if (threadIdx.x == 1023) cmin[blockIdx.x] = sh[1023];
}
// Chunk minima -> minimum of the chunks after each (one workgroup; each
// thread a contiguous run).
__global__ void k_cols_top(int32_t *cmin, int64_t nc) {
    __shared__ int32_t sh[1024];
    const int t = threadIdx.x;
    const int64_t per = (nc + 1023) / 1024, lo = min<int64_t>(t * per, nc), hi = min<int64_t>(lo + per, nc);
    int32_t m = INT32_MAX;
    for (int64_t c = lo; c < hi; ++c) m = min(m, cmin[c]);
    sh[t] = m;
    __syncthreads();
    // suffix minimum over the threads' runs (reversed inclusive scan)
    for (int o = 1; o < 1024; o <<= 1) {
        const int32_t y = t + o < 1024 ? sh[t + o] : INT32_MAX;
        __syncthreads();
        sh[t] = min(sh[t], y);
        __syncthreads();
    }
    int32_t run = t + 1 < 1024 ? sh[t + 1] : INT32_MAX;   // minimum of the runs after this one
    for (int64_t c = hi - 1; c >= lo; --c) {
        const int32_t v = cmin[c];
        cmin[c] = run;
        run = min(run, v);
    }
}
__global__ void k_cols_apply(int32_t *cf, int64_t m, const int32_t *cmin) {
    const int64_t i = (int64_t) blockIdx.x * 1024 + threadIdx.x;
    if (i < m) cf[i] = min(cf[i], cmin[blockIdx.x]);
}

struct SLook {
    const int64_t *ids;
    const int32_t *cf;
    int64_t cx, cy, cz;   // cells per axis
};

// Slot of cell (x, y, z), or -1.
__device__ __forceinline__ int64_t find_cell(const SLook &L, int64_t x, int64_t y, int64_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= L.cx || y >= L.cy || z >= L.cz) return -1;
    const int64_t col = x * L.cy + y, id = col * L.cz + z;
    int lo = L.cf[col];
    const int end = L.cf[col + 1];
    int hi = end;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L.ids[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && L.ids[lo] == id) ? lo : -1;
}

// Cell edge from corner c (its bit along axis a clear) along axis a.
__device__ __forceinline__ int edge_of(int c, int a) {
    const int k = a * 8 + c;
    return k < 16 ? (int) ((0xff57ff13ffffab98ull >> (4 * k)) & 15) : (int) ((0xf6f4f2f0u >> (4 * (k - 16))) & 15);
}

// Position of the vertex cell slot s (origin (x, y, z)) places on its edge e,
// if that edge crosses the level in the cell (same arithmetic as spos).
__device__ __forceinline__ bool inst_pos(const SparseMcIn &in, int64_t s, int64_t x, int64_t y, int64_t z, int e,
                                         float &px, float &py, float &pz) {
    const int ca = edge_a(e), cb = edge_b(e);
    const float va = in.values[8 * s + ca], vb = in.values[8 * s + cb];
    if (inside(va, in.level) == inside(vb, in.level)) return false;
    const float t = edge_t(va, vb, in.level);
    px = lerp1(t, grid_coord(x + (ca >> 2), in.g.X, in.g.amin[0], in.g.amax[0]),
               grid_coord(x + (cb >> 2), in.g.X, in.g.amin[0], in.g.amax[0]));
    py = lerp1(t, grid_coord(y + ((ca >> 1) & 1), in.g.Y, in.g.amin[1], in.g.amax[1]),
               grid_coord(y + ((cb >> 1) & 1), in.g.Y, in.g.amin[1], in.g.amax[1]));
    pz = lerp1(t, grid_coord(z + (ca & 1), in.g.Z, in.g.amin[2], in.g.amax[2]),
               grid_coord(z + (cb & 1), in.g.Z, in.g.amin[2], in.g.amax[2]));
    return true;
}

// The 9 cells sharing an edge with a cell that come before it in id order
// (offsets with a negative id step): index of offset (dx, dy, dz), or -1.
__device__ __forceinline__ int lower_nb(int dx, int dy, int dz) {
    const int k = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);   // 0..26
    // (0,0,-1) (0,-1,0) (-1,0,0) (0,-1,-1) (0,-1,1) (-1,0,-1) (-1,0,1) (-1,-1,0) (-1,1,0)
    switch (k) {
    case 12: return 0;
    case 10: return 1;
    case 4: return 2;
    case 9: return 3;
    case 11: return 4;
    case 3: return 5;
    case 5: return 6;
    case 1: return 7;
    case 7: return 8;
    default: return -1;
    }
}
__device__ __forceinline__ void lower_off(int k, int &dx, int &dy, int &dz) {
    constexpr int8_t t[9][3] = {{0, 0, -1}, {0, -1, 0}, {-1, 0, 0}, {0, -1, -1}, {0, -1, 1},
                                {-1, 0, -1}, {-1, 0, 1}, {-1, -1, 0}, {-1, 1, 0}};
    dx = t[k][0];
    dy = t[k][1];
    dz = t[k][2];
}

// Pass 2: each instance's position and representative key (slot * 12 +
// edge of the first instance in (slot, edge) order with the same position).
// A vertex more than 16 ulps from both endpoints along its edge can only
// share its position with the same grid edge in the (at most 3) other cells
// around it, and only those before this cell can lower the key: their
// slots are looked up once per cell (independent searches), and an edge's
// copy in such a cell is compared by its two end values first.  A vertex
// nearer an endpoint P can share its position with any edge through P in
// the 8 cells around P, which are searched directly (rare).


// Pass 2a: each cell's first instance (block scan of its crossing-edge
// count) and, per instance, its cell.
__global__ void __launch_bounds__(kB) k_smc_cells(int64_t n, const int64_t *bpre, SparseMcWork w) {
    __shared__ int64_t sh[kB];
    const int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    const int st = i < n ? (int) (w.cst[i] & 0xfffu) : 0;
    const int64_t off = bpre[blockIdx.x] + block_excl(__popc(st), sh);
    if (!st) return;
    w.cell_inst[i] = off;
    for (int j = 0; j < __popc(st); ++j) w.inst_cell[off + j] = (int32_t) i;
}

// Slot of the cell with id `key` among the slots [lo, hi) of its column, or -1.
__device__ __forceinline__ int64_t col_find(const int64_t *ids, int64_t lo, int64_t hi, int64_t key) {
    int64_t a = lo, b = hi;
    while (a < b) {
        const int64_t m = (a + b) >> 1;
        if (ids[m] < key) a = m + 1;
        else b = m;
    }
    return (a < hi && ids[a] == key) ? a : -1;
}

// Pass 2b: one lane per instance (cell slot i, its j-th crossing edge e):
// position and representative key (slot * 12 + edge of the first instance
// in (slot, edge) order with the same position).  A vertex more than 16 ulps
// from both endpoints along its edge can only share its position with the
// same grid edge in the (at most 3) other cells around that edge, and only
// those before this cell can lower the key; their slots come from the
// column table (the previous slot for z - 1, a short search in a
// neighbouring column otherwise), and an edge's copy is compared by its two
// end values first.  A vertex nearer an endpoint P can share its position
// with any edge through P in the 8 cells around P (rare).
__global__ void __launch_bounds__(kB) k_smc_inst(SparseMcIn in, SLook L, SparseMcWork w, int64_t ninst) {
    const int64_t q = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (q >= ninst) return;
    const int64_t i = w.inst_cell[q];
    const int st = (int) (w.cst[i] & 0xfffu);
    int e = 0;
    {
        int r = (int) (q - w.cell_inst[i]);
        uint32_t m = (uint32_t) st;
        while (r-- > 0) m &= m - 1;
        e = __ffs(m) - 1;
    }
    SCell s;
    load_cell(in, i, s);
    float px, py, pz;
    spos(s, in.level, e, px, py, pz);
    int64_t x, y, z;
    cell_xyz(in.ids[i], in.g.Y, in.g.Z, x, y, z);
    const int ca = edge_a(e), cb = edge_b(e), ax = edge_axis(e);
    const int sa = ax == 0 ? (ca >> 2) : (ax == 1 ? ((ca >> 1) & 1) : (ca & 1));
    const int sb = ax == 0 ? (cb >> 2) : (ax == 1 ? ((cb >> 1) & 1) : (cb & 1));
    const float ra = ax == 0 ? px : (ax == 1 ? py : pz);
    int64_t rep = i * 12 + e;
    if (!near_endpoint(ra, s.c[ax][sa], s.c[ax][sb])) {
        const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
        const int bu = (ca >> (2 - u)) & 1, bv = (ca >> (2 - v)) & 1;
        const int64_t col = x * L.cy + y;
        for (int d = 0; d < 4; ++d) {
            const int du = d & 1, dv = d >> 1;
            int dd[3] = {0, 0, 0};
            dd[u] = bu - du;
            dd[v] = bv - dv;
            if (lower_nb(dd[0], dd[1], dd[2]) < 0) continue;   // this cell, or a later one
            const int64_t xx = x + dd[0], yy = y + dd[1], zz = z + dd[2];
            if (xx < 0 || yy < 0 || zz < 0 || yy >= L.cy || zz >= L.cz) continue;
            int64_t s2;
            if (dd[0] == 0 && dd[1] == 0) {   // (0, 0, -1): the previous slot
                s2 = (i > 0 && in.ids[i - 1] == in.ids[i] - 1) ? i - 1 : -1;
            } else {
                const int64_t c2 = col + dd[0] * L.cy + dd[1];
                s2 = col_find(L.ids, L.cf[c2], L.cf[c2 + 1], c2 * L.cz + zz);
            }
            if (s2 < 0) continue;
            const int cc = (du << (2 - u)) | (dv << (2 - v));
            const int e2 = edge_of(cc, ax);
            const float va = in.values[8 * s2 + cc], vb = in.values[8 * s2 + (cc | (4 >> ax))];
            bool same;
            if (va == s.cv[ca] && vb == s.cv[cb]) {
                same = true;   // the same end values: the same vertex
            } else {
                if (inside(va, in.level) == inside(vb, in.level)) continue;
                const float t = edge_t(va, vb, in.level);
                const float qx = lerp1(t, s.c[0][ca >> 2], s.c[0][cb >> 2]);
                const float qy = lerp1(t, s.c[1][(ca >> 1) & 1], s.c[1][(cb >> 1) & 1]);
                const float qz = lerp1(t, s.c[2][ca & 1], s.c[2][cb & 1]);
                same = qx == px && qy == py && qz == pz;
            }
            if (same) rep = min(rep, s2 * 12 + e2);
        }
    } else {
        // the endpoint it is near, and the 3 edges through it in each cell around it
        const int64_t kr = ordered_key(ra), ka = ordered_key(s.c[ax][sa]), kb2 = ordered_key(s.c[ax][sb]);
        const int pc = (llabs(kr - ka) <= llabs(kr - kb2)) ? ca : cb;
        const int64_t P[3] = {x + (pc >> 2), y + ((pc >> 1) & 1), z + (pc & 1)};
        for (int d = 0; d < 8; ++d) {
            const int64_t c0 = P[0] - (d >> 2), c1 = P[1] - ((d >> 1) & 1), c2 = P[2] - (d & 1);
            const int64_t s2 = find_cell(L, c0, c1, c2);
            if (s2 < 0 || s2 > i) continue;
            for (int a2 = 0; a2 < 3; ++a2) {
                const int e2 = edge_of(d & ~(4 >> a2), a2);
                if (s2 == i && e2 == e) continue;
                float qx, qy, qz;
                if (inst_pos(in, s2, c0, c1, c2, e2, qx, qy, qz) && qx == px && qy == py && qz == pz)
                    rep = min(rep, s2 * 12 + e2);
            }
        }
    }
    w.ipos[3 * q + 0] = px;
    w.ipos[3 * q + 1] = py;
    w.ipos[3 * q + 2] = pz;
    w.isrc[q] = i * 12 + e;
    // the representative as an instance index (its cell's first instance +
    // its rank among the cell's crossing edges)
    if (rep != i * 12 + e) {
        const int64_t s2 = rep / 12;
        const int e2 = (int) (rep - s2 * 12);
        w.rkey[q] = w.cell_inst[s2] + __popc((int) (w.cst[s2] & 0xfffu) & ((1 << e2) - 1));
    } else {
        w.rkey[q] = q;
    }
}

__device__ __forceinline__ int64_t inst_of(int st, int e, int64_t base) {
    return base + __popc(st & ((1 << e) - 1));
}


// Pass 3: kept faces per cell (a triangle with two corners at one position is
// dropped); the representatives of the kept faces' corners are marked used.
__global__ void k_smc_faces(SparseMcIn in, SparseMcWork w, int64_t *bsum) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    int64_t nf = 0;
    if (i < in.n) {
        const uint32_t cw = w.cst[i];
        const int st = (int) (cw & 0xfffu);
        if (st) {
            const int cs = (int) ((cw >> 16) & 0xffu);
            const int64_t base = w.cell_inst[i];
            const uint64_t tw = cSTris[in.method][cs];
            const int nt = (int) (tw & 15);
            uint32_t kept = 0;
            for (int q = 0; q < nt; ++q) {
                const int64_t r0 = w.rkey[inst_of(st, (int) ((tw >> (4 + 12 * q)) & 15), base)];
                const int64_t r1 = w.rkey[inst_of(st, (int) ((tw >> (8 + 12 * q)) & 15), base)];
                const int64_t r2 = w.rkey[inst_of(st, (int) ((tw >> (12 + 12 * q)) & 15), base)];
                if (r0 == r1 || r0 == r2 || r1 == r2) continue;
                ++nf;
                kept |= 1u << q;
                w.used[r0] = 1;
                w.used[r1] = 1;
                w.used[r2] = 1;
            }
            w.cst[i] = cw | (kept << 24);   // kept triangles for the emit pass
        }
    }
    __shared__ int64_t sh[kB];
    sh[threadIdx.x] = nf;
    __syncthreads();
    for (int o = kB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned) o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

// Output vertices: used representatives, one bit per instance (64 per wave).
__global__ void k_smc_vbits(SparseMcWork w, int64_t ninst, uint32_t *bits) {
    const int64_t q = (int64_t) blockIdx.x * kB + threadIdx.x;
    const bool is_v = q < ninst && w.used[q] && w.rkey[q] == q;
    const uint64_t m = __ballot(is_v);
    const int lane = threadIdx.x & 63;
    const int64_t w0 = (q - lane) >> 5;
    if (lane == 0 && w0 < (ninst + 31) / 32) bits[w0] = (uint32_t) m;
    if (lane == 32 && w0 + 1 < (ninst + 31) / 32) bits[w0 + 1] = (uint32_t) (m >> 32);
}

__device__ __forceinline__ int64_t rank_of(const uint32_t *bits, const int64_t *pre, int64_t k) {
    uint32_t wd = bits[k >> 5];
    uint32_t below = (k & 31) ? (wd & ((1u << (k & 31)) - 1)) : 0u;
    return pre[k >> 5] + __popc(below);
}

// Vertex id of each instance's representative.
__global__ void k_smc_vid(SparseMcWork w, int64_t ninst, const uint32_t *bits, const int64_t *wpre) {
    const int64_t q = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (q >= ninst) return;
    const int64_t r = w.rkey[q];
    w.vid[q] = (int32_t) rank_of(bits, wpre, r);
}

// Pass 4: faces in (slot, LUT slot) order (the kept triangles recorded by
// the faces pass), corners as vertex ids of their representatives.
__global__ void k_smc_emit(SparseMcIn in, SparseMcWork w, const int64_t *fpre, SparseMcOut o) {
    __shared__ int64_t sh[kB];
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    const uint32_t cw = i < in.n ? w.cst[i] : 0u;
    const int st = (int) (cw & 0xfffu), cs = (int) ((cw >> 16) & 0xffu);
    const uint32_t kept = st ? (cw >> 24) & 31u : 0u;
    int64_t f = fpre[blockIdx.x] + block_excl(__popc(kept), sh);
    if (!kept) return;
    const int64_t base = w.cell_inst[i];
    const uint64_t tw = cSTris[in.method][cs];
    for (uint32_t m = kept; m; m &= m - 1) {
        const int q = __ffs(m) - 1;
        const int64_t i0 = inst_of(st, (int) ((tw >> (4 + 12 * q)) & 15), base);
        const int64_t i1 = inst_of(st, (int) ((tw >> (8 + 12 * q)) & 15), base);
        const int64_t i2 = inst_of(st, (int) ((tw >> (12 + 12 * q)) & 15), base);
        if (f < o.cap_f) {
            o.f[3 * f + 0] = w.vid[i0];
            o.f[3 * f + 1] = w.vid[i1];
            o.f[3 * f + 2] = w.vid[i2];
        }
        ++f;
    }
}

// Output vertices: each representative that a kept face uses writes its
// position (and its source edge) at its rank.
__global__ void k_smc_vwrite(SparseMcWork w, int64_t ninst, const uint32_t *bits, const int64_t *wpre,
                             SparseMcOut o) {
    const int64_t q = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (q >= ninst || !((bits[q >> 5] >> (q & 31)) & 1u)) return;
    const int64_t v = rank_of(bits, wpre, q);
    if (v >= o.cap_v) return;
    o.v[3 * v + 0] = w.ipos[3 * q + 0];
    o.v[3 * v + 1] = w.ipos[3 * q + 1];
    o.v[3 * v + 2] = w.ipos[3 * q + 2];
    if (o.vsrc) o.vsrc[v] = w.isrc[q];
}

__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v) {
    for (int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t) gridDim.x * kB) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t) gridDim.x * kB) p[i] = v;
}

// single-block exclusive scan of int64 block sums; total in *tot
__global__ void k_scan_i64(int64_t *a, int64_t n, int64_t *tot) {
    __shared__ int64_t sh[1024];
    int64_t per = (n + 1023) / 1024;
    int64_t lo = threadIdx.x * per, hi = min(lo + per, n);
    int64_t sum = 0;
    for (int64_t i = lo; i < hi; ++i) sum += a[i];
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t y = threadIdx.x >= (unsigned) o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    int64_t run = sh[threadIdx.x] - sum;
    for (int64_t i = lo; i < hi; ++i) {
        int64_t v = a[i];
        a[i] = run;
        run += v;
    }
    if (threadIdx.x == 1023) *tot = sh[1023];
}

// popcount prefix over bitmap words (blocks of 1024 words, then block sums)
__global__ void k_bit_block(const uint32_t *bits, int64_t nw, int64_t *pre, int64_t *bsum) {
    __shared__ int64_t sh[1024];
    int64_t i = (int64_t) blockIdx.x * 1024 + threadIdx.x;
    int64_t v = i < nw ? __popc(bits[i]) : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t y = threadIdx.x >= (unsigned) o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    if (i < nw) pre[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == 1023) bsum[blockIdx.x] = sh[1023];
}

__global__ void k_bit_add(int64_t *pre, int64_t nw, const int64_t *bsum) {
    int64_t i = (int64_t) blockIdx.x * 1024 + threadIdx.x;
    if (i < nw) pre[i] += bsum[blockIdx.x];
}

// Sum of n int64 entries (one workgroup, coalesced) to *tot.
__global__ void k_sum_i64(const int64_t *a, int64_t n, int64_t *tot) {
    __shared__ int64_t sh[1024];
    int64_t v = 0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) v += a[i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned) o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *tot = sh[0];
}

// In-place exclusive scan of each 1024-entry chunk of a; chunk sums to bsum.
__global__ void k_scan_chunk(int64_t *a, int64_t n, int64_t *bsum) {
    __shared__ int64_t sh[1024];
    int64_t i = (int64_t) blockIdx.x * 1024 + threadIdx.x;
    int64_t v = i < n ? a[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t y = threadIdx.x >= (unsigned) o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    if (i < n) a[i] = sh[threadIdx.x] - v;
    if (threadIdx.x == 1023) bsum[blockIdx.x] = sh[1023];
}

// ---- population: sorted-key set operations (src/grid/sparse.cu:71-121) ----------

// Bitmap of the elements a set operation keeps, one bit per element, written
// a 64-bit wave ballot at a time (element i -> bit i).
__device__ __forceinline__ bool found_sorted(const int64_t *a, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == key;
}

// union: b (sorted) elements that start a run and are absent from a
// difference: a elements absent from b (sorted)
__global__ void k_setop_flags(const int64_t *a, int64_t na, const int64_t *b, int64_t nb, int difference,
                              unsigned long long *bits) {
    const int64_t n = difference ? na : nb;
    for (int64_t i0 = (int64_t) blockIdx.x * kB; i0 < n; i0 += (int64_t) gridDim.x * kB) {
        const int64_t i = i0 + threadIdx.x;
        bool keep = false;
        if (i < n) {
            if (difference) keep = !found_sorted(b, nb, a[i]);
            else keep = (i == 0 || b[i] != b[i - 1]) && !found_sorted(a, na, b[i]);
        }
        const unsigned long long m = __ballot(keep);
        if ((threadIdx.x & 63) == 0 && i < n) bits[i >> 6] = m;
    }
}

// Kept elements of `src` (bitmap + per-32-bit-word exclusive prefix) to
// consecutive positions of `out`; `slot` (optional) gets the source index.
__global__ void k_setop_compact(const int64_t *src, int64_t n, const uint32_t *bits, const int64_t *pre,
                                int64_t *out, int64_t *slot) {
    for (int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t) gridDim.x * kB) {
        const uint32_t w = bits[i >> 5];
        if (!((w >> (i & 31)) & 1)) continue;
        const int64_t k = pre[i >> 5] + __popc(w & ((1u << (i & 31)) - 1));
        out[k] = src[i];
        if (slot) slot[k] = i;
    }
}

// Merge of two sorted, disjoint arrays: output k takes a[i] or b[k - i], i the
// number of a elements among the first k outputs (merge-path search on the
// diagonal); slot[k] = i when it comes from a, -1 from b.
__global__ void k_merge_disjoint(const int64_t *a, int64_t na, const int64_t *b, const int64_t *nb_p, int64_t *out,
                                 int64_t *slot, int64_t *n_out) {
    const int64_t nb = *nb_p, n = na + nb;
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = n;
    for (int64_t k = (int64_t) blockIdx.x * kB + threadIdx.x; k < n; k += (int64_t) gridDim.x * kB) {
        int64_t lo = max((int64_t) 0, k - nb), hi = min(k, na);
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (a[mid] < b[k - 1 - mid]) lo = mid + 1;
            else hi = mid;
        }
        const int64_t i = lo, j = k - lo;
        const bool from_a = j >= nb || (i < na && a[i] < b[j]);
        out[k] = from_a ? a[i] : b[j];
        if (slot) slot[k] = from_a ? i : -1;
    }
}

// Per-cell values after a set operation: the old row of the cell's source
// slot, or the default value for a new cell.
__global__ void k_gather_rows(const float *old_values, const int64_t *slot, int64_t n, float dflt, float *out) {
    for (int64_t t = (int64_t) blockIdx.x * kB + threadIdx.x; t < 8 * n; t += (int64_t) gridDim.x * kB) {
        const int64_t s = slot[t >> 3];
        out[t] = s >= 0 ? old_values[8 * s + (t & 7)] : dflt;
    }
}

unsigned nb(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + kB - 1) / kB); }
unsigned gs(int64_t n) { return (unsigned) std::max<int64_t>(1, std::min<int64_t>(8192, (n + kB - 1) / kB)); }

}   // namespace

#define ISO_TRY(x)                       \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

// Scratch of a set operation on na (or nb) keys: the bitmap, its word prefix,
// the block sums, the compacted b and two counters.
int64_t sparse_setop_scratch_bytes(int64_t na, int64_t nb) {
    const int64_t n = std::max(na, nb), nw64 = n / 64 + 2, nw32 = 2 * nw64;
    return 8 * (nw64 + nw32 + (nw32 + 1023) / 1024 + 8 + nb + 8) + 256;
}

hipError_t sparse_setop(const int64_t *a, int64_t na, const int64_t *b, int64_t nb, bool difference, void *scratch,
                        int64_t *out, int64_t *slot, int64_t *n_out, hipStream_t st) {
    const int64_t n = difference ? na : nb, nw64 = std::max(na, nb) / 64 + 2, nw32 = 2 * nw64;
    char *p = (char *) scratch;
    unsigned long long *bits = (unsigned long long *) p;
    int64_t *pre = (int64_t *) (p + 8 * nw64);
    int64_t *bsum = pre + nw32;
    const int64_t nblk = (nw32 + 1023) / 1024;
    int64_t *tot = bsum + nblk + 8;
    int64_t *bc = tot + 8;   // compacted b (union)
    hipLaunchKernelGGL(k_fill_u64, dim3(gs(nw64)), dim3(kB), 0, st, bits, nw64, 0ull);
    if (n > 0)
        hipLaunchKernelGGL(k_setop_flags, dim3(gs(n)), dim3(kB), 0, st, a, na, b, nb, difference ? 1 : 0, bits);
    hipLaunchKernelGGL(k_bit_block, dim3((unsigned) nblk), dim3(1024), 0, st, (const uint32_t *) bits, nw32, pre, bsum);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, bsum, nblk, tot);
    hipLaunchKernelGGL(k_bit_add, dim3((unsigned) nblk), dim3(1024), 0, st, pre, nw32, bsum);
    if (difference) {
        if (na > 0)
            hipLaunchKernelGGL(k_setop_compact, dim3(gs(na)), dim3(kB), 0, st, a, na, (const uint32_t *) bits, pre,
                               out, slot);
        ISO_TRY(hipMemcpyAsync(n_out, tot, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    } else {
        if (nb > 0)
            hipLaunchKernelGGL(k_setop_compact, dim3(gs(nb)), dim3(kB), 0, st, b, nb, (const uint32_t *) bits, pre, bc,
                               (int64_t *) nullptr);
        hipLaunchKernelGGL(k_merge_disjoint, dim3(gs(na + nb)), dim3(kB), 0, st, a, na, bc, tot, out, slot, n_out);
    }
    return hipGetLastError();
}

hipError_t sparse_gather_rows(const float *old_values, const int64_t *slot, int64_t n, float dflt, float *out,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_rows, dim3(gs(8 * n)), dim3(kB), 0, st, old_values, slot, n, dflt, out);
    return hipGetLastError();
}

hipError_t sparse_corner_points(const int64_t *ids, int64_t n, const SparseGeom &g, float *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_corner_points, dim3(nb(n)), dim3(kB), 0, st, ids, n, g, out);
    return hipGetLastError();
}

hipError_t sparse_crossing_flags(const float *values, int64_t n, float level, uint8_t *flags, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_crossing_flags, dim3(nb(n)), dim3(kB), 0, st, values, n, level, flags);
    return hipGetLastError();
}

hipError_t sparse_mc_luts() {
    // once per device (each GPU holds its own copy of the constant table);
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
    for (int i = 0; i < 256; ++i) {
        h[0][i] = kNagaeTris[i];
        h[1][i] = kLorensenTris[i];
    }
    e = hipMemcpyToSymbol(HIP_SYMBOL(cSTris), h, sizeof(h));
    if (e == hipSuccess) done[dev].store(true, std::memory_order_release);
    return e;
}

size_t sparse_mc_work_bytes(int64_t n, int64_t ninst) {
    const int64_t nw = (ninst + 31) / 32 + 1, nblk = nb(n);
    size_t b = 0;
    auto add = [&](size_t x) { b += (x + 255) & ~size_t(255); };
    add(4 * n);                              // cst
    add(8 * n);                              // cell_inst
    add(4 * ninst + 8);                      // inst_cell
    add(12 * ninst + 12);                    // ipos
    add(8 * ninst + 8);                      // isrc
    add(8 * ninst + 8);                      // rkey
    add(ninst + 8);                          // used
    add(4 * ninst + 8);                      // vid
    add(8 * (nblk + 8));                     // block sums (instances)
    add(8 * (nblk + 8));                     // block sums (faces)
    add(8 * ((nblk + 1023) / 1024 + 8));     // chunk sums (instances)
    add(8 * ((nblk + 1023) / 1024 + 8));     // chunk sums (faces)
    add(4 * nw);                             // bits
    add(8 * nw);                             // word prefix
    add(8 * ((nw + 1023) / 1024 + 8));
    add(64);
    return b;
}

// Exclusive scan of m int64 entries in place (chunks of 1024, then the
// chunk sums in one workgroup); the total to *tot.
static void scan_i64(int64_t *a, int64_t m, int64_t *chunk, int64_t *tot, hipStream_t st) {
    const int64_t nc = (m + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_chunk, dim3((unsigned) nc), dim3(1024), 0, st, a, m, chunk);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, chunk, nc, tot);
    hipLaunchKernelGGL(k_bit_add, dim3((unsigned) nc), dim3(1024), 0, st, a, m, (const int64_t *) chunk);
}

hipError_t sparse_mc_count(const SparseMcIn &in, int64_t *scratch, int64_t &ninst, hipStream_t st) {
    ninst = 0;
    if (in.n == 0) return hipSuccess;
    int64_t nblk = nb(in.n);
    hipLaunchKernelGGL(k_smc_count, dim3((unsigned) nblk), dim3(kB), 0, st, in, scratch, (uint32_t *) nullptr);
    hipLaunchKernelGGL(k_sum_i64, dim3(1), dim3(1024), 0, st, (const int64_t *) scratch, nblk, scratch + nblk);
    ISO_TRY(hipGetLastError());
    ISO_TRY(hipMemcpyAsync(&ninst, scratch + nblk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    return hipSuccess;
}

namespace {
// The column table of the calling thread (device memory, grown on demand,
// freed when the thread ends).
struct ColTable {
    int32_t *p = nullptr;
    size_t entries = 0;
    int dev = -1;
    ~ColTable() {
        if (p) (void) hipFree(p);
    }
};
thread_local ColTable tl_cols;
}   // namespace

hipError_t sparse_mc(const SparseMcIn &in, int64_t ninst, void *work, size_t work_bytes, SparseMcOut &o,
                     int64_t &nv, int64_t &nf, hipStream_t st) {
    nv = nf = 0;
    if (in.n == 0 || ninst == 0) return hipSuccess;
    if (in.n >= (int64_t) INT32_MAX || ninst >= (int64_t) INT32_MAX) return hipErrorInvalidValue;
    ISO_TRY(sparse_mc_luts());
    const int64_t nw = (ninst + 31) / 32 + 1;
    const int64_t nblk = nb(in.n);
    char *p = (char *) work;
    auto take = [&](size_t n) {
        char *r = p;
        p += (n + 255) & ~size_t(255);
        return r;
    };
    SparseMcWork w;
    w.cst = (uint32_t *) take(4 * in.n);
    w.cell_inst = (int64_t *) take(8 * in.n);
    w.inst_cell = (int32_t *) take(4 * ninst + 8);
    w.ipos = (float *) take(12 * ninst + 12);
    w.isrc = (int64_t *) take(8 * ninst + 8);
    w.rkey = (int64_t *) take(8 * ninst + 8);
    w.used = (uint8_t *) take(ninst + 8);
    w.vid = (int32_t *) take(4 * ninst + 8);
    int64_t *ibs = (int64_t *) take(8 * (nblk + 8));
    int64_t *fbs = (int64_t *) take(8 * (nblk + 8));
    int64_t *ich = (int64_t *) take(8 * ((nblk + 1023) / 1024 + 8));
    int64_t *fch = (int64_t *) take(8 * ((nblk + 1023) / 1024 + 8));
    uint32_t *bits = (uint32_t *) take(4 * nw);
    int64_t *wpre = (int64_t *) take(8 * nw);
    const int64_t nwb = (nw + 1023) / 1024;
    int64_t *wbs = (int64_t *) take(8 * (nwb + 8));
    int64_t *misc = (int64_t *) take(64);
    w.err = (int *) (misc + 4);
    if ((size_t) (p - (char *) work) > work_bytes) return hipErrorInvalidValue;
    // column table: (X - 1) (Y - 1) columns
    const int64_t ncol = (in.g.X - 1) * (in.g.Y - 1);
    int dev = -1;
    ISO_TRY(hipGetDevice(&dev));
    const int64_t ncf = ncol + 1, nchunk = (ncf + 1023) / 1024;
    const size_t entries = (size_t) (ncf + nchunk);
    if (tl_cols.entries < entries || tl_cols.dev != dev) {
        if (tl_cols.p) ISO_TRY(hipFree(tl_cols.p));
        tl_cols.p = nullptr;
        tl_cols.entries = 0;
        ISO_TRY(hipMalloc((void **) &tl_cols.p, sizeof(int32_t) * entries));
        tl_cols.entries = entries;
        tl_cols.dev = dev;
    }
    int32_t *cf = tl_cols.p, *cmin = tl_cols.p + ncf;
    SLook L{in.ids, cf, in.g.X - 1, in.g.Y - 1, in.g.Z - 1};

    hipLaunchKernelGGL(k_fill_u32, dim3(gs(ncf)), dim3(kB), 0, st, (uint32_t *) cf, ncf, (uint32_t) in.n);
    hipLaunchKernelGGL(k_fill_u32, dim3(gs((ninst + 8) / 4 + 1)), dim3(kB), 0, st, (uint32_t *) w.used,
                       (ninst + 8) / 4 + 1, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kB), 0, st, (unsigned long long *) misc, 8, 0ull);
    hipLaunchKernelGGL(k_smc_cols, dim3(nb(in.n)), dim3(kB), 0, st, in.ids, in.n, in.g.Z - 1, cf);
    hipLaunchKernelGGL(k_cols_chunk, dim3((unsigned) nchunk), dim3(1024), 0, st, cf, ncf, cmin);
    hipLaunchKernelGGL(k_cols_top, dim3(1), dim3(1024), 0, st, cmin, nchunk);
    hipLaunchKernelGGL(k_cols_apply, dim3((unsigned) nchunk), dim3(1024), 0, st, cf, ncf, (const int32_t *) cmin);
    hipLaunchKernelGGL(k_smc_count, dim3((unsigned) nblk), dim3(kB), 0, st, in, ibs, w.cst);
    scan_i64(ibs, nblk, ich, misc + 2, st);
    hipLaunchKernelGGL(k_smc_cells, dim3((unsigned) nblk), dim3(kB), 0, st, in.n, (const int64_t *) ibs, w);
    hipLaunchKernelGGL(k_smc_inst, dim3(nb(ninst)), dim3(kB), 0, st, in, L, w, ninst);
    hipLaunchKernelGGL(k_smc_faces, dim3((unsigned) nblk), dim3(kB), 0, st, in, w, fbs);
    scan_i64(fbs, nblk, fch, misc + 0, st);
    hipLaunchKernelGGL(k_smc_vbits, dim3(nb(ninst)), dim3(kB), 0, st, w, ninst, bits);
    hipLaunchKernelGGL(k_bit_block, dim3((unsigned) nwb), dim3(1024), 0, st, bits, nw - 1, wpre, wbs);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, wbs, nwb, misc + 1);
    hipLaunchKernelGGL(k_bit_add, dim3((unsigned) nwb), dim3(1024), 0, st, wpre, nw - 1, wbs);
    hipLaunchKernelGGL(k_smc_vid, dim3(nb(ninst)), dim3(kB), 0, st, w, ninst, bits, wpre);
    ISO_TRY(hipGetLastError());
    int64_t h[5];
    ISO_TRY(hipMemcpyAsync(h, misc, sizeof(h), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    if (h[4] != 0) return hipErrorUnknown;
    nf = h[0];
    nv = h[1];
    if (nv > o.cap_v || nf > o.cap_f) return hipSuccess;   // caller re-allocates and retries
    hipLaunchKernelGGL(k_smc_emit, dim3((unsigned) nblk), dim3(kB), 0, st, in, w, (const int64_t *) fbs, o);
    hipLaunchKernelGGL(k_smc_vwrite, dim3(nb(ninst)), dim3(kB), 0, st, w, ninst, (const uint32_t *) bits,
                       (const int64_t *) wpre, o);
    ISO_TRY(hipGetLastError());
    ISO_TRY(hipStreamSynchronize(st));
    return hipSuccess;
}

}   // namespace iso
