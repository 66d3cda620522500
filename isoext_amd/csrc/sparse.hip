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
// Sparse MC therefore keys vertices by position (hash on the float triple,
// equality on floats), with the representative of a position = the smallest
// (slot * 12 + edge) that produces it; vertex ids follow representative order
// and faces follow (slot, LUT slot) order, so output is deterministic.
// Degenerate triangles (two equal corners) are dropped as in
// src/mc/nagae.cu:71; a position is emitted only if a kept triangle uses it.
// The per-vertex representative is exported for the backward pass.

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

// Pass 1: crossing-edge instances per cell -> block sums.
__global__ void k_smc_count(SparseMcIn in, int64_t *bsum) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    int64_t k = 0;
    if (i < in.n) {
        int cs = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) cs |= inside(in.values[8 * i + j], in.level) << j;
        if (cs != 0 && cs != 255) k = __popc(scross(cs));
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

// Pass 2: write each instance's position; instance ids are (slot, edge) ordered.
__global__ void k_smc_inst(SparseMcIn in, const int64_t *bpre, SparseMcWork w) {
    __shared__ int64_t sh[kB];
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    SCell s;
    bool act = i < in.n && load_cell(in, i, s);
    int st = act ? scross(s.cs) : 0;
    int64_t off = bpre[blockIdx.x] + block_excl(__popc(st), sh);
    if (!act) return;
    w.cell_inst[i] = off;
    for (int e = 0; e < 12; ++e) {
        if (!((st >> e) & 1)) continue;
        float x, y, z;
        spos(s, in.level, e, x, y, z);
        w.ipos[3 * off + 0] = x;
        w.ipos[3 * off + 1] = y;
        w.ipos[3 * off + 2] = z;
        w.isrc[off] = i * 12 + e;
        ++off;
    }
}

// Pass 3: group instances by position (hash slot holds the first writer; the
// representative is the smallest instance id with that position).
__global__ void k_smc_hash(SparseMcWork w, int64_t ninst) {
    int64_t q = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (q >= ninst) return;
    float x = w.ipos[3 * q], y = w.ipos[3 * q + 1], z = w.ipos[3 * q + 2];
    uint64_t h = f3_hash(x, y, z) & (uint64_t) w.hmask;
    for (int64_t probe = 0; probe <= w.hmask; ++probe) {
        int64_t sl = (int64_t) ((h + probe) & (uint64_t) w.hmask);
        unsigned long long cur = w.hkey[sl];
        if (cur == ~0ull) {
            unsigned long long prev = atomicCAS(&w.hkey[sl], ~0ull, (unsigned long long) q);
            if (prev == ~0ull) cur = (unsigned long long) q;
            else cur = prev;
        }
        if (w.ipos[3 * cur] == x && w.ipos[3 * cur + 1] == y && w.ipos[3 * cur + 2] == z) {
            w.islot[q] = sl;
            atomicMin(&w.hmin[sl], (unsigned long long) q);
            return;
        }
    }
    atomicOr(w.err, 1);
}

__device__ __forceinline__ int64_t inst_of(int st, int e, int64_t base) {
    return base + __popc(st & ((1 << e) - 1));
}

// Pass 4: kept faces per cell (a triangle with two corners at one position is
// dropped); positions used by kept faces are marked.
__global__ void k_smc_faces(SparseMcIn in, SparseMcWork w, int64_t *bsum) {
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    int64_t nf = 0;
    if (i < in.n) {
        int cs = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) cs |= inside(in.values[8 * i + j], in.level) << j;
        if (cs != 0 && cs != 255) {
            const int st = scross(cs);
            const int64_t base = w.cell_inst[i];
            const uint64_t tw = cSTris[in.method][cs];
            const int nt = (int) (tw & 15);
            for (int q = 0; q < nt; ++q) {
                int64_t s0 = w.islot[inst_of(st, (int) ((tw >> (4 + 12 * q)) & 15), base)];
                int64_t s1 = w.islot[inst_of(st, (int) ((tw >> (8 + 12 * q)) & 15), base)];
                int64_t s2 = w.islot[inst_of(st, (int) ((tw >> (12 + 12 * q)) & 15), base)];
                if (s0 == s1 || s0 == s2 || s1 == s2) continue;
                ++nf;
                w.hused[s0] = 1u;
                w.hused[s1] = 1u;
                w.hused[s2] = 1u;
            }
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

// Used positions mark their representative instance in a bitmap.
__global__ void k_smc_mark(SparseMcWork w, int64_t nslots, uint32_t *bits) {
    int64_t sl = (int64_t) blockIdx.x * kB + threadIdx.x;
    if (sl >= nslots || w.hkey[sl] == ~0ull || !w.hused[sl]) return;
    unsigned long long k = w.hmin[sl];
    atomicOr(&bits[k >> 5], 1u << (k & 31));
}

__device__ __forceinline__ int64_t rank_of(const uint32_t *bits, const int64_t *pre, int64_t k) {
    uint32_t wd = bits[k >> 5];
    uint32_t below = (k & 31) ? (wd & ((1u << (k & 31)) - 1)) : 0u;
    return pre[k >> 5] + __popc(below);
}

// Pass 5: faces in (slot, LUT slot) order, representatives write vertices.
__global__ void k_smc_emit(SparseMcIn in, SparseMcWork w, const int64_t *fpre, const uint32_t *bits,
                           const int64_t *wpre, SparseMcOut o) {
    __shared__ int64_t sh[kB];
    int64_t i = (int64_t) blockIdx.x * kB + threadIdx.x;
    int cs = 0;
    if (i < in.n) {
#pragma unroll
        for (int j = 0; j < 8; ++j) cs |= inside(in.values[8 * i + j], in.level) << j;
    }
    const bool act = i < in.n && cs != 0 && cs != 255;
    const int st = act ? scross(cs) : 0;
    const int64_t base = act ? w.cell_inst[i] : 0;
    const uint64_t tw = act ? cSTris[in.method][cs] : 0;
    const int nt = (int) (tw & 15);
    int64_t nf = 0;
    for (int q = 0; q < nt; ++q) {
        int64_t s0 = w.islot[inst_of(st, (int) ((tw >> (4 + 12 * q)) & 15), base)];
        int64_t s1 = w.islot[inst_of(st, (int) ((tw >> (8 + 12 * q)) & 15), base)];
        int64_t s2 = w.islot[inst_of(st, (int) ((tw >> (12 + 12 * q)) & 15), base)];
        if (!(s0 == s1 || s0 == s2 || s1 == s2)) ++nf;
    }
    int64_t f = fpre[blockIdx.x] + block_excl(nf, sh);
    if (!act) return;
    for (int q = 0; q < nt; ++q) {
        int64_t s0 = w.islot[inst_of(st, (int) ((tw >> (4 + 12 * q)) & 15), base)];
        int64_t s1 = w.islot[inst_of(st, (int) ((tw >> (8 + 12 * q)) & 15), base)];
        int64_t s2 = w.islot[inst_of(st, (int) ((tw >> (12 + 12 * q)) & 15), base)];
        if (s0 == s1 || s0 == s2 || s1 == s2) continue;
        if (f < o.cap_f) {
            o.f[3 * f + 0] = (int32_t) rank_of(bits, wpre, (int64_t) w.hmin[s0]);
            o.f[3 * f + 1] = (int32_t) rank_of(bits, wpre, (int64_t) w.hmin[s1]);
            o.f[3 * f + 2] = (int32_t) rank_of(bits, wpre, (int64_t) w.hmin[s2]);
        }
        ++f;
    }
    int64_t q = base;
    for (int e = 0; e < 12; ++e) {
        if (!((st >> e) & 1)) continue;
        int64_t sl = w.islot[q];
        if ((int64_t) w.hmin[sl] == q && w.hused[sl]) {
            int64_t v = rank_of(bits, wpre, q);
            if (v < o.cap_v) {
                o.v[3 * v + 0] = w.ipos[3 * q + 0];
                o.v[3 * v + 1] = w.ipos[3 * q + 1];
                o.v[3 * v + 2] = w.ipos[3 * q + 2];
                if (o.vsrc) o.vsrc[v] = w.isrc[q];
            }
        }
        ++q;
    }
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
    static bool done = false;
    if (done) return hipSuccess;
    uint64_t h[2][256];
    for (int i = 0; i < 256; ++i) {
        h[0][i] = kNagaeTris[i];
        h[1][i] = kLorensenTris[i];
    }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cSTris), h, sizeof(h));
    if (e == hipSuccess) done = true;
    return e;
}

size_t sparse_mc_work_bytes(int64_t n, int64_t ninst) {
    int64_t hs = 1;
    while (hs < 2 * ninst + 64) hs <<= 1;
    int64_t nw = (ninst + 31) / 32 + 1;
    size_t b = 0;
    auto add = [&](size_t x) { b += (x + 255) & ~size_t(255); };
    add(8 * n);                      // cell_inst
    add(12 * ninst + 12);            // ipos
    add(8 * ninst + 8);              // isrc
    add(8 * ninst + 8);              // islot
    add(8 * hs);                     // hkey
    add(8 * hs);                     // hmin
    add(4 * hs);                     // hused
    add(8 * (nb(n) + 8));            // block sums (instances)
    add(8 * (nb(n) + 8));            // block sums (faces)
    add(4 * nw);                     // bits
    add(8 * nw);                     // word prefix
    add(8 * ((nw + 1023) / 1024 + 8));
    add(64);
    return b;
}

hipError_t sparse_mc_count(const SparseMcIn &in, int64_t *scratch, int64_t &ninst, hipStream_t st) {
    ninst = 0;
    if (in.n == 0) return hipSuccess;
    int64_t nblk = nb(in.n);
    hipLaunchKernelGGL(k_smc_count, dim3((unsigned) nblk), dim3(kB), 0, st, in, scratch);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, scratch, nblk, scratch + nblk);
    ISO_TRY(hipGetLastError());
    ISO_TRY(hipMemcpyAsync(&ninst, scratch + nblk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    return hipSuccess;
}

hipError_t sparse_mc(const SparseMcIn &in, int64_t ninst, void *work, size_t work_bytes, SparseMcOut &o,
                     int64_t &nv, int64_t &nf, hipStream_t st) {
    nv = nf = 0;
    if (in.n == 0 || ninst == 0) return hipSuccess;
    ISO_TRY(sparse_mc_luts());
    int64_t hs = 1;
    while (hs < 2 * ninst + 64) hs <<= 1;
    const int64_t nw = (ninst + 31) / 32 + 1;
    const int64_t nblk = nb(in.n);
    char *p = (char *) work;
    auto take = [&](size_t n) {
        char *r = p;
        p += (n + 255) & ~size_t(255);
        return r;
    };
    SparseMcWork w;
    w.hmask = hs - 1;
    w.cell_inst = (int64_t *) take(8 * in.n);
    w.ipos = (float *) take(12 * ninst + 12);
    w.isrc = (int64_t *) take(8 * ninst + 8);
    w.islot = (int64_t *) take(8 * ninst + 8);
    w.hkey = (unsigned long long *) take(8 * hs);
    w.hmin = (unsigned long long *) take(8 * hs);
    w.hused = (uint32_t *) take(4 * hs);
    int64_t *ibs = (int64_t *) take(8 * (nblk + 8));
    int64_t *fbs = (int64_t *) take(8 * (nblk + 8));
    uint32_t *bits = (uint32_t *) take(4 * nw);
    int64_t *wpre = (int64_t *) take(8 * nw);
    const int64_t nwb = (nw + 1023) / 1024;
    int64_t *wbs = (int64_t *) take(8 * (nwb + 8));
    int64_t *misc = (int64_t *) take(64);
    w.err = (int *) (misc + 4);
    if ((size_t) (p - (char *) work) > work_bytes) return hipErrorInvalidValue;

    hipLaunchKernelGGL(k_fill_u64, dim3(gs(hs)), dim3(kB), 0, st, w.hkey, hs, ~0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(gs(hs)), dim3(kB), 0, st, w.hmin, hs, ~0ull);
    hipLaunchKernelGGL(k_fill_u32, dim3(gs(hs)), dim3(kB), 0, st, w.hused, hs, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(gs(nw)), dim3(kB), 0, st, bits, nw, 0u);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kB), 0, st, (unsigned long long *) misc, 8, 0ull);
    hipLaunchKernelGGL(k_smc_count, dim3((unsigned) nblk), dim3(kB), 0, st, in, ibs);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, ibs, nblk, misc + 2);
    hipLaunchKernelGGL(k_smc_inst, dim3((unsigned) nblk), dim3(kB), 0, st, in, ibs, w);
    hipLaunchKernelGGL(k_smc_hash, dim3(nb(ninst)), dim3(kB), 0, st, w, ninst);
    hipLaunchKernelGGL(k_smc_faces, dim3((unsigned) nblk), dim3(kB), 0, st, in, w, fbs);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, fbs, nblk, misc + 0);
    hipLaunchKernelGGL(k_smc_mark, dim3(nb(hs)), dim3(kB), 0, st, w, hs, bits);
    hipLaunchKernelGGL(k_bit_block, dim3((unsigned) nwb), dim3(1024), 0, st, bits, nw, wpre, wbs);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, wbs, nwb, misc + 1);
    hipLaunchKernelGGL(k_bit_add, dim3((unsigned) nwb), dim3(1024), 0, st, wpre, nw, wbs);
    ISO_TRY(hipGetLastError());
    int64_t h[5];
    ISO_TRY(hipMemcpyAsync(h, misc, sizeof(h), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    if (h[4] != 0) return hipErrorUnknown;
    nf = h[0];
    nv = h[1];
    if (nv > o.cap_v || nf > o.cap_f) return hipSuccess;   // caller re-allocates and retries
    hipLaunchKernelGGL(k_smc_emit, dim3((unsigned) nblk), dim3(kB), 0, st, in, w, fbs, bits, wpre, o);
    ISO_TRY(hipGetLastError());
    ISO_TRY(hipStreamSynchronize(st));
    return hipSuccess;
}

}   // namespace iso
