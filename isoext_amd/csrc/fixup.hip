// Exact weld fix-up for marching cubes (dense and sparse share it).
//
// The reference welds triangle corners by bitwise position equality and
// drops triangles with two equal corners (src/utils.cu:32-59,
// src/mc/nagae.cu:71).  The fused kernels number vertices by grid edge, which
// is the same thing except where a vertex lands (within a few ulps) on a grid
// corner: there, vertices of different edges can coincide, triangles can
// degenerate and a vertex can lose every triangle that used it.  The kernels
// record those rare vertices and the faces of the cells holding them; this
// pass resolves them exactly:
//   1. group candidate vertices by position (hash on the float triple, equality
//      checked on the floats) -> representative = smallest provisional id;
//   2. a vertex is used if a kept face of a special cell references it, or if
//      some cell around its edge is not special (such a cell keeps all of its
//      triangles);
//   3. remove non-representatives and unused vertices, drop degenerate faces,
//      remap faces, compact (stable, so order is preserved).

#include <vector>

#include "common.h"
#include "fixup.h"
#include "mc_dense.h"

namespace iso {

namespace {

constexpr int64_t kEmpty = -1;

struct VInfo {
    int count_s;      // number of special-cell entries for this vertex
    int referenced;   // referenced by a kept face of a special cell
    int is_cand;
    int group;        // position-group slot (candidates)
    int64_t edge;
};

struct PGroup {
    int64_t rep;   // smallest provisional vid with this position
    int used;
    int pad;
};

__device__ int64_t h1_find_or_insert(int64_t *keys, int64_t mask, int64_t vid, bool insert) {
    if (vid < 0) return -1;
    uint64_t h = mix64((uint64_t) vid) & (uint64_t) mask;
    for (int64_t probe = 0; probe <= mask; ++probe) {
        int64_t slot = (int64_t) ((h + probe) & (uint64_t) mask);
        int64_t k = keys[slot];
        if (k == vid) return slot;
        if (k == kEmpty) {
            if (!insert) return -1;
            unsigned long long prev = atomicCAS((unsigned long long *) &keys[slot],
                                                (unsigned long long) kEmpty, (unsigned long long) vid);
            if ((int64_t) prev == kEmpty || (int64_t) prev == vid) return slot;
        }
    }
    return -1;
}

// Number of cells around edge (owner point, axis) that exist in the grid.
__device__ int edge_cell_count(int64_t edge, const FixupGrid &g) {
    int axis = (int) (edge % 3);
    int64_t pidx = edge / 3;
    int64_t z = pidx % g.Z, y = (pidx / g.Z) % g.Y, x = pidx / ((int64_t) g.Y * g.Z);
    int64_t c[3] = {x, y, z}, n[3] = {g.X - 1, g.Y - 1, g.Z - 1};
    int cnt = 1;
    for (int a = 0; a < 3; ++a) {
        if (a == axis) continue;
        int here = (c[a] < n[a]) ? 1 : 0;       // cell at c[a]
        int below = (c[a] >= 1) ? 1 : 0;        // cell at c[a]-1
        cnt *= here + below;
    }
    return cnt;
}


__device__ void insert_scell(const SpecialEdge &e, int64_t *keys, VInfo *info, int64_t mask) {
    int64_t s = h1_find_or_insert(keys, mask, e.vid, true);
    if (s < 0) return;
    atomicAdd(&info[s].count_s, 1);
    info[s].edge = e.edge;
}

__device__ void insert_cand(const SpecialVertex *cv, int64_t i, int64_t *keys, VInfo *info, int64_t mask,
                            int32_t *pslot, PGroup *groups, int64_t pmask) {
    const SpecialVertex v = cv[i];
    int64_t s = h1_find_or_insert(keys, mask, v.vid, true);
    // position group: slots hold the index of the first candidate seen
    uint64_t h = f3_hash(v.x, v.y, v.z) & (uint64_t) pmask;
    int64_t g = -1;
    for (int64_t probe = 0; probe <= pmask; ++probe) {
        int64_t slot = (int64_t) ((h + probe) & (uint64_t) pmask);
        int32_t cur = pslot[slot];
        if (cur < 0) {
            int32_t prev = atomicCAS(&pslot[slot], -1, (int32_t) i);
            if (prev < 0) {
                g = slot;
                break;
            }
            cur = prev;
        }
        const SpecialVertex o = cv[cur];
        if (o.x == v.x && o.y == v.y && o.z == v.z) {
            g = slot;
            break;
        }
    }
    if (g >= 0) atomicMin((long long *) &groups[g].rep, (long long) v.vid);
    if (s >= 0) {
        info[s].is_cand = 1;
        info[s].group = (int) g;
        info[s].edge = v.edge;
    }
}

// Crossing edges of special cells, then candidate vertices (one index space).
__global__ void k_fix_insert(const SpecialEdge *se, int64_t nscell, const SpecialVertex *cv, int64_t ncand,
                             int64_t *keys, VInfo *info, int64_t mask, int32_t *pslot, PGroup *groups,
                             int64_t pmask) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nscell + ncand;
         i += (int64_t) gridDim.x * blockDim.x) {
        if (i < nscell) insert_scell(se[i], keys, info, mask);
        else insert_cand(cv, i - nscell, keys, info, mask, pslot, groups, pmask);
    }
}

// Kept faces of special cells: mark the vertices they reference, then point
// their corners at the position group's representative (only these faces
// can reference a non-representative).
__global__ void k_kept_refs(const int64_t *kept, int64_t n, int32_t *faces, int64_t nf, int64_t *keys, VInfo *info,
                            int64_t mask, const PGroup *groups) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        int64_t f = kept[i];
        if (f < 0 || f >= nf) continue;
        for (int j = 0; j < 3; ++j) {
            const int64_t v = faces[3 * f + j];
            int64_t s = h1_find_or_insert(keys, mask, v, false);
            if (s < 0) continue;
            info[s].referenced = 1;
            if (info[s].is_cand && info[s].group >= 0) {
                const int64_t rep = groups[info[s].group].rep;
                if (rep != v) faces[3 * f + j] = (int32_t) rep;
            }
        }
    }
}

__global__ void k_used(const int64_t *keys, VInfo *info, int64_t nslots, PGroup *groups, FixupGrid g) {
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t) gridDim.x * blockDim.x) {
        if (keys[s] == kEmpty) continue;
        VInfo &v = info[s];
        bool used = v.referenced || v.count_s < edge_cell_count(v.edge, g);
        if (v.is_cand && v.group >= 0) {
            if (used) atomicOr(&groups[v.group].used, 1);
        } else {
            v.is_cand = 0;
            info[s].referenced = used ? 1 : 0;   // reuse: final "used" flag
        }
    }
}

// Clears the keep bits of removed vertices (counted in *nremoved) and of
// dropped faces.
__global__ void k_removed(const int64_t *keys, const VInfo *info, int64_t nslots, const PGroup *groups,
                          uint32_t *vkeep, int64_t nv, unsigned long long *nremoved, const int64_t *hole,
                          int64_t nhole, uint32_t *fkeep, int64_t nf) {
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < max(nslots, nhole);
         s += (int64_t) gridDim.x * blockDim.x) {
        if (s < nhole) {
            const int64_t f = hole[s];
            if (f >= 0 && f < nf) atomicAnd(&fkeep[f >> 5], ~(1u << (f & 31)));
        }
        if (s >= nslots) continue;
        int64_t vid = keys[s];
        if (vid < 0 || vid >= nv) continue;
        const VInfo &v = info[s];
        bool removed;
        if (v.is_cand && v.group >= 0)
            removed = (vid != groups[v.group].rep) || !groups[v.group].used;
        else
            removed = !v.referenced;
        if (removed) {
            atomicAnd(&vkeep[vid >> 5], ~(1u << (vid & 31)));
            atomicAdd(nremoved, 1ull);
        }
    }
}

__device__ __forceinline__ int64_t bit_rank(const uint32_t *bits, const int64_t *prefix, int64_t i) {
    uint32_t w = bits[i >> 5];
    uint32_t below = (i & 31) ? (w & ((1u << (i & 31)) - 1)) : 0u;
    return prefix[i >> 5] + __popc(below);
}




// Resets the fix-up state in one launch: the vertex hash, the position
// groups, the keep bitmaps (bits set for [0, n), clear beyond) and the
// counters.
__global__ void k_fix_init(int64_t *keys, VInfo *info, int64_t nslots, int32_t *pslot, PGroup *groups,
                           int64_t npslots, uint32_t *vkeep, int64_t vw, int64_t nv, uint32_t *fkeep, int64_t fw,
                           int64_t nf, int64_t *misc) {
    const int64_t n = max(max(nslots, npslots), max(vw, fw));
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        if (i < nslots) {
            keys[i] = kEmpty;
            info[i].count_s = 0;
            info[i].referenced = 0;
            info[i].is_cand = 0;
            info[i].group = -1;
            info[i].edge = 0;
        }
        if (i < npslots) {
            pslot[i] = -1;
            groups[i].rep = INT64_MAX;
            groups[i].used = 0;
            groups[i].pad = 0;
        }
        if (i < vw) vkeep[i] = (i + 1) * 32 <= nv ? ~0u : (i * 32 < nv ? ((1u << (nv & 31)) - 1) : 0u);
        if (i < fw) fkeep[i] = (i + 1) * 32 <= nf ? ~0u : (i * 32 < nf ? ((1u << (nf & 31)) - 1) : 0u);
        if (i < 8) misc[i] = 0;
    }
}

// popcounts of bitmap words -> exclusive prefix (int64), 3-phase scan
__global__ void k_scan_blocks(const uint32_t *bits, int64_t nw, int64_t *pre, int64_t *bsum) {
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

__global__ void k_scan_sums(int64_t *bsum, int64_t nb, int64_t *total) {
    // single block of 1024 threads; sequential chunks per thread
    __shared__ int64_t sh[1024];
    int64_t per = (nb + 1023) / 1024;
    int64_t lo = threadIdx.x * per, hi = min(lo + per, nb);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += bsum[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t y = threadIdx.x >= (unsigned) o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += y;
        __syncthreads();
    }
    int64_t run = sh[threadIdx.x] - s;
    for (int64_t i = lo; i < hi; ++i) {
        int64_t b = bsum[i];
        bsum[i] = run;
        run += b;
    }
    if (threadIdx.x == 1023) *total = sh[1023];
}

__global__ void k_scan_add(int64_t *pre, int64_t nw, const int64_t *bsum) {
    int64_t i = (int64_t) blockIdx.x * 1024 + threadIdx.x;
    if (i < nw) pre[i] += bsum[blockIdx.x];
}

// i-th set bit (ones = true) or i-th clear bit (ones = false) of a bitmap
// whose per-word exclusive popcount prefix is pre (nw words).
__device__ int64_t bit_select(const uint32_t *bits, const int64_t *pre, int64_t nw, int64_t i, bool ones) {
    int64_t lo = 0, hi = nw;   // last word w with count_before(w) <= i
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        int64_t before = ones ? pre[mid] : 32 * mid - pre[mid];
        if (before <= i) lo = mid;
        else hi = mid;
    }
    int64_t before = ones ? pre[lo] : 32 * lo - pre[lo];
    uint32_t w = ones ? bits[lo] : ~bits[lo];
    int64_t r = i - before;
    for (int b = 0; b < 32; ++b) {
        if ((w >> b) & 1) {
            if (r == 0) return 32 * lo + b;
            --r;
        }
    }
    return -1;
}

// bit_select by one whole wave (every lane returns the same value): a 64-ary
// search over the word prefix -- each round the lanes test 64 words spread
// over the remaining range and a ballot keeps the span between the last word
// whose prefix is <= i and the next -- so a bitmap of 2^17 words takes three
// rounds of independent loads instead of 17 dependent ones.
__device__ int64_t bit_select_wave(const uint32_t *bits, const int64_t *pre, int64_t nw, int64_t i, bool ones) {
    const int lane = threadIdx.x & 63;
    int64_t lo = 0, hi = nw;   // before(lo) <= i; hi == nw or before(hi) > i
    while (hi - lo > 1) {
        const int64_t span = hi - lo - 1;                  // candidates lo + 1 .. hi - 1
        const int64_t w = lo + 1 + (span * lane) / 64;     // non-decreasing in the lane
        const bool ok = w < hi && (ones ? pre[w] : 32 * w - pre[w]) <= i;
        const uint64_t m = __ballot(ok);                    // lanes 0 .. c-1 (the predicate is monotone)
        const int c = __popcll(m);
        const int64_t lo2 = c > 0 ? lo + 1 + (span * (c - 1)) / 64 : lo;
        const int64_t hi2 = c < 64 ? lo + 1 + (span * c) / 64 : hi;
        lo = lo2;
        hi = hi2 > lo2 ? hi2 : lo2 + 1;
    }
    const int64_t before = ones ? pre[lo] : 32 * lo - pre[lo];
    const uint32_t wd = ones ? bits[lo] : ~bits[lo];
    int64_t r = i - before;
    for (int b = 0; b < 32; ++b) {
        if ((wd >> b) & 1) {
            if (r == 0) return 32 * lo + b;
            --r;
        }
    }
    return -1;
}

// Tail relocation: the i-th removed slot below n_final takes the i-th kept
// element at or above n_final (disjoint ranges, so the moves do not race);
// n_final is read from the device.  The pairs are recorded sorted by source
// (src = INT64_MAX past the last pair) for the face remap.  Moves vertex
// triples (and keys) when v is given, face triples when f is given.
__global__ void k_reloc(const uint32_t *keep, const int64_t *pre, int64_t nw, int64_t n, const int64_t *n_final_p,
                        int64_t kmax, int64_t *src, int64_t *dst, float *v, int64_t *vkeys, int32_t *f,
                        int64_t *overflow) {
    const int64_t n_final = *n_final_p, k = n - n_final;
    if (k > kmax) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *overflow = 1;
        return;
    }
    const uint32_t wd = keep[n_final >> 5];
    const uint32_t below = (n_final & 31) ? (wd & ((1u << (n_final & 31)) - 1)) : 0u;
    const int64_t holes_low = n_final - (pre[n_final >> 5] + __popc(below));
    // one wave per element (the searches are wave-cooperative)
    const int64_t wave = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwave = ((int64_t) gridDim.x * blockDim.x) >> 6;
    const bool lead = (threadIdx.x & 63) == 0;
    for (int64_t i = wave; i < k; i += nwave) {
        if (i < holes_low) {
            const int64_t d = bit_select_wave(keep, pre, nw, i, false);
            const int64_t sidx = bit_select_wave(keep, pre, nw, (n_final - holes_low) + i, true);
            if (!lead) continue;
            dst[i] = d;
            src[i] = sidx;
            if (v) {
                v[3 * d + 0] = v[3 * sidx + 0];
                v[3 * d + 1] = v[3 * sidx + 1];
                v[3 * d + 2] = v[3 * sidx + 2];
                if (vkeys) vkeys[d] = vkeys[sidx];
            }
            if (f) {
                f[3 * d + 0] = f[3 * sidx + 0];
                f[3 * d + 1] = f[3 * sidx + 1];
                f[3 * d + 2] = f[3 * sidx + 2];
            }
        } else if (lead) {
            dst[i] = -1;
            src[i] = INT64_MAX;
        }
    }
}

// First face that can reference a vertex with id >= v under dense
// point-major numbering: vertices owned by points of x-plane x are
// referenced only by cells of x-planes x - 1 and x.
__device__ int64_t first_face_for_tail(const FixupGrid &g, int64_t v) {
    if (!g.blk_v) return 0;
    auto prefix = [&](const int64_t *blk, const uint32_t *cnt, int64_t x) {
        const int64_t e = x * g.Y * g.S;
        return blk[e / kMcScanBlock] + (int64_t) cnt[e];
    };
    int64_t lo = 0, hi = g.X;   // prefix(lo) <= v
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (prefix(g.blk_v, g.cnt_v, mid) <= v) lo = mid;
        else hi = mid;
    }
    return lo >= 1 ? prefix(g.blk_f, g.cnt_f, lo - 1) : 0;
}

// Face ids of relocated tail vertices -> their new slot (faces from the
// first one that can reference the tail).
__global__ void k_remap_tail(int32_t *f, int64_t nf, const int64_t *nv_final_p, int64_t nv, int64_t kmax,
                             const int64_t *rsrc, const int64_t *rdst, FixupGrid g) {
    const int64_t nv_final = *nv_final_p, k = nv - nv_final;
    if (k <= 0 || k > kmax) return;   // k > kmax: flagged by k_reloc
    __shared__ int64_t f0s;
    if (threadIdx.x == 0) f0s = first_face_for_tail(g, nv_final);
    __syncthreads();
    const int64_t f0 = f0s;
    for (int64_t i = f0 + (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nf;
         i += (int64_t) gridDim.x * blockDim.x) {
        for (int j = 0; j < 3; ++j) {
            const int64_t v0 = f[3 * i + j];
            if (v0 < nv_final) continue;
            int64_t lo = 0, hi = k;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (rsrc[mid] < v0) lo = mid + 1;
                else hi = mid;
            }
            f[3 * i + j] = (lo < k && rsrc[lo] == v0) ? (int32_t) rdst[lo] : -1;
        }
    }
}

unsigned grid_for(int64_t n) { return (unsigned) std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)); }
// k_reloc: one wave (of four per workgroup) per element.
unsigned reloc_grid(int64_t n) { return (unsigned) std::max<int64_t>(1, std::min<int64_t>(4096, (n + 3) / 4)); }

int64_t pow2_at_least(int64_t n) {
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Bitmap prefix: pre[w] = number of set bits before word w.
hipError_t bitmap_prefix(const uint32_t *bits, int64_t nw, int64_t *pre, int64_t *scratch,
                         int64_t *total, hipStream_t st) {
    int64_t nb = (nw + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned) nb), dim3(1024), 0, st, bits, nw, pre, scratch);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, scratch, nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned) nb), dim3(1024), 0, st, pre, nw, scratch);
    return hipGetLastError();
}

}   // namespace

#define ISO_TRY(x)                               \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return e_;         \
    } while (0)

// ---- single-workgroup fix-up --------------------------------------------------

namespace {

constexpr int kFixNT = 1024;

// Ascending bitonic sort of a[0, n) in LDS, padded with INT32_MAX to m (a
// power of two >= n, <= 2 * kFixSmallMax).
__device__ void lds_sort(int32_t *a, int n, int m) {
    for (int i = n + (int) threadIdx.x; i < m; i += kFixNT) a[i] = INT32_MAX;
    __syncthreads();
    for (int k = 2; k <= m; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < m; i += kFixNT) {
                const int l = i ^ j;
                if (l > i) {
                    const int32_t x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    if (up ? (x > y) : (x < y)) {
                        a[i] = y;
                        a[l] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

__device__ int lower_bound_lds(const int32_t *a, int n, int64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int64_t) a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ int pow2_at_least_dev(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

}   // namespace

__global__ void __launch_bounds__(kFixNT) k_fix_small(const FixSmallItem *items) {
    const FixSmallItem &it = items[blockIdx.x];
    __shared__ int32_t rv[2 * kFixSmallMax];    // removed vertex ids, sorted
    __shared__ int32_t hf[2 * kFixSmallMax];    // dropped face slots, sorted
    __shared__ int32_t psrc[kFixSmallMax], pdst[kFixSmallMax];   // vertex relocation pairs by source
    __shared__ int nrv, nhf, bad;
    __shared__ int64_t f0;
    const int tid = threadIdx.x;
    const SpecialLists &sp = it.sp;
    const int64_t ncand = (int64_t) sp.n_cand, nscell = (int64_t) sp.n_scell;
    const int64_t nkept = (int64_t) sp.n_kept, nhole = (int64_t) sp.n_hole;
    const int64_t nv = it.nv, nf = it.nf;
    int64_t *keys = it.keys;
    VInfo *info = (VInfo *) it.info;
    PGroup *groups = (PGroup *) it.groups;
    const int64_t mask = it.nslots - 1, pmask = it.npslots - 1;
    if (tid == 0) {
        nrv = 0;
        nhf = 0;
        bad = 0;
    }
    for (int64_t i = tid; i < it.nslots; i += kFixNT) {
        keys[i] = kEmpty;
        info[i] = VInfo{0, 0, 0, -1, 0};
    }
    for (int64_t i = tid; i < it.npslots; i += kFixNT) {
        it.pslot[i] = -1;
        groups[i] = PGroup{INT64_MAX, 0, 0};
    }
    __syncthreads();
    for (int64_t i = tid; i < nscell + ncand; i += kFixNT) {
        if (i < nscell) insert_scell(sp.scell[i], keys, info, mask);
        else insert_cand(sp.cand, i - nscell, keys, info, mask, it.pslot, groups, pmask);
    }
    __syncthreads();
    // kept faces: mark their vertices (their corners are pointed at the
    // representatives below, once the lists are known to fit)
    for (int64_t i = tid; i < nkept; i += kFixNT) {
        const int64_t fs = sp.kept[i];
        if (fs < 0 || fs >= nf) continue;
        for (int j = 0; j < 3; ++j) {
            const int64_t sl = h1_find_or_insert(keys, mask, it.f[3 * fs + j], false);
            if (sl >= 0) info[sl].referenced = 1;
        }
    }
    __syncthreads();
    for (int64_t sl = tid; sl < it.nslots; sl += kFixNT) {
        if (keys[sl] == kEmpty) continue;
        VInfo &vi = info[sl];
        const bool used = vi.referenced || vi.count_s < edge_cell_count(vi.edge, it.g);
        if (vi.is_cand && vi.group >= 0) {
            if (used) atomicOr(&groups[vi.group].used, 1);
        } else {
            vi.is_cand = 0;
            vi.referenced = used ? 1 : 0;
        }
    }
    __syncthreads();
    // removed vertices and dropped faces into LDS lists
    for (int64_t sl = tid; sl < it.nslots; sl += kFixNT) {
        const int64_t vid = keys[sl];
        if (vid < 0 || vid >= nv) continue;
        const VInfo &vi = info[sl];
        const bool removed = (vi.is_cand && vi.group >= 0)
                                 ? (vid != groups[vi.group].rep || !groups[vi.group].used)
                                 : !vi.referenced;
        if (removed) {
            const int k = atomicAdd(&nrv, 1);
            if (k < kFixSmallMax) rv[k] = (int32_t) vid;
            else bad = 1;
        }
    }
    for (int64_t i = tid; i < nhole; i += kFixNT) {
        const int64_t fs = sp.hole[i];
        if (fs < 0 || fs >= nf) continue;
        const int k = atomicAdd(&nhf, 1);
        if (k < kFixSmallMax) hf[k] = (int32_t) fs;
        else bad = 1;
    }
    __syncthreads();
    if (bad) {   // the lists do not fit in LDS: outputs untouched, the multi-kernel fix-up takes over
        if (tid == 0) {
            it.final_dev[0] = nv;
            it.final_dev[1] = nf;
            it.final_dev[2] = kFixSmallOverflow;
        }
        return;
    }
    // kept faces: corners to the representatives
    for (int64_t i = tid; i < nkept; i += kFixNT) {
        const int64_t fs = sp.kept[i];
        if (fs < 0 || fs >= nf) continue;
        for (int j = 0; j < 3; ++j) {
            const int64_t vv = it.f[3 * fs + j];
            const int64_t sl = h1_find_or_insert(keys, mask, vv, false);
            if (sl >= 0 && info[sl].is_cand && info[sl].group >= 0) {
                const int64_t rep = groups[info[sl].group].rep;
                if (rep != vv) it.f[3 * fs + j] = (int32_t) rep;
            }
        }
    }
    __syncthreads();
    const int kv = nrv, kf = nhf;
    lds_sort(rv, kv, pow2_at_least_dev(max(kv, 1)));
    lds_sort(hf, kf, pow2_at_least_dev(max(kf, 1)));
    const int64_t nv2 = nv - kv, nf2 = nf - kf;
    // vertices: the i-th removed slot below nv2 takes the i-th surviving
    // vertex at or above nv2 (sorted order on both sides)
    const int hv = lower_bound_lds(rv, kv, nv2);   // removed slots below nv2
    for (int64_t t = nv2 + tid; t < nv; t += kFixNT) {
        const int lo = lower_bound_lds(rv, kv, t);
        if (lo < kv && rv[lo] == t) continue;   // removed itself
        const int rank = (int) (t - nv2) - (lo - hv);
        const int64_t d = rv[rank];
        it.v[3 * d + 0] = it.v[3 * t + 0];
        it.v[3 * d + 1] = it.v[3 * t + 1];
        it.v[3 * d + 2] = it.v[3 * t + 2];
        if (it.vkeys) it.vkeys[d] = it.vkeys[t];
        psrc[rank] = (int32_t) t;
        pdst[rank] = (int32_t) d;
    }
    if (tid == 0) f0 = first_face_for_tail(it.g, nv2);
    __syncthreads();
    // face ids of relocated tail vertices (faces from the first that can
    // reference the tail); the pairs are sorted by source
    if (hv > 0) {
        for (int64_t i = f0 + tid; i < nf; i += kFixNT) {
            for (int j = 0; j < 3; ++j) {
                const int64_t v0 = it.f[3 * i + j];
                if (v0 < nv2) continue;
                const int lo = lower_bound_lds(psrc, hv, v0);
                it.f[3 * i + j] = (lo < hv && psrc[lo] == v0) ? pdst[lo] : -1;
            }
        }
    }
    __syncthreads();
    // faces: the i-th dropped slot below nf2 takes the i-th kept face at or above nf2
    const int hfl = lower_bound_lds(hf, kf, nf2);
    for (int64_t t = nf2 + tid; t < nf; t += kFixNT) {
        const int lo = lower_bound_lds(hf, kf, t);
        if (lo < kf && hf[lo] == t) continue;
        const int rank = (int) (t - nf2) - (lo - hfl);
        const int64_t d = hf[rank];
        it.f[3 * d + 0] = it.f[3 * t + 0];
        it.f[3 * d + 1] = it.f[3 * t + 1];
        it.f[3 * d + 2] = it.f[3 * t + 2];
    }
    if (tid == 0) {
        it.final_dev[0] = nv2;
        it.final_dev[1] = nf2;
        it.final_dev[2] = 0;
    }
}

bool fixup_small_ok(const SpecialLists &sp) {
    return (int64_t) (sp.n_cand + sp.n_scell) <= kFixSmallSlots && (int64_t) sp.n_hole <= kFixSmallMax;
}

size_t fixup_small_scratch(const SpecialLists &sp, int64_t &nslots, int64_t &npslots) {
    nslots = pow2_at_least(2 * (int64_t) (sp.n_cand + sp.n_scell) + 16);
    npslots = pow2_at_least(2 * (int64_t) sp.n_cand + 16);
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    return up(sizeof(int64_t) * nslots) + up(sizeof(VInfo) * nslots) + up(sizeof(int32_t) * npslots) +
           up(sizeof(PGroup) * npslots);
}

hipError_t fixup_small_launch(const FixSmallItem *items_dev, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fix_small, dim3((unsigned) n), dim3(kFixNT), 0, st, items_dev);
    return hipGetLastError();
}

size_t weld_fixup_small_batch_bytes(int n, const SpecialLists *sp, size_t *host_bytes) {
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t bytes = up(sizeof(FixSmallItem) * (size_t) n) + up(sizeof(int64_t) * 3 * (size_t) n);
    for (int i = 0; i < n; ++i) {
        int64_t a, b;
        bytes += fixup_small_scratch(sp[i], a, b);
    }
    if (host_bytes) *host_bytes = up(sizeof(FixSmallItem) * (size_t) n) + sizeof(int64_t) * 3 * (size_t) n;
    return bytes;
}

hipError_t weld_fixup_small_batch(int n, const SpecialLists *sp, float *const *v, int32_t *const *f,
                                  int64_t *const *vkeys, const int64_t *nv, const int64_t *nf, const FixupGrid *g,
                                  hipStream_t st, int64_t *final_host, const FixupScratch *ext) {
    if (n <= 0) return hipSuccess;
    if (n > kMaxBatch) return hipErrorInvalidValue;
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    // device: item records, per-item tables, final counts; host: pinned staging
    int64_t ns[kMaxBatch], nps[kMaxBatch];
    size_t off[kMaxBatch];
    size_t bytes = up(sizeof(FixSmallItem) * (size_t) n) + up(sizeof(int64_t) * 3 * (size_t) n);
    for (int i = 0; i < n; ++i) {
        off[i] = bytes;
        bytes += fixup_small_scratch(sp[i], ns[i], nps[i]);
    }
    struct Cache {
        char *dev = nullptr;
        size_t dev_bytes = 0;
        char *host = nullptr;
        size_t host_bytes = 0;
        ~Cache() {   // at thread exit; the HIP runtime outlives thread-local objects
            if (host) (void) hipHostFree(host);
            if (dev) (void) hipFree(dev);
        }
    };
    const size_t hb = up(sizeof(FixSmallItem) * (size_t) n) + sizeof(int64_t) * 3 * (size_t) n;
    char *dev, *host;
    if (ext) {   // the caller's scratch: nothing is allocated here
        if (ext->dev_bytes < bytes || ext->host_bytes < hb) return hipErrorOutOfMemory;
        dev = ext->dev;
        host = ext->host;
    } else {
        // one cache per (thread, device): a thread that moves between GPUs
        // never hands one GPU's scratch to another
        int devid = 0;
        ISO_TRY(hipGetDevice(&devid));
        if (devid < 0 || devid >= kFixMaxDevices) return hipErrorInvalidDevice;
        thread_local Cache caches[kFixMaxDevices];
        Cache &c = caches[devid];
        if (c.dev_bytes < bytes) {
            if (c.dev) {
                ISO_TRY(hipStreamSynchronize(st));
                ISO_TRY(hipFree(c.dev));
                c.dev = nullptr;
                c.dev_bytes = 0;
            }
            ISO_TRY(hipMalloc((void **) &c.dev, bytes + bytes / 2));
            c.dev_bytes = bytes + bytes / 2;
        }
        if (c.host_bytes < hb) {
            if (c.host) {
                ISO_TRY(hipStreamSynchronize(st));
                ISO_TRY(hipHostFree(c.host));
                c.host = nullptr;
                c.host_bytes = 0;
            }
            ISO_TRY(hipHostMalloc((void **) &c.host, hb + hb / 2, 0));
            c.host_bytes = hb + hb / 2;
        }
        dev = c.dev;
        host = c.host;
    }
    FixSmallItem *hi = (FixSmallItem *) host;
    int64_t *fin_dev = (int64_t *) (dev + up(sizeof(FixSmallItem) * (size_t) n));
    for (int i = 0; i < n; ++i) {
        char *base = dev + off[i];
        FixSmallItem &it = hi[i];
        it.sp = sp[i];
        it.v = v[i];
        it.f = f[i];
        it.vkeys = vkeys ? vkeys[i] : nullptr;
        it.nv = nv[i];
        it.nf = nf[i];
        it.g = g[i];
        it.nslots = ns[i];
        it.npslots = nps[i];
        it.keys = (int64_t *) base;
        base += up(sizeof(int64_t) * ns[i]);
        it.info = base;
        base += up(sizeof(VInfo) * ns[i]);
        it.pslot = (int32_t *) base;
        base += up(sizeof(int32_t) * nps[i]);
        it.groups = base;
        it.final_dev = fin_dev + 3 * i;
    }
    ISO_TRY(hipMemcpyAsync(dev, hi, sizeof(FixSmallItem) * (size_t) n, hipMemcpyHostToDevice, st));
    ISO_TRY(fixup_small_launch((const FixSmallItem *) dev, n, st));
    int64_t *fin_host = (int64_t *) (host + up(sizeof(FixSmallItem) * (size_t) n));
    ISO_TRY(hipMemcpyAsync(fin_host, fin_dev, sizeof(int64_t) * 3 * (size_t) n, hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < 3 * n; ++i) final_host[i] = fin_host[i];
    return hipSuccess;
}

hipError_t weld_fixup_impl(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                           const FixupGrid &g, hipStream_t st, int64_t *vkeys, int64_t *final_dev,
                           const FixupScratch *ext);

__global__ void k_fix_final(const int64_t *misc, int64_t *final_dev) {
    if (threadIdx.x == 0) {
        final_dev[0] = misc[2];
        final_dev[1] = misc[3];
        final_dev[2] = misc[4];
    }
}

// Resolves the special lists in place.  nv / nf hold the provisional counts on
// entry and the final counts on return.  Synchronises the stream once, at
// the end.
hipError_t weld_fixup(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                      const FixupGrid &g, hipStream_t st, int64_t *vkeys, const FixupScratch *ext) {
    return weld_fixup_impl(sp, v, f, nv, nf, g, st, vkeys, nullptr, ext);
}

hipError_t weld_fixup_async(const SpecialLists &sp, float *v, int32_t *f, int64_t nv, int64_t nf,
                            const FixupGrid &g, hipStream_t st, int64_t *vkeys, int64_t *final_dev) {
    return weld_fixup_impl(sp, v, f, nv, nf, g, st, vkeys, final_dev, nullptr);
}

// final_dev == nullptr: read the final counts back (one sync) into nv / nf;
// else leave them in final_dev.
// Scratch bytes of the multi-kernel fix-up for these lists and provisional counts.
static size_t fixup_layout(const SpecialLists &sp, int64_t nv, int64_t nf, size_t o[12]) {
    const int64_t ncand = (int64_t) sp.n_cand, nscell = (int64_t) sp.n_scell, nhole = (int64_t) sp.n_hole;
    const int64_t nslots = pow2_at_least(2 * (ncand + nscell) + 16);
    const int64_t npslots = pow2_at_least(2 * ncand + 16);
    const int64_t vw = nv / 32 + 2, fw = nf / 32 + 2;
    size_t bytes = 0;
    auto take = [&](size_t n) {
        size_t r = bytes;
        bytes += (n + 255) & ~size_t(255);
        return r;
    };
    const int64_t nrel = ncand + nscell + 16, nfrel = nhole + 16;
    o[0] = take(sizeof(int64_t) * nslots);
    o[1] = take(sizeof(VInfo) * nslots);
    o[2] = take(sizeof(int32_t) * npslots);
    o[3] = take(sizeof(PGroup) * npslots);
    o[4] = take(sizeof(uint32_t) * vw);
    o[5] = take(sizeof(uint32_t) * fw);
    o[6] = take(sizeof(int64_t) * vw);
    o[7] = take(sizeof(int64_t) * fw);
    o[8] = take(sizeof(int64_t) * ((vw + 1023) / 1024 + 1));
    o[9] = take(sizeof(int64_t) * ((fw + 1023) / 1024 + 1));
    o[10] = take(sizeof(int64_t) * 8);
    o[11] = take(sizeof(int64_t) * 2 * nrel);
    take(sizeof(int64_t) * 2 * nfrel);
    return bytes;
}

size_t weld_fixup_bytes(const SpecialLists &sp, int64_t nv, int64_t nf) {
    size_t o[12];
    return fixup_layout(sp, nv, nf, o);
}

hipError_t weld_fixup_impl(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                           const FixupGrid &g, hipStream_t st, int64_t *vkeys, int64_t *final_dev,
                           const FixupScratch *ext) {
    const int64_t ncand = (int64_t) sp.n_cand, nscell = (int64_t) sp.n_scell;
    const int64_t nkept = (int64_t) sp.n_kept, nhole = (int64_t) sp.n_hole;
    if (ncand == 0 && nhole == 0) {
        if (final_dev) {
            const int64_t h[3] = {nv, nf, 0};
            ISO_TRY(hipMemcpyAsync(final_dev, h, sizeof(h), hipMemcpyHostToDevice, st));
            ISO_TRY(hipStreamSynchronize(st));   // h lives on this stack frame
        }
        return hipSuccess;
    }
    const int64_t nslots = pow2_at_least(2 * (ncand + nscell) + 16), mask = nslots - 1;
    const int64_t npslots = pow2_at_least(2 * ncand + 16), pmask = npslots - 1;
    const int64_t vw = nv / 32 + 2, fw = nf / 32 + 2;
    const int64_t nrel = ncand + nscell + 16;   // removed vertices are hashed vertices
    const int64_t nfrel = nhole + 16;           // dropped faces are listed holes
    size_t lo[12];
    const size_t bytes = fixup_layout(sp, nv, nf, lo);
    const size_t o_keys = lo[0], o_info = lo[1], o_pslot = lo[2], o_groups = lo[3], o_vkeep = lo[4],
                 o_fkeep = lo[5], o_vpre = lo[6], o_fpre = lo[7], o_vscr = lo[8], o_fscr = lo[9], o_misc = lo[10],
                 o_rv = lo[11];
    const size_t o_rf = o_rv + ((sizeof(int64_t) * 2 * nrel + 255) & ~size_t(255));
    char *base;
    if (ext) {   // the caller's scratch: nothing is allocated here
        if (ext->dev_bytes < bytes || (!final_dev && ext->host_bytes < 8 * sizeof(int64_t)))
            return hipErrorOutOfMemory;
        base = ext->dev;
    } else {   // a cached per-thread buffer, grown on demand
        struct DevCache {   // freed when the thread ends
            char *p = nullptr;
            size_t bytes = 0;
            ~DevCache() {
                if (p) (void) hipFree(p);
            }
        };
        int devid = 0;   // one cache per (thread, device)
        ISO_TRY(hipGetDevice(&devid));
        if (devid < 0 || devid >= kFixMaxDevices) return hipErrorInvalidDevice;
        static thread_local DevCache caches[kFixMaxDevices];
        DevCache &cache = caches[devid];
        if (cache.bytes < bytes) {
            if (cache.p) {
                ISO_TRY(hipStreamSynchronize(st));
                ISO_TRY(hipFree(cache.p));
                cache.p = nullptr;
                cache.bytes = 0;
            }
            size_t want = bytes + bytes / 2;
            ISO_TRY(hipMalloc((void **) &cache.p, want));
            cache.bytes = want;
        }
        base = cache.p;
    }
    int64_t *keys = (int64_t *) (base + o_keys);
    VInfo *info = (VInfo *) (base + o_info);
    int32_t *pslot = (int32_t *) (base + o_pslot);
    PGroup *groups = (PGroup *) (base + o_groups);
    uint32_t *vkeep = (uint32_t *) (base + o_vkeep), *fkeep = (uint32_t *) (base + o_fkeep);
    int64_t *vpre = (int64_t *) (base + o_vpre), *fpre = (int64_t *) (base + o_fpre);
    int64_t *vscr = (int64_t *) (base + o_vscr), *fscr = (int64_t *) (base + o_fscr);
    int64_t *misc = (int64_t *) (base + o_misc);   // 0 removed, 2 nv final, 3 nf final, 4 overflow
    int64_t *rv_src = (int64_t *) (base + o_rv), *rv_dst = rv_src + nrel;
    int64_t *rf_src = (int64_t *) (base + o_rf), *rf_dst = rf_src + nfrel;

    hipLaunchKernelGGL(k_fix_init, dim3(grid_for(std::max(std::max(nslots, npslots), std::max(vw, fw)))), dim3(256),
                       0, st, keys, info, nslots, pslot, groups, npslots, vkeep, vw, nv, fkeep, fw, nf, misc);
    if (nscell + ncand)
        hipLaunchKernelGGL(k_fix_insert, dim3(grid_for(nscell + ncand)), dim3(256), 0, st, sp.scell, nscell, sp.cand,
                           ncand, keys, info, mask, pslot, groups, pmask);
    if (nkept)
        hipLaunchKernelGGL(k_kept_refs, dim3(grid_for(nkept)), dim3(256), 0, st, sp.kept, nkept, f, nf, keys, info,
                           mask, groups);
    hipLaunchKernelGGL(k_used, dim3(grid_for(nslots)), dim3(256), 0, st, keys, info, nslots, groups, g);
    hipLaunchKernelGGL(k_removed, dim3(grid_for(std::max(nslots, nhole))), dim3(256), 0, st, keys, info, nslots,
                       groups, vkeep, nv, (unsigned long long *) misc, sp.hole, nhole, fkeep, nf);
    ISO_TRY(hipGetLastError());
    ISO_TRY(bitmap_prefix(vkeep, vw, vpre, vscr, misc + 2, st));
    ISO_TRY(bitmap_prefix(fkeep, fw, fpre, fscr, misc + 3, st));
    // vertices: tail survivors fill the removed slots; faces: representatives
    // were applied above, relocated tail vertices now; then tail faces fill
    // the dropped slots
    hipLaunchKernelGGL(k_reloc, dim3(reloc_grid(nrel)), dim3(256), 0, st, vkeep, vpre, vw, nv, misc + 2, nrel, rv_src,
                       rv_dst, v, vkeys, (int32_t *) nullptr, misc + 4);
    hipLaunchKernelGGL(k_remap_tail, dim3(grid_for(nf)), dim3(256), 0, st, f, nf, misc + 2, nv, nrel, rv_src, rv_dst,
                       g);
    hipLaunchKernelGGL(k_reloc, dim3(reloc_grid(nfrel)), dim3(256), 0, st, fkeep, fpre, fw, nf, misc + 3, nfrel,
                       rf_src, rf_dst, (float *) nullptr, (int64_t *) nullptr, f, misc + 4);
    ISO_TRY(hipGetLastError());
    if (final_dev) {
        hipLaunchKernelGGL(k_fix_final, dim3(1), dim3(64), 0, st, misc, final_dev);
        return hipGetLastError();
    }
    // final counts through a small pinned buffer of the calling thread
    struct PinnedCounts {
        int64_t *p = nullptr;
        ~PinnedCounts() {
            if (p) (void) hipHostFree(p);
        }
    };
    int64_t *hp;
    if (ext) {
        hp = (int64_t *) ext->host;
    } else {
        thread_local PinnedCounts pc;
        if (!pc.p) ISO_TRY(hipHostMalloc((void **) &pc.p, 8 * sizeof(int64_t), 0));
        hp = pc.p;
    }
    ISO_TRY(hipMemcpyAsync(hp, misc, 5 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    int64_t h[5];
    for (int i = 0; i < 5; ++i) h[i] = hp[i];
    if (h[4]) return hipErrorInvalidValue;   // more relocations than the lists allow: inconsistent lists
    nv = h[2];
    nf = h[3];
    return hipSuccess;
}

}   // namespace iso
