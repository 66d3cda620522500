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

__global__ void k_clear64(int64_t *p, int64_t n, int64_t v) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_clear32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_scell(const SpecialEdge *se, int64_t n, int64_t *keys, VInfo *info, int64_t mask) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        int64_t s = h1_find_or_insert(keys, mask, se[i].vid, true);
        if (s < 0) continue;
        atomicAdd(&info[s].count_s, 1);
        info[s].edge = se[i].edge;
    }
}

__global__ void k_cand(const SpecialVertex *cv, int64_t n, int64_t *keys, VInfo *info, int64_t mask,
                       int32_t *pslot, PGroup *groups, int64_t pmask) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
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
}

__global__ void k_kept_refs(const int64_t *kept, int64_t n, const int32_t *faces, int64_t nf,
                            int64_t *keys, VInfo *info, int64_t mask, uint32_t *special_face_bits) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        int64_t f = kept[i];
        if (f < 0 || f >= nf) continue;
        atomicOr(&special_face_bits[f >> 5], 1u << (f & 31));
        for (int j = 0; j < 3; ++j) {
            int64_t s = h1_find_or_insert(keys, mask, (int64_t) faces[3 * f + j], false);
            if (s >= 0) info[s].referenced = 1;
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

__global__ void k_removed(const int64_t *keys, const VInfo *info, int64_t nslots, const PGroup *groups,
                          uint32_t *vkeep, int64_t nv, unsigned long long *nremoved) {
    for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < nslots;
         s += (int64_t) gridDim.x * blockDim.x) {
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

__global__ void k_holes(const int64_t *hole, int64_t n, uint32_t *fkeep, int64_t nf) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        int64_t f = hole[i];
        if (f < 0 || f >= nf) continue;
        atomicAnd(&fkeep[f >> 5], ~(1u << (f & 31)));
    }
}

__device__ __forceinline__ int64_t bit_rank(const uint32_t *bits, const int64_t *prefix, int64_t i) {
    uint32_t w = bits[i >> 5];
    uint32_t below = (i & 31) ? (w & ((1u << (i & 31)) - 1)) : 0u;
    return prefix[i >> 5] + __popc(below);
}

__global__ void k_tail_bits(uint32_t *vkeep, int64_t nv, uint32_t *fkeep, int64_t nf) {
    if (threadIdx.x == 0) {
        vkeep[nv >> 5] = (nv & 31) ? ((1u << (nv & 31)) - 1) : 0u;
        vkeep[(nv >> 5) + 1] = 0u;
        fkeep[nf >> 5] = (nf & 31) ? ((1u << (nf & 31)) - 1) : 0u;
        fkeep[(nf >> 5) + 1] = 0u;
    }
}

__global__ void k_init_groups(PGroup *g, int64_t n) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        g[i].rep = INT64_MAX;
        g[i].used = 0;
        g[i].pad = 0;
    }
}

__global__ void k_init_info(VInfo *v, int64_t n) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        v[i].count_s = 0;
        v[i].referenced = 0;
        v[i].is_cand = 0;
        v[i].group = -1;
        v[i].edge = 0;
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

// Relocation pairs: the i-th removed slot below n_final takes the i-th kept
// element at or above n_final (disjoint ranges, so moves do not race).
__global__ void k_reloc_pairs(const uint32_t *keep, const int64_t *pre, int64_t nw, int64_t n_final, int64_t k,
                              int64_t *src, int64_t *dst) {
    // removed slots below n_final (the rest of the removed ones are in the tail)
    const uint32_t wd = keep[n_final >> 5];
    const uint32_t below = (n_final & 31) ? (wd & ((1u << (n_final & 31)) - 1)) : 0u;
    const int64_t holes_low = n_final - (pre[n_final >> 5] + __popc(below));
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t) gridDim.x * blockDim.x) {
        if (i < holes_low) {
            dst[i] = bit_select(keep, pre, nw, i, false);
            src[i] = bit_select(keep, pre, nw, (n_final - holes_low) + i, true);
        } else {
            dst[i] = -1;
            src[i] = INT64_MAX;   // keeps src sorted for the remap's binary search
        }
    }
}

__global__ void k_move_v(float *v, int64_t *vkeys, const int64_t *src, const int64_t *dst, int64_t k) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t) gridDim.x * blockDim.x) {
        int64_t s = src[i], d = dst[i];
        if (d < 0 || s == INT64_MAX) continue;
        v[3 * d + 0] = v[3 * s + 0];
        v[3 * d + 1] = v[3 * s + 1];
        v[3 * d + 2] = v[3 * s + 2];
        if (vkeys) vkeys[d] = vkeys[s];
    }
}

__global__ void k_move_f(int32_t *f, const int64_t *src, const int64_t *dst, int64_t k) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t) gridDim.x * blockDim.x) {
        int64_t s = src[i], d = dst[i];
        if (d < 0 || s == INT64_MAX) continue;
        f[3 * d + 0] = f[3 * s + 0];
        f[3 * d + 1] = f[3 * s + 1];
        f[3 * d + 2] = f[3 * s + 2];
    }
}

// Rewrites face ids in place: non-representatives -> representative (only
// special faces can reference them), then tail vertices -> relocated slot.
__global__ void k_remap_inplace(int32_t *f, int64_t nf, const uint32_t *fkeep, const uint32_t *special_face_bits,
                                const int64_t *keys, const VInfo *info, int64_t mask, const PGroup *groups,
                                int64_t nv_final, const int64_t *rsrc, const int64_t *rdst, int64_t k) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += (int64_t) gridDim.x * blockDim.x) {
        if (!((fkeep[i >> 5] >> (i & 31)) & 1)) continue;
        const bool sp = (special_face_bits[i >> 5] >> (i & 31)) & 1;
        for (int j = 0; j < 3; ++j) {
            int64_t v0 = f[3 * i + j], v = v0;
            if (sp) {
                int64_t s = h1_find_or_insert(const_cast<int64_t *>(keys), mask, v, false);
                if (s >= 0 && info[s].is_cand && info[s].group >= 0) v = groups[info[s].group].rep;
            }
            if (v >= nv_final) {
                int64_t lo = 0, hi = k;
                while (lo < hi) {
                    int64_t mid = (lo + hi) >> 1;
                    if (rsrc[mid] < v) lo = mid + 1;
                    else hi = mid;
                }
                v = (lo < k && rsrc[lo] == v) ? rdst[lo] : -1;
            }
            if (v != v0) f[3 * i + j] = (int32_t) v;
        }
    }
}

unsigned grid_for(int64_t n) { return (unsigned) std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)); }

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

// Resolves the special lists in place.  nv / nf hold the provisional counts on
// entry and the final counts on return.  Synchronises the stream.
hipError_t weld_fixup(const SpecialLists &sp, float *v, int32_t *f, int64_t &nv, int64_t &nf,
                      const FixupGrid &g, hipStream_t st, int64_t *vkeys) {
    const int64_t ncand = (int64_t) sp.n_cand, nscell = (int64_t) sp.n_scell;
    const int64_t nkept = (int64_t) sp.n_kept, nhole = (int64_t) sp.n_hole;
    if (ncand == 0 && nhole == 0) return hipSuccess;
    const int64_t nslots = pow2_at_least(2 * (ncand + nscell) + 16), mask = nslots - 1;
    const int64_t npslots = pow2_at_least(2 * ncand + 16), pmask = npslots - 1;
    const int64_t vw = nv / 32 + 2, fw = nf / 32 + 2;
    // scratch layout (a cached per-thread buffer, grown on demand)
    size_t bytes = 0;
    auto take = [&](size_t n) {
        size_t o = bytes;
        bytes += (n + 255) & ~size_t(255);
        return o;
    };
    const int64_t nrel = ncand + nscell + 16;
    size_t o_keys = take(sizeof(int64_t) * nslots), o_info = take(sizeof(VInfo) * nslots);
    size_t o_pslot = take(sizeof(int32_t) * npslots), o_groups = take(sizeof(PGroup) * npslots);
    size_t o_vkeep = take(sizeof(uint32_t) * vw), o_fkeep = take(sizeof(uint32_t) * fw);
    size_t o_sfb = take(sizeof(uint32_t) * fw);
    size_t o_vpre = take(sizeof(int64_t) * vw), o_fpre = take(sizeof(int64_t) * fw);
    size_t o_scr = take(sizeof(int64_t) * ((std::max(vw, fw) + 1023) / 1024 + 1));
    size_t o_misc = take(sizeof(int64_t) * 8);
    size_t o_rv = take(sizeof(int64_t) * 2 * nrel);
    size_t o_rf = take(sizeof(int64_t) * 2 * (nhole + 16));
    static thread_local char *cache = nullptr;
    static thread_local size_t cache_bytes = 0;
    if (cache_bytes < bytes) {
        if (cache) {
            ISO_TRY(hipStreamSynchronize(st));
            ISO_TRY(hipFree(cache));
            cache = nullptr;
            cache_bytes = 0;
        }
        size_t want = bytes + bytes / 2;
        ISO_TRY(hipMalloc((void **) &cache, want));
        cache_bytes = want;
    }
    char *base = cache;
    int64_t *keys = (int64_t *) (base + o_keys);
    VInfo *info = (VInfo *) (base + o_info);
    int32_t *pslot = (int32_t *) (base + o_pslot);
    PGroup *groups = (PGroup *) (base + o_groups);
    uint32_t *vkeep = (uint32_t *) (base + o_vkeep), *fkeep = (uint32_t *) (base + o_fkeep);
    uint32_t *sfb = (uint32_t *) (base + o_sfb);
    int64_t *vpre = (int64_t *) (base + o_vpre), *fpre = (int64_t *) (base + o_fpre);
    int64_t *scr = (int64_t *) (base + o_scr), *misc = (int64_t *) (base + o_misc);
    int64_t *rv_src = (int64_t *) (base + o_rv), *rv_dst = rv_src + nrel;
    int64_t *rf_src = (int64_t *) (base + o_rf), *rf_dst = rf_src + (nhole + 16);

    hipLaunchKernelGGL(k_clear64, dim3(grid_for(nslots)), dim3(256), 0, st, keys, nslots, kEmpty);
    hipLaunchKernelGGL(k_init_info, dim3(grid_for(nslots)), dim3(256), 0, st, info, nslots);
    hipLaunchKernelGGL(k_clear32, dim3(grid_for(npslots)), dim3(256), 0, st, (uint32_t *) pslot, npslots,
                       0xffffffffu);
    hipLaunchKernelGGL(k_init_groups, dim3(grid_for(npslots)), dim3(256), 0, st, groups, npslots);
    // keep bitmaps: bits set for [0, n), clear beyond
    hipLaunchKernelGGL(k_clear32, dim3(grid_for(vw)), dim3(256), 0, st, vkeep, vw, 0xffffffffu);
    hipLaunchKernelGGL(k_clear32, dim3(grid_for(fw)), dim3(256), 0, st, fkeep, fw, 0xffffffffu);
    hipLaunchKernelGGL(k_clear32, dim3(grid_for(fw)), dim3(256), 0, st, sfb, fw, 0u);
    hipLaunchKernelGGL(k_clear64, dim3(1), dim3(64), 0, st, misc, 8, 0);
    hipLaunchKernelGGL(k_tail_bits, dim3(1), dim3(64), 0, st, vkeep, nv, fkeep, nf);
    ISO_TRY(hipGetLastError());
    if (nscell)
        hipLaunchKernelGGL(k_scell, dim3(grid_for(nscell)), dim3(256), 0, st, sp.scell, nscell, keys, info,
                           mask);
    if (ncand)
        hipLaunchKernelGGL(k_cand, dim3(grid_for(ncand)), dim3(256), 0, st, sp.cand, ncand, keys, info, mask,
                           pslot, groups, pmask);
    if (nkept)
        hipLaunchKernelGGL(k_kept_refs, dim3(grid_for(nkept)), dim3(256), 0, st, sp.kept, nkept, f, nf,
                           keys, info, mask, sfb);
    hipLaunchKernelGGL(k_used, dim3(grid_for(nslots)), dim3(256), 0, st, keys, info, nslots, groups, g);
    hipLaunchKernelGGL(k_removed, dim3(grid_for(nslots)), dim3(256), 0, st, keys, info, nslots, groups,
                       vkeep, nv, (unsigned long long *) misc);
    if (nhole) hipLaunchKernelGGL(k_holes, dim3(grid_for(nhole)), dim3(256), 0, st, sp.hole, nhole, fkeep, nf);
    ISO_TRY(hipGetLastError());
    ISO_TRY(bitmap_prefix(vkeep, vw, vpre, scr, misc + 2, st));
    ISO_TRY(bitmap_prefix(fkeep, fw, fpre, scr, misc + 3, st));
    int64_t h[4];
    ISO_TRY(hipMemcpyAsync(h, misc, sizeof(h), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    const int64_t nv2 = h[2], nf2 = h[3];
    const int64_t kv = nv - nv2, kf = nf - nf2;   // removed vertices / dropped faces
    if (kv > nrel || kf > nhole + 16) return hipErrorInvalidValue;
    // vertex relocation: tail survivors fill the removed slots
    if (kv > 0) {
        hipLaunchKernelGGL(k_reloc_pairs, dim3(grid_for(kv)), dim3(256), 0, st, vkeep, vpre, vw, nv2, kv, rv_src,
                           rv_dst);
        hipLaunchKernelGGL(k_move_v, dim3(grid_for(kv)), dim3(256), 0, st, v, vkeys, rv_src, rv_dst, kv);
    }
    // face ids: representatives, then relocated tail vertices
    hipLaunchKernelGGL(k_remap_inplace, dim3(grid_for(nf)), dim3(256), 0, st, f, nf, fkeep, sfb, keys, info, mask,
                       groups, nv2, rv_src, rv_dst, kv);
    // face relocation: tail faces fill the dropped slots
    if (kf > 0) {
        hipLaunchKernelGGL(k_reloc_pairs, dim3(grid_for(kf)), dim3(256), 0, st, fkeep, fpre, fw, nf2, kf, rf_src,
                           rf_dst);
        hipLaunchKernelGGL(k_move_f, dim3(grid_for(kf)), dim3(256), 0, st, f, rf_src, rf_dst, kf);
    }
    ISO_TRY(hipGetLastError());
    ISO_TRY(hipStreamSynchronize(st));
    nv = nv2;
    nf = nf2;
    return hipSuccess;
}

}   // namespace iso
