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

__global__ void k_compact_v(const float *in, float *out, int64_t n, const uint32_t *vkeep,
                            const int64_t *vpre, const int64_t *kin, int64_t *kout) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        if (!((vkeep[i >> 5] >> (i & 31)) & 1)) continue;
        int64_t j = bit_rank(vkeep, vpre, i);
        out[3 * j + 0] = in[3 * i + 0];
        out[3 * j + 1] = in[3 * i + 1];
        out[3 * j + 2] = in[3 * i + 2];
        if (kin) kout[j] = kin[i];
    }
}

__global__ void k_compact_f(const int32_t *in, int32_t *out, int64_t n, const uint32_t *fkeep,
                            const int64_t *fpre, const uint32_t *vkeep, const int64_t *vpre,
                            const uint32_t *special_face_bits, const int64_t *keys, const VInfo *info,
                            int64_t mask, const PGroup *groups, int64_t nv) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        if (!((fkeep[i >> 5] >> (i & 31)) & 1)) continue;
        int64_t j = bit_rank(fkeep, fpre, i);
        bool sp = (special_face_bits[i >> 5] >> (i & 31)) & 1;
        for (int k = 0; k < 3; ++k) {
            int64_t v = in[3 * i + k];
            if (sp) {
                int64_t s = h1_find_or_insert(const_cast<int64_t *>(keys), mask, v, false);
                if (s >= 0 && info[s].is_cand && info[s].group >= 0) v = groups[info[s].group].rep;
            }
            out[3 * j + k] = (v >= 0 && v < nv) ? (int32_t) bit_rank(vkeep, vpre, v) : -1;
        }
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
    const int64_t vw = (nv + 31) / 32 + 1, fw = (nf + 31) / 32 + 1;
    // scratch layout
    size_t bytes = 0;
    auto take = [&](size_t n) {
        size_t o = bytes;
        bytes += (n + 255) & ~size_t(255);
        return o;
    };
    size_t o_keys = take(sizeof(int64_t) * nslots), o_info = take(sizeof(VInfo) * nslots);
    size_t o_pslot = take(sizeof(int32_t) * npslots), o_groups = take(sizeof(PGroup) * npslots);
    size_t o_vkeep = take(sizeof(uint32_t) * vw), o_fkeep = take(sizeof(uint32_t) * fw);
    size_t o_sfb = take(sizeof(uint32_t) * fw);
    size_t o_vpre = take(sizeof(int64_t) * vw), o_fpre = take(sizeof(int64_t) * fw);
    size_t o_scr = take(sizeof(int64_t) * ((std::max(vw, fw) + 1023) / 1024 + 1));
    size_t o_misc = take(sizeof(int64_t) * 8);
    size_t o_v2 = take(sizeof(float) * 3 * std::max<int64_t>(nv, 1));
    size_t o_f2 = take(sizeof(int32_t) * 3 * std::max<int64_t>(nf, 1));
    size_t o_k2 = take(sizeof(int64_t) * std::max<int64_t>(vkeys ? nv : 1, 1));
    char *base = nullptr;
    ISO_TRY(hipMallocAsync((void **) &base, bytes, st));
    int64_t *keys = (int64_t *) (base + o_keys);
    VInfo *info = (VInfo *) (base + o_info);
    int32_t *pslot = (int32_t *) (base + o_pslot);
    PGroup *groups = (PGroup *) (base + o_groups);
    uint32_t *vkeep = (uint32_t *) (base + o_vkeep), *fkeep = (uint32_t *) (base + o_fkeep);
    uint32_t *sfb = (uint32_t *) (base + o_sfb);
    int64_t *vpre = (int64_t *) (base + o_vpre), *fpre = (int64_t *) (base + o_fpre);
    int64_t *scr = (int64_t *) (base + o_scr), *misc = (int64_t *) (base + o_misc);
    float *v2 = (float *) (base + o_v2);
    int32_t *f2 = (int32_t *) (base + o_f2);
    int64_t *k2 = (int64_t *) (base + o_k2);

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
    ISO_TRY(hipGetLastError());
    // tail bits beyond nv / nf
    {
        uint32_t tailv[2] = {nv % 32 ? ((1u << (nv % 32)) - 1) : 0xffffffffu, 0u};
        uint32_t tailf[2] = {nf % 32 ? ((1u << (nf % 32)) - 1) : 0xffffffffu, 0u};
        if (nv % 32) ISO_TRY(hipMemcpyAsync(vkeep + nv / 32, tailv, 8, hipMemcpyHostToDevice, st));
        else ISO_TRY(hipMemcpyAsync(vkeep + nv / 32, tailv + 1, 4, hipMemcpyHostToDevice, st));
        if (nf % 32) ISO_TRY(hipMemcpyAsync(fkeep + nf / 32, tailf, 8, hipMemcpyHostToDevice, st));
        else ISO_TRY(hipMemcpyAsync(fkeep + nf / 32, tailf + 1, 4, hipMemcpyHostToDevice, st));
        ISO_TRY(hipStreamSynchronize(st));
    }
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
    hipLaunchKernelGGL(k_compact_v, dim3(grid_for(nv)), dim3(256), 0, st, v, v2, nv, vkeep, vpre, vkeys, k2);
    hipLaunchKernelGGL(k_compact_f, dim3(grid_for(nf)), dim3(256), 0, st, f, f2, nf, fkeep, fpre, vkeep,
                       vpre, sfb, keys, info, mask, groups, nv);
    ISO_TRY(hipGetLastError());
    int64_t h[4];
    ISO_TRY(hipMemcpyAsync(h, misc, sizeof(h), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    const int64_t nv2 = h[2], nf2 = h[3];
    ISO_TRY(hipMemcpyAsync(v, v2, sizeof(float) * 3 * nv2, hipMemcpyDeviceToDevice, st));
    ISO_TRY(hipMemcpyAsync(f, f2, sizeof(int32_t) * 3 * nf2, hipMemcpyDeviceToDevice, st));
    if (vkeys) ISO_TRY(hipMemcpyAsync(vkeys, k2, sizeof(int64_t) * nv2, hipMemcpyDeviceToDevice, st));
    ISO_TRY(hipFreeAsync(base, st));
    ISO_TRY(hipStreamSynchronize(st));
    nv = nv2;
    nf = nf2;
    return hipSuccess;
}

}   // namespace iso
