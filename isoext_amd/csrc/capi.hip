// C-ABI entry points of libisoext_amd.so.  Plain pointers and int64 sizes
// only; no exceptions cross this boundary.  Each function returns an
// ISOEXT_* status; isoext_last_error() holds the message of the last failure
// on the calling thread.  Declarations and the reference interfaces they
// replace: include/isoext_amd.h.

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/isoext_amd.h"
#include "common.h"
#include "fixup.h"
#include "its_dc.h"
#include "sparse.h"
#include "backward.h"
#include "mc_dense.h"
#include "csg.h"

namespace {

thread_local std::string g_err;

// Optional HIP-event timing of the main kernel of the last call.
bool g_timing = false;
thread_local float g_kernel_ms = -1.0f;
thread_local iso::McTimes g_times = {-1.0f, -1.0f, -1.0f, -1.0f, -1.0f};

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(ISOEXT_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_CHECK(x, what)                              \
    do {                                                \
        hipError_t e_ = (x);                            \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// The aabb must be finite: with a NaN or infinite box no grid coordinate
// equals another, which the position-keyed weld passes rely on.
int check_aabb(const float *amin, const float *amax) {
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(amin[a]) || !std::isfinite(amax[a]))
            return fail(ISOEXT_E_ARG, "aabb_min and aabb_max must be finite");
    return ISOEXT_OK;
}

#define AABB_CHECK(amin, amax)                          \
    do {                                                \
        const int rc_ = check_aabb((amin), (amax));     \
        if (rc_ != ISOEXT_OK) return rc_;               \
    } while (0)

// Workspace layout for dense marching cubes.
struct McLayout {
    int tx, ty, nzc;
    int64_t W, S, NI, NJ, NK, nblk;
    size_t o_bits, o_cnt_v, o_cnt_f, o_bsum_v, o_bsum_f, o_blk_v, o_blk_f;
    size_t o_ctr, o_scand, o_sscell, o_skept, o_shole;
    size_t o_cand, o_scell, o_kept, o_hole, o_tflag, o_tlist, o_tcnt, total;
    int64_t cap, seg_cap[4];
    // read-once pipeline (tile configuration iso::kStrTileConfig): regions,
    // look-back words, region-major offsets, wave bases
    bool stream;
    int64_t NRX, NRY, NR;
    size_t o_vofs, o_fofs, o_wcv, o_wcf, o_wbv, o_wbf, o_csum, o_cofs, o_actr, o_aoff, o_wover, o_alist, o_acnt;
    size_t o_dslots;
    int64_t dslots;   // candidate position table (a power of two, >= 2 x the candidate capacity)
};

constexpr size_t kHeader = 4096;   // err, totals, SpecialLists
constexpr size_t kOffErr = 8, kOffTotals = 16, kOffSp = 64;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

bool mc_layout(const int64_t shape[3], int cfg, int64_t cap, McLayout &L) {
    if (!iso::mc_config_geom(cfg, L.tx, L.ty, L.nzc)) return false;
    const int64_t X = shape[0], Y = shape[1], Z = shape[2];
    const int64_t R = X * Y;
    L.W = (Z + 63) / 64;
    L.S = (L.W + L.nzc - 1) / L.nzc;
    const int64_t cell_chunks = (Z - 1 + 63) / 64;
    L.NI = (X - 1 + L.tx - 1) / L.tx;
    L.NJ = (Y - 1 + L.ty - 1) / L.ty;
    L.NK = (cell_chunks + L.nzc - 1) / L.nzc;
    L.nblk = (R * L.S + iso::kMcScanBlock - 1) / iso::kMcScanBlock;
    L.cap = cap;
    size_t o = kHeader;
    auto take = [&](size_t n) {
        size_t r = o;
        o = align256(o + n);
        return r;
    };
    // the read-once pipeline serves tile configuration kStrTileConfig; its
    // entries are region-major (whole regions, so at least R * W)
    L.stream = cfg == iso::kStrTileConfig;
    L.NRX = (X + iso::kStrRX - 1) / iso::kStrRX;
    L.NRY = (Y + iso::kStrRY - 1) / iso::kStrRY;
    L.NR = L.NRX * L.NRY * L.W;
    const int64_t nent = L.stream ? L.NR * iso::kStrWaves * iso::kStrRX * iso::kStrRows : 0;
    L.o_bits = take(sizeof(uint64_t) * std::max<int64_t>(R * L.W, nent));
    L.o_cnt_v = take(sizeof(uint32_t) * R * L.S);
    L.o_cnt_f = take(sizeof(uint32_t) * R * L.S);
    L.o_bsum_v = take(sizeof(uint32_t) * L.nblk);
    L.o_bsum_f = take(sizeof(uint32_t) * L.nblk);
    L.o_blk_v = take(sizeof(int64_t) * L.nblk);
    L.o_blk_f = take(sizeof(int64_t) * L.nblk);
    for (int k = 0; k < 4; ++k) L.seg_cap[k] = iso::sp_seg_cap(cap, k, iso::kSpSeg);
    L.o_ctr = take(sizeof(unsigned long long) * 4 * iso::kSpSeg * iso::kSpStride);
    L.o_scand = take(sizeof(iso::SpecialVertex) * iso::kSpSeg * L.seg_cap[0]);
    L.o_sscell = take(sizeof(iso::SpecialEdge) * iso::kSpSeg * L.seg_cap[1]);
    L.o_skept = take(sizeof(int64_t) * iso::kSpSeg * L.seg_cap[2]);
    L.o_shole = take(sizeof(int64_t) * iso::kSpSeg * L.seg_cap[3]);
    L.o_cand = take(sizeof(iso::SpecialVertex) * iso::kSpSeg * L.seg_cap[0]);
    L.o_scell = take(sizeof(iso::SpecialEdge) * iso::kSpSeg * L.seg_cap[1]);
    L.o_kept = take(sizeof(int64_t) * iso::kSpSeg * L.seg_cap[2]);
    L.o_hole = take(sizeof(int64_t) * iso::kSpSeg * L.seg_cap[3]);
    L.o_tflag = take(sizeof(int32_t) * L.NI * L.NJ * L.NK);
    L.o_tlist = take(sizeof(int32_t) * L.NI * L.NJ * L.NK);
    L.o_tcnt = take(sizeof(int32_t));
    L.dslots = 1;
    while (L.dslots < 2 * iso::kSpSeg * L.seg_cap[0] + 16) L.dslots <<= 1;
    L.o_dslots = take(sizeof(int32_t) * L.dslots);
    L.o_vofs = L.o_fofs = L.o_wcv = L.o_wcf = L.o_wbv = L.o_wbf = L.o_csum = L.o_cofs = 0;
    L.o_actr = L.o_aoff = L.o_wover = L.o_alist = L.o_acnt = 0;
    if (L.stream) {
        const int64_t nw = L.NR * iso::kStrWaves;
        const int64_t nchunk = (nw + iso::kStrScanChunk - 1) / iso::kStrScanChunk;
        L.o_vofs = take(sizeof(uint32_t) * nent);
        L.o_fofs = take(sizeof(uint32_t) * nent);
        L.o_wcv = take(sizeof(uint32_t) * nw);
        L.o_wcf = take(sizeof(uint32_t) * nw);
        L.o_wbv = take(sizeof(int64_t) * nw);
        L.o_wbf = take(sizeof(int64_t) * nw);
        L.o_csum = take(sizeof(int64_t) * 2 * nchunk);
        L.o_cofs = take(sizeof(int64_t) * 2 * nchunk);
        L.o_actr = take(sizeof(unsigned long long) * iso::kStrArenaSeg * iso::kSpStride);
        L.o_aoff = take(sizeof(int32_t) * nw * iso::kStrRX);
        L.o_wover = take(sizeof(uint8_t) * nw);
        L.o_alist = take(sizeof(int64_t) * nw);
        L.o_acnt = take(sizeof(int));
    }
    L.total = o;
    return true;
}

// Grid coordinates of neighbouring points must be many ulps apart for the
// corner-proximity test to be exact; otherwise every vertex is welded by
// position.
bool grid_well_separated(const int64_t shape[3], const float *amin, const float *amax) {
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 2) continue;
        float prev = iso::grid_coord(0, shape[a], amin[a], amax[a]);
        for (int64_t i = 1; i < shape[a]; ++i) {
            float cur = iso::grid_coord(i, shape[a], amin[a], amax[a]);
            int64_t d = (int64_t) iso::ordered_key(cur) - (int64_t) iso::ordered_key(prev);
            if (d < 0) d = -d;
            if (d < 256) return false;
            prev = cur;
        }
    }
    return true;
}

// Bound on the interpolation parameter t of a vertex that lies within 16
// ulps of an edge endpoint, with a 2x margin: |r - a| <= 32 ulp(M) near an
// endpoint, |r - (a + t h)| <= 4 ulp(M) from the rounding of the lerp (M the
// largest coordinate magnitude of the axis, h its smallest spacing), so
// t <= 36 ulp(M) / h, and 1 - t is known to within 2^-22 from the operands
// the kernel compares.  Infinite when an axis has no positive spacing.
float weld_cand_eps(const int64_t shape[3], const float *amin, const float *amax) {
    float eps = 0.0f;
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 2) continue;
        float m = std::max(std::fabs(amin[a]), std::fabs(amax[a])), hmin = INFINITY;
        float prev = iso::grid_coord(0, shape[a], amin[a], amax[a]);
        for (int64_t i = 1; i < shape[a]; ++i) {
            const float cur = iso::grid_coord(i, shape[a], amin[a], amax[a]);
            hmin = std::min(hmin, cur - prev);
            m = std::max(m, std::fabs(cur));
            prev = cur;
        }
        if (!(hmin > 0.0f) || !std::isfinite(m)) return INFINITY;
        const float ulp = std::nextafter(m, INFINITY) - m;
        eps = std::max(eps, 72.0f * ulp / hmin + 0x1p-20f);
    }
    return eps;
}

// Weld parameters of a grid geometry (corner separation, candidate bound),
// cached for the last geometry seen by the calling thread.
struct WeldGeom {
    int all_special;
    float cand_eps;
};
WeldGeom weld_geom(const int64_t shape[3], const float *amin, const float *amax) {
    thread_local bool valid = false;
    thread_local int64_t c_shape[3];
    thread_local float c_min[3], c_max[3];
    thread_local WeldGeom c_geom;
    bool hit = valid;
    for (int a = 0; a < 3 && hit; ++a)
        hit = c_shape[a] == shape[a] && c_min[a] == amin[a] && c_max[a] == amax[a];
    if (!hit) {
        c_geom.all_special = grid_well_separated(shape, amin, amax) ? 0 : 1;
        c_geom.cand_eps = weld_cand_eps(shape, amin, amax);
        for (int a = 0; a < 3; ++a) {
            c_shape[a] = shape[a];
            c_min[a] = amin[a];
            c_max[a] = amax[a];
        }
        valid = true;
    }
    return c_geom;
}

// Mask-pipeline parameters of the intersection pass on a dense grid (the
// scratch holds the masks, count arrays and totals; cap_* bound the writes).
iso::McParams its_params(const isoext_grid *g, float level, char *ws, int64_t cap_cells, int64_t cap_pts) {
    McLayout L;
    mc_layout(g->shape, iso::mc_auto_config(g->shape), 1, L);
    iso::McParams p{};
    p.values = g->values;
    p.X = g->shape[0];
    p.Y = g->shape[1];
    p.Z = g->shape[2];
    for (int a = 0; a < 3; ++a) {
        p.amin[a] = g->aabb_min[a];
        p.amax[a] = g->aabb_max[a];
    }
    p.level = level;
    p.its = 1;
    p.vec4 = (p.Z % 4 == 0 && ((uintptr_t) g->values & 15) == 0) ? 1 : 0;
    p.W = L.W;
    p.nzc = L.nzc;
    p.S = L.S;
    p.NI = L.NI;
    p.NJ = L.NJ;
    p.NK = L.NK;
    p.bits = (uint64_t *) (ws + L.o_bits);
    p.cnt_v = (uint32_t *) (ws + L.o_cnt_v);
    p.cnt_f = (uint32_t *) (ws + L.o_cnt_f);
    p.bsum_v = (uint32_t *) (ws + L.o_bsum_v);
    p.bsum_f = (uint32_t *) (ws + L.o_bsum_f);
    p.blk_v = (int64_t *) (ws + L.o_blk_v);
    p.blk_f = (int64_t *) (ws + L.o_blk_f);
    p.nblk = L.nblk;
    p.cap_v = cap_cells;
    p.cap_f = cap_pts;
    p.err = (int *) (ws + kOffErr);
    p.totals = (int64_t *) (ws + kOffTotals);
    // tiles without an active cell are skipped: their lookup-table entries
    // are never read (only cells next to a crossing edge are looked up, and
    // those are active)
    p.tile_flag = (int *) (ws + L.o_tflag);
    p.tile_list = (int *) (ws + L.o_tlist);
    p.tile_count = (int *) (ws + L.o_tcnt);
    p.tx = L.tx;
    p.ty = L.ty;
    return p;
}

iso::GridDesc to_desc(const isoext_grid *g) {
    iso::GridDesc d;
    d.values = g->values;
    d.sparse_ids = g->sparse_cell_ids;
    d.n_sparse = g->n_sparse;
    d.X = g->shape[0];
    d.Y = g->shape[1];
    d.Z = g->shape[2];
    for (int a = 0; a < 3; ++a) {
        d.amin[a] = g->aabb_min[a];
        d.amax[a] = g->aabb_max[a];
    }
    return d;
}

iso::ItsIn to_in(const isoext_its *t) {
    iso::ItsIn in;
    in.ncells = t->ncells;
    in.npts = t->npts;
    in.cell_indices = t->cell_indices;
    in.cell_dense = t->cell_dense;
    in.cell_offsets = t->cell_offsets;
    in.points = t->points;
    in.normals = t->normals;
    in.edge_local = t->edge_local;
    in.is_out = t->is_out;
    in.entry_cell = t->entry_cell;
    in.amask_tab = t->amask_tab;
    in.apre_tab = t->apre_tab;
    return in;
}

bool grid_has_cells(const isoext_grid *g) {
    if (g->shape[0] < 2 || g->shape[1] < 2 || g->shape[2] < 2) return false;
    if (g->sparse_cell_ids) return g->n_sparse > 0;
    return true;
}

struct DcLayout {
    size_t o_dual, o_cond, o_cnt, o_used, o_bnd, o_qn, o_qm, o_bsum, o_part, o_tot, o_tri, o_hs, o_hm, o_rep, o_fin,
        o_ak, o_av, total;
    int64_t hslots;
};

DcLayout dc_layout(int64_t ncells, int64_t npts) {
    DcLayout L;
    int64_t nb = (std::max(ncells, npts) + 255) / 256 + 1;
    int64_t hs = 1;
    while (hs < 2 * ncells + 16) hs <<= 1;
    L.hslots = hs;
    size_t o = 0;
    auto take = [&](size_t n) {
        size_t r = o;
        o = align256(o + n);
        return r;
    };
    L.o_dual = take(sizeof(float) * 3 * std::max<int64_t>(ncells, 1));
    L.o_cond = take(sizeof(float) * std::max<int64_t>(ncells, 1));
    L.o_cnt = take(64);
    L.o_used = take(std::max<int64_t>(ncells, 1));
    L.o_bnd = take(std::max<int64_t>(ncells, 1));
    L.o_qn = take(std::max<int64_t>(ncells, 1));
    L.o_qm = take(std::max<int64_t>(ncells, 1));
    L.o_bsum = take(sizeof(int64_t) * 2 * nb);
    L.o_part = take(sizeof(int64_t) * 2 * iso::dc_scan_part_pairs(nb));
    L.o_tot = take(64);
    L.o_tri = take(sizeof(int32_t) * 6 * std::max<int64_t>(npts, 1));
    L.o_hs = take(sizeof(int32_t) * hs);
    L.o_hm = take(sizeof(int32_t) * hs);
    L.o_rep = take(sizeof(int32_t) * std::max<int64_t>(ncells, 1));
    L.o_fin = take(sizeof(int32_t) * std::max<int64_t>(ncells, 1));
    L.o_ak = take(sizeof(int64_t) * hs);
    L.o_av = take(sizeof(int32_t) * hs);
    L.total = o;
    return L;
}


// ---- ragged batches of dense marching cubes --------------------------------------

constexpr int kBatchSeg = iso::kMaxSpSegBatch;   // special-list segments per batch item

struct ItemLayout {
    bool cells;
    int cfg, nzc;   // the item's emit tile configuration
    int64_t W, S, NI, NJ, NK, nblk, nctask, ngtask, ntiles, tile0, seg_cap[4];
    size_t o_hdr, o_bits, o_cnt_v, o_cnt_f, o_bsum_v, o_bsum_f, o_blk_v, o_blk_f;
    size_t o_ctr, o_scand, o_sscell, o_skept, o_shole, o_cand, o_scell, o_kept, o_hole;
};


struct BatchLayout {
    size_t o_hdr, o_items, o_order, o_summary, o_fixcnt, o_tflag, o_tlist, o_tcnt, total;
    std::vector<ItemLayout> it;
    iso::McBatchHdr h;
    // emit tiles are laid out by configuration: items of configuration c own
    // tiles [tile_begin[c], tile_end[c]); order lists the items by tile start
    std::vector<int64_t> tile_begin, tile_end;
    std::vector<int32_t> order;
};

// cfg >= 0: every item uses that tile configuration; cfg < 0: each item its
// own (iso::mc_auto_config of its shape).
bool batch_layout(int n, const isoext_mc_item *items, int cfg, int64_t special_cap, BatchLayout &B) {
    int tx, ty, nzc;
    if (cfg >= 0 && !iso::mc_config_geom(cfg, tx, ty, nzc)) return false;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t r = o;
        o = align256(o + bytes);
        return r;
    };
    B.o_hdr = take(sizeof(iso::McBatchHdr));
    B.o_items = take(sizeof(iso::McBatchItem) * (size_t) n);
    B.o_order = take(sizeof(int32_t) * (size_t) n);
    B.o_summary = take(sizeof(iso::McBatchSummary) * (size_t) n);
    B.o_fixcnt = take(sizeof(int64_t) * 4 * (size_t) n);
    B.it.assign(n, ItemLayout{});
    B.h = iso::McBatchHdr{n, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        ItemLayout &L = B.it[i];
        const int64_t X = items[i].shape[0], Y = items[i].shape[1], Z = items[i].shape[2];
        L.cfg = items[i].sdf ? iso::kProgramTileConfig : (cfg >= 0 ? cfg : iso::mc_auto_config(items[i].shape));
        iso::mc_config_geom(L.cfg, tx, ty, nzc);
        L.nzc = nzc;
        L.cells = X >= 2 && Y >= 2 && Z >= 2;
        L.W = L.cells ? (Z + 63) / 64 : 0;
        L.S = L.cells ? (L.W + nzc - 1) / nzc : 0;
        L.NI = L.cells ? (X - 1 + tx - 1) / tx : 0;
        L.NJ = L.cells ? (Y - 1 + ty - 1) / ty : 0;
        L.NK = L.cells ? ((Z - 1 + 63) / 64 + nzc - 1) / nzc : 0;
        L.nblk = L.cells ? (X * Y * L.S + iso::kMcScanBlock - 1) / iso::kMcScanBlock : 0;
        L.nctask = L.cells && !items[i].sdf ? (X * Y * Z + iso::kFlatPts - 1) / iso::kFlatPts : 0;
        // program blocks, padded to whole 256-block chunks (one item per workgroup)
        L.ngtask = L.cells && items[i].sdf ? (iso::csg_classify_tasks(X, Y, L.W) + 255) / 256 * 256 : 0;
        L.ntiles = L.NI * L.NJ * L.NK;
        B.h.nctask += L.nctask;
        B.h.ngtask += L.ngtask;
        if (items[i].sdf) ++B.h.n_prog;
        B.h.nblk += L.nblk;
        B.h.ntiles += L.ntiles;
        const int64_t R = L.cells ? X * Y : 0;
        for (int k = 0; k < 4; ++k) L.seg_cap[k] = iso::sp_seg_cap(special_cap, k, kBatchSeg);
        L.o_hdr = take(kHeader);
        L.o_bits = take(sizeof(uint64_t) * R * L.W);
        L.o_cnt_v = take(sizeof(uint32_t) * R * L.S);
        L.o_cnt_f = take(sizeof(uint32_t) * R * L.S);
        L.o_bsum_v = take(sizeof(uint32_t) * std::max<int64_t>(L.nblk, 1));
        L.o_bsum_f = take(sizeof(uint32_t) * std::max<int64_t>(L.nblk, 1));
        L.o_blk_v = take(sizeof(int64_t) * std::max<int64_t>(L.nblk, 1));
        L.o_blk_f = take(sizeof(int64_t) * std::max<int64_t>(L.nblk, 1));
        L.o_ctr = take(sizeof(unsigned long long) * 4 * kBatchSeg * iso::kSpStride);
        L.o_scand = take(sizeof(iso::SpecialVertex) * kBatchSeg * L.seg_cap[0]);
        L.o_sscell = take(sizeof(iso::SpecialEdge) * kBatchSeg * L.seg_cap[1]);
        L.o_skept = take(sizeof(int64_t) * kBatchSeg * L.seg_cap[2]);
        L.o_shole = take(sizeof(int64_t) * kBatchSeg * L.seg_cap[3]);
        L.o_cand = take(sizeof(iso::SpecialVertex) * kBatchSeg * L.seg_cap[0]);
        L.o_scell = take(sizeof(iso::SpecialEdge) * kBatchSeg * L.seg_cap[1]);
        L.o_kept = take(sizeof(int64_t) * kBatchSeg * L.seg_cap[2]);
        L.o_hole = take(sizeof(int64_t) * kBatchSeg * L.seg_cap[3]);
    }
    const int ncfg = iso::mc_config_count();
    B.tile_begin.assign(ncfg, 0);
    B.tile_end.assign(ncfg, 0);
    B.order.clear();
    int64_t t = 0;
    for (int c = 0; c < ncfg; ++c) {
        B.tile_begin[c] = t;
        for (int i = 0; i < n; ++i) {
            if (B.it[i].cfg != c) continue;
            B.it[i].tile0 = t;
            t += B.it[i].ntiles;
            B.order.push_back(i);
        }
        B.tile_end[c] = t;
    }
    B.o_tflag = take(sizeof(int32_t) * (size_t) std::max<int64_t>(B.h.ntiles, 1));
    B.o_tlist = take(sizeof(int32_t) * (size_t) std::max<int64_t>(B.h.ntiles, 1));
    B.o_tcnt = take(sizeof(int32_t) * iso::kMaxTileConfigs);
    B.total = o;
    return true;
}

// Marching-cubes batches emit only the tiles the count pass listed as
// active, when every grid's chunk indices fit 32 bits and the batch's tile
// ids fit an int (the same decision in the count and the emit call).
bool batch_tile_list(int n, const isoext_mc_item *items, const BatchLayout &B) {
    if (B.h.ntiles > (int64_t) INT32_MAX) return false;
    for (int i = 0; i < n; ++i)
        if (B.it[i].cells && items[i].shape[0] * items[i].shape[1] * B.it[i].W > (int64_t) UINT32_MAX) return false;
    return true;
}

// Pinned staging for the descriptor upload and the summary read-back,
// grown on demand, one per host thread (reused after each phase's sync).
struct Pinned {
    void *ptr = nullptr;
    size_t bytes = 0;
    ~Pinned() {
        if (ptr) (void) hipHostFree(ptr);
    }
    void *get(size_t want) {
        if (want > bytes) {
            if (ptr) (void) hipHostFree(ptr);
            ptr = nullptr;
            bytes = 0;
            if (hipHostMalloc(&ptr, want, 0) != hipSuccess) return nullptr;
            bytes = want;
        }
        return ptr;
    }
};
thread_local Pinned g_pinned;
thread_local Pinned g_pinned_rb;   // single-call read-back (counts, list descriptors)
thread_local Pinned g_pinned_prog;   // SDF program upload

// Stage timing events of the calling thread, created once per device.
hipEvent_t *timing_events() {
    thread_local hipEvent_t ev[5] = {};
    thread_local int dev = -1;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return nullptr;
    if (dev != cur) {
        for (auto &e : ev) {
            if (e) (void) hipEventDestroy(e);
            e = nullptr;
        }
        for (auto &e : ev)
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
        dev = cur;
    }
    return ev;
}

void fill_item(const isoext_mc_item &in, const ItemLayout &L, char *ws, int64_t ctask0, int64_t blk0,
               int64_t tile0, int64_t gtask0, iso::McBatchItem &out) {
    iso::McParams &p = out.p;
    p = iso::McParams{};
    p.values = in.values;
    p.X = in.shape[0];
    p.Y = in.shape[1];
    p.Z = in.shape[2];
    for (int a = 0; a < 3; ++a) {
        p.amin[a] = in.aabb_min[a];
        p.amax[a] = in.aabb_max[a];
    }
    p.level = in.level;
    p.method = in.method;
    const WeldGeom wg = weld_geom(in.shape, in.aabb_min, in.aabb_max);
    p.all_special = L.cells ? wg.all_special : 0;
    p.cand_eps = wg.cand_eps;
    p.vec4 = (in.shape[2] % 4 == 0 && ((uintptr_t) in.values & 15) == 0) ? 1 : 0;
    p.W = L.W;
    p.S = L.S;
    p.NI = L.NI;
    p.NJ = L.NJ;
    p.NK = L.NK;
    p.nzc = 1;   // set by the caller from the tile config
    p.bits = (uint64_t *) (ws + L.o_bits);
    p.cnt_v = (uint32_t *) (ws + L.o_cnt_v);
    p.cnt_f = (uint32_t *) (ws + L.o_cnt_f);
    p.bsum_v = (uint32_t *) (ws + L.o_bsum_v);
    p.bsum_f = (uint32_t *) (ws + L.o_bsum_f);
    p.blk_v = (int64_t *) (ws + L.o_blk_v);
    p.blk_f = (int64_t *) (ws + L.o_blk_f);
    p.nblk = L.nblk;
    p.out_v = in.out_v;
    p.cap_v = in.out_v ? in.cap_v : 0;
    p.out_key = nullptr;
    p.out_f = in.out_f;
    p.cap_f = in.out_f ? in.cap_f : 0;
    p.err = (int *) (ws + L.o_hdr + kOffErr);
    p.totals = (int64_t *) (ws + L.o_hdr + kOffTotals);
    p.sp = (iso::SpecialLists *) (ws + L.o_hdr + kOffSp);
    p.seg.ctr = (unsigned long long *) (ws + L.o_ctr);
    for (int k = 0; k < 4; ++k) p.seg.cap[k] = L.seg_cap[k];
    p.seg.nseg = kBatchSeg;
    p.seg.cand = (iso::SpecialVertex *) (ws + L.o_scand);
    p.seg.scell = (iso::SpecialEdge *) (ws + L.o_sscell);
    p.seg.kept = (int64_t *) (ws + L.o_skept);
    p.seg.hole = (int64_t *) (ws + L.o_shole);
    iso::SpecialLists &sp = out.sp_init;
    sp = iso::SpecialLists{};
    sp.cap_cand = kBatchSeg * L.seg_cap[0];
    sp.cap_scell = kBatchSeg * L.seg_cap[1];
    sp.cap_kept = kBatchSeg * L.seg_cap[2];
    sp.cap_hole = kBatchSeg * L.seg_cap[3];
    sp.cand = (iso::SpecialVertex *) (ws + L.o_cand);
    sp.scell = (iso::SpecialEdge *) (ws + L.o_scell);
    sp.kept = (int64_t *) (ws + L.o_kept);
    sp.hole = (int64_t *) (ws + L.o_hole);
    if (in.sdf) {
        p.values = nullptr;
        p.vec4 = 0;
        p.csg = (const iso::CsgProgram *) in.sdf;
        p.csg_list = iso::csg_list_fits(p.X, p.Y, p.W, p.S) ? 1 : 0;
    }
    out.ctask0 = ctask0;
    out.blk0 = blk0;
    out.tile0 = tile0;
    out.gtask0 = gtask0;
    out.mode = p.vec4 ? (p.Z % 64 == 0 ? 0 : 1) : 2;
}

// Upload the descriptors of a batch phase (pinned staging) and return the
// staging area for the summaries.
// its_caps != nullptr: the intersection pass (p.its = 1) with per-item
// capacities (cells, entries) its_caps[2i], its_caps[2i + 1].
int batch_upload(int n, const isoext_mc_item *items, char *ws, const BatchLayout &B, hipStream_t st,
                 iso::McBatchSummary **summary_host, const int64_t *its_caps = nullptr) {
    const size_t desc = sizeof(iso::McBatchHdr) + sizeof(iso::McBatchItem) * (size_t) n + sizeof(int32_t) * (size_t) n;
    const size_t sum = sizeof(iso::McBatchSummary) * (size_t) n;
    char *host = (char *) g_pinned.get(desc + sum + 64);
    if (!host) return fail(ISOEXT_E_DEVICE, "pinned staging allocation failed");
    iso::McBatchHdr *h = (iso::McBatchHdr *) host;
    *h = B.h;
    const bool listed = batch_tile_list(n, items, B);
    if (listed) {
        h->tile_flag = (int *) (ws + B.o_tflag);
        h->tile_count = (int *) (ws + B.o_tcnt);
    }
    iso::McBatchItem *it = (iso::McBatchItem *) (host + sizeof(iso::McBatchHdr));
    int64_t w0 = 0, b0 = 0, g0 = 0;
    for (int i = 0; i < n; ++i) {
        fill_item(items[i], B.it[i], ws, w0, b0, B.it[i].tile0, g0, it[i]);
        it[i].p.nzc = B.it[i].nzc;
        if (listed) {
            iso::McParams &p = it[i].p;
            int tx, ty, nzc;
            iso::mc_config_geom(B.it[i].cfg, tx, ty, nzc);
            p.tile_flag = h->tile_flag;
            p.tile_list = (int *) (ws + B.o_tlist) + B.tile_begin[B.it[i].cfg];
            p.tile_count = h->tile_count + B.it[i].cfg;
            p.tile_base = (int) B.it[i].tile0;
            p.tx = tx;
            p.ty = ty;
        }
        iso::mc_params_divisors(it[i].p);
        if (listed && B.it[i].cells && !it[i].p.tile_list)
            return fail(ISOEXT_E_DEVICE, "batch item %d: active-tile list unavailable", i);
        if (its_caps) {
            iso::McParams &p = it[i].p;
            p.its = 1;
            p.out_v = nullptr;
            p.out_f = nullptr;
            p.cap_v = its_caps[2 * i];
            p.cap_f = its_caps[2 * i + 1];
        }
        w0 += B.it[i].nctask;
        b0 += B.it[i].nblk;
        g0 += B.it[i].ngtask;
    }
    int32_t *order = (int32_t *) (it + n);
    for (int k = 0; k < n; ++k) order[k] = B.order[k];
    HIP_CHECK(hipMemcpyAsync(ws + B.o_hdr, h, sizeof(iso::McBatchHdr), hipMemcpyHostToDevice, st), "upload batch");
    HIP_CHECK(hipMemcpyAsync(ws + B.o_items, it, sizeof(iso::McBatchItem) * (size_t) n, hipMemcpyHostToDevice, st),
              "upload batch");
    HIP_CHECK(hipMemcpyAsync(ws + B.o_order, order, sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice, st),
              "upload batch");
    *summary_host = (iso::McBatchSummary *) (host + ((desc + 63) & ~size_t(63)));
    return ISOEXT_OK;
}

int batch_check(int n, const isoext_mc_item *items, int cfg) {
    int tx, ty, nzc;
    if (n < 1) return fail(ISOEXT_E_ARG, "empty batch");
    if (n > iso::kMaxBatch) return fail(ISOEXT_E_ARG, "batch of %d grids exceeds %d", n, iso::kMaxBatch);
    if (cfg >= 0 && !iso::mc_config_geom(cfg, tx, ty, nzc)) return fail(ISOEXT_E_ARG, "unknown tile config %d", cfg);
    for (int i = 0; i < n; ++i) {
        if (items[i].method != 0 && items[i].method != 1)
            return fail(ISOEXT_E_ARG, "Unknown method: %d", items[i].method);
        for (int a = 0; a < 3; ++a)
            if (items[i].shape[a] < 1) return fail(ISOEXT_E_ARG, "grid shape must be positive");
        if (!items[i].values && !items[i].sdf) return fail(ISOEXT_E_ARG, "values is null");
        if (items[i].sdf && ((uintptr_t) items[i].sdf & 15))
            return fail(ISOEXT_E_ARG, "item %d: SDF program must be 16-byte aligned", i);
        AABB_CHECK(items[i].aabb_min, items[i].aabb_max);
    }
    return ISOEXT_OK;
}

// Scratch of a batched dual contouring run: the grids' active cells (C in
// all; no intersection entries are stored: each cell's edges are recomputed
// from its corner values), their lookup tables, the per-grid descriptors and
// counts, and the DC work arrays for C cells.
struct DcBatchLayout {
    DcLayout dc;
    size_t o_dc, o_cd, o_outs, o_items, o_citem, o_vb, o_qb, o_tm, o_tp, total;
    std::vector<int64_t> tab0;   // per grid: first lookup-table entry
};

// Lookup-table entries of a dense grid: (X-1)(Y-1) cell rows x ceil(Z/64) words.
int64_t dc_table_entries(const isoext_mc_item &it) {
    const int64_t X = it.shape[0], Y = it.shape[1], Z = it.shape[2];
    return (X >= 2 && Y >= 2 && Z >= 2) ? (X - 1) * (Y - 1) * ((Z + 63) / 64) : 0;
}

DcBatchLayout dc_batch_layout(int n, const isoext_mc_item *items, int64_t C, int64_t P) {
    DcBatchLayout L;
    L.tab0.assign(n, 0);
    int64_t T = 0;
    for (int i = 0; i < n; ++i) {
        L.tab0[i] = T;
        T += dc_table_entries(items[i]);
    }
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t r = o;
        o = align256(o + bytes);
        return r;
    };
    (void) P;
    const size_t c1 = (size_t) std::max<int64_t>(C, 1);
    L.o_cd = take(sizeof(int64_t) * c1);
    L.o_outs = take(sizeof(iso::ItsOut) * (size_t) n);
    L.o_items = take(iso::dc_batch_item_bytes() * (size_t) n);
    L.o_citem = take(sizeof(int32_t) * c1);
    L.o_vb = take(sizeof(int64_t) * (size_t) (n + 1));
    L.o_qb = take(sizeof(int64_t) * (size_t) (n + 1));
    L.o_tm = take(sizeof(uint64_t) * (size_t) std::max<int64_t>(T, 1));
    L.o_tp = take(sizeof(int32_t) * (size_t) std::max<int64_t>(T, 1));
    L.dc = dc_layout(C, 0);
    L.o_dc = take(L.dc.total);
    L.total = o;
    return L;
}

thread_local Pinned g_pinned_its;   // batched DC: intersection output descriptors

}   // namespace

extern "C" {

int64_t isoext_its_scratch_bytes(const isoext_grid *g) {
    if (!grid_has_cells(g)) return 64;
    if (!g->sparse_cell_ids) {
        McLayout L;
        mc_layout(g->shape, iso::mc_auto_config(g->shape), 1, L);
        return (int64_t) L.total;
    }
    iso::GridDesc d = to_desc(g);
    return (int64_t) (sizeof(int64_t) * 2 * (iso::its_scratch_pairs(d) + 4));
}

int isoext_its_count(const isoext_grid *g, float level, void *scratch, int64_t scratch_bytes, int64_t counts[2],
                     void *stream) {
    counts[0] = counts[1] = 0;
    AABB_CHECK(g->aabb_min, g->aabb_max);
    if (!grid_has_cells(g)) return ISOEXT_OK;
    if (scratch_bytes < isoext_its_scratch_bytes(g)) return fail(ISOEXT_E_WORKSPACE, "intersection scratch too small");
    hipStream_t st = (hipStream_t) stream;
    int64_t h[2];
    if (!g->sparse_cell_ids) {
        iso::McParams p = its_params(g, level, (char *) scratch, 0, 0);
        HIP_CHECK(iso::its_dense_count_launch(p, st), "intersection count");
        HIP_CHECK(hipMemcpyAsync(h, p.totals, sizeof(h), hipMemcpyDeviceToHost, st), "read counts");
    } else {
        iso::GridDesc d = to_desc(g);
        int64_t nb = iso::its_scratch_pairs(d);
        int64_t *bsum = (int64_t *) scratch;
        int64_t *tot = bsum + 2 * nb;
        HIP_CHECK(iso::its_count(d, level, bsum, tot, st), "intersection count");
        HIP_CHECK(hipMemcpyAsync(h, tot, sizeof(h), hipMemcpyDeviceToHost, st), "read counts");
    }
    HIP_CHECK(hipStreamSynchronize(st), "intersection count");
    counts[0] = h[0];
    counts[1] = h[1];
    return ISOEXT_OK;
}

int64_t isoext_its_table_entries(const int64_t shape[3]) {
    if (shape[0] < 2 || shape[1] < 2 || shape[2] < 2) return 0;
    return (shape[0] - 1) * (shape[1] - 1) * ((shape[2] + 63) / 64);
}

int isoext_its_emit(const isoext_grid *g, float level, const void *scratch, int64_t ncells, int64_t npts,
                    int64_t *cell_indices, int64_t *cell_dense, int64_t *cell_offsets, float *points, float *normals,
                    uint8_t *edge_local, uint8_t *is_out, int32_t *entry_cell, uint64_t *amask_tab,
                    int32_t *apre_tab, void *stream) {
    hipStream_t st = (hipStream_t) stream;
    if (grid_has_cells(g) && ncells > 0) {
        iso::ItsOut o{cell_indices, cell_dense, cell_offsets, points, normals, edge_local, is_out, entry_cell};
        if (!g->sparse_cell_ids && amask_tab && apre_tab) {
            if (ncells > 0x7fffffffll) return fail(ISOEXT_E_TOO_LARGE, "more than 2^31-1 active cells");
            o.amask_tab = amask_tab;
            o.apre_tab = apre_tab;
        }
        if (!g->sparse_cell_ids) {
            iso::McParams p = its_params(g, level, (char *) scratch, ncells, npts);
            HIP_CHECK(iso::its_dense_emit_launch(iso::mc_auto_config(g->shape), p, o, st), "intersection emit");
        } else {
            HIP_CHECK(iso::its_emit(to_desc(g), level, (const int64_t *) scratch, o, st), "intersection emit");
        }
    }
    HIP_CHECK(iso::fill_i64(cell_offsets + ncells, 1, npts, st), "write offsets");
    return ISOEXT_OK;
}

int isoext_dc_backward(const isoext_grid *g, const isoext_its *its, float reg, float svd_tol, float level,
                       int64_t nv, const int32_t *vcell, const float *grad_v, float *grad_values,
                       float *grad_normals, void *stream) {
    if (!grid_has_cells(g) || nv == 0) return ISOEXT_OK;
    if (!its->normals || !vcell || !grad_v || !grad_values) return fail(ISOEXT_E_ARG, "missing argument");
    HIP_CHECK(iso::dc_backward(to_desc(g), to_in(its), reg, svd_tol, level, nv, vcell, grad_v, grad_values,
                               grad_normals, (hipStream_t) stream),
              "dual contouring backward");
    return ISOEXT_OK;
}

int isoext_its_normals(const isoext_grid *g, const isoext_its *its, float *normals, void *stream) {
    if (!grid_has_cells(g) || its->npts == 0) return ISOEXT_OK;
    HIP_CHECK(iso::its_normals(to_desc(g), to_in(its), normals, (hipStream_t) stream), "normals");
    return ISOEXT_OK;
}

int64_t isoext_dc_workspace_bytes(int64_t ncells, int64_t npts) { return (int64_t) dc_layout(ncells, npts).total; }

int isoext_dc(const isoext_grid *g, const isoext_its *its, float reg, float svd_tol, void *workspace,
              int64_t workspace_bytes, float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f, float *dual_v,
              int32_t *out_vcell, int64_t counts[4], void *stream) {
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    AABB_CHECK(g->aabb_min, g->aabb_max);
    if (!grid_has_cells(g) || its->ncells == 0) return ISOEXT_OK;
    if (!its->normals) return fail(ISOEXT_E_ARG, "intersection has no normals");
    DcLayout L = dc_layout(its->ncells, its->npts);
    if ((int64_t) L.total > workspace_bytes) return fail(ISOEXT_E_WORKSPACE, "dual contouring workspace too small");
    if (cap_v < its->ncells || cap_f < 2 * its->npts) return fail(ISOEXT_E_CAPACITY, "output capacity too small");
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace;
    iso::DcScratch w;
    w.dual_v = dual_v ? dual_v : (float *) (ws + L.o_dual);
    w.cond = (float *) (ws + L.o_cond);
    w.n_fp64 = (unsigned long long *) (ws + L.o_cnt);
    w.used = (uint8_t *) (ws + L.o_used);
    w.bnd = (uint8_t *) (ws + L.o_bnd);
    w.qn = (uint8_t *) (ws + L.o_qn);
    w.qmask = (uint8_t *) (ws + L.o_qm);
    w.bsum = (int64_t *) (ws + L.o_bsum);
    w.part = (int64_t *) (ws + L.o_part);
    w.tot = (int64_t *) (ws + L.o_tot);
    w.tri_cells = (int32_t *) (ws + L.o_tri);
    w.cap_tri = 6 * std::max<int64_t>(its->npts, 1);
    w.hslots = (int32_t *) (ws + L.o_hs);
    w.slot_min = (int32_t *) (ws + L.o_hm);
    w.hmask = L.hslots - 1;
    w.akey = (int64_t *) (ws + L.o_ak);
    w.aval = (int32_t *) (ws + L.o_av);
    w.amask = L.hslots - 1;
    w.rep_slot = (int32_t *) (ws + L.o_rep);
    w.final_id = (int32_t *) (ws + L.o_fin);
    w.out_v = out_v;
    w.out_f = out_f;
    w.out_vcell = out_vcell;
    // re-solve in FP64 where the FP32 QEF is ill-conditioned, and always in the
    // regime where the reference truncates singular values (cond >= 1/svd_tol)
    float cond_fp64 = std::min(1e6f, 0.5f / std::max(svd_tol, 1e-30f));
    int64_t nv = 0, nf = 0, nfp = 0;
    HIP_CHECK(iso::dc_run(to_desc(g), to_in(its), reg, svd_tol, cond_fp64, w, nv, nf, nfp, st), "dual contouring");
    counts[0] = nv;
    counts[1] = nf;
    counts[2] = (int64_t) nfp;
    counts[3] = w.nquads;
    return ISOEXT_OK;
}


const char *isoext_last_error(void) { return g_err.c_str(); }

void isoext_set_timing(int on) { g_timing = on != 0; }

float isoext_last_kernel_ms(void) { return g_kernel_ms; }

void isoext_mc_last_stage_ms(float out[5]) {
    out[0] = g_times.classify;
    out[1] = g_times.count;
    out[2] = g_times.scan;
    out[3] = g_times.emit;
    out[4] = g_times.total;
}

int isoext_version(void) { return ISOEXT_ABI_VERSION; }

int isoext_mc_config_count(void) { return iso::mc_config_count(); }

int isoext_mc_auto_config(const int64_t shape[3]) { return iso::mc_auto_config(shape); }

int64_t isoext_mc_dense_workspace_bytes(const int64_t shape[3], int cfg, int64_t special_cap) {
    McLayout L;
    if (!mc_layout(shape, cfg, special_cap, L)) return -1;
    return (int64_t) L.total;
}

// A captured marching-cubes pipeline: the arguments of every run and the
// instantiated graph of its kernel launches (reset, classify, count, scan,
// emit, special-list gather and dup-check).  The read-back of the counts and
// the weld fix-up run outside the graph, after each replay.
struct isoext_mc_plan {
    const float *values;
    int64_t shape[3];
    float amin[3], amax[3];
    float level;
    int method, cfg;
    void *workspace;
    int64_t workspace_bytes, special_cap;
    float *out_v;
    int64_t cap_v;
    int32_t *out_f;
    int64_t cap_f;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

enum { kPlanNone = 0, kPlanCapture = 1, kPlanRun = 2 };

// Capture on a private stream (the legacy default stream cannot capture).
static hipError_t plan_capture(isoext_mc_plan *plan, int cfg, iso::McParams &p, const iso::SpecialLists &sp) {
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        const hipError_t le = iso::mc_dense_launch(cfg, p, sp, cs, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(cs, &g);
        e = le != hipSuccess ? le : ce;
        if (e == hipSuccess) {
            plan->graph = g;
            e = hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0);
        } else if (g) {
            (void) hipGraphDestroy(g);
        }
    }
    (void) hipStreamDestroy(cs);
    return e;
}

static int mc_dense_impl(const float *values, const int64_t shape[3], const float aabb_min[3],
                         const float aabb_max[3], float level, int method, int cfg, void *workspace,
                         int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v,
                         int32_t *out_f, int64_t cap_f, int64_t *out_key, int64_t counts[6], void *stream,
                         const float *spheres = nullptr, int nsph = 0, const iso::FixupScratch *na = nullptr,
                         const iso::CsgProgram *csg = nullptr, isoext_mc_plan *plan = nullptr, int plan_mode = 0);

int isoext_mc_spheres(const float *spheres, int nsph, const int64_t shape[3], const float aabb_min[3],
                      const float aabb_max[3], float level, int method, int cfg, void *workspace,
                      int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v, int32_t *out_f,
                      int64_t cap_f, int64_t counts[6], void *stream) {
    if (!spheres) return fail(ISOEXT_E_ARG, "spheres is null");
    return mc_dense_impl(nullptr, shape, aabb_min, aabb_max, level, method, cfg, workspace, workspace_bytes,
                         special_cap, out_v, cap_v, out_f, cap_f, nullptr, counts, stream, spheres, nsph);
}

int isoext_mc_sdf(const void *program, const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                  float level, int method, int cfg, void *workspace, int64_t workspace_bytes, int64_t special_cap,
                  float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f, int64_t counts[6], void *stream) {
    if (!program) return fail(ISOEXT_E_ARG, "program is null");
    return mc_dense_impl(nullptr, shape, aabb_min, aabb_max, level, method, cfg, workspace, workspace_bytes,
                         special_cap, out_v, cap_v, out_f, cap_f, nullptr, counts, stream, nullptr, 0, nullptr,
                         (const iso::CsgProgram *) program);
}

// ---- captured marching-cubes plans ----------------------------------------------------

int isoext_mc_plan_create(const float *values, const int64_t shape[3], const float aabb_min[3],
                          const float aabb_max[3], float level, int method, int cfg, void *workspace,
                          int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v, int32_t *out_f,
                          int64_t cap_f, isoext_mc_plan **plan) {
    if (!plan) return fail(ISOEXT_E_ARG, "plan is null");
    *plan = nullptr;
    if (!values) return fail(ISOEXT_E_ARG, "values is null");
    std::unique_ptr<isoext_mc_plan> pl(new isoext_mc_plan);
    pl->values = values;
    for (int a = 0; a < 3; ++a) {
        pl->shape[a] = shape[a];
        pl->amin[a] = aabb_min[a];
        pl->amax[a] = aabb_max[a];
    }
    pl->level = level;
    pl->method = method;
    pl->cfg = cfg;
    pl->workspace = workspace;
    pl->workspace_bytes = workspace_bytes;
    pl->special_cap = special_cap;
    pl->out_v = out_v;
    pl->cap_v = cap_v;
    pl->out_f = out_f;
    pl->cap_f = cap_f;
    int64_t counts[6];
    const int rc = mc_dense_impl(values, shape, aabb_min, aabb_max, level, method, cfg, workspace, workspace_bytes,
                                 special_cap, out_v, cap_v, out_f, cap_f, nullptr, counts, nullptr, nullptr, 0,
                                 nullptr, nullptr, pl.get(), kPlanCapture);
    if (rc != ISOEXT_OK) return rc;
    *plan = pl.release();
    return ISOEXT_OK;
}

int isoext_mc_plan_run(isoext_mc_plan *pl, int64_t counts[6], void *stream) {
    if (!pl) return fail(ISOEXT_E_ARG, "plan is null");
    for (int i = 0; i < 6; ++i) counts[i] = 0;
    if (!pl->exec) return ISOEXT_OK;   // a grid without cells
    return mc_dense_impl(pl->values, pl->shape, pl->amin, pl->amax, pl->level, pl->method, pl->cfg, pl->workspace,
                         pl->workspace_bytes, pl->special_cap, pl->out_v, pl->cap_v, pl->out_f, pl->cap_f, nullptr,
                         counts, stream, nullptr, 0, nullptr, nullptr, pl, kPlanRun);
}

void isoext_mc_plan_destroy(isoext_mc_plan *pl) {
    if (!pl) return;
    if (pl->exec) (void) hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void) hipGraphDestroy(pl->graph);
    delete pl;
}

// ---- SDF programs --------------------------------------------------------------------

int64_t isoext_sdf_program_bytes(void) { return (int64_t) sizeof(iso::CsgProgram); }

int isoext_sdf_upload(const int32_t *ops, int32_t n_ops, const float *params, int32_t n_params, void *program,
                      void *stream) {
    if (!program) return fail(ISOEXT_E_ARG, "program is null");
    if ((uintptr_t) program & 15) return fail(ISOEXT_E_ARG, "program must be 16-byte aligned");
    // validate first (host only), then stage in pinned memory for the copy
    thread_local std::unique_ptr<iso::CsgProgram> built(new iso::CsgProgram);
    if (const char *err = iso::csg_build(ops, n_ops, params, n_params, *built))
        return fail(ISOEXT_E_ARG, "%s", err);
    iso::CsgProgram *h = (iso::CsgProgram *) g_pinned_prog.get(sizeof(iso::CsgProgram));
    if (!h) return fail(ISOEXT_E_DEVICE, "pinned staging allocation failed");
    std::memcpy(h, built.get(), sizeof(iso::CsgProgram));
    hipStream_t st = (hipStream_t) stream;
    HIP_CHECK(hipMemcpyAsync(program, h, sizeof(iso::CsgProgram), hipMemcpyHostToDevice, st), "upload program");
    HIP_CHECK(hipStreamSynchronize(st), "upload program");   // the staging buffer is reused by the next call
    return ISOEXT_OK;
}

int isoext_sdf_grid(const void *program, const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                    float *out, void *stream) {
    if (!program || !out) return fail(ISOEXT_E_ARG, "program or out is null");
    for (int a = 0; a < 3; ++a)
        if (shape[a] < 1) return fail(ISOEXT_E_ARG, "grid shape must be positive");
    AABB_CHECK(aabb_min, aabb_max);
    HIP_CHECK(iso::sdf_grid_launch((const iso::CsgProgram *) program, shape, aabb_min, aabb_max, out,
                                   (hipStream_t) stream),
              "SDF grid");
    return ISOEXT_OK;
}

int isoext_sdf_points(const void *program, const float *points, int64_t n, float *values, float *normals,
                      void *stream) {
    if (!program) return fail(ISOEXT_E_ARG, "program is null");
    if (n < 0 || (n > 0 && !points)) return fail(ISOEXT_E_ARG, "points is null");
    HIP_CHECK(iso::sdf_points_launch((const iso::CsgProgram *) program, points, n, values, normals,
                                     (hipStream_t) stream),
              "SDF points");
    return ISOEXT_OK;
}

int isoext_mc_dense(const float *values, const int64_t shape[3], const float aabb_min[3],
                    const float aabb_max[3], float level, int method, int cfg, void *workspace,
                    int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v,
                    int32_t *out_f, int64_t cap_f, int64_t counts[6], void *stream) {
    return mc_dense_impl(values, shape, aabb_min, aabb_max, level, method, cfg, workspace, workspace_bytes,
                         special_cap, out_v, cap_v, out_f, cap_f, nullptr, counts, stream);
}

int isoext_mc_dense_keyed(const float *values, const int64_t shape[3], const float aabb_min[3],
                          const float aabb_max[3], float level, int method, int cfg, void *workspace,
                          int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v,
                          int32_t *out_f, int64_t cap_f, int64_t *out_key, int64_t counts[6], void *stream) {
    return mc_dense_impl(values, shape, aabb_min, aabb_max, level, method, cfg, workspace, workspace_bytes,
                         special_cap, out_v, cap_v, out_f, cap_f, out_key, counts, stream);
}

// na != nullptr: the allocation-free variant -- the weld fix-up runs in
// na's device scratch, read-backs go through stack memory, no timing events.
static int mc_dense_impl(const float *values, const int64_t shape[3], const float aabb_min[3],
                         const float aabb_max[3], float level, int method, int cfg, void *workspace,
                         int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v,
                         int32_t *out_f, int64_t cap_f, int64_t *out_key, int64_t counts[6], void *stream,
                         const float *spheres, int nsph, const iso::FixupScratch *na, const iso::CsgProgram *csg,
                         isoext_mc_plan *plan, int plan_mode) {
    for (int i = 0; i < 6; ++i) counts[i] = 0;
    if (method != 0 && method != 1) return fail(ISOEXT_E_ARG, "Unknown method: %d", method);
    AABB_CHECK(aabb_min, aabb_max);
    for (int a = 0; a < 3; ++a)
        if (shape[a] < 1) return fail(ISOEXT_E_ARG, "grid shape must be positive");
    if (shape[0] < 2 || shape[1] < 2 || shape[2] < 2) return ISOEXT_OK;   // no cells
    McLayout L;
    if (!mc_layout(shape, cfg, special_cap, L)) return fail(ISOEXT_E_ARG, "unknown tile config %d", cfg);
    if ((int64_t) L.total > workspace_bytes)
        return fail(ISOEXT_E_WORKSPACE, "workspace too small: %lld < %lld bytes",
                    (long long) workspace_bytes, (long long) L.total);
    if (!values && !spheres && !csg) return fail(ISOEXT_E_ARG, "values is null");
    if (csg && ((uintptr_t) csg & 15)) return fail(ISOEXT_E_ARG, "SDF program must be 16-byte aligned");
    if (spheres && (nsph < 1 || nsph > iso::kMaxFusedSpheres))
        return fail(ISOEXT_E_ARG, "sphere count %d outside [1, %d]", nsph, iso::kMaxFusedSpheres);
    if (spheres && ((uintptr_t) spheres & 15)) return fail(ISOEXT_E_ARG, "spheres must be 16-byte aligned");
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace;

    iso::SpecialLists sp_host{};
    sp_host.dup = -1;
    sp_host.cap_cand = iso::kSpSeg * L.seg_cap[0];
    sp_host.cap_scell = iso::kSpSeg * L.seg_cap[1];
    sp_host.cap_kept = iso::kSpSeg * L.seg_cap[2];
    sp_host.cap_hole = iso::kSpSeg * L.seg_cap[3];
    sp_host.cand = (iso::SpecialVertex *) (ws + L.o_cand);
    sp_host.scell = (iso::SpecialEdge *) (ws + L.o_scell);
    sp_host.kept = (int64_t *) (ws + L.o_kept);
    sp_host.hole = (int64_t *) (ws + L.o_hole);

    iso::McParams p{};
    p.values = values;
    p.X = shape[0];
    p.Y = shape[1];
    p.Z = shape[2];
    for (int a = 0; a < 3; ++a) {
        p.amin[a] = aabb_min[a];
        p.amax[a] = aabb_max[a];
    }
    p.level = level;
    p.method = method;
    const WeldGeom wg = weld_geom(shape, aabb_min, aabb_max);
    p.all_special = wg.all_special;
    p.cand_eps = wg.cand_eps;
    p.vec4 = (shape[2] % 4 == 0 && ((uintptr_t) values & 15) == 0) ? 1 : 0;
    p.W = L.W;
    p.nzc = L.nzc;
    p.S = L.S;
    p.NI = L.NI;
    p.NJ = L.NJ;
    p.NK = L.NK;
    p.bits = (uint64_t *) (ws + L.o_bits);
    p.cnt_v = (uint32_t *) (ws + L.o_cnt_v);
    p.cnt_f = (uint32_t *) (ws + L.o_cnt_f);
    p.bsum_v = (uint32_t *) (ws + L.o_bsum_v);
    p.bsum_f = (uint32_t *) (ws + L.o_bsum_f);
    p.blk_v = (int64_t *) (ws + L.o_blk_v);
    p.blk_f = (int64_t *) (ws + L.o_blk_f);
    p.nblk = L.nblk;
    p.out_v = out_v;
    p.cap_v = out_v ? cap_v : 0;
    p.out_key = out_key;
    p.out_f = out_f;
    p.cap_f = out_f ? cap_f : 0;
    p.err = (int *) (ws + kOffErr);
    p.totals = (int64_t *) (ws + kOffTotals);
    p.sp = (iso::SpecialLists *) (ws + kOffSp);
    p.seg.ctr = (unsigned long long *) (ws + L.o_ctr);
    for (int k = 0; k < 4; ++k) p.seg.cap[k] = L.seg_cap[k];
    p.seg.nseg = iso::kSpSeg;
    p.seg.cand = (iso::SpecialVertex *) (ws + L.o_scand);
    p.seg.scell = (iso::SpecialEdge *) (ws + L.o_sscell);
    p.seg.kept = (int64_t *) (ws + L.o_skept);
    p.seg.hole = (int64_t *) (ws + L.o_shole);
    p.tile_flag = (int *) (ws + L.o_tflag);
    p.tile_list = (int *) (ws + L.o_tlist);
    p.tile_count = (int *) (ws + L.o_tcnt);
    p.dslots = (int32_t *) (ws + L.o_dslots);
    p.dmask = L.dslots - 1;
    p.tx = L.tx;
    p.ty = L.ty;
    if (spheres) {
        p.values = nullptr;
        p.vec4 = 0;
        p.spheres = (const float4 *) spheres;
        p.nsph = nsph;
    } else if (csg) {
        p.values = nullptr;
        p.vec4 = 0;
        p.csg = csg;
        p.csg_list = iso::csg_list_fits(p.X, p.Y, p.W, p.S) ? 1 : 0;
    }
    hipEvent_t *ev = (g_timing && !na && !plan) ? timing_events() : nullptr;
    if (g_timing && !na && !plan && !ev) return fail(ISOEXT_E_DEVICE, "timing events unavailable");
    if (plan && L.stream) return fail(ISOEXT_E_ARG, "plans take the four-stage tile configurations, not %d", cfg);
    if (plan_mode == kPlanCapture) {   // record the pipeline's launches into the plan's graph; no run
        HIP_CHECK(iso::mc_upload_luts(), "upload tables");
        HIP_CHECK(plan_capture(plan, cfg, p, sp_host), "capture marching cubes");
        return ISOEXT_OK;
    }
    const bool read_once = L.stream && !spheres && !csg && L.NR <= (int64_t) INT32_MAX && shape[2] < (int64_t(1) << 26);
    if (read_once) {
        p.NRX = L.NRX;
        p.NRY = L.NRY;
        p.NR = L.NR;
        p.vofs = (uint32_t *) (ws + L.o_vofs);
        p.fofs = (uint32_t *) (ws + L.o_fofs);
        p.wcnt_v = (uint32_t *) (ws + L.o_wcv);
        p.wcnt_f = (uint32_t *) (ws + L.o_wcf);
        p.wb_v = (int64_t *) (ws + L.o_wbv);
        p.wb_f = (int64_t *) (ws + L.o_wbf);
        const int64_t nchunk = (L.NR * iso::kStrWaves + iso::kStrScanChunk - 1) / iso::kStrScanChunk;
        p.csum_v = (int64_t *) (ws + L.o_csum);
        p.csum_f = p.csum_v + nchunk;
        p.cofs_v = (int64_t *) (ws + L.o_cofs);
        p.cofs_f = p.cofs_v + nchunk;
        p.arena_ctr = (unsigned long long *) (ws + L.o_actr);
        p.arena_off = (int32_t *) (ws + L.o_aoff);
        p.wover = (uint8_t *) (ws + L.o_wover);
        p.alist = (int64_t *) (ws + L.o_alist);
        p.acount = (int *) (ws + L.o_acnt);
        // the vertex arena: the face buffer, written by the faces kernel only
        // after the flush has emptied it (16-byte entries, 16-byte aligned)
        p.arena = nullptr;
        p.arena_seg_cap = 0;
        if (out_f && cap_f > 0) {
            const uintptr_t a = ((uintptr_t) out_f + 15) & ~(uintptr_t) 15;
            const int64_t bytes = (int64_t) ((uintptr_t) out_f + 12 * (uintptr_t) cap_f - a);
            const int64_t entries = std::min<int64_t>(bytes / 16, (int64_t) INT32_MAX);
            p.arena = (void *) a;
            p.arena_seg_cap = entries / iso::kStrArenaSeg;
        }
        HIP_CHECK(iso::mc_stream_launch(p, sp_host, st, ev), "launch marching cubes");
    } else if (plan_mode == kPlanRun) {
        HIP_CHECK(hipGraphLaunch(plan->exec, st), "launch the plan's graph");
    } else {
        HIP_CHECK(iso::mc_dense_launch(cfg, p, sp_host, st, ev), "launch marching cubes");
    }

    // read back counts and the special lists' descriptor (one pinned copy of
    // the workspace header: err, totals, SpecialLists)
    struct Header {
        int err;
        int pad;
        int64_t totals[2];
    };
    static_assert(kOffTotals == kOffErr + 8 && kOffSp >= kOffErr + (int) sizeof(Header), "header layout");
    const size_t rb_bytes = kOffSp - kOffErr + sizeof(iso::SpecialLists);
    alignas(64) char rb_stack[kOffSp + sizeof(iso::SpecialLists)];
    char *rb = na ? rb_stack : (char *) g_pinned_rb.get(rb_bytes);
    if (!rb) return fail(ISOEXT_E_DEVICE, "pinned read-back allocation failed");
    HIP_CHECK(hipMemcpyAsync(rb, ws + kOffErr, rb_bytes, hipMemcpyDeviceToHost, st), "read counts");
    HIP_CHECK(hipStreamSynchronize(st), "marching cubes");
    Header hdr;
    iso::SpecialLists sp_dev;
    std::memcpy(&hdr, rb, sizeof(hdr));
    std::memcpy(&sp_dev, rb + (kOffSp - kOffErr), sizeof(sp_dev));
    if (ev) {
        float ms[5] = {-1.0f, -1.0f, -1.0f, -1.0f, -1.0f};
        for (int i = 0; i < 4; ++i) (void) hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
        (void) hipEventElapsedTime(&ms[4], ev[0], ev[4]);
        g_times = {ms[0], ms[1], ms[2], ms[3], ms[4]};
        g_kernel_ms = ms[4];
    }
    int64_t nv = hdr.totals[0], nf = hdr.totals[1];
    counts[0] = counts[4] = nv;
    counts[1] = counts[5] = nf;
    int64_t need = std::max<int64_t>(
        sp_dev.need,
        (int64_t) std::max<unsigned long long>(
            sp_dev.n_cand, std::max<unsigned long long>((sp_dev.n_scell + 11) / 12,
                                                        std::max((sp_dev.n_kept + 4) / 5, (sp_dev.n_hole + 4) / 5))));
    counts[2] = need;
    counts[3] = hdr.err;
    if (hdr.err & iso::ISO_ERR_TOO_LARGE)
        return fail(ISOEXT_E_TOO_LARGE, "mesh exceeds 2^31-1 vertices or faces");
    if (nv > cap_v || nf > cap_f || (hdr.err & iso::ISO_ERR_CAPACITY))
        return fail(ISOEXT_E_CAPACITY, "output capacity too small: need %lld vertices, %lld faces",
                    (long long) nv, (long long) nf);
    if (need > special_cap)
        return fail(ISOEXT_E_WORKSPACE, "special-vertex lists need capacity %lld", (long long) need);
    // no dropped face and no two candidates at one position: every vertex
    // stays referenced and none has a twin, so there is nothing to weld
    const bool weld_noop = sp_dev.dup == 0 && sp_dev.n_hole == 0;
    if ((sp_dev.n_cand > 0 || sp_dev.n_hole > 0) && !weld_noop) {
        iso::FixupGrid g{shape[0], shape[1], shape[2], p.blk_v, p.blk_f, p.cnt_v, p.cnt_f, p.S};
        if (read_once) g = iso::FixupGrid{shape[0], shape[1], shape[2]};   // region-major numbering: no skip tables
        // the fixed scratch of the allocation-free variant must hold the
        // fix-up the lists call for
        const size_t multi_bytes = iso::weld_fixup_bytes(sp_dev, nv, nf);
        if (na && multi_bytes > na->dev_bytes)
            return fail(ISOEXT_E_WORKSPACE, "weld fix-up needs %zu scratch bytes, the fixed limit is %zu",
                        multi_bytes, na->dev_bytes);
        if (iso::fixup_small_ok(sp_dev)) {   // one workgroup, one launch
            int64_t fin[3];
            int64_t *keys1[1] = {out_key};
            size_t hb = 0;
            const size_t sb = iso::weld_fixup_small_batch_bytes(1, &sp_dev, &hb);
            if (na && (sb > na->dev_bytes || hb > na->host_bytes))
                return fail(ISOEXT_E_WORKSPACE, "weld fix-up needs %zu scratch bytes, the fixed limit is %zu", sb,
                            na->dev_bytes);
            HIP_CHECK(iso::weld_fixup_small_batch(1, &sp_dev, &out_v, &out_f, keys1, &nv, &nf, &g, st, fin, na),
                      "weld fix-up");
            if (fin[2] == iso::kFixSmallOverflow) {
                HIP_CHECK(iso::weld_fixup(sp_dev, out_v, out_f, nv, nf, g, st, out_key, na), "weld fix-up");
            } else if (fin[2]) {
                return fail(ISOEXT_E_DEVICE, "weld fix-up: inconsistent special lists");
            } else {
                nv = fin[0];
                nf = fin[1];
            }
        } else {
            HIP_CHECK(iso::weld_fixup(sp_dev, out_v, out_f, nv, nf, g, st, out_key, na), "weld fix-up");
        }
        counts[0] = nv;
        counts[1] = nf;
    }
    return ISOEXT_OK;
}

// UniformGrid::get_points (src/grid/uniform.cu:22-30): (X, Y, Z, 3) positions.
__global__ void k_uniform_points(float *out, int64_t X, int64_t Y, int64_t Z, float ax, float ay, float az,
                                 float bx, float by, float bz) {
    int64_t n = X * Y * Z;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        int64_t z = i % Z, y = (i / Z) % Y, x = i / (Y * Z);
        out[3 * i + 0] = iso::grid_coord(x, X, ax, bx);
        out[3 * i + 1] = iso::grid_coord(y, Y, ay, by);
        out[3 * i + 2] = iso::grid_coord(z, Z, az, bz);
    }
}

int isoext_uniform_points(const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                          float *out, void *stream) {
    int64_t n = shape[0] * shape[1] * shape[2];
    if (n == 0) return ISOEXT_OK;
    unsigned grid = (unsigned) std::min<int64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(k_uniform_points, dim3(grid), dim3(256), 0, (hipStream_t) stream, out, shape[0],
                       shape[1], shape[2], aabb_min[0], aabb_min[1], aabb_min[2], aabb_max[0], aabb_max[1],
                       aabb_max[2]);
    HIP_CHECK(hipGetLastError(), "launch uniform points");
    return ISOEXT_OK;
}

__global__ void k_fill(float *p, int64_t n, float v) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- allocation-free marching cubes ----------------------------------------------
// isoext_mc_dense_noalloc runs the dense pipeline of isoext_mc_dense in one
// fixed device array per GPU instead of a caller workspace and per-thread
// caches: the mask workspace, the special lists and the weld fix-up's tables
// all live in `scratch`.  Host code never passes `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the
// first call on a device looks up that device's copy with
// hipGetSymbolAddress, clears it with a kernel, and keeps the address.
constexpr size_t kNoAllocScratchBytes = size_t(96) << 20;   // per GPU, held until the process ends
constexpr size_t kNoAllocFixupBytes = size_t(16) << 20;     // the weld fix-up's part of it
constexpr int64_t kNoAllocSpecialCap = 32768;               // special-list capacity (counts[2] reports a call's need)
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

static std::mutex g_noalloc_mu;
static char *g_noalloc_ptr[kNoAllocMaxDevices];

__global__ void k_zero_u64(uint64_t *p, int64_t n) {
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        p[i] = 0;
}

// The current device's scratch address, set up on the first call there.
static int noalloc_scratch(hipStream_t st, char **out) {
    int dev = -1;
    HIP_CHECK(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(ISOEXT_E_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(ISOEXT_E_DEVICE, "scratch has no device address");
        hipLaunchKernelGGL(k_zero_u64, dim3(1024), dim3(256), 0, st, (uint64_t *) ptr,
                           (int64_t) (kNoAllocScratchBytes / sizeof(uint64_t)));
        HIP_CHECK(hipGetLastError(), "clear scratch");
        HIP_CHECK(hipStreamSynchronize(st), "clear scratch");
        g_noalloc_ptr[dev] = (char *) ptr;
    }
    *out = g_noalloc_ptr[dev];
    return ISOEXT_OK;
}

int64_t isoext_mc_dense_noalloc_limit(void) { return (int64_t) (kNoAllocScratchBytes - kNoAllocFixupBytes); }

int isoext_mc_dense_noalloc(const float *values, const int64_t shape[3], const float aabb_min[3],
                            const float aabb_max[3], float level, int method, float *out_v, int64_t cap_v,
                            int32_t *out_f, int64_t cap_f, int64_t counts[6], void *stream) {
    for (int i = 0; i < 6; ++i) counts[i] = 0;
    for (int a = 0; a < 3; ++a)
        if (shape[a] < 1) return fail(ISOEXT_E_ARG, "grid shape must be positive");
    const int cfg = iso::mc_auto_config(shape);
    McLayout L;
    if (shape[0] >= 2 && shape[1] >= 2 && shape[2] >= 2) {
        if (!mc_layout(shape, cfg, kNoAllocSpecialCap, L)) return fail(ISOEXT_E_ARG, "no tile configuration");
        const size_t limit = kNoAllocScratchBytes - kNoAllocFixupBytes;
        if (L.total > limit)
            return fail(ISOEXT_E_TOO_LARGE,
                        "grid %lld x %lld x %lld needs %zu bytes of scratch; the allocation-free entry point "
                        "holds %zu (use isoext_mc_dense)",
                        (long long) shape[0], (long long) shape[1], (long long) shape[2], L.total, limit);
    }
    hipStream_t st = (hipStream_t) stream;
    char *base = nullptr;
    const int rc = noalloc_scratch(st, &base);
    if (rc != ISOEXT_OK) return rc;
    alignas(256) char host_stage[4096];
    const iso::FixupScratch fx{base + (kNoAllocScratchBytes - kNoAllocFixupBytes), kNoAllocFixupBytes, host_stage,
                               sizeof(host_stage)};
    return mc_dense_impl(values, shape, aabb_min, aabb_max, level, method, cfg, base,
                         (int64_t) (kNoAllocScratchBytes - kNoAllocFixupBytes), kNoAllocSpecialCap, out_v, cap_v,
                         out_f, cap_f, nullptr, counts, stream, nullptr, 0, &fx);
}

int isoext_fill_f32(float *ptr, int64_t n, float value, void *stream) {
    if (n <= 0) return ISOEXT_OK;
    unsigned grid = (unsigned) std::min<int64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, (hipStream_t) stream, ptr, n, value);
    HIP_CHECK(hipGetLastError(), "launch fill");
    return ISOEXT_OK;
}


static iso::SparseGeom geom_of(const int64_t shape[3], const float amin[3], const float amax[3]) {
    iso::SparseGeom g;
    g.X = shape[0];
    g.Y = shape[1];
    g.Z = shape[2];
    for (int a = 0; a < 3; ++a) {
        g.amin[a] = amin[a];
        g.amax[a] = amax[a];
    }
    return g;
}

int isoext_sparse_corner_points(const int64_t *cell_ids, int64_t n, const int64_t shape[3], const float aabb_min[3],
                                const float aabb_max[3], float *out, void *stream) {
    if (n <= 0) return ISOEXT_OK;
    if (shape[1] < 2 || shape[2] < 2) return fail(ISOEXT_E_ARG, "sparse grid shape must be >= 2 along y and z");
    HIP_CHECK(iso::sparse_corner_points(cell_ids, n, geom_of(shape, aabb_min, aabb_max), out, (hipStream_t) stream),
              "corner points");
    return ISOEXT_OK;
}

int isoext_sparse_crossing_flags(const float *values, int64_t n, float level, uint8_t *flags, void *stream) {
    if (n <= 0) return ISOEXT_OK;
    HIP_CHECK(iso::sparse_crossing_flags(values, n, level, flags, (hipStream_t) stream), "crossing flags");
    return ISOEXT_OK;
}

int64_t isoext_sparse_setop_scratch_bytes(int64_t na, int64_t nb) {
    if (na < 0 || nb < 0) return fail(ISOEXT_E_ARG, "negative sizes");
    return iso::sparse_setop_scratch_bytes(na, nb);
}

static int sparse_setop_impl(const int64_t *a, int64_t na, const int64_t *b, int64_t nb, bool difference,
                             void *scratch, int64_t scratch_bytes, int64_t *out, int64_t *slot, int64_t *n_out,
                             void *stream) {
    if (na < 0 || nb < 0 || (na > 0 && !a) || (nb > 0 && !b) || !out || !n_out)
        return fail(ISOEXT_E_ARG, "invalid set-operation arguments");
    if (scratch_bytes < iso::sparse_setop_scratch_bytes(na, nb))
        return fail(ISOEXT_E_WORKSPACE, "set-operation scratch too small");
    HIP_CHECK(iso::sparse_setop(a, na, b, nb, difference, scratch, out, slot, n_out, (hipStream_t) stream),
              difference ? "sparse difference" : "sparse union");
    return ISOEXT_OK;
}

int isoext_sparse_union(const int64_t *a, int64_t na, const int64_t *b_sorted, int64_t nb, void *scratch,
                        int64_t scratch_bytes, int64_t *out, int64_t *slot, int64_t *n_out, void *stream) {
    return sparse_setop_impl(a, na, b_sorted, nb, false, scratch, scratch_bytes, out, slot, n_out, stream);
}

int isoext_sparse_difference(const int64_t *a, int64_t na, const int64_t *b_sorted, int64_t nb, void *scratch,
                             int64_t scratch_bytes, int64_t *out, int64_t *slot, int64_t *n_out, void *stream) {
    return sparse_setop_impl(a, na, b_sorted, nb, true, scratch, scratch_bytes, out, slot, n_out, stream);
}

int isoext_sparse_gather_values(const float *old_values, const int64_t *slot, int64_t n, float default_value,
                                float *new_values, void *stream) {
    if (n < 0 || (n > 0 && (!slot || !new_values))) return fail(ISOEXT_E_ARG, "invalid gather arguments");
    HIP_CHECK(iso::sparse_gather_rows(old_values, slot, n, default_value, new_values, (hipStream_t) stream),
              "sparse gather values");
    return ISOEXT_OK;
}

int64_t isoext_mc_sparse_workspace_bytes(int64_t n, int64_t ninst) {
    return (int64_t) iso::sparse_mc_work_bytes(n, ninst);
}

static iso::SparseMcIn smc_in(const isoext_grid *g, float level, int method) {
    iso::SparseMcIn in;
    in.values = g->values;
    in.ids = g->sparse_cell_ids;
    in.n = g->n_sparse;
    in.g = geom_of(g->shape, g->aabb_min, g->aabb_max);
    in.level = level;
    in.method = method;
    return in;
}

int isoext_mc_sparse_count(const isoext_grid *g, float level, int64_t *scratch, int64_t *ninst, void *stream) {
    *ninst = 0;
    AABB_CHECK(g->aabb_min, g->aabb_max);
    if (!g->sparse_cell_ids || g->n_sparse == 0) return ISOEXT_OK;
    HIP_CHECK(iso::sparse_mc_count(smc_in(g, level, 0), scratch, *ninst, (hipStream_t) stream), "sparse count");
    return ISOEXT_OK;
}

int isoext_mc_sparse(const isoext_grid *g, float level, int method, int64_t ninst, void *workspace,
                     int64_t workspace_bytes, float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f,
                     int64_t *vsrc, int64_t counts[2], void *stream) {
    counts[0] = counts[1] = 0;
    if (method != 0 && method != 1) return fail(ISOEXT_E_ARG, "Unknown method: %d", method);
    AABB_CHECK(g->aabb_min, g->aabb_max);
    if (!g->sparse_cell_ids || g->n_sparse == 0 || ninst == 0) return ISOEXT_OK;
    if (workspace_bytes < isoext_mc_sparse_workspace_bytes(g->n_sparse, ninst))
        return fail(ISOEXT_E_WORKSPACE, "sparse marching cubes workspace too small");
    iso::SparseMcOut o{out_v, out_v ? cap_v : 0, out_f, out_f ? cap_f : 0, vsrc};
    int64_t nv = 0, nf = 0;
    HIP_CHECK(iso::sparse_mc(smc_in(g, level, method), ninst, workspace, (size_t) workspace_bytes, o, nv, nf,
                             (hipStream_t) stream),
              "sparse marching cubes");
    counts[0] = nv;
    counts[1] = nf;
    if (nv > cap_v || nf > cap_f)
        return fail(ISOEXT_E_CAPACITY, "output capacity too small: need %lld vertices, %lld faces", (long long) nv,
                    (long long) nf);
    return ISOEXT_OK;
}


int isoext_mc_backward(const isoext_grid *g, float level, const float *grad_v, const int64_t *vertex_src, int64_t nv,
                       float *grad_values, void *stream) {
    if (nv <= 0) return ISOEXT_OK;
    iso::DenseGeom d;
    d.X = g->shape[0];
    d.Y = g->shape[1];
    d.Z = g->shape[2];
    for (int a = 0; a < 3; ++a) {
        d.amin[a] = g->aabb_min[a];
        d.amax[a] = g->aabb_max[a];
    }
    hipStream_t st = (hipStream_t) stream;
    if (g->sparse_cell_ids)
        HIP_CHECK(iso::mc_backward_sparse(grad_v, vertex_src, nv, g->values, g->sparse_cell_ids, d, level, grad_values,
                                          st),
                  "marching cubes backward");
    else
        HIP_CHECK(iso::mc_backward_dense(grad_v, vertex_src, nv, g->values, d, level, grad_values, st),
                  "marching cubes backward");
    return ISOEXT_OK;
}

int64_t isoext_mc_batch_workspace_bytes(int n, const isoext_mc_item *items, int cfg, int64_t special_cap) {
    if (batch_check(n, items, cfg) != ISOEXT_OK) return -1;
    BatchLayout B;
    if (!batch_layout(n, items, cfg, special_cap, B)) return -1;
    return (int64_t) B.total;
}

int isoext_mc_batch_count(int n, const isoext_mc_item *items, int cfg, void *workspace, int64_t workspace_bytes,
                          int64_t special_cap, int64_t *counts, void *stream) {
    int rc = batch_check(n, items, cfg);
    if (rc != ISOEXT_OK) return rc;
    for (int i = 0; i < 2 * n; ++i) counts[i] = 0;
    BatchLayout B;
    batch_layout(n, items, cfg, special_cap, B);
    if ((int64_t) B.total > workspace_bytes)
        return fail(ISOEXT_E_WORKSPACE, "workspace too small: %lld < %lld bytes", (long long) workspace_bytes,
                    (long long) B.total);
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace;
    iso::McBatchSummary *sum = nullptr;
    if ((rc = batch_upload(n, items, ws, B, st, &sum)) != ISOEXT_OK) return rc;
    HIP_CHECK(iso::mc_batch_count_launch((const iso::McBatchItem *) (ws + B.o_items),
                                         (const iso::McBatchHdr *) (ws + B.o_hdr), B.h,
                                         (iso::McBatchSummary *) (ws + B.o_summary), st),
              "launch batch count");
    HIP_CHECK(hipMemcpyAsync(sum, ws + B.o_summary, sizeof(iso::McBatchSummary) * (size_t) n, hipMemcpyDeviceToHost,
                             st),
              "read batch counts");
    HIP_CHECK(hipStreamSynchronize(st), "batch count");
    for (int i = 0; i < n; ++i) {
        if (sum[i].err & iso::ISO_ERR_TOO_LARGE)
            return fail(ISOEXT_E_TOO_LARGE, "item %d: mesh exceeds 2^31-1 vertices or faces", i);
        counts[2 * i] = sum[i].nv;
        counts[2 * i + 1] = sum[i].nf;
    }
    return ISOEXT_OK;
}

int isoext_mc_batch_emit(int n, const isoext_mc_item *items, int cfg, void *workspace, int64_t workspace_bytes,
                         int64_t special_cap, int64_t *counts, void *stream) {
    int rc = batch_check(n, items, cfg);
    if (rc != ISOEXT_OK) return rc;
    BatchLayout B;
    batch_layout(n, items, cfg, special_cap, B);
    if ((int64_t) B.total > workspace_bytes)
        return fail(ISOEXT_E_WORKSPACE, "workspace too small: %lld < %lld bytes", (long long) workspace_bytes,
                    (long long) B.total);
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace;
    iso::McBatchSummary *sum = nullptr;
    if ((rc = batch_upload(n, items, ws, B, st, &sum)) != ISOEXT_OK) return rc;
    const bool listed = batch_tile_list(n, items, B);
    HIP_CHECK(iso::mc_batch_emit_launch(B.tile_begin.data(), B.tile_end.data(),
                                        (const iso::McBatchItem *) (ws + B.o_items),
                                        (const int32_t *) (ws + B.o_order), (const iso::McBatchHdr *) (ws + B.o_hdr),
                                        B.h, (iso::McBatchSummary *) (ws + B.o_summary), st,
                                        listed ? (const int *) (ws + B.o_tlist) : nullptr,
                                        listed ? (const int *) (ws + B.o_tcnt) : nullptr),
              "launch batch emit");
    HIP_CHECK(hipMemcpyAsync(sum, ws + B.o_summary, sizeof(iso::McBatchSummary) * (size_t) n, hipMemcpyDeviceToHost,
                             st),
              "read batch counts");
    HIP_CHECK(hipStreamSynchronize(st), "batch emit");
    // weld fix-ups: items with small lists in one launch (a workgroup each),
    // the others back to back on the stream, then one read of their counts
    int64_t *fix_dev = (int64_t *) (ws + B.o_fixcnt);
    std::vector<int> fixed, small;
    std::vector<iso::SpecialLists> s_sp;
    std::vector<float *> s_v;
    std::vector<int32_t *> s_f;
    std::vector<int64_t> s_nv, s_nf;
    std::vector<iso::FixupGrid> s_g;
    for (int i = 0; i < n; ++i) {
        const int64_t nv = sum[i].nv, nf = sum[i].nf;
        counts[2 * i] = nv;
        counts[2 * i + 1] = nf;
        if (nv > items[i].cap_v || nf > items[i].cap_f)
            return fail(ISOEXT_E_CAPACITY, "item %d: output capacity too small", i);
        if (sum[i].sp.need > special_cap) {   // special lists overflowed: the caller re-runs this item alone
            counts[2 * i] = counts[2 * i + 1] = -1;
            continue;
        }
        if (sum[i].sp.n_cand > 0 || sum[i].sp.n_hole > 0) {
            const ItemLayout &IL = B.it[i];
            iso::FixupGrid g{items[i].shape[0], items[i].shape[1], items[i].shape[2],
                             (const int64_t *) (ws + IL.o_blk_v), (const int64_t *) (ws + IL.o_blk_f),
                             (const uint32_t *) (ws + IL.o_cnt_v), (const uint32_t *) (ws + IL.o_cnt_f), IL.S};
            if (iso::fixup_small_ok(sum[i].sp)) {
                small.push_back(i);
                s_sp.push_back(sum[i].sp);
                s_v.push_back(items[i].out_v);
                s_f.push_back(items[i].out_f);
                s_nv.push_back(nv);
                s_nf.push_back(nf);
                s_g.push_back(g);
                continue;
            }
            HIP_CHECK(iso::weld_fixup_async(sum[i].sp, items[i].out_v, items[i].out_f, nv, nf, g, st, nullptr,
                                            fix_dev + 4 * i),
                      "weld fix-up");
            fixed.push_back(i);
        }
    }
    if (!small.empty()) {   // synchronises the stream, so the fix-ups above are done too
        std::vector<int64_t> fin(3 * small.size());
        HIP_CHECK(iso::weld_fixup_small_batch((int) small.size(), s_sp.data(), s_v.data(), s_f.data(), nullptr,
                                              s_nv.data(), s_nf.data(), s_g.data(), st, fin.data()),
                  "weld fix-up");
        for (size_t k = 0; k < small.size(); ++k) {
            const int i = small[k];
            if (fin[3 * k + 2] == iso::kFixSmallOverflow) {   // lists too long for one workgroup
                HIP_CHECK(iso::weld_fixup_async(s_sp[k], items[i].out_v, items[i].out_f, s_nv[k], s_nf[k], s_g[k], st,
                                                nullptr, fix_dev + 4 * i),
                          "weld fix-up");
                fixed.push_back(i);
                continue;
            }
            if (fin[3 * k + 2]) return fail(ISOEXT_E_DEVICE, "item %d: inconsistent special lists", i);
            counts[2 * i] = fin[3 * k];
            counts[2 * i + 1] = fin[3 * k + 1];
        }
    }
    if (!fixed.empty()) {
        int64_t *fc = (int64_t *) g_pinned_rb.get(sizeof(int64_t) * 4 * (size_t) n);
        if (!fc) return fail(ISOEXT_E_DEVICE, "pinned read-back allocation failed");
        HIP_CHECK(hipMemcpyAsync(fc, fix_dev, sizeof(int64_t) * 4 * (size_t) n, hipMemcpyDeviceToHost, st),
                  "read fix-up counts");
        HIP_CHECK(hipStreamSynchronize(st), "weld fix-up");
        for (int i : fixed) {
            if (fc[4 * i + 2]) return fail(ISOEXT_E_DEVICE, "item %d: inconsistent special lists", i);
            counts[2 * i] = fc[4 * i];
            counts[2 * i + 1] = fc[4 * i + 1];
        }
    }
    return ISOEXT_OK;
}


// ---- batched dual contouring of dense grids ----------------------------------------

int64_t isoext_dc_batch_workspace_bytes(int n, const isoext_mc_item *items) {
    if (n > iso::kMaxDcBatchGrids || batch_check(n, items, -1) != ISOEXT_OK) return -1;
    BatchLayout B;
    if (!batch_layout(n, items, iso::kBatchTileConfig, 0, B)) return -1;
    return (int64_t) B.total;
}

int isoext_dc_batch_count(int n, const isoext_mc_item *items, void *workspace, int64_t workspace_bytes,
                          int64_t *counts, void *stream) {
    if (n > iso::kMaxDcBatchGrids) return fail(ISOEXT_E_ARG, "batch of %d grids exceeds %d", n, iso::kMaxDcBatchGrids);
    int rc = batch_check(n, items, -1);
    if (rc != ISOEXT_OK) return rc;
    for (int i = 0; i < 2 * n; ++i) counts[i] = 0;
    BatchLayout B;
    batch_layout(n, items, iso::kBatchTileConfig, 0, B);
    if ((int64_t) B.total > workspace_bytes)
        return fail(ISOEXT_E_WORKSPACE, "workspace too small: %lld < %lld bytes", (long long) workspace_bytes,
                    (long long) B.total);
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace;
    iso::McBatchSummary *sum = nullptr;
    if ((rc = batch_upload(n, items, ws, B, st, &sum, counts)) != ISOEXT_OK) return rc;
    HIP_CHECK(iso::mc_batch_count_launch((const iso::McBatchItem *) (ws + B.o_items),
                                         (const iso::McBatchHdr *) (ws + B.o_hdr), B.h,
                                         (iso::McBatchSummary *) (ws + B.o_summary), st),
              "launch batch intersection count");
    HIP_CHECK(hipMemcpyAsync(sum, ws + B.o_summary, sizeof(iso::McBatchSummary) * (size_t) n, hipMemcpyDeviceToHost,
                             st),
              "read batch counts");
    HIP_CHECK(hipStreamSynchronize(st), "batch intersection count");
    for (int i = 0; i < n; ++i) {
        if (sum[i].err & iso::ISO_ERR_TOO_LARGE)
            return fail(ISOEXT_E_TOO_LARGE, "item %d: more than 2^31-1 active cells or intersections", i);
        counts[2 * i] = sum[i].nv;
        counts[2 * i + 1] = sum[i].nf;
    }
    return ISOEXT_OK;
}

int64_t isoext_dc_batch_scratch_bytes(int n, const isoext_mc_item *items, const int64_t *counts) {
    if (n < 1 || n > iso::kMaxDcBatchGrids) return -1;
    int64_t C = 0, P = 0;
    for (int i = 0; i < n; ++i) {
        C += counts[2 * i];
        P += counts[2 * i + 1];
    }
    return (int64_t) dc_batch_layout(n, items, C, P).total;
}

int isoext_dc_batch_run(int n, const isoext_mc_item *items, const int64_t *counts, float reg, float svd_tol,
                        void *workspace, int64_t workspace_bytes, void *scratch, int64_t scratch_bytes, float *out_v,
                        int64_t cap_v, int32_t *out_f, int64_t cap_f, int64_t *out_counts, void *stream) {
    if (n > iso::kMaxDcBatchGrids) return fail(ISOEXT_E_ARG, "batch of %d grids exceeds %d", n, iso::kMaxDcBatchGrids);
    int rc = batch_check(n, items, -1);
    if (rc != ISOEXT_OK) return rc;
    for (int i = 0; i < 4 * n + 1; ++i) out_counts[i] = 0;
    BatchLayout B;
    batch_layout(n, items, iso::kBatchTileConfig, 0, B);
    if ((int64_t) B.total > workspace_bytes) return fail(ISOEXT_E_WORKSPACE, "workspace too small");
    int64_t C = 0, P = 0;
    std::vector<int64_t> cell0(n), ent0(n);
    for (int i = 0; i < n; ++i) {
        if (counts[2 * i] < 0 || counts[2 * i + 1] < 0) return fail(ISOEXT_E_ARG, "negative count");
        cell0[i] = C;
        ent0[i] = P;
        C += counts[2 * i];
        P += counts[2 * i + 1];
    }
    // cell and corner indices of the batch are int32 on the device
    if (C > 0x3fffffffll || 6 * P > 0x7fffffffll)
        return fail(ISOEXT_E_TOO_LARGE, "batch of %lld cells / %lld intersections is too large", (long long) C,
                    (long long) P);
    const DcBatchLayout L = dc_batch_layout(n, items, C, P);
    if ((int64_t) L.total > scratch_bytes) return fail(ISOEXT_E_WORKSPACE, "dual contouring scratch too small");
    if (cap_v < C || cap_f < 2 * P) return fail(ISOEXT_E_CAPACITY, "output capacity too small");
    if (C == 0) return ISOEXT_OK;
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace, *sc = (char *) scratch;
    // 1. intersections with normals, concatenated in grid order
    iso::McBatchSummary *sum = nullptr;
    if ((rc = batch_upload(n, items, ws, B, st, &sum, counts)) != ISOEXT_OK) return rc;
    iso::ItsOut *ho = (iso::ItsOut *) g_pinned_its.get(sizeof(iso::ItsOut) * (size_t) n);
    if (!ho) return fail(ISOEXT_E_DEVICE, "pinned staging allocation failed");
    for (int i = 0; i < n; ++i) {
        iso::ItsOut o{};   // active cells and lookup tables only
        o.cell_dense = (int64_t *) (sc + L.o_cd) + cell0[i];
        o.a_base = cell0[i];
        o.e_base = ent0[i];
        o.amask_tab = (uint64_t *) (sc + L.o_tm) + L.tab0[i];
        o.apre_tab = (int32_t *) (sc + L.o_tp) + L.tab0[i];
        ho[i] = o;
    }
    HIP_CHECK(hipMemcpyAsync(sc + L.o_outs, ho, sizeof(iso::ItsOut) * (size_t) n, hipMemcpyHostToDevice, st),
              "upload intersection descriptors");
    const bool listed = batch_tile_list(n, items, B);
    HIP_CHECK(iso::its_batch_emit_launch(B.tile_begin.data(), B.tile_end.data(),
                                         (const iso::McBatchItem *) (ws + B.o_items),
                                         (const int32_t *) (ws + B.o_order), (const iso::ItsOut *) (sc + L.o_outs),
                                         (const iso::McBatchHdr *) (ws + B.o_hdr), st,
                                         listed ? (const int *) (ws + B.o_tlist) : nullptr,
                                         listed ? (const int *) (ws + B.o_tcnt) : nullptr),
              "batch intersections");
    // 2. dual contouring of all grids at once, edges recomputed from the corner values
    iso::ItsIn in{};
    in.ncells = C;
    in.npts = P;
    in.cell_dense = (const int64_t *) (sc + L.o_cd);
    char *dw = sc + L.o_dc;
    const DcLayout &D = L.dc;
    iso::DcScratch w;
    w.dual_v = (float *) (dw + D.o_dual);
    w.cond = (float *) (dw + D.o_cond);
    w.n_fp64 = (unsigned long long *) (dw + D.o_cnt);
    w.used = (uint8_t *) (dw + D.o_used);
    w.bnd = (uint8_t *) (dw + D.o_bnd);
    w.qn = (uint8_t *) (dw + D.o_qn);
    w.qmask = (uint8_t *) (dw + D.o_qm);
    w.bsum = (int64_t *) (dw + D.o_bsum);
    w.part = (int64_t *) (dw + D.o_part);
    w.tot = (int64_t *) (dw + D.o_tot);
    w.tri_cells = (int32_t *) (dw + D.o_tri);
    w.cap_tri = 6;
    w.hslots = (int32_t *) (dw + D.o_hs);
    w.slot_min = (int32_t *) (dw + D.o_hm);
    w.hmask = D.hslots - 1;
    w.akey = (int64_t *) (dw + D.o_ak);
    w.aval = (int32_t *) (dw + D.o_av);
    w.amask = D.hslots - 1;
    w.rep_slot = (int32_t *) (dw + D.o_rep);
    w.final_id = (int32_t *) (dw + D.o_fin);
    w.out_v = out_v;
    w.out_f = out_f;
    w.out_vcell = nullptr;
    const iso::DcBatchDev dd{sc + L.o_items, (int32_t *) (sc + L.o_citem), (int64_t *) (sc + L.o_vb),
                             (int64_t *) (sc + L.o_qb)};
    std::vector<iso::GridDesc> g(n);
    std::vector<const uint64_t *> tm(n);
    std::vector<const int32_t *> tp(n);
    for (int i = 0; i < n; ++i) {
        tm[i] = (const uint64_t *) (sc + L.o_tm) + L.tab0[i];
        tp[i] = (const int32_t *) (sc + L.o_tp) + L.tab0[i];
        isoext_grid gi{items[i].values, nullptr, 0, {items[i].shape[0], items[i].shape[1], items[i].shape[2]},
                       {items[i].aabb_min[0], items[i].aabb_min[1], items[i].aabb_min[2]},
                       {items[i].aabb_max[0], items[i].aabb_max[1], items[i].aabb_max[2]}};
        g[i] = to_desc(&gi);
        g[i].csg = (const iso::CsgProgram *) items[i].sdf;
        g[i].csg_normals = items[i].sdf ? items[i].sdf_normals : 0;
    }
    const float cond_fp64 = std::min(1e6f, 0.5f / std::max(svd_tol, 1e-30f));
    std::vector<int64_t> nv(n), nq(n);
    std::vector<float> levels(n);
    for (int i = 0; i < n; ++i) levels[i] = items[i].level;
    int64_t nfp = 0;
    HIP_CHECK(iso::dc_run_batch(n, g.data(), cell0.data(), tm.data(), tp.data(), levels.data(), in, reg, svd_tol,
                                cond_fp64, w, dd, nv.data(), nq.data(), nfp, st),
              "batched dual contouring");
    int64_t v0 = 0, f0 = 0;
    for (int i = 0; i < n; ++i) {
        out_counts[4 * i] = nv[i];
        out_counts[4 * i + 1] = 2 * nq[i];
        out_counts[4 * i + 2] = v0;
        out_counts[4 * i + 3] = f0;
        v0 += nv[i];
        f0 += 2 * nq[i];
    }
    out_counts[4 * n] = nfp;
    return ISOEXT_OK;
}
}   // extern "C"

// Error reporting for the library's other translation units.
namespace iso {
int set_error(int code, const char *what) { return fail(code, "%s", what); }
}   // namespace iso
