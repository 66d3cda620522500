// C-ABI entry points of libisoext_amd.so.  Plain pointers and int64 sizes
// only; no exceptions cross this boundary.  Each function returns an
// ISOEXT_* status; isoext_last_error() holds the message of the last failure
// on the calling thread.  Declarations and the reference interfaces they
// replace: include/isoext_amd.h.

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/isoext_amd.h"
#include "common.h"
#include "fixup.h"
#include "mc_dense.h"

namespace {

thread_local std::string g_err;

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

// Workspace layout for dense marching cubes.
struct McLayout {
    int ni, nj, nk;
    int64_t ntiles, pub_per_tile;
    size_t o_status, o_pub, o_cand, o_scell, o_kept, o_hole, total;
    int64_t cap;
};

constexpr size_t kHeader = 4096;   // ticket, epoch, err, totals, SpecialLists
constexpr size_t kOffTicket = 0, kOffEpoch = 4, kOffErr = 8, kOffTotals = 16, kOffSp = 64;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

bool mc_layout(const int64_t shape[3], int cfg, int64_t cap, McLayout &L) {
    if (!iso::mc_config_tiles(cfg, shape, L.ni, L.nj, L.nk, L.pub_per_tile)) return false;
    L.ntiles = (int64_t) L.ni * L.nj * L.nk;
    L.cap = cap;
    size_t o = kHeader;
    L.o_status = o;
    o = align256(o + sizeof(uint64_t) * L.ntiles);
    L.o_pub = o;
    o = align256(o + sizeof(uint64_t) * L.ntiles * L.pub_per_tile);
    L.o_cand = o;
    o = align256(o + sizeof(iso::SpecialVertex) * cap);
    L.o_scell = o;
    o = align256(o + sizeof(iso::SpecialEdge) * cap * 12);
    L.o_kept = o;
    o = align256(o + sizeof(int64_t) * cap * 5);
    L.o_hole = o;
    o = align256(o + sizeof(int64_t) * cap * 5);
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

}   // namespace

extern "C" {

const char *isoext_last_error(void) { return g_err.c_str(); }

int isoext_version(void) { return ISOEXT_ABI_VERSION; }

int isoext_mc_config_count(void) { return iso::mc_config_count(); }

int64_t isoext_mc_dense_workspace_bytes(const int64_t shape[3], int cfg, int64_t special_cap) {
    McLayout L;
    if (!mc_layout(shape, cfg, special_cap, L)) return -1;
    return (int64_t) L.total;
}

int isoext_mc_dense(const float *values, const int64_t shape[3], const float aabb_min[3],
                    const float aabb_max[3], float level, int method, int cfg, void *workspace,
                    int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v,
                    int32_t *out_f, int64_t cap_f, int64_t counts[4], void *stream) {
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    if (method != 0 && method != 1) return fail(ISOEXT_E_ARG, "Unknown method: %d", method);
    for (int a = 0; a < 3; ++a)
        if (shape[a] < 1) return fail(ISOEXT_E_ARG, "grid shape must be positive");
    if (shape[0] < 2 || shape[1] < 2 || shape[2] < 2) return ISOEXT_OK;   // no cells
    McLayout L;
    if (!mc_layout(shape, cfg, special_cap, L)) return fail(ISOEXT_E_ARG, "unknown tile config %d", cfg);
    if ((int64_t) L.total > workspace_bytes)
        return fail(ISOEXT_E_WORKSPACE, "workspace too small: %lld < %lld bytes",
                    (long long) workspace_bytes, (long long) L.total);
    if (!values) return fail(ISOEXT_E_ARG, "values is null");
    hipStream_t st = (hipStream_t) stream;
    char *ws = (char *) workspace;

    iso::SpecialLists sp_host{};
    sp_host.cap_cand = special_cap;
    sp_host.cap_scell = special_cap * 12;
    sp_host.cap_kept = special_cap * 5;
    sp_host.cap_hole = special_cap * 5;
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
    p.all_special = grid_well_separated(shape, aabb_min, aabb_max) ? 0 : 1;
    p.NI = L.ni;
    p.NJ = L.nj;
    p.NK = L.nk;
    p.ntiles = L.ntiles;
    p.out_v = out_v;
    p.cap_v = out_v ? cap_v : 0;
    p.out_f = out_f;
    p.cap_f = out_f ? cap_f : 0;
    p.status = (uint64_t *) (ws + L.o_status);
    p.ticket = (uint32_t *) (ws + kOffTicket);
    p.epoch_ptr = (uint32_t *) (ws + kOffEpoch);
    p.pub = (uint64_t *) (ws + L.o_pub);
    p.err = (int *) (ws + kOffErr);
    p.totals = (int64_t *) (ws + kOffTotals);
    p.sp = (iso::SpecialLists *) (ws + kOffSp);
    HIP_CHECK(iso::mc_dense_launch(cfg, p, sp_host, st), "launch marching cubes");

    // read back counts and the special lists' descriptor
    struct {
        int err;
        int pad;
        int64_t totals[2];
    } hdr;
    iso::SpecialLists sp_dev;
    HIP_CHECK(hipMemcpyAsync(&hdr, ws + kOffErr, sizeof(hdr), hipMemcpyDeviceToHost, st), "read counts");
    HIP_CHECK(hipMemcpyAsync(&sp_dev, ws + kOffSp, sizeof(sp_dev), hipMemcpyDeviceToHost, st),
              "read special lists");
    HIP_CHECK(hipStreamSynchronize(st), "marching cubes");
    int64_t nv = hdr.totals[0], nf = hdr.totals[1];
    counts[0] = nv;
    counts[1] = nf;
    int64_t need = (int64_t) std::max<unsigned long long>(
        sp_dev.n_cand, std::max<unsigned long long>((sp_dev.n_scell + 11) / 12,
                                                    std::max((sp_dev.n_kept + 4) / 5, (sp_dev.n_hole + 4) / 5)));
    counts[2] = need;
    counts[3] = hdr.err;
    if (hdr.err & iso::ISO_ERR_SPIN) return fail(ISOEXT_E_DEVICE, "inter-workgroup wait timed out");
    if (hdr.err & iso::ISO_ERR_TOO_LARGE)
        return fail(ISOEXT_E_TOO_LARGE, "mesh exceeds 2^31-1 vertices or faces");
    if (nv > cap_v || nf > cap_f || (hdr.err & iso::ISO_ERR_CAPACITY))
        return fail(ISOEXT_E_CAPACITY, "output capacity too small: need %lld vertices, %lld faces",
                    (long long) nv, (long long) nf);
    if (need > special_cap)
        return fail(ISOEXT_E_WORKSPACE, "special-vertex lists need capacity %lld", (long long) need);
    if (sp_dev.n_cand > 0 || sp_dev.n_hole > 0) {
        iso::FixupGrid g{shape[0], shape[1], shape[2]};
        HIP_CHECK(iso::weld_fixup(sp_dev, out_v, out_f, nv, nf, g, st), "weld fix-up");
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

int isoext_fill_f32(float *ptr, int64_t n, float value, void *stream) {
    if (n <= 0) return ISOEXT_OK;
    unsigned grid = (unsigned) std::min<int64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(k_fill, dim3(grid), dim3(256), 0, (hipStream_t) stream, ptr, n, value);
    HIP_CHECK(hipGetLastError(), "launch fill");
    return ISOEXT_OK;
}

}   // extern "C"
