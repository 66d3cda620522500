// Edge intersections, trilinear normals and dual contouring (gfx950).
//
// Reference path replaced:
//   get_intersection            src/its.cu:93-159   (edge status, compaction,
//                                                   popcount offsets, get_its_op)
//   compute_intersection_normals src/its.cu:170-284 (trilinear central differences)
//   dual_contouring              src/dc.cu:161-218  (QEF, batched SVD solve,
//                                                   clip, quads, weld)
//   get_dual_quads               src/grid/uniform.cu:60-83, src/grid/sparse.cu:181-246
//
// Intersection layout (same entry order as the reference): active cells in
// ascending cell order; inside a cell its crossing edges in edge-index order.
// Per entry: point, normal, owning active cell, local edge index, is_out.
//
// Cells come from a dense grid (cell id = linear (x, y, z) over (X-1, Y-1,
// Z-1)) or from a sparse grid (slot s: dense id ids[s], 8 values at
// values[8 s ..]); `sparse` selects the source.  Counting and emission use a
// reduce-then-scan over fixed blocks of cells so the output is ordered.

#include <type_traits>

#include "common.h"
#include "its_dc.h"
#include "csg.h"

namespace iso {

namespace {

constexpr int kBlock = 256;

struct CellSrc {
    const float *values;
    const int64_t *ids;   // sparse: sorted dense cell ids (nullptr for dense)
    int64_t n;            // number of cells (dense: (X-1)(Y-1)(Z-1); sparse: slots)
    int64_t X, Y, Z;
    float amin[3], amax[3];
    const CsgProgram *csg;   // dense grid-mode dual contouring: program field (values unused)
    int csg_normals;         // 1: analytic program normals
};

__device__ __forceinline__ int64_t cell_dense_id(const CellSrc &s, int64_t i) {
    return s.ids ? s.ids[i] : i;
}

__device__ __forceinline__ void cell_xyz(const CellSrc &s, int64_t id, int64_t &x, int64_t &y, int64_t &z) {
    int64_t cz = s.Z - 1, cy = s.Y - 1;
    if ((uint64_t) id <= 0xffffffffull) {   // 32-bit division when the id fits
        const uint32_t i32 = (uint32_t) id, cz32 = (uint32_t) cz, cy32 = (uint32_t) cy;
        const uint32_t r = i32 / cz32;
        z = i32 - r * cz32;
        y = r % cy32;
        x = r / cy32;
        return;
    }
    const int64_t r = div_fast(id, cz);
    z = id - r * cz;
    x = div_fast(r, cy);
    y = r - x * cy;
}

__device__ __forceinline__ void cell_values(const CellSrc &s, int64_t i, int64_t x, int64_t y, int64_t z,
                                            float cv[8]) {
    if (s.ids) {
        const float *v = s.values + 8 * i;
#pragma unroll
        for (int k = 0; k < 8; ++k) cv[k] = v[k];
    } else {
        const float *r00 = s.values + (x * s.Y + y) * s.Z + z;
        const float *r01 = r00 + s.Z;
        const float *r10 = r00 + s.Y * s.Z;
        const float *r11 = r10 + s.Z;
        cv[0] = r00[0];
        cv[1] = r00[1];
        cv[2] = r01[0];
        cv[3] = r01[1];
        cv[4] = r10[0];
        cv[5] = r10[1];
        cv[6] = r11[0];
        cv[7] = r11[1];
    }
}

__device__ __forceinline__ int case_bits(const float cv[8], float level) {
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) c |= inside(cv[k], level) << k;
    return c;
}

__device__ __forceinline__ int crossing12(int cs) {
    int st = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e)
        if (((cs >> edge_a(e)) & 1) != ((cs >> edge_b(e)) & 1)) st |= 1 << e;
    return st;
}

// Block-wide exclusive scan of two int64 counters (kBlock threads).
__device__ __forceinline__ void block_scan2(int64_t a, int64_t b, int64_t &ea, int64_t &eb, int64_t &ta,
                                            int64_t &tb) {
    __shared__ int64_t sa[kBlock / 64], sb[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t ia = a, ib = b;
    for (int o = 1; o < 64; o <<= 1) {
        int64_t ya = __shfl_up(ia, o), yb = __shfl_up(ib, o);
        if (lane >= o) {
            ia += ya;
            ib += yb;
        }
    }
    if (lane == 63) {
        sa[wave] = ia;
        sb[wave] = ib;
    }
    __syncthreads();
    int64_t wa = 0, wb = 0;
    ta = tb = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        if (w < wave) {
            wa += sa[w];
            wb += sb[w];
        }
        ta += sa[w];
        tb += sb[w];
    }
    ea = wa + ia - a;
    eb = wb + ib - b;
    __syncthreads();
}

// ---- intersections ----------------------------------------------------------

// Cells per thread in the intersection passes (consecutive ids, so a
// thread's entries are contiguous in the reference's entry order).
constexpr int kCpt = 8;

// Next cell in linear id order (dense grids): z fastest, then y, then x.
__device__ __forceinline__ void next_cell(const CellSrc &s, int64_t &x, int64_t &y, int64_t &z) {
    if (++z == s.Z - 1) {
        z = 0;
        if (++y == s.Y - 1) {
            y = 0;
            ++x;
        }
    }
}

// Cases of a thread's kCpt consecutive cells (8 bits each; 0 past the end),
// with the active-cell and intersection counts.  Dense grids decode the
// first cell's coordinates once and step; sparse grids decode each id.
__device__ __forceinline__ uint64_t thread_cases(const CellSrc &s, int64_t i0, float level, int64_t &na,
                                                 int64_t &ni) {
    uint64_t cases = 0;
    na = ni = 0;
    if (i0 >= s.n) return 0;
    int64_t x, y, z;
    cell_xyz(s, cell_dense_id(s, i0), x, y, z);
    for (int j = 0; j < kCpt; ++j) {
        const int64_t i = i0 + j;
        if (i >= s.n) break;
        if (s.ids && j > 0) cell_xyz(s, s.ids[i], x, y, z);
        float cv[8];
        cell_values(s, i, x, y, z, cv);
        const int cs = case_bits(cv, level);
        cases |= (uint64_t) cs << (8 * j);
        if (cs != 0 && cs != 255) {
            na += 1;
            ni += __popc(crossing12(cs));
        }
        if (!s.ids) next_cell(s, x, y, z);
    }
    return cases;
}

__global__ void k_its_count(CellSrc s, float level, int64_t *bsum) {
    const int64_t i0 = ((int64_t) blockIdx.x * kBlock + threadIdx.x) * kCpt;
    int64_t na, ni;
    (void) thread_cases(s, i0, level, na, ni);
    int64_t ea, ei, ta, ti;
    block_scan2(na, ni, ea, ei, ta, ti);
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = ta;
        bsum[2 * blockIdx.x + 1] = ti;
    }
}

__global__ void k_its_emit(CellSrc s, float level, const int64_t *bsum_pre, ItsOut o) {
    const int64_t i0 = ((int64_t) blockIdx.x * kBlock + threadIdx.x) * kCpt;
    int64_t na, ni;
    const uint64_t cases = thread_cases(s, i0, level, na, ni);
    int64_t ea, ei, ta, ti;
    block_scan2(na, ni, ea, ei, ta, ti);
    if (!na) return;
    int64_t a = bsum_pre[2 * blockIdx.x] + ea;
    int64_t e = bsum_pre[2 * blockIdx.x + 1] + ei;
    for (int j = 0; j < kCpt; ++j) {
        const int64_t i = i0 + j;
        if (i >= s.n) break;
        const int cs = (int) ((cases >> (8 * j)) & 255);
        if (cs == 0 || cs == 255) continue;
        int64_t id = cell_dense_id(s, i), x, y, z;
        cell_xyz(s, id, x, y, z);
        float cv[8];
        cell_values(s, i, x, y, z, cv);
        o.cell_indices[a] = s.ids ? i : id;
        o.cell_dense[a] = id;
        o.cell_offsets[a] = e;
        float lo[3] = {grid_coord(x, s.X, s.amin[0], s.amax[0]), grid_coord(y, s.Y, s.amin[1], s.amax[1]),
                       grid_coord(z, s.Z, s.amin[2], s.amax[2])};
        float hi[3] = {grid_coord(x + 1, s.X, s.amin[0], s.amax[0]),
                       grid_coord(y + 1, s.Y, s.amin[1], s.amax[1]),
                       grid_coord(z + 1, s.Z, s.amin[2], s.amax[2])};
        const int st = crossing12(cs);
        for (int k = 0; k < 12; ++k) {
            if (!((st >> k) & 1)) continue;
            int ca = edge_a(k), cb = edge_b(k);
            float t = edge_t(cv[ca], cv[cb], level);
            float ax = (ca >> 2) ? hi[0] : lo[0], ay = ((ca >> 1) & 1) ? hi[1] : lo[1], az = (ca & 1) ? hi[2] : lo[2];
            float bx = (cb >> 2) ? hi[0] : lo[0], by = ((cb >> 1) & 1) ? hi[1] : lo[1], bz = (cb & 1) ? hi[2] : lo[2];
            float px = lerp1(t, ax, bx), py = lerp1(t, ay, by), pz = lerp1(t, az, bz);
            o.points[3 * e + 0] = px;
            o.points[3 * e + 1] = py;
            o.points[3 * e + 2] = pz;
            o.edge_local[e] = (uint8_t) k;
            o.is_out[e] = cv[ca] <= cv[cb] ? 1 : 0;
            o.entry_cell[e] = (int32_t) a;
            if (o.normals) {
                float nx, ny, nz;
                cell_normal(lo, hi, cv, px, py, pz, nx, ny, nz);
                o.normals[3 * e + 0] = nx;
                o.normals[3 * e + 1] = ny;
                o.normals[3 * e + 2] = nz;
            }
            ++e;
        }
        ++a;
    }
}

// Normals for an existing intersection set (compute_intersection_normals,
// src/its.cu:271-284): one thread per entry.
__global__ void k_its_normals(CellSrc s, ItsIn in, float *normals) {
    int64_t e = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (e >= in.npts) return;
    int32_t a = in.entry_cell[e];
    int64_t id = in.cell_dense[a];
    int64_t x, y, z;
    cell_xyz(s, id, x, y, z);
    float cv[8];
    cell_values(s, s.ids ? in.cell_indices[a] : id, x, y, z, cv);
    float lo[3] = {grid_coord(x, s.X, s.amin[0], s.amax[0]), grid_coord(y, s.Y, s.amin[1], s.amax[1]),
                   grid_coord(z, s.Z, s.amin[2], s.amax[2])};
    float hi[3] = {grid_coord(x + 1, s.X, s.amin[0], s.amax[0]),
                   grid_coord(y + 1, s.Y, s.amin[1], s.amax[1]),
                   grid_coord(z + 1, s.Z, s.amin[2], s.amax[2])};
    float nx, ny, nz;
    cell_normal(lo, hi, cv, in.points[3 * e], in.points[3 * e + 1], in.points[3 * e + 2], nx, ny, nz);
    normals[3 * e + 0] = nx;
    normals[3 * e + 1] = ny;
    normals[3 * e + 2] = nz;
}

// ---- batches of dense grids ----------------------------------------------------
// A batch concatenates the grids' intersection sets (cells, entries) into one
// index space; each cell knows its grid (cell_item, written by the QEF pass
// from the grids' first cells), dense cell ids are made unique across grids
// by a per-grid key base in the active-cell hash, and output vertices are
// renumbered per grid by the grids' vertex bases.  A grid's vertices and
// quads are contiguous in the batch-wide scans (cells and entries are in grid
// order), so its bases are the scans' prefixes at its first cell / entry,
// recorded by the thread that handles that cell / entry.
struct DcItem {
    CellSrc s;                // the grid
    int64_t cell0;            // first active cell of the grid in the batch
    int64_t key_base;         // added to the grid's dense cell ids in the active-cell hash
    const uint64_t *amask;    // lookup tables of the grid's active cells (instead of the hash)
    const int32_t *apre;
    int64_t W;                // words per cell row in the tables
    float level;              // without an intersection set: the grid's iso-level
};
struct DcBatch {
    const DcItem *items;   // nullptr: a single grid (the kernels' own CellSrc)
    int n;
    bool tables;           // every grid has lookup tables (no active-cell hash)
    int32_t *cell_item;    // per active cell: its grid
    int64_t *vbase;        // n + 1: first output vertex of each grid, total
    int64_t *qbase;        // n + 1: first quad of each grid, total
};

// ---- QEF solve ----------------------------------------------------------------

// Rotation parameters of a Jacobi step: t = tan, c = cos, s = sin.  The FP32
// solve uses the hardware reciprocal / square-root approximations (about 1
// ulp): its result is refined against an FP64 residual, which removes the
// solver's error; the FP64 solve stays IEEE.
__device__ __forceinline__ void jacobi_rot(float app, float aqq, float apq, float &c, float &s) {
    const float theta = (aqq - app) * __builtin_amdgcn_rcpf(2.0f * apq);
    const float t = (theta >= 0.0f ? 1.0f : -1.0f) *
                    __builtin_amdgcn_rcpf(fabsf(theta) + __builtin_amdgcn_sqrtf(theta * theta + 1.0f));
    c = __builtin_amdgcn_rsqf(t * t + 1.0f);
    s = t * c;
}
__device__ __forceinline__ void jacobi_rot(double app, double aqq, double apq, double &c, double &s) {
    const double theta = (aqq - app) / (2.0 * apq);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    c = 1.0 / sqrt(t * t + 1.0);
    s = t * c;
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations (in registers).
template <typename T>
__device__ __forceinline__ void jacobi_sym3(T a[3][3], T w[3], T v[3][3], int sweeps) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? T(1) : T(0);
    for (int sw = 0; sw < sweeps; ++sw) {
        T off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        if (off == T(0)) break;
#pragma unroll
        for (int pq = 0; pq < 3; ++pq) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            T apq = a[p][q];
            if (apq == T(0)) continue;
            T c, s;
            jacobi_rot(a[p][p], a[q][q], apq, c, s);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                T akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                T apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                T vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) w[i] = a[i][i];
}

// Truncated pseudo-inverse x = sum_j (v_j . b / w_j) v_j over |w_j| > tol*max|w|
// (the reference's rule, src/batched_la.cu:160-175).  Returns cond = max/min |w|.
template <typename T>
__device__ __forceinline__ T pinv3_apply(const T w[3], const T v[3][3], const T b[3], T tol, T x[3]) {
    T wmax = fmax(fabs(w[0]), fmax(fabs(w[1]), fabs(w[2])));
    T wmin = fmin(fabs(w[0]), fmin(fabs(w[1]), fabs(w[2])));
    T thr = tol * wmax;
    x[0] = x[1] = x[2] = T(0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        if (!(fabs(w[j]) > thr)) continue;
        T proj = v[0][j] * b[0] + v[1][j] * b[1] + v[2][j] * b[2];
        T c = proj / w[j];
#pragma unroll
        for (int i = 0; i < 3; ++i) x[i] += c * v[i][j];
    }
    return wmin > T(0) ? wmax / wmin : T(INFINITY);
}

template <typename T>
__device__ __forceinline__ T pinv3(T A[3][3], const T b[3], T tol, T x[3], int sweeps) {
    T w[3], v[3][3];
    jacobi_sym3(A, w, v, sweeps);
    return pinv3_apply(w, v, b, tol, x);
}

// Entries of one cell read from an intersection set.
struct ItsEdges {
    const float *pts, *nrm;
    int64_t lo;
    int n;
    __device__ __forceinline__ void point(int i, float p[3]) const {
        for (int k = 0; k < 3; ++k) p[k] = pts[3 * (lo + i) + k];
    }
    __device__ __forceinline__ void point_normal(int i, float p[3], float nv[3]) const {
        for (int k = 0; k < 3; ++k) {
            p[k] = pts[3 * (lo + i) + k];
            nv[k] = nrm[3 * (lo + i) + k];
        }
    }
};

// Entries of one cell recomputed from its 8 corner values with the
// intersection pass's formulas (get_its_op, src/its.cu:37-90; trilinear
// normals, src/its.cu:186-268) in the same edge order: the same floats as an
// intersection set of the grid with computed normals holds.
struct GridEdges {
    float cv[8], lo[3], hi[3], level;
    int8_t e[12];
    int n;
    __device__ __forceinline__ void point(int i, float p[3]) const {
        const int ca = edge_a(e[i]), cb = edge_b(e[i]);
        const float t = edge_t(cv[ca], cv[cb], level);
        p[0] = lerp1(t, (ca >> 2) ? hi[0] : lo[0], (cb >> 2) ? hi[0] : lo[0]);
        p[1] = lerp1(t, ((ca >> 1) & 1) ? hi[1] : lo[1], ((cb >> 1) & 1) ? hi[1] : lo[1]);
        p[2] = lerp1(t, (ca & 1) ? hi[2] : lo[2], (cb & 1) ? hi[2] : lo[2]);
    }
    __device__ __forceinline__ void point_normal(int i, float p[3], float nv[3]) const {
        point(i, p);
        cell_normal(lo, hi, cv, p[0], p[1], p[2], nv[0], nv[1], nv[2]);
    }
};

// Entries of one cell of a program field: GridEdges' points with the
// program's analytic unit gradient at each point as its normal (the
// normals get_sdf_normal gives at the intersection points,
// src/isoext/sdf.py:39-50).
struct CsgEdges : GridEdges {
    const CsgProgram *prog;
    __device__ __forceinline__ void point_normal(int i, float p[3], float nv[3]) const {
        point(i, p);
        float g[3];
        csg_value_grad<false>(*prog, p[0], p[1], p[2], g);
        csg_normalize(g, nv);
    }
};

// The QEF solve of one cell over its n >= 1 entries: FP32, centred at the
// mass point and scaled by 1/h, one FP64-residual refinement step, FP64
// absolute re-solve when ill-conditioned; clipped to the cell (fix_dual_v_op,
// src/dc.cu:82-99).  Returns the condition number; on_box when the clipped
// vertex lies on the cell's box.
template <class Ent>
__device__ float qef_cell(const Ent &en, const float blo[3], const float bhi[3], float reg, float svd_tol,
                          float cond_fp64, unsigned long long *n_fp64, float out[3], bool &on_box) {
    float c[3] = {0.f, 0.f, 0.f};
    for (int j = 0; j < en.n; ++j) {
        float p[3];
        en.point(j, p);
        for (int k = 0; k < 3; ++k) c[k] += p[k];
    }
    const float inv_n = 1.0f / (float) en.n;
    for (int k = 0; k < 3; ++k) c[k] *= inv_n;
    const float h = fmaxf(bhi[0] - blo[0], fmaxf(bhi[1] - blo[1], bhi[2] - blo[2]));
    const float ih = 1.0f / h;
    // A = sum n n^T + reg I ; r = sum n (n . (p - c)) / h   (centred, scaled)
    float A[3][3] = {{reg, 0.f, 0.f}, {0.f, reg, 0.f}, {0.f, 0.f, reg}};
    float r[3] = {0.f, 0.f, 0.f};
    for (int j = 0; j < en.n; ++j) {
        float pp[3], n[3];
        en.point_normal(j, pp, n);
        float d = (n[0] * ((pp[0] - c[0]) * ih) + n[1] * ((pp[1] - c[1]) * ih)) + n[2] * ((pp[2] - c[2]) * ih);
        for (int p = 0; p < 3; ++p) {
            for (int q = 0; q < 3; ++q) A[p][q] += n[p] * n[q];
            r[p] += n[p] * d;
        }
    }
    float Ac[3][3];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) Ac[p][q] = A[p][q];
    float xs[3], ew[3], ev[3][3];
    jacobi_sym3<float>(Ac, ew, ev, 6);
    float cond = pinv3_apply<float>(ew, ev, r, svd_tol, xs);
    // one step of iterative refinement: residual r - A x in FP64 (removes the
    // eigen-solver's error), correction solved in FP32 with the same
    // decomposition
    {
        float res[3];
        for (int p = 0; p < 3; ++p)
            res[p] = (float) ((double) r[p] - ((double) A[p][0] * xs[0] + (double) A[p][1] * xs[1] +
                                               (double) A[p][2] * xs[2]));
        float dx[3];
        pinv3_apply<float>(ew, ev, res, svd_tol, dx);
        for (int p = 0; p < 3; ++p) xs[p] += dx[p];
    }
    for (int k = 0; k < 3; ++k) out[k] = c[k] + xs[k] * h;
    if (!(cond <= cond_fp64)) {
        // FP64 re-solve in absolute coordinates exactly as the reference
        // poses it (minimum-norm when singular values are truncated).
        double Ad[3][3] = {{reg, 0, 0}, {0, reg, 0}, {0, 0, reg}}, bd[3] = {0, 0, 0}, pa[3] = {0, 0, 0};
        for (int j = 0; j < en.n; ++j) {
            float pf[3], nf[3];
            en.point_normal(j, pf, nf);
            double n[3] = {nf[0], nf[1], nf[2]};
            double p[3] = {pf[0], pf[1], pf[2]};
            double d = n[0] * p[0] + n[1] * p[1] + n[2] * p[2];
            for (int u = 0; u < 3; ++u) {
                for (int v = 0; v < 3; ++v) Ad[u][v] += n[u] * n[v];
                bd[u] += n[u] * d;
                pa[u] += p[u];
            }
        }
        for (int u = 0; u < 3; ++u) bd[u] += (double) reg * (pa[u] / (double) en.n);
        double xd[3];
        pinv3<double>(Ad, bd, (double) svd_tol, xd, 12);
        for (int k = 0; k < 3; ++k) out[k] = (float) xd[k];
        atomicAdd(n_fp64, 1ull);
    }
    on_box = false;
    for (int k = 0; k < 3; ++k) {
        const float v = fminf(fmaxf(out[k], blo[k]), bhi[k]);
        out[k] = v;
        on_box |= v == blo[k] || v == bhi[k];
    }
    return cond;
}

// Grid of active cell a in a batch (binary search over the grids' first
// cells staged in LDS); writes cell_item[a].
__device__ __forceinline__ int batch_item_of(const DcBatch &bt, const int64_t *starts, int64_t a) {
    int lo_i = 0, hi_i = bt.n;
    while (hi_i - lo_i > 1) {
        const int mid = (lo_i + hi_i) >> 1;
        if (starts[mid] <= a) lo_i = mid;
        else hi_i = mid;
    }
    bt.cell_item[a] = lo_i;
    return lo_i;
}

// Whether the quad of edge k of cell (x, y, z) has all four cells inside the
// grid: across the edge, cells c - 1 and c (corner offset 0) or c and c + 1
// (offset 1) in each of the other two axes.
__device__ __forceinline__ bool edge_interior(int k, int64_t x, int64_t y, int64_t z, const CellSrc &s) {
    const int ax = edge_axis(k), ca = edge_a(k);
    const int off[3] = {ca >> 2, (ca >> 1) & 1, ca & 1};
    const int64_t cc[3] = {x, y, z}, N[3] = {s.X, s.Y, s.Z};
    bool ok = true;
    for (int d = 0; d < 3; ++d)
        if (d != ax) ok = ok && (off[d] ? cc[d] + 3 <= N[d] : cc[d] >= 1);
    return ok;
}

// QEF pass over an intersection set.  Also sets bnd[a] when the clipped
// vertex lies on its cell's box (the only vertices another cell's vertex can
// coincide with), and used[a]: with dense_used (a dense grid whose
// intersection set is the grid's own), whether the cell belongs to a quad --
// one of its crossing edges has all four cells around it inside the grid --
// and qn[a], the quads the cell emits (its canonical crossing edges 0, 3, 8
// whose lower neighbours exist); otherwise used[a] = 0, set later by the
// quad pass.
__global__ void k_dc_qef(const ItsIn in, const CellSrc s1, float reg, float svd_tol, float cond_fp64,
                         float *dual_v, float *cond_out, unsigned long long *n_fp64, uint8_t *used, uint8_t *bnd,
                         uint8_t *qn, bool dense_used, DcBatch bt) {
    __shared__ int64_t starts[kMaxDcBatchGrids];
    if (bt.items) {
        for (int i = threadIdx.x; i < bt.n; i += kBlock) starts[i] = bt.items[i].cell0;
        __syncthreads();
    }
    int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (a >= in.ncells) return;
    const int item = bt.items ? batch_item_of(bt, starts, a) : 0;
    const CellSrc &s = bt.items ? bt.items[item].s : s1;
    const int64_t lo = in.cell_offsets[a], hi = in.cell_offsets[a + 1];
    int64_t x, y, z;
    cell_xyz(s, in.cell_dense[a], x, y, z);
    const float blo[3] = {grid_coord(x, s.X, s.amin[0], s.amax[0]), grid_coord(y, s.Y, s.amin[1], s.amax[1]),
                          grid_coord(z, s.Z, s.amin[2], s.amax[2])};
    const float bhi[3] = {grid_coord(x + 1, s.X, s.amin[0], s.amax[0]), grid_coord(y + 1, s.Y, s.amin[1], s.amax[1]),
                          grid_coord(z + 1, s.Z, s.amin[2], s.amax[2])};
    const ItsEdges en{in.points, in.normals, lo, (int) (hi - lo)};
    float out[3];
    bool on_box;
    const float cond = qef_cell(en, blo, bhi, reg, svd_tol, cond_fp64, n_fp64, out, on_box);
    for (int k = 0; k < 3; ++k) dual_v[3 * a + k] = out[k];
    uint8_t u = 0, nq = 0;
    if (dense_used) {
        for (int64_t j = lo; j < hi; ++j) {
            const int k = in.edge_local[j];
            const bool ok = edge_interior(k, x, y, z, s);
            u |= ok ? 1 : 0;
            // canonical edges sit at the cell's lowest corner (offsets 0): this cell emits their quad
            nq += (ok && edge_a(k) == 0) ? 1 : 0;
        }
        qn[a] = nq;
    }
    used[a] = u;
    bnd[a] = on_box ? 1 : 0;
    if (cond_out) cond_out[a] = cond;
}

// QEF pass of a dense grid without an intersection set: each active cell's
// crossing edges, points and normals are recomputed from its corner values.
// Writes dual_v, used, bnd, qn as k_dc_qef with dense_used, and qmask[a]:
// bit b (b = 0, 1, 2 for canonical edges 0, 3, 8) when the edge emits a quad,
// bit b + 3 its orientation (value at the lower corner <= value at the upper).
// PROG: the cells of grids given by an SDF program (corner values evaluated
// at the grid coordinates, as the stored grid holds them; normals trilinear
// or analytic); the other variant takes the stored grids' cells.
template <bool PROG>
__global__ void k_dc_qef_grid(const int64_t *cell_dense, int64_t ncells, const CellSrc s1, float level, float reg,
                              float svd_tol, float cond_fp64, float *dual_v, unsigned long long *n_fp64,
                              uint8_t *used, uint8_t *bnd, uint8_t *qn, uint8_t *qmask, DcBatch bt) {
    __shared__ int64_t starts[kMaxDcBatchGrids];
    if (bt.items) {
        for (int i = threadIdx.x; i < bt.n; i += kBlock) starts[i] = bt.items[i].cell0;
        __syncthreads();
    }
    const int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (a >= ncells) return;
    const int item = bt.items ? batch_item_of(bt, starts, a) : 0;
    const CellSrc &s = bt.items ? bt.items[item].s : s1;
    if ((s.csg != nullptr) != PROG) return;
    if (bt.items) level = bt.items[item].level;
    int64_t x, y, z;
    cell_xyz(s, cell_dense[a], x, y, z);
    std::conditional_t<PROG, CsgEdges, GridEdges> en;
    en.lo[0] = grid_coord(x, s.X, s.amin[0], s.amax[0]);
    en.lo[1] = grid_coord(y, s.Y, s.amin[1], s.amax[1]);
    en.lo[2] = grid_coord(z, s.Z, s.amin[2], s.amax[2]);
    en.hi[0] = grid_coord(x + 1, s.X, s.amin[0], s.amax[0]);
    en.hi[1] = grid_coord(y + 1, s.Y, s.amin[1], s.amax[1]);
    en.hi[2] = grid_coord(z + 1, s.Z, s.amin[2], s.amax[2]);
    if constexpr (PROG) {
        en.prog = s.csg;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            en.cv[q] = csg_value<false>(*s.csg, (q >> 2) ? en.hi[0] : en.lo[0], ((q >> 1) & 1) ? en.hi[1] : en.lo[1],
                                        (q & 1) ? en.hi[2] : en.lo[2]);
    } else {
        const float *r00 = s.values + (x * s.Y + y) * s.Z + z;
        const float *r01 = r00 + s.Z;
        const float *r10 = r00 + s.Y * s.Z;
        const float *r11 = r10 + s.Z;
        en.cv[0] = r00[0];
        en.cv[1] = r00[1];
        en.cv[2] = r01[0];
        en.cv[3] = r01[1];
        en.cv[4] = r10[0];
        en.cv[5] = r10[1];
        en.cv[6] = r11[0];
        en.cv[7] = r11[1];
    }
    en.level = level;
    const int st = crossing12(case_bits(en.cv, level));
    en.n = 0;
    uint8_t u = 0, nq = 0, qm = 0;
    for (int k = 0; k < 12; ++k) {
        if (!((st >> k) & 1)) continue;
        en.e[en.n++] = (int8_t) k;
        const bool ok = edge_interior(k, x, y, z, s);
        u |= ok ? 1 : 0;
        if (ok && edge_a(k) == 0) {
            const int b = k == 0 ? 0 : (k == 3 ? 1 : 2);
            ++nq;
            qm |= (uint8_t) (1u << b);
            if (en.cv[edge_a(k)] <= en.cv[edge_b(k)]) qm |= (uint8_t) (8u << b);
        }
    }
    float out[3];
    bool on_box = false;
    if (en.n > 0) {
        if (PROG && s.csg_normals) qef_cell(en, en.lo, en.hi, reg, svd_tol, cond_fp64, n_fp64, out, on_box);
        else qef_cell(static_cast<const GridEdges &>(en), en.lo, en.hi, reg, svd_tol, cond_fp64, n_fp64, out, on_box);
    } else {
        out[0] = out[1] = out[2] = 0.0f;   // not reached: active cells have a crossing edge
    }
    for (int k = 0; k < 3; ++k) dual_v[3 * a + k] = out[k];
    used[a] = u;
    bnd[a] = on_box ? 1 : 0;
    qn[a] = nq;
    qmask[a] = qm;
}

// ---- quads -------------------------------------------------------------------

// edge_neighbors_table (src/shared_luts.cu:78-82): offsets subtracted from an
// edge's lower point to reach its 4 cells, per axis.
__constant__ const int8_t kNb[3][4][3] = {{{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}},
                                          {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}},
                                          {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}};

__device__ __forceinline__ int64_t lower_bound64(const int64_t *a, int64_t n, int64_t key) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Active-cell hash: dense cell id -> active index (open addressing, linear
// probing, at most half full).  Replaces a binary search per neighbour.
struct ActHash {
    const int64_t *key;
    const int32_t *val;
    int64_t mask;
};

__global__ void k_act_insert(const int64_t *cell_dense, int64_t ncells, int64_t *akey, int32_t *aval,
                             int64_t amask, DcBatch bt) {
    const int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (a >= ncells) return;
    const int64_t key = cell_dense[a] + (bt.items ? bt.items[bt.cell_item[a]].key_base : 0);
    int64_t h = (int64_t) (mix64((uint64_t) key) & (uint64_t) amask);
    while (true) {
        const unsigned long long prev =
            atomicCAS((unsigned long long *) &akey[h], ~0ull, (unsigned long long) key);
        if (prev == ~0ull) {
            aval[h] = (int32_t) a;
            return;
        }
        h = (h + 1) & amask;
    }
}

// Active index of the dense cell id `cid`, or -1.
__device__ __forceinline__ int64_t active_of(const ActHash &t, int64_t cid) {
    int64_t h = (int64_t) (mix64((uint64_t) cid) & (uint64_t) t.mask);
    while (true) {
        const int64_t k = t.key[h];
        if (k == cid) return t.val[h];
        if (k == -1) return -1;
        h = (h + 1) & t.mask;
    }
}

// Quad of entry e, if e is the canonical copy of an interior crossing edge:
// the copy held by the cell whose min corner owns the edge (local edges 0,
// 3, 8).  These copies are in (p0, p1) order, the reference's quad order.
// The 4 cells around a crossing edge all hold it, so on a dense grid they
// are all active when they exist: counting there needs only the bounds
// (lookups = false; a sparse grid may lack some of them, so it always looks
// them up).  The entry's own cell is q[0] (its active index is
// entry_cell[e]); the other three come from the active-cell hash.
// Active index of active cell (x, y, z) of a dense grid from its lookup tables.
__device__ __forceinline__ int64_t active_of_tab(const DcItem &it, int64_t x, int64_t y, int64_t z) {
    const int64_t t = (x * (it.s.Y - 1) + y) * it.W + (z >> 6);
    return (int64_t) it.apre[t] + __popcll(it.amask[t] & ((1ull << (z & 63)) - 1ull));
}

__device__ bool entry_quad(const ItsIn &in, const CellSrc &s, const ActHash &ah, int64_t e, int64_t q[4],
                           bool lookups, int64_t key_base = 0, const DcItem *tab = nullptr) {
    int k = in.edge_local[e];
    if (k != 0 && k != 3 && k != 8) return false;
    const int axis = edge_axis(k);
    const int64_t own = in.entry_cell[e];
    int64_t x, y, z;
    cell_xyz(s, in.cell_dense[own], x, y, z);
    // offsets subtracted are 0 or 1: in bounds iff the lower cells exist
    if ((axis != 0 && x < 1) || (axis != 1 && y < 1) || (axis != 2 && z < 1)) return false;
    if (!lookups && !s.ids) return true;
    q[0] = own;
    for (int j = 1; j < 4; ++j) {
        int64_t cx = x - kNb[axis][j][0], cy = y - kNb[axis][j][1], cz = z - kNb[axis][j][2];
        q[j] = tab ? active_of_tab(*tab, cx, cy, cz) : active_of(ah, (cx * (s.Y - 1) + cy) * (s.Z - 1) + cz + key_base);
        if (q[j] < 0) return false;
    }
    if (!in.is_out[e]) {
        int64_t t0 = q[0], t1 = q[1];
        q[0] = q[3];
        q[1] = q[2];
        q[2] = t1;
        q[3] = t0;
    }
    return true;
}

__global__ void k_quad_count(ItsIn in, CellSrc s1, ActHash ah, int64_t *bsum, DcBatch bt) {
    int64_t e = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    int64_t nq = 0;
    int64_t q[4];
    if (e < in.npts) {
        const DcItem *it = bt.items ? &bt.items[bt.cell_item[in.entry_cell[e]]] : nullptr;
        if (entry_quad(in, it ? it->s : s1, ah, e, q, false, it ? it->key_base : 0)) nq = 1;
    }
    int64_t eq, dummy, tq, td;
    block_scan2(nq, 0, eq, dummy, tq, td);
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = tq;
        bsum[2 * blockIdx.x + 1] = 0;
    }
}

__device__ __forceinline__ float dist3(const float *a, const float *b) {
    float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    float s = x * x + y * y;
    s = s + z * z;
    return sqrtf(s);
}

// get_triangles_op (src/dc.cu:101-157): split along the shorter diagonal.
// Vertex id of cell a's dual vertex after the weld.
__device__ __forceinline__ int32_t welded_id(const uint8_t *bnd, const int32_t *rep_slot, const int32_t *slot_min,
                                             const int32_t *final_id, int32_t a) {
    return final_id[bnd[a] ? slot_min[rep_slot[a]] : a];
}

// Quad corners go to tri_cells (cells; used[] marked here), or, when the
// weld ran first (wd.final_id != nullptr), straight to out_f as vertex ids.
struct WeldIds {
    const uint8_t *bnd;
    const int32_t *rep_slot, *slot_min, *final_id;
    int32_t *out_f;
};

__global__ void k_quad_emit(ItsIn in, CellSrc s1, ActHash ah, const int64_t *bsum_pre, const float *dual_v,
                            int32_t *tri_cells, uint8_t *used, WeldIds wd, DcBatch bt) {
    int64_t e = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    int64_t nq = 0;
    int64_t q[4] = {0, 0, 0, 0};
    int item = 0;
    if (e < in.npts) {
        if (bt.items) {
            item = bt.cell_item[in.entry_cell[e]];
            const DcItem *it = &bt.items[item];
            if (entry_quad(in, it->s, ah, e, q, true, it->key_base, it->amask ? it : nullptr)) nq = 1;
        } else {
            const DcItem one{s1, 0, 0, in.amask_tab, in.apre_tab, (s1.Z + 63) / 64, 0.0f};
            if (entry_quad(in, s1, ah, e, q, true, 0, in.amask_tab && in.apre_tab && !s1.ids ? &one : nullptr)) nq = 1;
        }
    }
    int64_t eq, dummy, tq, td;
    block_scan2(nq, 0, eq, dummy, tq, td);
    const int64_t slot = bsum_pre[2 * blockIdx.x] + eq;
    // batch: the grid's first entry records the grid's first quad
    if (bt.items && e < in.npts && e == in.cell_offsets[bt.items[item].cell0]) bt.qbase[item] = slot;
    if (!nq) return;
    const float *v0 = dual_v + 3 * q[0], *v1 = dual_v + 3 * q[1], *v2 = dual_v + 3 * q[2], *v3 = dual_v + 3 * q[3];
    int32_t t[6];
    if (dist3(v0, v2) > dist3(v1, v3)) {
        int32_t u[6] = {(int32_t) q[1], (int32_t) q[3], (int32_t) q[0], (int32_t) q[3], (int32_t) q[1], (int32_t) q[2]};
        for (int k = 0; k < 6; ++k) t[k] = u[k];
    } else {
        int32_t u[6] = {(int32_t) q[2], (int32_t) q[0], (int32_t) q[1], (int32_t) q[0], (int32_t) q[2], (int32_t) q[3]};
        for (int k = 0; k < 6; ++k) t[k] = u[k];
    }
    if (wd.final_id) {
        for (int k = 0; k < 6; ++k) wd.out_f[6 * slot + k] = welded_id(wd.bnd, wd.rep_slot, wd.slot_min, wd.final_id, t[k]);
        return;
    }
    for (int k = 0; k < 6; ++k) tri_cells[6 * slot + k] = t[k];
    for (int k = 0; k < 4; ++k) used[q[k]] = 1;
}

// ---- weld of dual vertices by position ------------------------------------------
// vertex_welding (src/dc.cu:205-208) merges equal positions.  Vertices are
// clipped to their cells' boxes, and cell boxes meet only on their faces, so
// two cells' vertices can be equal only when both lie on their boxes (bnd):
// only those go through the position hash; every other used vertex is its
// own representative.

__global__ void k_pos_insert(const float *dv, const uint8_t *used, const uint8_t *bnd, int64_t n, int32_t *slots,
                             int64_t mask, int32_t *rep, const int32_t *cell_item) {
    int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (a >= n || !used[a] || !bnd[a]) return;
    float x = dv[3 * a], y = dv[3 * a + 1], z = dv[3 * a + 2];
    // batches: vertices of different grids never weld (the grid joins the hash and the test)
    const int32_t item = cell_item ? cell_item[a] : 0;
    uint64_t h = (f3_hash(x, y, z) ^ (uint64_t) mix64((uint64_t) item)) & (uint64_t) mask;
    for (int64_t probe = 0; probe <= mask; ++probe) {
        int64_t sl = (int64_t) ((h + probe) & (uint64_t) mask);
        int32_t cur = slots[sl];
        if (cur < 0) {
            int32_t prev = atomicCAS(&slots[sl], -1, (int32_t) a);
            if (prev < 0) {
                rep[a] = (int32_t) sl;
                return;
            }
            cur = prev;
        }
        if (dv[3 * cur] == x && dv[3 * cur + 1] == y && dv[3 * cur + 2] == z &&
            (!cell_item || cell_item[cur] == item)) {
            rep[a] = (int32_t) sl;
            return;
        }
    }
}

__global__ void k_slot_min(const uint8_t *used, const uint8_t *bnd, int64_t n, const int32_t *rep_slot,
                           int32_t *slot_min) {
    int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (a >= n || !used[a] || !bnd[a]) return;
    atomicMin(&slot_min[rep_slot[a]], (int32_t) a);
}

// Used vertex a is kept when it represents its position.
__device__ __forceinline__ bool kept_vertex(const uint8_t *used, const uint8_t *bnd, const int32_t *rep_slot,
                                            const int32_t *slot_min, int64_t a) {
    return used[a] && (!bnd[a] || slot_min[rep_slot[a]] == (int32_t) a);
}

// Block sums of (kept vertices, quads per cell from qn, or 0 without it).
__global__ void k_keep_flags(const uint8_t *used, const uint8_t *bnd, int64_t n, const int32_t *rep_slot,
                             const int32_t *slot_min, const uint8_t *qn, int64_t *bsum) {
    int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    int64_t k = (a < n && kept_vertex(used, bnd, rep_slot, slot_min, a)) ? 1 : 0;
    int64_t q = (qn && a < n) ? qn[a] : 0;
    int64_t ek, eq, tk, tq;
    block_scan2(k, q, ek, eq, tk, tq);
    if (threadIdx.x == 0) {
        bsum[2 * blockIdx.x] = tk;
        bsum[2 * blockIdx.x + 1] = tq;
    }
}

__global__ void k_keep_emit(const float *dv, const uint8_t *used, const uint8_t *bnd, int64_t n,
                            const int32_t *rep_slot, const int32_t *slot_min, const int64_t *bsum_pre,
                            int32_t *final_id, float *out_v, int32_t *out_vcell, DcBatch bt) {
    int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    int64_t k = (a < n && kept_vertex(used, bnd, rep_slot, slot_min, a)) ? 1 : 0;
    int64_t ek, d, tk, td;
    block_scan2(k, 0, ek, d, tk, td);
    if (!k) return;
    const int64_t id = bsum_pre[2 * blockIdx.x] + ek;
    // batch: face corners use grid-local vertex ids
    final_id[a] = (int32_t) (bt.items ? id - bt.vbase[bt.cell_item[a]] : id);
    out_v[3 * id + 0] = dv[3 * a + 0];
    out_v[3 * id + 1] = dv[3 * a + 1];
    out_v[3 * id + 2] = dv[3 * a + 2];
    if (out_vcell) out_vcell[id] = (int32_t) a;
}

// Face corners: 6 per quad; the quad count is read from the device.
__global__ void k_remap_faces(const int32_t *tri_cells, const int64_t *nquads, const uint8_t *bnd,
                              const int32_t *rep_slot, const int32_t *slot_min, const int32_t *final_id,
                              int32_t *out_f) {
    int64_t i = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (i >= 6 * *nquads) return;
    out_f[i] = welded_id(bnd, rep_slot, slot_min, final_id, tri_cells[i]);
}

// Batch: first output vertex of each grid = kept vertices before its first
// cell (block prefix of the keep scan + the kept cells of that block before
// it); one workgroup per grid, vbase[n] = the total.
// With qn (quads counted per cell), the first quads too: bsum_pre's second
// component and tot[3] hold the quad scan.
__global__ void k_item_vbase(DcBatch bt, const uint8_t *used, const uint8_t *bnd, const int32_t *rep_slot,
                             const int32_t *slot_min, const uint8_t *qn, const int64_t *bsum_pre, int64_t ncells,
                             const int64_t *tot) {
    __shared__ int part[kBlock / 64], partq[kBlock / 64];
    const int i = blockIdx.x;
    if (i == 0 && threadIdx.x == 0) {
        bt.vbase[bt.n] = tot[2];
        if (qn) bt.qbase[bt.n] = tot[3];
    }
    const int64_t c0 = bt.items[i].cell0;
    if (c0 >= ncells) {
        if (threadIdx.x == 0) {
            bt.vbase[i] = tot[2];
            if (qn) bt.qbase[i] = tot[3];
        }
        return;
    }
    const int64_t blk = c0 / kBlock, a = blk * kBlock + threadIdx.x;
    int k = (a < c0 && kept_vertex(used, bnd, rep_slot, slot_min, a)) ? 1 : 0;
    int q = (qn && a < c0) ? qn[a] : 0;
    for (int o = 32; o > 0; o >>= 1) {
        k += __shfl_down(k, o);
        q += __shfl_down(q, o);
    }
    if ((threadIdx.x & 63) == 0) {
        part[threadIdx.x >> 6] = k;
        partq[threadIdx.x >> 6] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0, tq = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            t += part[w];
            tq += partq[w];
        }
        bt.vbase[i] = bsum_pre[2 * blk] + t;
        if (qn) bt.qbase[i] = bsum_pre[2 * blk + 1] + tq;
    }
}

// Quads of dense sets, one thread per active cell: the cell's canonical
// entries in entry order (the per-entry kernel's order), at the cell's
// offset in the quad scan (second component of bsum_pre); corners are
// written as welded vertex ids.
__global__ void k_quad_emit_cells(ItsIn in, CellSrc s1, const uint8_t *qn, const int64_t *bsum_pre,
                                  const float *dual_v, WeldIds wd, DcBatch bt) {
    const int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    const int64_t nq = a < in.ncells ? qn[a] : 0;
    int64_t ed, eq, td, tq;
    block_scan2(0, nq, ed, eq, td, tq);
    if (!nq) return;
    int64_t slot = bsum_pre[2 * blockIdx.x + 1] + eq;
    const DcItem *it = nullptr;
    DcItem one;
    if (bt.items) {
        it = &bt.items[bt.cell_item[a]];
    } else {
        one = DcItem{s1, 0, 0, in.amask_tab, in.apre_tab, (s1.Z + 63) / 64, 0.0f};
        it = &one;
    }
    const ActHash ah{nullptr, nullptr, 0};
    const int64_t lo = in.cell_offsets[a], hi = in.cell_offsets[a + 1];
    for (int64_t e = lo; e < hi; ++e) {
        int64_t q[4];
        if (!entry_quad(in, it->s, ah, e, q, true, it->key_base, it)) continue;
        const float *v0 = dual_v + 3 * q[0], *v1 = dual_v + 3 * q[1], *v2 = dual_v + 3 * q[2],
                    *v3 = dual_v + 3 * q[3];
        int32_t t[6];
        if (dist3(v0, v2) > dist3(v1, v3)) {
            const int32_t u[6] = {(int32_t) q[1], (int32_t) q[3], (int32_t) q[0],
                                  (int32_t) q[3], (int32_t) q[1], (int32_t) q[2]};
            for (int k = 0; k < 6; ++k) t[k] = u[k];
        } else {
            const int32_t u[6] = {(int32_t) q[2], (int32_t) q[0], (int32_t) q[1],
                                  (int32_t) q[0], (int32_t) q[2], (int32_t) q[3]};
            for (int k = 0; k < 6; ++k) t[k] = u[k];
        }
        for (int k = 0; k < 6; ++k)
            wd.out_f[6 * slot + k] = welded_id(wd.bnd, wd.rep_slot, wd.slot_min, wd.final_id, t[k]);
        ++slot;
    }
}

// Quads of a dense grid without an intersection set, one thread per active
// cell: its canonical edges from qmask in edge order (0, 3, 8: the entry
// order of the intersection set), neighbours from the lookup tables, the
// reference's orientation and shorter-diagonal split, corners as welded
// vertex ids at the cell's offset in the quad scan.
__global__ void k_quad_emit_grid(const int64_t *cell_dense, int64_t ncells, CellSrc s1, const uint64_t *amask_tab,
                                 const int32_t *apre_tab, const uint8_t *qn, const uint8_t *qmask,
                                 const int64_t *bsum_pre, const float *dual_v, WeldIds wd, DcBatch bt) {
    const int64_t a = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    const int64_t nq = a < ncells ? qn[a] : 0;
    int64_t ed, eq, td, tq;
    block_scan2(0, nq, ed, eq, td, tq);
    if (!nq) return;
    int64_t slot = bsum_pre[2 * blockIdx.x + 1] + eq;
    DcItem one;
    const DcItem *it;
    if (bt.items) {
        it = &bt.items[bt.cell_item[a]];
    } else {
        one = DcItem{s1, 0, 0, amask_tab, apre_tab, (s1.Z + 63) / 64, 0.0f};
        it = &one;
    }
    int64_t x, y, z;
    cell_xyz(it->s, cell_dense[a], x, y, z);
    const int qm = qmask[a];
    for (int b = 0; b < 3; ++b) {
        if (!((qm >> b) & 1)) continue;
        const int axis = edge_axis(b == 0 ? 0 : (b == 1 ? 3 : 8));
        int64_t q[4];
        q[0] = a;
        for (int j = 1; j < 4; ++j)
            q[j] = active_of_tab(*it, x - kNb[axis][j][0], y - kNb[axis][j][1], z - kNb[axis][j][2]);
        if (!((qm >> (b + 3)) & 1)) {
            const int64_t t0 = q[0], t1 = q[1];
            q[0] = q[3];
            q[1] = q[2];
            q[2] = t1;
            q[3] = t0;
        }
        const float *v0 = dual_v + 3 * q[0], *v1 = dual_v + 3 * q[1], *v2 = dual_v + 3 * q[2],
                    *v3 = dual_v + 3 * q[3];
        int32_t t[6];
        if (dist3(v0, v2) > dist3(v1, v3)) {
            const int32_t u[6] = {(int32_t) q[1], (int32_t) q[3], (int32_t) q[0],
                                  (int32_t) q[3], (int32_t) q[1], (int32_t) q[2]};
            for (int k = 0; k < 6; ++k) t[k] = u[k];
        } else {
            const int32_t u[6] = {(int32_t) q[2], (int32_t) q[0], (int32_t) q[1],
                                  (int32_t) q[0], (int32_t) q[2], (int32_t) q[3]};
            for (int k = 0; k < 6; ++k) t[k] = u[k];
        }
        for (int k = 0; k < 6; ++k)
            wd.out_f[6 * slot + k] = welded_id(wd.bnd, wd.rep_slot, wd.slot_min, wd.final_id, t[k]);
        ++slot;
    }
}

__global__ void k_copy_i64(int64_t *dst, const int64_t *src) {
    if (threadIdx.x == 0) *dst = *src;
}

// First quads of grids without cells (no thread recorded them: they equal
// the next grid's) and the total; tot = {quads, ...}.
__global__ void k_item_qbase(DcBatch bt, int64_t ncells, const int64_t *tot) {
    if (threadIdx.x == 0) {
        bt.qbase[bt.n] = tot[0];
        for (int i = bt.n - 1; i >= 0; --i) {
            const int64_t next = i + 1 < bt.n ? bt.items[i + 1].cell0 : ncells;
            if (bt.items[i].cell0 == next) bt.qbase[i] = bt.qbase[i + 1];
        }
    }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t) blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t) gridDim.x * kBlock) p[i] = v;
}
__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    for (int64_t i = (int64_t) blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t) gridDim.x * kBlock) p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
    for (int64_t i = (int64_t) blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t) gridDim.x * kBlock) p[i] = v;
}

// exclusive scan of block sums (pairs), single block, in place; totals out
// Parallel exclusive scan of the block-sum pairs for large inputs: chunk
// sums, one workgroup over the chunk sums, then each chunk rescanned with its
// prefix (kScanChunk pairs per workgroup).
constexpr int kScanChunk = 1024;

template <int NT>
__device__ __forceinline__ void block_scan2_nt(int64_t a, int64_t b, int64_t &ea, int64_t &eb, int64_t &ta,
                                               int64_t &tb) {
    __shared__ int64_t sa[NT / 64], sb[NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t ia = a, ib = b;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t ya = __shfl_up(ia, o), yb = __shfl_up(ib, o);
        if (lane >= o) {
            ia += ya;
            ib += yb;
        }
    }
    if (lane == 63) {
        sa[wave] = ia;
        sb[wave] = ib;
    }
    __syncthreads();
    int64_t wa = 0, wb = 0;
    ta = tb = 0;
    for (int w = 0; w < NT / 64; ++w) {
        if (w < wave) {
            wa += sa[w];
            wb += sb[w];
        }
        ta += sa[w];
        tb += sb[w];
    }
    ea = wa + ia - a;
    eb = wb + ib - b;
    __syncthreads();
}

__global__ void __launch_bounds__(kScanChunk) k_scan_chunk_sums(const int64_t *bsum, int64_t nb, int64_t *part) {
    const int64_t i = (int64_t) blockIdx.x * kScanChunk + threadIdx.x;
    const int64_t a = i < nb ? bsum[2 * i] : 0, b = i < nb ? bsum[2 * i + 1] : 0;
    int64_t ea, eb, ta, tb;
    block_scan2_nt<kScanChunk>(a, b, ea, eb, ta, tb);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ta;
        part[2 * blockIdx.x + 1] = tb;
    }
}

// Exclusive scan of np chunk sums in one workgroup (loop with a carry); totals to tot.
__global__ void __launch_bounds__(kScanChunk) k_scan_chunk_tops(int64_t *part, int64_t np, int64_t *tot) {
    __shared__ int64_t carry[2];
    if (threadIdx.x == 0) carry[0] = carry[1] = 0;
    __syncthreads();
    for (int64_t base = 0; base < np; base += kScanChunk) {
        const int64_t i = base + threadIdx.x;
        const int64_t a = i < np ? part[2 * i] : 0, b = i < np ? part[2 * i + 1] : 0;
        int64_t ea, eb, ta, tb;
        block_scan2_nt<kScanChunk>(a, b, ea, eb, ta, tb);
        const int64_t ca = carry[0], cb = carry[1];
        if (i < np) {
            part[2 * i] = ca + ea;
            part[2 * i + 1] = cb + eb;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            carry[0] = ca + ta;
            carry[1] = cb + tb;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot[0] = carry[0];
        tot[1] = carry[1];
    }
}

__global__ void __launch_bounds__(kScanChunk) k_scan_chunk_down(int64_t *bsum, int64_t nb, const int64_t *part) {
    const int64_t i = (int64_t) blockIdx.x * kScanChunk + threadIdx.x;
    const int64_t a = i < nb ? bsum[2 * i] : 0, b = i < nb ? bsum[2 * i + 1] : 0;
    int64_t ea, eb, ta, tb;
    block_scan2_nt<kScanChunk>(a, b, ea, eb, ta, tb);
    if (i < nb) {
        bsum[2 * i] = part[2 * blockIdx.x] + ea;
        bsum[2 * i + 1] = part[2 * blockIdx.x + 1] + eb;
    }
}

// exclusive scan of block sums (pairs), single block, in place; totals out
__global__ void k_scan_pairs(int64_t *bsum, int64_t nb, int64_t *tot) {
    __shared__ int64_t sa[1024], sb[1024];
    int64_t per = (nb + 1023) / 1024;
    int64_t lo = threadIdx.x * per, hi = min(lo + per, nb);
    int64_t a = 0, b = 0;
    for (int64_t i = lo; i < hi; ++i) {
        a += bsum[2 * i];
        b += bsum[2 * i + 1];
    }
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t ya = threadIdx.x >= (unsigned) o ? sa[threadIdx.x - o] : 0;
        int64_t yb = threadIdx.x >= (unsigned) o ? sb[threadIdx.x - o] : 0;
        __syncthreads();
        sa[threadIdx.x] += ya;
        sb[threadIdx.x] += yb;
        __syncthreads();
    }
    int64_t ra = sa[threadIdx.x] - a, rb = sb[threadIdx.x] - b;
    for (int64_t i = lo; i < hi; ++i) {
        int64_t x = bsum[2 * i], y = bsum[2 * i + 1];
        bsum[2 * i] = ra;
        bsum[2 * i + 1] = rb;
        ra += x;
        rb += y;
    }
    if (threadIdx.x == 1023) {
        tot[0] = sa[1023];
        tot[1] = sb[1023];
    }
}

// ---- backward -------------------------------------------------------------------
// x = M^+ (sum_i n_i (n_i . p_i) + reg c), M = sum_i n_i n_i^T + reg I,
// c = mean p_i (the reference's regularised QEF, src/dc.cu:14-65, solved by
// src/batched_la.cu:151-178).  With y = M^+ g:
//   dL/dp_i = n_i (n_i . y) + (reg / k) y
//   dL/dn_i = (n_i . (p_i - x)) y + (p_i - x) (n_i . y)
// components of x clipped to the cell (src/dc.cu:82-99) pass no gradient;
// p_i = (1 - t) Pa + t Pb with t = (L - va) / (vb - va) carries dL/dp_i to va, vb.
__global__ void k_dc_backward(const ItsIn in, const CellSrc s, float reg, float svd_tol, float level, int64_t nv,
                              const int32_t *vcell, const float *grad_v, float *grad_values, float *grad_normals) {
    const int64_t v = (int64_t) blockIdx.x * kBlock + threadIdx.x;
    if (v >= nv) return;
    const int64_t a = vcell[v];
    const int64_t lo = in.cell_offsets[a], hi = in.cell_offsets[a + 1];
    const int64_t id = in.cell_dense[a];
    int64_t cx, cy, cz;
    cell_xyz(s, id, cx, cy, cz);
    const double blo[3] = {grid_coord(cx, s.X, s.amin[0], s.amax[0]), grid_coord(cy, s.Y, s.amin[1], s.amax[1]),
                           grid_coord(cz, s.Z, s.amin[2], s.amax[2])};
    const double bhi[3] = {grid_coord(cx + 1, s.X, s.amin[0], s.amax[0]),
                           grid_coord(cy + 1, s.Y, s.amin[1], s.amax[1]),
                           grid_coord(cz + 1, s.Z, s.amin[2], s.amax[2])};
    double M[3][3] = {{reg, 0, 0}, {0, reg, 0}, {0, 0, reg}}, b[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    const double k = (double) (hi - lo);
    for (int64_t j = lo; j < hi; ++j) {
        const double n[3] = {in.normals[3 * j], in.normals[3 * j + 1], in.normals[3 * j + 2]};
        const double p[3] = {in.points[3 * j], in.points[3 * j + 1], in.points[3 * j + 2]};
        const double d = n[0] * p[0] + n[1] * p[1] + n[2] * p[2];
        for (int u = 0; u < 3; ++u) {
            for (int w = 0; w < 3; ++w) M[u][w] += n[u] * n[w];
            b[u] += n[u] * d;
            c[u] += p[u] / k;
        }
    }
    for (int u = 0; u < 3; ++u) b[u] += reg * c[u];
    double Mc[3][3], x[3], y[3];
    for (int u = 0; u < 3; ++u)
        for (int w = 0; w < 3; ++w) Mc[u][w] = M[u][w];
    pinv3<double>(Mc, b, (double) svd_tol, x, 12);
    double g[3] = {grad_v[3 * v], grad_v[3 * v + 1], grad_v[3 * v + 2]};
    for (int u = 0; u < 3; ++u)
        if (x[u] < blo[u] || x[u] > bhi[u]) g[u] = 0.0;
    for (int u = 0; u < 3; ++u)
        for (int w = 0; w < 3; ++w) Mc[u][w] = M[u][w];
    pinv3<double>(Mc, g, (double) svd_tol, y, 12);
    for (int64_t j = lo; j < hi; ++j) {
        const double n[3] = {in.normals[3 * j], in.normals[3 * j + 1], in.normals[3 * j + 2]};
        const double p[3] = {in.points[3 * j], in.points[3 * j + 1], in.points[3 * j + 2]};
        const double ny = n[0] * y[0] + n[1] * y[1] + n[2] * y[2];
        double gp[3];
        for (int u = 0; u < 3; ++u) gp[u] = n[u] * ny + (reg / k) * y[u];
        if (grad_normals) {
            const double r[3] = {p[0] - x[0], p[1] - x[1], p[2] - x[2]};
            const double nr = n[0] * r[0] + n[1] * r[1] + n[2] * r[2];
            for (int u = 0; u < 3; ++u) grad_normals[3 * j + u] = (float) (nr * y[u] + r[u] * ny);
        }
        const int e = in.edge_local[j];
        const int ca = edge_a(e), cb = edge_b(e);
        const int64_t ax = cx + (ca >> 2), ay = cy + ((ca >> 1) & 1), az = cz + (ca & 1);
        const int64_t bx = cx + (cb >> 2), by = cy + ((cb >> 1) & 1), bz = cz + (cb & 1);
        int64_t ia, ib;
        if (s.ids) {
            const int64_t slot = in.cell_indices[a];
            ia = 8 * slot + ca;
            ib = 8 * slot + cb;
        } else {
            ia = (ax * s.Y + ay) * s.Z + az;
            ib = (bx * s.Y + by) * s.Z + bz;
        }
        const double va = s.values[ia], vb = s.values[ib];
        const double dv = vb - va;
        if (!(dv != 0.0)) continue;
        const double Pa[3] = {grid_coord(ax, s.X, s.amin[0], s.amax[0]), grid_coord(ay, s.Y, s.amin[1], s.amax[1]),
                              grid_coord(az, s.Z, s.amin[2], s.amax[2])};
        const double Pb[3] = {grid_coord(bx, s.X, s.amin[0], s.amax[0]), grid_coord(by, s.Y, s.amin[1], s.amax[1]),
                              grid_coord(bz, s.Z, s.amin[2], s.amax[2])};
        const double sd = gp[0] * (Pb[0] - Pa[0]) + gp[1] * (Pb[1] - Pa[1]) + gp[2] * (Pb[2] - Pa[2]);
        const double ga = sd * ((double) level - vb) / (dv * dv), gb = -sd * ((double) level - va) / (dv * dv);
        if (isfinite(ga) && ga != 0.0) atomicAdd(&grad_values[ia], (float) ga);
        if (isfinite(gb) && gb != 0.0) atomicAdd(&grad_values[ib], (float) gb);
    }
}

unsigned nblocks(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

CellSrc make_src(const GridDesc &g) {
    CellSrc s;
    s.values = g.values;
    s.ids = g.sparse_ids;
    s.X = g.X;
    s.Y = g.Y;
    s.Z = g.Z;
    s.n = g.sparse_ids ? g.n_sparse : (g.X - 1) * (g.Y - 1) * (g.Z - 1);
    for (int a = 0; a < 3; ++a) {
        s.amin[a] = g.amin[a];
        s.amax[a] = g.amax[a];
    }
    s.csg = g.csg;
    s.csg_normals = g.csg_normals;
    return s;
}

}   // namespace

#define ISO_TRY(x)                       \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

unsigned its_blocks(int64_t n) { return (unsigned) std::max<int64_t>(1, (n + kBlock * kCpt - 1) / (kBlock * kCpt)); }

hipError_t its_count(const GridDesc &g, float level, int64_t *bsum, int64_t *tot, hipStream_t st) {
    CellSrc s = make_src(g);
    unsigned nb = its_blocks(s.n);
    hipLaunchKernelGGL(k_its_count, dim3(nb), dim3(kBlock), 0, st, s, level, bsum);
    hipLaunchKernelGGL(k_scan_pairs, dim3(1), dim3(1024), 0, st, bsum, (int64_t) nb, tot);
    return hipGetLastError();
}

int64_t its_scratch_pairs(const GridDesc &g) {
    CellSrc s = make_src(g);
    return its_blocks(s.n);
}

hipError_t its_emit(const GridDesc &g, float level, const int64_t *bsum_pre, const ItsOut &o, hipStream_t st) {
    CellSrc s = make_src(g);
    hipLaunchKernelGGL(k_its_emit, dim3(its_blocks(s.n)), dim3(kBlock), 0, st, s, level, bsum_pre, o);
    return hipGetLastError();
}

hipError_t its_normals(const GridDesc &g, const ItsIn &in, float *normals, hipStream_t st) {
    CellSrc s = make_src(g);
    if (in.npts == 0) return hipSuccess;
    hipLaunchKernelGGL(k_its_normals, dim3(nblocks(in.npts)), dim3(kBlock), 0, st, s, in, normals);
    return hipGetLastError();
}

hipError_t fill_i64(int64_t *p, int64_t n, int64_t v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i64, dim3((unsigned) std::min<int64_t>(nblocks(n), 4096)), dim3(kBlock), 0, st, p, n,
                       v);
    return hipGetLastError();
}

// Launches the DC kernels (no synchronisation).  bt.items == nullptr: one
// grid (s); otherwise the batch's grids.
// Exclusive scan of nb block-sum pairs in place, totals to tot; part holds
// dc_scan_part_pairs(nb) pairs.
static void scan_pairs(int64_t *bsum, int64_t nb, int64_t *tot, int64_t *part, hipStream_t st) {
    if (nb <= 4 * kScanChunk || !part) {
        hipLaunchKernelGGL(k_scan_pairs, dim3(1), dim3(1024), 0, st, bsum, nb, tot);
        return;
    }
    const int64_t np = (nb + kScanChunk - 1) / kScanChunk;
    hipLaunchKernelGGL(k_scan_chunk_sums, dim3((unsigned) np), dim3(kScanChunk), 0, st, (const int64_t *) bsum, nb,
                       part);
    hipLaunchKernelGGL(k_scan_chunk_tops, dim3(1), dim3(kScanChunk), 0, st, part, np, tot);
    hipLaunchKernelGGL(k_scan_chunk_down, dim3((unsigned) np), dim3(kScanChunk), 0, st, bsum, nb,
                       (const int64_t *) part);
}

static hipError_t dc_launch(const CellSrc &s, const ItsIn &in, float reg, float svd_tol, float cond_fp64,
                            DcScratch &w, const DcBatch &bt, hipStream_t st) {
    // Dense grids with lookup tables (the intersection set is the grid's
    // own): no active-cell hash; the QEF pass also gives each cell's quad
    // membership and quad count, so the weld runs first and the quads are
    // written per cell as vertex ids, with vertex and quad offsets from one
    // scan.  Otherwise: hash, per-entry quads (marking used[]), weld, face
    // remap.
    const bool tables = bt.items ? bt.tables : (in.amask_tab && in.apre_tab && !s.ids);
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, st, (int64_t *) w.n_fp64, (int64_t) 1, (int64_t) 0);
    hipLaunchKernelGGL(k_dc_qef, dim3(nblocks(in.ncells)), dim3(kBlock), 0, st, in, s, reg, svd_tol, cond_fp64,
                       w.dual_v, w.cond, w.n_fp64, w.used, w.bnd, w.qn, tables, bt);
    const ActHash ah{w.akey, w.aval, w.amask};
    const unsigned nbq = nblocks(in.npts), nbv = nblocks(in.ncells);
    if (!tables) {   // active-cell hash for the quads' neighbour lookups, then the quads
        const unsigned nbh = (unsigned) std::min<int64_t>(nblocks(w.amask + 1), 4096);
        hipLaunchKernelGGL(k_fill_i64, dim3(nbh), dim3(kBlock), 0, st, w.akey, w.amask + 1, (int64_t) -1);
        hipLaunchKernelGGL(k_act_insert, dim3(nblocks(in.ncells)), dim3(kBlock), 0, st, in.cell_dense, in.ncells,
                           w.akey, w.aval, w.amask, bt);
        hipLaunchKernelGGL(k_quad_count, dim3(nbq), dim3(kBlock), 0, st, in, s, ah, w.bsum, bt);
        scan_pairs(w.bsum, (int64_t) nbq, w.tot, w.part, st);
        hipLaunchKernelGGL(k_quad_emit, dim3(nbq), dim3(kBlock), 0, st, in, s, ah, w.bsum, w.dual_v, w.tri_cells,
                           w.used, WeldIds{}, bt);
        if (bt.items) hipLaunchKernelGGL(k_item_qbase, dim3(1), dim3(64), 0, st, bt, in.ncells, (const int64_t *) w.tot);
    }
    // weld dual vertices by position
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(w.hmask + 1)), dim3(kBlock), 0, st, w.hslots, w.hmask + 1, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(w.hmask + 1)), dim3(kBlock), 0, st, w.slot_min, w.hmask + 1,
                       INT32_MAX);
    hipLaunchKernelGGL(k_pos_insert, dim3(nblocks(in.ncells)), dim3(kBlock), 0, st, w.dual_v, w.used, w.bnd,
                       in.ncells, w.hslots, w.hmask, w.rep_slot, (const int32_t *) bt.cell_item);
    hipLaunchKernelGGL(k_slot_min, dim3(nblocks(in.ncells)), dim3(kBlock), 0, st, w.used, w.bnd, in.ncells,
                       w.rep_slot, w.slot_min);
    const uint8_t *qn = tables ? w.qn : nullptr;
    hipLaunchKernelGGL(k_keep_flags, dim3(nbv), dim3(kBlock), 0, st, w.used, w.bnd, in.ncells, w.rep_slot,
                       w.slot_min, qn, w.bsum);
    // tot[2] = kept vertices, tot[3] = quads (tables); tot[0] = quads after the copy
    scan_pairs(w.bsum, (int64_t) nbv, w.tot + 2, w.part, st);
    if (tables) hipLaunchKernelGGL(k_copy_i64, dim3(1), dim3(64), 0, st, w.tot, (const int64_t *) (w.tot + 3));
    if (bt.items)
        hipLaunchKernelGGL(k_item_vbase, dim3(bt.n), dim3(kBlock), 0, st, bt, w.used, w.bnd, w.rep_slot, w.slot_min,
                           qn, w.bsum, in.ncells, (const int64_t *) w.tot);
    hipLaunchKernelGGL(k_keep_emit, dim3(nbv), dim3(kBlock), 0, st, w.dual_v, w.used, w.bnd, in.ncells,
                       w.rep_slot, w.slot_min, w.bsum, w.final_id, w.out_v, w.out_vcell, bt);
    if (tables)
        hipLaunchKernelGGL(k_quad_emit_cells, dim3(nbv), dim3(kBlock), 0, st, in, s, w.qn, w.bsum, w.dual_v,
                           WeldIds{w.bnd, w.rep_slot, w.slot_min, w.final_id, w.out_f}, bt);
    else
        hipLaunchKernelGGL(k_remap_faces, dim3(nblocks(6 * in.npts)), dim3(kBlock), 0, st, w.tri_cells, w.tot,
                           w.bnd, w.rep_slot, w.slot_min, w.final_id, w.out_f);
    return hipGetLastError();
}

// Dense grids without an intersection set (active cells and lookup tables
// only): QEF from the corner values, weld, one scan for vertex and quad
// offsets, quads per cell.
static hipError_t dc_launch_grid(const CellSrc &s, const int64_t *cell_dense, int64_t ncells, const uint64_t *amask,
                                 const int32_t *apre, float level, float reg, float svd_tol, float cond_fp64,
                                 DcScratch &w, const DcBatch &bt, hipStream_t st, bool has_grid = true,
                                 bool has_prog = false) {
    const unsigned nbv = nblocks(ncells);
    hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, st, (int64_t *) w.n_fp64, (int64_t) 1, (int64_t) 0);
    if (has_grid)
        hipLaunchKernelGGL(k_dc_qef_grid<false>, dim3(nbv), dim3(kBlock), 0, st, cell_dense, ncells, s, level, reg,
                           svd_tol, cond_fp64, w.dual_v, w.n_fp64, w.used, w.bnd, w.qn, w.qmask, bt);
    if (has_prog)
        hipLaunchKernelGGL(k_dc_qef_grid<true>, dim3(nbv), dim3(kBlock), 0, st, cell_dense, ncells, s, level, reg,
                           svd_tol, cond_fp64, w.dual_v, w.n_fp64, w.used, w.bnd, w.qn, w.qmask, bt);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(w.hmask + 1)), dim3(kBlock), 0, st, w.hslots, w.hmask + 1, -1);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(w.hmask + 1)), dim3(kBlock), 0, st, w.slot_min, w.hmask + 1,
                       INT32_MAX);
    hipLaunchKernelGGL(k_pos_insert, dim3(nbv), dim3(kBlock), 0, st, w.dual_v, w.used, w.bnd, ncells, w.hslots,
                       w.hmask, w.rep_slot, (const int32_t *) bt.cell_item);
    hipLaunchKernelGGL(k_slot_min, dim3(nbv), dim3(kBlock), 0, st, w.used, w.bnd, ncells, w.rep_slot, w.slot_min);
    hipLaunchKernelGGL(k_keep_flags, dim3(nbv), dim3(kBlock), 0, st, w.used, w.bnd, ncells, w.rep_slot, w.slot_min,
                       (const uint8_t *) w.qn, w.bsum);
    // tot[2] = kept vertices, tot[3] = quads; tot[0] = quads after the copy
    scan_pairs(w.bsum, (int64_t) nbv, w.tot + 2, w.part, st);
    hipLaunchKernelGGL(k_copy_i64, dim3(1), dim3(64), 0, st, w.tot, (const int64_t *) (w.tot + 3));
    if (bt.items)
        hipLaunchKernelGGL(k_item_vbase, dim3(bt.n), dim3(kBlock), 0, st, bt, w.used, w.bnd, w.rep_slot, w.slot_min,
                           (const uint8_t *) w.qn, w.bsum, ncells, (const int64_t *) w.tot);
    hipLaunchKernelGGL(k_keep_emit, dim3(nbv), dim3(kBlock), 0, st, w.dual_v, w.used, w.bnd, ncells, w.rep_slot,
                       w.slot_min, w.bsum, w.final_id, w.out_v, w.out_vcell, bt);
    hipLaunchKernelGGL(k_quad_emit_grid, dim3(nbv), dim3(kBlock), 0, st, cell_dense, ncells, s, amask, apre,
                       (const uint8_t *) w.qn, (const uint8_t *) w.qmask, w.bsum, w.dual_v,
                       WeldIds{w.bnd, w.rep_slot, w.slot_min, w.final_id, w.out_f}, bt);
    return hipGetLastError();
}

hipError_t dc_run(const GridDesc &g, const ItsIn &in, float reg, float svd_tol, float cond_fp64, DcScratch &w,
                  int64_t &nv, int64_t &nf, int64_t &n_fp64, hipStream_t st) {
    CellSrc s = make_src(g);
    nv = nf = n_fp64 = 0;
    if (in.ncells == 0) return hipSuccess;
    ISO_TRY(dc_launch(s, in, reg, svd_tol, cond_fp64, w, DcBatch{}, st));
    ISO_TRY(hipGetLastError());
    // totals and the re-solve counter through a small pinned buffer of the calling thread
    struct PinnedTot {
        int64_t *p = nullptr;
        ~PinnedTot() {
            if (p) (void) hipHostFree(p);
        }
    };
    thread_local PinnedTot pt;
    if (!pt.p) ISO_TRY(hipHostMalloc((void **) &pt.p, 8 * sizeof(int64_t), 0));
    ISO_TRY(hipMemcpyAsync(pt.p, w.tot, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipMemcpyAsync(pt.p + 4, w.n_fp64, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    const int64_t nq = pt.p[0];
    w.nquads = nq;
    nv = pt.p[2];
    nf = 2 * nq;
    n_fp64 = pt.p[4];
    return hipSuccess;
}

size_t dc_batch_item_bytes() { return sizeof(DcItem); }

hipError_t dc_run_batch(int n, const GridDesc *g, const int64_t *cell0, const uint64_t *const *amask,
                        const int32_t *const *apre, const float *levels, const ItsIn &in, float reg, float svd_tol,
                        float cond_fp64, DcScratch &w, const DcBatchDev &d, int64_t *nv, int64_t *nq,
                        int64_t &n_fp64, hipStream_t st) {
    n_fp64 = 0;
    for (int i = 0; i < n; ++i) nv[i] = nq[i] = 0;
    if (n < 1 || n > kMaxDcBatchGrids) return hipErrorInvalidValue;
    if (in.ncells == 0) return hipSuccess;
    // descriptors up and counts back through one pinned buffer of the calling thread
    struct PinnedStage {
        char *p = nullptr;
        size_t bytes = 0;
        ~PinnedStage() {
            if (p) (void) hipHostFree(p);
        }
    };
    thread_local PinnedStage ps;
    const size_t up = sizeof(DcItem) * (size_t) n, down = sizeof(int64_t) * (2 * (size_t) n + 3);
    const size_t want = ((up + 255) & ~size_t(255)) + down;
    if (ps.bytes < want) {
        if (ps.p) ISO_TRY(hipHostFree(ps.p));
        ps.p = nullptr;
        ps.bytes = 0;
        ISO_TRY(hipHostMalloc((void **) &ps.p, want, 0));
        ps.bytes = want;
    }
    DcItem *hi = (DcItem *) ps.p;
    int64_t key_base = 0;
    bool tables = amask && apre, has_grid = false, has_prog = false;
    for (int i = 0; i < n; ++i) {
        hi[i].s = make_src(g[i]);
        hi[i].cell0 = cell0[i];
        hi[i].key_base = key_base;
        key_base += hi[i].s.n;
        hi[i].amask = amask ? amask[i] : nullptr;
        hi[i].apre = apre ? apre[i] : nullptr;
        hi[i].W = (g[i].Z + 63) / 64;
        hi[i].level = levels ? levels[i] : 0.0f;
        if (g[i].sparse_ids) tables = false;
        if (g[i].csg) has_prog = true;
        else has_grid = true;
    }
    if (has_prog && !levels) return hipErrorInvalidValue;   // program fields have no intersection set
    if (levels && !tables) return hipErrorInvalidValue;   // the grid path needs the tables
    int64_t *rb = (int64_t *) (ps.p + ((up + 255) & ~size_t(255)));
    ISO_TRY(hipMemcpyAsync(d.items, hi, up, hipMemcpyHostToDevice, st));
    const DcBatch bt{(const DcItem *) d.items, n, tables, d.cell_item, d.vbase, d.qbase};
    if (levels)
        ISO_TRY(dc_launch_grid(hi[0].s, in.cell_dense, in.ncells, nullptr, nullptr, 0.0f, reg, svd_tol, cond_fp64, w,
                               bt, st, has_grid, has_prog));
    else
        ISO_TRY(dc_launch(hi[0].s, in, reg, svd_tol, cond_fp64, w, bt, st));
    ISO_TRY(hipMemcpyAsync(rb, d.vbase, sizeof(int64_t) * (size_t) (n + 1), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipMemcpyAsync(rb + n + 1, d.qbase, sizeof(int64_t) * (size_t) (n + 1), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipMemcpyAsync(rb + 2 * n + 2, w.n_fp64, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    ISO_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
        nv[i] = rb[i + 1] - rb[i];
        nq[i] = rb[n + 2 + i] - rb[n + 1 + i];
    }
    w.nquads = rb[2 * n + 1];
    n_fp64 = rb[2 * n + 2];
    return hipSuccess;
}

hipError_t dc_backward(const GridDesc &g, const ItsIn &in, float reg, float svd_tol, float level, int64_t nv,
                       const int32_t *vcell, const float *grad_v, float *grad_values, float *grad_normals,
                       hipStream_t st) {
    if (nv == 0) return hipSuccess;
    CellSrc s = make_src(g);
    hipLaunchKernelGGL(k_dc_backward, dim3(nblocks(nv)), dim3(kBlock), 0, st, in, s, reg, svd_tol, level, nv, vcell,
                       grad_v, grad_values, grad_normals);
    return hipGetLastError();
}

}   // namespace iso
