// CPU restatement of the reference iso-surface algorithms -- TEST INFRASTRUCTURE.
//
// This file is the parity oracle and the CPU baseline.  Only tests/, the
// smoke check and bench.py's cpu_baseline leg may load it; the product path
// (isoext_amd) never does.
//
// Every function below restates a reference routine with 64-bit indexing and
// an explicit, uncontracted float operation order (compile with
// -ffp-contract=off).  Reference citations are to GuangyanCai/isoext v0.5.1.
//
// Corner k of a cell is at (k>>2, (k>>1)&1, k&1) relative to the cell's min
// corner (include/shared_luts.cuh:30-32); edge e joins kEdge[e][0] < kEdge[e][1]
// (src/shared_luts.cu:3-4).

#pragma STDC FP_CONTRACT OFF

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <parallel/algorithm>
#include <vector>

#include "../isoext_amd/csrc/mc_luts.h"

namespace {

const int kEdge[12][2] = {{0, 1}, {1, 3}, {2, 3}, {0, 2}, {4, 5}, {5, 7},
                          {6, 7}, {4, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};

struct F3 {
    float x, y, z;
};

inline bool eq3(const F3 &a, const F3 &b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool less3(const F3 &a, const F3 &b) {
    // thrust::tuple lexicographic compare (include/math.cuh:112-118)
    if (a.x < b.x) return true;
    if (b.x < a.x) return false;
    if (a.y < b.y) return true;
    if (b.y < a.y) return false;
    return a.z < b.z;
}

// get_vtx_pos_op (include/utils.cuh:62-80): coordinate of grid index i on an
// axis with res points: aabb_min + (i / (res-1)) * (aabb_max - aabb_min).
inline float grid_coord(int64_t i, int64_t res, float amin, float amax) {
    float size = amax - amin;
    float p = (float) i / (float) (res - 1);
    float prod = p * size;
    return amin + prod;
}

// lerp (include/math.cuh:77-81): (1 - t) * a + t * b, one component.
inline float lerp1(float t, float a, float b) {
    float s = 1.0f - t;
    float u = s * a;
    float v = t * b;
    return u + v;
}

// Edge interpolation parameter (src/mc/nagae.cu:55-57).  A NaN parameter
// (only possible with NaN/inf inputs, which the reference handles
// inconsistently) is mapped to 0.
inline float edge_t(float v0, float v1, float level) {
    float denom = v1 - v0;
    float t = (denom != 0.0f) ? (level - v0) / denom : 0.0f;
    if (t != t) t = 0.0f;
    return t;
}

struct Mesh {
    std::vector<float> v;   // 3 per vertex
    std::vector<int32_t> f; // 3 per face
};

// vertex_welding (src/utils.cu:32-59): sort positions lexicographically,
// unique by equality, face index = lower_bound rank.  With more than one
// thread the sort is libstdc++'s parallel multiway mergesort and the
// lower_bound pass is split over threads; the result is the same.
void weld(const std::vector<F3> &tri_v, Mesh &out, int threads) {
    std::vector<F3> s(tri_v);
    const int nth = std::max(1, threads);
    if (nth > 1 && s.size() > (size_t) 1 << 16)
        __gnu_parallel::sort(s.begin(), s.end(), less3, __gnu_parallel::multiway_mergesort_tag(nth));
    else
        std::sort(s.begin(), s.end(), less3);
    s.erase(std::unique(s.begin(), s.end(), eq3), s.end());
    out.v.resize(s.size() * 3);
    for (size_t i = 0; i < s.size(); ++i) {
        out.v[3 * i + 0] = s[i].x;
        out.v[3 * i + 1] = s[i].y;
        out.v[3 * i + 2] = s[i].z;
    }
    out.f.resize(tri_v.size());
    const int64_t n = (int64_t) tri_v.size();
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int64_t i = 0; i < n; ++i)
        out.f[i] = (int32_t) (std::lower_bound(s.begin(), s.end(), tri_v[i], less3) - s.begin());
}

// process_cube_op (src/mc/nagae.cu:34-80, src/mc/lorensen.cu:34-80): emit the
// non-degenerate triangles of one cell given its corner positions/values.
void emit_cell(const F3 *cp, const float *cv, int case_num, const uint64_t *tris, float level,
               std::vector<F3> &out, uint64_t *n_degenerate) {
    int status = 0;
    for (int e = 0; e < 12; ++e)
        if (((case_num >> kEdge[e][0]) & 1) != ((case_num >> kEdge[e][1]) & 1)) status |= 1 << e;
    F3 ev[12];
    for (int e = 0; e < 12; ++e) {
        if (!(status & (1 << e))) continue;
        int a = kEdge[e][0], b = kEdge[e][1];
        float t = edge_t(cv[a], cv[b], level);
        ev[e].x = lerp1(t, cp[a].x, cp[b].x);
        ev[e].y = lerp1(t, cp[a].y, cp[b].y);
        ev[e].z = lerp1(t, cp[a].z, cp[b].z);
    }
    uint64_t w = tris[case_num];
    int nt = (int) (w & 15);
    for (int k = 0; k < nt; ++k) {
        int e0 = (int) ((w >> (4 + 12 * k)) & 15);
        int e1 = (int) ((w >> (8 + 12 * k)) & 15);
        int e2 = (int) ((w >> (12 + 12 * k)) & 15);
        const F3 &v0 = ev[e0], &v1 = ev[e1], &v2 = ev[e2];
        if (!eq3(v0, v1) && !eq3(v0, v2) && !eq3(v1, v2)) {
            out.push_back(v0);
            out.push_back(v1);
            out.push_back(v2);
        } else if (n_degenerate) {
            ++*n_degenerate;
        }
    }
}

inline int case_of(const float *cv, float level) {
    // get_case_num_op (include/utils.cuh:91-101)
    int c = 0;
    for (int k = 0; k < 8; ++k) c |= ((cv[k] - level) < 0.0f) << k;
    return c;
}

struct DenseGrid {
    int64_t nx, ny, nz;
    float amin[3], amax[3];
    const float *values;

    int64_t pid(int64_t x, int64_t y, int64_t z) const { return (x * ny + y) * nz + z; }
    void corner(int64_t cx, int64_t cy, int64_t cz, int k, F3 &p, float &v) const {
        int64_t x = cx + (k >> 2), y = cy + ((k >> 1) & 1), z = cz + (k & 1);
        p.x = grid_coord(x, nx, amin[0], amax[0]);
        p.y = grid_coord(y, ny, amin[1], amax[1]);
        p.z = grid_coord(z, nz, amin[2], amax[2]);
        v = values[pid(x, y, z)];
    }
};

// Active cells of a dense grid in ascending linear cell order
// (src/mc/mc.cu:26-42), with their cases.  OpenMP over x-slabs keeps order.
void dense_active(const DenseGrid &g, float level, std::vector<int64_t> &cells,
                  std::vector<uint8_t> &cases, int threads) {
    int64_t cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
    cells.clear();
    cases.clear();
    if (cx <= 0 || cy <= 0 || cz <= 0) return;
    std::vector<std::vector<int64_t>> part_c(cx);
    std::vector<std::vector<uint8_t>> part_k(cx);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int64_t x = 0; x < cx; ++x) {
        for (int64_t y = 0; y < cy; ++y) {
            const float *r00 = g.values + g.pid(x, y, 0);
            const float *r01 = g.values + g.pid(x, y + 1, 0);
            const float *r10 = g.values + g.pid(x + 1, y, 0);
            const float *r11 = g.values + g.pid(x + 1, y + 1, 0);
            for (int64_t z = 0; z < cz; ++z) {
                float cv[8] = {r00[z], r00[z + 1], r01[z], r01[z + 1],
                               r10[z], r10[z + 1], r11[z], r11[z + 1]};
                int c = case_of(cv, level);
                if (c != 0 && c != 255) {
                    part_c[x].push_back((x * cy + y) * cz + z);
                    part_k[x].push_back((uint8_t) c);
                }
            }
        }
    }
    for (int64_t x = 0; x < cx; ++x) {
        cells.insert(cells.end(), part_c[x].begin(), part_c[x].end());
        cases.insert(cases.end(), part_k[x].begin(), part_k[x].end());
    }
}

template <class T> T *dup(const std::vector<T> &v) {
    T *p = (T *) std::malloc(std::max<size_t>(1, v.size() * sizeof(T)));
    if (!v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

// ---- symmetric 3x3 eigen-solver in FP64 (cyclic Jacobi) ---------------------
void jacobi3(double a[3][3], double w[3], double v[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if (off <= 1e-36 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p) {
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < 3; ++i) w[i] = a[i][i];
}

// Truncated pseudo-inverse solve (src/batched_la.cu:151-178): singular values
// of a symmetric matrix are |eigenvalues|; keep s_j > tol * s_max.
void pinv_solve(const double A[9], const double b[3], double tol, double x[3], double *cond) {
    double a[3][3], w[3], v[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) a[i][j] = A[3 * i + j];
    jacobi3(a, w, v);
    double smax = std::max(std::fabs(w[0]), std::max(std::fabs(w[1]), std::fabs(w[2])));
    double smin = std::min(std::fabs(w[0]), std::min(std::fabs(w[1]), std::fabs(w[2])));
    if (cond) *cond = smin > 0 ? smax / smin : INFINITY;
    double thr = tol * smax;
    x[0] = x[1] = x[2] = 0.0;
    for (int j = 0; j < 3; ++j) {
        if (!(std::fabs(w[j]) > thr)) continue;
        double proj = v[0][j] * b[0] + v[1][j] * b[1] + v[2][j] * b[2];
        double c = proj / w[j];
        for (int i = 0; i < 3; ++i) x[i] += c * v[i][j];
    }
}

}   // namespace

extern "C" {

// ---- results ----------------------------------------------------------------
struct OracleMesh {
    int64_t nv, nf;
    float *v;
    int32_t *f;
    uint64_t n_degenerate;
};

void oracle_free(void *p) { std::free(p); }

void oracle_free_mesh(OracleMesh *m) {
    if (!m) return;
    std::free(m->v);
    std::free(m->f);
    std::free(m);
}

static OracleMesh *to_result(const std::vector<F3> &tri_v, uint64_t ndeg, int threads = 1) {
    Mesh m;
    weld(tri_v, m, threads);
    OracleMesh *r = (OracleMesh *) std::calloc(1, sizeof(OracleMesh));
    r->nv = (int64_t) m.v.size() / 3;
    r->nf = (int64_t) m.f.size() / 3;
    r->v = dup(m.v);
    r->f = dup(m.f);
    r->n_degenerate = ndeg;
    return r;
}

// mc::marching_cubes on a UniformGrid (src/mc/mc.cu:17-68).  method 0=nagae,
// 1=lorensen.  Values are (nx, ny, nz) C-order, z fastest.
OracleMesh *oracle_mc_dense(const float *values, int64_t nx, int64_t ny, int64_t nz,
                            const float *amin, const float *amax, float level, int method,
                            int threads) {
    DenseGrid g{nx, ny, nz, {amin[0], amin[1], amin[2]}, {amax[0], amax[1], amax[2]}, values};
    std::vector<int64_t> cells;
    std::vector<uint8_t> cases;
    dense_active(g, level, cells, cases, threads);
    const uint64_t *tris = method == 1 ? kLorensenTris : kNagaeTris;
    int64_t cy = ny - 1, cz = nz - 1;
    int nth = std::max(1, threads);
    std::vector<std::vector<F3>> part(nth);
    std::vector<uint64_t> pdeg(nth, 0);
    int64_t n = (int64_t) cells.size();
    // Ordered emission: split active cells into contiguous chunks.
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int t = 0; t < nth; ++t) {
        int64_t lo = n * t / nth, hi = n * (t + 1) / nth;
        for (int64_t i = lo; i < hi; ++i) {
            int64_t c = cells[i];
            int64_t x = c / (cy * cz), y = (c / cz) % cy, z = c % cz;
            F3 cp[8];
            float cv[8];
            for (int k = 0; k < 8; ++k) g.corner(x, y, z, k, cp[k], cv[k]);
            emit_cell(cp, cv, cases[i], tris, level, part[t], &pdeg[t]);
        }
    }
    std::vector<F3> tri_v;
    uint64_t ndeg = 0;
    for (int t = 0; t < nth; ++t) {
        tri_v.insert(tri_v.end(), part[t].begin(), part[t].end());
        ndeg += pdeg[t];
    }
    return to_result(tri_v, ndeg, nth);
}

// Cases and active cells only (classification stage, for count checks and
// timing the reference's per-cell classification).
int64_t oracle_count_active(const float *values, int64_t nx, int64_t ny, int64_t nz,
                            float level, int threads) {
    float z3[3] = {0, 0, 0}, o3[3] = {1, 1, 1};
    DenseGrid g{nx, ny, nz, {z3[0], z3[1], z3[2]}, {o3[0], o3[1], o3[2]}, values};
    std::vector<int64_t> cells;
    std::vector<uint8_t> cases;
    dense_active(g, level, cells, cases, threads);
    return (int64_t) cells.size();
}

// marching_cubes on a SparseGrid (src/mc/mc.cu:17-68 with
// SparseGrid::get_points/get_cells, src/grid/sparse.cu:21-69): each listed
// cell carries its own 8 corner values; corner positions come from the cell's
// linear index (x, y, z decomposed with (nx-1, ny-1, nz-1), utils.cuh:44-47).
OracleMesh *oracle_mc_sparse(const int64_t *cell_ids, const float *values, int64_t n,
                             int64_t nx, int64_t ny, int64_t nz, const float *amin,
                             const float *amax, float level, int method) {
    const uint64_t *tris = method == 1 ? kLorensenTris : kNagaeTris;
    std::vector<F3> tri_v;
    uint64_t ndeg = 0;
    for (int64_t i = 0; i < n; ++i) {
        const float *cv = values + 8 * i;
        int c = case_of(cv, level);
        if (c == 0 || c == 255) continue;
        int64_t id = cell_ids[i];
        int64_t z = id % (nz - 1);
        int64_t r = id / (nz - 1);
        int64_t y = r % (ny - 1);
        int64_t x = r / (ny - 1);
        F3 cp[8];
        for (int k = 0; k < 8; ++k) {
            cp[k].x = grid_coord(x + (k >> 2), nx, amin[0], amax[0]);
            cp[k].y = grid_coord(y + ((k >> 1) & 1), ny, amin[1], amax[1]);
            cp[k].z = grid_coord(z + (k & 1), nz, amin[2], amax[2]);
        }
        emit_cell(cp, cv, c, tris, level, tri_v, &ndeg);
    }
    return to_result(tri_v, ndeg);
}

// ---- intersections (src/its.cu) ---------------------------------------------
struct OracleIts {
    int64_t ncells, npts;
    int64_t *cell_indices;   // active cells: dense linear cell id, or sparse slot
    int64_t *cell_offsets;   // ncells + 1
    float *points;           // npts * 3
    float *normals;          // npts * 3 (valid if computed)
    int64_t *edges;          // npts * 2 point ids (dense ids, or sparse 8*slot+k)
    uint8_t *is_out;         // npts
};

void oracle_free_its(OracleIts *r) {
    if (!r) return;
    std::free(r->cell_indices);
    std::free(r->cell_offsets);
    std::free(r->points);
    std::free(r->normals);
    std::free(r->edges);
    std::free(r->is_out);
    std::free(r);
}

// trilinear_interp (src/its.cu:170-184)
static float trilinear(float tx, float ty, float tz, const float *c) {
    float a = 1.0f - tz;
    float c00 = c[0] * a + c[1] * tz;
    float c01 = c[2] * a + c[3] * tz;
    float c10 = c[4] * a + c[5] * tz;
    float c11 = c[6] * a + c[7] * tz;
    float b = 1.0f - ty;
    float c0 = c00 * b + c01 * ty;
    float c1 = c10 * b + c11 * ty;
    float d = 1.0f - tx;
    return c0 * d + c1 * tx;
}

// compute_normals_op (src/its.cu:186-268) for one intersection point.
static void cell_normal(const F3 *cp, const float *cv, F3 p, F3 &n) {
    F3 mn = cp[0];
    F3 sz{cp[7].x - cp[0].x, cp[7].y - cp[0].y, cp[7].z - cp[0].z};
    float tx = (p.x - mn.x) / sz.x, ty = (p.y - mn.y) / sz.y, tz = (p.z - mn.z) / sz.z;
    tx = std::fmax(0.01f, std::fmin(0.99f, tx));
    ty = std::fmax(0.01f, std::fmin(0.99f, ty));
    tz = std::fmax(0.01f, std::fmin(0.99f, tz));
    const float eps = 0.02f;
    float pxx = std::fmin(tx + eps, 0.99f), mxx = std::fmax(tx - eps, 0.01f);
    float pyy = std::fmin(ty + eps, 0.99f), myy = std::fmax(ty - eps, 0.01f);
    float pzz = std::fmin(tz + eps, 0.99f), mzz = std::fmax(tz - eps, 0.01f);
    float dx = (trilinear(pxx, ty, tz, cv) - trilinear(mxx, ty, tz, cv)) / ((pxx - mxx) * sz.x);
    float dy = (trilinear(tx, pyy, tz, cv) - trilinear(tx, myy, tz, cv)) / ((pyy - myy) * sz.y);
    float dz = (trilinear(tx, ty, pzz, cv) - trilinear(tx, ty, mzz, cv)) / ((pzz - mzz) * sz.z);
    float d2 = dx * dx + dy * dy;
    d2 = d2 + dz * dz;
    float len = std::sqrt(d2);
    if (len > 1e-8f) {
        n.x = dx / len;
        n.y = dy / len;
        n.z = dz / len;
    } else {
        n.x = 0.0f;
        n.y = 0.0f;
        n.z = 1.0f;
    }
}

// get_intersection (src/its.cu:93-159) for a UniformGrid (sparse == 0) or a
// SparseGrid (sparse == 1: values are (n, 8), cell_ids lists the n cells).
OracleIts *oracle_intersection(const float *values, const int64_t *cell_ids, int64_t n_sparse,
                               int sparse, int64_t nx, int64_t ny, int64_t nz,
                               const float *amin, const float *amax, float level,
                               int compute_normals) {
    std::vector<int64_t> ci, off;
    std::vector<float> pts, nrm;
    std::vector<int64_t> edg;
    std::vector<uint8_t> out;
    int64_t cy = ny - 1, cz = nz - 1;
    int64_t ncell = sparse ? n_sparse : (nx - 1) * cy * cz;
    if (nx < 2 || ny < 2 || nz < 2) ncell = sparse ? n_sparse : 0;
    DenseGrid g{nx, ny, nz, {amin[0], amin[1], amin[2]}, {amax[0], amax[1], amax[2]}, values};
    for (int64_t i = 0; i < ncell; ++i) {
        int64_t id = sparse ? cell_ids[i] : i;
        int64_t x = id / (cy * cz), y = (id / cz) % cy, z = id % cz;
        F3 cp[8];
        float cv[8];
        int64_t pidk[8];
        for (int k = 0; k < 8; ++k) {
            int64_t px = x + (k >> 2), py = y + ((k >> 1) & 1), pz = z + (k & 1);
            cp[k].x = grid_coord(px, nx, amin[0], amax[0]);
            cp[k].y = grid_coord(py, ny, amin[1], amax[1]);
            cp[k].z = grid_coord(pz, nz, amin[2], amax[2]);
            if (sparse) {
                cv[k] = values[8 * i + k];
                pidk[k] = 8 * i + k;
            } else {
                pidk[k] = g.pid(px, py, pz);
                cv[k] = values[pidk[k]];
            }
        }
        int c = case_of(cv, level);
        if (c == 0 || c == 255) continue;
        ci.push_back(sparse ? i : id);
        off.push_back((int64_t) out.size());
        for (int e = 0; e < 12; ++e) {
            int a = kEdge[e][0], b = kEdge[e][1];
            if (((c >> a) & 1) == ((c >> b) & 1)) continue;
            edg.push_back(pidk[a]);
            edg.push_back(pidk[b]);
            out.push_back(cv[a] <= cv[b]);
            float t = edge_t(cv[a], cv[b], level);
            F3 p{lerp1(t, cp[a].x, cp[b].x), lerp1(t, cp[a].y, cp[b].y), lerp1(t, cp[a].z, cp[b].z)};
            pts.push_back(p.x);
            pts.push_back(p.y);
            pts.push_back(p.z);
            if (compute_normals) {
                F3 nn;
                cell_normal(cp, cv, p, nn);
                nrm.push_back(nn.x);
                nrm.push_back(nn.y);
                nrm.push_back(nn.z);
            }
        }
    }
    off.push_back((int64_t) out.size());
    OracleIts *r = (OracleIts *) std::calloc(1, sizeof(OracleIts));
    r->ncells = (int64_t) ci.size();
    r->npts = (int64_t) out.size();
    r->cell_indices = dup(ci);
    r->cell_offsets = dup(off);
    r->points = dup(pts);
    if (!compute_normals) nrm.assign(pts.size(), 0.0f);
    r->normals = dup(nrm);
    r->edges = dup(edg);
    r->is_out = dup(out);
    return r;
}

// Single-point entry to compute_normals_op (src/its.cu:186-268) for the
// closed-form checks of the normal: cp = 8 corner positions (x, y, z each),
// cv = 8 corner values, p = the point; writes the unit normal to n.
void oracle_cell_normal(const float *cp, const float *cv, const float *p, float *n) {
    F3 c[8];
    for (int k = 0; k < 8; ++k) c[k] = F3{cp[3 * k], cp[3 * k + 1], cp[3 * k + 2]};
    F3 r;
    cell_normal(c, cv, F3{p[0], p[1], p[2]}, r);
    n[0] = r.x;
    n[1] = r.y;
    n[2] = r.z;
}

// Single-system entry to the truncated pseudo-inverse solve
// (src/batched_la.cu:151-178) used by oracle_dual_contouring: A is a
// symmetric 3x3 (row-major), singular values s_j <= tol * s_max are dropped.
// Returns the condition number s_max / s_min.
double oracle_pinv_solve(const double *A, const double *b, double tol, double *x) {
    double cond = 0.0;
    pinv_solve(A, b, tol, x, &cond);
    return cond;
}

// ---- dual contouring (src/dc.cu:161-218) ------------------------------------
struct OracleDC {
    OracleMesh *mesh;       // welded mesh built from the FP32-rounded dual vertices
    int64_t ncells;
    double *dual_v64;       // ncells * 3: FP64 QEF solution, clipped
    double *cond;           // ncells: condition number of the regularised QEF
    int64_t nquads;
    int64_t *quads;         // nquads * 4 active-cell indices, orientation applied
    uint8_t *split02;       // nquads: 1 when the quad was split along v0-v2
};

void oracle_free_dc(OracleDC *r) {
    if (!r) return;
    oracle_free_mesh(r->mesh);
    std::free(r->dual_v64);
    std::free(r->cond);
    std::free(r->quads);
    std::free(r->split02);
    std::free(r);
}

// edge_neighbors_table (src/shared_luts.cu:78-82): offsets subtracted from an
// edge's lower point to get its 4 cells, per axis (x, y, z).
static const int kEdgeNb[3][4][3] = {{{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}},
                                     {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}},
                                     {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}};

// Dual contouring from an intersection set.  Inputs mirror Intersection:
// per-entry points/normals/edge point ids (dense point ids, or for sparse
// grids 8*slot+k), per-entry is_out, per active cell its cell id (dense) or
// sparse slot, and offsets.  For sparse grids, slot_cell maps a slot to its
// dense cell id (SparseGrid::get_dual_quads, src/grid/sparse.cu:181-246).
// QEF accumulation is FP64 from the FP32 inputs; the solve is the FP64
// truncated pseudo-inverse; quads whose 4 cells do not all exist are skipped
// (the reference aliases past the upper grid boundary, utils.cuh:130-134).
OracleDC *oracle_dual_contouring(const float *points, const float *normals, const int64_t *edges,
                                 const uint8_t *is_out, int64_t npts,
                                 const int64_t *cell_indices, const int64_t *cell_offsets,
                                 int64_t ncells, int sparse, const int64_t *slot_cell,
                                 int64_t nslots, int64_t nx, int64_t ny, int64_t nz,
                                 const float *amin, const float *amax, float reg, float svd_tol) {
    int64_t cy = ny - 1, cz = nz - 1;
    OracleDC *r = (OracleDC *) std::calloc(1, sizeof(OracleDC));
    r->ncells = ncells;
    r->dual_v64 = (double *) std::malloc(sizeof(double) * 3 * std::max<int64_t>(1, ncells));
    r->cond = (double *) std::malloc(sizeof(double) * std::max<int64_t>(1, ncells));
    std::vector<F3> dv(ncells);
    // dense cell id of each active cell
    std::vector<int64_t> dense_id(ncells);
    for (int64_t i = 0; i < ncells; ++i)
        dense_id[i] = sparse ? slot_cell[cell_indices[i]] : cell_indices[i];
    for (int64_t i = 0; i < ncells; ++i) {
        double A[9] = {0}, b[3] = {0}, pa[3] = {0};
        int64_t lo = cell_offsets[i], hi = cell_offsets[i + 1];
        for (int64_t j = lo; j < hi; ++j) {
            double n[3] = {normals[3 * j], normals[3 * j + 1], normals[3 * j + 2]};
            double p[3] = {points[3 * j], points[3 * j + 1], points[3 * j + 2]};
            double d = n[0] * p[0] + n[1] * p[1] + n[2] * p[2];
            for (int a = 0; a < 3; ++a) {
                for (int c = 0; c < 3; ++c) A[3 * a + c] += n[a] * n[c];
                b[a] += n[a] * d;
                pa[a] += p[a];
            }
        }
        double cnt = (double) (hi - lo);
        for (int a = 0; a < 3; ++a) {
            pa[a] /= cnt;
            A[4 * a] += (double) reg;
            b[a] += (double) reg * pa[a];
        }
        double x[3];
        pinv_solve(A, b, (double) svd_tol, x, &r->cond[i]);
        int64_t id = dense_id[i];
        int64_t cxi = id / (cy * cz), cyi = (id / cz) % cy, czi = id % cz;
        float lo3[3] = {grid_coord(cxi, nx, amin[0], amax[0]), grid_coord(cyi, ny, amin[1], amax[1]),
                        grid_coord(czi, nz, amin[2], amax[2])};
        float hi3[3] = {grid_coord(cxi + 1, nx, amin[0], amax[0]),
                        grid_coord(cyi + 1, ny, amin[1], amax[1]),
                        grid_coord(czi + 1, nz, amin[2], amax[2])};
        for (int a = 0; a < 3; ++a) {
            double v = std::fmin(std::fmax(x[a], (double) lo3[a]), (double) hi3[a]);
            r->dual_v64[3 * i + a] = v;
        }
        dv[i] = F3{(float) r->dual_v64[3 * i], (float) r->dual_v64[3 * i + 1],
                   (float) r->dual_v64[3 * i + 2]};
    }
    // unique crossing edges in (p0, p1) order, with is_out (uniform.cu:60-83)
    struct E {
        int64_t p0, p1;
        uint8_t out;
    };
    std::vector<E> es;
    es.reserve(npts);
    for (int64_t j = 0; j < npts; ++j) {
        int64_t a = edges[2 * j], b = edges[2 * j + 1];
        if (sparse) {
            // sparse point id 8*slot+k -> dense point id
            int64_t sa = a / 8, ka = a % 8, sb = b / 8, kb = b % 8;
            int64_t ida = slot_cell[sa], idb = slot_cell[sb];
            int64_t xa = ida / (cy * cz), ya = (ida / cz) % cy, za = ida % cz;
            int64_t xb = idb / (cy * cz), yb = (idb / cz) % cy, zb = idb % cz;
            a = ((xa + (ka >> 2)) * ny + ya + ((ka >> 1) & 1)) * nz + za + (ka & 1);
            b = ((xb + (kb >> 2)) * ny + yb + ((kb >> 1) & 1)) * nz + zb + (kb & 1);
        }
        es.push_back(E{a, b, is_out[j]});
    }
    std::stable_sort(es.begin(), es.end(), [](const E &u, const E &w) {
        return u.p0 < w.p0 || (u.p0 == w.p0 && u.p1 < w.p1);
    });
    std::vector<E> ue;
    for (auto &e : es)
        if (ue.empty() || ue.back().p0 != e.p0 || ue.back().p1 != e.p1) ue.push_back(e);
    // cell id -> active index
    std::vector<std::pair<int64_t, int64_t>> amap(ncells);
    for (int64_t i = 0; i < ncells; ++i) amap[i] = {dense_id[i], i};
    std::sort(amap.begin(), amap.end());
    auto find_active = [&](int64_t id) -> int64_t {
        auto it = std::lower_bound(amap.begin(), amap.end(), std::make_pair(id, (int64_t) -1));
        if (it == amap.end() || it->first != id) return -1;
        return it->second;
    };
    std::vector<int64_t> qv;
    std::vector<uint8_t> split;
    std::vector<F3> tri_v;
    for (auto &e : ue) {
        int64_t px = e.p0 / (ny * nz), py = (e.p0 / nz) % ny, pz = e.p0 % nz;
        int64_t d = e.p1 - e.p0;
        int axis = d == ny * nz ? 0 : (d == nz ? 1 : 2);
        int64_t q[4];
        bool ok = true;
        for (int k = 0; k < 4 && ok; ++k) {
            int64_t x = px - kEdgeNb[axis][k][0], y = py - kEdgeNb[axis][k][1],
                    z = pz - kEdgeNb[axis][k][2];
            if (x < 0 || y < 0 || z < 0 || x >= nx - 1 || y >= cy || z >= cz) {
                ok = false;
                break;
            }
            q[k] = find_active((x * cy + y) * cz + z);
            if (q[k] < 0) ok = false;
        }
        if (!ok) continue;
        if (!e.out) {
            std::swap(q[0], q[3]);
            std::swap(q[1], q[2]);
        }
        // get_triangles_op (src/dc.cu:101-157) on FP32-rounded vertices
        const F3 &v0 = dv[q[0]], &v1 = dv[q[1]], &v2 = dv[q[2]], &v3 = dv[q[3]];
        auto nrm = [](F3 a, F3 b) {
            float x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
            float s = x * x + y * y;
            s = s + z * z;
            return std::sqrt(s);
        };
        bool s13 = nrm(v0, v2) > nrm(v1, v3);
        for (int k = 0; k < 4; ++k) qv.push_back(q[k]);
        split.push_back(s13 ? 0 : 1);
        if (s13) {
            F3 t[6] = {v1, v3, v0, v3, v1, v2};
            tri_v.insert(tri_v.end(), t, t + 6);
        } else {
            F3 t[6] = {v2, v0, v1, v0, v2, v3};
            tri_v.insert(tri_v.end(), t, t + 6);
        }
    }
    r->nquads = (int64_t) split.size();
    r->quads = dup(qv);
    r->split02 = dup(split);
    r->mesh = to_result(tri_v, 0);
    return r;
}

}   // extern "C"
