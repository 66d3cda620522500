/*
 * libisoext_amd -- C ABI of the MI355X iso-surface engine.
 *
 * Every pointer named values/out/ids/... is a device pointer (HBM) unless
 * documented otherwise; sizes are int64 element counts; `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  Functions return
 * ISOEXT_OK or an ISOEXT_E_* code and never throw; isoext_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * The reference (GuangyanCai/isoext v0.5.1) exposes this path through the
 * nanobind module isoext_ext (src/isoext_ext.cu:93-384); each declaration
 * below names the reference interface it replaces.  Grids are dense (X, Y, Z)
 * float32 arrays in C order (z fastest), `shape` counts POINTS per axis
 * (src/grid/uniform.cu:11-12), positions follow get_vtx_pos_op
 * (include/utils.cuh:62-80).
 */
#ifndef ISOEXT_AMD_H
#define ISOEXT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ISOEXT_ABI_VERSION 2

enum {
    ISOEXT_OK = 0,
    ISOEXT_E_ARG = -1,        /* invalid argument (maps to ValueError/TypeError) */
    ISOEXT_E_DEVICE = -2,     /* HIP runtime or kernel failure (RuntimeError) */
    ISOEXT_E_CAPACITY = -3,   /* an output buffer is too small; counts hold the need */
    ISOEXT_E_WORKSPACE = -4,  /* workspace too small; counts[2] holds the need */
    ISOEXT_E_TOO_LARGE = -5,  /* more than 2^31-1 vertices or faces */
};

/* Message of the last failure on this thread (empty if none). */
const char *isoext_last_error(void);
int isoext_version(void);

/* Instrumentation (no reference counterpart): when enabled, dense marching
 * cubes brackets its kernels with HIP events on the caller's stream;
 * isoext_last_kernel_ms() returns the extraction's device time (first to
 * last kernel, weld fix-up excluded) and isoext_mc_last_stage_ms() the
 * stages: classify, count, scan, emit, total for the four-stage pipeline;
 * reset, stream, 0, faces, total for the read-once pipeline (tile
 * configuration 3).  Per calling thread. */
void isoext_set_timing(int on);
float isoext_last_kernel_ms(void);
void isoext_mc_last_stage_ms(float out[5]);

/* ---- grids ---------------------------------------------------------------- */

/* UniformGrid.get_points (src/grid/uniform.cu:22-30, bound at
 * src/isoext_ext.cu:147-152): writes X*Y*Z*3 floats. */
int isoext_uniform_points(const int64_t shape[3], const float aabb_min[3],
                          const float aabb_max[3], float *out, void *stream);

/* UniformGrid default fill / SparseGrid value reset
 * (src/grid/uniform.cu:18-19, src/grid/sparse.cu:94-95). */
int isoext_fill_f32(float *ptr, int64_t n, float value, void *stream);

/* ---- marching cubes on a UniformGrid --------------------------------------- */

/* Number of tile configurations compiled into the library (autotuner range). */
int isoext_mc_config_count(void);

/* Default tile configuration for a grid shape: the largest tiles that still
 * give every CU several workgroups (no reference counterpart). */
int isoext_mc_auto_config(const int64_t shape[3]);

/* Workspace bytes for isoext_mc_dense with tile config `cfg` and room for
 * `special_cap` weld-candidate vertices.  The workspace must be zero-filled
 * before its first use and may be reused by later calls on the same stream
 * (calls that overlap on one workspace race). */
int64_t isoext_mc_dense_workspace_bytes(const int64_t shape[3], int cfg, int64_t special_cap);

/* marching_cubes(grid: UniformGrid, level, method) -- replaces
 * mc::marching_cubes (include/mc/mc.cuh:14-15, src/mc/mc.cu:17-68) bound at
 * src/isoext_ext.cu:95-109.  method: 0 = "nagae", 1 = "lorensen"
 * (src/mc/base.cu:19-27).  Writes up to cap_v vertices (3 floats each) and
 * cap_f faces (3 int32 each).  counts[0..1] = vertices, faces of the final
 * mesh; counts[2] = special-list capacity needed; counts[3] = device error
 * bits; counts[4..5] = provisional vertices, faces (before the weld fix-up
 * removes duplicates): the capacities a call on the same input needs.  On
 * ISOEXT_E_CAPACITY counts[0..1] also hold the provisional sizes to allocate
 * before retrying.  Synchronises `stream`. */
int isoext_mc_dense(const float *values, const int64_t shape[3], const float aabb_min[3],
                    const float aabb_max[3], float level, int method, int cfg,
                    void *workspace, int64_t workspace_bytes, int64_t special_cap,
                    float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f,
                    int64_t counts[6], void *stream);

/* Marching cubes of a field that is never stored (no reference counterpart;
 * the step before the path: the reference's users build the grid with
 * UniformGrid.set_values(sdf(points))).  The field is the union of nsph
 * spheres (1..256), v(p) = min_k sqrt((dx*dx + dy*dy) + dz*dz) - r_k in fp32
 * at the grid points of (shape, aabb); `spheres` is a device array of nsph
 * float4 (cx, cy, cz, r), 16-byte aligned.  The classify pass evaluates only
 * the spheres that can reach below `level` in each block of points and
 * writes inside bits; the emit pass evaluates the corners of active cells
 * over the spheres that can be the minimum in its tile.  The result is
 * identical to isoext_mc_dense on the grid such a generator writes
 * (isoext_sdf_spheres), without writing or reading that grid.  Same
 * workspace (isoext_mc_dense_workspace_bytes), outputs and counts as
 * isoext_mc_dense.  Synchronises `stream`. */
int isoext_mc_spheres(const float *spheres, int nsph, const int64_t shape[3], const float aabb_min[3],
                      const float aabb_max[3], float level, int method, int cfg, void *workspace,
                      int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v, int32_t *out_f,
                      int64_t cap_f, int64_t counts[6], void *stream);

/* Allocation-free marching cubes: no reference counterpart.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same result, outputs and
 * counts as isoext_mc_dense with the default tile configuration
 * (isoext_mc_auto_config), but every intermediate -- inside masks, counts,
 * special lists, the weld fix-up's tables -- lives in one fixed device array
 * per GPU (96 MiB, held until the process ends); no hipMalloc, hipFree or
 * other allocation happens in the call.  Grids whose workspace exceeds
 * isoext_mc_dense_noalloc_limit() bytes (a cube of about 620^3 points) return
 * ISOEXT_E_TOO_LARGE; a special-list need (counts[2]: weld candidates, spread
 * over 256 segments) above 32,768 returns ISOEXT_E_WORKSPACE.
 * The caller supplies the outputs (on ISOEXT_E_CAPACITY counts[0..1] hold the
 * sizes to allocate).  Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams or threads) race.  Synchronises
 * `stream`. */
int isoext_mc_dense_noalloc(const float *values, const int64_t shape[3], const float aabb_min[3],
                            const float aabb_max[3], float level, int method, float *out_v, int64_t cap_v,
                            int32_t *out_f, int64_t cap_f, int64_t counts[6], void *stream);

/* Workspace bytes the allocation-free entry point can give one grid. */
int64_t isoext_mc_dense_noalloc_limit(void);

/* ---- grid descriptor used by intersections, dual contouring, sparse MC ---- */

/* A UniformGrid (sparse_cell_ids == NULL: values is (X, Y, Z)) or a
 * SparseGrid (values is (n_sparse, 8) per-cell corner values and
 * sparse_cell_ids the n_sparse sorted, unique linear cell ids over
 * (X-1, Y-1, Z-1); src/grid/sparse.cu:21-69). */
typedef struct isoext_grid {
    const float *values;
    const int64_t *sparse_cell_ids;
    int64_t n_sparse;
    int64_t shape[3];
    float aabb_min[3];
    float aabb_max[3];
} isoext_grid;

/* An intersection set in the reference's entry order (src/its.cu:93-159,
 * include/its.cuh:6-26): active cells ascending; per cell its crossing edges
 * in edge-index order.  cell_indices is the linear cell id (UniformGrid) or
 * the sparse slot (SparseGrid); cell_dense is always the linear cell id. */
typedef struct isoext_its {
    int64_t ncells, npts;
    const int64_t *cell_indices;
    const int64_t *cell_dense;
    const int64_t *cell_offsets; /* ncells + 1 */
    const float *points;         /* npts x 3 */
    const float *normals;        /* npts x 3 */
    const uint8_t *edge_local;   /* edge index 0..11 inside its cell */
    const uint8_t *is_out;       /* value at the lower corner <= value at the upper */
    const int32_t *entry_cell;   /* active-cell index of each entry */
    /* Optional (UniformGrid sets written by isoext_its_emit, else NULL):
     * per (cell row x * (Y-1) + y, 64-cell word z / 64) the word's active-cell
     * mask and the active index of its first active cell.  isoext_dc then
     * finds a quad's neighbour cells from them instead of a hash. */
    const uint64_t *amask_tab;
    const int32_t *apre_tab;
} isoext_its;

/* Scratch bytes shared by isoext_its_count / isoext_its_emit for `g`. */
int64_t isoext_its_scratch_bytes(const isoext_grid *g);

/* get_intersection pass 1 (src/its.cu:105-135): counts[0] = active cells,
 * counts[1] = intersection points.  Synchronises `stream`. */
int isoext_its_count(const isoext_grid *g, float level, void *scratch, int64_t scratch_bytes,
                     int64_t counts[2], void *stream);

/* Entries of the lookup tables of a UniformGrid intersection set:
 * (X-1) * (Y-1) * ceil(Z / 64). */
int64_t isoext_its_table_entries(const int64_t shape[3]);

/* get_intersection pass 2 (get_its_op src/its.cu:37-90, plus
 * compute_normals_op src/its.cu:186-268 when normals != NULL), using the
 * scratch of the preceding isoext_its_count on the same grid and level.
 * amask_tab / apre_tab (UniformGrid only, isoext_its_table_entries each, or
 * NULL) receive the set's lookup tables (see isoext_its). */
int isoext_its_emit(const isoext_grid *g, float level, const void *scratch, int64_t ncells,
                    int64_t npts, int64_t *cell_indices, int64_t *cell_dense, int64_t *cell_offsets,
                    float *points, float *normals, uint8_t *edge_local, uint8_t *is_out,
                    int32_t *entry_cell, uint64_t *amask_tab, int32_t *apre_tab, void *stream);

/* compute_intersection_normals (src/its.cu:271-284) for an existing set. */
int isoext_its_normals(const isoext_grid *g, const isoext_its *its, float *normals, void *stream);

/* Workspace bytes for isoext_dc on an intersection set of this size. */
int64_t isoext_dc_workspace_bytes(int64_t ncells, int64_t npts);

/* dual_contouring(grid, level, intersection, reg, svd_tol) -- replaces
 * ::dual_contouring (include/dc.cuh:9-11, src/dc.cu:161-218) with the
 * BatchedLASolver (src/batched_la.cu:151-178) and get_dual_quads
 * (src/grid/uniform.cu:60-83, src/grid/sparse.cu:181-246), bound at
 * src/isoext_ext.cu:345-378.  Needs cap_v >= ncells and cap_f >= 2*npts.
 * counts = {vertices, faces, cells re-solved in FP64, quads}.  Synchronises. */
int isoext_dc(const isoext_grid *g, const isoext_its *its, float reg, float svd_tol, void *workspace,
              int64_t workspace_bytes, float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f,
              float *dual_v, int32_t *out_vcell, int64_t counts[4], void *stream);

/* Backward of dual contouring (no reference counterpart): accumulates
 * dL/dvalues += dL/dv * dv/dvalues through the regularised QEF solve of each
 * output vertex's cell and the edge interpolation of its intersection
 * points (out_vcell of isoext_dc maps vertices to cells; clipped components
 * pass no gradient); writes dL/dnormals per intersection when grad_normals
 * is not NULL.  Normals are treated as inputs. */
int isoext_dc_backward(const isoext_grid *g, const isoext_its *its, float reg, float svd_tol, float level,
                       int64_t nv, const int32_t *vcell, const float *grad_v, float *grad_values,
                       float *grad_normals, void *stream);

/* Same as isoext_mc_dense, also writing each vertex's grid edge to out_key
 * (owner point linear id * 3 + axis, 0 = x) for the backward pass.  No
 * reference counterpart: the reference has no autograd. */
int isoext_mc_dense_keyed(const float *values, const int64_t shape[3], const float aabb_min[3],
                          const float aabb_max[3], float level, int method, int cfg,
                          void *workspace, int64_t workspace_bytes, int64_t special_cap,
                          float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f,
                          int64_t *out_key, int64_t counts[6], void *stream);

/* Backward of marching cubes (no reference counterpart): accumulates
 * dL/dvalues += dL/dv * dv/dvalues through each vertex's edge interpolation.
 * Dense: keys from isoext_mc_dense_keyed, grad_values is (X, Y, Z).
 * Sparse: vsrc from isoext_mc_sparse, grad_values is (n_sparse, 8). */
int isoext_mc_backward(const isoext_grid *g, float level, const float *grad_v, const int64_t *vertex_src,
                       int64_t nv, float *grad_values, void *stream);

/* Benchmark input generation (no reference counterpart; the reference's SDFs
 * are torch code, src/isoext/sdf.py): out[x,y,z] = min_k |p - c_k| - r_k over
 * k <= 256 spheres given as device float4 (cx, cy, cz, r). */
int isoext_sdf_spheres(const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                       const float *spheres, int k, float *out, void *stream);

/* ---- SparseGrid ------------------------------------------------------------ */

/* SparseGrid.get_points / get_points_by_cell_indices
 * (src/grid/sparse.cu:21-36,144-148): (n, 8, 3) corner positions of cells
 * given by linear cell id over (X-1, Y-1, Z-1). */
int isoext_sparse_corner_points(const int64_t *cell_ids, int64_t n, const int64_t shape[3],
                                const float aabb_min[3], const float aabb_max[3], float *out,
                                void *stream);

/* SparseGrid.filter_cell_indices (src/grid/sparse.cu:150-179): flags[i] = 1
 * when cell i's 8 values (values is (n, 8)) straddle `level`. */
int isoext_sparse_crossing_flags(const float *values, int64_t n, float level, uint8_t *flags,
                                 void *stream);

/* ---- SparseGrid population (src/grid/sparse.cu:71-121) ----------------------- */

/* Scratch bytes of isoext_sparse_union / isoext_sparse_difference. */
int64_t isoext_sparse_setop_scratch_bytes(int64_t na, int64_t nb);

/* SparseGrid::add_cells (src/grid/sparse.cu:71-96): sorted union of the
 * grid's sorted unique ids a[na] with b_sorted[nb] (sorted, duplicates
 * allowed).  out (capacity na + nb) receives the result, slot[k] the index in
 * a of out[k] or -1 for a new id (slot may be NULL), *n_out (device memory)
 * the result length.  Stream-ordered, no host synchronisation. */
int isoext_sparse_union(const int64_t *a, int64_t na, const int64_t *b_sorted, int64_t nb, void *scratch,
                        int64_t scratch_bytes, int64_t *out, int64_t *slot, int64_t *n_out, void *stream);

/* SparseGrid::remove_cells (src/grid/sparse.cu:98-121): a minus b_sorted,
 * same conventions (out capacity na). */
int isoext_sparse_difference(const int64_t *a, int64_t na, const int64_t *b_sorted, int64_t nb, void *scratch,
                             int64_t scratch_bytes, int64_t *out, int64_t *slot, int64_t *n_out, void *stream);

/* Values after a set operation: row k of new_values (n x 8) = row slot[k] of
 * old_values, or default_value where slot[k] < 0.  The reference resets
 * every value instead (src/grid/sparse.cu:94-95,119-120); this keeps the
 * values of surviving cells when the caller asks for it. */
int isoext_sparse_gather_values(const float *old_values, const int64_t *slot, int64_t n, float default_value,
                                float *new_values, void *stream);

/* Workspace bytes for isoext_mc_sparse on n cells with `ninst` crossing-edge
 * instances (from isoext_mc_sparse_count). */
int64_t isoext_mc_sparse_workspace_bytes(int64_t n, int64_t ninst);

/* Pass 1 of marching_cubes on a SparseGrid: number of crossing-edge
 * instances.  scratch needs 8 * (n / 256 + 16) bytes.  Synchronises. */
int isoext_mc_sparse_count(const isoext_grid *g, float level, int64_t *scratch, int64_t *ninst,
                           void *stream);

/* marching_cubes(grid: SparseGrid, level, method) -- replaces
 * mc::marching_cubes (src/mc/mc.cu:17-68) over SparseGrid::get_cells
 * (src/grid/sparse.cu:61-69), bound at src/isoext_ext.cu:95-109.  Faces in
 * cell then LUT-slot order; vertices welded by exact position.  vsrc
 * (optional) receives slot * 12 + edge of each vertex.  counts = {vertices,
 * faces}; ISOEXT_E_CAPACITY when cap_v/cap_f are too small (counts hold the
 * need).  Synchronises. */
int isoext_mc_sparse(const isoext_grid *g, float level, int method, int64_t ninst, void *workspace,
                     int64_t workspace_bytes, float *out_v, int64_t cap_v, int32_t *out_f,
                     int64_t cap_f, int64_t *vsrc, int64_t counts[2], void *stream);

/* ---- captured plans (no reference counterpart: repeated extractions of one
 * grid buffer, e.g. an optimisation loop that updates the values in place) ----
 * isoext_mc_plan_create records isoext_mc_dense's kernel launches for these
 * exact arguments (values pointer, shape, level, method, tile configuration
 * other than the read-once one, workspace, output buffers and capacities)
 * into a HIP graph; isoext_mc_plan_run replays it on `stream`, reads the
 * counts back and runs the weld fix-up when needed, with isoext_mc_dense's
 * results and error codes (ISOEXT_E_CAPACITY / _WORKSPACE: create a plan
 * with larger buffers).  The buffers must stay allocated while the plan
 * lives; a plan's runs must not overlap. */
typedef struct isoext_mc_plan isoext_mc_plan;
int isoext_mc_plan_create(const float *values, const int64_t shape[3], const float aabb_min[3],
                          const float aabb_max[3], float level, int method, int cfg, void *workspace,
                          int64_t workspace_bytes, int64_t special_cap, float *out_v, int64_t cap_v, int32_t *out_f,
                          int64_t cap_f, isoext_mc_plan **plan);
int isoext_mc_plan_run(isoext_mc_plan *plan, int64_t counts[6], void *stream);
void isoext_mc_plan_destroy(isoext_mc_plan *plan);

/* ---- ragged batches of dense marching cubes (no reference counterpart:
 * the reference extracts one grid per call, src/mc/mc.cu:17-68) ---------- */

/* One grid of a batch.  out_v / cap_v / out_f / cap_f are only read by
 * isoext_mc_batch_emit. */
typedef struct isoext_mc_item {
    const float *values;      /* (X, Y, Z) float32, C order */
    int64_t shape[3];
    float aabb_min[3], aabb_max[3];
    float level;
    int method;               /* 0 = "nagae", 1 = "lorensen" */
    float *out_v;
    int64_t cap_v;
    int32_t *out_f;
    int64_t cap_f;
    const void *sdf;          /* optional: SDF program (isoext_sdf_upload) giving the field; values unused */
    int sdf_normals;          /* dual contouring of an sdf item: 1 analytic program normals, 0 trilinear */
} isoext_mc_item;

/* Workspace for a batch (zero-filled before first use); special_cap is the
 * weld-candidate capacity per item.  cfg >= 0: every item uses that emit tile
 * configuration; cfg < 0: each item its own (isoext_mc_auto_config of its
 * shape).  Items with an SDF program always use configuration 5 (8 x 8 x 1
 * cells).  The same cfg must be passed to both phases. */
int64_t isoext_mc_batch_workspace_bytes(int n, const isoext_mc_item *items, int cfg, int64_t special_cap);

/* Phase 1: classify, count and scan every item (one launch per stage for the
 * whole batch).  counts[2i], counts[2i+1] = provisional vertices, faces of
 * item i: the capacities its outputs need.  Synchronises `stream`. */
int isoext_mc_batch_count(int n, const isoext_mc_item *items, int cfg, void *workspace, int64_t workspace_bytes,
                          int64_t special_cap, int64_t *counts, void *stream);

/* Phase 2 (same items and workspace, outputs set): emit every item, then the
 * weld fix-up of the items that need it.  counts[2i], counts[2i+1] = final
 * vertices, faces; -1, -1 when item i overflowed its special capacity (run it
 * alone with isoext_mc_dense).  Results equal isoext_mc_dense's per item.
 * Synchronises `stream`. */
int isoext_mc_batch_emit(int n, const isoext_mc_item *items, int cfg, void *workspace, int64_t workspace_bytes,
                         int64_t special_cap, int64_t *counts, void *stream);

/* ---- batched dual contouring of dense grids (no counterpart: the reference
 * runs get_intersection + dual_contouring once per grid, src/its.cu:93-159,
 * src/dc.cu:161-218) ------------------------------------------------------------
 * Per item: get_intersection(grid, level, compute_normals=True) followed by
 * dual_contouring(grid, level, its) with the given reg / svd_tol.  Items use
 * values, shape, aabb, level (method must be 0 or 1 and is not used; out_* are
 * not read).  At most 256 items.  Results equal isoext_its_count/emit +
 * isoext_dc per item. */

/* Workspace (zero-filled before first use) shared by both phases. */
int64_t isoext_dc_batch_workspace_bytes(int n, const isoext_mc_item *items);

/* Phase 1: counts[2i] = active cells, counts[2i + 1] = intersection points of
 * item i.  Synchronises `stream`. */
int isoext_dc_batch_count(int n, const isoext_mc_item *items, void *workspace, int64_t workspace_bytes,
                          int64_t *counts, void *stream);

/* Scratch for phase 2 given phase 1's counts. */
int64_t isoext_dc_batch_scratch_bytes(int n, const isoext_mc_item *items, const int64_t *counts);

/* Phase 2 (same workspace as phase 1, untouched in between): intersections,
 * QEF solve, quads and the vertex weld of every grid, one launch per stage.
 * out_v (cap_v >= total active cells rows) and out_f (cap_f >= 2 x total
 * intersections rows) receive the meshes back to back: out_counts[4i..4i+3] =
 * vertices, faces, first vertex row, first face row of item i (face corners
 * are item-local vertex ids); out_counts[4n] = FP64 re-solves.  Synchronises
 * `stream`. */
int isoext_dc_batch_run(int n, const isoext_mc_item *items, const int64_t *counts, float reg, float svd_tol,
                        void *workspace, int64_t workspace_bytes, void *scratch, int64_t scratch_bytes, float *out_v,
                        int64_t cap_v, int32_t *out_f, int64_t cap_f, int64_t *out_counts, void *stream);

/* ---- SDF programs (the step before the path: UniformGrid.set_values(sdf(p)),
 * src/grid/uniform.cu:37-40, with the SDF built from the primitives and CSG
 * operators of src/isoext/sdf.py:70-221) ------------------------------------------
 * A program is n_ops ops of two int32 each -- a code word
 * op | dst << 8 | src << 12 | acc << 16 and a parameter offset -- over 8 value
 * slots and 4 point slots (point slot 0 = the evaluation point, result =
 * value slot 0), with n_params float parameters:
 *   1 SPHERE r | 2 SPHERE_AT cx cy cz r | 3 TORUS R r | 4 CUBOID hx hy hz |
 *   5 CYLINDER_Z r   (primitives: value slot dst from point slot src,
 *                     acc 0 store, 1 min, 2 max into the slot)
 *   6 VALUE (v[src], acc as above) | 7 SMOOTH_MIN k (v[dst] with v[src]) |
 *   8 NEG (v[dst]) | 9 TRANSLATE ox oy oz | 10 ROTATE R00..R22 (point slot
 *   dst = point slot src - o, or @ R).
 * Formulas keep the op order of the reference's torch code.  At most 256 ops
 * and 1024 parameters. */

/* Device bytes of an uploaded program. */
int64_t isoext_sdf_program_bytes(void);

/* Validate a program (ops, params in HOST memory) and write its device form
 * to `program` (isoext_sdf_program_bytes() bytes of device memory, 16-byte
 * aligned).  Synchronises `stream`. */
int isoext_sdf_upload(const int32_t *ops, int32_t n_ops, const float *params, int32_t n_params, void *program,
                      void *stream);

/* The program's values at the grid points of (shape, aabb): the stored grid
 * of UniformGrid.set_values(sdf(grid.get_points())). */
int isoext_sdf_grid(const void *program, const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                    float *out, void *stream);

/* Values (optional) and analytic unit normals (optional: get_sdf_normal,
 * src/isoext/sdf.py:39-50) at n points (n x 3 floats). */
int isoext_sdf_points(const void *program, const float *points, int64_t n, float *values, float *normals,
                      void *stream);

/* Marching cubes of a program's field without storing the grid (same
 * arguments and results as isoext_mc_dense on the grid isoext_sdf_grid
 * writes; identical meshes).  No reference counterpart: the reference always
 * stores the grid first (src/grid/uniform.cu:37-40, src/mc/mc.cu:17-68). */
int isoext_mc_sdf(const void *program, const int64_t shape[3], const float aabb_min[3], const float aabb_max[3],
                  float level, int method, int cfg, void *workspace, int64_t workspace_bytes, int64_t special_cap,
                  float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f, int64_t counts[6], void *stream);

/* ---- OBJ output ------------------------------------------------------------- */

/* write_obj (src/isoext/utils.py:42-63): writes "v x y z" per vertex and
 * "f a b c" per face (indices + 1) to `path`, byte-identical to the
 * reference's text: each coordinate is printed as Python's repr of the float
 * value (shortest round-trip decimal; fixed notation for decimal exponents in
 * (-4, 16], else d.ddde+XX).  v: nv x 3 doubles (float32 vertices widened
 * exactly), f: nf x 3 zero-based indices, both in host memory.  An empty mesh
 * (nv or nf 0) truncates the file and writes nothing.  Lines are formatted on
 * `threads` host threads (<= 0: 8). */
int isoext_write_obj(const char *path, const double *v, int64_t nv, const int64_t *f, int64_t nf, int threads);

/* Python's repr of x into out (NUL-terminated); returns its length, or minus
 * the length when cap is too small.  The formatter isoext_write_obj uses. */
int64_t isoext_format_float(double x, char *out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* ISOEXT_AMD_H */
