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

#define ISOEXT_ABI_VERSION 1

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
 * bits.  On ISOEXT_E_CAPACITY counts[0..1] hold the provisional sizes to
 * allocate before retrying.  Synchronises `stream`. */
int isoext_mc_dense(const float *values, const int64_t shape[3], const float aabb_min[3],
                    const float aabb_max[3], float level, int method, int cfg,
                    void *workspace, int64_t workspace_bytes, int64_t special_cap,
                    float *out_v, int64_t cap_v, int32_t *out_f, int64_t cap_f,
                    int64_t counts[4], void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ISOEXT_AMD_H */
