/*
 * hdbscan_rocm -- C ABI of the MI355X (gfx950) batched HDBSCAN hot path.
 *
 * Every pointer is a device pointer owned by the caller; the library
 * allocates no memory and enqueues work only on the given HIP stream
 * (hipStream_t passed as void*; NULL = the null stream).  Calls on different
 * streams may overlap; calls that share a workspace must not.
 *
 * Precision codes: HDB_F32 (float) and HDB_F64 (double).  x is a dense
 * row-major [batch][n][d] array of that type.
 *
 * Status: *status (device int32) is OR-accumulated; the caller zeroes it
 * before the call and reads it after synchronising.  HDB_STATUS_NONFINITE
 * outranks HDB_STATUS_OVERFLOW when reporting.  Labels of a call whose
 * status is non-zero are undefined.
 *
 * Return value: 0 on success, otherwise a negative code; hdb_last_error()
 * then describes the failure (invalid argument or HIP launch error).
 *
 * The functions replace the reference's CUDA backend entry
 *   cluster(x, min_cluster_size, min_samples, epsilon, precision)
 *   (reference hdbscan_gpu/_cuda.py:91-128, called from
 *    hdbscan_gpu/_api.py:92-102)
 * and its kernel launches K1/K2 and the library k-selection
 *   (reference hdbscan_gpu/_cuda.py:110-127).
 */
#ifndef HDBSCAN_ROCM_H
#define HDBSCAN_ROCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HDB_F32 0
#define HDB_F64 1

#define HDB_MAX_SAMPLES 4096

#define HDB_STATUS_OVERFLOW 1  /* a pairwise distance is not finite */
#define HDB_STATUS_NONFINITE 2 /* an input coordinate is not finite */

/* Flags for hdb_cluster / hdb_workspace_bytes (tuning and test knobs). */
#define HDB_FLAG_GLOBAL_ARENA 1 /* tree arena in device memory, not LDS */
#define HDB_FLAG_MATRIX 2       /* materialise distances, Prim reads rows */
#define HDB_FLAG_NO_POINT_LDS 4 /* current vertex read from device memory */
#define HDB_FLAG_MATRIX_INPUT 8 /* sizing for hdb_cluster_matrix (distances given) */
#define HDB_FLAG_SERIAL_SORT 16 /* one-lane introsort (the parallel one is default) */
#define HDB_FLAG_PREFILTER 32   /* Prim: certified FMA Gram bound skips exact distances */
#define HDB_FLAG_PLAIN 64       /* no MFMA Gram filter: every core distance exact */
#define HDB_FLAG_FILTER 128     /* MFMA Gram filter also on the matrix path */
#define HDB_FLAG_ONE_STEP 65536 /* matrix-path Prim: one vertex per step, no row speculation */
#define HDB_FLAG_GRAM_CHUNKED 131072 /* core candidates by the chunked Gram kernel also for windows of <= 1024 points, <= 16 dims */
#define HDB_FLAG_GRAM_GRAPH 262144 /* float32 matrix path: MFMA Gram matrix + interval-key Prim (no exact matrix) */
#define HDB_FLAG_NO_GRAPH 524288    /* ragged calls: launch directly instead of capturing / replaying a HIP graph */
#define HDB_FLAG_SPLIT_TREE 1048576 /* tree stage as a sort kernel + a narrow hierarchy kernel (sorted-edge transport) */
#define HDB_FLAG_FUSED_TREE 2097152 /* one tree kernel per window even when windows outnumber the CUs */
#define HDB_FLAG_SERIAL_LINKAGE 4194304 /* linkage by one wave in sort order (the parallel run form is default) */
#define HDB_FLAG_NO_GRAM_EXACT 8388608 /* float32 matrix path: evaluate every pair exactly (no MFMA-bounded near/far matrix) */
/* Bits 8..11: vertices per Prim thread (1, 2 or 4; 8 on the speculative
 * matrix path only); 0 = automatic.
 * Bits 12..13: 16-row sets per wave of the Gram core kernel (1 -> 1, 2 -> 2,
 * 3 -> 4; 0 = 4 for windows of 1025..2048 points, else 2; clamped to what
 * the dimension allows).
 * Bits 14..15: Prim repacking interval of the points path (0 -> 128 steps,
 * 1 -> never, 2 -> 32, 3 -> 64). */

/* Library version (major*10000 + minor*100 + patch). */
int hdb_version(void);

/* Text of the last error of the calling thread ("" if none). */
const char *hdb_last_error(void);

/* Stage timers of the calling thread (no counterpart in the reference; the
 * reference times stages from Python, scripts/profile_cuda.py:17-146).
 * With hdb_profile(1), every hdb_cluster / hdb_cluster_matrix call records a
 * HIP event after each kernel stage on its stream.  After synchronising,
 * hdb_stage_ms(i) is the duration of stage i of the last call on this thread
 * and hdb_stage_name(i) its kernel name. */
int hdb_profile(int enable);
int hdb_stage_count(void);
const char *hdb_stage_name(int i);
float hdb_stage_ms(int i);

/* Measured VALU issue rate of the exact distance's instruction mix (sub,
 * mul, add; no FMA, no packed fp32), the attainable ceiling of the
 * exact-distance kernels.  No counterpart in the reference (a measurement
 * diagnostic, like the stage timers).  Runs `blocks` workgroups of 256
 * threads, each thread 24 * iters operations, on `stream` (one untimed
 * warm-up launch); out: a device buffer of blocks * 256 floats; *ms: the
 * timed launch's duration.  Synchronises the stream. */
int hdb_probe_valu(void *stream, int blocks, int iters, float *out, float *ms);

/* Bytes of device workspace hdb_cluster needs for these arguments.
 * Replaces the reference's global-workspace sizing
 * (reference hdbscan_gpu/_cuda.py:62-64, 119-121). */
size_t hdb_workspace_bytes(int64_t batch, int64_t n, int64_t d, int dtype,
                           int32_t min_samples, int flags);

/* Labels-only Euclidean EOM HDBSCAN of `batch` independent windows.
 * labels: [batch][n] int32, -1 = noise.  min_samples is self-inclusive;
 * epsilon is converted to the working precision.
 * Replaces reference hdbscan_gpu/_cuda.py:91-128 (cluster). */
int hdb_cluster(void *stream, const void *x, int64_t batch, int64_t n, int64_t d,
                int dtype, int64_t min_cluster_size, int32_t min_samples,
                double epsilon, int32_t *labels, void *workspace,
                size_t workspace_bytes, int32_t *status, int flags);

/* hdb_cluster without a workspace argument: it allocates nothing.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its workspace is a fixed array of
 * HDB_NOALLOC_BYTES in the library's device memory, one per GPU, held until
 * the process ends.  A call whose hdb_workspace_bytes(batch, n, d, dtype,
 * min_samples, flags) exceeds HDB_NOALLOC_BYTES fails with an error (B=256
 * windows of 1024 points in 16 float32 dimensions fit).  Every call on a GPU
 * shares that GPU's array: calls enqueued on one stream run one after the
 * other, but calls that overlap on one GPU (different streams, or threads
 * without ordering) race and give undefined labels.  Arguments, status and
 * results as hdb_cluster. */
#define HDB_NOALLOC_BYTES (64u * 1024u * 1024u)
int hdb_cluster_noalloc(void *stream, const void *x, int64_t batch, int64_t n,
                        int64_t d, int dtype, int64_t min_cluster_size,
                        int32_t min_samples, double epsilon, int32_t *labels,
                        int32_t *status, int flags);

/* hdb_cluster with every intermediate exported (layout of hdb_cluster_matrix
 * below) and an optional forced introsort depth: the white-box view of the
 * production path.  Replaces the reference's diagnostic kernels that export
 * the cluster kernel's workspace (reference tests/test_kernels.py:197-270). */
int hdb_cluster_export(void *stream, const void *x, int64_t batch, int64_t n,
                       int64_t d, int dtype, int64_t min_cluster_size,
                       int32_t min_samples, double epsilon, int sort_depth,
                       int32_t *labels, void *workspace, size_t workspace_bytes,
                       int32_t *status, int flags, int32_t *export_ints,
                       void *export_reals);

/* Dense pairwise distances out[batch][n][n] (diagonal +0), bitwise equal to
 * the reference's triangle-and-mirror kernel.
 * Replaces reference hdbscan_gpu/_kernels.py:45-88 (distances_kernel). */
int hdb_distances(void *stream, const void *x, int64_t batch, int64_t n,
                  int64_t d, int dtype, void *out, int32_t *status);

/* The MFMA-bounded matrix of the float32 matrix path, out[batch][n][n]:
 * for each pair either its exact distance (bitwise hdb_distances' value,
 * sign bit clear) or minus a certified lower bound of it (sign bit set),
 * as hdb_cluster's float32 matrix path stores it for min_samples <= 16
 * (which pairs are exact depends on min_samples).  Workspace:
 * hdb_workspace_bytes(batch, n, d, HDB_F32, min_samples, HDB_FLAG_MATRIX).
 * A diagnostic with no counterpart in the reference (whose matrix is always
 * exact, hdbscan_gpu/_kernels.py:45-88); it exposes the bounds for tests and
 * for measuring their error against a float64 evaluation. */
int hdb_bounded_distances(void *stream, const void *x, int64_t batch, int64_t n,
                          int64_t d, int dtype, int32_t min_samples, void *out,
                          void *workspace, size_t workspace_bytes,
                          int32_t *status);

/* Core distances core[batch][n]: the (min_samples-1)-th smallest distance of
 * each row, self included.  Workspace: hdb_workspace_bytes(batch, n, d,
 * dtype, min_samples, flags | HDB_FLAG_MATRIX | HDB_FLAG_FILTER |
 * HDB_FLAG_NO_GRAM_EXACT) bytes always suffice (none is used when
 * min_samples <= 64 and the MFMA filter does not apply).  The filter is used
 * unless flags has HDB_FLAG_PLAIN: its per-window form for windows of up to
 * 1024 points and 16 dimensions (unless HDB_FLAG_GRAM_CHUNKED), the chunked
 * form otherwise.
 * Replaces reference hdbscan_gpu/_cuda.py:114-118 (cp.partition). */
int hdb_core_distances(void *stream, const void *x, int64_t batch, int64_t n,
                       int64_t d, int dtype, int32_t min_samples, void *core,
                       void *workspace, size_t workspace_bytes, int32_t *status,
                       int flags);

/* Ragged batch: `windows` windows of uneven size packed point after point in
 * x[total][d] (row-major), window w holding points offsets[w] ..
 * offsets[w+1]-1 (offsets[0] = 0, each size 2..4096, min_samples <= size).
 * `offsets` is the host copy (planning), `offsets_dev` the same values in
 * device memory (read by the kernels).  Labels are written per point into
 * labels[total], numbered within each window as hdb_cluster numbers them.
 * Windows are grouped by size into buckets (up to 64, 128, ..., 4096 points);
 * the buckets run concurrently on internal streams forked from and joined
 * into `stream` (one set of internal streams per caller stream, so ragged
 * calls on different streams may overlap).  Each bucket's Prim and tree
 * launches are persistent: their workgroups claim the bucket's windows from
 * a device counter, largest first.  The first call of a signature (buffers,
 * sizes, parameters, offsets) is captured into a HIP graph that later calls
 * replay (HDB_FLAG_NO_GRAPH: always launch directly).  Windows whose
 * coordinates do not fit the points path's registers (float32 above 16
 * dimensions over 1024 points) take a materialised matrix per window.
 * The reference API is dense only (hdbscan_gpu/_api.py:71-78: one [B, N, D]
 * array); this entry serves uneven windows in one call. */
size_t hdb_ragged_workspace_bytes(const int64_t *offsets, int64_t windows,
                                  int64_t d, int dtype, int32_t min_samples,
                                  int flags);
int hdb_cluster_ragged(void *stream, const void *x, const int64_t *offsets,
                       const int64_t *offsets_dev, int64_t windows, int64_t d,
                       int dtype, int64_t min_cluster_size,
                       int32_t min_samples, double epsilon, int32_t *labels,
                       void *workspace, size_t workspace_bytes,
                       int32_t *status, int flags);

/* Tuning table of ragged calls (no counterpart in the reference; written by
 * the autotuner, hdbscan_amd/autotune.py): for size bucket k (windows of up
 * to 64, 128, 256, 512, 1024, 2048, 4096 points), bucket_flags[k] replaces
 * the call's tuning bits (vertices per Prim thread, Gram row sets, repacking
 * interval, HDB_FLAG_GLOBAL_ARENA, HDB_FLAG_NO_POINT_LDS) for that bucket.
 * count = 0 clears the table.  Process-wide; results do not depend on it. */
int hdb_ragged_set_tuning(const int32_t *bucket_flags, int count);

/* Graph and tree stages from a given distance matrix [batch][n][n] and core
 * distances [batch][n] (core may be NULL when unit_core != 0, meaning every
 * mutual-reachability weight equals the distance).  Distances and core
 * distances must be +0 or positive (weights are compared as IEEE bit
 * patterns, whose order is the values' order only there), as the
 * reference's distance kernel produces them.  When export_ints /
 * export_reals are non-NULL every intermediate is written per window:
 *   export_ints  [batch][10][n] int32: edge_u, edge_v, order, left, right,
 *                sizes, parent, selected, point_parent, {next_label, ...}
 *   export_reals [batch][4][n] Real:  edge_w, distance, birth, stability
 * (cluster-indexed arrays hold cluster id - n).  sort_depth < -1000 keeps
 * NumPy's depth; other values force the introsort depth.
 * Replaces reference hdbscan_gpu/_cuda.py:122-127 (clusters_kernel launch)
 * and the diagnostic workspace export of reference tests/test_kernels.py:197-270. */
int hdb_cluster_matrix(void *stream, const void *distances, const void *core,
                       int64_t batch, int64_t n, int dtype, int64_t min_cluster_size,
                       int unit_core, double epsilon, int sort_depth,
                       int32_t *labels, void *workspace, size_t workspace_bytes,
                       int32_t *status, int flags, int32_t *export_ints,
                       void *export_reals);

#ifdef __cplusplus
}
#endif

#endif /* HDBSCAN_ROCM_H */
