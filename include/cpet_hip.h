/*
 * libcpet_hip.so — C-ABI of the MI355X (gfx950) implementation of PyCPET's
 * two data-parallel hot paths:
 *   1. the direct Coulomb field / electrostatic potential over box grids
 *      (options-file methods "volume", "volume_ESP"), and
 *   2. electric-field streamline propagation with distance / curvature
 *      features and a 2-D histogram ("topo", "topo_GPU").
 *
 * Every entry point returns an int status: 0 on success, a negative
 * CPET_ERR_* code otherwise; cpet_last_error() then describes the failure
 * (thread-local).  Nothing aborts the process.
 *
 * Part 1 keeps the exact names, argument order and argument meaning of the
 * reference's native module CPET/utils/math_module.c, so the reference's
 * ctypes binding (CPET/utils/c_ops.py: CDLL + ndpointer argtypes, restype
 * None) binds this library unchanged.  Those entry points take host arrays.
 *
 * Part 2 is the batched, device-resident API: a context (device + stream +
 * scratch) and whole-grid / whole-streamline-set calls whose array arguments
 * are host pointers, or device pointers when CPET_F_DEVICE is set (for
 * example torch.Tensor.data_ptr() of tensors on that device).
 */
#ifndef CPET_HIP_H
#define CPET_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define CPET_OK 0
#define CPET_ERR_ARG (-1)    /* invalid argument (sizes, NULLs, non-finite) */
#define CPET_ERR_HIP (-2)    /* a HIP runtime call failed */
#define CPET_ERR_NODEV (-3)  /* no usable GPU */
#define CPET_ERR_NOMEM (-4)  /* device allocation failed */
#define CPET_ERR_BUDGET (-5) /* request exceeds the admission budget */

/* ----------------------------------------------------------------- flags */
#define CPET_F_DEVICE 0x1u   /* array arguments are device pointers */
#define CPET_F_OUT_F64 0x2u  /* field/phi outputs are double, not float */
#define CPET_F_NO_SYNC 0x4u  /* device-pointer calls: do not wait for the stream */
#define CPET_F_DIPOLE_PHYSICAL 0x10u /* cpet_topo_dipole: physical 3(mu.R)R - mu r^2 term */
#define CPET_F_TOPO_RK4 0x20u  /* cpet_topo: classical RK4 steps on E/|E| for the walk and the
                                  curvature steps; no counterpart in the reference, which
                                  integrates with Euler (math_module.c:489-503) */
#define CPET_F_COUNTS 0x40u  /* cpet_cluster_histograms: the raw bin counts (as doubles)
                                instead of count / total */
/* precision of the field kernels, bits 8..10 */
#define CPET_F_PREC_AUTO (0u << 8)  /* E: FP32 tiled, points with strong
                                       cancellation recomputed as MIXED;
                                       phi: MIXED (default) */
#define CPET_F_PREC_NAIVE (1u << 8) /* one FP32 running sum */
#define CPET_F_PREC_KAHAN (2u << 8) /* Kahan FP32 within a 64-charge tile, FP64 across */
#define CPET_F_PREC_F64 (3u << 8)   /* FP64 throughout */
#define CPET_F_PREC_TILED (4u << 8) /* FP32 per 64-charge tile, FP64 across */
#define CPET_F_PREC_MIXED (5u << 8) /* FP64 displacement and r^2, FP32 rsqrt seed +
                                       one FP64 Newton step, FP64 sums */
#define CPET_F_PREC_MASK (7u << 8)

/* ============ Part 1: reference-compatible entry points (host arrays) ==== */

/* E[m] = k * sum_j Q[j] (x_0[m] - x[j]) * max(|x_0[m]-x[j]|^2, 1e-6)^(-3/2)
 * for m < total_points.  x_0 (total_points,3), x (n_charges,3), Q
 * (n_charges), E (total_points,3), all float32 C-order.
 * Replaces compute_looped_field, CPET/utils/math_module.c:405-451
 * (binding CPET/utils/c_ops.py:151-159, wrapper :250-263). */
int compute_looped_field(int total_points, int n_charges, const float *x_0,
                         const float *x, const float *Q, float *E);

/* ESP[0] += k * sum_j Q[j] / |x_init - x[j]| (no softening; r = 0 gives inf).
 * Replaces calc_esp_base, CPET/utils/math_module.c:453-486
 * (binding CPET/utils/c_ops.py:131-138, wrapper :326-341). */
int calc_esp_base(float *ESP, const float *x_init, int n_charges,
                  const float *x, const float *Q);

/* E = k * sum_j Q[j] (x_init - x[j]) / |x_init - x[j]|^3 (no softening).
 * Replaces calc_field, CPET/utils/math_module.c:255-293
 * (binding CPET/utils/c_ops.py:113-120, wrapper :360-375). */
int calc_field(float *E, const float *x_init, int n_charges, const float *x,
               const float *Q);

/* As calc_field but accumulates into E (E += field), as the reference does.
 * Replaces calc_field_base, CPET/utils/math_module.c:296-333
 * (binding CPET/utils/c_ops.py:122-129, wrapper :343-358). */
int calc_field_base(float *E, const float *x_init, int n_charges,
                    const float *x, const float *Q);

/* One streamline from x_0: up to n_iter Euler steps x += h E/|E|, stopping
 * after the first point strictly outside +-dimensions; ret = [distance,
 * mean curvature] (float32).  Replaces thread_operation,
 * CPET/utils/math_module.c:523-591 (binding CPET/utils/c_ops.py:85-95,
 * wrapper :281-302). */
int thread_operation(int n_charges, int n_iter, float step_size,
                     const float *x_0, const float *dimensions,
                     const float *x, const float *Q, float *ret);

/* Dipole streamline from x_0 (the dipole variant of thread_operation,
 * marked "in development" in the reference): the field of n_dipoles point
 * dipoles x (n,3) with moments mu (n,3) as calc_field_base_dipole computes
 * it, E = k sum r^-5 ((3 mu_x R_x + mu_y R_y + mu_z R_z) R - mu r^2).
 * Replaces thread_operation_dipole, CPET/utils/math_module.c:593-659
 * (binding CPET/utils/c_ops.py:97-110, wrapper :304-325). */
int thread_operation_dipole(int n_dipoles, int n_iter, float step_size, const float *x_0,
                            const float *dimensions, const float *x, const float *mu,
                            float *ret);

/* E[m] = k * sum_j Q[j] (x_0[m] - x[j]) / |x_0[m] - x[j]|^3 (no softening)
 * for m < total_points; batch_size (> 0) is the reference's chunk length
 * and does not change the result.  Runs as one launch of the unsoftened
 * points kernel.  Replaces compute_batched_field,
 * CPET/utils/math_module.c:374-403 (binding CPET/utils/c_ops.py:140-149,
 * wrapper :265-279). */
int compute_batched_field(int total_points, int batch_size, int n_charges, const float *x_0,
                          const float *x, const float *Q, float *E);

/* The remaining helpers the reference binding sets argtypes on
 * (CPET/utils/c_ops.py:26-77).  They are off the hot path and run on the
 * host, sequentially, in the element order and arithmetic types of the
 * reference loops; no GPU is needed. */

/* ret[r] = sum_{j = indptr[r]}^{indptr[r+1]-1} A[j] * B[indA[j]] over the
 * indptrlen - 1 rows of a CSR matrix (FP64).  Replaces sparse_dot,
 * CPET/utils/math_module.c:15-46 (binding CPET/utils/c_ops.py:26-36). */
int sparse_dot(double *ret, const int *indptr, int indptrlen, const int *indA, int lenindA,
               const double *A, int lenA, const double *B, int size_array);

/* ret[i] = sum_j A[i*cols + j] * B[j] (FP64).  Replaces dot,
 * CPET/utils/math_module.c:49-68 (binding CPET/utils/c_ops.py:44-50). */
int dot(double *ret, const double *A, const double *B, int rows, int cols);

/* ret = A + B (float32).  Replaces vecaddn, CPET/utils/math_module.c:124-132
 * (binding CPET/utils/c_ops.py:37-42). */
int vecaddn(float *ret, const float *A, const float *B, int lenA);

/* result[k] = sum_i (k * Q[i] * r_mag[k][i]) R[k][i][:] with r_mag
 * (batch,rows), Q (rows), R (batch,rows,3), result (batch,3), float32.
 * Replaces einsum_operation_batch, CPET/utils/math_module.c:135-172
 * (binding CPET/utils/c_ops.py:66-74). */
int einsum_operation_batch(int batch, int rows, const float *r_mag, const float *Q,
                           const float *R, float *result);

/* einsum_operation_batch for one point: result (3).  Replaces
 * einsum_operation, CPET/utils/math_module.c:175-209 (binding
 * CPET/utils/c_ops.py:76-83). */
int einsum_operation(int rows, const float *r_mag, const float *Q, const float *R,
                     float *result);

/* ret[k][i] = sum_j A[k][i][j], A (batch,rows,cols), float32.  Replaces
 * einsum_ij_i_batch, CPET/utils/math_module.c:212-232 (binding
 * CPET/utils/c_ops.py:58-64). */
int einsum_ij_i_batch(int batch, int rows, int cols, const float *A, float *ret);

/* ret[i] = sum_j A[i][j], A (rows,cols), float32.  Replaces einsum_ij_i,
 * CPET/utils/math_module.c:235-252 (binding CPET/utils/c_ops.py:52-57). */
int einsum_ij_i(int rows, int cols, const float *A, float *ret);

/* ============ Part 2: batched / device-resident API ====================== */

typedef struct cpet_ctx cpet_ctx;

/* Description of the last failure on this thread ("" if none). */
const char *cpet_last_error(void);
/* Number of visible GPUs (0 when none). */
int cpet_device_count(int *n);

/* A context owns one device, one HIP stream and grow-only scratch buffers.
 * Calls on one context are serialised; contexts run concurrently.
 * No counterpart in the reference (which has no device state). */
int cpet_ctx_create(int device, cpet_ctx **out);
int cpet_ctx_destroy(cpet_ctx *ctx);
/* Queue the context's work on the caller's stream (e.g.
 * torch.cuda.current_stream().cuda_stream); NULL is the null (legacy
 * default) stream, the handle torch reports for its default stream.  With
 * device pointers the caller's allocator sees the work in its own stream
 * order.  cpet_ctx_use_own_stream returns to the context's private stream. */
int cpet_ctx_set_stream(cpet_ctx *ctx, void *hip_stream);
int cpet_ctx_use_own_stream(cpet_ctx *ctx);
int cpet_ctx_get_stream(cpet_ctx *ctx, void **hip_stream);
int cpet_ctx_sync(cpet_ctx *ctx);

/* Field and/or potential at M arbitrary points pts (M,3).
 *   E_out  (M,3) or NULL: softened field as compute_looped_field
 *                  (math_module.c:405-451) when soften != 0, else the
 *                  unsoftened field of calc_field (math_module.c:255-293);
 *   phi_out (M)  or NULL: k sum q/r as calc_esp_base (math_module.c:453-486).
 * Outputs are float, or double with CPET_F_OUT_F64.
 * Batched replacement of compute_field_on_grid / compute_ESP_on_grid
 * (CPET/utils/calculator.py:430-475), which call the two C kernels above. */
int cpet_field_points(cpet_ctx *ctx, int M, int N, const float *pts,
                      const float *x, const float *q, void *E_out,
                      void *phi_out, int soften, uint32_t flags);

/* The same on the tensor-product grid ax[nx] x ay[ny] x az[nz] with
 * row-major ('ij', z fastest) point order — the mesh built by
 * initialize_box_points_uniform (CPET/utils/calculator.py:171-248) and
 * flattened by compute_field_on_grid (CPET/utils/calculator.py:442).
 * Same per-pair arithmetic as cpet_field_points on the flattened mesh; the
 * two agree to FP32 rounding (bitwise when both use the same charge split). */
int cpet_field_grid(cpet_ctx *ctx, int nx, int ny, int nz, const float *ax,
                    const float *ay, const float *az, int N, const float *x,
                    const float *q, void *E_out, void *phi_out, int soften,
                    uint32_t flags);

/* L streamlines (thread_operation semantics, math_module.c:523-591; the
 * topo_GPU path CPET/source/calculator.py:793-978 computes the same).
 *   starts (L,3) float, n_iter (L) int32, h, dims (3) float
 *   feats  (L,2) double out: [distance, mean curvature] per start, input order
 *   steps  (L) int32 out or NULL: steps taken (Euler, or RK4 with CPET_F_TOPO_RK4)
 *   exited (L) int32 out or NULL: 1 if stopped by the box, 0 by n_iter
 *   hist   (nb_d*nb_c) uint32 out or NULL: 2-D histogram of the float32-
 *          rounded features with numpy.histogramdd edge rules over the
 *          given edges (nb_d+1 and nb_c+1 doubles, increasing), as
 *          np.histogram2d does in CPET/utils/calculator.py:702-707.
 * Positions are integrated in FP64 with FP32 tiled field sums; the four
 * curvature-defining evaluations run in FP64.  With CPET_F_PREC_F64 the
 * field sums of the walk are FP64 as well (a precision reference, slower);
 * the other precision modes give the FP32 tiled sums.  CPET_F_TOPO_RK4
 * selects RK4 steps (see the flag).
 * dims (like h) is a host array in every mode; with CPET_F_DEVICE the other
 * arrays, the edges included, are device pointers, and with CPET_F_NO_SYNC
 * as well the call neither waits on nor reads back from the stream (so it
 * can be captured into a HIP graph once its scratch has grown). */
int cpet_topo(cpet_ctx *ctx, int N, int L, const float *starts,
              const int32_t *n_iter, float h, const float *dims,
              const float *x, const float *q, double *feats, int32_t *steps,
              int32_t *exited, int nb_d, const double *edges_d, int nb_c,
              const double *edges_c, uint32_t *hist, uint32_t flags);

/* Bytes of the fixed device scratch array of cpet_topo_noalloc (one copy
 * per GPU, held until the process ends). */
#define CPET_NOALLOC_SCRATCH_BYTES (32u << 20)

/* cpet_topo that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * array but dims (host, 3 floats) is a device pointer of the current device
 * (CPET_F_DEVICE is implied); the work is queued on hip_stream (NULL: the
 * null stream), and the call waits for it unless flags has CPET_F_NO_SYNC.
 * All scratch comes from the library's fixed array of
 * CPET_NOALLOC_SCRATCH_BYTES: about 60 bytes per streamline plus 16 per
 * charge (e.g. 100,000 streamlines over 50,000 charges take 6.8 MB); a
 * call that needs more returns CPET_ERR_BUDGET.  Default precision and the
 * Euler walk only (CPET_F_PREC_F64 and CPET_F_TOPO_RK4 are rejected).
 * Calls on one GPU share the array: calls that overlap on the device (on
 * different streams) race.  Results equal cpet_topo's bitwise. */
int cpet_topo_noalloc(void *hip_stream, int N, int L, const float *starts,
                      const int32_t *n_iter, float h, const float *dims, const float *x,
                      const float *q, double *feats, int32_t *steps, int32_t *exited, int nb_d,
                      const double *edges_d, int nb_c, const double *edges_c, uint32_t *hist,
                      uint32_t flags);

/* L dipole streamlines: cpet_topo with point dipoles x (N,3), mu (N,3) as
 * the sources (calc_field_base_dipole, math_module.c:335-371; walk and
 * features of thread_operation_dipole, :593-659).  The pair term is the
 * reference's unless flags has CPET_F_DIPOLE_PHYSICAL (3 (mu.R) R - mu r^2).
 * dims is a host array; the other arrays follow CPET_F_DEVICE. */
int cpet_topo_dipole(cpet_ctx *ctx, int N, int L, const float *starts, const int32_t *n_iter,
                     float h, const float *dims, const float *x, const float *mu,
                     double *feats, int32_t *steps, int32_t *exited, uint32_t flags);

/* ---- many frames per launch (batching) ---------------------------------
 * Per-frame sizes and scalars (n_charges, n_starts, grid_dims, h, dims)
 * are host arrays; the concatenated per-frame data (x, q, axes, starts,
 * n_iter) and the outputs are host or device pointers (CPET_F_DEVICE).
 * Frame f's charges are rows [sum_{g<f} n_charges[g], ...) of x and q.
 * No single counterpart in the reference, which loops over frames in Python
 * (CPET/source/CPET.py:96-229) and calls the per-frame kernels above. */

/* Grids: frame f has the mesh of grid_dims[3f..3f+2] axis lengths whose
 * values follow in `axes` (x values, then y, then z, frame after frame).
 * Outputs are concatenated in frame order (E (sum M_f, 3), phi (sum M_f)).
 * One launch for all frames. */
int cpet_batch_volume(cpet_ctx *ctx, int F, const int32_t *n_charges,
                      const float *x, const float *q, const int32_t *grid_dims,
                      const float *axes, void *E_out, void *phi_out, int soften,
                      uint32_t flags);

/* Streamlines of F frames in one persistent launch: frame f has n_starts[f]
 * starts (concatenated in `starts` / `n_iter`), step h[f] and half-widths
 * dims[3f..3f+2].  feats (sum L, 2), steps, exited as cpet_topo in frame
 * order; hist (F, nb_d, nb_c) per-frame histograms over shared edges. */
int cpet_batch_topo(cpet_ctx *ctx, int F, const int32_t *n_charges,
                    const float *x, const float *q, const int32_t *n_starts,
                    const float *starts, const int32_t *n_iter, const float *h,
                    const float *dims, double *feats, int32_t *steps,
                    int32_t *exited, int nb_d, const double *edges_d, int nb_c,
                    const double *edges_c, uint32_t *hist, uint32_t flags);

/* ---- frame preparation (filters, box-frame transform, compaction) ----
 * dtype codes: 0 = float32, 1 = float64.  `x` holds N rows of 3 values in
 * x_dtype; arithmetic on (x - center) runs in sub_dtype (numpy's promotion of
 * the two operands); center / Tinv / box_shift / dims are passed as doubles
 * holding values of that type.  Outputs are caller-allocated; host or device
 * pointers as for the other calls (CPET_F_DEVICE). */

/* keep[i] = 1 iff row i's residue run (runs[i], rows with equal (resid,
 * resnum) that are adjacent) has an atom with sqrt((d0*d0+d1*d1)+d2*d2) <
 * radius, d = x - center, no FMA.  Replaces the decision of
 * filter_radius_whole_residue, CPET/utils/io.py:283-338. */
int cpet_filter_radius_residue(cpet_ctx *ctx, int N, const void *x, int x_dtype,
                               const double *center, int sub_dtype, double radius,
                               const int32_t *runs, int n_runs, uint8_t *keep,
                               uint32_t flags);

/* out = ((x - center) @ Tinv) - box_shift (Tinv row-major 3x3; NULL Tinv
 * means no rotation, NULL box_shift no shift).  x - center runs in
 * sub_dtype, the dot products (d0*T0j + d1*T1j) + d2*T2j (no FMA) in
 * dot_dtype, the shift in shift_dtype; out is stored in out_dtype.
 * Replaces the box-frame steps of calculator.__init__,
 * CPET/source/calculator.py:341-346,369-371. */
int cpet_frame_transform(cpet_ctx *ctx, int N, const void *x, int x_dtype,
                         const double *center, int sub_dtype, const double *Tinv,
                         int dot_dtype, const double *box_shift, int shift_dtype,
                         void *out, int out_dtype, uint32_t flags);

/* keep[i] = 0 iff every (x - center)_k lies strictly inside (-dims_k,
 * dims_k), compared in float64.  Replaces the decision of filter_in_box,
 * CPET/utils/io.py:422-475. */
int cpet_filter_in_box(cpet_ctx *ctx, int N, const void *x, int x_dtype,
                       const double *center, int sub_dtype, const double *dims,
                       uint8_t *keep, uint32_t flags);

/* Stable compaction of the kept rows: x_out (count,3), q_out (count) and
 * idx_out (count, original row indices; may be NULL); *count = kept rows.
 * The np.delete / boolean-mask step of the same filters. */
int cpet_compact_charges(cpet_ctx *ctx, int N, const uint8_t *keep, const float *x,
                         const float *q, float *x_out, float *q_out,
                         int32_t *idx_out, int *count, uint32_t flags);

/* Opt-in per-phase timing with HIP events recorded on the context's stream
 * (the stream the kernels run on).  After a call returns (or after
 * cpet_ctx_sync for CPET_F_NO_SYNC calls), cpet_ctx_phase_times gives the
 * durations in ms of the last call's phases:
 *   cpet_topo:          [order, propagate, epilogue(+histogram)]
 *   cpet_field_*:       [field kernel, split reduction, FP64 recompute]
 *   cpet_batch_volume:  [field kernel, FP64 recompute]
 * No counterpart in the reference (its timing is time.time() around calls,
 * CPET/source/calculator.py:686,706-711). */
int cpet_ctx_set_profiling(cpet_ctx *ctx, int on);
int cpet_ctx_phase_times(cpet_ctx *ctx, float *ms, int n);

/* Number of grid / points the last field call on this context recomputed in
 * FP64 because the auto-precision check found strong cancellation there
 * (u * sum|q|/r^2 > 4e-6 * |E|); 0 when that call had no check (precision
 * modes other than auto, or no E requested).  Waits for the context's
 * stream.  No counterpart in the reference (its sums are FP32 throughout,
 * CPET/utils/math_module.c:405-451). */
int cpet_ctx_refined_points(cpet_ctx *ctx, long long *n);

/* Tuning knobs (autotuner): 0 restores the built-in heuristic.
 *   cpet_set_tuning(kind, a, b): kind 0 = grid field (a = points per lane
 *   along z, b = charge splits), kind 1 = points field (a = points per lane,
 *   b = splits), kind 2 = topo (a = waves per workgroup, b = workgroups),
 *   kind 3 = batch topo (a = waves per workgroup, b = persistent
 *   workgroups), kind 4 = batch volume (a = points per lane), kind 5 =
 *   streamline pair term (0 or 1: the direct form; 2: the expanded
 *   far-field form for frames whose charges are all far from the walk
 *   region, the direct form for the others; plain charges, 4 or more waves
 *   per workgroup).  Streamline results are the same bit for bit with 16
 *   waves and with 1 wave per workgroup (the one wave walks the 16 waves'
 *   charge slices in order); 2, 4 and 8 waves split the charges
 *   differently and agree to FP32 rounding.  The default is 1 wave for a
 *   batch whose frames all have at most 128 charges and for a single frame
 *   of at most 16 charges and at least 50,000 starts, else 16. */
int cpet_set_tuning(cpet_ctx *ctx, int kind, int a, int b);

/* ------------------------------------------------------ cluster step
 * The step after "topo" (CPET/source/cluster.py:126-130): per-frame
 * histograms of the streamline features and their chi-square distances. */

/* Per-frame normalised 2-D histograms, as make_histograms
 * (CPET/utils/calculator.py:596-718) builds them: for frame f (feature rows
 * offsets[f] .. offsets[f+1]-1 of feats (P,2) = [distance, curvature]),
 * counts over the shared edges with numpy.histogramdd's rule (searchsorted
 * right; a value equal to the last edge in the last bin; others outside are
 * dropped), then divided by the frame's own in-range count (0/0 = NaN, as
 * numpy), or the counts themselves with CPET_F_COUNTS (make_histograms_mem,
 * CPET/utils/calculator.py:721-845, normalises numpy's density on the host).
 * hists (F, nb_d*nb_c) double, row-major [bd][bc].  offsets (F+1,
 * int64, starting at 0) and the edges (nb_d+1, nb_c+1 doubles, increasing)
 * are host arrays; feats and hists are device pointers with CPET_F_DEVICE. */
int cpet_cluster_histograms(cpet_ctx *ctx, int F, const int64_t *offsets, const double *feats,
                            int nb_d, const double *edges_d, int nb_c, const double *edges_c,
                            double *hists, uint32_t flags);

/* Order statistics of the columns of x (n, cols) row-major double, for the
 * shared bins of make_histograms: its np.max / np.min and the 25th / 75th
 * percentiles inside scipy.stats.iqr (CPET/utils/calculator.py:664-670) are
 * all k-th smallest values of a column.  values[k] (host, n_sel doubles) =
 * the sel_ranks[k]-th smallest (0-based) value of column sel_cols[k], an
 * element of that column, exact; NaNs order after every number, as in
 * numpy's sort, and -0.0 is returned as +0.0.  nan_counts (host, cols int64,
 * or NULL) = the NaNs of each column, for numpy's rule that a NaN anywhere
 * makes max, min and percentile NaN.  1 <= cols <= 8, 1 <= n_sel <= 16,
 * 1 <= n < 2^32; x is a device pointer with CPET_F_DEVICE; the call waits
 * for the stream. */
int cpet_order_stats(cpet_ctx *ctx, int64_t n, int cols, const double *x, int n_sel,
                     const int32_t *sel_cols, const int64_t *sel_ranks, double *values,
                     int64_t *nan_counts, uint32_t flags);

/* out (F,F) double: out[i][j] = sum over bins with h_i + h_j != 0 of
 * (h_i - h_j)^2 / (h_i + h_j), halved; zero diagonal.  Replaces
 * distance_numpy + construct_distance_matrix (CPET/utils/calculator.py:
 * 975-978, 1000-1015; sklearn pairwise_distances over that metric).  FP64;
 * bins are summed in order (numpy sums pairwise: agreement to FP64
 * rounding). */
int cpet_chi2_matrix(cpet_ctx *ctx, int F, int B, const double *hists, double *out,
                     uint32_t flags);

/* ------------------------------------------------------ frame text I/O
 * Host-only (no GPU needed).  These feed frames to the kernels and write
 * their results, at the rate a serving loop needs. */

/* Counts ATOM/HETATM records of a PDB file (CPET/utils/io.py:596-598). */
int cpet_pdb_scan(const char *path, int *n_atoms);

/* Parses the n_atoms ATOM/HETATM records of a PDB file exactly as
 * CPET/utils/io.py:574-636 (parse_pdb) does: fixed columns serial [6:11],
 * name [12:16], resName [17:20], chain [21], resSeq [22:26], x y z
 * [30:38] [38:46] [46:54] and, when get_charges, the charge [55:64]; fields
 * are whitespace-stripped, numbers parsed to correctly rounded doubles.
 * Outputs: xyz[3*n_atoms], q[n_atoms] (get_charges), atom_number, resnum,
 * atom_type[4*n_atoms], resname[3*n_atoms] (NUL-padded), chain[n_atoms].
 * Returns CPET_ERR_ARG, naming the line, where Python's int()/float() would
 * raise; CPET_ERR_ARG with a message starting "non-ASCII" when a record
 * holds non-ASCII bytes (byte columns then differ from Python's character
 * columns). */
int cpet_pdb_parse(const char *path, int n_atoms, int get_charges, double *xyz, double *q,
                   int32_t *atom_number, int32_t *resnum, char *atom_type, char *resname,
                   char *chain);

/* Parses the n_atoms ATOM/HETATM records of a PQR file as
 * CPET/utils/io.py:478-571 (parse_pqr) does: whitespace tokens give the
 * serial, atom name, residue name and residue number (a first token longer
 * than 6 characters holds the serial from its 7th character and shifts the
 * rest by one; a residue-number token longer than 4 characters is replaced
 * by the previous token without its first character); residue numbers above
 * 999 shift the coordinate columns [30:38] [38:46] [46:54] and the charge
 * column [54:61] right by (digits - 3).  Outputs: xyz[3*n_atoms],
 * q[n_atoms], atom_number, resnum, and atom_type / resname in `width`-byte
 * NUL-padded slots.  Count the records with cpet_pdb_scan (same rule).
 * Returns CPET_ERR_ARG, naming the line, for a record Python would reject
 * and for one this parser does not read exactly as Python would (non-ASCII
 * bytes, digit-group underscores, a token of width or more bytes). */
int cpet_pqr_parse(const char *path, int n_atoms, int width, double *xyz, double *q,
                   int32_t *atom_number, int32_t *resnum, char *atom_type, char *resname);

/* Writes an optional header then rows x cols values, byte-identical to
 * numpy.savetxt(path, data, fmt) for fmt 0 = "%.3f" (the volume .dat rows,
 * CPET/utils/io.py:97-101) and fmt 1 = "%.18e" (the .top histogram,
 * CPET/source/CPET.py:123).  dtype 0 = float32, 1 = float64, 2 = float16.
 * Rows are formatted by n_threads threads (0 = all cores). */
int cpet_write_rows(const char *path, const char *header, const void *data, int dtype,
                    int64_t rows, int cols, int fmt, int n_threads);

/* Reads the numeric rows of a text file as make_histograms reads a .top
 * file (CPET/utils/calculator.py:603-633): lines starting with '#' are
 * skipped; every other line gives its first `cols` whitespace-separated
 * values, parsed as Python's float() does (correctly rounded).  With out ==
 * NULL only counts the rows (*n_rows); otherwise fills out (cap rows x cols)
 * and sets *n_rows.  A line with fewer than cols values, or an unparsable
 * value, is an error naming the line. */
int cpet_read_rows(const char *path, int cols, double *out, int64_t cap, int64_t *n_rows);

#ifdef __cplusplus
}
#endif

#endif /* CPET_HIP_H */
