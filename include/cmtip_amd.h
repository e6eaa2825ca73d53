/*
 * cmtip_amd C ABI: the hot path of cartesian MTIP reconstruction (apeck12/cmtip) as
 * hand-written HIP kernels for MI355X (gfx950).
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t passed as
 * void*; none allocates host-visible state beyond the plan/solver objects, none
 * synchronises the device unless its comment says so, and every one returns an int
 * status (CMTIP_OK on success).  No exception crosses the ABI.  After a failure,
 * cmtip_last_error() returns a message for the calling thread.
 *
 * Array conventions (the reference's CPU path, cmtip/nufft/forward.py:40 and
 * cmtip/nufft/adjoint.py:33, is normative):
 *   - Uniform grids are (M, M, M) row-major (last axis fastest).  Array index i along
 *     an axis holds Fourier mode k = i - floor(M/2).
 *   - Point coordinate x pairs with array axis 0, y with axis 1, z with axis 2.
 *   - Type 2 ("forward"):  c_j = sum_k f_k exp(-i k.x_j)       (iflag -1)
 *     Type 1 ("adjoint"):  f_k = sum_j c_j exp(+i k.x_j)       (iflag +1)
 *   - Coordinates must satisfy |x| < 3*pi (forward.py:27, adjoint.py:29).
 *   - Complex arrays are interleaved float pairs (numpy complex64 / torch.complex64).
 *
 * Plans and solvers are not thread-safe; use one per stream.
 */
#ifndef CMTIP_AMD_H
#define CMTIP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define CMTIP_OK 0
#define CMTIP_ERR_ARG 1     /* bad size, pointer or flag combination */
#define CMTIP_ERR_ALLOC 2   /* device allocation failed */
#define CMTIP_ERR_HIP 3     /* a HIP runtime call failed */
#define CMTIP_ERR_LAUNCH 4  /* a kernel launch failed */
#define CMTIP_ERR_RANGE 5   /* a nonuniform point lies outside [-3pi, 3pi) */
#define CMTIP_ERR_STATE 6   /* call order violated (e.g. execute before setpts) */

/* Message for the last failure on the calling thread ("" if none). */
const char* cmtip_last_error(void);
/* ABI version: major*10000 + minor*100 + patch. */
int cmtip_version(void);

/* ---------------------------------------------------------------- NUFFT plans
 * Replace the cufinufft v1 plan object used by the reference GPU path
 *   cufinufft(type, n_modes, n_trans=1, isign, eps, dtype=float32, gpu_method=1)
 *   -> set_pts(...) -> execute(c, f)                (cmtip/nufft/forward.py:77-81,
 *                                                    cmtip/nufft/adjoint.py:73-76)
 * and finufft.nufft3d2 / finufft.nufft3d1 on the CPU path (forward.py:40,
 * adjoint.py:33).  FP32 arithmetic; the exponential-of-semicircle kernel width w
 * follows from eps (w = ceil(log10(1/eps)) + 1, 2 <= w <= 16, upsampling 2).
 */
typedef struct cmtip_nufft_plan_s* cmtip_nufft_plan;

/* type: 1 or 2.  n_modes: M (cubic).  iflag: sign of the exponent (+1 / -1).
 * The plan owns its fine grid and FFT scratch.  Allocates device memory. */
int cmtip_nufft_makeplan(int type, int n_modes, int iflag, double eps, cmtip_nufft_plan* plan);

/* Kernel width, fine-grid edge and ES shape parameter of a plan (any pointer may be NULL). */
int cmtip_nufft_info(cmtip_nufft_plan plan, int* w, int* n_fine, double* beta);

/* Register npts nonuniform points (device float arrays, caller-owned, kept by
 * reference until the next setpts).  Replaces cufinufft set_pts (forward.py:78,
 * adjoint.py:74).  The points are also sorted into the spreading / interpolation
 * bins here (counting sort on the device). */
int cmtip_nufft_setpts(cmtip_nufft_plan plan, int64_t npts, const float* x, const float* y,
                       const float* z, void* stream);

/* Type-1 plans: register the points together with real strengths c (npts floats).  Each
 * point's bin-sorted record carries its strength in place of its index, and max |c| and
 * the sign of c are recorded while the points are binned, so a following execute with
 * CMTIP_PTS_BOUND (c ignored) spreads without gathering c by index and without a pass
 * over c.  No counterpart in the reference: cufinufft takes the strengths at execute
 * (adjoint.py:75); this serves solve_ac's right-hand side (autocorrelation.py:162-168),
 * whose strengths are known when the points are registered.  Until the next setpts the
 * plan executes only CMTIP_PTS_BOUND or CMTIP_PTS_ONES (CMTIP_ERR_STATE otherwise). */
int cmtip_nufft_setpts_strengths(cmtip_nufft_plan plan, int64_t npts, const float* x, const float* y,
                                 const float* z, const float* c, void* stream);

/* Register the n_orient * P rotated pixels of cmtip_rotate_coords (quats: n_orient x 4
 * float, s: 3 x P float64, recip_extent as there; point j = o * P + p) without storing
 * their coordinate arrays: the coordinates are generated where the points are binned
 * and copied into the bin-sorted records, the same floats cmtip_rotate_coords writes.
 * c (type-1 plans, optional, n_orient * P floats): strengths bound as by
 * cmtip_nufft_setpts_strengths.  Replaces the gen_nonuniform_positions + set_pts pair of
 * solve_ac (autocorrelation.py:208-209 then adjoint.py:74); the range check is available
 * from cmtip_nufft_points_max_abs.  The plan cannot be autotuned until the next setpts
 * (CMTIP_ERR_STATE), which needs arrays to bin again.  The quaternions and s are read
 * while this call's kernels run on the stream. */
int cmtip_nufft_setpts_rotated(cmtip_nufft_plan plan, const float* quats, int64_t n_orient, const double* s,
                               int64_t P, float recip_extent, const float* c, void* stream);

/* Largest |coordinate| of the points registered by the last setpts (NaN counts as
 * +inf), recorded while the points were binned: the range check of forward.py:27,64 and
 * adjoint.py:29,66 without another pass over the points.  Synchronises the stream.
 * CMTIP_ERR_STATE when the plan's points were replaced by a slicing route. */
int cmtip_nufft_points_max_abs(cmtip_nufft_plan plan, float* max_abs, void* stream);

/* execute flags */
#define CMTIP_GRID_REAL 1   /* the uniform grid f is float32 (real); else complex64 */
#define CMTIP_PTS_REAL 2    /* the point values c are float32 (real); else complex64 */
#define CMTIP_PTS_ONES 4    /* type 1 only: every strength is 1 (c is ignored), as the
                               all-ones kernel input of autocorrelation.py:143,146 */
#define CMTIP_PTS_IMAG 8    /* type 2 only: c (float) receives Im(T2(f)), e.g. the
                               coordinate gradient of a real slice */
#define CMTIP_PTS_BOUND 32  /* type 1 only: the real strengths registered with the points by
                               cmtip_nufft_setpts_strengths or cmtip_nufft_setpts_rotated
                               (c is ignored) */
#define CMTIP_EXEC_ADJOINT 16 /* run the adjoint of the plan's transform on its registered
                               points: a type-2 plan spreads (type 1, sign flipped), a
                               type-1 plan interpolates (type 2, sign flipped); c and f
                               keep their meaning (c: values at the points, f: the
                               uniform grid) and the flags above apply to the direction
                               executed.  The bin size must suit that direction (the
                               defaults do). */

/* Type 2: c = T2(support (.) f); the support (float32 M^3, may be NULL) multiplies f
 * on load (forward.py:34-35); with CMTIP_PTS_REAL only Re(c) is written.
 * Type 1: f = scale * support (.) T1(c); with CMTIP_GRID_REAL only Re(f) is written;
 * CMTIP_PTS_REAL means the strengths c are real.  The reference's adjoint is
 * scale = 1/M^3 with the support applied after the transform (adjoint.py:35-43).
 * Replaces cufinufft execute (forward.py:79, adjoint.py:75). */
int cmtip_nufft_execute(cmtip_nufft_plan plan, void* c, void* f, int flags, const float* support,
                        float scale, void* stream);

int cmtip_nufft_destroy(cmtip_nufft_plan plan);

/* Bin and launch settings of a plan, used from the next setpts on: bin_size = edge
 * of a spatial bin in fine cells (its LDS subgrid is bin_size + w - 1 per side),
 * chunk = points per unit of work (64..4096), wg_per_cu = persistent workgroups per
 * CU of the spread / interpolation kernels; 0 selects the default of each.  Replaces
 * cufinufft's plan options (gpu_method=1 with its default bin sizes and subproblem
 * size, forward.py:77, adjoint.py:73), which the reference leaves at their defaults. */
int cmtip_nufft_set_bins(cmtip_nufft_plan plan, int bin_size, int chunk, int wg_per_cu);
int cmtip_nufft_get_bins(cmtip_nufft_plan plan, int* bin_size, int* chunk, int* wg_per_cu);

/* Autotuner for the settings above: with the plan's points registered, for each of
 * the n_cand candidates (cands: n_cand x {bin_size, chunk, wg_per_cu}) re-bins the
 * points and runs cmtip_nufft_execute(c, f, flags, support, scale) once untimed and
 * reps times timed (HIP events on the stream; the call synchronises).  times_ms
 * (2 x n_cand floats, may be NULL) receives {setpts ms, execute ms} per candidate, -1
 * for candidates that do not fit the plan.  The plan keeps the candidate with the
 * least setpts_weight * setpts + execute, re-binned; *best (may be NULL) is its index.
 * Outputs of execute (f for type 1, c for type 2) are overwritten.  The reference has
 * no counterpart: it is the bin / tile-size selection behind the spreading kernels. */
int cmtip_nufft_autotune(cmtip_nufft_plan plan, void* c, void* f, int flags, const float* support, float scale,
                         const int* cands, int n_cand, int reps, float setpts_weight, float* times_ms, int* best,
                         void* stream);

/* ---------------------------------------------------------------- rotated q-vectors
 * Coordinates of every pixel under every orientation, orientation-major
 * (j = o*P + p):  x_j = fl32(fl32((R(q_o)^-1 s_p)_0) / fl32(recip_extent)) * fl32(pi),
 * likewise y (row 1) and z (row 2).  R(q) is the scalar-first quaternion rotation of
 * cmtip/eval/geom.py:37-45; its inverse is taken exactly (adjugate / determinant, FP64)
 * as np.linalg.inv does in cmtip/alignment/align.py:51 and
 * cmtip/autocorrelation/autocorrelation.py:24.  The rotated vector is computed in FP64
 * and rounded once, as the reference's einsum + astype(float32) does (align.py:52-57,
 * autocorrelation.py:27,133-135,210).
 * quats: n_orient x 4 float32; s: 3 x P float64 (pixel_position_reciprocal).
 * recip_extent = 0 writes the unscaled fl32(R^-1 s) (gen_nonuniform_positions). */
int cmtip_rotate_coords(const float* quats, int64_t n_orient, const double* s, int64_t P,
                        float recip_extent, float* x, float* y, float* z, void* stream);

/* ---------------------------------------------------------------- fused slicing
 * compute_slices (cmtip/alignment/align.py:39-65) for a real autocorrelation, in two
 * steps.  cmtip_slice_prepare loads ac (float32 M^3) into the type-2 plan's fine grid
 * (deconvolution, zero padding and three FFT passes; the grid kept is Re(FFT), which
 * is exactly what Re(c) needs for a real ac).  cmtip_slice_real then evaluates
 * Re(T2(ac)) at the rotated pixel coordinates, generating them in-kernel from
 * (quats, s, recip_extent) exactly as cmtip_rotate_coords does, and writes
 * out[o*P + p] (float32).  order (optional, P int32, a permutation of the pixels)
 * is the order in which pixels are processed, 64 at a time: pixels that are close on
 * the detector should be close in it (8 x 8 blocks of a square panel), which lets each
 * group of 64 stage its fine-grid footprint in LDS.  NULL processes pixels in index
 * order.  The result does not depend on it.  out_bf16 (optional, uint16 bfloat16 bits, round to
 * nearest even) receives the same values.  moments (optional, 2 doubles, accumulated,
 * not cleared) receives sum(out) and sum(out^2) for the reference's per-batch std
 * scaling (align.py:122-124).
 * Three routes compute the same sums (FP32 additions in different orders): fused
 * tiles (coordinates, an LDS box per 64 pixels and interpolation in one kernel),
 * binned tiles (tiles of 64 pixels sorted by fine-grid bin; one LDS subgrid serves all
 * tiles of a bin) and binned points (pixels sorted by bin, for large M where tile
 * footprints are wide).  The binned routes replace the plan's registered points: call
 * cmtip_nufft_setpts again before cmtip_nufft_execute.  With the automatic route the
 * first call for a detector geometry (s, P, recip_extent, order: compared by pointer
 * and value; cmtip_slice_set_route clears it) probes the fused kernel on up to 16
 * orientations and synchronises the stream once. */
int cmtip_slice_prepare(cmtip_nufft_plan plan, const float* ac, const float* support, void* stream);
int cmtip_slice_real(cmtip_nufft_plan plan, const float* quats, int64_t n_orient, const double* s,
                     const int* order, int64_t P, float recip_extent, float* out, uint16_t* out_bf16,
                     double* moments, void* stream);

/* Slicing route of a plan: 0 automatic (default: binned tiles when at least 95% of the
 * probed tiles fit the tile route's margin, the call has at least 64 tiles per bin and
 * the fine grid is at least 128 cells wide; else fused tiles for calls of at most 2^20
 * points when at least 80% of their boxes fit LDS; else binned points), 1 fused tiles,
 * 2 binned points, 3 binned tiles.
 * A fixed route never
 * synchronises (use one under stream capture).  get_route reports the request and the
 * route the last cmtip_slice_real used (0 before the first).  No counterpart in the
 * reference, whose compute_slices (align.py:51-63) always takes the generic NUFFT. */
int cmtip_slice_set_route(cmtip_nufft_plan plan, int route);
int cmtip_slice_get_route(cmtip_nufft_plan plan, int* requested, int* last_used);

/* Kernel timer of a plan's slicing: start = 1 starts recording HIP events on the
 * stream around the main kernel of each following cmtip_slice_real call (the
 * interpolation kernel of the fused and tile routes, the whole binned-points route), up
 * to 64 calls, without synchronising; start = 0 waits for the recorded events and
 * returns their total time and count, then stops.  A measurement hook with no
 * counterpart in the reference. */
int cmtip_slice_timer(cmtip_nufft_plan plan, int start, float* total_ms, int* count);

/* ---------------------------------------------------------------- matching
 * Orientation matching (cmtip/alignment/align.py:113-131): for every image i, the
 * reference r minimising the weighted Euclidean distance
 *     d2[r,i] = sum_p w_p^2 (gamma a[r,p] - b[i,p])^2
 * (sklearn euclidean_distances, align.py:35, on model_slices*gamma*w and slices*w).
 * Both operands are first centred on the same per-pixel vector c (the mean image,
 * cmtip_image_stats), which leaves every d2 unchanged but removes the common
 * intensity profile, so the expanded form
 *     d2 = an[r] - 2 <a'_r, b'_i> + bn[i],  a' = gamma a - c,  b' = w^2 (.) (b - c),
 *     an = sum w^2 (gamma a - c)^2,  bn = sum w^2 (b - c)^2
 * does not cancel catastrophically in FP32 / BF16 when the images have a large
 * dynamic range.  a' and an come from cmtip_center_rows, b' from
 * cmtip_match_prepare_images_centered and bn from cmtip_center_rows with gamma = 1.
 * best[i] packs (float bits of max(d2,0)) << 32 | r and is reduced with atomicMin,
 * so ties keep the smallest reference index (np.argmin, align.py:131).  Callers
 * initialise best to all ones bits before the first batch; ref_offset is added to r
 * (the batch start, align.py:116).  gamma may be NULL (1). */
/* Per-pixel sums of the images (colsum, 2P doubles of workspace: sums then sums of
 * squares), their mean image c (float32, may be NULL) and mom = (sum, sum of squares)
 * over all entries (2 doubles, may be NULL; for slices_.std(), align.py:122), all in
 * FP64 with a fixed summation order. */
int cmtip_image_stats(const float* X, int64_t rows, int64_t P, double* colsum, float* c, double* mom,
                      void* stream);
/* cmtip_image_stats without the colsum workspace argument: the 2P doubles live in a
 * file-scope device array of the library (4 MiB per device, held until the process
 * ends), so the call allocates nothing and needs no caller workspace.  No counterpart
 * in the reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  P at most
 * 262144 pixels (CMTIP_ERR_ARG beyond).  All calls on one GPU share the array, so calls
 * that overlap on one GPU (other streams, other threads) race; calls ordered on one
 * stream do not.  Same kernels and results as cmtip_image_stats. */
int cmtip_image_stats_noalloc(const float* X, int64_t rows, int64_t P, float* c, double* mom, void* stream);
/* Data/model standard-deviation ratio of one reference batch (align.py:122-124: both
 * population stds, the ratio rounded to float32 as numpy keeps it), written to the n
 * doubles of gamma.  img_mom = (sum, sum of squares[, count]) of the images with
 * img_count entries, or with the count in img_mom[2] (at least 1) when img_count <= 0
 * (masked images); model_mom = (sum, sum of squares) of model_count slice entries.
 * FP64 in the order sqrt(max(s2 / n - (s1 / n)^2, 0)), without contraction. */
int cmtip_match_gamma(const double* img_mom, double img_count, const double* model_mom, double model_count,
                      int64_t n, double* gamma, void* stream);
/* y = gamma x - c per row (c may be NULL), norms = sum_p w2 y^2; y written to Y (float32,
 * may alias X) and/or Y16 (bfloat16 bits) when given. */
int cmtip_center_rows(const float* X, int64_t rows, int64_t P, const double* gamma, const float* c,
                      const float* w2, float* Y, uint16_t* Y16, float* norms, void* stream);
/* Bw = w2 (.) (B - c) as float32 and/or bfloat16 (c may be NULL). */
int cmtip_match_prepare_images_centered(const float* B, int64_t n_img, int64_t P, const float* w2,
                                        const float* c, float* Bw, uint16_t* Bw_bf16, void* stream);
/* cmtip_match_prepare_images_centered and the image norms of cmtip_center_rows (gamma
 * = 1: norms = sum w2 (b - c)^2) in one pass over the images when P is a multiple of 4
 * and the rows are aligned (two passes otherwise); the same values either way. */
int cmtip_match_prepare_images_norms(const float* B, int64_t n_img, int64_t P, const float* w2,
                                     const float* c, float* Bw, uint16_t* Bw_bf16, float* norms, void* stream);
int cmtip_weighted_norms(const float* X, int64_t rows, int64_t P, const float* w2, float* out,
                         void* stream);
int cmtip_match_prepare_images(const float* B, int64_t n_img, int64_t P, const float* w2,
                               float* Bw, uint16_t* Bw_bf16, void* stream);
int cmtip_fill_u64(unsigned long long* dst, int64_t count, unsigned long long value, void* stream);

/* FP32 path: cross term on MFMA (v_mfma_f32_16x16x4f32), argmin fused in the epilogue. */
int cmtip_match_argmin_f32(const float* A, int64_t n_ref, int64_t ref_offset, const float* Bw,
                           int64_t n_img, int64_t P, const float* an, const float* bn,
                           const double* gamma, unsigned long long* best, void* stream);

/* calc_eudist (align.py:27-36): D[r, i] = sqrt(max(an[r] - 2 <a_r, b_i> + bn[i], 0)) as
 * float32 (n_a x n_b, row-major), with an, bn the plain squared row norms
 * (cmtip_weighted_norms with w2 = 1).  Same MFMA kernel, full-matrix epilogue. */
int cmtip_eudist_f32(const float* A, int64_t n_a, const float* B, int64_t n_b, int64_t P,
                     const float* an, const float* bn, float* D, void* stream);

/* BF16 path: cross term in bfloat16 on MFMA with FP32 accumulation (this library's
 * kernels: 256 x 256 tiles, direct-to-LDS operand pipeline for aligned rows with K a
 * multiple of 64, register-staged otherwise); keeps the k (1 <= k <= 16) smallest
 * approximate distances per image in cand (n_img x k packed
 * keys as above, initialised to all ones bits before the first batch), then
 * cmtip_match_rescore recomputes each candidate's distance in FP32 from the float32
 * slices (d2 = sum_p w_p^2 (gamma a - b)^2 accumulated in FP32) and writes the argmin
 * to best.  gamma_of_ref maps a reference index to its batch gamma (n_ref_total
 * doubles). */
int cmtip_match_topk_bf16(const uint16_t* A, int64_t n_ref, int64_t ref_offset,
                          const uint16_t* Bw, int64_t n_img, int64_t P, const float* an,
                          const float* bn, const double* gamma, int k,
                          unsigned long long* cand, void* stream);
/* The same BF16 cross term in two steps, for large blocks: cmtip_match_cross_bf16
 * writes the FP32 cross-term block Dt[i * n_ref + r] = <A_r, B_i> (bfloat16 rows of K
 * elements, FP32 accumulation) with the hipBLASLt GEMM; cmtip_match_topk_cross then
 * merges d2 = gamma^2 an[r] - 2 gamma Dt[i, r] + bn[i] into the image's k candidates
 * (1 <= k <= 16 here, in cmtip_match_topk_bf16 and in cmtip_match_rescore[_masked])
 * (one wave per image, no atomics; ties keep the smallest reference index).  Replaces
 * the same sklearn euclidean_distances + argmin (align.py:35,128-131). */
int cmtip_match_cross_bf16(const uint16_t* A, int64_t n_ref, const uint16_t* B, int64_t n_img, int64_t K,
                           float* Dt, void* stream);
int cmtip_match_topk_cross(const float* Dt, int64_t n_ref, int64_t ref_offset, int64_t n_img, const float* an,
                           const float* bn, const double* gamma, int k, unsigned long long* cand, void* stream);
int cmtip_match_rescore(const float* A_all, int64_t n_ref_total, const float* B, int64_t n_img,
                        int64_t P, const float* w2, const double* gamma_of_ref, int k,
                        const unsigned long long* cand, unsigned long long* best, void* stream);

/* Per-image pixel masks (uint8, n_img x P, 1 = valid; an extension for detectors with
 * beamstops, absent from the reference).  d2 = sum_p m_ip w_p^2 (gamma a - b)^2 is one
 * cross term: float32 rows over 2P columns, model [y^2, y] with y = gamma a - c
 * (cmtip_center_rows_sq) and image [-m w^2 / 2, m w^2 (b - c)] (cmtip_match_prepare_masked),
 * K = 2P; bfloat16 rows over 3P columns, model [hi(y^2), lo(y^2), y] (y^2 split into two
 * bfloat16 parts) and image [-m w^2 / 2, -m w^2 / 2, m w^2 (b - c)], K = 3P.  bn =
 * sum m w^2 (b - c)^2; the matching kernels run with an = 0.
 * cmtip_image_stats_masked: colsum (3P doubles workspace), c = masked column mean,
 * mom = (sum m x, sum m x^2, sum m). */
int cmtip_image_stats_masked(const float* X, const uint8_t* mask, int64_t rows, int64_t P, double* colsum,
                             float* c, double* mom, void* stream);
int cmtip_match_prepare_masked(const float* B, const uint8_t* mask, int64_t n_img, int64_t P, const float* w2,
                               const float* c, float* Bw2, uint16_t* Bw2_bf16, float* bn, void* stream);
int cmtip_center_rows_sq(const float* X, int64_t rows, int64_t P, const double* gamma, const float* c,
                         float* Y2, uint16_t* Y2_bf16, void* stream);
int cmtip_match_rescore_masked(const float* A_all, int64_t n_ref_total, const float* B, const uint8_t* mask,
                               int64_t n_img, int64_t P, const float* w2, const double* gamma_of_ref, int k,
                               const unsigned long long* cand, unsigned long long* best, void* stream);

/* ---------------------------------------------------------------- autocorrelation solve
 * The W x = d system of cmtip/autocorrelation/autocorrelation.py:113-173 and its scipy
 * CG solve (:239).  A solver holds FFT plans for the M^3 grid and the L^3 Toeplitz
 * embedding (L even, L >= 2M; the reference uses L = 2M, :141), the kernel spectrum,
 * the support and the CG work vectors.  Vectors are float32 M^3; CG scalars are FP64. */
typedef struct cmtip_acsolver_s* cmtip_acsolver;

int cmtip_acsolver_create(int M, int L, cmtip_acsolver* out);
int cmtip_acsolver_destroy(cmtip_acsolver s);

/* support = (gaussian_filter(ac_estimate, 0.5) > 1e-12) as float32 (autocorrelation.py:218-219;
 * scipy.ndimage defaults: mode 'reflect', truncate 4, i.e. a 5-tap kernel), FP64 math.
 * ac_estimate: float64 M^3 on the device. */
int cmtip_ac_support(const double* ac_estimate, int M, float* support, void* stream);
/* As cmtip_ac_support; round_f32 = 1 rounds the filtered volume to float32 after each
 * axis, as scipy does when ac_estimate is a float32 array. */
int cmtip_ac_support_ex(const double* ac_estimate, int M, int round_f32, float* support, void* stream);

/* Keep the support used by the operators (float32 M^3, device, caller-owned). */
int cmtip_acsolver_set_support(cmtip_acsolver s, const float* support);

/* F_ugrid_conv_ = fftn(ifftshift(ugrid_conv)) (autocorrelation.py:148) from the real
 * type-1 of ones on the L^3 grid (float32 L^3).  The spectrum of a real kernel that is
 * even modulo L is real; its real part is kept. */
int cmtip_acsolver_set_kernel(cmtip_acsolver s, const float* ugrid_conv, void* stream);
/* (set_kernel's scratch grids belong to the solver: the call neither allocates after
 * its first use nor waits for the device.) */
/* Install a spectrum given directly (real part, float32 L^3 in FFT position order, as
 * the F_ugrid_conv_ argument of core_problem_convolution, autocorrelation.py:32). */
int cmtip_acsolver_set_spectrum(cmtip_acsolver s, const float* F_real, void* stream);

/* gen_F_antisupport (autocorrelation.py:66-85) as float32 0/1 in array order (M^3). */
int cmtip_gen_F_antisupport(int M, float* out, void* stream);
/* Install the antisupport used by fourier_reg (float32 M^3, array order; the solver
 * keeps it ifftshifted, autocorrelation.py:105).  acsolver_create installs
 * gen_F_antisupport(M) itself. */
int cmtip_acsolver_set_antisupport(cmtip_acsolver s, const float* F_antisupport, void* stream);

/* core_problem_convolution (autocorrelation.py:32-63): out = S (.) crop(ifft(F . fft(pad(S (.) u)))). */
int cmtip_core_problem_convolution(cmtip_acsolver s, const float* u, float* out, void* stream);

/* fourier_reg (autocorrelation.py:88-110) with the antisupport of gen_F_antisupport (:66-85). */
int cmtip_fourier_reg(cmtip_acsolver s, const float* u, float* out, void* stream);

/* W_matvec (autocorrelation.py:150-155): out = alpha*conv(u) + rho*u + phi*reg(u). */
int cmtip_acsolver_matvec(cmtip_acsolver s, const float* u, float* out, double alpha, double rho,
                          double phi, void* stream);

/* scipy.sparse.linalg.cg (scipy/sparse/linalg/_isolve/iterative.py:305-422) on W with
 * rtol, atol = 0, the given maxiter, x0 = x (overwritten by the solution).  Returns
 * the number of completed iterations (the reference's callback count,
 * autocorrelation.py:228-246) in *iters and scipy's info in *info.  Iterations run in
 * blocks of 5 (one captured hipGraph replayed on the solver's stream; the iterations
 * after convergence inside a block are switched off on the device); the host
 * synchronises once per block to read the FP64 convergence state (rtol > 0), or once
 * after all ceil(maxiter / 5) blocks (rtol = 0).  x and the support run in solver-owned
 * copies, so the captured block is kept for later calls with the same rho and phi. */
int cmtip_acsolver_cg(cmtip_acsolver s, const float* d, float* x, double alpha, double rho,
                      double phi, int maxiter, double rtol, int* iters, int* info, void* stream);

/* ---------------------------------------------------------------- phasing
 * Phase retrieval of the oversampled autocorrelation (cmtip/phasing/phase.py): error
 * reduction, hybrid input-output, shrink-wrap and recentring on the device, every step
 * in FP32 with complex64 FFTs (hand-written axis passes) and the Gaussian filters in
 * FP64.  Volumes are float32 M^3; those whose names end in an underscore are in FFT
 * position order (DC at index 0, the reference's convention, phase.py:4-9), the others
 * centred.  Supports are float32 0 / 1.  A phaser owns the amplitudes and an M^3
 * complex scratch volume; it is not thread-safe. */
typedef struct cmtip_phaser_s* cmtip_phaser;

int cmtip_phaser_create(int M, cmtip_phaser* out);
int cmtip_phaser_destroy(cmtip_phaser p);

/* From the centred autocorrelation ac: ac_filt = gaussian_filter(max(ac, 0),
 * mode='constant', sigma=1, truncate=2) (FP64), the phaser's amplitudes_ =
 * sqrt(|fftn(ifftshift(ac_filt))|) with the DC entry masked (kept by the projection)
 * (phase.py:186-195).  ac_filt_ (may be NULL) receives ifftshift(ac_filt); support_
 * (may be NULL) receives create_support_(ac_filt_, M, Mquat) (phase.py:47-64). */
int cmtip_phase_setup(cmtip_phaser p, const float* ac, float* ac_filt_, float* support_, int Mquat, void* stream);
/* Copy the amplitudes out / in (a negative entry marks a value the projection keeps,
 * the DC entry after setup). */
int cmtip_phase_get_amplitudes(cmtip_phaser p, float* amplitudes_, void* stream);
int cmtip_phase_set_amplitudes(cmtip_phaser p, const float* amplitudes_, void* stream);
/* n steps of error reduction (mode 1, ER phase.py:83-99) or hybrid input-output (mode 2,
 * HIO phase.py:102-119, feedback beta) on rho_ in place, each with the amplitude
 * projection of step_phase (phase.py:122-144); rho_max = +inf disables the threshold as
 * the reference's driver does.  Mode 0 replaces rho_ by the projected density rho_mod_
 * of step_phase itself (support_ unused, may be NULL). */
int cmtip_phase_iterate(cmtip_phaser p, float* rho_, const float* support_, int mode, int n, double beta,
                        double rho_max, void* stream);
/* shrink_wrap (phase.py:147-162): support_ = gaussian_filter(|rho_|, mode='wrap', sigma,
 * truncate=2) > max|rho_| * cutoff. */
int cmtip_phase_shrink_wrap(cmtip_phaser p, const float* rho_, double cutoff, double sigma, float* support_,
                            void* stream);
/* recenter (phase.py:11-44): rolls rho_ and support_ in place so the centre of mass of
 * |rho_| moves to the origin; shift (3 ints, may be NULL) receives the vector rolled by.
 * Synchronises the stream once (the shift is computed on the host). */
int cmtip_phase_recenter(cmtip_phaser p, float* rho_, float* support_, int* shift, void* stream);
/* ac_phased = fftshift(|ifftn(|fftn(rho_)|^2)|) (phase.py:214-217), centred. */
int cmtip_phase_ac(cmtip_phaser p, const float* rho_, float* ac_phased, void* stream);

/* ---------------------------------------------------------------- data layer
 * Which pixels enter the path and at what binning (cmtip/prep_data.py:96-262).
 * Copy / integer work bound by HBM. */

/* bin_data (prep_data.py:197-226): out[s, i, j] = mean of in[s] over the b x b block at
 * (b i, b j), for n_stack stacked (nx, ny) panels (float32; nx, ny divisible by b,
 * 1 <= b <= 128).  numpy's float32 order of operations: per block row the b values are
 * summed in numpy's pairwise order and divided by b, the b row means are added in
 * order and divided by b (bit-identical to reshape(...).mean(-1).mean(3)). */
int cmtip_bin_images(const float* in, int64_t n_stack, int nx, int ny, int bin_factor, float* out, void* stream);
/* The pixel selection of clip_data (prep_data.py:166-167): flags[p] (uint8, P) = 1 where
 * (1 / sqrt((x^2 + y^2) + z^2)) * 1e10 > res_limit, all in float32 (res_limit rounded to
 * float32, as numpy compares a float32 array with a Python float), for pixel p of
 * s = (x[P], y[P], z[P]) with components stride floats apart; idx receives the
 * selected pixel indices in increasing order and *count (device int64) their number. */
int cmtip_resolution_select(const float* s, int64_t stride, int64_t P, double res_limit, uint8_t* flags,
                            int64_t* idx, int64_t* count, void* stream);
/* out[r, k] = in[r, idx[k]] (0 where idx[k] < 0) for rows of P_in elements in and P_out
 * out, elements of 4 or 8 bytes: clip_data's arr[:, :, indices] (prep_data.py:168) and
 * trim_dataset's masked gathers (:142-148). */
int cmtip_gather_columns(const void* in, int elem_bytes, int64_t rows, int64_t P_in, const int64_t* idx,
                         int64_t P_out, void* out, void* stream);
/* assemble_image_stack_batch (prep_data.py:245-262) as an index: src[x * Y + y] = the
 * flat pixel index (panel-major) that the index map (n_pixels x 2 int64) sends to (x, y),
 * the last one when several do (numpy's fancy assignment), or -1. */
int cmtip_assemble_index(const int64_t* index_map, int64_t n_pixels, int64_t X, int64_t Y, int64_t* src,
                         void* stream);
/* bin_pixel_index_map (prep_data.py:229-242): out[l, i, j, c] = min over a, e in {0, 1} of
 * in[l, b i + a, b j + e, c], floor-divided by b (int64, (n_panels, px, py, 2) in). */
int cmtip_bin_index_map(const int64_t* in, int64_t n_panels, int px, int py, int bin_factor, int64_t* out,
                        void* stream);

/* ---------------------------------------------------------------- measurement
 * No counterpart in the reference.  workgroups x 512 threads each read a 64 KB LDS
 * buffer iters times with conflict-free ds_read_b128 (16 B per lane per read) and
 * write one float per thread to out (workgroups x 512 floats): the LDS read ceiling
 * that the interpolation kernels are measured against, timed by the caller. */
int cmtip_lds_read_probe(int workgroups, int iters, float* out, void* stream);
/* No counterpart in the reference.  The same launch shape with ds_add_u64 (each lane its
 * own 8-byte counter, a wave on consecutive counters): the 64-bit LDS atomic rate the
 * spreading kernel's accumulation is measured against. */
int cmtip_lds_atomic_probe(int workgroups, int iters, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CMTIP_AMD_H */
