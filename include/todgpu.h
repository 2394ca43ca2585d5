/*
 * todgpu.h -- C ABI of the MI355X (gfx950) beam-convolution TOD library
 * (libtodgpu.so).
 *
 * All array arguments are DEVICE pointers (HIP global memory) unless stated
 * otherwise; `stream` is a hipStream_t (NULL = the default stream).  Calls are
 * stream-ordered and asynchronous: the caller keeps every buffer alive until
 * the stream has passed the call.  The library keeps no reference to caller
 * buffers after a call returns.  Every entry point returns TODG_OK (0) or a
 * negative TODG_ERR_* code; todg_last_error() gives the calling thread's last
 * message.  No C++ exception crosses this boundary.
 *
 * Layouts (row-major, as in the reference):
 *   vec_orig   (S, 3) float32  beam-frame unit vectors
 *   beam_vals  (S,)   float32  beam weights
 *   axes       (B, 3) float32  first Rodrigues axis       } the six arrays of
 *   cos_a/sin_a (B,)  float32  first Rodrigues angle      } _rotation_params,
 *   ax_pts     (B, 3) float32  boresight unit vector      } tod_rotations.py:306-333
 *   cos_p/sin_p (B,)  float32  polarisation-roll angle    }
 *   tod        (C, B) float64  accumulated in place (+=), as the reference
 *   map_packed (npix, CS) float32, pixel-major, produced by todg_pack_map
 *              from the reference's (C, npix) field-major mp_stacked.
 */
#ifndef TODGPU_H
#define TODGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TODG_ABI_VERSION 1

#define TODG_OK 0
#define TODG_ERR_ARG -1 /* invalid argument (nside, C, c_q/c_u, null, range) */
#define TODG_ERR_HIP -2 /* HIP runtime error (allocation, launch)          */

/* ABI version of the loaded library (TODG_ABI_VERSION). */
int todg_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated,
 * truncated to n-1 bytes); returns its length. */
int todg_last_error(char *buf, size_t n);

/* Index-check counters of the checked build (libtodgpu_checked.so, the same
 * sources compiled with -DTODG_CHECKED; no counterpart in the reference,
 * whose Numba kernels run with boundscheck off).  In that build every map
 * gather, weight-grid deposit and ring-slot lookup tests its index, counts a
 * violation by kind and still clamps, so nothing is accessed out of bounds.
 * Synchronises the device, writes the counts of kinds 0 (map pixel outside
 * [0, npix)), 1 (weight-grid cell outside the grid) and 2 (slot outside its
 * ring or window), summed over the library's modules on the current device,
 * into counts[0..n) (host memory; kinds beyond the third read 0), clears them
 * when reset != 0 and returns the number of kinds (3).  The release library
 * returns TODG_ERR_ARG. */
int todg_check_counts(unsigned long long *counts, int n, int reset);

/* Packed-map layout for a beam entry with C channels and Q/U positions c_q,
 * c_u (-1 = absent).  Writes the source channel of each packed slot (perm4,
 * -1 = zero) and the tod row each slot accumulates into (row4); returns the
 * slot stride CS (1, 2 or 4) or a negative error.  Replaces the channel
 * bookkeeping of tod_core.py:105-112 and tod_bilinear.py:275-285. */
int todg_map_slots(int C, int c_q, int c_u, int *perm4, int *row4);

/* Converts the reference's field-major mp_stacked (C, npix) float32 (the
 * array built at sample_based_tod_generation_gridint.py:447-450) into the
 * pixel-major packed layout read by the gather kernels.  packed holds
 * npix * CS floats (CS from todg_map_slots). */
int todg_pack_map(const float *mp_stacked, int C, int64_t npix, int c_q, int c_u,
                  float *packed, void *stream);

/* Bilinear beam-convolution TOD with spin-2 Q/U correction.
 * Replaces _gather_accum_fused_jit (tod_bilinear.py:205-222, same argument
 * order; mp_stacked becomes map_packed + C).  variant selects the kernel:
 * 0 = footprint kernel (per-sample LDS weight grid, 4 waves/SIMD; the
 * default), 8 = the footprint kernel compiled for 2 waves/SIMD, 1/2/4 =
 * direct kernel with that many samples per wavefront.  Results agree to
 * rounding; neighbour indices are identical.  S <= 2^28 (status -1 beyond). */
int todg_gather_accum_fused(const float *vec_orig, const float *axes, const float *cos_a,
                            const float *sin_a, const float *ax_pts, const float *cos_p,
                            const float *sin_p, int64_t nside, const float *map_packed, int C,
                            const float *beam_vals, int64_t B, int64_t S, double *tod,
                            int c_q, int c_u, int variant, void *stream);

/* variant (all three kernels below and todg_gather_accum): 0 = footprint
 * kernel; 8 = the same for 2 waves/SIMD; 1/2/4 = direct kernel (bilinear and
 * nearest); 16 + 3 i + j (bilinear and nearest) = footprint kernel with claims
 * of {128, 256, 512}[i] beam pixels and {3, 4, 5}[j] 64-cell chunks per
 * lane in the footprint sweep (variant 20 = variant 0's tile; built for I/Q/U
 * with Q/U and for one field, other channel layouts run variant 0's tile).
 * Any other value: status -1.  A sample whose rotation operands are not all
 * finite is skipped (its tod entries are left as they are).  The same call
 * repeated gives bit-identical results (weight-grid sums are exact, pixels
 * outside a sample's footprint are summed in a fixed order), except for a
 * sample whose laid-out footprint misses more than 32 of its neighbour
 * pixels and, in bilinear mode, a sample whose footprint exceeds the grid
 * (4608 cells or 128 rings: a beam over ~3.7 deg in radius at nside 2048) or
 * a beam whose sum of |beam_vals| is zero or not finite (it runs without a
 * weight grid): those agree to rounding. */

/* Nearest-pixel beam-convolution TOD.  Replaces _gather_accum_nearest_jit
 * (tod_nearest.py:21-38, same argument order). */
int todg_gather_accum_nearest(const float *vec_orig, const float *axes, const float *cos_a,
                              const float *sin_a, const float *ax_pts, const float *cos_p,
                              const float *sin_p, int64_t nside, const float *map_packed,
                              int C, const float *beam_vals, int64_t B, int64_t S,
                              double *tod, int c_q, int c_u, int variant, void *stream);

/* Beam-convolution TOD with the bicubic judge: the same rotation and spin-2
 * as todg_gather_accum_fused, each beam pixel interpolated with Keys
 * (Catmull-Rom, a = -0.5) weights on the gnomonic tangent plane over the
 * reference's 7-ring x 5-pixel stencil (numba_healpy.py:892-977; the
 * weighting follows README.md:120 -- the reference's own bicubic module,
 * tod_bicubic.py on its bicubic branch, is not in its snapshot).  Same
 * arguments as todg_gather_accum_fused; variant 0 or 8 (footprint kernel). */
int todg_gather_accum_bicubic(const float *vec_orig, const float *axes, const float *cos_a,
                              const float *sin_a, const float *ax_pts, const float *cos_p,
                              const float *sin_p, int64_t nside, const float *map_packed,
                              int C, const float *beam_vals, int64_t B, int64_t S,
                              double *tod, int c_q, int c_u, int variant, void *stream);

/* The three kernels behind one entry (mode 0 = bilinear, 1 = nearest, 2 =
 * bicubic), the C-ABI form of the reference's call surface sketched for this
 * path (tod_core.py:129-164 dispatches to the kernels by interp_mode), with
 * one more output: status (optional device int) gets bit 0 set when a
 * sample's rotation operands were not all finite (that sample is skipped).
 * The reference checks nothing (Numba kernels without bounds checks). */
int todg_gather_accum(int mode, const float *vec_orig, const float *axes, const float *cos_a,
                      const float *sin_a, const float *ax_pts, const float *cos_p,
                      const float *sin_p, int64_t nside, const float *map_packed, int C,
                      const float *beam_vals, int64_t B, int64_t S, double *tod, int c_q,
                      int c_u, int variant, int *status, void *stream);

/* Index dump of the production footprint kernel (variant 0's tile; mode 0 =
 * bilinear, 1 = nearest; C = 3 with Q/U or C = 1): the same launch as
 * todg_gather_accum (tod accumulated likewise) built to also record, for
 * every (b, s), i = b S + s, what the kernel deposited: dump holds
 * todg_fp_dump_bytes(B, S) bytes laid out as pix (4 B S int64: the bilinear
 * neighbours in the order of numba_healpy.py:473-480, or the nearest pixel
 * in pix[4 i] and -1 in the other three), w (4 B S float64: beam value x
 * weight as deposited) and path (B S bytes: 0 = the certified fast path,
 * 1 = the exact path).  B S <= 2^31.  Diagnostic entry (tests); the
 * reference counterparts of the indices are _ring_interp_with_angles_jit
 * (numba_healpy.py:358-499) and tod_nearest.py:108-141. */
size_t todg_fp_dump_bytes(int64_t B, int64_t S);
int todg_fp_dump(int mode, const float *vec_orig, const float *axes, const float *cos_a,
                 const float *sin_a, const float *ax_pts, const float *cos_p, const float *sin_p,
                 int64_t nside, const float *map_packed, int C, const float *beam_vals, int64_t B,
                 int64_t S, double *tod, int c_q, int c_u, int64_t *dump, void *stream);

/* Point values of one map field (map: npix float32, RING order) at N
 * directions (z, phi) (device float64 arrays): mode 0 = bilinear
 * (numba_healpy.py:358-499), 1 = the bicubic judge above.  out: N float64. */
int todg_interp_points(int mode, int64_t nside, const float *map, int64_t N, const double *z,
                       const double *phi, double *out, void *stream);

/* One detector (or one detector's scan chunk) of a batched launch.  All
 * pointers are device pointers with the layouts above; `first` is the
 * global index of this item's first sample, i.e. the sum of B over the
 * items before it (items are ordered by `first`). */
typedef struct todg_work_item {
    const float *vec_orig, *beam_vals;
    const float *axes, *cos_a, *sin_a, *ax_pts, *cos_p, *sin_p;
    double *tod;
    int64_t B, S, first;
} todg_work_item;

/* Many detectors with their own beams (uneven S) and scan chunks (uneven B)
 * in one persistent launch (the reference calls beam_tod_batch once per
 * beam entry per batch: sample_based_tod_generation_gridint.py:284-296).
 * mode 0 = bilinear, 1 = nearest, 2 = bicubic; items: device array of n_items (S <= 2^28
 * each); the map,
 * C and c_q/c_u are shared.  chunk = samples per work-queue claim (1..65536);
 * at most 2^33 samples per launch.
 * queue: device scratch of >= 8 unsigned ints (zeroed by the call itself,
 * so a captured graph replays correctly; concurrent calls need separate
 * queues).  item_status: optional device int[n_items], bit 0 set for an
 * item with a non-finite pointing operand (that sample is skipped).
 * variant: 0 or 8 as for todg_gather_accum_fused. */
int todg_gather_accum_batch(int mode, const todg_work_item *items, int n_items,
                            int64_t total_samples, int64_t nside, const float *map_packed,
                            int C, int c_q, int c_u, int chunk, unsigned int *queue,
                            int *item_status, int variant, void *stream);

/* Exact-path index dumps for samples [b0, b0+nb): the four bilinear
 * neighbours and weights of every (b, s) as the kernels' exact path (the
 * direct kernel, and the footprint kernel for the beam pixels its certified
 * fast path does not take) evaluates them (pix/wgt: (nb, S, 4) int64/
 * float64), i.e. _rodrigues_apply_one_jit + _ring_interp_with_angles_jit
 * (tod_rotations.py:83-134, numba_healpy.py:358-499).  What the production
 * footprint kernel itself deposits, fast path included: todg_fp_dump. */
int todg_bilinear_indices(const float *vec_orig, const float *axes, const float *cos_a,
                          const float *sin_a, const float *ax_pts, const float *cos_p,
                          const float *sin_p, int64_t nside, int64_t B, int64_t S,
                          int64_t b0, int64_t nb, int64_t *pix, double *wgt, void *stream);

/* The nearest pixel of every (b, s) as the nearest kernels' exact path
 * chooses it (pix: (nb, S) int64), i.e. tod_nearest.py:91-141; the
 * production footprint kernel's own choices: todg_fp_dump. */
int todg_nearest_indices(const float *vec_orig, const float *axes, const float *cos_a,
                         const float *sin_a, const float *ax_pts, const float *cos_p,
                         const float *sin_p, int64_t nside, int64_t B, int64_t S,
                         int64_t b0, int64_t nb, int64_t *pix, void *stream);

/* HOST pointers: the per-ring constants the kernels use, rows 0..4*nside
 * (ring index ir; rows 0 and 4*nside are zero padding).  Replaces
 * _ring_info_jit / _ring_z_jit (numba_healpy.py:73-114); any output may be
 * NULL. */
int todg_ring_table(int64_t nside, double *z, double *theta, double *phi0, double *dphi,
                    int64_t *first, int64_t *n);


/* Beam transfer function B_ell = sum_i w_i P_ell(x_i), ell = 0..lmax, by the
 * Legendre three-term recurrence per pixel (replaces the loop of compute_bell,
 * tod_utils.py:147-251; selection, normalisation and the cos-distance x_i
 * stay on the host).  x, w: N float64; ws: todg_beam_bell_workspace(N, lmax)
 * float64 of device scratch; bell: lmax + 1 float64 (device). */
size_t todg_beam_bell_workspace(int64_t N, int lmax);
int todg_beam_bell(const double *x, const double *w, int64_t N, int lmax, double *ws,
                   double *bell, void *stream);

/* todg_beam_bell without a workspace argument, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight loops such as the clustering
 * calibration sweep).  The per-block partial sums live in a library-owned
 * device array of 2^20 doubles (8 MiB on each device it runs on, held until
 * the process ends); calls with ceil(N / 256) * (lmax + 1) beyond that return
 * TODG_ERR_ARG.  Every call on one device shares that array: calls that
 * overlap on one device (different streams, or host threads) race.  Results
 * are bit-identical to todg_beam_bell's. */
int todg_beam_bell_noalloc(const double *x, const double *w, int64_t N, int lmax, double *bell,
                           void *stream);

/* Ring sums of a spin-0 spherical-harmonic synthesis,
 *   F_m(theta_r) = sum_{l=m}^{lmax} a_lm lambda_lm(theta_r),
 * for K <= 4 fields at once (lambda_lm orthonormal associated Legendre
 * functions, three-term recurrence in l with a rescaled mantissa so polar
 * rings at high m neither underflow nor overflow: the mantissa is divided by
 * `big` when it exceeds it and the exponent grows by logbig = log(big)).  The
 * step of healpy's alm2map / synfast before the per-ring FFT; the reference
 * reads its input skies from files and takes such synthetic ones in its
 * configurations and tests, so this has no counterpart in its own code.
 * Device pointers: alm K x (lmax+1) x (lmax+1) complex binary64 as (re, im)
 * pairs, laid out [k][m][l] (entries with l < m unused); x = cos theta_r and
 * st = sin theta_r, R each; logmm: lmax + 1 values log sqrt((2m+1)/(4 pi)
 * prod_{j<=m} (2j-1)/(2j)); F: K x (lmax+1) x R complex binary64, [k][m][r]. */
int todg_ring_sums(const double *alm, int K, int lmax, const double *x, const double *st, int64_t R,
                   const double *logmm, double big, double logbig, double *F, void *stream);

/* ------------------------------------------------------------------------
 * Learned interpolation corrector (no counterpart in the reference: the
 * build's corrector of bilinear interpolation toward the bicubic judge).
 * An MLP 21 -> 64 (GELU) -> 64 (GELU) -> 16 with FP32 master parameters in
 * one flat array of TODG_CORR_NPARAMS floats: W1 (64 x 21), b1, W2 (64 x 64),
 * b2, W3 (16 x 64), b3 (row-major, output x input).  Inputs: the query's
 * sub-pixel coordinates in its 4 x 4 ring neighbourhood (fa, v, fb, fm, fp,
 * each minus 1/2: slot fractions in rings ira, ira+1, ira-1, ira+2 and the
 * bilinear ring weight v) and the 16 neighbourhood values normalised by
 * their mean and RMS; outputs: 16 weight corrections (their mean is
 * removed).  Matrix products run on BF16 MFMA with FP32 accumulation.  All
 * pointers are device pointers. */
#define TODG_CORR_NPARAMS 6608

/* Features of N query directions (z, phi: float64) on one RING-ordered map
 * field (npix float32): X (N x 21: fa, v, fb, fm, fp minus 1/2, the 4 x 4
 * neighbourhood normalised by its mean and RMS), wbil (N x 4: bilinear
 * weights of neighbourhood cells 5, 6, 9, 10), mu, sigma (N), pix (N x 16
 * pixel indices) and ok (N: 0 where the neighbourhood leaves the normal
 * ring range and the corrector does not apply). */
int todg_corr_features(int64_t nside, const float *field, int64_t N, const double *z,
                       const double *phi, float *X, float *wbil, float *mu, float *sigma,
                       int *pix, int *ok, void *stream);

/* Float count of the workspace for N rows of todg_corr_forward / _loss_grad. */
size_t todg_corr_workspace_floats(int64_t N);

/* O (N x 16) = the MLP's raw outputs for feature rows X. */
int todg_corr_forward(const float *params, const float *X, int64_t N, float *ws, float *O,
                      void *stream);

/* Loss sum_i (sum_k w_ik xhat_ik - that_i)^2 over the N rows (into *loss,
 * float64, may be NULL), its gradient with respect to the parameters
 * divided by N (into grad, TODG_CORR_NPARAMS floats, may be NULL) and the
 * normalised predictions (into pred, N floats, may be NULL). */
int todg_corr_loss_grad(const float *params, const float *X, const float *wbil,
                        const float *that, int64_t N, float *ws, float *grad, double *loss,
                        float *pred, void *stream);

/* One Adam step (bias-corrected, step >= 1) on the FP32 master parameters. */
int todg_corr_adam(float *params, const float *grad, float *m, float *v, int step, float lr,
                   float beta1, float beta2, float eps, void *stream);

/* Beam-convolution TOD with the corrector fused into the gather: arguments
 * of todg_gather_accum_fused plus the corrector's parameters.  Runs on the
 * footprint kernel: per beam pixel the corrector reads the 4 x 4
 * neighbourhood of the first field outside the Q/U pair (I; Q when the map
 * holds only Q and U), runs the MLP (BF16 MFMA, FP32 accumulation) and
 * scatters the corrected weights, which every field then uses (Q/U rotated
 * per pixel by the reference's spin-2 angle, tod_bilinear.py:403-432).
 * Where the corrector does not apply (the rings next to the poles) or its
 * correction exceeds `bound` (in units of the neighbourhood RMS; INFINITY:
 * no bound) the plain bilinear weights are used, and so they are for beam
 * pixels n_corrected .. S-1 (callers that correct only the strongest beam
 * pixels sort the beam by value first; n_corrected >= S: all).  stats
 * (optional, 2 counters, accumulated): corrected and plain beam-pixel
 * samples. */
int todg_gather_accum_learned(const float *vec_orig, const float *axes, const float *cos_a,
                              const float *sin_a, const float *ax_pts, const float *cos_p,
                              const float *sin_p, int64_t nside, const float *map_packed, int C,
                              const float *beam_vals, int64_t B, int64_t S, double *tod,
                              int c_q, int c_u, const float *params, float bound,
                              int64_t n_corrected, unsigned long long *stats, void *stream);

/* The six per-sample kernel operands of a batch of boresight pointings
 * (device float64 theta, phi, psi, B each), computed on the GPU: the
 * driver's float32 casts and psi_b - beta (sample_based_tod_generation_
 * gridint.py:277-281), precompute_rotation_vector_batch (tod_rotations.py:
 * 219-303) and _rotation_params (tod_rotations.py:306-333), each step in the
 * precision the reference's numpy code takes it.  centre: HOST pointer to 6
 * doubles, the beam centre's unit vector c and its local north n_c.  Outputs
 * (device float32): axes (B, 3), cos_a, sin_a (B), ax_pts (B, 3), cos_p,
 * sin_p (B).  Binary32 sines and cosines are the binary64 ones rounded once;
 * numpy's float32 routines differ from these by at most one ulp. */
int todg_rotation_params(const double *theta, const double *phi, const double *psi, int64_t B,
                         const double *centre, float *axes, float *cos_a, float *sin_a,
                         float *ax_pts, float *cos_p, float *sin_p, void *stream);

/* Factorised corrector, map side: the two stencil channels of one map field
 * (field: npix float32, RING order).  Per pixel p on ring r, the 16 features
 * are F_0 = m(p) and, for (dr, ds) in {0..3}^2 minus (0, 0) in row-major
 * order, the mean of the values at rings r -+ dr, longitudes phi_p -+ ds
 * dphi_r (each linearly interpolated between the two ring pixels around it)
 * minus m(p); features with dr > 0 are 0 on the 3 rings next to each pole.
 * out0[p * stride] = sum_j coef[j] F_j and out1[p * stride] = sum_j coef[16 +
 * j] F_j (coef: HOST pointer to 32 doubles; out0 / out1 device float32, e.g.
 * one slot of a packed map: todg_map_slots).  No reference counterpart: the
 * corrector is this library's own (the learned interpolation corrector's
 * fast form). */
int todg_fcorr_channels(int64_t nside, const float *field, const double *coef, float *out0,
                        float *out1, int64_t stride, void *stream);

/* Factorised corrector, TOD side: arguments of todg_gather_accum_fused with
 * two packed maps of the same slot layout -- map_packed holding the main
 * channel A0 and map1_packed the correction channel A1 (both from
 * todg_fcorr_channels) -- and the status word of todg_gather_accum.  Each
 * beam pixel's bilinear neighbours p_k and weights w_k (the reference's,
 * numba_healpy.py:358-499) give bv sum_k w_k (A0(p_k) + (1 - g) g A1(p_k)), g
 * = w_2 + w_3 the lower ring's weight; Q/U are rotated per pixel by the
 * reference's spin-2 angle (tod_bilinear.py:403-432) after both channels are
 * combined.  With A0 = map and A1 = 0 this is the plain bilinear TOD. */
int todg_gather_accum_fcorr(const float *vec_orig, const float *axes, const float *cos_a,
                            const float *sin_a, const float *ax_pts, const float *cos_p,
                            const float *sin_p, int64_t nside, const float *map_packed,
                            const float *map1_packed, int C, const float *beam_vals, int64_t B,
                            int64_t S, double *tod, int c_q, int c_u, int *status, void *stream);

/* C (M x N) = alpha op(A) op(B) + beta C on BF16 MFMA (row-major float32
 * operands rounded to bfloat16; ta / tb transpose A / B). */
int todg_corr_gemm_bf16(int M, int N, int K, const float *A, int lda, int ta, const float *B,
                        int ldb, int tb, float *C, int ldc, float alpha, float beta,
                        void *stream);

/* MLP stencil corrector (no reference counterpart: the learned interpolation
 * corrector with its network evaluated once per map pixel).  Per pixel p of
 * a map field, with the features F_0 .. F_15 of todg_fcorr_channels and
 * sigma = sqrt(sum_{j>=1} F_j^2 / 15):
 *   x  = [2 |z_p| - 1, F_1 / sigma, .., F_15 / sigma]   (0 where sigma = 0)
 *   o  = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3          (16 -> 64 -> 64 -> 32)
 *   A0 = F_0 + sum_{j>=1} o_j F_j,  A1 = sum_{j>=1} o_{16+j} F_j
 * (gelu as the MLP corrector's).  A0 and A1 are the two channels of
 * todg_gather_accum_fcorr.  Parameters: TODG_SMLP_NPARAMS float32 in one
 * flat device array W1 (64 x 16), b1, W2 (64 x 64), b2, W3 (32 x 64), b3.
 * Matrix products on BF16 MFMA with FP32 accumulation. */
#define TODG_SMLP_NPARAMS 7328

/* Channels A0, A1 of one field (npix float32) in one fused kernel (features,
 * MLP, channels) into out0[p * stride], out1[p * stride]. */
int todg_smlp_channels(int64_t nside, const float *field, const float *params, float *out0,
                       float *out1, int64_t stride, void *stream);

/* Training side: features F (npix x 16) and MLP inputs X (npix x 16). */
int todg_smlp_features(int64_t nside, const float *field, float *F, float *X, void *stream);

/* Float count of the workspace of N rows for todg_smlp_forward / _backward. */
size_t todg_smlp_workspace_floats(int64_t N);

/* MLP outputs O (N x 32) of input rows X (N x 16); keeps the activations in
 * the workspace for todg_smlp_backward. */
int todg_smlp_forward(const float *params, const float *X, int64_t N, float *ws, float *O,
                      void *stream);

/* Gradient (TODG_SMLP_NPARAMS floats) of sum_i dO_i . o_i over the rows of
 * the preceding todg_smlp_forward (whose workspace it consumes). */
int todg_smlp_backward(const float *params, const float *X, int64_t N, float *ws, const float *dO,
                       float *grad, void *stream);

/* A0, A1 (N) from MLP outputs O (N x 32) and features F (N x 16). */
int todg_smlp_head(int64_t N, const float *F, const float *O, float *A0, float *A1, void *stream);

/* Backward of todg_smlp_head: dO (N x 32) from the gradients G0, G1 (N) of
 * the loss with respect to A0, A1. */
int todg_smlp_head_backward(int64_t N, const float *F, const float *G0, const float *G1, float *dO,
                            void *stream);

/* Two-channel sparse operator, one row per TOD sample (CSR: rowptr int64,
 * col int32, values v0 / v1 float32): y_r = sum_e v0_e x0[col_e] + v1_e
 * x1[col_e] (binary64, in row order), e_r = y_r - target_r (target may be
 * NULL), resid_r = scale e_r (float32, may be NULL), y (may be NULL), and
 * *sumsq = sum_r e_r^2 (may be NULL; needs part, todg_csr_rows2_parts(nrows)
 * doubles). */
size_t todg_csr_rows2_parts(int64_t nrows);
int todg_csr_rows2(int64_t nrows, const int64_t *rowptr, const int *col, const float *v0,
                   const float *v1, const float *x0, const float *x1, const double *target,
                   double scale, double *y, float *resid, double *part, double *sumsq,
                   void *stream);

/* The transpose of such an operator stored by columns (one CSR row per map
 * pixel, col = sample): out0_p = sum_e v0_e r[col_e], out1_p = sum_e v1_e
 * r[col_e] (binary64 in row order, stored float32). */
int todg_csr_adjoint2(int64_t nrows, const int64_t *rowptr, const int *col, const float *v0,
                      const float *v1, const float *r, float *out0, float *out1, void *stream);

/* One Adam step (bias-corrected, step >= 1) on the stencil MLP's parameters:
 * the hidden layers (W1, b1, W2, b2) at learning rate lr, the output layer
 * (W3, b3: the stencil weights themselves) at lr_out. */
int todg_smlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr,
                   float lr_out, float beta1, float beta2, float eps, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TODGPU_H */
