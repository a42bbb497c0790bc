/* zcsmi.h — C ABI of libzcsmi.so, zcmir's SMI hot path on AMD MI355X (gfx950).
 *
 * Drop-in boundary for the reference's native library (src/exports.zig, src/native_api.zig) as
 * bound by python/zcmir/_native.py:79-133. Plain pointers and sizes; no C++ or torch types.
 * Every call returns 0 (or a size / count) on success and -1 on failure; zc_error() returns the
 * reason (exports.zig:1-16). An engine is one HIP device and stream; calls on one engine are
 * serialized by the caller (python/zcmir/api.py:29). Pose H: row-major 3x3 f64, moving → fixed
 * pixels, 1-based (python/zcmir/__init__.py:18-20). Outputs are caller-allocated with explicit
 * lengths; a too-small buffer fails with MapTooSmall / BufferTooSmall (engine.zig:802,
 * exports.zig:269).
 */
#ifndef ZCSMI_H
#define ZCSMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* engine.zig:32-53 ShiftResult (88 bytes) */
typedef struct {
  uint32_t n, cw, ch, canvas_w, canvas_h;
  int32_t ox, oy;
  uint32_t peak_index;
  double dx, dy, peak, zero, corr_area_px;
  uint32_t exact, calibrated, n_y, _pad;
} zc_shift_result;

/* engine.zig:55-73 RsResult (96 bytes) */
typedef struct {
  uint32_t n, n_th, n_lam, peak_index;
  double dlam, r0, r1, cx, cy, peak, zero, dth, dsg;
  uint32_t symmetric, calibrated;
} zc_rs_result;

/* engine.zig:76-85 GradResult (616 bytes) */
typedef struct {
  double score, fwd, inv, n;
  double grad[8];
  double hess[64];
  uint32_t has_hess, nk;
} zc_grad_result;

/* search.zig:30-41 Result */
typedef struct {
  double H[9];
  double score, prev;
  uint32_t improved, count_a, count_b, _pad;
} zc_search_result;

/* climb.zig:7-12 Result */
typedef struct {
  double H[9];
  double score;
  uint32_t iterations, moved;
} zc_climb_result;

/* ── lifetime (native_api.zig:15-52) ── */
/* native_api.zig:20 zc_version */
const char* zc_version(void);
/* native_api.zig:26 zc_create: `dev` is a HIP device index as text ("0"); any other text
 * (the reference's wgpu-native path) selects device 0 */
void* zc_create(const char* dev);
/* native_api.zig:41 */
size_t zc_create_error(char* buf, size_t cap);
/* native_api.zig:47 */
void zc_destroy(void* e);
/* exports.zig:22: the device name and gfx arch */
size_t zc_adapter(void* e, char* buf, size_t cap);
/* exports.zig:30 */
size_t zc_error(void* e, char* buf, size_t cap);

/* ── settings (exports.zig:35, 42; settings.zig:14-199) ── */
int32_t zc_configure(void* e, const char* json, size_t len);
size_t zc_settings(void* e, char* buf, size_t cap);

/* ── images and features ── */
/* exports.zig:50: a work image, no preprocessing */
int32_t zc_set_image(void* e, uint32_t side, const float* data, uint32_t w, uint32_t h);
/* exports.zig:57; preprocess=1: the settings' CLAHE, band pass and invert (gls.zig:462-559), then
 * the f16 work image when half is on; preprocess=0: the image is the work image unchanged */
int32_t zc_set_image_gray(void* e, uint32_t side, const float* data, uint32_t w, uint32_t h, uint32_t preprocess);
/* exports.zig:63: RGBA8 → gray, always preprocessed as zc_set_image_gray with preprocess=1 */
int32_t zc_set_image_rgba(void* e, uint32_t side, const uint8_t* rgba, uint32_t w, uint32_t h);
/* exports.zig:238 */
void zc_image_size(void* e, uint32_t side, uint32_t out[2]);
/* exports.zig:320 */
int32_t zc_work_image(void* e, uint32_t side, float* out, size_t len);
/* exports.zig:264: 4 x h x w plane-major whitened features */
int32_t zc_features(void* e, uint32_t side, float* out, size_t len);
/* exports.zig:226 */
void zc_swap(void* e);

/* ── scores and maps ── */
/* exports.zig:276: the 45 overlap sums of moving onto fixed at H */
int32_t zc_moments(void* e, const double H[9], double out[45]);
/* No counterpart in the reference: zc_moments for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops). feat_a / feat_b: device pointers to 4 planes of
 * wa x ha / wb x hb float32 features (zc_features' layout); H on the host; out: a device pointer
 * to 45 doubles, written when the work queued on `stream` (NULL: the default stream) has run. The
 * partial sums live in a fixed 360 KiB device array of the library (1024 rows of 45 doubles, one
 * per device, kept until the process ends): no allocation happens, and calls that overlap on one
 * device race on that array. Returns 0, -1 for an invalid argument (null pointer, empty image,
 * more than 2^30 pixels), -2 for a device or launch failure. */
int32_t zc_moments_noalloc(const float* feat_a, uint32_t wa, uint32_t ha, const float* feat_b, uint32_t wb,
                           uint32_t hb, const double H[9], double* out, void* stream);
/* exports.zig:417: [mean, forward, inverse] (inverse NaN when not symmetric) */
int32_t zc_score(void* e, const double H[9], double out[3]);
/* exports.zig:424: score, overlap and tangent gradient at H (hess: also the Gauss-Newton
 * matrix, nk x nk row-major), symmetric and exact per the settings; fwd, inv left 0 */
int32_t zc_gradient(void* e, const double H[9], uint32_t hess, zc_grad_result* out);
/* exports.zig:436: flags 1 exact | 2 calibrated, cap the FFT length cap (0: 1024), no
 * overlap floor; map (nullable) receives n_y x n floats */
int32_t zc_shift_map(void* e, const double H[9], uint32_t flags, uint32_t cap, zc_shift_result* out, float* map,
                     size_t map_len);
/* exports.zig:442: the settings' score, resolution (map_res) and overlap floor (min_overlap) */
int32_t zc_shift_map_set(void* e, const double H[9], zc_shift_result* out, float* map, size_t map_len);
/* exports.zig:449: the roto-scale map at H about c (null: the moving centroid), symmetric: averaged
 * with the inverse-frame map (engine.zig:836-879); map / fwd / inv nullable, each len >= n*n floats.
 * Follows the settings exact, calibrated, map_res (grid <= 2048 angles), rs_smin, rs_smax. */
int32_t zc_rs_map(void* e, const double H[9], const double c[2], uint32_t symmetric, zc_rs_result* out, float* map,
                  float* fwd, float* inv, size_t len);
/* exports.zig:79: one 2-D FFT of `planes` n x n complex planes in place (interleaved f32;
 * inverse scaled by 1/n^2); n a power of two 32..4096; four_step ignored */
int32_t zc_fft2(void* e, float* data, uint32_t n, uint32_t planes, uint32_t inverse, uint32_t four_step);

/* ── pose, search, climb ── */
/* exports.zig:388 */
void zc_set_pose(void* e, const double H[9]);
/* exports.zig:392 */
void zc_get_pose(void* e, double H[9]);
/* exports.zig:344: the sweep search (search_mode sweep, sw_* settings): sweep, NMS, a
 * 4-iteration Gauss-Newton screen, 25 iterations on the finalists; the best pose becomes the
 * engine's pose when it beats the current score */
int32_t zc_search(void* e, zc_search_result* out);
/* exports.zig:405 (gclimb.zig:313-489): the climb from the engine's pose: per iteration a
 * shift-map hop (setting hop), a Gauss-Newton block, and with hop_fm a roto-scale hop and another
 * block; at most g_iters iterations, two without a move end it; no spline stage in this library.
 * The result becomes the engine's pose */
int32_t zc_climb(void* e, zc_climb_result* out);
/* exports.zig:221 */
void zc_cancel(void* e);

/* ── events, profiling, counters ── */
/* exports.zig:412 (events.zig:13-86): queued progress events as JSON lines, kinds "log",
 * "trail", "pose" (zc_search, zc_climb) and "sweep" (zc_search's landscape); whole lines up to
 * cap bytes, removed from the queue; 0 when none are left */
size_t zc_events(void* e, char* buf, size_t cap);
/* exports.zig:210: per-kernel timing on (1) / off (0); resets the report */
int32_t zc_profile(void* e, uint32_t on);
/* exports.zig:215: lines "kernel launches total_ms mean_ms" */
size_t zc_profile_report(void* e, char* buf, size_t cap);
/* exports.zig:110: dispatches, submits, reads, waits since the last call */
void zc_gpu_counts(void* e, uint32_t out[4]);
/* exports.zig:497 */
void zc_result_sizes(uint32_t out[4]);
/* exports.zig:95: tilings of the batched FFTs for rows / columns (0: default), for FFTs run from
 * now on; process-wide. rows: the inverse row pass over lines of >= 256 points (0: 8 points per
 * lane at 256, two lines per workgroup at 1024, 32 points per lane from 2048; 1: 32 points per
 * lane; 2: 8; 3: 16; 4: 16 with twice the lines per workgroup).
 * cols: the column-FFT tiling code */
void zc_fft_tune(uint32_t rows, uint32_t cols);
/* this library's: the column-FFT tiling used for column length L (zc_fft_tune's value, else the
 * autotuned one when the setting fft_autotune is on, timing the tilings now if not yet done,
 * else 0 = built-in default); no counterpart in the reference */
int32_t zc_fft_tiling(void* e, uint32_t L, uint32_t* cols);

/* ── this library's own hooks (no counterpart in the reference) ── */
/* the rank transform of a side: rank (w*h f32, nullable), hist[257], cdf[258], stats[2] =
 * (mean, spread) — for bit-exact checks of the rank transform */
int32_t zc_rank(void* e, uint32_t side, float* rank, size_t len, uint32_t* hist, uint32_t* cdf, float* stats);
/* one untimed, unprofiled map (left out with flags bit 4: the caller has already made one with
 * the same settings), then `iters` shift maps back to back (flags bits 1 and 2, cap as
 * zc_shift_map); mean ms per map, HIP events */
int32_t zc_bench_shift_map(void* e, const double H[9], uint32_t flags, uint32_t cap, uint32_t iters,
                           double* ms_per_map);

/* the rotation x scale sweep alone (the first stage of zc_search, search.zig:300-494 Sweep.sweep,
 * with the candidate grid and options of the current settings; sw_full = 1 extends sw_grid to
 * one canvas cell per fixed pixel) */
typedef struct {
  uint32_t nc, n, grid, n_peaks; /* candidates, largest plane length, canvas cells, raw peaks */
  double cell, D, ox, oy;        /* cell size and canvas side (fixed pixels), canvas origin */
} zc_sweep_info;
typedef struct {                 /* search.zig:145 Peak */
  double score;                  /* n in fixed pixels */
  uint32_t cand, it, is, ik, ia; /* candidate (theta-major), theta / sigma / shear / stretch index */
  uint32_t _pad;                 /* 0 */
  double th, s;
  int32_t sx, sy;                /* canvas shift in cells */
  double H[9];                   /* corrected pose, moving -> fixed */
} zc_sweep_peak;
/* H0: prior pose (NULL: the settings' rule). Writes up to cap raw peaks (every candidate's top 4
 * local maxima above the floor) and nbest per-candidate best scores; returns the peak count. */
int32_t zc_sweep(void* e, const double H0[9], zc_sweep_info* info, zc_sweep_peak* peaks, size_t cap, float* best,
                 size_t nbest);
/* zc_search restricted to k (rotation index, scale index) cells of the sweep grid (cells: 2k
 * indices, e.g. the learned prior's proposals); best_peak (nullable): the sweep's best raw peak
 * score over the fixed image's area (the learned search's fallback test) */
int32_t zc_search_cells(void* e, const uint32_t* cells, uint32_t k, zc_search_result* out, double* best_peak);
/* zc_sweep over k (rotation index, scale index) cells of the grid only (cells: 2k indices; NULL:
 * the whole grid); best in the cells' order */
int32_t zc_sweep_cells(void* e, const double H0[9], const uint32_t* cells, uint32_t k, zc_sweep_info* info,
                       zc_sweep_peak* peaks, size_t cap, float* best, size_t nbest);
/* the score map of the last sweep batch's first candidate; dims = (nx, ny) */
int32_t zc_sweep_map(void* e, float* map, size_t len, uint32_t dims[2]);
/* candidates per sweep batch (0: with the fft_autotune setting on, the size timed fastest per
 * plane size, at most 256, else 256; the reference's batches hold 64); process-wide; results do
 * not depend on it */
void zc_sweep_tune(uint32_t batch);
/* this library's: full-resolution maps at integer-translation poses from now on, 1 (default)
 * through the translation plans (the moving weight plane is a rectangle too: the correlations
 * with it come from summed-area tables, 16 of 45 through the FFT), 0 through the rectangle plans,
 * 2 through the translation plans with every lag reading its own four table cells per sum (1
 * shares the cells of four lags where both images have the same size and the plane is twice as
 * large; 1 and 2 give identical maps); process-wide; 0 and 1 agree within FP32 rounding */
void zc_map_path(uint32_t translation_plans);

/* stacks of uneven slice pairs (the role of python/zcmir/api.py:522-580 register_stack, batched):
 * slice i = moving[i] (moving_wh[2i] x moving_wh[2i+1] f32 gray, taken as the work image) onto
 * fixed[i]; features stay resident until the next zc_stack_set */
int32_t zc_stack_set(void* e, uint32_t n, const float* const* moving, const uint32_t* moving_wh,
                     const float* const* fixed, const uint32_t* fixed_wh);
/* H0: n x 9 start poses (NULL: centred); init_map: 1 correct each by its SMI shift map's peak
 * first, 2 by the best peak of a rotation x scale sweep about it (+-12 deg, 0.88-1.12), 0 neither;
 * then `iters` Gauss-Newton iterations on all slices at once. Writes H_out (n x 9), scores (n) and
 * H_init (n x 9, nullable: the poses after the map); returns the iterations run */
int32_t zc_stack_register(void* e, const double* H0, uint32_t iters, uint32_t init_map, double* H_out,
                          double* scores, double* H_init);

/* zc_set_image_gray (preprocess 0) from device memory; the work writing `data` must be complete */
int32_t zc_set_image_device(void* e, uint32_t side, const float* data, uint32_t w, uint32_t h);

/* ── the learned rotation x scale prior (no counterpart in the reference): device pointers, work
 *    on `stream` (a hipStream_t; NULL: the null stream); 0 on success, -1 on a bad shape ── */
/* y = relu(conv3x3(x) + b): x NHWC [n][h][w][cin], w [9*cin][cout] (ky, kx, cin major), stride 1|2,
 * H axis periodic, W axis zero-padded; BF16 MFMA operands, FP32 accumulation */
int32_t zc_cnn_conv_fwd(const float* x, const float* w, const float* b, float* y, int n, int h, int wd, int cin,
                        int cout, int stride, int relu, void* stream);
/* gradients of zc_cnn_conv_fwd from dy (through the relu via y): dx (nullable) written, dw and db
 * accumulated */
int32_t zc_cnn_conv_bwd(const float* x, const float* w, const float* y, const float* dy, float* dx, float* dw,
                        float* db, int n, int h, int wd, int cin, int cout, int stride, int relu, void* stream);
/* out[n][j] = f[n][:] . w[:][j] + b[j] */
int32_t zc_cnn_fc_fwd(const float* f, const float* w, const float* b, float* out, int n, int k, int j, void* stream);
/* df (nullable) written, dw and db accumulated */
int32_t zc_cnn_fc_bwd(const float* f, const float* w, const float* dout, float* df, float* dw, float* db, int n, int k,
                      int j, void* stream);
/* softmax cross-entropy: loss[n], dlogits = (p - onehot) / n */
int32_t zc_cnn_xent(const float* logits, const int* labels, float* loss, float* dlogits, int n, int j, void* stream);
/* one Adam step (bias-corrected, step >= 1) */
int32_t zc_cnn_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
                    int step, void* stream);
/* network input of `count` gray images: S x S downsample, rank, Sobel magnitude x Hann, 2-D FFT,
 * log(1 + |F|) on T angles x R log radii (r0..r1) into channel ch of NHWC [count][T][R][C],
 * normalized per map once ch == C - 1; work: count * S * S * 12 bytes */
int32_t zc_lp_features(const float* imgs, int count, int w, int h, int S, int T, int R, int C, int ch, float r0,
                       float r1, float* out, void* work, void* stream);
/* bilinear warp of `count` images by His (count x 9, destination -> source, 1-based), zero
 * outside, then remap (0 none, 1 1-x^2, 2 sqrt x, 3 sin pi x) */
int32_t zc_warp_gray(const float* src, int count, int w, int h, const double* His, const int* remap, float* dst,
                     void* stream);

/* ── reference entry points outside the SMI hot path (exports.zig:95-138, 232, 245, 284-376,
 *    382, 397, 430, 458-489): exported so the reference's full ctypes symbol list binds; each
 *    fails with zc_error "Unsupported" (or is a no-op tuning hook) ── */
int32_t zc_compile_wgsl(void* e, const char* code, size_t code_len, const char* entry, size_t entry_len, char* msg,
                        size_t cap);
int32_t zc_pg_info(void* e, double out[5]);
int32_t zc_pg_maps(void* e, uint32_t side, float* out, size_t len);
int32_t zc_pg_descriptors(void* e, uint32_t side, float* out, size_t len);
int32_t zc_bench_moments(void* e, uint32_t n_poses, uint32_t iters);
int32_t zc_match_scores(void* e, float* out, size_t cap);
int32_t zc_detect(void* e, void* out);
int32_t zc_match(void* e, void* out);
int32_t zc_auto(void* e, void* out);
int32_t zc_keypoints(void* e, uint32_t side, float* out, size_t cap_points);
int32_t zc_keypoint_frames(void* e, uint32_t side, float* out, size_t cap_points);
int32_t zc_matches(void* e, uint32_t* out, size_t cap);
int32_t zc_set_ffd(void* e, const float* cps, size_t n);
size_t zc_get_ffd(void* e, float* cps, size_t cap);
uint32_t zc_fit_check(void* e, const double H[9]);
int32_t zc_ffd_gradient(void* e, const double H[9], double* score, double* grad, size_t len);
int32_t zc_canvas(void* e, const double H[9], uint32_t inverse, void* out);
int32_t zc_overlay(void* e, const double H[9], uint32_t inverse, void* out, float* moving, float* fixed, size_t len);
int32_t zc_tile_heat(void* e, const double H[9], uint32_t grid, uint32_t min_n, double* out, float* dest, float* src,
                     size_t len);
int64_t zc_export_warp(void* e, const char* stem, size_t stem_len, void* out, size_t cap);
int32_t zc_displacement(void* e, float* out, size_t len);
void zc_tune_parts(uint32_t px_per_thread, uint32_t fill_wg, uint32_t tg_px, uint32_t tg_cap);
void zc_fft_tune_wg(uint32_t rows, uint32_t cols, uint32_t balanced);

#ifdef __cplusplus
}
#endif

#endif /* ZCSMI_H */
