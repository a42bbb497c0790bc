// Host-side launchers of the SMI kernels (k_prep, k_pack, k_fft, k_ifft, k_score, k_pose, k_sweep .hip). All launches go on `st`; device
// pointers only. Sizes are validated by the engine before any launch.
#pragma once
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>

namespace zc {

// Tiling of the inverse row pass for lines of >= 256 points (zc_fft_tune's rows value): 0 8
// points per lane at 256, two lines per workgroup at 1024, 32 points per lane from 2048; 1 32
// points per lane; 2 8; 3 16; 4 16 with twice the lines per workgroup; process-wide.
struct FftRowTiling {
  static inline std::atomic<int> v{0};
};

struct WhitenArgs {
  double X[16];   // lower-triangular L^-1, row-major
  double mu[4];
};

// pose sampling: Hinv (fixed pixel → moving pixel, 1-based homogeneous), row-major
struct PoseArgs {
  double Hi[9];
};

// ── prepare ──
void launch_round_half(float* x, int64_t n, hipStream_t st);
void launch_maxabs(const float* x, int64_t n, uint32_t* out, hipStream_t st);
void launch_fixsum(const float* x, int64_t n, double scale, float mean, int squared, long long* hi,
                   unsigned long long* lo, hipStream_t st);
void launch_hist(const float* x, int64_t n, float mean, float spread, uint32_t* hist, hipStream_t st);
void launch_rank(const float* x, int64_t n, float mean, float spread, const uint32_t* hist, const uint32_t* cdf,
                 float* rank, hipStream_t st);
// features + per-block partial moments (14 doubles per block); returns the block count
int launch_feats(const float* rank, int w, int h, int mode, float* feat, float* mag, double* part,
                 hipStream_t st);
int feats_blocks(int w, int h);
// batch: blockIdx.y selects consecutive groups of nparts x m partials and m outputs
void launch_sum_parts(const double* part, int nparts, int m, double* out, hipStream_t st, int batch = 1);
void launch_whiten(float* feat, float* stack, int64_t n, const WhitenArgs& a, hipStream_t st);
int acf_parts(int w, int h);   // partial-sum rows of launch_acf (1156 doubles each)
void launch_acf(const float* feat, int w, int h, double* part, hipStream_t st);

// ── maps ──
void launch_warp_stack(const float* stack, int sw, int sh, const double* Hi, float* canvas, int cw, int ch,
                       hipStream_t st);
// the same with Hi (9 doubles) in device memory
void launch_warp_stack_dev(const float* stack, int sw, int sh, const double* Hi_dev, float* canvas, int cw, int ch,
                           hipStream_t st);
void launch_paste_stack(const float* stack, int sw, int sh, int ox, int oy, float* canvas, int cw, int ch,
                        hipStream_t st);
// grid: nx × ny FFT, cw × ch cells resampled from a W × H canvas; writes n_cplx planes of
// ch rows × nx into spec (plane stride nx·ny). Returns false for an unsupported length. cells
// (5 · cw · ch floats, optional): where a resampled grid's cell values are gathered first.
bool launch_pack_rowfft(int plan, int side, const float* canvas, int W, int H, int cw, int ch, int nx, int ny,
                        float2* spec, const float2* tw_x, hipStream_t st, float* cells = nullptr);
// The same for one cell per canvas pixel when the canvas holds an image (a stack of sw × sh
// planes) moved by (sx, sy) and nothing else: the image is read in place, no canvas is written.
bool launch_pack_rowfft_shifted(int plan, int side, const float* stack, int sw, int sh, int sx, int sy, int cw,
                                int ch, int nx, int ny, float2* spec, const float2* tw_x, hipStream_t st);
// column FFTs (length ny) of `planes` planes of nx columns, in place; rows >= rows_in read as 0
// lines: columns per workgroup (0: the default for the length)
// skip (nullable): per entry of `per` consecutive planes (x, y, ., .), output rows [x, y) that no
// later pass reads; the plain tilings of columns of up to 1024 points and the plain and wide
// tilings of 2^a·3^b points leave them unstored
bool launch_colfft(float2* data, int planes, int nx, int ny, int rows_in, bool inverse, const float2* tw_y,
                   hipStream_t st, int lines = 0, const int4* skip = nullptr, int per = 1);
// batch: spec_a and work hold `batch` consecutive entries (n_cplx / n_out planes each); spec_b
// is shared
void launch_cmul(int plan, const float2* spec_a, const float2* spec_b, float2* work, int nx, int ny,
                 hipStream_t st, int batch = 1);
// dead (nullable): per batch entry (x, y, z, w), the lag rows [x, y) and lag columns [z, w) at
// which the two windows' footprints do not overlap (every correlation is 0 there): a workgroup
// whose rows all lie there stores n = 0 in the count plane and nothing else (n stays under any
// overlap floor, so the score reads no other plane there); lines of up to 1024 points store
// only n in those columns, and n = 0 in those rows of a workgroup that also holds other rows
// (the column pass leaves such rows unstored at those lengths). Per-store checks cost the
// longer lines' tilings registers and occupancy: they store every column and transform every
// row of a workgroup that holds rows with overlap.
bool launch_rowifft_extract(int plan, const float2* work, float* corr, int nx, int ny, const float2* tw_x,
                            hipStream_t st, int batch = 1, const int4* dead = nullptr);
// Rectangle and translation plans (4..11), lines of 32..4096: the correlations as quads
// (correlation k: component k & 3 of float4 plane k >> 2, (n_corr + 3) / 4 planes of nx · ny).
bool launch_rowifft_quads(int plan, const float2* work, float4* corr4, int nx, int ny, const float2* tw_x,
                          hipStream_t st);
// Maps at integer-translation poses (full resolution): 1 (default) the translation plans, 0 the
// rectangle plans every other full-resolution pose takes, 2 the translation plans with every lag
// reading its own table cells (no quad form); process-wide (zc_map_path)
struct MapPath {
  static inline std::atomic<int> trans{1};
};
struct CombineArgs {
  int exact, calib, nf;
  float min_n;
  double area, ridge;
};
int combine_blocks(int64_t nn);
// Summed-area tables of the rectangle-sum planes of a rectangle plan (4..7): table k of plane
// kPlans[plan].sat_plane[k].
int sat_bands(int H);
// canvas: 5 planes W x H; band: n_sat · sat_bands(H) · W doubles; sat: sat_pairs(n_sat) ·
// (W+1)(H+1) double2 (tables 2p, 2p+1 interleaved)
// side 0: side A's tables of a rectangle plan (canvas A); side 2: an image's own 15 tables (its
// stack as the canvas, W x H = the image)
bool launch_sat_build(int plan, int side, const float* canvas, int W, int H, double* band, double2* sat,
                      hipStream_t st);
inline constexpr int sat_pairs(int n_tables) { return (n_tables + 1) / 2; }
// The tables and the rectangle [bx0, bx1) x [by0, by1) in table cells; nx = 1 << lx.
// sign +1: the rectangle moves with the lag (side A's sums, R = the fixed rectangle: R + t);
// -1: against it (side B's sums of a translation plan, R = the moving rectangle: R − t).
// The moved rectangle is clipped to [cx0, cx1] x [cy0, cy1] (the canvas, or the part of an image's
// own tables that lies on the canvas); a table row holds `stride` cells (its width + 1) and a
// table `cells` cells ((width + 1) x (height + 1)).
struct SatView {
  const double2* sat;   // the table pairs (k_sat_rows' layout)
  int nx, ny, lx, bx0, bx1, by0, by1;
  int sign;
  int cx0, cx1, cy0, cy1, stride;
  uint32_t cells;
};
// The quad form of a translation plan's table view (k_combine_quad); quad_view fills it when the
// view allows it.
struct QuadAxis {
  int b, s, c0, c1, hi, patB;
};
struct QuadView {
  const double2* sat;
  uint32_t stride, cells;
  QuadAxis y, x;
};
bool quad_view(const SatView& v, QuadView* q);
// sv: side A's tables (rectangle plans); svb: side B's (translation plans); rectangle plans read
// corr in launch_rowifft_quads' layout, the others one plane per correlation
void launch_combine(int plan, const float* corr, float* map, int64_t nn, const CombineArgs& a, float* arg_part,
                    hipStream_t st, const SatView* sv = nullptr, const SatView* svb = nullptr);
// out (ry x rx) = the lags of a wrapped map (ny x nx, nx >= rx, ny >= ry) on a grid of rx x ry:
// column j holds lag j for 2j <= rx, else lag j - rx (rows likewise)
void launch_map_remap(const float* in, int nx, int ny, float* out, int rx, int ry, hipStream_t st);
void launch_argmax_finish(const float* arg_part, int nparts, const float* map, float* res, hipStream_t st);

// ── roto-scale maps ──
// The canvas stack (5 planes W x H) on the n_th x n_lam log-polar grid about (cx, cy) (canvas
// pixels): rho, cth, sth tables of launch_rs_tables.
struct LogPolarArgs {
  const float* canvas;
  int W, H, n_th, n_lam;
  double cx, cy;
  const double* rho;
  const double* cth;
  const double* sth;
};
void launch_rs_tables(int n_th, int n_lam, double log0, double span, double* rho, double* cth, double* sth,
                      hipStream_t st);
// plans 0-3; nx >= n_th, ny >= n_lam powers of two; writes rows < n_lam of n_cplx spectra
bool launch_pack_logpolar(int plan, const LogPolarArgs& a, int nx, int ny, float2* spec, const float2* tw_x,
                          hipStream_t st);
// The n x n map from correlations on a px x py grid (px = n, or a power of two >= 2n folded).
struct RsArgs {
  int n, px, py;
  double dlam, dth;
};
void launch_combine_rs(int plan, const float* corr, float* map, const RsArgs& r, const CombineArgs& a,
                       hipStream_t st);
// arg_part: combine_blocks(n·n) pairs; then launch_argmax_finish over them
void launch_rs_finish(const float* fwd, const float* inv, int n, bool mirror, double dlam, double lo, double hi,
                      float* out, float* arg_part, hipStream_t st);

// plain 2-D FFT of `planes` n × n planes (tests, benchmarks)
bool launch_fft2(float2* data, int planes, int n, bool inverse, const float2* tw, hipStream_t st, int lines = 0);
// the same with a process-wide twiddle table per length
bool launch_fft2_twiddled(float2* data, int planes, int n, bool inverse, hipStream_t st);

// ── pose moments (45 sums per pose, per-block partials) ──
int moments_blocks(int w, int h);
void launch_pose_moments(const float* feat_a, int wa, int ha, const float* feat_b, int wb, int hb, const PoseArgs& p,
                         double* part, hipStream_t st);

// ── batched pose sums and gradients (one pose per job) ──
// moving features fa (wa × ha) sampled at Hi · X for every pixel X of fb (wb × hb); fa / fb: the
// sides' interleaved features (one float4 per pixel)
struct PoseJob {
  const float4* fa;
  const float4* fb;
  int wa, ha, wb, hb;
  double Hi[9];
};
// the tangent pass of one pose: Hi, M_k = −Hi · G_k (pixel generators of fb's image, zero past
// the group's), the score's coefficients (W, M̂, Q row-major, the means, g = ∂S/∂sums)
struct TangentJob {
  const float4* fa;
  const float4* fb;
  int wa, ha, wb, hb, hess, _pad;
  double Hi[9];
  float Mk[72];
  float W[16], Mh[16], Q[16], ma[4], mb[4], g[45];
};
// ── Gauss–Newton batches on the device (k_gn.hip) ──
inline constexpr int kGnMaxTries = 17;   // the ascent ladder (at most 12) and 5 step lengths
// a pose's image pair: moving side a, fixed side b (interleaved features)
struct GnPair {
  const float4* fa;
  const float4* fb;
  int wa, ha, wb, hb;
};
// a pose in the batch: its current pose and score; active while it moves
struct GnState {
  double H[9];
  double score;
  int active, moved;
};
struct GnArgs {
  int n, homography, exact, symmetric, min_n, n_amps;
  int n_try;   // try slots per pose: n_amps + 5 (at most kGnMaxTries)
  double ridge;
  double amps[12];
};
// device buffers of a batch of n poses (nd = n, or 2n when symmetric; T = GnArgs::n_try)
struct GnBuffers {
  const GnPair* pairs;
  GnState* state;
  PoseJob* gjobs;        // nd
  double* gsums;         // nd · 45
  TangentJob* tjobs;     // nd
  double* tsums;         // nd · 44
  double* dir_score;     // nd
  double* dir_n;         // nd
  double* step;          // n · 18: unit ascent direction, Gauss-Newton step, its norm, flags
  double* try_H;         // n · T · 9
  int* try_ok;           // n · T: the slot holds a try
  PoseJob* try_jobs;     // nd · T
  double* try_sums;      // nd · T · 45
  double* try_score;     // nd · T
  double* part;          // nd · T · blocks · 45
  int* moved_flag;       // 1
  int* iters_done;       // 1
};
// one iteration for every active pose (kernel launches only: capturable)
void launch_gn_iteration(const GnBuffers& b, const GnArgs& a, int blocks_grad, int blocks_try, hipStream_t st);

// ── the climb on the device (gclimb.zig:313-489 run; climb.wgsl hop_cand, rs_cand, gn_cands,
// climb_select, iter_begin / iter_end): pose 0 of a one-pose GnBuffers batch is the climb's pose ──
struct ClimbDev {
  int done, moved, stall, iter, n_rec;
  int step_ok;   // the Gauss–Newton step at the current pose was tried and beat nothing
};
struct ClimbArgs {
  int hop, hop_fm, n_iters, homography;
  int ox, oy;                 // the batch's canvas origin (the maps' frame)
  int map_nx, map_ny;         // the shift map's grid
  double cell_x, cell_y;      // canvas pixels per map cell
  int rs_n, rs_nth;           // the roto-scale map's grid
  double rs_dlam, rs_c0, rs_c1;   // its log-radius step and centre (1-based fixed pixels)
};
// iteration start: no move yet; hi = the canvas → moving inverse of the pose (the warp's Hi)
void launch_climb_begin(ClimbDev* c, GnState* s, const ClimbArgs& a, double* hi, bool reset_moved, hipStream_t st);
// a hop: try slot 0 of pose 0 from the map's peak (res: [peak, index bits, map[0]]); rs: from the
// roto-scale map's peak, else from the shift map's
void launch_climb_hop(const GnBuffers& b, const GnArgs& ga, ClimbDev* c, const ClimbArgs& a, const float* res,
                      bool rs, hipStream_t st);
// the tries' scores (one-pose batch)
void launch_climb_try_scores(const GnBuffers& b, const GnArgs& ga, hipStream_t st);
// the best try that beats the pose by 1e-6 of its score replaces it; gn: a step with no such try
// ends the Gauss–Newton block (the pose leaves the batch)
void launch_climb_select(const GnBuffers& b, const GnArgs& ga, ClimbDev* c, bool gn, hipStream_t st);
// a Gauss–Newton block: up to three steps of five lengths
void launch_climb_gn_block(const GnBuffers& b, const GnArgs& ga, ClimbDev* c, int blocks_grad, int blocks_try,
                           hipStream_t st);
void launch_climb_end(ClimbDev* c, const ClimbArgs& a, hipStream_t st);
// lengths the FFT kernels take: powers of two from 32 to 4096, and 1536, 2304 and 3072
inline bool fft_length_ok(int n) {
  return (n >= 32 && n <= 4096 && (n & (n - 1)) == 0) || n == 1536 || n == 2304 || n == 3072;
}
// workgroups per job of a launch of njobs pose jobs of at most max_pixels destination pixels
int jobs_blocks(int64_t max_pixels, int njobs);
// 45 sums per job into sums (part: njobs · blocks · 45 doubles)
void launch_pose_moments_jobs(const PoseJob* jobs, int njobs, int blocks, double* part, double* sums,
                              hipStream_t st);
// 44 sums per job (8 gradient + 36 upper-triangle matrix) into sums (part: njobs · blocks · 44)
void launch_tangent_jobs(const TangentJob* jobs, int njobs, int blocks, double* part, double* sums,
                         hipStream_t st);

// ── rotation × scale sweep ──
// A candidate: dest → source homography (row-major, 1-based pixels), taps per cell side and
// the window's first canvas cell (sweep.wgsl:106-110).
struct SweepPose {
  float h[9];
  int taps, gx0, gy0;
};
// One pack launch: the canvas (origin ox, oy in 1-based dest pixels, `cell` pixels per cell,
// grid × grid cells), the source image (sw × sh) and the window (win_x × win_y cells) in an
// nx × ny plane.
struct SweepPack {
  float ox, oy, cell;
  int grid, sw, sh, win_x, win_y, ny;
};
void launch_interleave(const float* feat, int64_t n, float4* il, hipStream_t st);
bool launch_pack_sweep(int plan, const float4* il, const SweepPose* poses, int batch, const SweepPack& u, int nx,
                       float2* spec, const float2* tw_x, hipStream_t st);
void launch_combine_sweep(int plan, const float* corr, float* ns, int64_t nn, int batch, const CombineArgs& a,
                          const int* zero_lag, hipStream_t st);
int peaks_chunks(int64_t nn);
// nx, ny powers of two; part: batch · peaks_chunks(nx·ny) · 256 entries; peaks: batch · n_peak
// (value, lag bits)
void launch_peaks(const float* ns, int nx, int ny, int batch, int n_peak, float2* part, float2* peaks,
                  hipStream_t st);

}  // namespace zc
