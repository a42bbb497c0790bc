// The SMI engine behind the C ABI: one HIP device and stream, an image pair whose features stay
// resident in HBM, and the map pipeline. Mirrors the reference engine's responsibilities for
// the hot path (engine.zig:136-873): images → features, pose, maps, scores.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "events.hpp"
#include "geometry.hpp"
#include "kernels.hpp"
#include "search.hpp"
#include "settings.hpp"

namespace zc {

// a profiled launch on the engine's stream
#define LAUNCH(name, call) \
  do {                     \
    pre();                 \
    call;                  \
    post(name);            \
  } while (0)

#define ZC_HIP(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + #expr); \
  } while (0)

// A device buffer that grows on demand (contents are not preserved).
struct DBuf {
  void* p = nullptr;
  size_t cap = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      cap = o.cap;
      o.p = nullptr;
      o.cap = 0;
    }
    return *this;
  }
  // a buffer an owner forgot to release is freed with it (the owner makes its device current)
  ~DBuf() { release(); }
  // counts every allocation in the process: a captured graph whose key holds it is not replayed
  // after any buffer moved
  static inline std::atomic<uint64_t> gen{0};
  void ensure(size_t bytes) {
    if (bytes <= cap) return;
    if (p) ZC_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    ZC_HIP(hipMalloc(&p, bytes));
    cap = bytes;
    ++gen;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// ABI result structs: layouts identical to the reference's (engine.zig:32-87, search.zig:30-41,
// climb.zig:7-12; python/zcmir/_native.py:14-58).
struct ShiftResult {
  uint32_t n, cw, ch, canvas_w, canvas_h;
  int32_t ox, oy;
  uint32_t peak_index;
  double dx, dy, peak, zero, corr_area_px;
  uint32_t exact, calibrated, n_y, _pad;
};
static_assert(sizeof(ShiftResult) == 88, "ShiftResult layout");

struct RsResult {
  uint32_t n, n_th, n_lam, peak_index;
  double dlam, r0, r1, cx, cy, peak, zero, dth, dsg;
  uint32_t symmetric, calibrated;
};
static_assert(sizeof(RsResult) == 96, "RsResult layout");

struct GradResult {
  double score, fwd, inv, n;
  double grad[8];
  double hess[64];
  uint32_t has_hess, nk;
};
static_assert(sizeof(GradResult) == 616, "GradResult layout");

struct SearchResult {
  double H[9];
  double score, prev;
  uint32_t improved, count_a, count_b, _pad;
};

struct ClimbResult {
  double H[9];
  double score;
  uint32_t iterations, moved;
};

// Per-kernel timing with HIP events on the engine's stream (zc_profile / zc_profile_report).
// Launches back to back share events: a launch's start is the previous launch's end event, so a
// profiled map records one event per kernel (its duration includes the launch gap before it).
// Anything else put on the stream (copies, host waits) breaks the chain.
class Prof {
 public:
  bool on = false;
  void begin(hipStream_t st);
  void end(const char* name, hipStream_t st);
  void break_chain() { last_ = nullptr; }
  void collect();   // after a synchronization
  std::string report() const;
  void reset() {
    acc_.clear();
    bytes_.clear();
    work_.clear();
  }
  // algorithmic HBM bytes of a launch (minimal traffic of its role), summed per kernel name
  void add_bytes(const char* name, double b) { bytes_[name] += b; }
  // work units of a launch (pose kernels: destination pixels × poses), summed per kernel name
  void add_work(const char* name, double u) { work_[name] += u; }
  ~Prof();

 private:
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<hipEvent_t> pool_;
  std::vector<Pending> pending_;
  hipEvent_t cur_ = nullptr;
  hipEvent_t last_ = nullptr;   // the end event of the previous launch, when nothing came between
  std::map<std::string, std::pair<long, double>> acc_;
  std::map<std::string, double> bytes_;
  std::map<std::string, double> work_;
  hipEvent_t get();
};

inline constexpr double kRidge = 1e-4;   // smi.zig:18
inline constexpr int kMinN = 12;          // smi.zig:17
inline constexpr int kMaxFft = 4096;      // longest FFT line the kernels implement
// Largest image (pixels): the pose kernels address a feature plane with 32-bit byte offsets.
inline constexpr int64_t kMaxImagePixels = int64_t(1) << 30;

struct Side {
  int w = 0, h = 0;
  bool has = false;
  bool preprocess = false;
  DBuf raw;          // the image as given (f32 gray), kept to bake again when settings change
  std::string baked; // the settings the work image and features were made with
  DBuf gray, rank, feat, mag, stack;
  // features interleaved per pixel (float4) for the sweep's and the pose kernels' gathers, made
  // on first use (a cache of feat: const users may fill it)
  mutable DBuf il;
  mutable bool il_ok = false;
  // the image's own summed-area tables of its 14 stack planes other than the weight (pairs), made
  // on first use by a translation-plan map
  mutable DBuf tab;
  mutable bool tab_ok = false;
  float mean = 0, spread = 1;
  uint32_t hist[257] = {};
  uint32_t cdf[258] = {};
  double acf[289 * 4] = {};
  std::string mode;
  bool half = false;
};

// Column-FFT tiling forced by zc_fft_tune (0: the autotuned or built-in one); the row tiling is
// FftRowTiling (kernels.hpp).
struct FftTune {
  static inline std::atomic<int> cols{0};
};

// Candidates per sweep batch by default: 256 (the reference uses 64; a batch only schedules the
// candidates' independent maps, and 256 measured 47.0K against 44.7K poses/s on the 128² canvas).
inline constexpr uint32_t kSweepBatch = 256;
// Sweep tuning: candidates per batch (0: the timed best per plane size up to kSweepBatch when the
// fft_autotune setting is on, else kSweepBatch).
struct SweepTune {
  static inline std::atomic<int> batch{0};
  // work buffers of one batch, GiB (at most half the free device memory): 96 lets a
  // full-resolution sweep of a 1024² pair batch 56, 44 and 25 candidates at 2048², 2304² and
  // 3072² planes, 5 % faster than the 18, 14 and 8 of 32 GiB
  static inline std::atomic<int> budget_gb{96};
};

struct SweepOpts {
  Mat3 H0;
  const SweepGrid* gv;
  uint32_t grid = 128;   // canvas cells per side
  bool full = false;     // one cell per fixed pixel (grid = ceil(D))
  bool smooth = true;    // planes of 1536, 2304 or 3072 points where they are the shortest that fit
  uint32_t mode = 1;     // 0 global, 1 exact SMI
  uint32_t n_peak = 4;
  uint32_t batch = kSweepBatch;   // the most candidates per batch
  bool tune_batch = false;        // time batch sizes up to `batch` per plane size on first use
  double min_overlap = 0.05;
  // only these (rotation, scale) index pairs of the grid (shear and stretch index 0), in order
  const std::vector<std::array<uint32_t, 2>>* cells = nullptr;
};

struct SweepOut {
  std::vector<SweepPeak> peaks;
  std::vector<float> best;   // each candidate's map maximum (θ-major, then σ, shear, stretch)
  uint32_t n = 0, nc = 0, grid = 0;
  double cell = 0, D = 0, ox = 0, oy = 0;
};

// Score and gradient coefficients of one direction's 45 sums (moments.zig:24-28 Stats).
struct Coefs {
  double score, n;
  double W[16], Mh[16], Q[16], ma[4], mb[4], g[45];
};
Coefs score_coefs(const double* st, bool exact, double ridge);

// A pose's score, overlap, tangent gradient and Gauss–Newton matrix (nk × nk row-major)
// (pose.zig:76-82 Grad).
struct GradOut {
  double score, n;
  double grad[8];
  double hess[64];
  bool has_hess;
};
struct GnOut {
  Mat3 H;
  double score;
  bool moved;
};

struct MapOut {
  Frame fr;
  GridPlan g;    // the grid the kernels ran on (power-of-two FFT lengths)
  GridPlan gr;   // the reference's plan (2^a·3^b lengths): the map and peak index the ABI returns
  bool exact, calib;
  float res[3];
};

// One direction of a pose: side a's features sampled onto side b's pixels at H (a → b).
struct DirJob {
  const Side* a;
  const Side* b;
  Mat3 H;
};
// An image pair: moving side a, fixed side b.
struct PairSides {
  const Side* a;
  const Side* b;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();

  int dev;
  hipStream_t st = nullptr;
  // a second stream (default priority) for work that runs beside the main sequence: the map's
  // summed-area tables; ev_fork / ev_join order it against st
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  Settings set;
  Side sides[2];
  Mat3 H = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::string err;
  std::atomic<bool> cancel{false};
  Prof prof;
  Events events;   // progress events (searches, climbs), drained by zc_events
  uint32_t counts[4] = {0, 0, 0, 0};   // dispatches, submits, reads, waits
  std::string device_name;

  void set_image(int side, const float* data, int w, int h, bool preprocess);
  bool has_images() const { return sides[0].has && sides[1].has; }
  // Images ready: re-bakes a side whose work image or features were made with other settings
  // (engine.zig:296-303 needImages → ensureBaked / syncFeatures).
  void need_images();
  // The settings a side's work image and features depend on.
  std::string bake_key(const Side& sd) const;
  // Work image (preprocessing per the settings when sd.preprocess) from sd.raw, then features
  // (engine.zig:321-340 bake).
  void bake(Side& sd);
  void set_image_rgba(int s, const uint8_t* rgba, int w, int h);
  double pair_area() const { return pair_area(sides[0], sides[1]); }
  double pair_area(const Side& A, const Side& B) const;

  // map of H into the engine's map buffer; returns its plan and [peak, idx, map[0]]
  // wait=false: enqueue only (the result lands in res_host at the next synchronization)
  MapOut shift_map(const Mat3& H, bool exact, bool calib, uint32_t cap, double min_overlap, bool wait = true);
  // frame: the canvas (default: H's own, reg_frame); the climb keeps one frame for a batch of
  // iterations (gclimb.zig:384-387)
  // dev_hi: the canvas → moving inverse pose in device memory (9 doubles, a pose device code
  // chose): the moving stack is warped from it, the plan is the one of a general pose (no
  // translation plan), and nothing is copied to the host (the peak stays in `res`); wait = false.
  MapOut shift_map_sides(Side& A, Side& B, const Mat3& H, bool exact, bool calib, uint32_t cap, double min_overlap,
                         bool wait = true, const Frame* frame = nullptr, const double* dev_hi = nullptr);
  float* res_host = nullptr;   // pinned
  void fill_shift_result(const MapOut& m, ShiftResult* out) const;
  void read_map(float* out, size_t len, const MapOut& m);

  // Roto-scale map at H about c (null: the moving centroid), smi.zig:577-651, engine.zig:836-879.
  // Optional outputs (len floats each): the finished map, the forward and inverse-frame maps.
  RsResult rs_map(const Mat3& H, const double* c, bool symmetric, float* map, float* fwd, float* inv, size_t len);
  struct RsPlan {
    int n, n_th, n_lam, px, py;
    double log0, span, dlam, dth, r0, r1, cx, cy;
  };
  // One direction: side A warped by H onto B's canvas, both on the log-polar grid about c
  // (1-based fixed pixels); r0 < 0 / r1 <= 0: from the canvas. The n x n map into dst.
  // dev_hi: as for shift_map_sides
  RsPlan rs_map_dir(Side& A, Side& B, const Mat3& H, double c0, double c1, double r0, double r1, bool exact,
                    bool calib, uint32_t cap, DBuf& dst, const Frame* frame = nullptr, const double* dev_hi = nullptr);
  // Finish (symmetric: average with rs_inv) into rs_out and its peak into res (and res_host unless
  // to_host is false).
  void rs_finish(const RsPlan& p, bool symmetric, bool mirror, bool to_host = true);
  // The climb from the engine's pose (gclimb.zig:313-489): shift-map hops, roto-scale hops and
  // Gauss–Newton blocks per iteration; the engine's pose becomes the result.
  ClimbResult climb(uint32_t n_iters, bool hop, bool hop_fm, bool homography, bool exact, bool symmetric);
  // The same climb with its decisions on the device (climb_graph): one captured graph per iteration
  // (map hop, Gauss–Newton blocks, roto-scale hop, stall and stop rule over device state); the host
  // reads the state once per iteration and sets each batch's frame.
  ClimbResult climb_device(uint32_t n_iters, bool hop, bool hop_fm, bool homography, bool exact, bool symmetric);

  void moments(int a_side, int b_side, const Mat3& H, double out[45]);
  // SMI from 45 sums (moments.zig:91-134), FP64
  static double score_from_moments(const double* st, bool exact, double ridge);

  void fft2(float* data, int n, int planes, bool inverse);
  // every candidate's shift map reduced to its top local maxima (search.zig:300-494)
  // A captured launch sequence, replayed while its key (every argument of every launch in it)
  // is unchanged; `kernels` = the launches it holds (for the dispatch counter).
  struct GraphCache {
    std::vector<uint8_t> key;
    hipGraphExec_t exec = nullptr;
    uint32_t kernels = 0;
    void reset();
  };
  GraphCache sweep_gc;
  GraphCache gn_gc;   // one Gauss-Newton iteration of the device batch
  GraphCache climb_gc;   // one iteration of the device climb
  // Issues `issue` (kernel launches only on st: no copies, fills, syncs or allocations) through
  // gc's graph: replayed when the key matches, else captured, instantiated and launched. Runs
  // `issue` eagerly when `eager` (profiling needs per-launch events).
  void run_captured(GraphCache& gc, std::vector<uint8_t>&& key, bool eager, const std::function<void()>& issue);
  SweepOut sweep(const SweepOpts& o);
  SweepOut sweep_sides(Side& A, Side& B, const SweepOpts& o);
  // batched pose evaluation (gn.cpp)
  void moments_jobs(const std::vector<DirJob>& d, std::vector<std::array<double, 45>>& out);
  void tangent_jobs(const std::vector<DirJob>& d, const std::vector<Coefs>& co, bool homography, bool hess,
                    std::vector<std::array<double, 44>>& out);
  std::vector<double> pair_scores(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, bool exact,
                                  bool symmetric);
  // the engine pair's score at H: mean, forward, inverse (NaN when not symmetric)
  std::array<double, 3> score_parts(const Mat3& H, bool exact, bool symmetric);
  std::vector<GradOut> pair_grads(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, bool exact,
                                  bool homography, bool symmetric, bool hess);
  std::vector<GnOut> gn_batch(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, uint32_t iters,
                              bool homography, bool exact, bool symmetric, const std::vector<double>& amps,
                              uint32_t* iterations = nullptr);
  // the same batch with every iteration on the device: one captured graph replayed per iteration,
  // the state read back once at the end
  std::vector<GnOut> gn_batch_device(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, uint32_t iters,
                                     bool homography, bool exact, bool symmetric, const std::vector<double>& amps,
                                     uint32_t* iterations);
  std::vector<PairSides> own_pair(size_t n) const;
  // best_peak: the sweep's best raw peak score (n in fixed pixels), nullable
  SearchResult search(const SweepOpts& o, uint32_t keep, uint32_t n_fin, bool homography, bool exact,
                      bool symmetric, const std::vector<double>& amps, double* best_peak = nullptr);
  // stacks (stack.cpp): slice pairs with their own features, registered together
  std::vector<std::unique_ptr<Side>> stack;   // moving, fixed per slice
  void stack_set(uint32_t n, const float* const* mv, const uint32_t* mwh, const float* const* fx,
                 const uint32_t* fwh);
  // init: 0 the start poses, 1 each corrected by its shift map's peak, 2 by the best peak of a
  // rotation × scale sweep about it (rotation ±12° in 4° steps × scale 0.88 – 1.12 in 5, 64² canvas)
  std::vector<GnOut> stack_register(const std::vector<Mat3>& H0, uint32_t iters, int init, std::vector<Mat3>* Hinit,
                                    uint32_t* iterations);
  static void release_side(Side& s);
  // the map of the last sweep batch's first candidate (tests)
  void read_sweep_map(float* out, size_t len, uint32_t dims[2]);
  uint32_t sweep_last_nx = 0, sweep_last_ny = 0;

  void sync();
  // launch bookkeeping
  void pre() { if (prof.on) prof.begin(st); }
  void bytes(const char* name, double b) {
    if (prof.on) prof.add_bytes(name, b);
  }
  void count_work(const char* name, double u) {
    if (prof.on) prof.add_work(name, u);
  }
  void post(const char* name) {
    ++counts[0];
    if (prof.on) prof.end(name, st);
    ZC_HIP(hipGetLastError());
  }

  DBuf canvas_a, canvas_b, spec_a, spec_b, work, corr, map, arg_part, res, parts, sums, tmp_u32, tmp_i64;
  DBuf map_out;   // a map laid out on the reference plan's grid (read_map)
  DBuf cells_buf;   // a resampled map's cell values (launch_pack_rowfft)
  DBuf sat_band, sat;   // rectangle plans: band sums and summed-area tables
  DBuf rs_fwd, rs_inv, rs_out, rs_tab;   // roto-scale maps: both directions, finished, log-polar tables
  DBuf pre_buf, pre_hist, pre_cdf, pre_part, pre_rgba;   // preprocessing scratch
  DBuf sw_poses, sw_pose_fixed, sw_zero, sw_dead, sw_ns, sw_part, sw_peaks;
  DBuf gn_jobs, gn_part, gn_sums;
  // the device Gauss-Newton batch (GnBuffers)
  DBuf gd_pairs, gd_state, gd_gjobs, gd_gsums, gd_tjobs, gd_tsums, gd_dir, gd_step, gd_try_H, gd_try_ok, gd_try_jobs,
      gd_try_sums, gd_try_score, gd_part, gd_ctr;
  DBuf cl_state;   // the device climb: ClimbDev, then the warp's inverse pose (9 doubles)
  std::map<int, DBuf> tw;
  const float2* twiddles(int L);
  // Column-FFT tiling for column length L (launch_colfft's `lines`): zc_fft_tune's value when set,
  // else (fft_autotune) the fastest of the tilings timed once per length and process, else 0.
  int col_lines(int L);
  int autotune_cols(int L);
  DBuf tune_buf;
  void ensure_acf(Side& sd);
  void ensure_il(const Side& sd);
  void ensure_tables(const Side& sd);

 private:
  void prepare(Side& sd, bool acf);
};

}  // namespace zc
