// The C ABI (include/zcsmi.h). Names, argument meaning, result structs and error behaviour
// follow the reference's exports (src/exports.zig, src/native_api.zig): 0 (or a size) on
// success, -1 on failure with the reason for zc_error. Exceptions never cross this boundary.
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "engine.hpp"
#include "lie.hpp"

using zc::Engine;
using zc::Mat3;

namespace {

thread_local std::string g_create_error;

size_t copy_out(const std::string& s, char* buf, size_t cap) {
  const size_t n = std::min(s.size(), cap);
  if (buf && n) std::memcpy(buf, s.data(), n);
  return n;
}

Mat3 mat(const double* H) {
  Mat3 m;
  for (int i = 0; i < 9; ++i) m[i] = H[i];
  return m;
}

template <typename F>
int32_t guard(Engine* e, const char* what, F&& f) {
  if (!e) return -1;
  try {
    (void)hipSetDevice(e->dev);
    f();
    return 0;
  } catch (const std::exception& ex) {
    e->err = std::string(what) + ": " + ex.what();
    return -1;
  }
}

int32_t unsupported(Engine* e, const char* what) {
  if (e) e->err = std::string(what) + ": Unsupported (not part of the SMI hot path this library implements)";
  return -1;
}

struct SweepInfo {
  uint32_t nc, n, grid, n_peaks;
  double cell, D, ox, oy;
};
static_assert(sizeof(SweepInfo) == 48, "SweepInfo layout");
static_assert(sizeof(zc::SweepPeak) == 128, "SweepPeak layout");

// The sweep options of the current settings (search.zig:663-711 runSweep).
zc::SweepOpts sweep_opts(Engine* e, zc::SweepGrid& gv) {
  const auto& s = e->set;
  auto clampf = [](double x, double lo, double hi) { return std::min(std::max(x, lo), hi); };
  auto clampu = [](uint32_t x, uint32_t lo, uint32_t hi) { return std::min(std::max(x, lo), hi); };
  if (s.e("metric") != "smi") throw std::runtime_error("Unsupported: the sweep of this library scores metric smi");
  const uint32_t g = s.u("sw_grid");
  const double rot = clampf(s.f("sw_rot"), 0, 180);
  const bool range = s.b("sw_rot_range");
  const double lo = range ? clampf(s.f("sw_rot_lo"), -360, 360) : -rot;
  double hi = range ? clampf(s.f("sw_rot_hi"), -360, 360) : rot;
  if (hi < lo) hi += 360;
  hi = std::min(hi, lo + 360);
  gv = zc::sweep_grid(lo, hi, clampu(s.u("sw_nth"), 1, 720), clampf(s.f("sw_smin"), 0.05, 1),
                      clampf(s.f("sw_smax"), 1, 20), clampu(s.u("sw_ns"), 1, 129),
                      s.b("sw_shear") ? clampf(s.f("sw_shear_max"), 0, 1) : 0, clampu(s.u("sw_nshear"), 1, 15),
                      s.b("sw_aniso") ? clampf(s.f("sw_aniso_max"), 0, 1) : 0, clampu(s.u("sw_naniso"), 1, 15));
  zc::SweepOpts o{};
  const double lw = e->sides[0].w, lh = e->sides[0].h, rw = e->sides[1].w, rh = e->sides[1].h;
  o.H0 = s.b("sw_rel") ? e->H : Mat3{1, 0, (rw - lw) / 2, 0, 1, (rh - lh) / 2, 0, 0, 1};
  o.gv = &gv;
  o.grid = (g == 32 || g == 64 || g == 128 || g == 256) ? g : 128;
  o.full = s.b("sw_full");
  o.smooth = s.b("sweep_smooth");
  o.mode = s.b("exact") ? 1 : 0;
  o.n_peak = 4;
  const int tb = zc::SweepTune::batch;
  o.batch = tb > 0 ? (uint32_t)tb : zc::kSweepBatch;
  o.tune_batch = tb <= 0 && s.b("fft_autotune");
  o.min_overlap = clampf(s.f("min_overlap"), 0, 0.5);
  return o;
}

}  // namespace

extern "C" {

int32_t zc_sweep_cells(void* p, const double* H0, const uint32_t* cells, uint32_t k, SweepInfo* info,
                       zc::SweepPeak* peaks, size_t cap, float* best, size_t nbest);

const char* zc_version(void) { return "0.1.4+mi355x"; }

void* zc_create(const char* dev) {
  try {
    int d = 0;
    if (dev && *dev) {
      char* end = nullptr;
      const long v = std::strtol(dev, &end, 10);
      if (end && *end == '\0') d = (int)v;   // a device index; any other text (a path) is ignored
    }
    return new Engine(d);
  } catch (const std::exception& ex) {
    g_create_error = ex.what();
    return nullptr;
  }
}

size_t zc_create_error(char* buf, size_t cap) { return copy_out(g_create_error, buf, cap); }

void zc_destroy(void* p) { delete static_cast<Engine*>(p); }

size_t zc_adapter(void* p, char* buf, size_t cap) {
  Engine* e = static_cast<Engine*>(p);
  return e ? copy_out(e->device_name, buf, cap) : 0;
}

size_t zc_error(void* p, char* buf, size_t cap) {
  Engine* e = static_cast<Engine*>(p);
  if (!e) return 0;
  const size_t n = copy_out(e->err, buf, cap);
  e->err.clear();
  return n;
}

int32_t zc_configure(void* p, const char* json, size_t len) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "configure", [&] { e->set.apply(json, len); });
}

size_t zc_settings(void* p, char* buf, size_t cap) {
  Engine* e = static_cast<Engine*>(p);
  if (!e) return 0;
  const std::string j = e->set.to_json();
  copy_out(j, buf, cap);
  return j.size();
}

int32_t zc_set_image(void* p, uint32_t side, const float* data, uint32_t w, uint32_t h) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "set_image", [&] { e->set_image((int)side, data, (int)w, (int)h, false); });
}

int32_t zc_set_image_gray(void* p, uint32_t side, const float* data, uint32_t w, uint32_t h, uint32_t preprocess) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "set_image_gray", [&] { e->set_image((int)side, data, (int)w, (int)h, preprocess != 0); });
}

// zc_set_image_gray from device memory (work that writes `data` must be complete).
int32_t zc_set_image_device(void* p, uint32_t side, const float* data, uint32_t w, uint32_t h) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "set_image_device", [&] { e->set_image((int)side, data, (int)w, (int)h, false); });
}

int32_t zc_set_image_rgba(void* p, uint32_t side, const uint8_t* rgba, uint32_t w, uint32_t h) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "set_image_rgba", [&] { e->set_image_rgba((int)side, rgba, (int)w, (int)h); });
}

void zc_image_size(void* p, uint32_t side, uint32_t out[2]) {
  Engine* e = static_cast<Engine*>(p);
  if (e && side < 2 && e->sides[side].has) {
    out[0] = (uint32_t)e->sides[side].w;
    out[1] = (uint32_t)e->sides[side].h;
  } else {
    out[0] = out[1] = 0;
  }
}

int32_t zc_work_image(void* p, uint32_t side, float* out, size_t len) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "work_image", [&] {
    if (side > 1 || !e->sides[side].has) throw std::runtime_error("NoImage");
    if (e->sides[side].baked != e->bake_key(e->sides[side])) e->bake(e->sides[side]);
    const auto& s = e->sides[side];
    const size_t n = (size_t)s.w * s.h;
    if (len < n) throw std::runtime_error("BufferTooSmall");
    ZC_HIP(hipMemcpyAsync(out, s.gray.p, n * 4, hipMemcpyDeviceToHost, e->st));
    e->sync();
  });
}

int32_t zc_features(void* p, uint32_t side, float* out, size_t len) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "features", [&] {
    if (side > 1 || !e->sides[side].has) throw std::runtime_error("NoImage");
    if (e->sides[side].baked != e->bake_key(e->sides[side])) e->bake(e->sides[side]);
    const auto& s = e->sides[side];
    const size_t n = 4 * (size_t)s.w * s.h;
    if (len < n) throw std::runtime_error("BufferTooSmall");
    ZC_HIP(hipMemcpyAsync(out, s.feat.p, n * 4, hipMemcpyDeviceToHost, e->st));
    e->sync();
  });
}

int32_t zc_moments(void* p, const double* H, double* out) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "moments", [&] { e->moments(0, 1, mat(H), out); });
}

int32_t zc_score(void* p, const double* H, double* out) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "score", [&] {
    const bool exact = e->set.b("exact");
    double m[45];
    const Mat3 Hm = mat(H);
    e->moments(0, 1, Hm, m);
    const double fwd = Engine::score_from_moments(m, exact, 1e-4);
    if (!e->set.b("symmetric")) {
      out[0] = fwd;
      out[1] = fwd;
      out[2] = std::numeric_limits<double>::quiet_NaN();
      return;
    }
    e->moments(1, 0, zc::inv3(Hm), m);
    const double inv = Engine::score_from_moments(m, exact, 1e-4);
    out[0] = 0.5 * (fwd + inv);
    out[1] = fwd;
    out[2] = inv;
  });
}

int32_t zc_shift_map(void* p, const double* H, uint32_t flags, uint32_t cap, zc::ShiftResult* out, float* map,
                     size_t map_len) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "shift_map", [&] {
    const auto m = e->shift_map(mat(H), (flags & 1) != 0, (flags & 2) != 0, cap == 0 ? 1024 : cap, 0.0);
    if (map) e->read_map(map, map_len, m);
    e->fill_shift_result(m, out);
  });
}

int32_t zc_shift_map_set(void* p, const double* H, zc::ShiftResult* out, float* map, size_t map_len) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "shift_map", [&] {
    const std::string metric = e->set.e("metric");
    if (metric != "smi") throw std::runtime_error("Unsupported metric " + metric + " (smi only)");
    const auto m = e->shift_map(mat(H), e->set.b("exact"), e->set.b("calibrated"), e->set.u("map_res"),
                                e->set.f("min_overlap"));
    if (map) e->read_map(map, map_len, m);
    e->fill_shift_result(m, out);
  });
}

int32_t zc_rs_map(void* p, const double* H, const double* c, uint32_t symmetric, zc::RsResult* out, float* map,
                  float* fwd, float* inv, size_t len) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "rs_map", [&] { *out = e->rs_map(mat(H), c, symmetric != 0, map, fwd, inv, len); });
}

void zc_set_pose(void* p, const double* H) {
  Engine* e = static_cast<Engine*>(p);
  if (e) e->H = mat(H);
}

void zc_get_pose(void* p, double* H) {
  Engine* e = static_cast<Engine*>(p);
  if (e)
    for (int i = 0; i < 9; ++i) H[i] = e->H[i];
}

// exports.zig:410-414: queued progress events (searches, climbs), whole JSON lines up to cap
// bytes, removed from the queue; 0 when none are left (or the next line is longer than cap).
size_t zc_events(void* p, char* buf, size_t cap) {
  Engine* e = static_cast<Engine*>(p);
  return e ? e->events.take(buf, cap) : 0;
}

int32_t zc_profile(void* p, uint32_t on) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "profile", [&] {
    e->sync();
    e->prof.on = on != 0;
    e->prof.reset();
  });
}

size_t zc_profile_report(void* p, char* buf, size_t cap) {
  Engine* e = static_cast<Engine*>(p);
  if (!e) return 0;
  const std::string r = e->prof.report();
  copy_out(r, buf, cap);
  return r.size();
}

void zc_gpu_counts(void* p, uint32_t out[4]) {
  Engine* e = static_cast<Engine*>(p);
  if (!e) return;
  for (int i = 0; i < 4; ++i) {
    out[i] = e->counts[i];
    e->counts[i] = 0;
  }
}

void zc_cancel(void* p) {
  Engine* e = static_cast<Engine*>(p);
  if (e) e->cancel = true;
}

void zc_swap(void* p) {
  Engine* e = static_cast<Engine*>(p);
  if (!e) return;
  std::swap(e->sides[0], e->sides[1]);
}

int32_t zc_fft2(void* p, float* data, uint32_t n, uint32_t planes, uint32_t inverse, uint32_t four_step) {
  Engine* e = static_cast<Engine*>(p);
  (void)four_step;
  return guard(e, "fft2", [&] { e->fft2(data, (int)n, (int)planes, inverse != 0); });
}

void zc_result_sizes(uint32_t out[4]) {
  out[0] = sizeof(zc::ShiftResult);
  out[1] = sizeof(zc::RsResult);
  out[2] = sizeof(zc::GradResult);
  out[3] = 0;
}

// ── inspection and benchmark hooks of this library (no reference counterpart) ──
int32_t zc_rank(void* p, uint32_t side, float* rank, size_t len, uint32_t* hist, uint32_t* cdf, float* stats) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "rank", [&] {
    if (side > 1 || !e->sides[side].has) throw std::runtime_error("NoImage");
    if (e->sides[side].baked != e->bake_key(e->sides[side])) e->bake(e->sides[side]);
    const auto& s = e->sides[side];
    const size_t n = (size_t)s.w * s.h;
    if (rank) {
      if (len < n) throw std::runtime_error("BufferTooSmall");
      ZC_HIP(hipMemcpyAsync(rank, s.rank.p, n * 4, hipMemcpyDeviceToHost, e->st));
      e->sync();
    }
    if (hist) std::memcpy(hist, s.hist, sizeof s.hist);
    if (cdf) std::memcpy(cdf, s.cdf, sizeof s.cdf);
    if (stats) {
      stats[0] = s.mean;
      stats[1] = s.spread;
    }
  });
}

// `iters` shift maps of H back to back, no readback between them; the mean time per map (ms,
// HIP events on the engine's stream) into ms_per_map.
int32_t zc_bench_shift_map(void* p, const double* H, uint32_t flags, uint32_t cap, uint32_t iters,
                           double* ms_per_map) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "bench_shift_map", [&] {
    if (!(flags & 4)) {
      // one untimed map first (buffers, tilings, twiddles in place), outside any profile
      struct Pause {
        zc::Prof& p;
        bool was;
        ~Pause() {
          p.on = was;
          p.break_chain();
        }
      } pause{e->prof, e->prof.on};
      e->prof.on = false;
      (void)e->shift_map(mat(H), (flags & 1) != 0, (flags & 2) != 0, cap, 0.0);
    }
    hipEvent_t a, b;
    ZC_HIP(hipEventCreate(&a));
    ZC_HIP(hipEventCreate(&b));
    ZC_HIP(hipEventRecord(a, e->st));
    for (uint32_t i = 0; i < iters; ++i) e->shift_map(mat(H), (flags & 1) != 0, (flags & 2) != 0, cap, 0.0, false);
    ZC_HIP(hipEventRecord(b, e->st));
    e->sync();
    float ms = 0;
    ZC_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_map = iters ? ms / iters : 0.0;
  });
}

// ── the sweep on its own (search.zig:300-494): raw peaks of every candidate, no NMS or climb ──
// H0: the prior pose (null: the settings' rule, centred or the current pose with sw_rel).
// Writes up to `cap` raw peaks and each candidate's best score; returns the peak count.
int32_t zc_sweep(void* p, const double* H0, SweepInfo* info, zc::SweepPeak* peaks, size_t cap, float* best,
                 size_t nbest) {
  return zc_sweep_cells(p, H0, nullptr, 0, info, peaks, cap, best, nbest);
}

// zc_sweep over k given (rotation index, scale index) cells of the grid only (cells: 2·k; NULL:
// the whole grid); `best` in the cells' order.
int32_t zc_sweep_cells(void* p, const double* H0, const uint32_t* cells, uint32_t k, SweepInfo* info,
                       zc::SweepPeak* peaks, size_t cap, float* best, size_t nbest) {
  Engine* e = static_cast<Engine*>(p);
  int32_t count = 0;
  const int32_t rc = guard(e, "sweep", [&] {
    e->need_images();
    zc::SweepGrid gv;
    zc::SweepOpts o = sweep_opts(e, gv);
    std::vector<std::array<uint32_t, 2>> cl(cells ? k : 0);
    for (size_t i = 0; i < cl.size(); ++i) cl[i] = {cells[2 * i], cells[2 * i + 1]};
    if (cells) o.cells = &cl;
    if (H0) o.H0 = mat(H0);
    const zc::SweepOut r = e->sweep(o);
    if (info) *info = {r.nc, r.n, r.grid, (uint32_t)r.peaks.size(), r.cell, r.D, r.ox, r.oy};
    if (peaks) {
      const size_t n = std::min(cap, r.peaks.size());
      std::memcpy(peaks, r.peaks.data(), n * sizeof(zc::SweepPeak));
    }
    if (best) std::memcpy(best, r.best.data(), std::min(nbest, r.best.size()) * sizeof(float));
    count = (int32_t)r.peaks.size();
  });
  return rc < 0 ? rc : count;
}

// exports.zig:424-427: score and tangent gradient at H (hess: also the Gauss–Newton matrix),
// symmetric per the settings (engine.zig:786-797; fwd and inv are left 0 as there).
int32_t zc_gradient(void* p, const double* H, uint32_t hess, zc::GradResult* out) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "gradient", [&] {
    e->need_images();
    const bool homog = e->set.e("group") == "homography";
    const auto g = e->pair_grads(e->own_pair(1), {mat(H)}, e->set.b("exact"), homog, e->set.b("symmetric"),
                                 hess != 0)[0];
    std::memset(out, 0, sizeof *out);
    const int nk = zc::n_keys(homog);
    out->score = g.score;
    out->n = g.n;
    for (int k = 0; k < 8; ++k) out->grad[k] = g.grad[k];
    if (hess) {
      for (int k = 0; k < nk * nk; ++k) out->hess[k] = g.hess[k];
      out->has_hess = 1;
    }
    out->nk = (uint32_t)nk;
  });
}

// exports.zig:405-408: climb from the engine's pose (gclimb.zig:313-489): shift-map hops (`hop`),
// roto-scale hops (`hop_fm`) and Gauss–Newton blocks, g_iters iterations at most; the result is the
// engine's new pose. No B-spline stage (the spline is not part of this library).
int32_t zc_climb(void* p, zc::ClimbResult* out) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "climb", [&] {
    const auto& s = e->set;
    const std::string metric = s.e("metric");
    if (metric != "smi") throw std::runtime_error("Unsupported metric " + metric + " (smi only)");
    *out = e->climb(s.u("g_iters"), s.b("hop"), s.b("hop_fm"), s.e("group") == "homography", s.b("exact"),
                    s.b("symmetric"));
  });
}

// exports.zig:376-380 / search.zig:646-822: the sweep search with the current settings; the best
// pose replaces the engine's when it scores higher.
int32_t zc_search(void* p, zc::SearchResult* out) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "search", [&] {
    e->need_images();
    e->cancel = false;
    const auto& s = e->set;
    if (s.e("search_mode") != "sweep") throw std::runtime_error("Unsupported: search_mode cloud");
    zc::SweepGrid gv;
    const zc::SweepOpts o = sweep_opts(e, gv);
    const uint32_t k = s.u("sw_k");
    const uint32_t keep = (k == 8 || k == 16 || k == 32) ? k : 32;
    const uint32_t n_fin = std::min<uint32_t>(std::max<uint32_t>(s.u("sw_fin"), 1), 16);
    const auto amps = zc::geo_amps(s.f("g_a0"), s.f("g_decay"), s.u("g_steps") > 0 ? s.u("g_steps") : 7);
    *out = e->search(o, keep, n_fin, s.e("group") == "homography", s.b("exact"), s.b("symmetric"), amps);
  });
}

// ── stacks: many slice pairs registered together (the role of python/zcmir/api.py:522-580) ──
// Slice i: moving[i] (moving_wh[2i] × moving_wh[2i+1] f32 gray, used as the work image) and
// fixed[i]; their features stay resident until the next zc_stack_set.
int32_t zc_stack_set(void* p, uint32_t n, const float* const* moving, const uint32_t* moving_wh,
                     const float* const* fixed, const uint32_t* fixed_wh) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "stack_set", [&] { e->stack_set(n, moving, moving_wh, fixed, fixed_wh); });
}

// Register every slice of the stack: start poses H0 (n × 9, NULL: each moving image centred on
// its fixed one), init_map: 1 first correct each start by its SMI shift map's peak (map_res,
// min_overlap, exact), 2 by the best peak of a small rotation × scale sweep about it (±12°,
// 0.88 – 1.12), 0 neither; then `iters` Gauss–Newton iterations on all slices at once (group,
// symmetric, exact, g_a0 / g_decay / g_steps). Writes the poses (n × 9), their scores (n) and,
// if not NULL, the poses after the map (n × 9); returns the Gauss–Newton iterations run.
int32_t zc_stack_register(void* p, const double* H0, uint32_t iters, uint32_t init_map, double* H_out,
                          double* scores, double* H_init) {
  Engine* e = static_cast<Engine*>(p);
  uint32_t it = 0;
  const int32_t rc = guard(e, "stack_register", [&] {
    const size_t n = e->stack.size() / 2;
    std::vector<Mat3> h0(n);
    for (size_t i = 0; i < n; ++i) {
      if (H0) {
        h0[i] = mat(H0 + 9 * i);
      } else {
        const double lw = e->stack[2 * i]->w, lh = e->stack[2 * i]->h;
        const double rw = e->stack[2 * i + 1]->w, rh = e->stack[2 * i + 1]->h;
        h0[i] = {1, 0, (rw - lw) / 2, 0, 1, (rh - lh) / 2, 0, 0, 1};
      }
    }
    std::vector<Mat3> hinit;
    e->cancel = false;
    if (init_map > 2) throw std::runtime_error("BadArgument: init 0 (none), 1 (shift map) or 2 (sweep)");
    const auto r = e->stack_register(h0, iters, (int)init_map, &hinit, &it);
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 9; ++k) {
        H_out[9 * i + k] = r[i].H[k];
        if (H_init) H_init[9 * i + k] = hinit[i][k];
      }
      if (scores) scores[i] = r[i].score;
    }
  });
  return rc < 0 ? rc : (int32_t)it;
}

// zc_search over `k` given (rotation index, scale index) cells of the sweep grid only (cells:
// 2·k indices; the learned prior's proposals); best_peak (nullable): the sweep's best raw peak
// score divided by the fixed image's area.
int32_t zc_search_cells(void* p, const uint32_t* cells, uint32_t k, zc::SearchResult* out, double* best_peak) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "search_cells", [&] {
    e->need_images();
    e->cancel = false;
    const auto& s = e->set;
    zc::SweepGrid gv;
    zc::SweepOpts o = sweep_opts(e, gv);
    std::vector<std::array<uint32_t, 2>> cl(k);
    for (uint32_t i = 0; i < k; ++i) cl[i] = {cells[2 * i], cells[2 * i + 1]};
    o.cells = &cl;
    const uint32_t kk = s.u("sw_k");
    const uint32_t keep = (kk == 8 || kk == 16 || kk == 32) ? kk : 32;
    const uint32_t n_fin = std::min<uint32_t>(std::max<uint32_t>(s.u("sw_fin"), 1), 16);
    const auto amps = zc::geo_amps(s.f("g_a0"), s.f("g_decay"), s.u("g_steps") > 0 ? s.u("g_steps") : 7);
    double bp = 0;
    *out = e->search(o, keep, n_fin, s.e("group") == "homography", s.b("exact"), s.b("symmetric"), amps, &bp);
    if (best_peak) *best_peak = bp / ((double)e->sides[1].w * e->sides[1].h);
  });
}

// The score map of the last sweep batch's first candidate (nx × ny into dims).
int32_t zc_sweep_map(void* p, float* map, size_t len, uint32_t* dims) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "sweep_map", [&] { e->read_sweep_map(map, len, dims); });
}

// Candidates per sweep batch (0: tuned per plane size up to kSweepBatch when fft_autotune is on).
void zc_sweep_tune(uint32_t batch) { zc::SweepTune::batch = (int)batch; }

// Full-resolution maps at integer-translation poses from now on: 1 (default) through the
// translation plans (both weight planes rectangles: 16 of 45 correlations by FFT), 0 through the
// rectangle plans (30 by FFT). The two agree within FP32 rounding; process-wide.
void zc_map_path(uint32_t translation_plans) { zc::MapPath::trans = translation_plans > 2 ? 1 : (int)translation_plans; }


// ── reference entry points outside the SMI hot path: present so a ctypes binding of the
//    reference's full symbol list loads, each failing with "Unsupported" ──
#define ZC_UNSUPPORTED(name, ...)                                  \
  int32_t name(__VA_ARGS__) { return unsupported((Engine*)e, #name); }

ZC_UNSUPPORTED(zc_compile_wgsl, void* e, const char*, size_t, const char*, size_t, char*, size_t)
ZC_UNSUPPORTED(zc_pg_info, void* e, double*)
ZC_UNSUPPORTED(zc_pg_maps, void* e, uint32_t, float*, size_t)
ZC_UNSUPPORTED(zc_pg_descriptors, void* e, uint32_t, float*, size_t)
ZC_UNSUPPORTED(zc_bench_moments, void* e, uint32_t, uint32_t)
ZC_UNSUPPORTED(zc_match_scores, void* e, float*, size_t)
ZC_UNSUPPORTED(zc_detect, void* e, void*)
ZC_UNSUPPORTED(zc_match, void* e, void*)
ZC_UNSUPPORTED(zc_auto, void* e, void*)
ZC_UNSUPPORTED(zc_keypoints, void* e, uint32_t, float*, size_t)
ZC_UNSUPPORTED(zc_keypoint_frames, void* e, uint32_t, float*, size_t)
ZC_UNSUPPORTED(zc_matches, void* e, uint32_t*, size_t)
ZC_UNSUPPORTED(zc_set_ffd, void* e, const float*, size_t)
ZC_UNSUPPORTED(zc_ffd_gradient, void* e, const double*, double*, double*, size_t)
ZC_UNSUPPORTED(zc_canvas, void* e, const double*, uint32_t, void*)
ZC_UNSUPPORTED(zc_overlay, void* e, const double*, uint32_t, void*, float*, float*, size_t)
ZC_UNSUPPORTED(zc_tile_heat, void* e, const double*, uint32_t, uint32_t, double*, float*, float*, size_t)
ZC_UNSUPPORTED(zc_displacement, void* e, float*, size_t)

int64_t zc_export_warp(void* e, const char*, size_t, void*, size_t) { return unsupported((Engine*)e, "zc_export_warp"); }
size_t zc_get_ffd(void* e, float*, size_t) {
  (void)e;
  return 0;
}
uint32_t zc_fit_check(void* e, const double*) {
  (void)e;
  return 7;   // "it fails the sanity check": keypoint fits are not part of this library
}
void zc_tune_parts(uint32_t, uint32_t, uint32_t, uint32_t) {}
// exports.zig:95: tilings of the batched FFTs for rows / columns (0: default), for FFTs run from
// now on. A column value a length cannot take makes zc_fft2 fail with UnsupportedFftLength; maps
// and sweeps then use the default tiling. A row value above 4 is the default.
void zc_fft_tune(uint32_t rows, uint32_t cols) {
  zc::FftRowTiling::v = rows <= 4 ? (int)rows : 0;
  zc::FftTune::cols = (int)cols;
}
void zc_fft_tune_wg(uint32_t, uint32_t, uint32_t) {}

// The column-FFT tiling the engine uses for column length L (autotuning it now if needed).
int32_t zc_fft_tiling(void* p, uint32_t L, uint32_t* cols) {
  Engine* e = static_cast<Engine*>(p);
  return guard(e, "fft_tiling", [&] {
    if (!zc::fft_length_ok((int)L)) throw std::runtime_error("FftLength: a power of two 32..4096, 1536, 2304 or 3072");
    *cols = (uint32_t)e->col_lines((int)L);
  });
}

}  // extern "C"
