// Engine implementation and the C ABI (extern "C" zc_* entry points).
//
// Every entry point returns 0 (or a size) on success and a negative value on failure, with the
// reason kept for zc_error (exports.zig:1-16). No C++ exception crosses the ABI.
#include "engine.hpp"
#include "plan.hpp"
#include "pre.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>

namespace zc {


// ── profiling ─────────────────────────────────────────────────────────────────────────────
hipEvent_t Prof::get() {
  if (pool_.empty()) {
    hipEvent_t e;
    ZC_HIP(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = pool_.back();
  pool_.pop_back();
  return e;
}
void Prof::begin(hipStream_t st) {
  if (last_) {
    cur_ = last_;
    return;
  }
  cur_ = get();
  ZC_HIP(hipEventRecord(cur_, st));
}
void Prof::end(const char* name, hipStream_t st) {
  hipEvent_t b = get();
  ZC_HIP(hipEventRecord(b, st));
  pending_.push_back({name, cur_, b});
  cur_ = nullptr;
  last_ = b;
}
void Prof::collect() {
  // an event may end one launch and start the next: return each to the pool once
  std::set<hipEvent_t> seen;
  for (auto& p : pending_) {
    float ms = 0;
    ZC_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& a = acc_[p.name];
    a.first += 1;
    a.second += ms;
    for (hipEvent_t e : {p.a, p.b})
      if (e != last_ && seen.insert(e).second) pool_.push_back(e);
  }
  pending_.clear();
}
std::string Prof::report() const {
  std::ostringstream o;
  for (const auto& kv : acc_) {
    char line[256];
    const auto b = bytes_.find(kv.first);
    const auto u = work_.find(kv.first);
    std::snprintf(line, sizeof line, "%s %ld %.6f %.6f %.0f %.0f\n", kv.first.c_str(), kv.second.first,
                  kv.second.second, kv.second.second / std::max(1L, kv.second.first),
                  b == bytes_.end() ? 0.0 : b->second, u == work_.end() ? 0.0 : u->second);
    o << line;
  }
  return o.str();
}
Prof::~Prof() {
  for (auto e : pool_) (void)hipEventDestroy(e);
  for (auto& p : pending_) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
}


// ── engine ────────────────────────────────────────────────────────────────────────────────
Engine::Engine(int device) : dev(device) {
  int n = 0;
  ZC_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw std::runtime_error("no HIP device " + std::to_string(device));
  ZC_HIP(hipSetDevice(dev));
  ZC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  ZC_HIP(hipStreamCreateWithFlags(&st2, hipStreamNonBlocking));
  ZC_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  ZC_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  hipDeviceProp_t p;
  ZC_HIP(hipGetDeviceProperties(&p, dev));
  device_name = std::string(p.name) + " (" + p.gcnArchName + ")";
  // [0, 64): map and roto-scale peaks; [64, 256): the device climb's state
  ZC_HIP(hipHostMalloc((void**)&res_host, 256, hipHostMallocDefault));
}

Engine::~Engine() {
  (void)hipSetDevice(dev);
  if (st) (void)hipStreamSynchronize(st);
  if (st2) (void)hipStreamSynchronize(st2);
  // the captured graphs first: their launches name the buffers below
  sweep_gc.reset();
  gn_gc.reset();
  climb_gc.reset();
  for (auto& s : sides) release_side(s);
  for (auto& s : stack) release_side(*s);
  for (DBuf* b : {&canvas_a, &canvas_b, &spec_a, &spec_b, &work, &corr, &map, &arg_part, &res, &parts, &sums,
                  &tmp_u32, &tmp_i64, &sw_poses, &sw_pose_fixed, &sw_zero, &sw_dead, &sw_ns, &sw_part, &sw_peaks, &gn_jobs, &gn_part,
                  &gn_sums, &sat_band, &sat, &rs_fwd, &rs_inv, &rs_out, &rs_tab, &pre_buf, &pre_hist, &pre_cdf,
                  &pre_part, &pre_rgba, &map_out, &gd_pairs, &gd_state, &gd_gjobs, &gd_gsums, &gd_tjobs, &gd_tsums,
                  &gd_dir, &gd_step, &gd_try_H, &gd_try_ok, &gd_try_jobs, &gd_try_sums, &gd_try_score, &gd_part, &gd_ctr,
                  &cl_state, &cells_buf})
    b->release();
  for (auto& kv : tw) kv.second.release();
  if (res_host) (void)hipHostFree(res_host);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (st2) (void)hipStreamDestroy(st2);
  if (st) (void)hipStreamDestroy(st);
}

void Engine::release_side(Side& s) {
  for (DBuf* b : {&s.raw, &s.gray, &s.rank, &s.feat, &s.mag, &s.stack, &s.il, &s.tab}) b->release();
  s.has = false;
  s.il_ok = false;
  s.tab_ok = false;
}

// The image's own summed-area tables of its 15 stack planes (translation plans' rectangle sums).
void Engine::ensure_tables(const Side& sd) {
  if (sd.tab_ok) return;
  sd.tab.ensure((size_t)sat_pairs(kImageTableCount) * (sd.w + 1) * (sd.h + 1) * 16);
  sat_band.ensure((size_t)kImageTableCount * sat_bands(sd.h) * sd.w * 8);
  bool ok = true;
  LAUNCH("image_tables", ok = launch_sat_build(0, 2, sd.stack.as<float>(), sd.w, sd.h, sat_band.as<double>(),
                                               sd.tab.as<double2>(), st));
  if (!ok) throw std::runtime_error("UnsupportedImageSize: image tables");
  sd.tab_ok = true;
}

void Engine::GraphCache::reset() {
  if (exec) (void)hipGraphExecDestroy(exec);
  exec = nullptr;
  key.clear();
  kernels = 0;
}

void Engine::run_captured(GraphCache& gc, std::vector<uint8_t>&& key, bool eager, const std::function<void()>& issue) {
  if (eager) {
    issue();
    return;
  }
  if (!gc.exec || gc.key != key) {
    gc.reset();
    const uint32_t c0 = counts[0];
    ZC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      issue();
    } catch (...) {
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      counts[0] = c0;
      throw;
    }
    ZC_HIP(hipStreamEndCapture(st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    ZC_HIP(e);
    gc.exec = ex;
    gc.key = std::move(key);
    gc.kernels = counts[0] - c0;
    counts[0] = c0;   // capturing ran nothing
  }
  ZC_HIP(hipGraphLaunch(gc.exec, st));
  counts[0] += gc.kernels;
  prof.break_chain();
}

void Engine::sync() {
  ZC_HIP(hipStreamSynchronize(st));
  ++counts[3];
  if (prof.on) prof.collect();
  prof.break_chain();
}

const float2* Engine::twiddles(int L) {
  auto it = tw.find(L);
  if (it != tw.end()) return it->second.as<float2>();
  std::vector<float2> h(L);
  for (int t = 0; t < L; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)L;
    h[t] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  DBuf& b = tw[L];
  b.ensure(sizeof(float2) * L);
  ZC_HIP(hipMemcpy(b.p, h.data(), sizeof(float2) * L, hipMemcpyHostToDevice));
  return b.as<float2>();
}

int Engine::col_lines(int L) {
  const int forced = FftTune::cols;
  if (forced) return forced;
  if (!set.b("fft_autotune")) return 0;
  static std::mutex mu;
  static std::map<int, int> best;
  {
    std::lock_guard<std::mutex> lk(mu);
    const auto it = best.find(L);
    if (it != best.end()) return it->second;
  }
  const int c = autotune_cols(L);
  std::lock_guard<std::mutex> lk(mu);
  best[L] = c;
  return c;
}

// Times every column tiling that launches at length L (plain 1–16 columns per workgroup, two
// groups sharing the LDS, streaming over planes, the wide tilings with 32 points per thread or a
// split exchange) on square planes of ≥ 64 MB, three runs each
// after a warm-up, and returns the fastest (0 keeps the built-in default when none is faster).
// Every tiling computes each column with the same stages, so the choice never changes results.
int Engine::autotune_cols(int L) {
  const int nx = L;
  const size_t plane = (size_t)nx * L * 8;
  const int planes = (int)std::min<size_t>(16, std::max<size_t>(1, ((size_t)64 << 20) / plane));
  tune_buf.ensure(plane * planes);
  ZC_HIP(hipMemsetAsync(tune_buf.p, 0, plane * planes, st));
  prof.break_chain();
  const float2* tw = twiddles(L);
  hipEvent_t e0, e1;
  ZC_HIP(hipEventCreate(&e0));
  ZC_HIP(hipEventCreate(&e1));
  auto time = [&](int lines) -> double {
    if (!launch_colfft(tune_buf.as<float2>(), planes, nx, L, L, true, tw, st, lines)) return INFINITY;
    float best = INFINITY;
    for (int r = 0; r < 3; ++r) {
      ZC_HIP(hipEventRecord(e0, st));
      launch_colfft(tune_buf.as<float2>(), planes, nx, L, L, true, tw, st, lines);
      ZC_HIP(hipEventRecord(e1, st));
      ZC_HIP(hipEventSynchronize(e1));
      float ms = 0;
      ZC_HIP(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    return best;
  };
  double t_best = time(0);
  int c_best = 0;
  for (int lines : {1, 2, 4, 8, 16, 9, 17, 33, 68, 72, 80, 41, 42, 43, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53}) {
    const double t = time(lines);
    if (t < 0.97 * t_best) {
      t_best = t;
      c_best = lines;
    }
  }
  ZC_HIP(hipEventDestroy(e0));
  ZC_HIP(hipEventDestroy(e1));
  tune_buf.release();
  prof.break_chain();
  return c_best;
}

void Engine::need_images() {
  if (!has_images()) throw std::runtime_error("NoImages");
  for (Side& sd : sides)
    if (sd.baked != bake_key(sd)) bake(sd);
}

std::string Engine::bake_key(const Side& sd) const {
  std::string k = set.e("smi_features") + (set.b("half") ? "|h" : "|f");
  if (sd.preprocess) {
    char b[160];
    std::snprintf(b, sizeof b, "|c%d,%u,%u|b%d,%.17g,%.17g|i%d", set.b("clahe") ? 1 : 0, set.u("clahe_grid"),
                  set.u("clahe_bins"), set.b("band") ? 1 : 0, set.f("bp_fine"), set.f("bp_coarse"),
                  set.b("invert") ? 1 : 0);
    k += b;
  }
  return k;
}

void Engine::bake(Side& sd) {
  const int w = sd.w, h = sd.h;
  const int64_t n = (int64_t)w * h;
  const bool half = set.b("half");
  float* raw = sd.raw.as<float>();
  float* work = sd.gray.as<float>();
  if (!sd.preprocess) {
    LAUNCH("store", launch_store(raw, work, n, half, st));
  } else {
    // five scratch planes: the work image in progress, fine, coarse and two for the passes
    pre_buf.ensure((size_t)n * 4 * 5);
    float* A = pre_buf.as<float>();
    float* Bf = A + n;
    float* Dc = Bf + n;
    float* P0 = Dc + n;
    float* T = P0 + n;
    if (set.b("clahe")) {
      const ClaheArgs a = clahe_args(w, h, 40.0, set.u("clahe_bins"), set.u("clahe_grid"));
      pre_hist.ensure((size_t)a.hg * a.hg * a.nb * 4);
      pre_cdf.ensure((size_t)a.hg * a.hg * a.nb * 4);
      LAUNCH("clahe", launch_clahe(raw, a, pre_hist.as<uint32_t>(), pre_cdf.as<float>(), A, half, st));
    } else {
      LAUNCH("store", launch_store(raw, A, n, half, st));
    }
    if (set.b("band")) {
      // gls.zig:519-545: fine = fine_n Gaussian passes, coarse = coarse_n − fine_n more on top
      auto rnd = [](double v) { return std::floor(v + 0.5); };
      const double fp = set.f("bp_fine"), cp = set.f("bp_coarse");
      const int fine_n = (int)std::max(1.0, std::min(16.0, rnd(std::pow(fp > 0 ? fp : 1.0, 2.0))));
      const int coarse_n = std::max(fine_n + 1, (int)std::min(64.0, rnd(std::pow(cp > 0 ? cp : 6.0, 2.0))));
      // passes of (x, y) Gaussian passes from src to dst, intermediates alternating p / q
      auto blur = [&](const float* src, float* dst, int passes, float* p, float* q) {
        const float* from = src;
        for (int i = 0; i < passes; ++i) {
          float* to = i == passes - 1 ? dst : (i % 2 == 0 ? p : q);
          LAUNCH("gauss", launch_gauss(from, T, w, h, 1, half, st));
          LAUNCH("gauss", launch_gauss(T, to, w, h, 0, half, st));
          from = to;
        }
      };
      blur(A, Bf, fine_n, P0, Dc);
      blur(Bf, Dc, coarse_n - fine_n, P0, A);   // A holds nothing needed any more
      pre_part.ensure((size_t)band_parts(n) * 8);
      LAUNCH("band", launch_band_finish(Bf, Dc, P0, pre_part.as<float>(), A, n, half, st));
    }
    if (set.b("invert")) LAUNCH("invert", launch_invert(A, n, half, st));
    ZC_HIP(hipMemcpyAsync(work, A, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    prof.break_chain();
  }
  sd.il_ok = false;
  sd.tab_ok = false;
  prepare(sd, true);
  sd.baked = bake_key(sd);
}

void Engine::set_image(int s, const float* data, int w, int h, bool preprocess) {
  if (s < 0 || s > 1 || w < 8 || h < 8 || !data || (int64_t)w * h > kMaxImagePixels) throw std::runtime_error("BadImage");
  Side& sd = sides[s];
  const size_t n = (size_t)w * h;
  sd.w = w;
  sd.h = h;
  sd.has = false;
  sd.il_ok = false;
  sd.tab_ok = false;
  sd.preprocess = preprocess;
  sd.raw.ensure(n * 4);
  sd.gray.ensure(n * 4);
  sd.rank.ensure(n * 4);
  sd.feat.ensure(n * 16);
  sd.mag.ensure(n * 4);
  sd.stack.ensure(n * 20);
  // host or device memory (the caller finishes any device work that writes `data` first)
  ZC_HIP(hipMemcpyAsync(sd.raw.p, data, n * 4, hipMemcpyDefault, st));
  prof.break_chain();
  bake(sd);
  sd.has = true;
}

// exports.zig:60-63 / engine.zig:422-427: RGBA8 → gray (0.299 R + 0.587 G + 0.114 B), then the
// work image always preprocessed per the settings.
void Engine::set_image_rgba(int s, const uint8_t* rgba, int w, int h) {
  if (s < 0 || s > 1 || w < 8 || h < 8 || !rgba || (int64_t)w * h > kMaxImagePixels)
    throw std::runtime_error("BadImage");
  const size_t n = (size_t)w * h;
  pre_rgba.ensure(n * 4);
  ZC_HIP(hipMemcpyAsync(pre_rgba.p, rgba, n * 4, hipMemcpyHostToDevice, st));
  prof.break_chain();
  Side& sd = sides[s];
  sd.raw.ensure(n * 4);
  LAUNCH("rgba_gray", launch_rgba_gray(pre_rgba.as<uint32_t>(), sd.raw.as<float>(), (int64_t)n, st));
  sd.w = w;
  sd.h = h;
  sd.has = false;
  sd.il_ok = false;
  sd.tab_ok = false;
  sd.preprocess = true;
  sd.gray.ensure(n * 4);
  sd.rank.ensure(n * 4);
  sd.feat.ensure(n * 16);
  sd.mag.ensure(n * 4);
  sd.stack.ensure(n * 20);
  bake(sd);
  sd.has = true;
}

// Smi.prepare (smi.zig:348-400) for SMI features, with exact image statistics (see the oracle's
// definition: order-independent integer sums, so the rank bins are reproducible bit for bit).
void Engine::prepare(Side& sd, bool acf) {
  const int64_t n = (int64_t)sd.w * sd.h;
  float* gray = sd.gray.as<float>();
  sd.half = set.b("half");
  sd.mode = set.e("smi_features");
  if (sd.half) LAUNCH("round_half", launch_round_half(gray, n, st));
  // max |x| → the fixed-point scale
  tmp_u32.ensure(4 * 1024);
  tmp_i64.ensure(8 * 4);
  uint32_t* acc32 = tmp_u32.as<uint32_t>();
  ZC_HIP(hipMemsetAsync(acc32, 0, 4, st));
  prof.break_chain();
  LAUNCH("maxabs", launch_maxabs(gray, n, acc32, st));
  uint32_t maxbits = 0;
  ZC_HIP(hipMemcpyAsync(&maxbits, acc32, 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  sync();
  float maxabs;
  std::memcpy(&maxabs, &maxbits, 4);
  int e = 0;
  if (maxabs > 0) std::frexp((double)maxabs, &e);
  auto fixed_sum = [&](int squared, float mean, int K) -> double {
    long long* hi = tmp_i64.as<long long>();
    unsigned long long* lo = reinterpret_cast<unsigned long long*>(hi + 1);
    ZC_HIP(hipMemsetAsync(hi, 0, 16, st));
    prof.break_chain();
    LAUNCH("fixsum", launch_fixsum(gray, n, std::ldexp(1.0, K), mean, squared, hi, lo, st));
    long long h2[2];
    ZC_HIP(hipMemcpyAsync(h2, hi, 16, hipMemcpyDeviceToHost, st));
    prof.break_chain();
    sync();
    const __int128 S = (__int128)h2[0] * ((__int128)1 << 32) + (__int128)(unsigned long long)h2[1];
    const long long shi = (long long)(S >> 64);
    const unsigned long long slo = (unsigned long long)S;
    const double v = ((double)shi * 18446744073709551616.0 + (double)slo) / (double)n;
    return std::ldexp(v, -K);
  };
  const int K = 62 - e;
  sd.mean = (float)fixed_sum(0, 0.0f, K);
  const double var = fixed_sum(1, sd.mean, 60 - 2 * e);
  sd.spread = (float)std::sqrt(std::max(var, 1e-12));
  // histogram → CDF (host, exact) → rank
  uint32_t* hist = acc32 + 16;
  uint32_t* cdf = acc32 + 16 + 272;
  ZC_HIP(hipMemsetAsync(hist, 0, 257 * 4, st));
  prof.break_chain();
  LAUNCH("hist", launch_hist(gray, n, sd.mean, sd.spread, hist, st));
  ZC_HIP(hipMemcpyAsync(sd.hist, hist, 257 * 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  sync();
  uint64_t accu = 0;
  for (int i = 0; i < 256; ++i) {
    sd.cdf[i] = (uint32_t)accu;
    accu += sd.hist[i];
  }
  sd.cdf[256] = (uint32_t)accu;
  sd.cdf[257] = 0;
  ZC_HIP(hipMemcpyAsync(cdf, sd.cdf, 258 * 4, hipMemcpyHostToDevice, st));
  prof.break_chain();
  LAUNCH("rank", launch_rank(gray, n, sd.mean, sd.spread, hist, cdf, sd.rank.as<float>(), st));
  // features and their moments
  const int mode = sd.mode == "laplacian" ? 0 : (sd.mode == "three" ? 1 : 2);
  const int nb = feats_blocks(sd.w, sd.h);
  parts.ensure((size_t)nb * 14 * 8);
  sums.ensure(2048 * 8);
  LAUNCH("feats", launch_feats(sd.rank.as<float>(), sd.w, sd.h, mode, sd.feat.as<float>(), sd.mag.as<float>(),
                               parts.as<double>(), st));
  LAUNCH("sum_parts", launch_sum_parts(parts.as<double>(), nb, 14, sums.as<double>(), st));
  double m[14];
  ZC_HIP(hipMemcpyAsync(m, sums.p, 14 * 8, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  sync();
  // global ridge whitening (smi.wgsl:554-599), FP64
  const double nn = (double)n;
  double mu[4], G[4][4];
  for (int i = 0; i < 4; ++i) mu[i] = m[i] / nn;
  auto tk = [](int i, int j) {
    const int a = std::min(i, j), b = std::max(i, j);
    return a * 4 - (a * (a + 1)) / 2 + b;
  };
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) G[i][j] = m[4 + tk(i, j)] / nn - mu[i] * mu[j];
  const double tr = G[0][0] + G[1][1] + G[2][2] + G[3][3];
  const double ridge = kRidge * std::max(tr * 0.25, 1e-12);
  for (int i = 0; i < 4; ++i) G[i][i] += ridge;
  double L[4][4] = {};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j <= i; ++j) {
      double sacc = G[i][j];
      for (int k = 0; k < j; ++k) sacc -= L[i][k] * L[j][k];
      if (i == j) L[i][i] = std::sqrt(std::max(sacc, 1e-12));
      else L[i][j] = sacc / L[j][j];
    }
  WhitenArgs wa{};
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < 4; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) v -= L[i][k] * wa.X[k * 4 + c];
      wa.X[i * 4 + c] = v / L[i][i];
    }
  for (int i = 0; i < 4; ++i) wa.mu[i] = mu[i];
  LAUNCH("whiten", launch_whiten(sd.feat.as<float>(), sd.stack.as<float>(), n, wa, st));
  if (acf) ensure_acf(sd);
}

// Normalized feature autocorrelation (smi.wgsl:936-975, 1574-1580).
void Engine::ensure_acf(Side& sd) {
  const int nb = acf_parts(sd.w, sd.h);
  parts.ensure((size_t)nb * 1156 * 8);
  sums.ensure(2048 * 8);
  LAUNCH("acf", launch_acf(sd.feat.as<float>(), sd.w, sd.h, parts.as<double>(), st));
  LAUNCH("sum_parts", launch_sum_parts(parts.as<double>(), nb, 1156, sums.as<double>(), st));
  std::vector<double> raw(1156);
  ZC_HIP(hipMemcpyAsync(raw.data(), sums.p, 1156 * 8, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  sync();
  for (int lag = 0; lag < 289; ++lag) {
    const int dx = lag % 17 - 8, dy = lag / 17 - 8;
    const double cnt = std::max(1.0, (double)(sd.w - std::abs(dx)) * (double)(sd.h - std::abs(dy)));
    for (int p = 0; p < 4; ++p) sd.acf[lag * 4 + p] = raw[lag * 4 + p] / cnt;
  }
  double zero[4];
  for (int p = 0; p < 4; ++p) zero[p] = sd.acf[144 * 4 + p];
  for (int lag = 0; lag < 289; ++lag)
    for (int p = 0; p < 4; ++p) sd.acf[lag * 4 + p] /= std::max(zero[p], 1e-12);
}

// Bartlett correlation area of the pair, px² (smi.wgsl:1582-1595).
double Engine::pair_area(const Side& A, const Side& B) const {
  double s = 0;
  for (int k = 0; k < 289; ++k) {
    double ra = 0, rb = 0;
    for (int i = 0; i < 4; ++i) {
      ra += A.acf[k * 4 + i];
      rb += B.acf[k * 4 + i];
    }
    s += ra * rb;
  }
  return std::max(1.0, s / 16.0);
}

// The dense shift map of pose H (smi.zig:540-572; engine.zig:795-819).
MapOut Engine::shift_map(const Mat3& Hp, bool exact, bool calib, uint32_t cap, double min_overlap, bool wait) {
  need_images();
  return shift_map_sides(sides[0], sides[1], Hp, exact, calib, cap, min_overlap, wait);
}

MapOut Engine::shift_map_sides(Side& A, Side& B, const Mat3& Hp, bool exact, bool calib, uint32_t cap,
                               double min_overlap, bool wait, const Frame* frame, const double* dev_hi) {
  if (A.mode != B.mode) throw std::runtime_error("FeatureModeMismatch: set both images with the same smi_features");
  const bool three = A.mode == "three";
  MapOut m{};
  m.exact = exact;
  m.calib = exact && calib;
  m.fr = frame ? *frame : reg_frame(Hp, A.w, A.h, B.w, B.h);
  const Mat3 Hc = canvas_h(Hp, m.fr.ox, m.fr.oy);
  const bool ladder = set.e("fft_sizes") == "ladder";
  if (cap == 0) {
    m.g = auto_shift_plan(m.fr.w, m.fr.h);
  } else {
    const uint32_t capn = largest_smooth(std::min<uint32_t>(cap, kMaxFft), ladder);
    m.g = linear_corr_plan_rect(m.fr.w, m.fr.h, capn, ladder);
  }
  // the kernels take power-of-two lengths: a longer FFT only pads (a linear correlation's
  // values per lag do not depend on the length once it is >= 2 cells - 1); the map returned
  // through the ABI is laid out on the reference plan's grid (read_map, fill_shift_result)
  m.gr = m.g;
  m.g.nx = pow2_ceil(m.g.nx);
  m.g.ny = pow2_ceil(m.g.ny);
  if (m.g.nx > (uint32_t)kMaxFft || m.g.ny > (uint32_t)kMaxFft) throw std::runtime_error("FftTooLong");
  const int W = m.fr.w, Hh = m.fr.h;
  const int nx = (int)m.g.nx, ny = (int)m.g.ny, cw = (int)m.g.cw, ch = (int)m.g.ch;
  // One canvas cell per pixel: the fixed weight plane is the indicator of the pasted rectangle, so
  // the correlations (a, 0) are rectangle sums (summed-area tables) instead of FFT correlations.
  const bool rect = cw == W && ch == Hh && W <= 2048;
  // An integer translation onto the canvas copies the moving stack exactly (every bilinear tap at a
  // pixel centre), so its weight plane is the indicator of the moving rectangle as well: the
  // correlations (0, b) become rectangle sums of the fixed side (translation plans).
  const bool int_shift = Hc[0] == 1.0 && Hc[1] == 0.0 && Hc[3] == 0.0 && Hc[4] == 1.0 && Hc[6] == 0.0 &&
                         Hc[7] == 0.0 && Hc[8] == 1.0 && Hc[2] == std::floor(Hc[2]) && Hc[5] == std::floor(Hc[5]) &&
                         std::fabs(Hc[2]) < 1e6 && std::fabs(Hc[5]) < 1e6;
  const bool trans = rect && int_shift && MapPath::trans != 0 && !dev_hi;
  const int plan = (exact ? 1 : 0) + (three ? 2 : 0) + (rect ? 4 : 0) + (trans ? 4 : 0);
  const CorrPlan& pi = kPlans[plan];
  const int64_t N = (int64_t)nx * ny;
  const size_t cn = (size_t)W * Hh;
  canvas_a.ensure(cn * 20);
  canvas_b.ensure(cn * 20);
  // both sides' spectra in one buffer (side B's planes right after side A's): one forward column
  // pass over all of them
  spec_a.ensure((size_t)N * 8 * (pi.n_cplx + pi.b_n_cplx));
  float2* const spa = spec_a.as<float2>();
  float2* const spb = spa + N * pi.n_cplx;
  if (rect && !trans) {
    sat_band.ensure((size_t)pi.n_sat * sat_bands(Hh) * W * 8);
    sat.ensure((size_t)sat_pairs(pi.n_sat) * (W + 1) * (Hh + 1) * 16);
  }
  // translation plans read each image's own tables (made once per image)
  if (trans) {
    ensure_tables(A);
    ensure_tables(B);
  }
  work.ensure((size_t)N * 8 * pi.n_out);
  // rectangle plans: correlation quads (k_rowifft_quads), else one plane per correlation
  corr.ensure((size_t)N * 4 * (rect ? 4 * ((pi.n_corr + 3) / 4) : pi.n_corr));
  map.ensure((size_t)N * 4);
  const int nparts = combine_blocks(N);
  arg_part.ensure((size_t)nparts * 8);
  res.ensure(16);
  // canvas stacks
  double Hi[9];
  warp_inverse(Hc, Hi);
  const float2* y_tw = twiddles(ny);
  // a translation plan's canvases are the images moved by integers: the pack reads the images in
  // place, so neither canvas is written
  if (!trans) {
    if (dev_hi)
      LAUNCH("warp_stack", launch_warp_stack_dev(A.stack.as<float>(), A.w, A.h, dev_hi, canvas_a.as<float>(), W, Hh, st));
    else
      LAUNCH("warp_stack", launch_warp_stack(A.stack.as<float>(), A.w, A.h, Hi, canvas_a.as<float>(), W, Hh, st));
    LAUNCH("paste_stack", launch_paste_stack(B.stack.as<float>(), B.w, B.h, m.fr.ox, m.fr.oy, canvas_b.as<float>(), W, Hh, st));
  }
  bool ok = true;
  // A rectangle plan's tables depend only on canvas A: unless launches are profiled (one event
  // chain on one stream), they are built on the second stream while the FFT stages run.
  const bool fork = rect && !trans && !prof.on && !dev_hi;
  hipStream_t sst = fork ? st2 : st;
  if (fork) {
    ZC_HIP(hipEventRecord(ev_fork, st));
    ZC_HIP(hipStreamWaitEvent(st2, ev_fork, 0));
  }
  SatView sv{}, svb{};
  int lx = 0;
  while ((1 << lx) < nx) ++lx;
  // the fixed image pasted at (-ox, -oy) (paste_stack): side A's sums run over this rectangle
  const int fx0 = std::max(0, -m.fr.ox), fx1 = std::min(W, -m.fr.ox + B.w);
  const int fy0 = std::max(0, -m.fr.oy), fy1 = std::min(Hh, -m.fr.oy + B.h);
  if (rect && !trans) {
    LAUNCH("sat_build", ok &= launch_sat_build(plan, 0, canvas_a.as<float>(), W, Hh, sat_band.as<double>(), sat.as<double2>(), sst));
    sv = {sat.as<double2>(), nx, ny, lx, fx0, fx1, fy0, fy1, 1, 0, W, 0, Hh, W + 1, (uint32_t)((W + 1) * (Hh + 1))};
  }
  if (trans) {
    // Side A's sums: the fixed rectangle moved by the lag, over the moving image's own tables (the
    // image is copied to canvas cells [tx, tx + w) x [ty, ty + h), so table cell = canvas cell −
    // (tx, ty)), clipped to the image and to the canvas.
    const int tx = (int)Hc[2], ty = (int)Hc[5];
    sv = {A.tab.as<double2>(), nx, ny, lx, fx0 - tx, fx1 - tx, fy0 - ty, fy1 - ty, 1,
          std::max(0, -tx), std::min(A.w, W - tx), std::max(0, -ty), std::min(A.h, Hh - ty), A.w + 1,
          (uint32_t)((A.w + 1) * (A.h + 1))};
    // Side B's sums: the moving rectangle (clipped to the canvas) moved against the lag, over the
    // fixed image's own tables (the image at (-ox, -oy)).
    const int mx0 = std::min(std::max(0, tx), W), mx1 = std::min(std::max(0, tx + A.w), W);
    const int my0 = std::min(std::max(0, ty), Hh), my1 = std::min(std::max(0, ty + A.h), Hh);
    const int bx = -m.fr.ox, by = -m.fr.oy;
    svb = {B.tab.as<double2>(), nx, ny, lx, mx0 - bx, mx1 - bx, my0 - by, my1 - by, -1,
           std::max(0, -bx), std::min(B.w, W - bx), std::max(0, -by), std::min(B.h, Hh - by), B.w + 1,
           (uint32_t)((B.w + 1) * (B.h + 1))};
  }
  if (fork) ZC_HIP(hipEventRecord(ev_join, st2));
  float* cells = nullptr;
  if (!rect) {
    cells_buf.ensure((size_t)cw * ch * 5 * 4);
    cells = cells_buf.as<float>();
  }
  if (trans) {
    LAUNCH("pack_rowfft", ok &= launch_pack_rowfft_shifted(plan, 0, A.stack.as<float>(), A.w, A.h, (int)Hc[2], (int)Hc[5],
                                                           cw, ch, nx, ny, spa, twiddles(nx), st));
    LAUNCH("pack_rowfft", ok &= launch_pack_rowfft_shifted(plan, 1, B.stack.as<float>(), B.w, B.h, -m.fr.ox, -m.fr.oy,
                                                           cw, ch, nx, ny, spb, twiddles(nx), st));
  } else {
    LAUNCH("pack_rowfft", ok &= launch_pack_rowfft(plan, 0, canvas_a.as<float>(), W, Hh, cw, ch, nx, ny, spa, twiddles(nx), st, cells));
    LAUNCH("pack_rowfft", ok &= launch_pack_rowfft(plan, 1, canvas_b.as<float>(), W, Hh, cw, ch, nx, ny, spb, twiddles(nx), st, cells));
  }
  LAUNCH("colfft_fwd", ok &= launch_colfft(spa, pi.n_cplx + pi.b_n_cplx, nx, ny, ch, false, y_tw, st, col_lines(ny)));
  LAUNCH("cmul", launch_cmul(plan, spa, spb, work.as<float2>(), nx, ny, st));
  LAUNCH("colfft_inv", ok &= launch_colfft(work.as<float2>(), pi.n_out, nx, ny, ny, true, y_tw, st, col_lines(ny)));
  if (rect)
    LAUNCH("rowifft_extract", ok &= launch_rowifft_quads(plan, work.as<float2>(), corr.as<float4>(), nx, ny, twiddles(nx), st));
  else
    LAUNCH("rowifft_extract", ok &= launch_rowifft_extract(plan, work.as<float2>(), corr.as<float>(), nx, ny, twiddles(nx), st));
  if (!ok) throw std::runtime_error("UnsupportedFftLength");
  const uint32_t floor_n = min_overlap > 0 ? (uint32_t)std::floor(min_overlap * (double)cw * (double)ch) : 0u;
  CombineArgs ca{};
  ca.exact = exact;
  ca.calib = m.calib;
  ca.nf = three ? 3 : 4;
  ca.min_n = (float)std::max<uint32_t>(kMinN, floor_n);
  ca.ridge = kRidge;
  const double gcell = (double)W / (double)cw;
  ca.area = std::max(1.0, pair_area(A, B) / (gcell * gcell));
  if (fork) ZC_HIP(hipStreamWaitEvent(st, ev_join, 0));
  LAUNCH("combine", launch_combine(plan, corr.as<float>(), map.as<float>(), N, ca, arg_part.as<float>(), st, rect ? &sv : nullptr, trans ? &svb : nullptr));
  LAUNCH("argmax_finish", launch_argmax_finish(arg_part.as<float>(), nparts, map.as<float>(), res.as<float>(), st));
  ++counts[1];
  if (dev_hi) return m;
  ZC_HIP(hipMemcpyAsync(res_host, res.p, 12, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  if (wait) {
    sync();
    std::memcpy(m.res, res_host, 12);
  }
  return m;
}

void Engine::read_map(float* out, size_t len, const MapOut& m) {
  const size_t nn = (size_t)m.gr.nx * m.gr.ny;
  if (len < nn) throw std::runtime_error("MapTooSmall");
  const void* src = map.p;
  if (m.gr.nx != m.g.nx || m.gr.ny != m.g.ny) {
    // every lag of the reference grid, read from the (longer) grid the kernels ran on
    map_out.ensure(nn * 4);
    LAUNCH("map_remap", launch_map_remap(map.as<float>(), (int)m.g.nx, (int)m.g.ny, map_out.as<float>(),
                                         (int)m.gr.nx, (int)m.gr.ny, st));
    src = map_out.p;
  }
  ZC_HIP(hipMemcpyAsync(out, src, nn * 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
}

void Engine::fill_shift_result(const MapOut& m, ShiftResult* r) const {
  std::memset(r, 0, sizeof *r);
  r->n = m.gr.nx;
  r->n_y = m.gr.ny;
  r->cw = m.g.cw;
  r->ch = m.g.ch;
  r->canvas_w = (uint32_t)m.fr.w;
  r->canvas_h = (uint32_t)m.fr.h;
  r->ox = m.fr.ox;
  r->oy = m.fr.oy;
  r->peak = m.res[0];
  r->zero = m.res[2];
  r->corr_area_px = pair_area();
  r->exact = m.exact;
  r->calibrated = m.calib;
  uint32_t idx;
  std::memcpy(&idx, &m.res[1], 4);
  const uint64_t nn = (uint64_t)m.g.nx * m.g.ny;
  if (idx < nn) {
    // peakLag (engine.zig:789-793)
    const int64_t py = idx / m.g.nx, px = idx % m.g.nx;
    const int64_t lx = px <= m.g.nx / 2 ? px : px - (int64_t)m.g.nx;
    const int64_t ly = py <= m.g.ny / 2 ? py : py - (int64_t)m.g.ny;
    // the peak's index on the reference grid (its lag wrapped to that grid's lengths)
    const int64_t jx = lx >= 0 ? lx : lx + (int64_t)m.gr.nx, jy = ly >= 0 ? ly : ly + (int64_t)m.gr.ny;
    r->peak_index = (uint32_t)(jy * m.gr.nx + jx);
    r->dx = (double)lx * ((double)m.fr.w / (double)m.g.cw);
    r->dy = (double)ly * ((double)m.fr.h / (double)m.g.ch);
  }
}

// The 45 overlap sums of side a's features onto side b's at H (a → b), smi.wgsl:980-1057.
void Engine::moments(int as, int bs, const Mat3& Hp, double out[45]) {
  need_images();
  const Side& A = sides[as];
  const Side& B = sides[bs];
  const Mat3 Hi = inv3(Hp);
  PoseArgs pa;
  for (int i = 0; i < 9; ++i) pa.Hi[i] = Hi[i];
  const int nb = moments_blocks(B.w, B.h);
  parts.ensure((size_t)nb * 45 * 8);
  sums.ensure(2048 * 8);
  LAUNCH("pose_moments", launch_pose_moments(A.feat.as<float>(), A.w, A.h, B.feat.as<float>(), B.w, B.h, pa,
                                             parts.as<double>(), st));
  LAUNCH("sum_parts", launch_sum_parts(parts.as<double>(), nb, 45, sums.as<double>(), st));
  ZC_HIP(hipMemcpyAsync(out, sums.p, 45 * 8, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
}

// moments.zig:91-134 scoreFromMoments (the score part), FP64.
double Engine::score_from_moments(const double* s, bool exact, double ridge) {
  const double n = s[0];
  if (!(n >= kMinN)) return 0.0;
  double ma[4], mb[4], C[4][4];
  for (int i = 0; i < 4; ++i) {
    ma[i] = s[1 + i] / n;
    mb[i] = s[5 + i] / n;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) C[i][j] = s[9 + i * 4 + j] / n - ma[i] * mb[j];
  if (!exact) {
    double phi = 0;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) phi += C[i][j] * C[i][j];
    const double sc = n * phi;
    return std::isfinite(sc) ? sc : 0.0;
  }
  auto tk = [](int i, int j) {
    const int a = std::min(i, j), b = std::max(i, j);
    return a * 4 - (a * (a + 1)) / 2 + b;
  };
  double GA[4][4], GB[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      GA[i][j] = s[25 + tk(i, j)] / n - ma[i] * ma[j];
      GB[i][j] = s[35 + tk(i, j)] / n - mb[i] * mb[j];
    }
  const double ra = ridge * std::max((GA[0][0] + GA[1][1] + GA[2][2] + GA[3][3]) / 4, 1e-12);
  const double rb = ridge * std::max((GB[0][0] + GB[1][1] + GB[2][2] + GB[3][3]) / 4, 1e-12);
  for (int i = 0; i < 4; ++i) {
    GA[i][i] += ra;
    GB[i][i] += rb;
  }
  auto chol = [](double (&G)[4][4], double (&L)[4][4]) {
    std::memset(L, 0, sizeof L);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j <= i; ++j) {
        double v = G[i][j];
        for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
        if (i == j) {
          if (!(v > 0)) return false;
          L[i][i] = std::sqrt(v);
        } else {
          L[i][j] = v / L[j][j];
        }
      }
    return true;
  };
  double LA[4][4], LB[4][4];
  if (!chol(GA, LA) || !chol(GB, LB)) return 0.0;
  // Φ = ‖L_A⁻¹ C L_B⁻ᵀ‖²_F = tr(G_A⁻¹ C G_B⁻¹ Cᵀ)
  double X[4][4];
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      double v = C[i][j];
      for (int k = 0; k < i; ++k) v -= LA[i][k] * X[k][j];
      X[i][j] = v / LA[i][i];
    }
  double phi = 0;
  for (int r = 0; r < 4; ++r) {
    double y[4];
    for (int i = 0; i < 4; ++i) {
      double v = X[r][i];
      for (int k = 0; k < i; ++k) v -= LB[i][k] * y[k];
      y[i] = v / LB[i][i];
      phi += y[i] * y[i];
    }
  }
  const double sc = n * phi;
  return std::isfinite(sc) ? sc : 0.0;
}

void Engine::fft2(float* data, int n, int planes, bool inverse) {
  if (!fft_length_ok(n)) throw std::runtime_error("FftLength: a power of two 32..4096, 1536, 2304 or 3072");
  const size_t bytes = (size_t)planes * n * n * 8;
  work.ensure(bytes);
  ZC_HIP(hipMemcpyAsync(work.p, data, bytes, hipMemcpyHostToDevice, st));
  prof.break_chain();
  bool ok = true;
  LAUNCH("fft2", ok = launch_fft2(work.as<float2>(), planes, n, inverse, twiddles(n), st, FftTune::cols));
  if (!ok) throw std::runtime_error("UnsupportedFftLength");
  ZC_HIP(hipMemcpyAsync(data, work.p, bytes, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  sync();
}

}  // namespace zc
