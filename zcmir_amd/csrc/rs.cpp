// Roto-scale (log-polar) maps: smi.zig:577-651 (rsMap, rsMapFrom, rsFinish) and engine.zig:836-879.
//
// Both canvas stacks are sampled on an n_th x n_lam log-polar grid about the centre (angles across,
// log radius down), every real plane is weighted by the radius, and the 45 / 28 correlations are
// circular in angle and linear in log radius on the reference's n x n grid. The HIP FFTs take
// power-of-two lengths: a power-of-two n is used as is; otherwise the angle axis is padded to a
// power of two >= 2n and the two linear lags of each circular lag are added in the score kernel,
// and the log-radius axis (n / 2 rows of data) is padded to a power of two >= n.
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "engine.hpp"
#include "plan.hpp"
#include "search.hpp"

namespace zc {

namespace {
constexpr uint32_t kRsMaxN = 2048;   // maxFftN (fft.zig:118-120): the longest roto-scale grid
}  // namespace

Engine::RsPlan Engine::rs_map_dir(Side& A, Side& B, const Mat3& Hp, double c0, double c1, double r0_opt,
                                  double r1_opt, bool exact, bool calib, uint32_t cap, DBuf& dst,
                                  const Frame* frame, const double* dev_hi) {
  if (A.mode != B.mode) throw std::runtime_error("FeatureModeMismatch: set both images with the same smi_features");
  const bool three = A.mode == "three";
  const int plan = (exact ? 1 : 0) + (three ? 2 : 0);
  const CorrPlan& pi = kPlans[plan];
  const Frame fr = frame ? *frame : reg_frame(Hp, A.w, A.h, B.w, B.h);
  const Mat3 Hc = canvas_h(Hp, fr.ox, fr.oy);
  const int W = fr.w, Hh = fr.h;
  RsPlan p{};
  p.cx = c0 - 1.0 - (double)fr.ox;
  p.cy = c1 - 1.0 - (double)fr.oy;
  double r1 = r1_opt > 0 ? r1_opt : 0.0;
  if (r1 == 0.0) {
    const double Wd = W, Hd = Hh;
    const double cs[4][2] = {{0, 0}, {Wd - 1, 0}, {Wd - 1, Hd - 1}, {0, Hd - 1}};
    for (const auto& q : cs) r1 = std::max(r1, std::hypot(q[0] - p.cx, q[1] - p.cy));
  }
  const double r0 = r0_opt >= 0 ? r0_opt : std::max(2.0, std::min(r1 / 16.0, 8.0));
  p.r1 = std::max(r1, 2.0);
  p.r0 = std::min(std::max(r0, 1e-3), p.r1 / 8.0);
  const uint32_t n_ring = std::max<uint32_t>(64, (uint32_t)std::ceil(2.0 * M_PI * p.r1));
  const bool ladder = set.e("fft_sizes") == "ladder";
  uint32_t n;
  if (cap == 0) {
    n = std::min<uint32_t>(std::max<uint32_t>(pow2_ceil(n_ring), 128), 512);
  } else {
    n = std::min(plan_size(n_ring, 1, ladder), largest_smooth(std::min(cap, kRsMaxN), ladder));
  }
  if (n < 32) throw std::runtime_error("FftTooShort: roto-scale grids take at least 32 angles");
  p.n = p.n_th = (int)n;
  p.n_lam = (int)std::max<uint32_t>(16, n >> 1);
  p.log0 = std::log(p.r0);
  p.span = std::log(p.r1) - p.log0;
  p.dlam = p.span / (double)p.n_lam;
  p.dth = 2.0 * M_PI / (double)p.n_th;
  const bool pow2 = (n & (n - 1)) == 0;
  p.px = (int)(pow2 ? n : pow2_ceil(2 * n - 1));
  p.py = (int)pow2_ceil(n);
  if (p.px > kMaxFft || p.py > kMaxFft) throw std::runtime_error("FftTooLong");
  const int nx = p.px, ny = p.py;
  const int64_t N = (int64_t)nx * ny;
  const size_t cn = (size_t)W * Hh;
  canvas_a.ensure(cn * 20);
  canvas_b.ensure(cn * 20);
  spec_a.ensure((size_t)N * 8 * pi.n_cplx);
  spec_b.ensure((size_t)N * 8 * pi.n_cplx);
  work.ensure((size_t)N * 8 * pi.n_out);
  corr.ensure((size_t)N * 4 * pi.n_corr);
  dst.ensure((size_t)n * n * 4);
  rs_tab.ensure((size_t)(p.n_lam + 2 * p.n_th) * 8);
  double Hi[9];
  warp_inverse(Hc, Hi);
  double* rho = rs_tab.as<double>();
  double* cth = rho + p.n_lam;
  double* sth = cth + p.n_th;
  if (dev_hi)
    LAUNCH("warp_stack", launch_warp_stack_dev(A.stack.as<float>(), A.w, A.h, dev_hi, canvas_a.as<float>(), W, Hh, st));
  else
    LAUNCH("warp_stack", launch_warp_stack(A.stack.as<float>(), A.w, A.h, Hi, canvas_a.as<float>(), W, Hh, st));
  LAUNCH("paste_stack", launch_paste_stack(B.stack.as<float>(), B.w, B.h, fr.ox, fr.oy, canvas_b.as<float>(), W, Hh, st));
  LAUNCH("rs_tables", launch_rs_tables(p.n_th, p.n_lam, p.log0, p.span, rho, cth, sth, st));
  LogPolarArgs la{};
  la.W = W;
  la.H = Hh;
  la.n_th = p.n_th;
  la.n_lam = p.n_lam;
  la.cx = p.cx;
  la.cy = p.cy;
  la.rho = rho;
  la.cth = cth;
  la.sth = sth;
  LogPolarArgs lb = la;
  la.canvas = canvas_a.as<float>();
  lb.canvas = canvas_b.as<float>();
  const float2* y_tw = twiddles(ny);
  bool ok = true;
  LAUNCH("pack_logpolar", ok &= launch_pack_logpolar(plan, la, nx, ny, spec_a.as<float2>(), twiddles(nx), st));
  LAUNCH("pack_logpolar", ok &= launch_pack_logpolar(plan, lb, nx, ny, spec_b.as<float2>(), twiddles(nx), st));
  LAUNCH("colfft_fwd", ok &= launch_colfft(spec_a.as<float2>(), pi.n_cplx, nx, ny, p.n_lam, false, y_tw, st, col_lines(ny)));
  LAUNCH("colfft_fwd", ok &= launch_colfft(spec_b.as<float2>(), pi.n_cplx, nx, ny, p.n_lam, false, y_tw, st, col_lines(ny)));
  LAUNCH("cmul", launch_cmul(plan, spec_a.as<float2>(), spec_b.as<float2>(), work.as<float2>(), nx, ny, st));
  LAUNCH("colfft_inv", ok &= launch_colfft(work.as<float2>(), pi.n_out, nx, ny, ny, true, y_tw, st, col_lines(ny)));
  LAUNCH("rowifft_extract", ok &= launch_rowifft_extract(plan, work.as<float2>(), corr.as<float>(), nx, ny, twiddles(nx), st));
  if (!ok) throw std::runtime_error("UnsupportedFftLength");
  CombineArgs ca{};
  ca.exact = exact;
  ca.calib = exact && calib;
  ca.nf = three ? 3 : 4;
  ca.min_n = (float)kMinN;
  ca.ridge = kRidge;
  ca.area = std::max(1.0, pair_area(A, B));   // log-polar cells: area scale 1 (smi.zig:633)
  RsArgs ra{};
  ra.n = p.n;
  ra.px = nx;
  ra.py = ny;
  ra.dlam = p.dlam;
  ra.dth = p.dth;
  LAUNCH("combine_rs", launch_combine_rs(plan, corr.as<float>(), dst.as<float>(), ra, ca, st));
  ++counts[1];
  return p;
}

void Engine::rs_finish(const RsPlan& p, bool symmetric, bool mirror, bool to_host) {
  const int n = p.n;
  const int64_t nn = (int64_t)n * n;
  const double smin = set.f("rs_smin"), smax = set.f("rs_smax");
  const double lo = std::log(std::max(smin > 0 ? smin : 0.2, 1e-6));
  const double hi = std::log(std::max(smax > 0 ? smax : 5.0, 1.000001));
  rs_out.ensure((size_t)nn * 4);
  const int nparts = combine_blocks(nn);
  arg_part.ensure((size_t)nparts * 8);
  res.ensure(16);
  LAUNCH("rs_finish", launch_rs_finish(rs_fwd.as<float>(), symmetric ? rs_inv.as<float>() : nullptr, n, mirror, p.dlam,
                                       lo, hi, rs_out.as<float>(), arg_part.as<float>(), st));
  LAUNCH("argmax_finish", launch_argmax_finish(arg_part.as<float>(), nparts, rs_out.as<float>(), res.as<float>(), st));
  if (!to_host) return;
  ZC_HIP(hipMemcpyAsync(res_host, res.p, 12, hipMemcpyDeviceToHost, st));
  prof.break_chain();
}

RsResult Engine::rs_map(const Mat3& H, const double* c_opt, bool symmetric, float* map, float* fwd, float* inv,
                        size_t len) {
  need_images();
  const std::string metric = set.e("metric");
  if (metric != "smi") throw std::runtime_error("Unsupported metric " + metric + " (smi only)");
  const bool exact = set.b("exact");
  const bool calib = exact && set.b("calibrated");
  const uint32_t cap = set.u("map_res");
  const std::array<double, 2> c =
      c_opt ? std::array<double, 2>{c_opt[0], c_opt[1]} : moving_centroid(H, sides[0].w, sides[0].h);
  const RsPlan p = rs_map_dir(sides[0], sides[1], H, c[0], c[1], -1.0, 0.0, exact, calib, cap, rs_fwd);
  if (symmetric) {
    const Mat3 Hi = inv3(H);
    const auto ci = map_pt(Hi, c[0], c[1]);
    const RsPlan q = rs_map_dir(sides[1], sides[0], Hi, ci[0], ci[1], p.r0, p.r1, exact, calib, cap, rs_inv);
    if (q.n != p.n) throw std::runtime_error("GridMismatch");
  }
  const int n = p.n;
  const int64_t nn = (int64_t)n * n;
  if ((map || fwd || (inv && symmetric)) && len < (size_t)nn) throw std::runtime_error("MapTooSmall");
  rs_finish(p, symmetric, H[0] * H[4] - H[1] * H[3] > 0);
  if (map) ZC_HIP(hipMemcpyAsync(map, rs_out.p, nn * 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  if (fwd) ZC_HIP(hipMemcpyAsync(fwd, rs_fwd.p, nn * 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  if (inv && symmetric) ZC_HIP(hipMemcpyAsync(inv, rs_inv.p, nn * 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
  RsResult r{};
  r.n = (uint32_t)n;
  r.n_th = (uint32_t)p.n_th;
  r.n_lam = (uint32_t)p.n_lam;
  r.dlam = p.dlam;
  r.r0 = p.r0;
  r.r1 = p.r1;
  r.cx = c[0];
  r.cy = c[1];
  r.peak = res_host[0];
  r.zero = res_host[2];
  r.symmetric = symmetric ? 1u : 0u;
  r.calibrated = calib ? 1u : 0u;
  uint32_t idx;
  std::memcpy(&idx, &res_host[1], 4);
  if (idx < (uint64_t)nn) {
    // engine.zig:864-870: the peak lag → the correction about c
    const int64_t py = idx / n, px = idx % n;
    const int64_t lx = px <= n / 2 ? px : px - n;
    const int64_t ly = py <= n / 2 ? py : py - n;
    r.peak_index = idx;
    r.dth = -((double)lx / (double)std::max(p.n_th, 1)) * 2.0 * M_PI;
    r.dsg = -(double)ly * p.dlam;
  }
  return r;
}

}  // namespace zc
