// The rotation × scale sweep (search.zig:300-494 Sweep.sweep): every candidate pose's shift map
// on a square canvas around the fixed image, reduced on the GPU to its top local maxima.
//
// Candidates are grouped by plane size and run in batches: per batch one pack + row FFT launch
// (all candidates' windows), one column FFT launch over every plane of the batch, one cross
// spectrum against the group's cached fixed spectra, the inverse column and row FFTs, the
// per-lag scores and the peak reduction. Plane lengths are powers of two (the linear
// correlation of an M-cell and an F-cell window has the same values per lag in any plane of
// at least M + F − 1 points).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include <map>
#include <mutex>

#include "engine.hpp"
#include "plan.hpp"

namespace zc {

void Engine::ensure_il(const Side& sd) {
  if (sd.il_ok) return;
  const int64_t n = (int64_t)sd.w * sd.h;
  sd.il.ensure((size_t)n * 16);
  LAUNCH("interleave", launch_interleave(sd.feat.as<float>(), n, sd.il.as<float4>(), st));
  sd.il_ok = true;
}

namespace {

// canvas cells [lo, hi) covering dest-pixel extent [a, b], one cell of margin (search.zig:329-334)
std::array<uint32_t, 2> cells_of(double a, double b, double o0, double c, int64_t g) {
  const int64_t lo = std::clamp<int64_t>((int64_t)std::floor((a - o0) / c) - 1, 0, g);
  const int64_t hi = std::clamp<int64_t>((int64_t)std::ceil((b - o0) / c) + 1, lo + 1, g);
  return {(uint32_t)lo, (uint32_t)hi};
}

// The shortest plane the kernels take that holds `need` points: a power of two, or with `smooth`
// 1536, 2304 or 3072 (2^a·3^b lengths of the reference's ladder in that range,
// search.zig:180-188).
uint32_t plane_len(uint32_t need, bool smooth) {
  uint32_t n = pow2_ceil(std::max<uint32_t>(need, 32));
  if (smooth)
    for (uint32_t s3 : {1536u, 2304u, 3072u})
      if (s3 >= need && s3 < n) n = s3;
  return n;
}

struct Cand {
  uint32_t it, is, ik, ia;
  double th, s;
  Mat3 H;
  uint32_t gx0, gy0, nx, ny;
  int taps;
  uint32_t mw, mh;   // the moving footprint's cells from (gx0, gy0): window cells beyond are 0
};

}  // namespace

SweepOut Engine::sweep(const SweepOpts& o) {
  need_images();
  return sweep_sides(sides[0], sides[1], o);
}

SweepOut Engine::sweep_sides(Side& A, Side& B, const SweepOpts& o) {
  if (A.mode != B.mode) throw std::runtime_error("FeatureModeMismatch: set both images with the same smi_features");
  if (o.mode > 1) throw std::runtime_error("Unsupported: the sweep scores SMI (global or exact) only");
  const bool exact = o.mode == 1;
  const bool three = A.mode == "three";
  const int plan = (exact ? 1 : 0) + (three ? 2 : 0);
  const CorrPlan& pi = kPlans[plan];
  const double fw = B.w, fh = B.h, mw = A.w, mh = A.h;
  const SweepGrid& gv = *o.gv;
  // canvas: a square around the fixed image, large enough for the moving image at the largest
  // stretch, capped at twice the fixed size
  const auto cm = map_pt(o.H0, (mw + 1) / 2, (mh + 1) / 2);
  const double cmv[2] = {cm[0], cm[1]};
  double s_max = -INFINITY;
  for (double s : gv.sigmas)
    for (double k : gv.shears)
      for (double a : gv.anisos) s_max = std::max(s_max, sv2(cand_linear(0, s, k, a))[0]);
  const double D = std::min(2 * std::max(fw, fh), std::max(std::max(fw, fh), s_max * std::hypot(mw, mh)));
  const uint32_t grid = o.full ? (uint32_t)std::ceil(D - 1e-9) : o.grid;
  const double cell = D / (double)grid;
  const double ox = (fw + 1) / 2 - D / 2, oy = (fh + 1) / 2 - D / 2;
  const int64_t G = grid;
  const auto fcx = cells_of(0.5, fw + 0.5, ox, cell, G), fcy = cells_of(0.5, fh + 0.5, oy, cell, G);
  const uint32_t Fx = fcx[1] - fcx[0], Fy = fcy[1] - fcy[0];
  // candidates, θ-major (or the given cells, in order)
  std::vector<std::array<uint32_t, 4>> idx;
  if (o.cells) {
    for (const auto& c : *o.cells) {
      if (c[0] >= gv.thetas.size() || c[1] >= gv.sigmas.size()) throw std::runtime_error("BadArgument: sweep cell");
      idx.push_back({c[0], c[1], 0, 0});
    }
  } else {
    for (uint32_t i = 0; i < gv.thetas.size(); ++i)
      for (uint32_t j = 0; j < gv.sigmas.size(); ++j)
        for (uint32_t ik = 0; ik < gv.shears.size(); ++ik)
          for (uint32_t ia = 0; ia < gv.anisos.size(); ++ia) idx.push_back({i, j, ik, ia});
  }
  const uint32_t nc = (uint32_t)idx.size();
  if (nc == 0) throw std::runtime_error("BadArgument: no sweep candidates");
  std::vector<Cand> cands(nc);
  const double s0_raw = std::sqrt(std::fabs(o.H0[0] * o.H0[4] - o.H0[1] * o.H0[3]));
  const double s0 = (s0_raw == 0 || std::isnan(s0_raw)) ? 1 : s0_raw;
  {
    size_t ci = 0;
    for (const auto& q : idx) {
      const uint32_t i = q[0], j = q[1], ik = q[2], ia = q[3];
      const auto L = cand_linear(gv.thetas[i], gv.sigmas[j], gv.shears[ik], gv.anisos[ia]);
      const Mat3 H = mul3(about_c(cmv, L), o.H0);
      double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
      const double qs[4][2] = {{0.5, 0.5}, {mw + 0.5, 0.5}, {mw + 0.5, mh + 0.5}, {0.5, mh + 0.5}};
      for (const auto& cq : qs) {
        const auto p = map_pt(H, cq[0], cq[1]);
        x0 = std::min(x0, p[0]);
        x1 = std::max(x1, p[0]);
        y0 = std::min(y0, p[1]);
        y1 = std::max(y1, p[1]);
      }
      const bool ok = std::isfinite(x0 + x1 + y0 + y1);
      const auto mcx = ok ? cells_of(x0, x1, ox, cell, G) : std::array<uint32_t, 2>{0, grid};
      const auto mcy = ok ? cells_of(y0, y1, oy, cell, G) : std::array<uint32_t, 2>{0, grid};
      Cand& c = cands[ci++];
      c = {i, j, ik, ia, gv.thetas[i], gv.sigmas[j], H, mcx[0], mcy[0],
           plane_len(mcx[1] - mcx[0] + Fx - 1, o.smooth), plane_len(mcy[1] - mcy[0] + Fy - 1, o.smooth), 1};
      // box-filter taps per cell side from the candidate's smallest stretch
      c.taps = (int)std::min(6.0, std::max(1.0, std::ceil(cell / (sv2(L)[1] * s0) - 1e-3)));
      c.mw = mcx[1] - mcx[0];
      c.mh = mcy[1] - mcy[0];
      if (c.nx > (uint32_t)kMaxFft || c.ny > (uint32_t)kMaxFft) throw std::runtime_error("FftTooLong");
    }
  }
  // slots grouped by plane size; perm[k] = the candidate in slot k
  std::vector<uint32_t> perm(nc);
  std::iota(perm.begin(), perm.end(), 0u);
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
    const uint64_t ka = ((uint64_t)cands[a].nx << 32) | cands[a].ny, kb = ((uint64_t)cands[b].nx << 32) | cands[b].ny;
    return ka < kb;
  });
  std::vector<SweepPose> poses(nc);
  std::vector<int> zero(nc);
  std::vector<int4> dead(nc);
  uint64_t nn_max = 0;
  for (uint32_t k = 0; k < nc; ++k) {
    const Cand& c = cands[perm[k]];
    const Mat3 Hi = inv3(c.H);
    for (int q = 0; q < 9; ++q) poses[k].h[q] = (float)Hi[q];
    poses[k].taps = c.taps;
    poses[k].gx0 = (int)c.gx0;
    poses[k].gy0 = (int)c.gy0;
    // the plane index of canvas displacement 0: lag = fixed origin − window origin (mod plane)
    const int64_t lx = (((int64_t)fcx[0] - c.gx0) % c.nx + c.nx) % c.nx;
    const int64_t ly = (((int64_t)fcy[0] - c.gy0) % c.ny + c.ny) % c.ny;
    zero[k] = (int)(ly * c.nx + lx);
    // lags (moving index − fixed index) outside [−(F − 1), footprint − 1] pair the fixed window
    // with the empty part of the moving one: plane rows (and columns) [footprint, plane − F + 1)
    dead[k] = make_int4((int)c.mh, (int)(c.ny - Fy + 1), (int)c.mw, (int)(c.nx - Fx + 1));
    nn_max = std::max<uint64_t>(nn_max, (uint64_t)c.nx * c.ny);
  }
  SweepPose pf{};
  pf.h[0] = pf.h[4] = pf.h[8] = 1.0f;
  pf.taps = (int)std::min(6.0, std::max(1.0, std::ceil(cell - 1e-3)));
  pf.gx0 = (int)fcx[0];
  pf.gy0 = (int)fcy[0];
  // work buffers from a memory budget: batches of up to o.batch candidates whose planes fit
  // (every buffer scales with the batch's plane points). The budget: SweepTune::budget_gb, at
  // most half the device memory free now, at least 8 GiB (at 4096² planes a candidate's buffers
  // take 7.3 GB, and 4 per batch measured 6 % faster than 1).
  const uint64_t budget_cfg = (uint64_t)SweepTune::budget_gb.load() << 30;
  uint64_t budget = budget_cfg;
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
      budget = std::max<uint64_t>(std::min<uint64_t>(budget, free_b / 2), 8ull << 30);
  }
  auto per_cand = [&](uint64_t nn) { return nn * 8 * (pi.n_cplx + pi.n_out) + nn * 4 * (pi.n_corr + 1); };
  const uint64_t B_max = std::max<uint64_t>(1, std::min<uint64_t>(o.batch, budget / per_cand(nn_max)));
  const uint64_t pts_max = B_max * nn_max;   // plane points per batch
  // bn: the batch bound now (the budget, capped by the free memory); bk: the bound of the
  // configured budget alone, which keys the batch-size trials (the free memory changes with
  // whatever else the process holds; the trials' result then applies up to bn)
  struct Group {
    uint32_t k0, k1, nx, ny, bn, bk;
  };
  std::vector<Group> groups;
  uint64_t part_max = 0, pts_need = 0;
  for (uint32_t k0 = 0; k0 < nc;) {
    const uint32_t nx = cands[perm[k0]].nx, ny = cands[perm[k0]].ny;
    uint32_t k1 = k0;
    while (k1 < nc && cands[perm[k1]].nx == nx && cands[perm[k1]].ny == ny) ++k1;
    const uint64_t nn = (uint64_t)nx * ny;
    const uint32_t bn = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(o.batch, pts_max / nn));
    const uint32_t bk = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(o.batch, budget_cfg / per_cand(nn)));
    groups.push_back({k0, k1, nx, ny, bn, bk});
    // no batch runs larger than its group (nor larger than bn, tuned or not): the buffers hold
    // what the sweep runs, not the budget
    const uint64_t b_run = std::min<uint64_t>(bn, k1 - k0);
    part_max = std::max<uint64_t>(part_max, b_run * peaks_chunks((int64_t)nn));
    pts_need = std::max<uint64_t>(pts_need, b_run * nn);
    k0 = k1;
  }
  const uint32_t n_peak = o.n_peak;
  sw_poses.ensure(sizeof(SweepPose) * nc);
  sw_pose_fixed.ensure(sizeof(SweepPose));
  sw_zero.ensure(sizeof(int) * nc);
  sw_dead.ensure(sizeof(int4) * nc);
  sw_peaks.ensure(sizeof(float2) * nc * n_peak);
  spec_b.ensure(nn_max * 8 * pi.n_cplx);
  spec_a.ensure(pts_need * 8 * pi.n_cplx);
  work.ensure(pts_need * 8 * pi.n_out);
  corr.ensure(pts_need * 4 * pi.n_corr);
  sw_ns.ensure(pts_need * 4);
  sw_part.ensure(sizeof(float2) * 256 * part_max);
  ZC_HIP(hipMemcpyAsync(sw_poses.p, poses.data(), sizeof(SweepPose) * nc, hipMemcpyHostToDevice, st));
  prof.break_chain();
  ZC_HIP(hipMemcpyAsync(sw_pose_fixed.p, &pf, sizeof pf, hipMemcpyHostToDevice, st));
  prof.break_chain();
  ZC_HIP(hipMemcpyAsync(sw_zero.p, zero.data(), sizeof(int) * nc, hipMemcpyHostToDevice, st));
  prof.break_chain();
  ZC_HIP(hipMemcpyAsync(sw_dead.p, dead.data(), sizeof(int4) * nc, hipMemcpyHostToDevice, st));
  prof.break_chain();
  ensure_il(A);
  ensure_il(B);
  const uint32_t min_n =
      (uint32_t)std::max<double>(kMinN, (double)js_round(o.min_overlap * (fw * fh) / (cell * cell)));
  CombineArgs ca{};
  ca.exact = exact;
  ca.calib = 0;
  ca.nf = three ? 3 : 4;
  ca.min_n = (float)min_n;
  ca.ridge = kRidge;
  ca.area = 1.0;
  const SweepPose* dposes = sw_poses.as<SweepPose>();
  // the fixed planes of a group's plane size: its footprint at the plane origin
  auto fixed_batch = [&](const Group& gr) {
    bool ok = true;
    const uint32_t nx = gr.nx, ny = gr.ny;
    const float2* tx = twiddles((int)nx);
    const float2* ty = twiddles((int)ny);
    SweepPack uf{(float)ox, (float)oy, (float)cell, (int)grid, B.w, B.h, (int)Fx, (int)Fy, (int)ny};
    const double NCd = pi.n_cplx, nnd = (double)nx * ny;
    bytes("pack_sweep_fixed", NCd * Fy * nx * 8 + 16.0 * B.w * B.h);
    bytes("colfft_fwd_fixed", NCd * (Fy * nx * 8.0 + nnd * 8));
    LAUNCH("pack_sweep_fixed", ok &= launch_pack_sweep(plan, B.il.as<float4>(), sw_pose_fixed.as<SweepPose>(), 1, uf,
                                                        (int)nx, spec_b.as<float2>(), tx, st));
    LAUNCH("colfft_fwd_fixed", ok &= launch_colfft(spec_b.as<float2>(), pi.n_cplx, (int)nx, (int)ny, (int)Fy, false,
                                                    ty, st, col_lines((int)ny)));
    if (!ok) throw std::runtime_error("UnsupportedFftLength");
  };
  // b candidates from slot c0 of a group: moving windows (as wide as the plane allows without
  // wraparound), spectra, cross spectra, inverse transforms, scores, peaks
  auto moving_batch = [&](const Group& gr, uint32_t c0, int b) {
    bool ok = true;
    const uint32_t nx = gr.nx, ny = gr.ny;
    const int64_t nn = (int64_t)nx * ny;
    const float2* tx = twiddles((int)nx);
    const float2* ty = twiddles((int)ny);
    // the window as wide and tall as the batch's largest footprint: the cells beyond it are 0 in
    // every candidate's window, and the column FFT reads the rows not written as 0
    uint32_t fw_b = 0, fh_b = 0;
    for (int i = 0; i < b; ++i) {
      fw_b = std::max(fw_b, cands[perm[c0 + i]].mw);
      fh_b = std::max(fh_b, cands[perm[c0 + i]].mh);
    }
    SweepPack um{(float)ox, (float)oy, (float)cell, (int)grid, A.w, A.h, (int)std::min(nx - Fx + 1, fw_b),
                 (int)std::min(ny - Fy + 1, fh_b), (int)ny};
    const double NCd = pi.n_cplx, Qd = pi.n_out, Cd = pi.n_corr, nnd = (double)nn;
    // algorithmic bytes of each launch: its role's minimal HBM traffic
    bytes("pack_sweep", b * (NCd * um.win_y * nx * 8 + 16.0 * A.w * A.h));
    bytes("colfft_fwd", b * NCd * (um.win_y * nx * 8.0 + nnd * 8));
    bytes("cmul", b * (2 * NCd + Qd) * nnd * 8);
    // The inverse columns: Q planes read, their lag rows with overlap written (the rows without
    // overlap are read by nothing). The inverse rows: lag rows with overlap read Q outputs and
    // store C planes (lines of up to 1024 points: n, and the other planes at the lag columns with
    // overlap); rows without overlap store n = 0. The score: n and the score at every lag, the
    // other planes at the lags with overlap (an upper bound of those above the floor).
    double cols_b = 0, rows_b = 0, comb_b = 0;
    for (int i = 0; i < b; ++i) {
      const Cand& cd = cands[perm[c0 + i]];
      const double ly = std::min<double>(ny, cd.mh + Fy - 1.0), lx = std::min<double>(nx, cd.mw + Fx - 1.0);
      cols_b += Qd * nnd * 8 + Qd * ly * nx * 8;
      rows_b += nx <= 1024 ? ly * nx * (Qd * 8 + 4) + ly * lx * (Cd - 1) * 4 + (ny - ly) * nx * 4
                           : ly * nx * (Qd * 8 + Cd * 4) + (ny - ly) * nx * 4;
      comb_b += 2 * nnd * 4 + ly * lx * (Cd - 1) * 4;
    }
    bytes("colfft_inv", cols_b);
    bytes("rowifft_extract", rows_b);
    bytes("combine_sweep", comb_b);
    bytes("peaks", b * nnd * 4);
    LAUNCH("pack_sweep", ok &= launch_pack_sweep(plan, A.il.as<float4>(), dposes + c0, b, um, (int)nx,
                                                  spec_a.as<float2>(), tx, st));
    LAUNCH("colfft_fwd", ok &= launch_colfft(spec_a.as<float2>(), pi.n_cplx * b, (int)nx, (int)ny, um.win_y, false,
                                            ty, st, col_lines((int)ny)));
    LAUNCH("cmul", launch_cmul(plan, spec_a.as<float2>(), spec_b.as<float2>(), work.as<float2>(), (int)nx, (int)ny,
                               st, b));
    LAUNCH("colfft_inv", ok &= launch_colfft(work.as<float2>(), pi.n_out * b, (int)nx, (int)ny, (int)ny, true, ty,
                                            st, col_lines((int)ny), sw_dead.as<int4>() + c0, pi.n_out));
    LAUNCH("rowifft_extract", ok &= launch_rowifft_extract(plan, work.as<float2>(), corr.as<float>(), (int)nx,
                                                           (int)ny, tx, st, b, sw_dead.as<int4>() + c0));
    LAUNCH("combine_sweep", launch_combine_sweep(plan, corr.as<float>(), sw_ns.as<float>(), nn, b, ca,
                                                 sw_zero.as<int>() + c0, st));
    LAUNCH("peaks", launch_peaks(sw_ns.as<float>(), (int)nx, (int)ny, b, (int)n_peak, sw_part.as<float2>(),
                                 sw_peaks.as<float2>() + (size_t)c0 * n_peak, st));
    if (!ok) throw std::runtime_error("UnsupportedFftLength");
  };
  // Candidates per batch, tuned per (plan, plane size, window) on first use unless forced: each
  // power of two up to the memory bound (and the bound itself) runs one batch untimed, then the
  // group's first max(4 b, 16) candidates (at most the group) in batches of b, timed twice; the
  // least time per candidate wins, and a larger batch within 1 % of it (its fixed spectra and
  // launches are shared by more candidates: at 3072² planes 8 per batch swept 4 % faster than 2
  // where the trials had timed them within 3 %). The runs write the same outputs the sweep then
  // writes again.
  auto tuned_batch = [&](const Group& gr) -> uint32_t {
    static std::mutex mu;
    static std::map<std::array<uint32_t, 6>, uint32_t> best;
    const std::array<uint32_t, 6> k{(uint32_t)plan, gr.nx, gr.ny, Fx, Fy, gr.bk};
    {
      std::lock_guard<std::mutex> lk(mu);
      const auto it = best.find(k);
      if (it != best.end()) return it->second;
    }
    const uint32_t n_grp = gr.k1 - gr.k0, avail = std::min(gr.bn, n_grp);
    std::vector<uint32_t> sizes;
    for (uint32_t b = 1; b < avail; b *= 2) sizes.push_back(b);
    sizes.push_back(avail);
    hipEvent_t e0, e1;
    ZC_HIP(hipEventCreate(&e0));
    ZC_HIP(hipEventCreate(&e1));
    fixed_batch(gr);
    std::vector<double> per_cand;
    for (const uint32_t b : sizes) {
      const uint32_t n = std::min(n_grp, std::max(4 * b, 16u));
      moving_batch(gr, gr.k0, (int)b);
      float t = INFINITY;
      for (int r = 0; r < 2; ++r) {
        ZC_HIP(hipEventRecord(e0, st));
        for (uint32_t c0 = gr.k0; c0 < gr.k0 + n; c0 += b) moving_batch(gr, c0, (int)std::min(b, gr.k0 + n - c0));
        ZC_HIP(hipEventRecord(e1, st));
        ZC_HIP(hipEventSynchronize(e1));
        float ms = 0;
        ZC_HIP(hipEventElapsedTime(&ms, e0, e1));
        t = std::min(t, ms);
      }
      per_cand.push_back((double)t / n);
    }
    const double t_min = *std::min_element(per_cand.begin(), per_cand.end());
    double t_best = t_min;
    uint32_t b_best = sizes[0];
    for (size_t i = 0; i < sizes.size(); ++i)
      if (per_cand[i] <= 1.01 * t_min) {
        b_best = sizes[i];
        t_best = per_cand[i];
      }
    ZC_HIP(hipEventDestroy(e0));
    ZC_HIP(hipEventDestroy(e1));
    prof.break_chain();
    events.log("sweep: %u candidates per batch at %ux%u planes (timed, %.3f ms per candidate)", b_best, gr.nx, gr.ny,
               t_best);
    std::lock_guard<std::mutex> lk(mu);
    best[k] = b_best;
    return b_best;
  };
  if (o.tune_batch && !prof.on)
    for (Group& gr : groups)
      if (gr.k1 - gr.k0 > 1) gr.bn = std::min(gr.bn, tuned_batch(gr));
  auto issue = [&] {
    for (const Group& gr : groups) {
      fixed_batch(gr);
      for (uint32_t c0 = gr.k0; c0 < gr.k1; c0 += gr.bn) moving_batch(gr, c0, (int)std::min(gr.bn, gr.k1 - c0));
    }
  };
  // everything the launches read from the host, resolved before a capture: twiddles and the
  // column tilings (which may time the tilings on first use), then the key of the launch sequence
  std::vector<uint8_t> key;
  auto put = [&key](const auto& v) {
    const auto* p = reinterpret_cast<const uint8_t*>(&v);
    key.insert(key.end(), p, p + sizeof v);
  };
  for (uint32_t k = 0; k < nc; ++k) {   // the footprints size each batch's pack window
    put(cands[perm[k]].mw);
    put(cands[perm[k]].mh);
  }
  for (const Group& gr : groups) {
    put(gr);
    put(twiddles((int)gr.nx));
    put(twiddles((int)gr.ny));
    put(col_lines((int)gr.ny));
  }
  put(plan);
  put(FftRowTiling::v.load());   // the inverse row pass's kernel
  put(n_peak);
  put(ox);
  put(oy);
  put(cell);
  put(grid);
  put(Fx);
  put(Fy);
  put(ca);
  for (const int v : {A.w, A.h, B.w, B.h}) put(v);
  for (const void* p : {A.il.p, B.il.p, spec_a.p, spec_b.p, work.p, corr.p, sw_ns.p, sw_part.p, sw_peaks.p,
                        sw_poses.p, sw_pose_fixed.p, sw_zero.p, sw_dead.p})
    put(p);
  run_captured(sweep_gc, std::move(key), prof.on || !set.b("sweep_graph"), issue);
  sweep_last_nx = groups.back().nx;
  sweep_last_ny = groups.back().ny;
  std::vector<float2> raw((size_t)nc * n_peak);
  ZC_HIP(hipMemcpyAsync(raw.data(), sw_peaks.p, raw.size() * sizeof(float2), hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  ++counts[1];
  sync();
  // decode (search.zig:462-492)
  SweepOut out;
  out.nc = nc;
  out.cell = cell;
  out.grid = grid;
  out.D = D;
  out.ox = ox;
  out.oy = oy;
  out.best.assign(nc, 0.0f);
  const double c2 = cell * cell;
  for (uint32_t slot = 0; slot < nc; ++slot) {
    const Cand& cd = cands[perm[slot]];
    for (uint32_t k = 0; k < n_peak; ++k) {
      const float2 e = raw[(size_t)slot * n_peak + k];
      const double v = (double)e.x * c2;
      uint32_t t;
      std::memcpy(&t, &e.y, 4);
      if (!(v > 0) || (uint64_t)t >= (uint64_t)cd.nx * cd.ny) continue;
      if (k == 0) out.best[perm[slot]] = (float)v;
      out.n = std::max(out.n, std::max(cd.nx, cd.ny));
      // lag (moving index − fixed index) in [−(F − 1), window − 1], then the canvas
      // displacement of the moving image: gx0 + lag − fx0 cells
      const int64_t py = t / cd.nx, px = t - py * cd.nx;
      const int64_t wx = (int64_t)cd.nx - Fx + 1, wy = (int64_t)cd.ny - Fy + 1;
      const int64_t lx = px < wx ? px : px - cd.nx, ly = py < wy ? py : py - cd.ny;
      const int32_t sx = (int32_t)((int64_t)cd.gx0 + lx - fcx[0]);
      const int32_t sy = (int32_t)((int64_t)cd.gy0 + ly - fcy[0]);
      // the map peak displaces the moving image by +d; the correction is T(−d)
      const double dx = sx * cell, dy = sy * cell;
      const Mat3 H = mul3(translate(-dx, -dy), cd.H);
      out.peaks.push_back({v, perm[slot], cd.it, cd.is, cd.ik, cd.ia, 0, cd.th, cd.s, sx, sy, H});
    }
  }
  return out;
}

void Engine::read_sweep_map(float* out, size_t len, uint32_t dims[2]) {
  const size_t nn = (size_t)sweep_last_nx * sweep_last_ny;
  if (nn == 0) throw std::runtime_error("NoSweep");
  if (len < nn) throw std::runtime_error("MapTooSmall");
  ZC_HIP(hipMemcpyAsync(out, sw_ns.p, nn * 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
  dims[0] = sweep_last_nx;
  dims[1] = sweep_last_ny;
}

}  // namespace zc
