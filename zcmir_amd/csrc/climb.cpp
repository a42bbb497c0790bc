// The climb from the engine's pose: gclimb.zig:313-489 (run) with climb.wgsl's stages (hop_cand,
// rs_cand, gn_cands, climb_select, iter_begin / iter_end), SMI family.
//
// Per iteration: a shift-map hop (the map's peak as a translation of the pose), a Gauss–Newton
// block (up to three steps, each trying five lengths of the step; the block ends at a step that
// beats nothing), and with hop_fm a roto-scale hop about the moving centroid followed by another
// block. A candidate replaces the pose when it beats its score by 1e-6 of it. A move resets the
// stall count; the climb ends after two iterations in a row without a move (one when no hop is
// on) or after n_iters. The maps' canvas and the roto-scale centre are fixed for batches of eight
// iterations, from the pose each batch starts at. The host drives the loop; every map, score and
// gradient runs on the device.
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "engine.hpp"
#include "lie.hpp"
#include "search.hpp"

namespace zc {

namespace {
constexpr int kBatch = 8;      // gclimb.zig:50 BATCH
constexpr int kGnSteps = 3;    // gclimb.zig:51 GN_STEPS
}  // namespace

ClimbResult Engine::climb(uint32_t n_iters, bool hop, bool hop_fm, bool homography, bool exact, bool symmetric) {
  if (set.b("climb_graph")) return climb_device(n_iters, hop, hop_fm, homography, exact, symmetric);
  need_images();
  cancel = false;
  Side& A = sides[0];
  Side& B = sides[1];
  const std::vector<PairSides> own = own_pair(1);
  const bool calib = exact && set.b("calibrated");
  const uint32_t cap = set.u("map_res");
  const double min_ov = set.f("min_overlap");
  const int nk = n_keys(homography);
  n_iters = std::max<uint32_t>(1, n_iters);

  Mat3 Hc = H;
  GradOut g = pair_grads(own, {Hc}, exact, homography, symmetric, true)[0];
  double cur = g.score;
  const double start = cur;
  uint32_t iter = 0, stall = 0, n_rec = 0;
  bool moved = false, done = false;

  // climb_select: the best candidate that beats the pose by 1e-6 of its score becomes the pose
  auto select = [&](const std::vector<Mat3>& cands) -> bool {
    if (cands.empty()) return false;
    const std::vector<double> sc = pair_scores(std::vector<PairSides>(cands.size(), own[0]), cands, exact, symmetric);
    int best = -1;
    double bs = cur;
    for (size_t i = 0; i < cands.size(); ++i) {
      if (!(sc[i] == sc[i])) continue;
      if (sc[i] > cur + std::max(1e-6 * std::fabs(cur), 1e-12) && (best < 0 || sc[i] > bs)) {
        best = (int)i;
        bs = sc[i];
      }
    }
    if (best < 0) return false;
    Hc = cands[best];
    cur = bs;
    moved = true;
    ++n_rec;
    g = pair_grads(own, {Hc}, exact, homography, symmetric, true)[0];
    return true;
  };
  // gn_cands: (A + 1e-4 max(A_ii, 1e-12)) δ = ½ g, the step at ×{0.5, 1, 2, 4, 8} capped at norm 0.25
  auto gn_cands = [&]() {
    std::vector<Mat3> out;
    double M[64], half[8], step[8];
    std::copy(g.hess, g.hess + nk * nk, M);
    for (int q = 0; q < nk; ++q) {
      M[q * nk + q] += 1e-4 * std::max(M[q * nk + q], 1e-12);
      half[q] = 0.5 * g.grad[q];
    }
    if (!solve_linear(M, half, nk, step)) return out;
    double sn = 0.0;
    for (int q = 0; q < nk; ++q) {
      if (!(std::fabs(step[q]) < 1e30)) return out;
      sn += step[q] * step[q];
    }
    sn = std::sqrt(sn);
    for (double mu : {0.5, 1.0, 2.0, 4.0, 8.0}) {
      const double nrm = sn * mu;
      const double sc = mu * (nrm > 0.25 ? 0.25 / nrm : 1.0);
      Coords x{};
      for (int q = 0; q < nk; ++q) x[q] = sc * step[q];
      out.push_back(compose_n(Hc, x, homography, B.w, B.h));
    }
    return out;
  };
  auto gn_block = [&]() {
    for (int s = 0; s < kGnSteps && !cancel; ++s)
      if (!select(gn_cands())) break;
  };

  while (!done && !cancel) {
    const uint32_t rec0 = n_rec;
    const Frame fr = reg_frame(Hc, A.w, A.h, B.w, B.h);
    const auto rs_c = moving_centroid(Hc, A.w, A.h);
    const bool mirror = Hc[0] * Hc[4] - Hc[1] * Hc[3] > 0;
    for (int b = 0; b < kBatch && !done && !cancel; ++b) {
      moved = false;
      if (hop) {
        // hop_cand: translate(−dx, −dy) · H from the map's peak
        const MapOut m = shift_map_sides(A, B, Hc, exact, calib, cap, min_ov, true, &fr);
        uint32_t idx;
        std::memcpy(&idx, &m.res[1], 4);
        const uint64_t nn = (uint64_t)m.g.nx * m.g.ny;
        if (idx < nn) {
          const int64_t py = idx / m.g.nx, px = idx % m.g.nx;
          const int64_t lx = px <= m.g.nx / 2 ? px : px - (int64_t)m.g.nx;
          const int64_t ly = py <= m.g.ny / 2 ? py : py - (int64_t)m.g.ny;
          const double dx = (double)lx * ((double)m.fr.w / (double)m.g.cw);
          const double dy = (double)ly * ((double)m.fr.h / (double)m.g.ch);
          select({mul3(translate(-dx, -dy), Hc)});
        }
      }
      if (hop || !hop_fm) gn_block();
      if (hop_fm && !cancel) {
        // rs_cand: simAbout(c, dth, dsg) · H from the forward roto-scale map's peak
        const RsPlan p = rs_map_dir(A, B, Hc, rs_c[0], rs_c[1], -1.0, 0.0, exact, calib, cap, rs_fwd, &fr);
        rs_finish(p, false, mirror);
        sync();
        uint32_t idx;
        std::memcpy(&idx, &res_host[1], 4);
        const int n = p.n;
        if (idx < (uint64_t)n * n) {
          const int64_t py = idx / n, px = idx % n;
          const int64_t lx = px <= n / 2 ? px : px - n;
          const int64_t ly = py <= n / 2 ? py : py - n;
          const double dth = -((double)lx / (double)std::max(p.n_th, 1)) * 2.0 * M_PI;
          const double dsg = -(double)ly * p.dlam;
          if (!(std::fabs(dth) < 2e-4 && std::fabs(dsg) < 2e-4))
            select({project_group(mul3(sim_about(rs_c, dth, dsg), Hc), homography)});
        }
        gn_block();
      }
      stall = moved ? 0 : stall + 1;
      ++iter;
      if ((!moved && (!(hop || hop_fm) || stall >= 2)) || iter >= n_iters) done = true;
    }
    // per batch (gclimb.zig:446-481): the pose it reached and where the climb stands
    if (n_rec > rec0) {
      const auto parts = score_parts(Hc, exact, symmetric);
      events.trail("ascent", parts[0], parts[1], parts[2], Hc);
      events.pose(Hc);
    }
    events.log("ascent %u/%u score %.4f  stall %u  (%u steps so far)", iter, n_iters, cur, stall, n_rec);
  }
  events.log("ascent done  score %.4f", cur);
  H = Hc;
  ClimbResult r{};
  for (int i = 0; i < 9; ++i) r.H[i] = Hc[i];
  r.score = cur;
  r.iterations = iter;
  r.moved = (n_rec > 0 && cur > start) ? 1u : 0u;
  return r;
}

// The climb above with every decision on the device (k_gn.hip k_climb_*): per iteration one
// sequence of launches — the pose's warp inverse, the shift map at that pose and its hop, a
// Gauss–Newton block, the roto-scale map and its hop, another block, the stall and stop rule —
// over device state (pose 0 of a one-pose Gauss–Newton batch and a ClimbDev), captured once into a
// graph and replayed per iteration; launches after the end change nothing (every decision kernel
// checks `done`). The host reads the state (136 bytes) after each iteration, stops on `done`, and
// sets each batch's frame, roto-scale centre and orientation from the pose the batch starts at.
// The first iteration of a batch whose graph is not cached runs eagerly (it sizes the buffers,
// twiddle tables and column tilings the capture then reuses).
ClimbResult Engine::climb_device(uint32_t n_iters, bool hop, bool hop_fm, bool homography, bool exact,
                                 bool symmetric) {
  need_images();
  cancel = false;
  Side& A = sides[0];
  Side& B = sides[1];
  const std::vector<PairSides> own = own_pair(1);
  const bool calib = exact && set.b("calibrated");
  const uint32_t cap = set.u("map_res");
  const double min_ov = set.f("min_overlap");
  n_iters = std::max<uint32_t>(1, n_iters);

  Mat3 Hc = H;
  const double start = pair_grads(own, {Hc}, exact, homography, symmetric, true)[0].score;
  double cur = start;
  ensure_il(A);
  ensure_il(B);
  const GnPair pair{A.il.as<float4>(), B.il.as<float4>(), A.w, A.h, B.w, B.h};
  GnState s0{};
  for (int k = 0; k < 9; ++k) s0.H[k] = Hc[k];
  s0.score = cur;
  s0.active = 1;
  GnArgs ga{};
  ga.n = 1;
  ga.homography = homography;
  ga.exact = exact;
  ga.symmetric = symmetric;
  ga.min_n = kMinN;
  ga.n_amps = 0;
  ga.n_try = 5;   // the step at ×{0.5, 1, 2, 4, 8}; a hop uses slot 0
  ga.ridge = kRidge;
  const int nd = symmetric ? 2 : 1, T = ga.n_try;
  const int64_t px = std::max<int64_t>((int64_t)A.w * A.h, (int64_t)B.w * B.h);
  const int blocks = jobs_blocks(px, nd), blocks_try = jobs_blocks(px, nd * T);
  gd_pairs.ensure(sizeof(GnPair));
  gd_state.ensure(sizeof(GnState));
  gd_gjobs.ensure(nd * sizeof(PoseJob));
  gd_gsums.ensure(nd * 45 * 8);
  gd_tjobs.ensure(nd * sizeof(TangentJob));
  gd_tsums.ensure(nd * 44 * 8);
  gd_dir.ensure(2 * nd * 8);
  gd_step.ensure(18 * 8);
  gd_try_H.ensure(T * 9 * 8);
  gd_try_ok.ensure(T * 4);
  gd_try_jobs.ensure((size_t)nd * T * sizeof(PoseJob));
  gd_try_sums.ensure((size_t)nd * T * 45 * 8);
  gd_try_score.ensure((size_t)nd * T * 8);
  gd_part.ensure((size_t)std::max(nd * blocks, nd * T * blocks_try) * 45 * 8);
  gd_ctr.ensure(2 * 4);
  cl_state.ensure(sizeof(ClimbDev) + 9 * 8);
  ClimbDev* cdev = cl_state.as<ClimbDev>();
  double* hi = reinterpret_cast<double*>(cl_state.as<uint8_t>() + sizeof(ClimbDev));
  const ClimbDev c0{};
  ZC_HIP(hipMemcpyAsync(gd_pairs.p, &pair, sizeof pair, hipMemcpyHostToDevice, st));
  ZC_HIP(hipMemcpyAsync(gd_state.p, &s0, sizeof s0, hipMemcpyHostToDevice, st));
  ZC_HIP(hipMemcpyAsync(cl_state.p, &c0, sizeof c0, hipMemcpyHostToDevice, st));
  prof.break_chain();
  const GnBuffers gb{gd_pairs.as<GnPair>(), gd_state.as<GnState>(), gd_gjobs.as<PoseJob>(), gd_gsums.as<double>(),
                     gd_tjobs.as<TangentJob>(), gd_tsums.as<double>(), gd_dir.as<double>(), gd_dir.as<double>() + nd,
                     gd_step.as<double>(), gd_try_H.as<double>(), gd_try_ok.as<int>(), gd_try_jobs.as<PoseJob>(),
                     gd_try_sums.as<double>(), gd_try_score.as<double>(), gd_part.as<double>(), gd_ctr.as<int>(),
                     gd_ctr.as<int>() + 1};
  auto* host_c = reinterpret_cast<ClimbDev*>(reinterpret_cast<uint8_t*>(res_host) + 64);
  auto* host_s = reinterpret_cast<GnState*>(reinterpret_cast<uint8_t*>(res_host) + 64 + 32);
  static_assert(sizeof(ClimbDev) <= 32 && 64 + 32 + sizeof(GnState) <= 256, "pinned readback layout");

  ClimbDev cs{};
  while (!cs.done && !cancel) {
    const uint32_t rec0 = (uint32_t)cs.n_rec;
    const Frame fr = reg_frame(Hc, A.w, A.h, B.w, B.h);
    const auto rs_c = moving_centroid(Hc, A.w, A.h);
    const bool mirror = Hc[0] * Hc[4] - Hc[1] * Hc[3] > 0;
    ClimbArgs ca{};
    ca.hop = hop;
    ca.hop_fm = hop_fm;
    ca.n_iters = (int)n_iters;
    ca.homography = homography;
    ca.ox = fr.ox;
    ca.oy = fr.oy;
    ca.rs_c0 = rs_c[0];
    ca.rs_c1 = rs_c[1];
    // one iteration (the maps' plans depend only on the frame, the centre and the settings)
    auto issue = [&] {
      LAUNCH("climb_begin", launch_climb_begin(cdev, gb.state, ca, hi, true, st));
      if (hop) {
        const MapOut m = shift_map_sides(A, B, Hc, exact, calib, cap, min_ov, false, &fr, hi);
        ca.map_nx = (int)m.g.nx;
        ca.map_ny = (int)m.g.ny;
        ca.cell_x = (double)m.fr.w / (double)m.g.cw;
        ca.cell_y = (double)m.fr.h / (double)m.g.ch;
        LAUNCH("climb_hop", launch_climb_hop(gb, ga, cdev, ca, res.as<float>(), false, st));
        LAUNCH("pose_moments", launch_pose_moments_jobs(gb.try_jobs, nd * T, blocks_try, gb.part, gb.try_sums, st));
        LAUNCH("gn_try_scores", launch_climb_try_scores(gb, ga, st));
        LAUNCH("climb_select", launch_climb_select(gb, ga, cdev, false, st));
      }
      if (hop || !hop_fm) LAUNCH("climb_gn_block", launch_climb_gn_block(gb, ga, cdev, blocks, blocks_try, st));
      if (hop_fm) {
        LAUNCH("climb_begin", launch_climb_begin(cdev, gb.state, ca, hi, false, st));
        const RsPlan p = rs_map_dir(A, B, Hc, rs_c[0], rs_c[1], -1.0, 0.0, exact, calib, cap, rs_fwd, &fr, hi);
        rs_finish(p, false, mirror, false);
        ca.rs_n = p.n;
        ca.rs_nth = p.n_th;
        ca.rs_dlam = p.dlam;
        LAUNCH("climb_hop", launch_climb_hop(gb, ga, cdev, ca, res.as<float>(), true, st));
        LAUNCH("pose_moments", launch_pose_moments_jobs(gb.try_jobs, nd * T, blocks_try, gb.part, gb.try_sums, st));
        LAUNCH("gn_try_scores", launch_climb_try_scores(gb, ga, st));
        LAUNCH("climb_select", launch_climb_select(gb, ga, cdev, false, st));
        LAUNCH("climb_gn_block", launch_climb_gn_block(gb, ga, cdev, blocks, blocks_try, st));
      }
      LAUNCH("climb_end", launch_climb_end(cdev, ca, st));
    };
    // the graph's key: every input of the launches above that can change between climbs, and the
    // process's allocation count (a moved buffer changes it)
    auto make_key = [&] {
      std::vector<uint8_t> key;
      auto put = [&key](const auto& v) {
        const auto* q = reinterpret_cast<const uint8_t*>(&v);
        key.insert(key.end(), q, q + sizeof v);
      };
      const void* ptrs[6] = {A.stack.p, B.stack.p, A.feat.p, B.feat.p, A.il.p, B.il.p};
      const int dims[4] = {A.w, A.h, B.w, B.h};
      const int flags[8] = {hop, hop_fm, homography, exact, calib, symmetric, FftTune::cols.load(),
                            FftRowTiling::v.load()};
      put(DBuf::gen.load());
      put(fr);
      put(rs_c);
      put(mirror);
      put(ptrs);
      put(dims);
      put(flags);
      put(ga);
      put(n_iters);
      put(cap);
      put(min_ov);
      const std::string plan_set = A.mode + "|" + set.e("fft_sizes") + "|" + std::to_string(set.f("rs_smin")) + "|" +
                                   std::to_string(set.f("rs_smax"));
      key.insert(key.end(), plan_set.begin(), plan_set.end());
      return key;
    };
    bool first = true;
    for (int b = 0; b < kBatch && !cs.done && !cancel; ++b) {
      std::vector<uint8_t> key = make_key();
      const bool cached = climb_gc.exec && climb_gc.key == key;
      if (prof.on || (first && !cached))
        issue();
      else
        run_captured(climb_gc, std::move(key), false, issue);
      first = false;
      ZC_HIP(hipMemcpyAsync(host_c, cl_state.p, sizeof(ClimbDev), hipMemcpyDeviceToHost, st));
      ZC_HIP(hipMemcpyAsync(host_s, gd_state.p, sizeof(GnState), hipMemcpyDeviceToHost, st));
      prof.break_chain();
      ++counts[2];
      sync();
      cs = *host_c;
      for (int k = 0; k < 9; ++k) Hc[k] = host_s->H[k];
      cur = host_s->score;
    }
    // per batch (gclimb.zig:446-481): the pose it reached and where the climb stands
    if ((uint32_t)cs.n_rec > rec0) {
      const auto parts = score_parts(Hc, exact, symmetric);
      events.trail("ascent", parts[0], parts[1], parts[2], Hc);
      events.pose(Hc);
    }
    events.log("ascent %u/%u score %.4f  stall %u  (%u steps so far)", (uint32_t)cs.iter, n_iters, cur,
               (uint32_t)cs.stall, (uint32_t)cs.n_rec);
  }
  events.log("ascent done  score %.4f", cur);
  H = Hc;
  ClimbResult r{};
  for (int i = 0; i < 9; ++i) r.H[i] = Hc[i];
  r.score = cur;
  r.iterations = (uint32_t)cs.iter;
  r.moved = (cs.n_rec > 0 && cur > start) ? 1u : 0u;
  return r;
}

}  // namespace zc
