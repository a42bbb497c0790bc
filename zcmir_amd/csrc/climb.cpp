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

}  // namespace zc
