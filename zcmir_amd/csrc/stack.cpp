// Stacks of uneven slice pairs registered together: every slice's features stay resident, an
// SMI shift map per slice gives its translation, then Gauss–Newton runs on all slices at once
// (each iteration one overlap-sum launch and one tangent launch over every slice's poses).
// The role of the reference's register_stack (python/zcmir/api.py:522-580), which registers
// slices one after another; here the slices are independent and batched.
#include <cstring>

#include "engine.hpp"
#include "lie.hpp"

namespace zc {

void Engine::stack_set(uint32_t n, const float* const* mv, const uint32_t* mwh, const float* const* fx,
                       const uint32_t* fwh) {
  for (auto& s : stack) release_side(*s);
  stack.clear();
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 2; ++k) {
      const float* data = k == 0 ? mv[i] : fx[i];
      const int w = (int)(k == 0 ? mwh[2 * i] : fwh[2 * i]), h = (int)(k == 0 ? mwh[2 * i + 1] : fwh[2 * i + 1]);
      if (w < 8 || h < 8 || !data || (int64_t)w * h > kMaxImagePixels)
        throw std::runtime_error("BadImage: slice " + std::to_string(i));
      auto sd = std::make_unique<Side>();
      const size_t px = (size_t)w * h;
      sd->w = w;
      sd->h = h;
      sd->gray.ensure(px * 4);
      sd->rank.ensure(px * 4);
      sd->feat.ensure(px * 16);
      sd->mag.ensure(px * 4);
      sd->stack.ensure(px * 20);
      ZC_HIP(hipMemcpyAsync(sd->gray.p, data, px * 4, hipMemcpyHostToDevice, st));
      prof.break_chain();
      prepare(*sd, false);
      sd->has = true;
      stack.push_back(std::move(sd));
    }
  }
  sync();
}

std::vector<GnOut> Engine::stack_register(const std::vector<Mat3>& H0, uint32_t iters, int init,
                                          std::vector<Mat3>* Hinit, uint32_t* iterations) {
  const size_t n = stack.size() / 2;
  if (n == 0) throw std::runtime_error("NoImages: set a stack first");
  if (H0.size() != n) throw std::runtime_error("BadPose: one start pose per slice");
  const bool exact = set.b("exact"), symmetric = set.b("symmetric");
  const bool homography = set.e("group") == "homography";
  std::vector<Mat3> Hs = H0;
  if (init == 2) {
    const SweepGrid gv = sweep_grid(-12, 12, 7, 0.88, 1.12, 5, 0, 1, 0, 1);
    for (size_t i = 0; i < n; ++i) {
      SweepOpts o{};
      o.H0 = Hs[i];
      o.gv = &gv;
      o.grid = 64;
      o.mode = exact ? 1 : 0;
      o.n_peak = 4;
      o.batch = kSweepBatch;
      o.min_overlap = std::min(std::max(set.f("min_overlap"), 0.0), 0.5);
      const SweepOut r = sweep_sides(*stack[2 * i], *stack[2 * i + 1], o);
      const SweepPeak* best = nullptr;
      for (const SweepPeak& p : r.peaks)
        if (!best || p.score > best->score) best = &p;
      if (best) Hs[i] = best->H;
    }
  } else if (init == 1) {
    // the translation of each slice from its shift map at the start pose (engine.zig:795-819:
    // the map's peak displaces the moving image by +d; the corrected pose is T(−d) H)
    const uint32_t cap = set.u("map_res");
    const double min_ov = std::min(std::max(set.f("min_overlap"), 0.0), 0.5);
    for (size_t i = 0; i < n; ++i) {
      const MapOut m = shift_map_sides(*stack[2 * i], *stack[2 * i + 1], Hs[i], exact, false, cap, min_ov, true);
      ShiftResult r;
      fill_shift_result(m, &r);
      if (r.peak > 0) Hs[i] = mul3(translate(-r.dx, -r.dy), Hs[i]);
    }
  }
  if (Hinit) *Hinit = Hs;
  std::vector<PairSides> pr(n);
  for (size_t i = 0; i < n; ++i) pr[i] = {stack[2 * i].get(), stack[2 * i + 1].get()};
  const auto amps = geo_amps(set.f("g_a0"), set.f("g_decay"), set.u("g_steps") > 0 ? set.u("g_steps") : 7);
  return gn_batch(pr, Hs, iters, homography, exact, symmetric, amps, iterations);
}

}  // namespace zc
