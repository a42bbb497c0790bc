// Correlation plans of the SMI maps (smi.zig:77-159 CorrPlan.make), built at compile time.
//
// Real planes of a premultiplied stack [w, w f0..w f3]: 0 the weight, 1-4 the features, 5-14
// their products w f_i f_j (upper triangle, smi.wgsl:643-649). An exact map correlates the 45
// pairs listed in `list` (global whitening: the first 25). With three features every plane and
// correlation of the fourth feature is left out (28 / 16 correlations of 10 / 4 planes); kept
// planes move up into slots, correlations keep their numbers. Two real planes share a complex
// FFT input (slot 2p in re, 2p+1 in im) and two correlations share a complex inverse FFT
// (the first in re, the second in im).
#pragma once

namespace zc {

struct CorrPlan {
  int n_real;        // real planes (slots)
  int n_cplx;        // complex spectra per side
  int n_corr;        // correlations kept
  int n_out;         // complex inverse FFTs
  int slot_plane[15];
  int plane_slot[15];       // -1: left out
  int ca[45], cb[45];       // kept correlation k: slot of A, slot of B
  int cnum[45];             // kept correlation k: its number (0..44)
  int kept_of[45];          // number → kept index, -1 if left out
};

constexpr CorrPlan make_plan(bool exact, bool three) {
  CorrPlan p{};
  int la[45] = {}, lb[45] = {};
  int len = 0;
  la[len] = 0; lb[len] = 0; ++len;
  for (int i = 0; i < 4; ++i) { la[len] = 1 + i; lb[len] = 0; ++len; }
  for (int j = 0; j < 4; ++j) { la[len] = 0; lb[len] = 1 + j; ++len; }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) { la[len] = 1 + i; lb[len] = 1 + j; ++len; }
  if (exact) {
    for (int k = 0; k < 10; ++k) { la[len] = 5 + k; lb[len] = 0; ++len; }
    for (int k = 0; k < 10; ++k) { la[len] = 0; lb[len] = 5 + k; ++len; }
  }
  const int n_planes = exact ? 15 : 5;
  int n_real = 0;
  for (int pl = 0; pl < 15; ++pl) p.plane_slot[pl] = -1;
  for (int pl = 0; pl < n_planes; ++pl) {
    const bool fourth = pl == 4 || pl == 8 || pl == 11 || pl == 13 || pl == 14;
    if (three && fourth) continue;
    p.plane_slot[pl] = n_real;
    p.slot_plane[n_real] = pl;
    ++n_real;
  }
  p.n_real = n_real;
  p.n_cplx = (n_real + 1) / 2;
  for (int c = 0; c < 45; ++c) p.kept_of[c] = -1;
  int kept = 0;
  for (int c = 0; c < len; ++c) {
    const int sa = p.plane_slot[la[c]], sb = p.plane_slot[lb[c]];
    if (sa < 0 || sb < 0) continue;
    p.ca[kept] = sa;
    p.cb[kept] = sb;
    p.cnum[kept] = c;
    p.kept_of[c] = kept;
    ++kept;
  }
  p.n_corr = kept;
  p.n_out = (kept + 1) / 2;
  return p;
}

// plan id: exact + 2 · three
constexpr CorrPlan kPlans[4] = {make_plan(false, false), make_plan(true, false), make_plan(false, true),
                                make_plan(true, true)};

template <int ID>
struct PlanT {
  static constexpr CorrPlan p = kPlans[ID];
};

}  // namespace zc
