// Correlation plans of the SMI maps (smi.zig:77-159 CorrPlan.make), built at compile time.
//
// Real planes of a premultiplied stack [w, w f0..w f3]: 0 the weight, 1-4 the features, 5-14
// their products w f_i f_j (upper triangle, smi.wgsl:643-649). An exact map correlates the 45
// pairs listed in `list` (global whitening: the first 25). With three features every plane and
// correlation of the fourth feature is left out (28 / 16 correlations of 10 / 4 planes); kept
// planes move up into slots, correlations keep their numbers. Two real planes share a complex
// FFT input (slot 2p in re, 2p+1 in im) and two correlations share a complex inverse FFT
// (the first in re, the second in im).
//
// Rectangle plans (`rect`): when the fixed side's weight plane is exactly the indicator of a
// rectangle in cell space (the fixed image pasted, cells = canvas pixels), every correlation
// (a, 0) of a moving plane with that weight is the sum of plane a over the rectangle shifted by
// the lag, read from a summed-area table of plane a in O(1) per lag. Those correlations leave the
// FFT path; the moving side then transforms only the planes of the remaining correlations (w and
// the features) and the fixed side only its planes other than w.
//
// Translation plans (`trans`, rectangle plans whose moving stack lands on the canvas by an integer
// translation): the moving weight plane is then the indicator of a rectangle too, so every
// correlation (0, b) is the sum of fixed plane b over the moving rectangle shifted by minus the
// lag — a read of the fixed side's tables. Only the cross terms (i, j) of the features stay in the
// FFT path: 16 correlations of 4 + 4 planes (9 of 3 + 3 with three features).
#pragma once

namespace zc {

struct CorrPlan {
  int n_real;               // side A: real planes (slots) transformed
  int n_cplx;               // side A: complex spectra
  int n_corr;               // correlations through the FFT
  int n_out;                // complex inverse FFTs
  int slot_plane[15];       // side A: slot → plane
  int plane_slot[15];       // side A: plane → slot, -1 not transformed
  int b_n_real, b_n_cplx;   // side B
  int b_slot_plane[15], b_plane_slot[15];
  int ca[45], cb[45];       // FFT correlation k: slot of A, slot of B
  int cnum[45];             // FFT correlation k: its number (0..44)
  int kept_of[45];          // number → FFT correlation index, -1 if not through the FFT
  int n_sat;                // rectangle-sum planes of side A
  int sat_plane[15];        // rectangle-sum index → plane of A
  int sat_of[45];           // number → rectangle-sum index, -1 if not one
  int n_bsat;               // translation plans: rectangle-sum planes of side B
  int bsat_plane[15];       // B rectangle-sum index → plane of B
  int bsat_of[45];          // number → B rectangle-sum index, -1 if not one
};

constexpr CorrPlan make_plan(bool exact, bool three, bool rect, bool trans = false) {
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
  bool valid[15] = {};
  for (int pl = 0; pl < n_planes; ++pl) {
    const bool fourth = pl == 4 || pl == 8 || pl == 11 || pl == 13 || pl == 14;
    valid[pl] = !(three && fourth);
  }
  // which correlations go where, and which planes each side transforms
  bool use_a[15] = {}, use_b[15] = {};
  for (int c = 0; c < 45; ++c) {
    p.kept_of[c] = -1;
    p.sat_of[c] = -1;
    p.bsat_of[c] = -1;
  }
  int sat_idx[15] = {}, bsat_idx[15] = {};
  for (int pl = 0; pl < 15; ++pl) sat_idx[pl] = bsat_idx[pl] = -1;
  for (int c = 0; c < len; ++c) {
    if (!valid[la[c]] || !valid[lb[c]]) continue;
    if (rect && lb[c] == 0) {
      if (sat_idx[la[c]] < 0) {
        sat_idx[la[c]] = p.n_sat;
        p.sat_plane[p.n_sat++] = la[c];
      }
      p.sat_of[c] = sat_idx[la[c]];
    } else if (trans && la[c] == 0) {
      if (bsat_idx[lb[c]] < 0) {
        bsat_idx[lb[c]] = p.n_bsat;
        p.bsat_plane[p.n_bsat++] = lb[c];
      }
      p.bsat_of[c] = bsat_idx[lb[c]];
    } else {
      use_a[la[c]] = true;
      use_b[lb[c]] = true;
    }
  }
  for (int pl = 0; pl < 15; ++pl) {
    p.plane_slot[pl] = -1;
    p.b_plane_slot[pl] = -1;
  }
  for (int pl = 0; pl < n_planes; ++pl) {
    if (use_a[pl]) {
      p.plane_slot[pl] = p.n_real;
      p.slot_plane[p.n_real++] = pl;
    }
    if (use_b[pl]) {
      p.b_plane_slot[pl] = p.b_n_real;
      p.b_slot_plane[p.b_n_real++] = pl;
    }
  }
  p.n_cplx = (p.n_real + 1) / 2;
  p.b_n_cplx = (p.b_n_real + 1) / 2;
  int kept = 0;
  for (int c = 0; c < len; ++c) {
    if (!valid[la[c]] || !valid[lb[c]] || p.sat_of[c] >= 0 || p.bsat_of[c] >= 0) continue;
    p.ca[kept] = p.plane_slot[la[c]];
    p.cb[kept] = p.b_plane_slot[lb[c]];
    p.cnum[kept] = c;
    p.kept_of[c] = kept;
    ++kept;
  }
  p.n_corr = kept;
  p.n_out = (kept + 1) / 2;
  return p;
}

// plan id: exact + 2 · three + 4 · rect + 4 · trans (translation plans are rectangle plans: 8..11)
constexpr CorrPlan kPlans[12] = {make_plan(false, false, false), make_plan(true, false, false),
                                 make_plan(false, true, false),  make_plan(true, true, false),
                                 make_plan(false, false, true),  make_plan(true, false, true),
                                 make_plan(false, true, true),   make_plan(true, true, true),
                                 make_plan(false, false, true, true), make_plan(true, false, true, true),
                                 make_plan(false, true, true, true),  make_plan(true, true, true, true)};

// Without rectangle sums both sides transform the same planes.
static_assert(kPlans[1].n_real == 15 && kPlans[1].b_n_real == 15 && kPlans[1].n_corr == 45, "exact plan");
static_assert(kPlans[3].n_real == 10 && kPlans[3].n_corr == 28, "three-feature plan");
static_assert(kPlans[5].n_sat == 15 && kPlans[5].n_real == 5 && kPlans[5].b_n_real == 14 && kPlans[5].n_corr == 30,
              "exact rectangle plan");
static_assert(kPlans[7].n_sat == 10 && kPlans[7].n_real == 4 && kPlans[7].b_n_real == 9 && kPlans[7].n_corr == 18,
              "three-feature rectangle plan");
static_assert(kPlans[4].n_sat == 5 && kPlans[4].n_corr == 20, "global rectangle plan");
static_assert(kPlans[9].n_sat == 15 && kPlans[9].n_bsat == 14 && kPlans[9].n_corr == 16 && kPlans[9].n_real == 4 &&
                  kPlans[9].b_n_real == 4,
              "exact translation plan");
static_assert(kPlans[11].n_sat == 10 && kPlans[11].n_bsat == 9 && kPlans[11].n_corr == 9 && kPlans[11].n_real == 3,
              "three-feature translation plan");
static_assert(kPlans[8].n_sat == 5 && kPlans[8].n_bsat == 4 && kPlans[8].n_corr == 16, "global translation plan");

// The planes of one side's summed-area tables: set s < 16 is side A of plan s, 16 + s side B.
struct SatSet {
  int n;
  int plane[15];
};

// sid < 16: side A's tables of plan sid; 16 + plan: side B's; kImageTables: an image's own tables,
// table k of plane k + 1 for the 14 planes other than the weight (translation plans read both
// sides' sums from these; an image's weight plane is 1 on every pixel, so its rectangle sums are
// the rectangles' areas and need no table).
inline constexpr int kImageTables = 32;
inline constexpr int kImageTableCount = 14;

constexpr SatSet sat_set(int sid) {
  SatSet t{};
  if (sid == kImageTables) {
    t.n = kImageTableCount;
    for (int k = 0; k < kImageTableCount; ++k) t.plane[k] = k + 1;
    return t;
  }
  const CorrPlan& p = kPlans[sid & 15];
  t.n = sid < 16 ? p.n_sat : p.n_bsat;
  for (int k = 0; k < 15; ++k) t.plane[k] = sid < 16 ? p.sat_plane[k] : p.bsat_plane[k];
  return t;
}

// The side tables the pack kernels use: index 2 · plan + side (side B of a plan without
// rectangle sums shares side A's table).
struct PackTable {
  int n_real, n_cplx;
  int slot_plane[15];
};

constexpr PackTable pack_table(int ti) {
  const CorrPlan& p = kPlans[ti / 2];
  PackTable t{};
  const bool b = (ti & 1) != 0;
  t.n_real = b ? p.b_n_real : p.n_real;
  t.n_cplx = b ? p.b_n_cplx : p.n_cplx;
  for (int s = 0; s < 15; ++s) t.slot_plane[s] = b ? p.b_slot_plane[s] : p.slot_plane[s];
  return t;
}

}  // namespace zc
