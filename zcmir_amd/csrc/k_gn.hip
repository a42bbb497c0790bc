// Gauss–Newton batches without the host (gnBatch, search.zig:539-614): every pose's iteration —
// gradient and Gauss–Newton matrix, the tries, their scores, the selection — runs as a fixed
// sequence of launches over device state, so one captured graph serves every iteration. The FP64
// algebra is gn_dev.hpp's, the same arithmetic as the host loop in gn.cpp.
#include "gn_dev.hpp"
#include "kdev.hpp"

namespace zc {

namespace {

__device__ __forceinline__ void job_of(PoseJob& j, const float4* fa, int wa, int ha, const float4* fb, int wb, int hb,
                                       const double* Hi) {
  j.fa = fa;
  j.fb = fb;
  j.wa = wa;
  j.ha = ha;
  j.wb = wb;
  j.hb = hb;
  for (int k = 0; k < 9; ++k) j.Hi[k] = Hi[k];
}

// a job with nothing to sum (moving side narrower than 5 px: the kernels skip it, its sums are 0)
__device__ __forceinline__ void idle_job(PoseJob& j) {
  j.fa = j.fb = nullptr;
  j.wa = j.ha = 0;
  j.wb = j.hb = 1;
  for (int k = 0; k < 9; ++k) j.Hi[k] = 0;
}

__device__ __forceinline__ double fin(double v) { return isfinite(v) ? v : 0.0; }

// the two directions of pose i: forward (moving → fixed) H, inverse H⁻¹, as in pair_scores
__device__ __forceinline__ void pose_jobs(const GnPair& p, const double* H, bool symmetric, PoseJob& f, PoseJob* b) {
  double Hinv[9], Hinv2[9];
  gd::inv3(H, Hinv);
  job_of(f, p.fa, p.wa, p.ha, p.fb, p.wb, p.hb, Hinv);
  if (symmetric && b) {
    gd::inv3(Hinv, Hinv2);   // the inverse direction's pose is H⁻¹; its job holds (H⁻¹)⁻¹
    job_of(*b, p.fb, p.wb, p.hb, p.fa, p.wa, p.ha, Hinv2);
  }
}

}  // namespace

// 1. the gradient pass's overlap-sum jobs of every active pose (inactive: idle jobs)
__global__ __launch_bounds__(64) void k_gn_grad_jobs(const GnPair* __restrict__ pairs, const GnState* __restrict__ state, GnArgs a,
                               PoseJob* jobs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  if (!state[i].active) {
    idle_job(jobs[i]);
    if (a.symmetric) idle_job(jobs[a.n + i]);
    return;
  }
  pose_jobs(pairs[i], state[i].H, a.symmetric, jobs[i], a.symmetric ? &jobs[a.n + i] : nullptr);
}

// 2. per direction: the score's coefficients from its sums, and its tangent job (directions with
// fewer than min_n overlapping pixels: an idle job, gradient 0)
__global__ __launch_bounds__(64) void k_gn_coefs(const GnPair* __restrict__ pairs, const GnState* __restrict__ state, GnArgs a,
                           const double* __restrict__ sums, double* dir_score, double* dir_n, TangentJob* tjobs) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int nd = a.symmetric ? 2 * a.n : a.n;
  if (d >= nd) return;
  const int i = d < a.n ? d : d - a.n;
  const bool back = d >= a.n;
  TangentJob& j = tjobs[d];
  j.fa = j.fb = nullptr;
  j.wa = j.ha = 0;
  j.wb = j.hb = 1;
  j.hess = 1;
  dir_score[d] = 0;
  dir_n[d] = 0;
  if (!state[i].active) return;
  gd::DCoefs co;
  gd::score_coefs(sums + (int64_t)d * 45, a.exact, a.ridge, a.min_n, co);
  dir_score[d] = co.score;
  dir_n[d] = co.n;
  if (!(co.n >= a.min_n)) return;
  const GnPair& p = pairs[i];
  double H[9], Hi[9];
  if (back) {
    gd::inv3(state[i].H, H);   // the direction's pose H⁻¹
  } else {
    for (int k = 0; k < 9; ++k) H[k] = state[i].H[k];
  }
  gd::inv3(H, Hi);
  j.fa = back ? p.fb : p.fa;
  j.fb = back ? p.fa : p.fb;
  j.wa = back ? p.wb : p.wa;
  j.ha = back ? p.hb : p.ha;
  j.wb = back ? p.wa : p.wb;
  j.hb = back ? p.ha : p.hb;
  for (int k = 0; k < 9; ++k) j.Hi[k] = Hi[k];
  const int nk = a.homography ? 8 : 6;
  for (int k = 0; k < 8; ++k) {
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, M[9];
    if (k < nk) gd::generator(k, a.homography, true, (double)j.wb, (double)j.hb, G);
    gd::mul3(Hi, G, M);
    for (int q = 0; q < 9; ++q) j.Mk[9 * k + q] = (float)-M[q];
  }
  for (int q = 0; q < 16; ++q) {
    j.W[q] = (float)co.W[q];
    j.Mh[q] = (float)co.Mh[q];
    j.Q[q] = (float)co.Q[q];
  }
  for (int q = 0; q < 4; ++q) {
    j.ma[q] = (float)co.ma[q];
    j.mb[q] = (float)co.mb[q];
  }
  for (int q = 0; q < 45; ++q) j.g[q] = (float)co.g[q];
}

// 3a. per pose (one workgroup of 64): the gradient and matrix — the symmetric mean with the
// inverse direction pulled back, one matrix entry per thread — the pose's score, the unit ascent
// direction and the Gauss–Newton step (A + 1e-4 max(A_ii, 1e-12)) δ = ½ g
__global__ __launch_bounds__(64) void k_gn_step(const GnPair* __restrict__ pairs, GnState* state, GnArgs a,
                                                const double* __restrict__ tsums, const double* __restrict__ dir_score,
                                                const double* __restrict__ dir_n, double* step_out) {
  const int i = blockIdx.x, t = threadIdx.x;
  if (i >= a.n) return;
  GnState& s = state[i];
  if (!s.active) return;   // uniform over the workgroup
  const int nk = a.homography ? 8 : 6;
  __shared__ double g[8], h[8][8], gb[8], hb[8][8], PT[8][8];
  // one direction's gradient and matrix from its 44 sums (0 below the overlap floor)
  auto dir_grad = [&](int d, double* gg, double (*hh)[8]) {
    const bool ok = dir_n[d] >= a.min_n;
    const double* acc = tsums + (int64_t)d * 44;
    const int x = t >> 3, y = t & 7;
    double v = 0;
    if (ok && x < nk && y < nk) {
      const int lo = x < y ? x : y, hi = x < y ? y : x;
      v = acc[8 + lo * 8 - lo * (lo - 1) / 2 + (hi - lo)];   // row lo of the upper triangle, column hi
    }
    hh[x][y] = v;
    if (t < 8) gg[t] = ok && t < nk ? acc[t] : 0.0;
  };
  dir_grad(i, g, h);
  if (a.symmetric) {
    dir_grad(a.n + i, gb, hb);
    if (t == 0) {
      const GnPair& p = pairs[i];
      double P[8][8];
      gd::adjoint_cols(s.H, a.homography, p.wa, p.ha, p.wb, p.hb, P);
      for (int x = 0; x < 8; ++x)
        for (int y = 0; y < 8; ++y) PT[x][y] = P[x][y];
    }
  }
  __syncthreads();
  double score = dir_score[i];
  if (a.symmetric) {
    score = 0.5 * (score + dir_score[a.n + i]);
    const int x = t >> 3, c = t & 7;
    double hm = 0, pulled = 0;
    if (x < nk && c < nk) {
      double v = 0;
      for (int u = 0; u < nk; ++u)
        for (int w = 0; w < nk; ++w) v += PT[x][u] * hb[u][w] * PT[c][w];
      hm = 0.5 * (h[x][c] + v);
    }
    if (t < nk) {
      for (int j = 0; j < nk; ++j) pulled += PT[t][j] * gb[j];
      pulled = 0.5 * (g[t] + pulled);
    }
    __syncthreads();
    h[x][c] = hm;
    if (t < nk) g[t] = pulled;
    __syncthreads();
  }
  // (A + 1e-4 max(A_ii, 1e-12)) δ = ½ g by the whole workgroup
  __shared__ double A[64], half[8], fac[8];
  __shared__ int piv[9];
  {
    const int x = t / nk, c = t - (t / nk) * nk;
    if (x < nk) A[x * nk + c] = h[x][c] + (x == c ? 1e-4 * fmax(h[x][c], 1e-12) : 0.0);
    if (t < nk) half[t] = 0.5 * g[t];
  }
  __syncthreads();
  const bool solved = gd::solve_linear_wg(A, half, nk, t, piv, fac);
  if (t != 0) return;
  s.score = score;
  // step_out[i]: u[8], step[8], |step|, flags (bit 0: ascent tries, bit 1: step tries)
  double* o = step_out + (int64_t)i * 18;
  double u[8] = {0, 0, 0, 0, 0, 0, 0, 0}, nrm = 0;
  for (int k = 0; k < nk; ++k) {
    u[k] = g[k];
    nrm += u[k] * u[k];
  }
  nrm = sqrt(nrm);
  if (!(nrm < 1e-18))
    for (int k = 0; k < nk; ++k) u[k] /= nrm;
  double nd = 0;
  for (int k = 0; k < nk; ++k) nd += u[k] * u[k];
  int flags = (a.n_amps > 0 && sqrt(nd) > 1e-12) ? 1 : 0;
  double step[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double sn = 0;
  if (solved) {
    flags |= 2;
    for (int q = 0; q < nk; ++q) {
      step[q] = half[q];
      sn += step[q] * step[q];
    }
    sn = sqrt(sn);
  }
  for (int k = 0; k < 8; ++k) {
    o[k] = u[k];
    o[8 + k] = step[k];
  }
  o[16] = sn;
  o[17] = (double)flags;
}

// 3b. one thread per try slot: slots [0, n_amps) the ascent ladder along u, then five lengths of
// the step (×{0.5, 1, 2, 4, 8}, capped at norm 0.25); the try's pose and its overlap-sum jobs, or
// an idle slot
__global__ __launch_bounds__(64) void k_gn_try_jobs(const GnPair* __restrict__ pairs, const GnState* __restrict__ state,
                                                    GnArgs a, const double* __restrict__ step_in, double* try_H,
                                                    int* try_ok, PoseJob* jobs) {
  const int T = a.n_try;
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= a.n * T) return;
  const int i = slot / T, k = slot % T;
  PoseJob& jf = jobs[slot];
  PoseJob* jb = a.symmetric ? &jobs[(int64_t)a.n * T + slot] : nullptr;
  idle_job(jf);
  if (jb) idle_job(*jb);
  try_ok[slot] = 0;
  const GnState& s = state[i];
  if (!s.active) return;
  const double* o = step_in + (int64_t)i * 18;
  const int flags = (int)o[17];
  const int nk = a.homography ? 8 : 6;
  double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (k < a.n_amps) {
    if (!(flags & 1)) return;
    for (int q = 0; q < nk; ++q) x[q] = a.amps[k] * o[q];
  } else {
    if (!(flags & 2)) return;
    const double mults[5] = {0.5, 1, 2, 4, 8};
    const double m = mults[k - a.n_amps], sn = o[16];
    const double sc = m * (sn * m > 0.25 ? 0.25 / (sn * m) : 1);
    for (int q = 0; q < nk; ++q) x[q] = sc * o[8 + q];
  }
  const GnPair& p = pairs[i];
  double* Ht = try_H + (int64_t)slot * 9;
  gd::compose_n(s.H, x, a.homography, p.wb, p.hb, Ht);
  try_ok[slot] = 1;
  pose_jobs(p, Ht, a.symmetric, jf, jb);
}

// 4. every try direction's score from its sums
__global__ __launch_bounds__(64) void k_gn_try_scores(const GnState* __restrict__ state, GnArgs a, const double* __restrict__ sums,
                                const int* __restrict__ try_ok, double* score) {
  const int T = a.n_try;
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int nd = (a.symmetric ? 2 : 1) * a.n * T;
  if (d >= nd) return;
  const int slot = d % (a.n * T), i = slot / T, k = slot % T;
  score[d] = 0;
  if (!state[i].active || !try_ok[slot]) return;
  (void)k;
  gd::DCoefs co;
  gd::score_coefs(sums + (int64_t)d * 45, a.exact, a.ridge, a.min_n, co);
  score[d] = fin(co.score);
}

// 5. per pose: a try replaces the pose when it beats it (in order, each against the pose as it
// stands); poses that did not move leave the batch
__global__ __launch_bounds__(64) void k_gn_select(GnState* state, GnArgs a, const double* __restrict__ try_H,
                            const int* __restrict__ try_ok, const double* __restrict__ score, int* moved_flag) {
  const int T = a.n_try;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  GnState& s = state[i];
  if (!s.active) return;
  bool moved = false;
  int t = 0;
  for (int k = 0; k < T; ++k) {
    if (!try_ok[(int64_t)i * T + k]) continue;
    ++t;
    const double f = score[(int64_t)i * T + k];
    const double v = a.symmetric ? 0.5 * (f + score[(int64_t)(a.n + i) * T + k]) : f;
    if (gd::beats(v, s.score)) {
      for (int q = 0; q < 9; ++q) s.H[q] = try_H[((int64_t)i * T + k) * 9 + q];
      s.score = v;
      moved = true;
    }
  }
  if (moved) s.moved = 1;
  s.active = moved ? 1 : 0;
  if (t > 0) atomicAdd(moved_flag, 1);
}

// 6. the iteration count: an iteration counts when some pose had a try
__global__ void k_gn_tick(int* moved_flag, int* iters_done) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (*moved_flag > 0) ++*iters_done;
    *moved_flag = 0;
  }
}

void launch_gn_iteration(const GnBuffers& b, const GnArgs& a, int blocks_grad, int blocks_try, hipStream_t st) {
  const int n = a.n, nd = (a.symmetric ? 2 : 1) * n, T = a.n_try;
  const int g1 = (n + 63) / 64;
  k_gn_grad_jobs<<<g1, 64, 0, st>>>(b.pairs, b.state, a, b.gjobs);
  launch_pose_moments_jobs(b.gjobs, nd, blocks_grad, b.part, b.gsums, st);
  k_gn_coefs<<<(nd + 63) / 64, 64, 0, st>>>(b.pairs, b.state, a, b.gsums, b.dir_score, b.dir_n, b.tjobs);
  launch_tangent_jobs(b.tjobs, nd, blocks_grad, b.part, b.tsums, st);
  k_gn_step<<<n, 64, 0, st>>>(b.pairs, b.state, a, b.tsums, b.dir_score, b.dir_n, b.step);
  k_gn_try_jobs<<<(n * T + 63) / 64, 64, 0, st>>>(b.pairs, b.state, a, b.step, b.try_H, b.try_ok, b.try_jobs);
  launch_pose_moments_jobs(b.try_jobs, nd * T, blocks_try, b.part, b.try_sums, st);
  k_gn_try_scores<<<(nd * T + 63) / 64, 64, 0, st>>>(b.state, a, b.try_sums, b.try_ok, b.try_score);
  k_gn_select<<<g1, 64, 0, st>>>(b.state, a, b.try_H, b.try_ok, b.try_score, b.moved_flag);
  k_gn_tick<<<1, 64, 0, st>>>(b.moved_flag, b.iters_done);
}

// ── the climb ────────────────────────────────────────────────────────────────────────────
// One thread each: the climb's decisions on pose 0 of a one-pose batch, gated by c->done so that
// a replayed iteration after the end changes nothing.

__global__ void k_climb_begin(ClimbDev* c, GnState* state, ClimbArgs a, double* hi, int reset_moved) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (c->done) return;
  if (reset_moved) c->moved = 0;
  // canvas pose T(−ox, −oy) · H and its inverse (geometry.hpp canvas_h, warp_inverse)
  const double T[9] = {1, 0, -(double)a.ox, 0, 1, -(double)a.oy, 0, 0, 1};
  double Hc[9];
  gd::mul3(T, state[0].H, Hc);
  const double det = Hc[0] * (Hc[4] * Hc[8] - Hc[5] * Hc[7]) - Hc[1] * (Hc[3] * Hc[8] - Hc[5] * Hc[6]) +
                     Hc[2] * (Hc[3] * Hc[7] - Hc[4] * Hc[6]);
  if (fabs(det) >= 1e-12) {
    hi[0] = (Hc[4] * Hc[8] - Hc[5] * Hc[7]) / det;
    hi[1] = (Hc[2] * Hc[7] - Hc[1] * Hc[8]) / det;
    hi[2] = (Hc[1] * Hc[5] - Hc[2] * Hc[4]) / det;
    hi[3] = (Hc[5] * Hc[6] - Hc[3] * Hc[8]) / det;
    hi[4] = (Hc[0] * Hc[8] - Hc[2] * Hc[6]) / det;
    hi[5] = (Hc[2] * Hc[3] - Hc[0] * Hc[5]) / det;
    hi[6] = (Hc[3] * Hc[7] - Hc[4] * Hc[6]) / det;
    hi[7] = (Hc[1] * Hc[6] - Hc[0] * Hc[7]) / det;
    hi[8] = (Hc[0] * Hc[4] - Hc[1] * Hc[3]) / det;
  } else {
    for (int k = 0; k < 9; ++k) hi[k] = 0;
  }
}

// hop_cand: T(−dx, −dy) · H from the shift map's peak; rs_cand: simAbout(c, dθ, dσ) · H from the
// roto-scale map's (none within 2e-4 of no change)
__global__ __launch_bounds__(64) void k_climb_hop(const GnPair* __restrict__ pairs, GnState* state, GnArgs ga,
                                                  ClimbDev* c, ClimbArgs a, const float* __restrict__ res, int rs,
                                                  double* try_H, int* try_ok, PoseJob* jobs) {
  const int T = ga.n_try, k = threadIdx.x;
  if (k >= T) return;
  PoseJob* jb = ga.symmetric ? &jobs[T + k] : nullptr;
  idle_job(jobs[k]);
  if (jb) idle_job(*jb);
  try_ok[k] = 0;
  if (k != 0) return;   // slot 0 (this thread's own, idled just above) may hold the hop
  GnState& s = state[0];
  s.active = c->done ? 0 : 1;
  if (c->done) return;
  const uint32_t idx = __float_as_uint(res[1]);
  double Ht[9];
  if (!rs) {
    const uint64_t nn = (uint64_t)a.map_nx * a.map_ny;
    if (idx >= nn) return;
    const int64_t py = idx / a.map_nx, px = idx % a.map_nx;
    const int64_t lx = px <= a.map_nx / 2 ? px : px - (int64_t)a.map_nx;
    const int64_t ly = py <= a.map_ny / 2 ? py : py - (int64_t)a.map_ny;
    const double dx = (double)lx * a.cell_x, dy = (double)ly * a.cell_y;
    const double Tm[9] = {1, 0, -dx, 0, 1, -dy, 0, 0, 1};
    gd::mul3(Tm, s.H, Ht);
  } else {
    const int n = a.rs_n;
    if (idx >= (uint64_t)n * n) return;
    const int64_t py = idx / n, px = idx % n;
    const int64_t lx = px <= n / 2 ? px : px - n;
    const int64_t ly = py <= n / 2 ? py : py - n;
    const double dth = -((double)lx / (double)max(a.rs_nth, 1)) * 2.0 * M_PI;
    const double dsg = -(double)ly * a.rs_dlam;
    if (fabs(dth) < 2e-4 && fabs(dsg) < 2e-4) return;
    // simAbout (search.hpp sim_about): X ↦ c + e^dσ R(dθ) (X − c)
    const double sc = exp(dsg);
    const double m00 = sc * cos(dth), m01 = -sc * sin(dth), m10 = sc * sin(dth), m11 = sc * cos(dth);
    const double S[9] = {m00, m01, a.rs_c0 - m00 * a.rs_c0 - m01 * a.rs_c1,
                         m10, m11, a.rs_c1 - m10 * a.rs_c0 - m11 * a.rs_c1, 0, 0, 1};
    gd::mul3(S, s.H, Ht);
    gd::project_group(Ht, a.homography != 0);
  }
  for (int q = 0; q < 9; ++q) try_H[q] = Ht[q];
  try_ok[0] = 1;
  pose_jobs(pairs[0], Ht, ga.symmetric, jobs[0], jb);
}

__global__ void k_climb_select(GnState* state, GnArgs ga, ClimbDev* c, const double* __restrict__ try_H,
                               const int* __restrict__ try_ok, const double* __restrict__ score, int gn) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  GnState& s = state[0];
  if (c->done || !s.active) return;
  const int T = ga.n_try;
  int best = -1;
  double bs = s.score;
  for (int k = 0; k < T; ++k) {
    if (!try_ok[k]) continue;
    const double f = score[k];
    const double v = ga.symmetric ? 0.5 * (f + score[T + k]) : f;
    if (gd::beats(v, s.score) && (best < 0 || v > bs)) {
      best = k;
      bs = v;
    }
  }
  if (best < 0) {
    if (gn) {
      s.active = 0;   // the block ends at a step that beats nothing
      c->step_ok = 1;
    }
    return;
  }
  for (int q = 0; q < 9; ++q) s.H[q] = try_H[best * 9 + q];
  s.score = bs;
  s.moved = 1;
  c->moved = 1;
  c->step_ok = 0;
  ++c->n_rec;
}

// A block at a pose whose Gauss–Newton step already beat nothing would try the same five poses
// with the same scores: it is skipped (the host loop runs it and finds nothing).
__global__ void k_climb_block(GnState* state, const ClimbDev* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) state[0].active = (c->done || c->step_ok) ? 0 : 1;
}

// iter_end: a move resets the stall count; two iterations in a row without one end the climb
// (one when no hop is on), as does n_iters. An iteration without a move leaves the pose, and with
// it the frame, the maps and the tries of the next iteration, unchanged, so the next one would
// repeat it move for move: it is counted and not run.
__global__ void k_climb_end(ClimbDev* c, ClimbArgs a) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || c->done) return;
  c->stall = c->moved ? 0 : c->stall + 1;
  ++c->iter;
  if (!c->moved && (a.hop || a.hop_fm) && c->stall < 2 && c->iter < a.n_iters) {
    ++c->iter;
    c->stall = 2;
  }
  if ((!c->moved && (!(a.hop || a.hop_fm) || c->stall >= 2)) || c->iter >= a.n_iters) c->done = 1;
}

void launch_climb_begin(ClimbDev* c, GnState* s, const ClimbArgs& a, double* hi, bool reset_moved, hipStream_t st) {
  k_climb_begin<<<1, 64, 0, st>>>(c, s, a, hi, reset_moved ? 1 : 0);
}

void launch_climb_hop(const GnBuffers& b, const GnArgs& ga, ClimbDev* c, const ClimbArgs& a, const float* res,
                      bool rs, hipStream_t st) {
  k_climb_hop<<<1, 64, 0, st>>>(b.pairs, b.state, ga, c, a, res, rs ? 1 : 0, b.try_H, b.try_ok, b.try_jobs);
}

void launch_climb_try_scores(const GnBuffers& b, const GnArgs& ga, hipStream_t st) {
  const int nd = (ga.symmetric ? 2 : 1) * ga.n_try;
  k_gn_try_scores<<<(nd + 63) / 64, 64, 0, st>>>(b.state, ga, b.try_sums, b.try_ok, b.try_score);
}

void launch_climb_select(const GnBuffers& b, const GnArgs& ga, ClimbDev* c, bool gn, hipStream_t st) {
  k_climb_select<<<1, 64, 0, st>>>(b.state, ga, c, b.try_H, b.try_ok, b.try_score, gn ? 1 : 0);
}

void launch_climb_gn_block(const GnBuffers& b, const GnArgs& ga, ClimbDev* c, int blocks_grad, int blocks_try,
                           hipStream_t st) {
  const int nd = ga.symmetric ? 2 : 1, T = ga.n_try;
  k_climb_block<<<1, 64, 0, st>>>(b.state, c);
  for (int step = 0; step < 3; ++step) {   // gclimb.zig:51 GN_STEPS
    k_gn_grad_jobs<<<1, 64, 0, st>>>(b.pairs, b.state, ga, b.gjobs);
    launch_pose_moments_jobs(b.gjobs, nd, blocks_grad, b.part, b.gsums, st);
    k_gn_coefs<<<1, 64, 0, st>>>(b.pairs, b.state, ga, b.gsums, b.dir_score, b.dir_n, b.tjobs);
    launch_tangent_jobs(b.tjobs, nd, blocks_grad, b.part, b.tsums, st);
    k_gn_step<<<1, 64, 0, st>>>(b.pairs, b.state, ga, b.tsums, b.dir_score, b.dir_n, b.step);
    k_gn_try_jobs<<<(T + 63) / 64, 64, 0, st>>>(b.pairs, b.state, ga, b.step, b.try_H, b.try_ok, b.try_jobs);
    launch_pose_moments_jobs(b.try_jobs, nd * T, blocks_try, b.part, b.try_sums, st);
    k_gn_try_scores<<<(nd * T + 63) / 64, 64, 0, st>>>(b.state, ga, b.try_sums, b.try_ok, b.try_score);
    k_climb_select<<<1, 64, 0, st>>>(b.state, ga, c, b.try_H, b.try_ok, b.try_score, 1);
  }
}

void launch_climb_end(ClimbDev* c, const ClimbArgs& a, hipStream_t st) { k_climb_end<<<1, 64, 0, st>>>(c, a); }

}  // namespace zc
