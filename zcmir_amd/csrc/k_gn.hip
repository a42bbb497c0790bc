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
  double A[64], half[8], step[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int x = 0; x < nk; ++x)
    for (int c = 0; c < nk; ++c) A[x * nk + c] = h[x][c];
  for (int q = 0; q < nk; ++q) {
    A[q * nk + q] += 1e-4 * fmax(A[q * nk + q], 1e-12);
    half[q] = 0.5 * g[q];
  }
  double sn = 0;
  if (gd::solve_linear(A, half, nk, step)) {
    flags |= 2;
    for (int q = 0; q < nk; ++q) sn += step[q] * step[q];
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

}  // namespace zc
