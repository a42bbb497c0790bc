// Pose scores, tangent gradients and the batched Gauss–Newton refinement (pair.zig:155-306,
// pose.zig:267-289, moments.zig:91-172, search.zig:539-614), plus the search flow on top of the
// sweep (search.zig:663-822 runSweep).
//
// Every evaluation is batched: one launch computes the 45 overlap sums of all poses of all
// directions in flight (PoseJob per pose), the score coefficients are formed on the host in FP64,
// and one launch computes every pose's tangent gradient and Gauss–Newton matrix (TangentJob).
// Each image pair is referenced by its two sides, so the same code serves many poses of one pair
// (the search) and one pose each of many pairs (a stack).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>

#include "engine.hpp"
#include "lie.hpp"

namespace zc {

namespace {

using M4 = std::array<double, 16>;

bool chol4(const M4& G, M4& L) {
  L.fill(0);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = G[i * 4 + j];
      for (int k = 0; k < j; ++k) s -= L[i * 4 + k] * L[j * 4 + k];
      if (i == j) {
        if (!(s > 0)) return false;
        L[i * 5] = std::sqrt(s);
      } else {
        L[i * 4 + j] = s / L[j * 5];
      }
    }
  return true;
}

// (L Lᵀ)⁻¹ column by column: forward then back substitution
M4 chol_inv(const M4& L) {
  M4 X{};
  for (int c = 0; c < 4; ++c) {
    double y[4];
    for (int i = 0; i < 4; ++i) {
      double s = i == c ? 1 : 0;
      for (int k = 0; k < i; ++k) s -= L[i * 4 + k] * y[k];
      y[i] = s / L[i * 5];
    }
    for (int i = 3; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 4; ++k) s -= L[k * 4 + i] * X[k * 4 + c];
      X[i * 4 + c] = s / L[i * 5];
    }
  }
  return X;
}

M4 mm(const M4& A, const M4& B) {
  M4 o{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 4 + j];
      o[i * 4 + j] = s;
    }
  return o;
}

M4 tr(const M4& A) {
  M4 o;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) o[i * 4 + j] = A[j * 4 + i];
  return o;
}

inline int tri(int i, int j) {
  const int a = std::min(i, j), b = std::max(i, j);
  return a * 4 - (a * (a + 1)) / 2 + b;
}

}  // namespace

// SMI score n Φ from the 45 sums and its gradient coefficients (moments.zig:91-172): global
// Φ = ‖C‖², exact Φ = tr(G_A⁻¹ C G_B⁻¹ Cᵀ) with ridge-regularized covariances; W, M̂ (the
// ridge's trace dependence included), the Gauss–Newton weight Q, the means and g = ∂S/∂sums.
Coefs score_coefs(const double* st, bool exact, double ridge) {
  Coefs o{};
  const double n = st[0];
  o.n = n;
  if (!(n >= kMinN)) return o;
  double ma[4], mb[4];
  M4 C;
  for (int i = 0; i < 4; ++i) {
    ma[i] = st[1 + i] / n;
    mb[i] = st[5 + i] / n;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) C[i * 4 + j] = st[9 + i * 4 + j] / n - ma[i] * mb[j];
  M4 W = C, Mh{}, Nh{};
  double phi = 0;
  if (!exact) {
    for (double v : C) phi += v * v;
  } else {
    M4 GA, GB;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        GA[i * 4 + j] = st[25 + tri(i, j)] / n - ma[i] * ma[j];
        GB[i * 4 + j] = st[35 + tri(i, j)] / n - mb[i] * mb[j];
      }
    const double ra = ridge * std::max((GA[0] + GA[5] + GA[10] + GA[15]) / 4, 1e-12);
    const double rb = ridge * std::max((GB[0] + GB[5] + GB[10] + GB[15]) / 4, 1e-12);
    for (int i = 0; i < 4; ++i) {
      GA[i * 5] += ra;
      GB[i * 5] += rb;
    }
    M4 LA, LB;
    if (!chol4(GA, LA) || !chol4(GB, LB)) return o;
    const M4 GAi = chol_inv(LA), GBi = chol_inv(LB);
    W = mm(mm(GAi, C), GBi);
    for (int i = 0; i < 16; ++i) phi += W[i] * C[i];
    Mh = mm(mm(W, tr(C)), GAi);
    Nh = mm(mm(tr(W), C), GBi);
    const double tm = (Mh[0] + Mh[5] + Mh[10] + Mh[15]) * ridge / 4;
    const double tn = (Nh[0] + Nh[5] + Nh[10] + Nh[15]) * ridge / 4;
    for (int i = 0; i < 4; ++i) {
      Mh[i * 5] += tm;
      Nh[i * 5] += tn;
    }
  }
  const double score = n * phi;
  if (!std::isfinite(score)) return o;
  o.score = score;
  for (int i = 0; i < 16; ++i) {
    o.W[i] = W[i];
    o.Mh[i] = Mh[i];
  }
  for (int i = 0; i < 4; ++i) {
    o.ma[i] = ma[i];
    o.mb[i] = mb[i];
  }
  double* g = o.g;
  for (int i = 0; i < 4; ++i) {
    double ga = 0, gb = 0;
    for (int j = 0; j < 4; ++j) {
      ga += -2 * W[i * 4 + j] * mb[j] + 2 * Mh[i * 4 + j] * ma[j];
      gb += -2 * W[j * 4 + i] * ma[j] + 2 * Nh[i * 4 + j] * mb[j];
      g[9 + i * 4 + j] = 2 * W[i * 4 + j];
    }
    g[1 + i] = ga;
    g[5 + i] = gb;
    for (int j = i; j < 4; ++j) {
      const double f = i == j ? 1 : 2;
      g[25 + tri(i, j)] = -f * Mh[i * 4 + j];
      g[35 + tri(i, j)] = -f * Nh[i * 4 + j];
    }
  }
  double rest = 0;
  for (int k = 1; k < 45; ++k) rest += g[k] * st[k];
  g[0] = (score - rest) / n;
  const M4 Q = exact ? Mh : mm(C, tr(C));
  for (int i = 0; i < 16; ++i) o.Q[i] = Q[i];
  return o;
}

// ── batched evaluations ─────────────────────────────────────────────────────────────────
void Engine::moments_jobs(const std::vector<DirJob>& d, std::vector<std::array<double, 45>>& out) {
  const size_t nj = d.size();
  out.assign(nj, {});
  if (!nj) return;
  std::vector<PoseJob> jobs(nj);
  int64_t px = 1;
  // work: destination pixels over all poses; algorithmic bytes: each side's 4 feature planes once
  double pixels = 0, feat_bytes = 0;
  std::set<const Side*> seen;
  for (size_t i = 0; i < nj; ++i) {
    const Side& a = *d[i].a;
    const Side& b = *d[i].b;
    const Mat3 Hi = inv3(d[i].H);
    ensure_il(*d[i].a);
    ensure_il(*d[i].b);
    jobs[i] = {a.il.as<float4>(), b.il.as<float4>(), a.w, a.h, b.w, b.h, {}};
    for (int k = 0; k < 9; ++k) jobs[i].Hi[k] = Hi[k];
    px = std::max<int64_t>(px, (int64_t)b.w * b.h);
    pixels += (double)b.w * b.h;
    seen.insert(&a);
    seen.insert(&b);
  }
  for (const Side* s : seen) feat_bytes += 16.0 * s->w * s->h;
  const int blocks = jobs_blocks(px, (int)nj);
  gn_jobs.ensure(nj * sizeof(PoseJob));
  gn_part.ensure(nj * blocks * 45 * sizeof(double));
  gn_sums.ensure(nj * 45 * sizeof(double));
  ZC_HIP(hipMemcpyAsync(gn_jobs.p, jobs.data(), nj * sizeof(PoseJob), hipMemcpyHostToDevice, st));
  prof.break_chain();
  bytes("pose_moments", feat_bytes);
  count_work("pose_moments", pixels);
  LAUNCH("pose_moments", launch_pose_moments_jobs(gn_jobs.as<PoseJob>(), (int)nj, blocks, gn_part.as<double>(),
                                                  gn_sums.as<double>(), st));
  ZC_HIP(hipMemcpyAsync(out.data(), gn_sums.p, nj * 45 * sizeof(double), hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
}

void Engine::tangent_jobs(const std::vector<DirJob>& d, const std::vector<Coefs>& co, bool homography, bool hess,
                          std::vector<std::array<double, 44>>& out) {
  const size_t nj = d.size();
  out.assign(nj, {});
  if (!nj) return;
  std::vector<TangentJob> jobs(nj);
  int64_t px = 1;
  double pixels = 0, feat_bytes = 0;
  std::set<const Side*> seen;
  for (size_t i = 0; i < nj; ++i) {
    const Side& a = *d[i].a;
    const Side& b = *d[i].b;
    TangentJob& j = jobs[i];
    std::memset(&j, 0, sizeof j);
    ensure_il(*d[i].a);
    ensure_il(*d[i].b);
    j.fa = a.il.as<float4>();
    j.fb = b.il.as<float4>();
    j.wa = a.w;
    j.ha = a.h;
    j.wb = b.w;
    j.hb = b.h;
    j.hess = hess ? 1 : 0;
    const Mat3 Hi = inv3(d[i].H);
    for (int k = 0; k < 9; ++k) j.Hi[k] = Hi[k];
    const auto gens = generators(homography, true, b.w, b.h);
    for (int k = 0; k < 8; ++k) {
      const Mat3 M = mul3(Hi, gens[k]);
      for (int q = 0; q < 9; ++q) j.Mk[9 * k + q] = (float)-M[q];
    }
    const Coefs& c = co[i];
    for (int q = 0; q < 16; ++q) {
      j.W[q] = (float)c.W[q];
      j.Mh[q] = (float)c.Mh[q];
      j.Q[q] = (float)c.Q[q];
    }
    for (int q = 0; q < 4; ++q) {
      j.ma[q] = (float)c.ma[q];
      j.mb[q] = (float)c.mb[q];
    }
    for (int q = 0; q < 45; ++q) j.g[q] = (float)c.g[q];
    px = std::max<int64_t>(px, (int64_t)b.w * b.h);
    pixels += (double)b.w * b.h;
    seen.insert(&a);
    seen.insert(&b);
  }
  for (const Side* s : seen) feat_bytes += 16.0 * s->w * s->h;
  const int blocks = jobs_blocks(px, (int)nj);
  gn_jobs.ensure(nj * sizeof(TangentJob));
  gn_part.ensure(nj * blocks * 44 * sizeof(double));
  gn_sums.ensure(nj * 44 * sizeof(double));
  ZC_HIP(hipMemcpyAsync(gn_jobs.p, jobs.data(), nj * sizeof(TangentJob), hipMemcpyHostToDevice, st));
  prof.break_chain();
  bytes("tangent_grad", feat_bytes);
  count_work("tangent_grad", pixels);
  LAUNCH("tangent_grad", launch_tangent_jobs(gn_jobs.as<TangentJob>(), (int)nj, blocks, gn_part.as<double>(),
                                             gn_sums.as<double>(), st));
  ZC_HIP(hipMemcpyAsync(out.data(), gn_sums.p, nj * 44 * sizeof(double), hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
}

// Pair.scorePoses (pair.zig:192-239): the forward score, or with symmetric the mean of the
// forward and inverse directions; non-finite direction scores count 0.
std::vector<double> Engine::pair_scores(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, bool exact,
                                        bool symmetric) {
  const size_t n = Hs.size();
  std::vector<DirJob> d;
  d.reserve(n * 2);
  for (size_t i = 0; i < n; ++i) d.push_back({pr[i].a, pr[i].b, Hs[i]});
  if (symmetric)
    for (size_t i = 0; i < n; ++i) d.push_back({pr[i].b, pr[i].a, inv3(Hs[i])});
  std::vector<std::array<double, 45>> st;
  moments_jobs(d, st);
  auto fin = [](double v) { return std::isfinite(v) ? v : 0.0; };
  std::vector<double> out(n);
  for (size_t i = 0; i < n; ++i) {
    const double f = fin(score_coefs(st[i].data(), exact, kRidge).score);
    out[i] = symmetric ? 0.5 * (f + fin(score_coefs(st[n + i].data(), exact, kRidge).score)) : f;
  }
  return out;
}

std::array<double, 3> Engine::score_parts(const Mat3& Hp, bool exact, bool symmetric) {
  std::vector<DirJob> d{{&sides[0], &sides[1], Hp}};
  if (symmetric) d.push_back({&sides[1], &sides[0], inv3(Hp)});
  std::vector<std::array<double, 45>> st;
  moments_jobs(d, st);
  auto fin = [](double v) { return std::isfinite(v) ? v : 0.0; };
  const double f = fin(score_coefs(st[0].data(), exact, kRidge).score);
  if (!symmetric) return {f, f, std::nan("")};
  const double i = fin(score_coefs(st[1].data(), exact, kRidge).score);
  return {0.5 * (f + i), f, i};
}

// Pair.gradBatch (pair.zig:276-306, pose.zig:267-289 gradFrom, pair.zig:241-268 sym).
std::vector<GradOut> Engine::pair_grads(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, bool exact,
                                        bool homography, bool symmetric, bool hess) {
  const size_t n = Hs.size();
  const int nk = n_keys(homography);
  std::vector<DirJob> d;
  d.reserve(n * 2);
  for (size_t i = 0; i < n; ++i) d.push_back({pr[i].a, pr[i].b, Hs[i]});
  if (symmetric)
    for (size_t i = 0; i < n; ++i) d.push_back({pr[i].b, pr[i].a, inv3(Hs[i])});
  std::vector<std::array<double, 45>> st;
  moments_jobs(d, st);
  std::vector<Coefs> co(d.size());
  for (size_t k = 0; k < d.size(); ++k) co[k] = score_coefs(st[k].data(), exact, kRidge);
  // the tangent pass only for directions with enough overlap
  std::vector<DirJob> dt;
  std::vector<Coefs> ct;
  std::vector<size_t> which;
  for (size_t k = 0; k < d.size(); ++k)
    if (co[k].n >= kMinN) {
      dt.push_back(d[k]);
      ct.push_back(co[k]);
      which.push_back(k);
    }
  std::vector<std::array<double, 44>> tg;
  tangent_jobs(dt, ct, homography, hess, tg);
  std::vector<GradOut> dir(d.size());
  for (size_t k = 0; k < d.size(); ++k) {
    dir[k] = GradOut{};
    dir[k].score = co[k].score;
    dir[k].n = co[k].n;
    dir[k].has_hess = hess;
  }
  for (size_t q = 0; q < which.size(); ++q) {
    GradOut& r = dir[which[q]];
    const auto& acc = tg[q];
    for (int k = 0; k < nk; ++k) r.grad[k] = acc[k];
    if (hess) {
      int h = 8;
      for (int a = 0; a < 8; ++a)
        for (int b = a; b < 8; ++b, ++h)
          if (a < nk && b < nk) {
            r.hess[a * nk + b] = acc[h];
            r.hess[b * nk + a] = acc[h];
          }
    }
  }
  if (!symmetric) return std::vector<GradOut>(dir.begin(), dir.begin() + n);
  std::vector<GradOut> out(n);
  for (size_t i = 0; i < n; ++i) {
    const GradOut& f = dir[i];
    const GradOut& b = dir[n + i];
    const double lw = pr[i].a->w, lh = pr[i].a->h, rw = pr[i].b->w, rh = pr[i].b->h;
    GradOut o = f;
    o.score = 0.5 * (f.score + b.score);
    const Coords pulled = adjoint_grad(b.grad, Hs[i], homography, lw, lh, rw, rh);
    for (int k = 0; k < nk; ++k) o.grad[k] = 0.5 * (f.grad[k] + pulled[k]);
    if (hess) {
      double PT[64] = {};
      for (int j = 0; j < nk; ++j) {
        double e[8] = {};
        e[j] = 1;
        const Coords col = adjoint_grad(e, Hs[i], homography, lw, lh, rw, rh);
        for (int k = 0; k < nk; ++k) PT[k * nk + j] = col[k];
      }
      for (int a = 0; a < nk; ++a)
        for (int c = 0; c < nk; ++c) {
          double s = 0;
          for (int x = 0; x < nk; ++x)
            for (int y = 0; y < nk; ++y) s += PT[a * nk + x] * b.hess[x * nk + y] * PT[c * nk + y];
          o.hess[a * nk + c] = 0.5 * (f.hess[a * nk + c] + s);
        }
    }
    out[i] = o;
  }
  return out;
}

// gnBatch (search.zig:539-614): per iteration, every active pose's gradient and Gauss–Newton
// matrix; tries: the ascent ladder along the unit gradient and 5 lengths of the step
// (A + 1e-4 diag) δ = ½ g (capped at norm 0.25); a try replaces its pose when it beats it; poses
// that did not move leave the batch. In input order.
std::vector<GnOut> Engine::gn_batch(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs, uint32_t iters,
                                    bool homography, bool exact, bool symmetric, const std::vector<double>& amps,
                                    uint32_t* iterations) {
  if (set.b("gn_graph") && !Hs.empty() && amps.size() <= 12)
    return gn_batch_device(pr, Hs, iters, homography, exact, symmetric, amps, iterations);
  const size_t n = Hs.size();
  const int nk = n_keys(homography);
  std::vector<GnOut> cur(n);
  for (size_t i = 0; i < n; ++i) cur[i] = {Hs[i], -INFINITY, false};
  std::vector<size_t> act(n);
  std::iota(act.begin(), act.end(), 0);
  const double mults[5] = {0.5, 1, 2, 4, 8};
  auto move = [&](size_t i, const Mat3& H, const double* v, double a) {
    Coords x{};
    for (int k = 0; k < nk; ++k) x[k] = a * v[k];
    return compose_n(H, x, homography, pr[i].b->w, pr[i].b->h);
  };
  uint32_t it = 0;
  for (; it < iters && !cancel && !act.empty(); ++it) {
    std::vector<PairSides> ap;
    std::vector<Mat3> ah;
    for (size_t i : act) {
      ap.push_back(pr[i]);
      ah.push_back(cur[i].H);
    }
    const std::vector<GradOut> gr = pair_grads(ap, ah, exact, homography, symmetric, true);
    std::vector<size_t> ti;
    std::vector<PairSides> tp;
    std::vector<Mat3> th;
    for (size_t k = 0; k < act.size(); ++k) {
      const size_t i = act[k];
      const GradOut& g = gr[k];
      cur[i].score = g.score;
      if (!amps.empty()) {
        const Coords d = unit_ascent(g.grad, homography);
        double nd = 0;
        for (int q = 0; q < nk; ++q) nd += d[q] * d[q];
        if (std::sqrt(nd) > 1e-12)
          for (double a : amps) {
            ti.push_back(i);
            tp.push_back(pr[i]);
            th.push_back(move(i, cur[i].H, d.data(), a));
          }
      }
      double A[64], half[8], step[8];
      std::copy(g.hess, g.hess + nk * nk, A);
      for (int q = 0; q < nk; ++q) {
        A[q * nk + q] += 1e-4 * std::max(A[q * nk + q], 1e-12);
        half[q] = 0.5 * g.grad[q];
      }
      if (!solve_linear(A, half, nk, step)) continue;
      double nrm = 0;
      for (int q = 0; q < nk; ++q) nrm += step[q] * step[q];
      nrm = std::sqrt(nrm);
      for (double m : mults) {
        const double s = m * (nrm * m > 0.25 ? 0.25 / (nrm * m) : 1);
        ti.push_back(i);
        tp.push_back(pr[i]);
        th.push_back(move(i, cur[i].H, step, s));
      }
    }
    if (th.empty()) break;
    const std::vector<double> sc = pair_scores(tp, th, exact, symmetric);
    std::vector<char> moved(n, 0);
    for (size_t k = 0; k < th.size(); ++k)
      if (beats(sc[k], cur[ti[k]].score)) {
        cur[ti[k]].H = th[k];
        cur[ti[k]].score = sc[k];
        cur[ti[k]].moved = true;
        moved[ti[k]] = 1;
      }
    std::vector<size_t> next;
    for (size_t i : act)
      if (moved[i]) next.push_back(i);
    act.swap(next);
  }
  if (iterations) *iterations = it;
  return cur;
}

// gnBatch with every iteration on the device (k_gn.hip): the pairs and start poses go up once,
// one captured graph (gradient sums, coefficients, tangent pass, tries, their sums and scores, the
// selection) is replayed per iteration without a host wait, and the poses, scores and the count
// of iterations come back once. Same arithmetic as the host loop above.
std::vector<GnOut> Engine::gn_batch_device(const std::vector<PairSides>& pr, const std::vector<Mat3>& Hs,
                                           uint32_t iters, bool homography, bool exact, bool symmetric,
                                           const std::vector<double>& amps, uint32_t* iterations) {
  const size_t n = Hs.size();
  std::vector<GnPair> pairs(n);
  std::vector<GnState> state(n);
  int64_t px = 1;
  for (size_t i = 0; i < n; ++i) {
    const Side& a = *pr[i].a;
    const Side& b = *pr[i].b;
    ensure_il(a);
    ensure_il(b);
    pairs[i] = {a.il.as<float4>(), b.il.as<float4>(), a.w, a.h, b.w, b.h};
    GnState& s = state[i];
    for (int k = 0; k < 9; ++k) s.H[k] = Hs[i][k];
    s.score = -INFINITY;
    s.active = 1;
    s.moved = 0;
    px = std::max<int64_t>(px, std::max<int64_t>((int64_t)a.w * a.h, (int64_t)b.w * b.h));
  }
  GnArgs ga{};
  ga.n = (int)n;
  ga.homography = homography;
  ga.exact = exact;
  ga.symmetric = symmetric;
  ga.min_n = kMinN;
  ga.n_amps = (int)amps.size();
  ga.ridge = kRidge;
  for (size_t k = 0; k < amps.size(); ++k) ga.amps[k] = amps[k];
  ga.n_try = (int)amps.size() + 5;
  const int nd = (symmetric ? 2 : 1) * (int)n, T = ga.n_try;
  const int blocks = jobs_blocks(px, nd), blocks_try = jobs_blocks(px, nd * T);
  gd_pairs.ensure(n * sizeof(GnPair));
  gd_state.ensure(n * sizeof(GnState));
  gd_gjobs.ensure(nd * sizeof(PoseJob));
  gd_gsums.ensure(nd * 45 * 8);
  gd_tjobs.ensure(nd * sizeof(TangentJob));
  gd_tsums.ensure(nd * 44 * 8);
  gd_dir.ensure(2 * nd * 8);
  gd_step.ensure(n * 18 * 8);
  gd_try_H.ensure(n * T * 9 * 8);
  gd_try_ok.ensure(n * T * 4);
  gd_try_jobs.ensure((size_t)nd * T * sizeof(PoseJob));
  gd_try_sums.ensure((size_t)nd * T * 45 * 8);
  gd_try_score.ensure((size_t)nd * T * 8);
  gd_part.ensure((size_t)std::max(nd * blocks, nd * T * blocks_try) * 45 * 8);
  gd_ctr.ensure(2 * 4);
  ZC_HIP(hipMemcpyAsync(gd_pairs.p, pairs.data(), n * sizeof(GnPair), hipMemcpyHostToDevice, st));
  ZC_HIP(hipMemcpyAsync(gd_state.p, state.data(), n * sizeof(GnState), hipMemcpyHostToDevice, st));
  ZC_HIP(hipMemsetAsync(gd_ctr.p, 0, 2 * 4, st));   // outside the graph: the counters start at 0
  prof.break_chain();
  GnBuffers gb{gd_pairs.as<GnPair>(), gd_state.as<GnState>(), gd_gjobs.as<PoseJob>(), gd_gsums.as<double>(),
               gd_tjobs.as<TangentJob>(), gd_tsums.as<double>(), gd_dir.as<double>(), gd_dir.as<double>() + nd,
               gd_step.as<double>(), gd_try_H.as<double>(), gd_try_ok.as<int>(), gd_try_jobs.as<PoseJob>(),
               gd_try_sums.as<double>(),
               gd_try_score.as<double>(), gd_part.as<double>(), gd_ctr.as<int>(), gd_ctr.as<int>() + 1};
  std::vector<uint8_t> key;
  auto put = [&key](const auto& v) {
    const auto* p = reinterpret_cast<const uint8_t*>(&v);
    key.insert(key.end(), p, p + sizeof v);
  };
  put(gb);
  put(ga);
  put(blocks);
  put(blocks_try);
  for (uint32_t it = 0; it < iters && !cancel; ++it)
    run_captured(gn_gc, std::vector<uint8_t>(key), prof.on,
                 [&] { LAUNCH("gn_iteration", launch_gn_iteration(gb, ga, blocks, blocks_try, st)); });
  int done = 0;
  ZC_HIP(hipMemcpyAsync(state.data(), gd_state.p, n * sizeof(GnState), hipMemcpyDeviceToHost, st));
  ZC_HIP(hipMemcpyAsync(&done, gd_ctr.as<int>() + 1, 4, hipMemcpyDeviceToHost, st));
  prof.break_chain();
  ++counts[2];
  sync();
  if (iterations) *iterations = (uint32_t)done;
  std::vector<GnOut> out(n);
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 9; ++k) out[i].H[k] = state[i].H[k];
    out[i].score = state[i].score;
    out[i].moved = state[i].moved != 0;
  }
  return out;
}

std::vector<PairSides> Engine::own_pair(size_t n) const { return std::vector<PairSides>(n, {&sides[0], &sides[1]}); }

// The search (search.zig:663-822 runSweep): sweep → NMS → a 4-iteration Gauss–Newton screen of
// the peaks → 25 iterations on the finalists (the best screened poses and the best coarse
// peaks); the best pose replaces the current one when it beats its score.
SearchResult Engine::search(const SweepOpts& o, uint32_t keep, uint32_t n_fin, bool homography, bool exact,
                            bool symmetric, const std::vector<double>& amps, double* best_peak) {
  need_images();
  const Mat3 Hcur = H;
  const double s_cur = pair_scores(own_pair(1), {Hcur}, exact, symmetric)[0];
  SearchResult out{};
  for (int i = 0; i < 9; ++i) out.H[i] = Hcur[i];
  out.score = out.prev = s_cur;
  const SweepOut sw = sweep(o);
  if (best_peak) {
    *best_peak = 0;
    for (const SweepPeak& p : sw.peaks) *best_peak = std::max(*best_peak, p.score);
  }
  const SweepGrid& gv = *o.gv;
  const std::vector<SweepPeak> top = nms4(sw.peaks, (uint32_t)gv.thetas.size(), keep, gv.wrap);
  out.count_a = (uint32_t)top.size();
  const double deg = 180.0 / M_PI;
  events.log("search  %zuθ %.0f°…%.0f° × %zuσ %.3g–%.3g  %u maps  %u² grid (%.1f px, FFT %u²)  %s  %zu peaks",
             gv.thetas.size(), gv.thetas.front() * deg, gv.thetas.back() * deg, gv.sigmas.size(), gv.sigmas.front(),
             gv.sigmas.back(), sw.nc, sw.grid, sw.cell, sw.n, exact ? "SMI" : "global SMI", top.size());
  // the sweep's landscape (search.zig:822-833 emitSweep): each (θ, σ) cell's best score over
  // shears and stretches, the NMS peaks, finalists, the result and the climbed candidates
  const size_t nsh = gv.shears.size() * gv.anisos.size();
  std::vector<double> land(gv.thetas.size() * gv.sigmas.size(), 0.0);
  for (size_t c = 0; c < land.size() && c * nsh < sw.best.size(); ++c)
    for (size_t k = 0; k < nsh && c * nsh + k < sw.best.size(); ++k) land[c] = std::max(land[c], (double)sw.best[c * nsh + k]);
  auto sweep_event = [&](const std::vector<size_t>& fin, const Mat3* result, const std::string& cands, bool applied) {
    std::string tops = "[", fins = "[";
    for (size_t i = 0; i < top.size(); ++i)
      tops += (i ? ",[" : "[") + std::to_string(top[i].it) + "," + std::to_string(top[i].is) + "]";
    for (size_t i = 0; i < fin.size(); ++i) fins += (i ? "," : "") + std::to_string(fin[i]);
    JsonObj ev;
    ev.str("kind", "sweep").nums("thetas", gv.thetas).nums("sigmas", gv.sigmas).boolean("wrap", gv.wrap);
    ev.nums("best", land).raw("top", tops + "]").raw("finalists", fins + "]");
    if (result) {
      const auto ts = sim_of(*result, o.H0);
      ev.nums("result", ts.data(), 2);
    } else {
      ev.null("result");
    }
    events.emit(ev.raw("cands", cands).integer("applied", applied ? 0 : -1));
  };
  if (top.empty()) {
    events.log("search  no peak above the overlap floor");
    sweep_event({}, nullptr, "[]", false);
    return out;
  }
  std::vector<Mat3> starts;
  for (const SweepPeak& p : top) starts.push_back(p.H);
  const std::vector<GnOut> screened = gn_batch(own_pair(starts.size()), starts, 4, homography, exact, symmetric, amps);
  std::vector<size_t> order(screened.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return screened[a].score > screened[b].score; });
  std::vector<size_t> fin;
  for (size_t k = 0; k < std::min<size_t>(n_fin, order.size()); ++k)
    if (std::find(fin.begin(), fin.end(), order[k]) == fin.end()) fin.push_back(order[k]);
  for (size_t k = 0; k < std::min<size_t>(n_fin, top.size()); ++k)
    if (std::find(fin.begin(), fin.end(), k) == fin.end()) fin.push_back(k);
  {
    std::string f;
    for (size_t i = 0; i < fin.size(); ++i) f += (i ? " #" : "#") + std::to_string(fin[i] + 1);
    events.log("search  screen %zu × 4 steps  best nS %.1f (peak #%zu)  finalists %s", top.size(),
               screened[order[0]].score, order[0] + 1, f.c_str());
  }
  std::vector<Mat3> fh;
  for (size_t k : fin) fh.push_back(screened[k].H);
  const std::vector<GnOut> finals = gn_batch(own_pair(fh.size()), fh, 25, homography, exact, symmetric, amps);
  size_t bi = 0;
  for (size_t i = 0; i < finals.size(); ++i)
    if (finals[i].score > finals[bi].score) bi = i;
  out.count_b = (uint32_t)finals.size();
  {
    std::string f;
    char b[32];
    for (size_t i = 0; i < finals.size(); ++i) {
      std::snprintf(b, sizeof b, "%s%.1f", i ? " / " : "", finals[i].score);
      f += b;
    }
    events.log("search  climb %zu × ≤25 GN  nS %s", finals.size(), f.c_str());
  }
  // the climbed finalists, best first
  std::vector<size_t> rank(finals.size());
  std::iota(rank.begin(), rank.end(), 0);
  std::stable_sort(rank.begin(), rank.end(), [&](size_t a, size_t b) { return finals[a].score > finals[b].score; });
  std::string cands = "[";
  for (size_t r = 0; r < rank.size(); ++r) {
    const size_t i = rank[r];
    const auto ts = sim_of(finals[i].H, o.H0);
    JsonObj c;
    c.mat("H", finals[i].H).num("score", finals[i].score).integer("peak", (long long)fin[i] + 1);
    cands += (r ? "," : "") + c.num("th", ts[0]).num("s", ts[1]).done();
  }
  const bool improved = beats(finals[bi].score, s_cur);
  sweep_event(fin, &finals[bi].H, cands + "]", improved);
  if (improved) {
    H = finals[bi].H;
    for (int i = 0; i < 9; ++i) out.H[i] = H[i];
    out.score = finals[bi].score;
    out.improved = 1;
    const auto parts = score_parts(H, exact, symmetric);
    events.trail("search", parts[0], parts[1], parts[2], H);
    events.pose(H);
    const auto ts = sim_of(H, o.H0);
    events.log("search done  nS %.1f (was %.1f)  θ %.1f° σ %.3f  peak #%zu", out.score, s_cur, ts[0] * deg, ts[1],
               fin[bi] + 1);
  } else {
    events.log("search done  kept the current pose (nS %.1f ≥ %.1f)", s_cur, finals[bi].score);
  }
  return out;
}

}  // namespace zc
