// Host-side pieces of the rotation × scale search (search.zig:48-137, 500-529): the candidate
// grid, each candidate's linear part, the pose about the moving centre, singular values, and
// non-maximum suppression over the raw peaks.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "geometry.hpp"

namespace zc {

struct SweepGrid {
  std::vector<double> thetas, sigmas, shears, anisos;
  bool wrap = false;
  size_t count() const { return thetas.size() * sigmas.size() * shears.size() * anisos.size(); }
};

// n points from lo to hi inclusive; one point: the midpoint (search.zig:78-86)
inline std::vector<double> linspace(double lo, double hi, uint32_t n) {
  std::vector<double> out(n);
  for (uint32_t i = 0; i < n; ++i)
    out[i] = n == 1 ? 0.5 * (lo + hi) : lo + ((double)i * (hi - lo)) / (double)(n - 1);
  return out;
}

// sweepGrid (search.zig:89-116): the full circle without repeating ±180°, a partial range with
// both ends; scales and stretches log-spaced, shears linear about 0.
inline SweepGrid sweep_grid(double rot_lo, double rot_hi, uint32_t n_theta, double s_min, double s_max,
                            uint32_t n_sigma, double shear_max, uint32_t n_shear, double aniso_max, uint32_t n_aniso) {
  SweepGrid g;
  const uint32_t nT = std::max<uint32_t>(1, n_theta);
  const double lo_d = std::min(rot_lo, rot_hi), span_d = std::fabs(rot_hi - rot_lo);
  g.wrap = span_d >= 360 - 1e-9;
  const double rad = M_PI / 180.0;
  g.thetas.resize(nT);
  for (uint32_t i = 0; i < nT; ++i) {
    const double fi = i, fn = nT;
    const double d = g.wrap ? -180 + 360 * fi / fn : (nT == 1 ? lo_d + span_d / 2 : lo_d + span_d * fi / (fn - 1));
    g.thetas[i] = d * rad;
  }
  const double llo = std::log(std::max(1e-3, std::min(s_min, s_max)));
  const double lhi = std::log(std::max(1e-3, std::max(s_min, s_max)));
  g.sigmas = linspace(llo, lhi, n_sigma);
  for (double& v : g.sigmas) v = std::exp(v);
  const double kmax = std::fabs(shear_max);
  g.shears = linspace(-kmax, kmax, kmax > 0 ? n_shear : 1);
  const double amax = std::log(1 + std::fabs(aniso_max));
  g.anisos = linspace(-amax, amax, amax > 0 ? n_aniso : 1);
  for (double& v : g.anisos) v = std::exp(v);
  return g;
}

// σ R(θ) diag(a, 1/a) [1 k; 0 1], row-major 2×2 (search.zig:119-127)
inline std::array<double, 4> cand_linear(double th, double s, double k, double a) {
  const double c = s * std::cos(th), sn = s * std::sin(th);
  const double p00 = a, p01 = a * k, p11 = 1 / a;
  return {c * p00, c * p01 - sn * p11, sn * p00, sn * p01 + c * p11};
}

// X ↦ c + L (X − c) (search.zig:130-132)
inline Mat3 about_c(const double c[2], const std::array<double, 4>& L) {
  return {L[0], L[1], c[0] - L[0] * c[0] - L[1] * c[1], L[2], L[3], c[1] - L[2] * c[0] - L[3] * c[1], 0, 0, 1};
}

// singular values of a 2×2, largest first (search.zig:135-143)
inline std::array<double, 2> sv2(const std::array<double, 4>& L) {
  const double a = L[0] * L[0] + L[2] * L[2], b = L[0] * L[1] + L[2] * L[3], d = L[1] * L[1] + L[3] * L[3];
  const double t = 0.5 * (a + d), r = std::sqrt(std::max(0.0, 0.25 * (a - d) * (a - d) + b * b));
  return {std::sqrt(t + r), std::sqrt(std::max(0.0, t - r))};
}

// lie.mapPt (lie.zig:311-317): a homogeneous divide with |Z| < 1e-8 taken as 1e-8
inline std::array<double, 2> map_pt(const Mat3& H, double x, double y) {
  const double X = H[0] * x + H[1] * y + H[2], Y = H[3] * x + H[4] * y + H[5], Z = H[6] * x + H[7] * y + H[8];
  const double z = std::fabs(Z) < 1e-8 ? 1e-8 : Z;
  return {X / z, Y / z};
}

// lie.zig:319-329 movingCentroid: the mean of the moving image's corners under H
inline std::array<double, 2> moving_centroid(const Mat3& H, double lw, double lh) {
  const double cs[4][2] = {{1, 1}, {lw, 1}, {lw, lh}, {1, lh}};
  std::array<double, 2> c{0.0, 0.0};
  for (const auto& q : cs) {
    const auto p = map_pt(H, q[0], q[1]);
    c[0] += 0.25 * p[0];
    c[1] += 0.25 * p[1];
  }
  return c;
}

// lie.zig:332-341 simAbout: X ↦ c + e^sg R_th (X − c)
inline Mat3 sim_about(const std::array<double, 2>& c, double th, double sg) {
  const double s = std::exp(sg);
  const double a = s * std::cos(th), b = -s * std::sin(th), d = s * std::sin(th), e = s * std::cos(th);
  return {a, b, c[0] - a * c[0] - b * c[1], d, e, c[1] - d * c[0] - e * c[1], 0, 0, 1};
}

// One raw peak of the sweep (search.zig:145): score (n in fixed pixels), candidate and grid
// indices, the candidate's θ and σ, the canvas shift in cells and the corrected pose.
// Rotation (radians) and scale of pose H relative to H0: the linear part of H · H0⁻¹ (the
// search's result as the sweep reports it; search.zig simOf's role).
inline std::array<double, 2> sim_of(const Mat3& H, const Mat3& H0) {
  const Mat3 M = mul3(H, inv3(H0));
  const double a = M[0] / M[8], b = M[1] / M[8], c = M[3] / M[8], d = M[4] / M[8];
  return {std::atan2(c - b, a + d), std::sqrt(std::fabs(a * d - b * c))};
}

struct SweepPeak {
  double score;
  uint32_t cand, it, is, ik, ia;
  uint32_t pad;   // zero: the record crosses the C ABI byte for byte
  double th, s;
  int32_t sx, sy;
  Mat3 H;
};

// nms4 (search.zig:500-529): strongest first; a peak survives unless a kept one lies within
// ±2 θ steps (circular when wrap), ±2 σ steps, ±1 shear and stretch step and 4 cells.
inline std::vector<SweepPeak> nms4(const std::vector<SweepPeak>& peaks, uint32_t n_theta, uint32_t keep, bool wrap) {
  std::vector<SweepPeak> sorted = peaks;
  std::stable_sort(sorted.begin(), sorted.end(), [](const SweepPeak& a, const SweepPeak& b) { return a.score > b.score; });
  std::vector<SweepPeak> kept;
  auto ad = [](uint32_t a, uint32_t b) { return a > b ? a - b : b - a; };
  for (const SweepPeak& p : sorted) {
    bool ok = true;
    for (const SweepPeak& q : kept) {
      const uint32_t a = ad(p.it, q.it);
      const uint32_t dt = wrap ? std::min(a, n_theta - a) : a;
      if (dt <= 2 && ad(p.is, q.is) <= 2 && ad(p.ik, q.ik) <= 1 && ad(p.ia, q.ia) <= 1 &&
          std::hypot((double)(p.sx - q.sx), (double)(p.sy - q.sy)) <= 4) {
        ok = false;
        break;
      }
    }
    if (ok) {
      kept.push_back(p);
      if (kept.size() >= keep) break;
    }
  }
  return kept;
}

}  // namespace zc
