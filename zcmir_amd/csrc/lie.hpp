// Pose algebra on the host, FP64 (lie.zig:22-376). Conventions: H is row-major 3×3, moving pixel
// → fixed pixel, 1-based; tangent coordinates ξ = [tx, ty, th, sg, al, ga, px, py] are left
// increments in normalized coordinates of the destination (destN); affine uses the first 6.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "geometry.hpp"

namespace zc {

using Coords = std::array<double, 8>;

inline int n_keys(bool homography) { return homography ? 8 : 6; }

// pixel → [-1, 1] coordinates of a w × h image (lie.zig:28-37)
inline Mat3 dest_n(double w0, double h0) {
  const double w = std::max(w0, 1.0), h = std::max(h0, 1.0);
  const double sx = 0.5 * w, sy = 0.5 * h, cx = 0.5 * (w + 1), cy = 0.5 * (h + 1);
  return {1 / sx, 0, -cx / sx, 0, 1 / sy, -cy / sy, 0, 0, 1};
}

// lie.zig:120-141
inline Mat3 hat(const Coords& x, bool homography) {
  double px = x[6], py = x[7], iso, z;
  if (!homography) {
    px = py = 0;
    z = 0;
    iso = x[3];
  } else {
    iso = x[3] / 3;
    z = -2 * iso;
  }
  return {iso + x[4], x[5] - x[2], x[0], x[5] + x[2], iso - x[4], x[1], px, py, z};
}

// scaling and squaring with a 13-term Taylor series (lie.zig:143-163)
inline Mat3 expm3(const Mat3& A) {
  double nrm = 0;
  for (double v : A) nrm = std::max(nrm, std::fabs(v));
  int s = 0;
  if (nrm > 0.5) s = (int)std::min(10.0, std::max(0.0, std::ceil(std::log2(nrm / 0.5))));
  const double sc = std::pow(0.5, s);
  Mat3 B;
  for (int i = 0; i < 9; ++i) B[i] = A[i] * sc;
  Mat3 T = {1, 0, 0, 0, 1, 0, 0, 0, 1}, P = T;
  for (int k = 1; k < 14; ++k) {
    P = mul3(P, B);
    for (double& v : P) v /= (double)k;
    for (int i = 0; i < 9; ++i) T[i] += P[i];
  }
  for (int i = 0; i < s; ++i) T = mul3(T, T);
  return T;
}

// lie.zig:165-175
inline Mat3 project_group(const Mat3& H, bool homography) {
  const double s = std::fabs(H[8]) > 1e-12 ? H[8] : 1;
  Mat3 o;
  for (int i = 0; i < 9; ++i) o[i] = H[i] / s;
  if (!homography) {
    o[6] = 0;
    o[7] = 0;
    o[8] = 1;
  }
  return o;
}

// exp(ξ̂) in normalized coordinates of a dw × dh destination, applied to H0 (lie.zig:182-185)
inline Mat3 compose_n(const Mat3& H0, const Coords& x, bool homography, double dw, double dh) {
  const Mat3 N = dest_n(dw, dh);
  return project_group(mul3(mul3(mul3(inv3(N), expm3(hat(x, homography))), N), H0), homography);
}

// lie.zig:187-207
inline Coords vee(const Mat3& A0, bool homography) {
  double a = A0[0], e = A0[4];
  const double b = A0[1], c = A0[2], d = A0[3], f = A0[5], g = A0[6], h = A0[7], i = A0[8];
  if (homography) {
    const double tr = (a + e + i) / 3;
    a -= tr;
    e -= tr;
    const double iso = (a + e) / 2;
    return {c, f, (d - b) / 2, iso * 3, (a - e) / 2, (d + b) / 2, g, h};
  }
  return {c, f, (d - b) / 2, (a + e) / 2, (a - e) / 2, (d + b) / 2, 0, 0};
}

// one generator per tangent (zero past the group's), in pixel coordinates of a dw × dh
// destination when dest (lie.zig:209-225)
inline std::array<Mat3, 8> generators(bool homography, bool dest = false, double dw = 0, double dh = 0) {
  std::array<Mat3, 8> out{};
  const int nk = n_keys(homography);
  for (int k = 0; k < nk; ++k) {
    Coords x{};
    x[k] = 1;
    out[k] = hat(x, homography);
  }
  if (dest) {
    const Mat3 N = dest_n(dw, dh), Ni = inv3(N);
    for (int k = 0; k < nk; ++k) out[k] = mul3(mul3(Ni, out[k]), N);
  }
  return out;
}

// dS/dξ on H from dS/dη on H⁻¹ (η: left increment in the moving frame) (lie.zig:227-241)
inline Coords adjoint_grad(const double* g_eta, const Mat3& H, bool homography, double mw, double mh, double fw,
                           double fh) {
  const Mat3 Hi = inv3(H);
  const Mat3 M = mul3(mul3(dest_n(mw, mh), Hi), inv3(dest_n(fw, fh)));
  const Mat3 Mi = inv3(M);
  const auto gens = generators(homography);
  const int nk = n_keys(homography);
  Coords out{};
  for (int i = 0; i < nk; ++i) {
    const Coords cc = vee(mul3(mul3(M, gens[i]), Mi), homography);
    for (int j = 0; j < nk; ++j) out[i] -= g_eta[j] * cc[j];
  }
  return out;
}

// lie.zig:243-255
inline Coords unit_ascent(const double* g, bool homography) {
  const int nk = n_keys(homography);
  Coords u{};
  double n = 0;
  for (int i = 0; i < nk; ++i) {
    u[i] = g[i];
    n += u[i] * u[i];
  }
  n = std::sqrt(n);
  if (n < 1e-18) return u;
  for (int i = 0; i < nk; ++i) u[i] /= n;
  return u;
}

// A x = b, n×n row-major, Gauss–Jordan with partial pivoting; false if singular or non-finite
// (lie.zig:348-376)
inline bool solve_linear(const double* A, const double* b, int n, double* x) {
  double M[64];
  std::copy(A, A + n * n, M);
  std::copy(b, b + n, x);
  for (int i = 0; i < n; ++i) {
    int piv = i;
    for (int r = i + 1; r < n; ++r)
      if (std::fabs(M[r * n + i]) > std::fabs(M[piv * n + i])) piv = r;
    if (piv != i) {
      for (int j = 0; j < n; ++j) std::swap(M[i * n + j], M[piv * n + j]);
      std::swap(x[i], x[piv]);
    }
    const double d = M[i * n + i];
    if (!std::isfinite(d) || std::fabs(d) < 1e-300) return false;
    for (int r = 0; r < n; ++r) {
      if (r == i) continue;
      const double f = M[r * n + i] / d;
      if (f == 0) continue;
      for (int j = i; j < n; ++j) M[r * n + j] -= f * M[i * n + j];
      x[r] -= f * x[i];
    }
  }
  for (int i = 0; i < n; ++i) {
    x[i] /= M[i * n + i];
    if (!std::isfinite(x[i])) return false;
  }
  return true;
}

// the ascent ladder a0, a0 r, a0 r², … (climb.zig:15-24)
inline std::vector<double> geo_amps(double a0, double ratio, uint32_t n) {
  double a = a0 > 0 ? a0 : 0.08;
  const double r = std::min(0.95, std::max(0.2, ratio > 0 ? ratio : 0.5));
  const uint32_t k = std::max<uint32_t>(2, std::min<uint32_t>(12, n));
  std::vector<double> out(k);
  for (uint32_t i = 0; i < k; ++i) {
    out[i] = a;
    a *= r;
  }
  return out;
}

// higher by more than 1e-6 of the old magnitude (pair.zig:51-53)
inline bool beats(double nw, double old) { return nw > old + std::max(1e-6 * std::fabs(old), 1e-12); }

}  // namespace zc
