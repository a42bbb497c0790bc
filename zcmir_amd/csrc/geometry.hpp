// Host geometry and grid plans of the shift map (FP64).
//   reg_frame / canvas_h: lie.zig:537-579; grid plans: fft.zig:20-43, 96-109, smi.zig:35-47.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <string>

namespace zc {

using Mat3 = std::array<double, 9>;

inline Mat3 mul3(const Mat3& a, const Mat3& b) {
  Mat3 o{};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o[r * 3 + c] = a[r * 3] * b[c] + a[r * 3 + 1] * b[3 + c] + a[r * 3 + 2] * b[6 + c];
  return o;
}

inline double det3(const Mat3& m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Inverse with one Newton refinement step (lie.zig:93-118 inv3).
inline Mat3 inv3(const Mat3& m) {
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  const double A = e * i - f * h, B = f * g - d * i, C = d * h - e * g;
  const double D = c * h - b * i, E = a * i - c * g, F = b * g - a * h;
  const double G = b * f - c * e, Hh = c * d - a * f, I = a * e - b * d;
  const double det = a * A + b * B + c * C;
  const double sg = det > 0 ? 1.0 : (det < 0 ? -1.0 : 0.0);
  const double s = 1.0 / (sg * std::max(std::fabs(det), 1e-18));
  const Mat3 inv = {A * s, D * s, G * s, B * s, E * s, Hh * s, C * s, F * s, I * s};
  const Mat3 p = mul3(m, inv);
  const Mat3 corr = {2 - p[0], -p[1], -p[2], -p[3], 2 - p[4], -p[5], -p[6], -p[7], 2 - p[8]};
  return mul3(inv, corr);
}

inline Mat3 translate(double tx, double ty) { return {1, 0, tx, 0, 1, ty, 0, 0, 1}; }

struct Frame {
  int ox, oy, w, h;
};

// Canvas of a pose: the fixed image plus the moving image's warped corners, clipped to ±2 max(size).
inline Frame reg_frame(const Mat3& H, int lw, int lh, int rw, int rh) {
  const double corners[4][2] = {{1, 1}, {(double)lw, 1}, {(double)lw, (double)lh}, {1, (double)lh}};
  double xs_min = 0, ys_min = 0, xs_max = rw, ys_max = rh;
  const double lim = std::max(std::max((double)lw, (double)rw), std::max(std::max((double)lh, (double)rh), 8.0)) * 2;
  for (const auto& p : corners) {
    const double x = H[0] * p[0] + H[1] * p[1] + H[2];
    const double y = H[3] * p[0] + H[4] * p[1] + H[5];
    const double z = H[6] * p[0] + H[7] * p[1] + H[8];
    const double zz = std::fabs(z) < 1e-8 ? 1e-8 : z;
    xs_min = std::min(xs_min, x / zz - 1);
    xs_max = std::max(xs_max, x / zz - 1);
    ys_min = std::min(ys_min, y / zz - 1);
    ys_max = std::max(ys_max, y / zz - 1);
  }
  auto clip = [](double v, double l) { return std::min(l, std::max(-l, v)); };
  const double ox = std::min(std::floor(clip(xs_min, lim)), 0.0);
  const double oy = std::min(std::floor(clip(ys_min, lim)), 0.0);
  const double x1 = std::max(std::ceil(clip(xs_max, lim)), (double)rw);
  const double y1 = std::max(std::ceil(clip(ys_max, lim)), (double)rh);
  return {(int)ox, (int)oy, (int)std::max(8.0, x1 - ox), (int)std::max(8.0, y1 - oy)};
}

inline Mat3 canvas_h(const Mat3& H, int ox, int oy) { return mul3(translate(-(double)ox, -(double)oy), H); }

// The inverse of a canvas pose for the warp kernels; a singular pose (|det| < 1e-12) gives all
// zeros (every q.z = 0: the warped stack is empty).
inline void warp_inverse(const Mat3& Hc, double Hi[9]) {
  const double det = det3(Hc);
  if (std::fabs(det) >= 1e-12) {
    const Mat3 inv = {(Hc[4] * Hc[8] - Hc[5] * Hc[7]) / det, (Hc[2] * Hc[7] - Hc[1] * Hc[8]) / det,
                      (Hc[1] * Hc[5] - Hc[2] * Hc[4]) / det, (Hc[5] * Hc[6] - Hc[3] * Hc[8]) / det,
                      (Hc[0] * Hc[8] - Hc[2] * Hc[6]) / det, (Hc[2] * Hc[3] - Hc[0] * Hc[5]) / det,
                      (Hc[3] * Hc[7] - Hc[4] * Hc[6]) / det, (Hc[1] * Hc[6] - Hc[0] * Hc[7]) / det,
                      (Hc[0] * Hc[4] - Hc[1] * Hc[3]) / det};
    for (int i = 0; i < 9; ++i) Hi[i] = inv[i];
  } else {
    for (int i = 0; i < 9; ++i) Hi[i] = 0;
  }
}

// ── FFT length plans ──
inline bool smooth(uint32_t n, bool ladder) {
  static const uint32_t primes[] = {2, 3, 5, 7, 11, 13};
  const int np = ladder ? 2 : 6;
  for (int i = 0; i < np; ++i)
    while (n % primes[i] == 0) n /= primes[i];
  return n == 1;
}

inline uint32_t plan_size(uint32_t target, uint32_t div, bool ladder) {
  uint32_t n = std::max(target, 1u);
  while (true) {
    if (n % div == 0 && smooth(n, ladder)) return n;
    ++n;
  }
}

inline uint32_t largest_smooth(uint32_t max_n, bool ladder) {
  uint32_t n = std::max(32u, max_n);
  while (n >= 32) {
    if (smooth(n, ladder)) return n;
    --n;
  }
  return 32;
}

inline uint32_t js_round(double x) { return (uint32_t)std::floor(x + 0.5); }

struct GridPlan {
  uint32_t nx, ny, cw, ch;
};

inline GridPlan linear_corr_plan_rect(uint32_t w, uint32_t h, uint32_t max_n0, bool ladder) {
  const uint32_t max_n = largest_smooth(max_n0, ladder);
  double s = std::min(1.0, (double)std::max(8u, max_n >> 1) / (double)std::max(std::max(w, h), 1u));
  for (int i = 0; i < 16; ++i) {
    const uint32_t cw = std::max(8u, js_round(w * s));
    const uint32_t ch = std::max(8u, js_round(h * s));
    const uint32_t nx = std::min(plan_size(std::max(cw + cw - 1, 32u), 1, ladder), max_n);
    const uint32_t ny = std::min(plan_size(std::max(ch + ch - 1, 32u), 1, ladder), max_n);
    if (nx >= cw + cw - 1 && ny >= ch + ch - 1) return {nx, ny, cw, ch};
    s *= 0.85;
  }
  const uint32_t c = std::max(8u, (max_n + 1) >> 1);
  return {max_n, max_n, c, c};
}

inline uint32_t pow2_ceil(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

inline GridPlan auto_shift_plan(uint32_t w, uint32_t h) {
  const uint32_t d = std::max({w, h, 1u});
  uint32_t g = 1;
  while (2 * ((d + g - 1) / g) - 1 > 512) g *= 2;
  const uint32_t cw = std::max(1u, (w + g - 1) / g), ch = std::max(1u, (h + g - 1) / g);
  auto p2 = [](uint32_t need) { return std::min(std::max(pow2_ceil(std::max(need, 1u)), 128u), 512u); };
  return {p2(2 * cw - 1), p2(2 * ch - 1), cw, ch};
}

}  // namespace zc
