// Device code shared by the kernel translation units (k_*.hip): block reductions, the
// canvas resample and real planes, the row-FFT tiling and the FP64 per-lag score.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>
#include <vector>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <cmath>
#include <algorithm>
#include <stdexcept>

#include "fft.hpp"
#include "fft_regs.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace zc {

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// ── block reductions (wave64 shuffles, then LDS across waves, fixed order) ──────────────────
template <typename V, int M, int T>
__device__ __forceinline__ void block_sum(V (&v)[M], V* out, int tid) {
  constexpr int W = T / 64;
  __shared__ V sm[W][M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    V x = v[m];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    v[m] = x;
  }
  const int lane = tid & 63, wv = tid >> 6;
  if (lane == 0) {
#pragma unroll
    for (int m = 0; m < M; ++m) sm[wv][m] = v[m];
  }
  __syncthreads();
  for (int m = tid; m < M; m += T) {
    V s = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) s += sm[w][m];
    out[m] = s;
  }
}


__device__ __forceinline__ constexpr int tri_k(int i, int j) {
  return (i < j ? i : j) * 4 - ((i < j ? i : j) * ((i < j ? i : j) + 1)) / 2 + (i < j ? j : i);
}


// ── maps: pack + row FFT ──────────────────────────────────────────────────────────────────
// identity: 1 the cells are the canvas pixels; 2 they are an image's pixels moved by (sx, sy) (a
// stack of W × H planes; cells outside the image are 0): an integer translation onto the canvas
// read without a canvas copy; 0 the k×k box resample
struct Resample {
  int W, H, cw, ch, kx, ky, identity;
  float fx, fy;
  int sx, sy;
};

__device__ __forceinline__ float sample_plane(const float* __restrict__ s, int W, int H, float px, float py) {
  const float x = px - 0.5f, y = py - 0.5f;
  const float x0f = floorf(x), y0f = floorf(y);
  const float ax = x - x0f, ay = y - y0f;
  const int x0 = (int)x0f, y0 = (int)y0f;
  const int xa = min(max(x0, 0), W - 1), xb = min(max(x0 + 1, 0), W - 1);
  const int ya = min(max(y0, 0), H - 1), yb = min(max(y0 + 1, 0), H - 1);
  const float a = s[(int64_t)ya * W + xa], b = s[(int64_t)ya * W + xb];
  const float c = s[(int64_t)yb * W + xa], d = s[(int64_t)yb * W + xb];
  return (a * (1.0f - ax) + b * ax) * (1.0f - ay) + (c * (1.0f - ax) + d * ax) * ay;
}

// The five stack values of grid cell (x, y): the canvas pixel, or the mean of k×k bilinear
// taps over the cell's footprint (smi.wgsl:678-705 pack_resample).
__device__ __forceinline__ void cell_values(const float* __restrict__ canvas, const Resample& r, int x, int y,
                                            float* s) {
  const int64_t n = (int64_t)r.W * r.H;
  if (r.identity == 2) {
    const int ix = x - r.sx, iy = y - r.sy;
    const bool in = ix >= 0 && iy >= 0 && ix < r.W && iy < r.H;
#pragma unroll
    for (int p = 0; p < 5; ++p) s[p] = in ? canvas[p * n + (int64_t)iy * r.W + ix] : 0.0f;
    return;
  }
  if (r.identity) {
#pragma unroll
    for (int p = 0; p < 5; ++p) s[p] = canvas[p * n + (int64_t)y * r.W + x];
    return;
  }
#pragma unroll
  for (int p = 0; p < 5; ++p) s[p] = 0.0f;
  for (int j = 0; j < r.ky; ++j) {
    const float py = ((float)y + ((float)j + 0.5f) / (float)r.ky) * r.fy;
    for (int i = 0; i < r.kx; ++i) {
      const float px = ((float)x + ((float)i + 0.5f) / (float)r.kx) * r.fx;
#pragma unroll
      for (int p = 0; p < 5; ++p) s[p] += sample_plane(canvas + p * n, r.W, r.H, px, py);
    }
  }
  const float inv = 1.0f / (float)(r.kx * r.ky);
#pragma unroll
  for (int p = 0; p < 5; ++p) s[p] *= inv;
}

// Real plane `pl` (0 w, 1-4 w f_i, 5-14 w f_i f_j) from stack samples (smi.wgsl:643-649).
__device__ __forceinline__ float real_plane(int pl, const float* s) {
  if (pl < 5) return s[pl];
  if (s[0] <= 1e-6f) return 0.0f;
  constexpr int TI[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
  constexpr int TJ[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
  const int k = pl - 5;
  return s[1 + TI[k]] * s[1 + TJ[k]] / s[0];
}

// Row kernels: LINES rows of length L per workgroup, M = 16 points per thread (L >= 1024: one
// row, L / 16 threads); lines of 2^a·3^b points (1536, 2304, 3072) take 12 per thread.
template <int L>
struct RowT {
  static constexpr int LINES = L >= 1024 ? 1 : 1024 / L;
  static constexpr int M = L % 3 == 0 ? 12 : 16;
  static constexpr int T = LINES * L / M;
  static constexpr int STR = padded_len(L);
};


// Row tilings of the inverse row pass (zc_fft_tune's rows value) for lines of >= 256 points:
// 0 (default) 8 points per lane at 256 (0.287 -> 0.270 ms for a 128²-canvas sweep batch), two
// 1024-point lines per workgroup at 16 points per lane, 32 points per lane from 2048 (0.237 ->
// 0.225 ms for the 15 planes of a 2048² map, 1.54 -> 1.33 ms for a sweep batch's 4096-point
// planes), RowT's tiling otherwise; 1 32 points per lane; 2 8 points
// per lane; 3 16 points per lane with RowT's lines per workgroup; 4 16 points per lane, twice
// RowT's lines. Every workgroup holds at least one wave. Shorter lines always take RowT's
// tiling. Lines whose radix-16/8 stages split the same way at 16 and 32 points per lane (2048,
// 4096) give the same results bit for bit at either.
template <int L, int V>
struct RowX {
  static constexpr bool WIDE = V == 1 || (V == 0 && L >= 2048);
  // lines of 2^a·3^b points: 48, 12 or 24 points per thread in place of 32, 8 or 16; a line of
  // 9 · 2^k points takes 12 (2304: 192 lanes and radices 12·12·16, where 48 points left a
  // 96-lane workgroup of two lines)
  static constexpr int M = L % 3 == 0 ? (L % 9 == 0 ? 12 : WIDE ? 48 : V == 2 ? 12 : 24)
                                      : (WIDE ? 32 : (V == 2 || (V == 0 && L == 256)) ? 8 : 16);
  static constexpr int TWO = (V == 4 || (V == 0 && L == 1024)) ? 2 : 1;
  static constexpr int MIN_LINES = (64 * M + L - 1) / L;
  static constexpr int LINES =
      RowT<L>::LINES * TWO > MIN_LINES ? RowT<L>::LINES * TWO : MIN_LINES;
  static constexpr int T = LINES * L / M;
  static constexpr int STR = RowT<L>::STR;
};

// ── per-lag score ─────────────────────────────────────────────────────────────────────────
__device__ __forceinline__ bool better(float v, uint32_t i, float bv, uint32_t bi) {
  return v > bv || (v == bv && i < bi);
}

// Wilson–Hilferty normal score of an F(d1, d2) value (smi.wgsl:816-822).
__device__ __forceinline__ double f_to_z(double f, double d1, double d2) {
  const double a = 2.0 / (9.0 * d1), b = 2.0 / (9.0 * d2);
  const double c = cbrt(fmax(f, 0.0));
  return ((1.0 - b) * c - (1.0 - a)) / sqrt(b * c * c + a);
}

// In-place Cholesky of the lower triangle of g (NF x NF); rd[i] = 1 / L_ii. False if not PD.
// FLAT: without branches — a pivot that is not positive marks the result false and is replaced by
// 1 so that the remaining steps stay finite (their values are then discarded); for a positive
// definite g every step is the plain one, so both forms give the same factor.
template <int NF, bool FLAT = false>
__device__ __forceinline__ bool chol(double (&g)[4][4], double (&rd)[4]) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < NF; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = g[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= g[i][k] * g[j][k];
      if (i == j) {
        if (!FLAT && !(s > 0.0)) return false;
        ok = ok && s > 0.0;
        s = s > 0.0 ? s : 1.0;
        const double r = rsqrt(s);   // 1 / L_ii without a division
        g[i][i] = s * r;
        rd[i] = r;
      } else {
        g[i][j] = s * rd[j];
      }
    }
  }
  return ok;
}

// The per-lag score (smi.wgsl:825-931 combine_cell) from the correlation planes read through
// C(number, lag): 0 below the overlap floor; FP64 throughout.
template <int PID, bool FLAT = false, typename CF>
__device__ __forceinline__ double lag_score(CF&& C, int64_t t, double floor_n, const CombineArgs& a, bool calib) {
  constexpr bool EXACT = PID & 1;
  constexpr int NF = (PID & 2) ? 3 : 4;
  const double n = C(0, t);
  const double neff = n / fmax(a.area, 1.0);
  double score = 0.0;
  if (n < floor_n || (calib && neff <= 10.0)) return 0.0;
  const double den = fmax(n, 1e-8), iden = 1.0 / den;
  double ma[4], mb[4];
#pragma unroll
  for (int i = 0; i < NF; ++i) {
    ma[i] = C(1 + i, t) * iden;
    mb[i] = C(5 + i, t) * iden;
  }
  if constexpr (!EXACT) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < NF; ++i)
#pragma unroll
      for (int j = 0; j < NF; ++j) {
        const double c = fmin(fmax(C(9 + i * 4 + j, t) * iden - ma[i] * mb[j], -1.0), 1.0);
        s += c * c;
      }
    score = n * s;
  } else {
    double ga[4][4], gb[4][4];
    double tra = 0.0, trb = 0.0;
#pragma unroll
    for (int i = 0; i < NF; ++i) {
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        ga[i][j] = C(25 + tri_k(i, j), t) * iden - ma[i] * ma[j];
        gb[i][j] = C(35 + tri_k(i, j), t) * iden - mb[i] * mb[j];
      }
      tra += ga[i][i];
      trb += gb[i][i];
    }
    const double ra = a.ridge * fmax(tra * 0.25, 1e-12), rb = a.ridge * fmax(trb * 0.25, 1e-12);
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      ga[i][i] += ra;
      gb[i][i] += rb;
    }
    double ra_[4], rb_[4];
    bool pd;
    if constexpr (FLAT) {
      const bool pda = chol<NF, true>(ga, ra_), pdb = chol<NF, true>(gb, rb_);
      pd = pda && pdb;
    } else {
      pd = chol<NF>(ga, ra_) && chol<NF>(gb, rb_);
    }
    if (pd) {
      // X = L_A⁻¹ C (columns), then Y = X L_B⁻ᵀ (rows), acc = ‖Y‖²
      double x[4][4];
#pragma unroll
      for (int j = 0; j < NF; ++j)
#pragma unroll
        for (int i = 0; i < NF; ++i) {
          double s = C(9 + i * 4 + j, t) * iden - ma[i] * mb[j];
#pragma unroll
          for (int k = 0; k < i; ++k) s -= ga[i][k] * x[k][j];
          x[i][j] = s * ra_[i];
        }
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < NF; ++r) {
        double y[4];
#pragma unroll
        for (int i = 0; i < NF; ++i) {
          double s = x[r][i];
#pragma unroll
          for (int k = 0; k < i; ++k) s -= gb[i][k] * y[k];
          y[i] = s * rb_[i];
          acc += y[i] * y[i];
        }
      }
      score = n * acc;
      if (calib) {
        const double q = (double)NF;
        const double v = fmin(acc, q - 1e-4);
        const double f = (neff - q - 1.0) / q * v / (q - v);
        score = fmax(f_to_z(f, q * q, q * (neff - q - 1.0)), 0.0);
      }
    }
  }
  if (!(score == score) || isinf(score)) score = 0.0;
  return score;
}


// The sweep's and the generic FFT's lengths: the powers of two, and the 2^a·3^b lengths 1536,
// 2304 and 3072 between 1024 and 4096 (2592 = 2^5·3^4 was measured slower per candidate than
// 3072: five radix stages, 432-lane column workgroups).
#define ZC_FFT_SWITCH(L, CALL)  \
  switch (L) {                  \
    case 32: CALL(32); break;   \
    case 64: CALL(64); break;   \
    case 128: CALL(128); break; \
    case 256: CALL(256); break; \
    case 512: CALL(512); break; \
    case 1024: CALL(1024); break; \
    case 1536: CALL(1536); break; \
    case 2048: CALL(2048); break; \
    case 2304: CALL(2304); break; \
    case 3072: CALL(3072); break; \
    case 4096: CALL(4096); break; \
    default: return false;      \
  }

#define ZC_POW2_SWITCH(L, CALL) \
  switch (L) {                  \
    case 32: CALL(32); break;   \
    case 64: CALL(64); break;   \
    case 128: CALL(128); break; \
    case 256: CALL(256); break; \
    case 512: CALL(512); break; \
    case 1024: CALL(1024); break; \
    case 2048: CALL(2048); break; \
    case 4096: CALL(4096); break; \
    default: return false;      \
  }

}  // namespace zc
