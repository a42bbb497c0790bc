// SMI hot-path kernels for gfx950 (CDNA4, wave64).
//
// Prepare (once per image): exact integer image statistics → bit-exact rank transform →
// Sobel/Laplacian features with per-block moments → FP64 ridge whitening → premultiplied stack;
// feature autocorrelation for calibrated maps.
// Map (per pose): canvas warp/paste → fused resample + real-plane pack + row FFT → column FFT →
// Hermitian-unpacking cross-spectrum (all correlations of a bin pair in one pass) → column IFFT →
// row IFFT + extract → FP64 per-lag whitened score + block argmax → argmax finish.
// Pose moments: the 45 overlap sums at a pose (FP32 two-sum compensated per thread, FP64 across).
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

// ── prepare ───────────────────────────────────────────────────────────────────────────────
__global__ void k_round_half(float* x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = __half2float(__float2half_rn(x[i]));
}

__global__ __launch_bounds__(256) void k_maxabs(const float* __restrict__ x, int64_t n, uint32_t* out) {
  uint32_t m = 0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    m = max(m, __float_as_uint(fabsf(x[i])));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_down((int)m, off, 64));
  // one atomic per workgroup (atomics on one address serialize)
  __shared__ uint32_t wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(out, max(max(wm[0], wm[1]), max(wm[2], wm[3])));
}

// Exact fixed-point sums: term = trunc(v · 2^K), v = x or f32(x - mean)^2 (exact in f64),
// split into a signed high and an unsigned low 32-bit part (integer atomics: order-free).
__global__ __launch_bounds__(256) void k_fixsum(const float* __restrict__ x, int64_t n, double scale, float mean,
                                                int squared, long long* hi_acc, unsigned long long* lo_acc) {
  long long hi = 0;
  unsigned long long lo = 0;
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double v;
    if (squared) {
      const double d = (double)__fsub_rn(x[i], mean);
      v = d * d;
    } else {
      v = (double)x[i];
    }
    const long long t = (long long)(v * scale);
    hi += t >> 32;
    lo += (unsigned long long)(t & 0xffffffffLL);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    hi += __shfl_down(hi, off, 64);
    lo += __shfl_down(lo, off, 64);
  }
  // one atomic pair per workgroup (the 64-bit atomics on one address serialize)
  __shared__ long long whi[4];
  __shared__ unsigned long long wlo[4];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    whi[w] = hi;
    wlo[w] = lo;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long*)hi_acc, (unsigned long long)(whi[0] + whi[1] + whi[2] + whi[3]));
    atomicAdd(lo_acc, wlo[0] + wlo[1] + wlo[2] + wlo[3]);
  }
}

__device__ __forceinline__ uint32_t rank_bin(float x, float mean, float spread) {
  const float d = __fsub_rn(x, mean);
  const float z = __fdiv_rn(d, fmaxf(spread, 1e-8f));
  const float t = __fmul_rn(__fadd_rn(z, 4.0f), 32.0f);
  return (uint32_t)fminf(fmaxf(t, 0.0f), 255.0f);
}

__global__ __launch_bounds__(256) void k_hist(const float* __restrict__ x, int64_t n, float mean, float spread,
                                              uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    atomicAdd(&h[rank_bin(x[i], mean, spread)], 1u);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_rank(const float* __restrict__ x, int64_t n, float mean, float spread,
                                              const uint32_t* __restrict__ hist, const uint32_t* __restrict__ cdf,
                                              float* rank) {
  const float tot = (float)max(cdf[256], 1u);
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const uint32_t b = rank_bin(x[i], mean, spread);
    const float c = (float)cdf[b], hb = (float)hist[b];
    const float r = __fdiv_rn(__fadd_rn(c, __fmul_rn(0.5f, hb)), tot);
    rank[i] = __fsub_rn(__fmul_rn(r, 2.0f), 1.0f);
  }
}

__device__ __forceinline__ constexpr int tri_k(int i, int j) {
  return (i < j ? i : j) * 4 - ((i < j ? i : j) * ((i < j ? i : j) + 1)) / 2 + (i < j ? j : i);
}

// [r, r², |∇r|, f3] (Sobel, clamped borders) and |∇r|; 16×16 tiles; per-block moments:
// Σ f_i (4) and Σ f_i f_j (i ≤ j, 10) in FP64.
__global__ __launch_bounds__(256) void k_feats(const float* __restrict__ rank, int w, int h, int mode,
                                               float* feat, float* mag, double* part) {
  __shared__ float tile[18][19];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
  for (int k = tid; k < 324; k += 256) {
    const int lx = k % 18, ly = k / 18;
    const int gx = min(max(x0 + lx - 1, 0), w - 1), gy = min(max(y0 + ly - 1, 0), h - 1);
    tile[ly][lx] = rank[(int64_t)gy * w + gx];
  }
  __syncthreads();
  double acc[14];
#pragma unroll
  for (int m = 0; m < 14; ++m) acc[m] = 0.0;
  const int gx = x0 + tx, gy = y0 + ty;
  if (gx < w && gy < h) {
    auto at = [&](int dx, int dy) { return tile[ty + 1 + dy][tx + 1 + dx]; };
    const float r = at(0, 0);
    const float sx = -at(-1, -1) + at(1, -1) - 2.0f * at(-1, 0) + 2.0f * at(1, 0) - at(-1, 1) + at(1, 1);
    const float sy = -at(-1, -1) - 2.0f * at(0, -1) - at(1, -1) + at(-1, 1) + 2.0f * at(0, 1) + at(1, 1);
    const float g = sqrtf(sx * sx + sy * sy);
    float f3 = 0.0f;
    if (mode == 0) f3 = at(-1, 0) + at(1, 0) + at(0, -1) + at(0, 1) - 4.0f * r;
    else if (mode == 2) f3 = r * r * r;
    const int64_t n = (int64_t)w * h, i = (int64_t)gy * w + gx;
    const float f[4] = {r, r * r, g, f3};
#pragma unroll
    for (int p = 0; p < 4; ++p) feat[p * n + i] = f[p];
    mag[i] = g;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      acc[a] = f[a];
#pragma unroll
      for (int b = a; b < 4; ++b) acc[4 + tri_k(a, b)] = (double)f[a] * (double)f[b];
    }
  }
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  block_sum<double, 14, 256>(acc, part + (int64_t)blk * 14, tid);
}

// Column sums of nparts rows of m doubles, in a fixed order (deterministic); blockIdx.y selects
// one of several such tables (table y at part + y · nparts · m, its sums at out + y · m).
// out[c] = Σ_r part[r · m + c]: one workgroup per column c (blockIdx.x); each thread sums rows
// t, t + 256, … in order, then a fixed tree.
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int nparts, int m,
                                                   double* out) {
  part += (int64_t)blockIdx.y * nparts * m;
  out += (int64_t)blockIdx.y * m;
  __shared__ double sm[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < nparts; r += 256) s += part[(int64_t)r * m + c];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sm[0];
}

// f ← L⁻¹ (f − μ) in FP64, and the stack [1, f0..f3].
__global__ __launch_bounds__(256) void k_whiten(float* feat, float* stack, int64_t n, WhitenArgs a) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double x[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) x[p] = (double)feat[p * n + i] - a.mu[p];
    stack[i] = 1.0f;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q <= p; ++q) s += a.X[p * 4 + q] * x[q];
      feat[p * n + i] = (float)s;
      stack[(p + 1) * n + i] = (float)s;
    }
  }
}

// Autocorrelation partial sums Σ_x f_i(x) f_i(x + τ), |τ| ≤ 8 (289 lags × 4 planes), over
// 32×32 tiles with an 8-px halo (zero outside the image, so whole tiles can be summed). Thread
// (plane p, row lag dy, third q of the tile's rows) slides a 17-value window of row y + dy along
// row y: each step reads one value of each row from LDS and updates all 17 column lags. One
// partial-sum row per (tile, third): part[(tile · ACF_Q + q) · 1156 + lag · 4 + p].
constexpr int ACF_T = 32, ACF_Q = 3;
__global__ __launch_bounds__(256) void k_acf(const float* __restrict__ feat, int w, int h, double* part) {
  constexpr int S = ACF_T + 16;
  __shared__ float t[4][S][S + 1];
  const int x0 = blockIdx.x * ACF_T, y0 = blockIdx.y * ACF_T;
  const int64_t n = (int64_t)w * h;
  for (int k = threadIdx.x; k < 4 * S * S; k += 256) {
    const int p = k / (S * S), r = k % (S * S), ly = r / S, lx = r % S;
    const int gx = x0 + lx - 8, gy = y0 + ly - 8;
    t[p][ly][lx] = (gx >= 0 && gy >= 0 && gx < w && gy < h) ? feat[p * n + (int64_t)gy * w + gx] : 0.0f;
  }
  __syncthreads();
  const int id = threadIdx.x;
  if (id >= 4 * 17 * ACF_Q) return;   // no barrier follows
  const int q = id % ACF_Q, r = id / ACF_Q, dyi = r % 17, p = r / 17;
  constexpr int ROWS = (ACF_T + ACF_Q - 1) / ACF_Q;
  const int ya = q * ROWS, yb = min(ACF_T, ya + ROWS);
  float acc[17];
#pragma unroll
  for (int i = 0; i < 17; ++i) acc[i] = 0.0f;
  for (int yy = ya; yy < yb; ++yy) {
    const float* r1 = &t[p][yy + 8][8];
    const float* r2 = &t[p][yy + dyi][0];   // row yy + 8 + dy, columns from x − 8
    float win[17];
#pragma unroll
    for (int i = 0; i < 17; ++i) win[i] = r2[i];
#pragma unroll
    for (int xx = 0; xx < ACF_T; ++xx) {
      const float a = r1[xx];
#pragma unroll
      for (int i = 0; i < 17; ++i) acc[i] += a * win[i];
      if (xx + 1 < ACF_T) {
#pragma unroll
        for (int i = 0; i < 16; ++i) win[i] = win[i + 1];
        win[16] = r2[xx + 17];
      }
    }
  }
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  double* out = part + ((int64_t)blk * ACF_Q + q) * 1156;
#pragma unroll
  for (int i = 0; i < 17; ++i) out[(dyi * 17 + i) * 4 + p] = (double)acc[i];
}

// ── maps: canvas ──────────────────────────────────────────────────────────────────────────
struct H9 {
  double m[9];
};

__global__ __launch_bounds__(256) void k_warp_stack(const float* __restrict__ stack, int sw, int sh, H9 Hi,
                                                    float* canvas, int cw, int ch) {
  const int gx = blockIdx.x * 16 + (threadIdx.x & 15), gy = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (gx >= cw || gy >= ch) return;
  const int64_t n = (int64_t)cw * ch, pix = (int64_t)gy * cw + gx, sn = (int64_t)sw * sh;
  const double X = gx + 1.0, Y = gy + 1.0;
  const double qx = Hi.m[0] * X + Hi.m[1] * Y + Hi.m[2];
  const double qy = Hi.m[3] * X + Hi.m[4] * Y + Hi.m[5];
  const double qz = Hi.m[6] * X + Hi.m[7] * Y + Hi.m[8];
  bool ok = fabs(qz) >= 1e-8;
  double fx = 0.0, fy = 0.0;
  if (ok) {
    fx = qx / qz - 1.0;
    fy = qy / qz - 1.0;
    ok = fx >= 0.0 && fy >= 0.0 && fx <= (double)(sw - 1) && fy <= (double)(sh - 1);   // false for NaN
  }
  int xa = 0, xb = 0, ya = 0, yb = 0;
  float tx = 0.f, ty = 0.f;
  if (ok) {
    const double x0 = floor(fx), y0 = floor(fy);
    tx = (float)(fx - x0);
    ty = (float)(fy - y0);
    xa = min(max((int)x0, 0), sw - 1);
    xb = min(max((int)x0 + 1, 0), sw - 1);
    ya = min(max((int)y0, 0), sh - 1);
    yb = min(max((int)y0 + 1, 0), sh - 1);
  }
#pragma unroll
  for (int p = 0; p < 5; ++p) {
    float v = 0.0f;
    if (ok) {
      const float* s = stack + p * sn;
      const float a = s[(int64_t)ya * sw + xa], b = s[(int64_t)ya * sw + xb];
      const float c = s[(int64_t)yb * sw + xa], d = s[(int64_t)yb * sw + xb];
      v = (a * (1.0f - tx) + b * tx) * (1.0f - ty) + (c * (1.0f - tx) + d * tx) * ty;
    }
    canvas[p * n + pix] = v;
  }
}

__global__ __launch_bounds__(256) void k_paste_stack(const float* __restrict__ stack, int sw, int sh, int ox, int oy,
                                                     float* canvas, int cw, int ch) {
  const int gx = blockIdx.x * 16 + (threadIdx.x & 15), gy = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (gx >= cw || gy >= ch) return;
  const int64_t n = (int64_t)cw * ch, pix = (int64_t)gy * cw + gx, sn = (int64_t)sw * sh;
  const int sx = gx + ox, sy = gy + oy;
  const bool inside = sx >= 0 && sy >= 0 && sx < sw && sy < sh;
#pragma unroll
  for (int p = 0; p < 5; ++p) canvas[p * n + pix] = inside ? stack[p * sn + (int64_t)sy * sw + sx] : 0.0f;
}

// ── maps: pack + row FFT ──────────────────────────────────────────────────────────────────
struct Resample {
  int W, H, cw, ch, kx, ky, identity;
  float fx, fy;
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
// row, L / 16 threads).
template <int L>
struct RowT {
  static constexpr int LINES = L >= 1024 ? 1 : 1024 / L;
  static constexpr int M = 16;
  static constexpr int T = LINES * L / M;
  static constexpr int STR = padded_len(L);
};

// Resample + real-plane pack + forward row FFT: rows y < ch of the grid; writes n_cplx spectra
// rows (plane stride L · ny). The stack values of a thread's 16 points are read once.
// TI: the side table (pack_table): which real planes this side transforms.
template <int L, int TI>
__global__ __launch_bounds__(RowT<L>::T) void k_pack_rowfft(const float* __restrict__ canvas, Resample r, int ny,
                                                            float2* spec, const float2* __restrict__ tw) {
  constexpr PackTable P = pack_table(TI);
  constexpr int T = RowT<L>::T, M = RowT<L>::M, LINES = RowT<L>::LINES, STR = RowT<L>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * ny;
  float s[M][5];
  {
    float2 dummy[M];
    int idx = 0;
    rload<L, M, T, LINES, false>(dummy, tid, [&](int line, int x) {
      const int y = y0 + line;
      if (x < r.cw && y < r.ch) cell_values(canvas, r, x, y, s[idx]);
      else {
#pragma unroll
        for (int p = 0; p < 5; ++p) s[idx][p] = 0.0f;
      }
      ++idx;
      return make_float2(0.0f, 0.0f);
    });
  }
#pragma unroll
  for (int c = 0; c < P.n_cplx; ++c) {
    float2 v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const float re = real_plane(P.slot_plane[2 * c], s[i]);
      const float im = (2 * c + 1 < P.n_real) ? real_plane(P.slot_plane[2 * c + 1], s[i]) : 0.0f;
      v[i] = make_float2(re, im);
    }
    rfft<L, M, T, LINES, false, -1>(v, lds, STR, tw, tid);
    float2* dst = spec + c * N;
    rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
      const int y = y0 + line;
      if (y < r.ch) dst[(int64_t)y * L + x] = val;
    });
    __syncthreads();   // the next plane's first exchange reuses the LDS
  }
}

// ── roto-scale maps: log-polar pack ───────────────────────────────────────────────────────
// Cell (x, y) of the log-polar grid samples the canvas stack at c + rho_y (cos th_x, sin th_x)
// (pixel i at i; 0 unless all four bilinear taps are inside: smi.wgsl:1418-1452), then every real
// plane is multiplied by rho_y (the area element, smi.wgsl:718-720), and the row FFT follows.
// Rows y < n_lam hold data; cells x >= n_th are 0 (a padded angle axis).
__device__ __forceinline__ void logpolar_values(const LogPolarArgs& a, int x, int y, float* s) {
  const double rho = a.rho[y];
  const double xs = a.cx + rho * a.cth[x], ys = a.cy + rho * a.sth[x];
  const double x0d = floor(xs), y0d = floor(ys);
  const int x0 = (int)x0d, y0 = (int)y0d;
  if (!(xs >= 0.0 && ys >= 0.0) || x0 + 1 >= a.W || y0 + 1 >= a.H) {
#pragma unroll
    for (int p = 0; p < 5; ++p) s[p] = 0.0f;
    return;
  }
  const float fx = (float)(xs - x0d), fy = (float)(ys - y0d);
  const int64_t n = (int64_t)a.W * a.H, o = (int64_t)y0 * a.W + x0;
#pragma unroll
  for (int p = 0; p < 5; ++p) {
    const float* c = a.canvas + p * n + o;
    const float v0 = c[0] * (1.0f - fx) + c[1] * fx;
    const float v1 = c[a.W] * (1.0f - fx) + c[a.W + 1] * fx;
    s[p] = v0 * (1.0f - fy) + v1 * fy;
  }
}

template <int L, int PID>
__global__ __launch_bounds__(RowT<L>::T) void k_pack_logpolar(LogPolarArgs a, int ny, float2* spec,
                                                              const float2* __restrict__ tw) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int T = RowT<L>::T, M = RowT<L>::M, LINES = RowT<L>::LINES, STR = RowT<L>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * ny;
  float s[M][5];
  float om[M];
  {
    float2 dummy[M];
    int idx = 0;
    rload<L, M, T, LINES, false>(dummy, tid, [&](int line, int x) {
      const int y = y0 + line;
      if (x < a.n_th && y < a.n_lam) {
        logpolar_values(a, x, y, s[idx]);
        om[idx] = (float)a.rho[y];
      } else {
#pragma unroll
        for (int p = 0; p < 5; ++p) s[idx][p] = 0.0f;
        om[idx] = 0.0f;
      }
      ++idx;
      return make_float2(0.0f, 0.0f);
    });
  }
#pragma unroll
  for (int c = 0; c < P.n_cplx; ++c) {
    float2 v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const float re = real_plane(P.slot_plane[2 * c], s[i]) * om[i];
      const float im = (2 * c + 1 < P.n_real) ? real_plane(P.slot_plane[2 * c + 1], s[i]) * om[i] : 0.0f;
      v[i] = make_float2(re, im);
    }
    rfft<L, M, T, LINES, false, -1>(v, lds, STR, tw, tid);
    float2* dst = spec + c * N;
    rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
      const int y = y0 + line;
      if (y < a.n_lam) dst[(int64_t)y * L + x] = val;
    });
    __syncthreads();
  }
}

// rho_y = exp(log0 + (y + 1/2) / n_lam · span), (cos, sin) of th_x = (x + 1/2) / n_th · 2 pi, FP64.
__global__ __launch_bounds__(256) void k_rs_tables(int n_th, int n_lam, double log0, double span, double* rho,
                                                   double* cth, double* sth) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n_lam) rho[i] = exp(log0 + ((double)i + 0.5) / (double)n_lam * span);
  if (i < n_th) {
    const double th = ((double)i + 0.5) / (double)n_th * 6.283185307179586476925;
    cth[i] = cos(th);
    sth[i] = sin(th);
  }
}

// Generic in-place row FFT of `rows` rows (tests, benchmarks).
template <int L, int S>
__global__ __launch_bounds__(RowT<L>::T) void k_rowfft(float2* data, int rows, float scale,
                                                       const float2* __restrict__ tw) {
  constexpr int T = RowT<L>::T, M = RowT<L>::M, LINES = RowT<L>::LINES, STR = RowT<L>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * LINES;
  float2 v[M];
  rload<L, M, T, LINES, false>(v, tid, [&](int line, int x) {
    const int y = y0 + line;
    return y < rows ? data[(int64_t)y * L + x] : make_float2(0.0f, 0.0f);
  });
  rfft<L, M, T, LINES, false, S>(v, lds, STR, tw, tid);
  rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
    const int y = y0 + line;
    if (y < rows) data[(int64_t)y * L + x] = make_float2(val.x * scale, val.y * scale);
  });
}

// ── column FFT ────────────────────────────────────────────────────────────────────────────
// C adjacent columns per workgroup (line-fastest mapping: a wave reads C·8-byte row segments),
// 16 points per thread; logical blocks remapped so column neighbours share an XCD's L2.
template <int L, int C>
struct ColT {
  static constexpr int M = 16;
  static constexpr int T = C * L / M;
  static constexpr int STR = line_stride(L, C);
  static constexpr bool fits = C * STR * 8 <= 160 * 1024 && T <= 1024 && T >= 16;
};

// lines (columns) per workgroup of the column FFT: the default, or a tuned value that fits
// the widest row segments whose LDS fits (measured best at 2048: 8 columns, 1 workgroup per CU)
inline int col_lines_default(int L) { return L >= 4096 ? 4 : (L >= 2048 ? 8 : 16); }
static_assert(ColT<32, 16>::fits && ColT<64, 16>::fits && ColT<128, 16>::fits && ColT<256, 16>::fits &&
                  ColT<512, 16>::fits && ColT<1024, 16>::fits && ColT<2048, 8>::fits && ColT<4096, 4>::fits,
              "every default column tiling must be launchable");

template <int L, int S, int C>
__global__ __launch_bounds__((ColT<L, C>::T)) void k_colfft(float2* data, int nx, int planes, int rows_in, float scale,
                                                          const float2* __restrict__ tw) {
  constexpr int T = ColT<L, C>::T, M = ColT<L, C>::M, STR = ColT<L, C>::STR;
  __shared__ float2 lds[C * STR];
  const int tid = threadIdx.x;
  const int groups = nx / C;
  const int lb = xcd_remap((int)blockIdx.x, groups * planes);
  const int plane = lb / groups, x0 = (lb - plane * groups) * C;
  float2* base = data + (int64_t)plane * nx * L + x0;
  float2 v[M];
  rload<L, M, T, C, true>(v, tid, [&](int c, int y) {
    return y < rows_in ? base[(unsigned)(y * nx + c)] : make_float2(0.0f, 0.0f);
  });
  rfft<L, M, T, C, true, S>(v, lds, STR, tw, tid);
  rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
    base[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
  });
}

// 2·C columns per workgroup in two groups of C sharing one LDS buffer (see rfft_groups).
template <int L, int S, int C>
__global__ __launch_bounds__((ColT<L, C>::T)) void k_colfft_g2(float2* data, int nx, int planes, int rows_in,
                                                               float scale, const float2* __restrict__ tw) {
  constexpr int T = ColT<L, C>::T, M = ColT<L, C>::M, STR = ColT<L, C>::STR;
  __shared__ float2 lds[C * STR];
  const int tid = threadIdx.x;
  const int groups = nx / (2 * C);
  const int lb = xcd_remap((int)blockIdx.x, groups * planes);
  const int plane = lb / groups, x0 = (lb - plane * groups) * 2 * C;
  float2* base = data + (int64_t)plane * nx * L + x0;
  float2 v[2][M];
#pragma unroll
  for (int g = 0; g < 2; ++g)
    rload<L, M, T, C, true>(v[g], tid, [&](int c, int y) {
      return y < rows_in ? base[(unsigned)(y * nx + g * C + c)] : make_float2(0.0f, 0.0f);
    });
  rfft_groups<L, M, T, C, true, S, 2>(v, lds, STR, tw, tid);
#pragma unroll
  for (int g = 0; g < 2; ++g)
    rstore<L, M, T, C, true>(v[g], tid, [&](int c, int y, float2 val) {
      base[(unsigned)(y * nx + g * C + c)] = make_float2(val.x * scale, val.y * scale);
    });
}

template <int L, int S, int C>
static bool colfft_g2_LC(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw,
                         hipStream_t st) {
  if constexpr (!ColT<L, C>::fits) {
    return false;
  } else {
    if (nx % (2 * C)) return false;
    const unsigned g = (unsigned)(nx / (2 * C) * planes);
    k_colfft_g2<L, S, C><<<g, ColT<L, C>::T, 0, st>>>(data, nx, planes, rows_in, scale, tw);
    return true;
  }
}

// Streaming column FFT: one workgroup per C-column group runs the FFTs of that group in every
// plane (planes [0, pa) of A, then [0, pb) of B), loading plane i+1 into registers while it
// transforms plane i, so HBM traffic overlaps the LDS exchanges of a 1-workgroup-per-CU kernel.
template <int L, int S, int C>
__global__ __launch_bounds__((ColT<L, C>::T)) void k_colfft_stream(float2* A, int pa, float2* B, int pb, int nx,
                                                                   int rows_in, float scale,
                                                                   const float2* __restrict__ tw) {
  constexpr int T = ColT<L, C>::T, M = ColT<L, C>::M, STR = ColT<L, C>::STR;
  __shared__ float2 lds[C * STR];
  const int tid = threadIdx.x;
  const int groups = nx / C;
  const int x0 = xcd_remap((int)blockIdx.x, groups) * C;
  const int64_t plane_sz = (int64_t)nx * L;
  const int total = pa + pb;
  auto base_of = [&](int i) { return (i < pa ? A + (int64_t)i * plane_sz : B + (int64_t)(i - pa) * plane_sz) + x0; };
  float2 v[M], nxt[M];
  {
    const float2* b0 = base_of(0);
    rload<L, M, T, C, true>(v, tid, [&](int c, int y) {
      return y < rows_in ? b0[(unsigned)(y * nx + c)] : make_float2(0.0f, 0.0f);
    });
  }
  for (int i = 0; i < total; ++i) {
    if (i + 1 < total) {
      const float2* bn = base_of(i + 1);
      rload<L, M, T, C, true>(nxt, tid, [&](int c, int y) {
        return y < rows_in ? bn[(unsigned)(y * nx + c)] : make_float2(0.0f, 0.0f);
      });
    }
    rfft<L, M, T, C, true, S>(v, lds, STR, tw, tid);
    float2* bi = base_of(i);
    rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
      bi[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
    });
    __syncthreads();   // the next plane's first exchange reuses the LDS
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = nxt[k];
  }
}

template <int L, int S, int C>
static bool colfft_stream_LC(float2* A, int pa, float2* B, int pb, int nx, int rows_in, float scale,
                             const float2* tw, hipStream_t st) {
  if constexpr (!ColT<L, C>::fits) {
    return false;
  } else {
    if (nx % C) return false;
    k_colfft_stream<L, S, C><<<(unsigned)(nx / C), ColT<L, C>::T, 0, st>>>(A, pa, B, pb, nx, rows_in, scale, tw);
    return true;
  }
}

// Wide column tilings for long columns (L >= 1024): M points per thread, and optionally the
// exchange split into real and imaginary halves (half the LDS: 16 columns of 2048 = 128-byte row
// segments, or 8 columns of 4096, in one CU's LDS).
template <int L, int C, int M, bool SPLIT>
struct ColW {
  static constexpr int T = C * L / M;
  static constexpr int STR = line_stride(L, C);
  static constexpr bool fits =
      L >= 1024 && C * STR * (SPLIT ? 4 : 8) <= 160 * 1024 && T <= 1024 && T >= 64 && (C * L) % M == 0 &&
      (M < 32 || T <= 512) &&   // 32 complex points per lane need more than the 128 VGPRs of a 1024-lane block
      (M < 64 || T == 512);     // 64 points: 128 VGPRs of data, two waves per SIMD
};

template <int L, int S, int C, int M, bool SPLIT>
__global__ __launch_bounds__((ColW<L, C, M, SPLIT>::T), (SPLIT && M == 16 ? 2 : 1)) void k_colfft_w(float2* data, int nx, int planes, int rows_in,
                                                                       float scale, const float2* __restrict__ tw) {
  constexpr int T = ColW<L, C, M, SPLIT>::T, STR = ColW<L, C, M, SPLIT>::STR;
  const int tid = threadIdx.x;
  const int groups = nx / C;
  const int lb = xcd_remap((int)blockIdx.x, groups * planes);
  const int plane = lb / groups, x0 = (lb - plane * groups) * C;
  float2* base = data + (int64_t)plane * nx * L + x0;
  float2 v[M];
  rload<L, M, T, C, true>(v, tid, [&](int c, int y) {
    return y < rows_in ? base[(unsigned)(y * nx + c)] : make_float2(0.0f, 0.0f);
  });
  if constexpr (SPLIT) {
    __shared__ float lds[C * STR];
    rfft_split<L, M, T, C, true, S>(v, lds, STR, tw, tid);
  } else {
    __shared__ float2 lds[C * STR];
    rfft<L, M, T, C, true, S>(v, lds, STR, tw, tid);
  }
  rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
    base[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
  });
}

template <int L, int S, int C, int M, bool SPLIT>
static bool colfft_w_LC(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw, hipStream_t st) {
  if constexpr (!ColW<L, C, M, SPLIT>::fits) {
    return false;
  } else {
    if (nx % C) return false;
    const unsigned g = (unsigned)(nx / C * planes);
    k_colfft_w<L, S, C, M, SPLIT><<<g, ColW<L, C, M, SPLIT>::T, 0, st>>>(data, nx, planes, rows_in, scale, tw);
    return true;
  }
}

template <int L, int S, int C>
static bool colfft_LC(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw, hipStream_t st) {
  if constexpr (!ColT<L, C>::fits) {
    return false;
  } else {
    if (nx % C) return false;
    const unsigned g = (unsigned)(nx / C * planes);
    k_colfft<L, S, C><<<g, ColT<L, C>::T, 0, st>>>(data, nx, planes, rows_in, scale, tw);
    return true;
  }
}

template <int L, int S>
static bool colfft_L(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw, int lines,
                     hipStream_t st) {
  switch (lines > 0 ? lines : col_lines_default(L)) {
    case 1: return colfft_LC<L, S, 1>(data, planes, nx, rows_in, scale, tw, st);
    case 2: return colfft_LC<L, S, 2>(data, planes, nx, rows_in, scale, tw, st);
    case 4: return colfft_LC<L, S, 4>(data, planes, nx, rows_in, scale, tw, st);
    case 8: return colfft_LC<L, S, 8>(data, planes, nx, rows_in, scale, tw, st);
    case 16: return colfft_LC<L, S, 16>(data, planes, nx, rows_in, scale, tw, st);
    // two groups sharing the LDS: 16 = 2 x 8, 32 = 2 x 16 columns
    case 17: return colfft_g2_LC<L, S, 8>(data, planes, nx, rows_in, scale, tw, st);
    case 33: return colfft_g2_LC<L, S, 16>(data, planes, nx, rows_in, scale, tw, st);
    case 9: return colfft_g2_LC<L, S, 4>(data, planes, nx, rows_in, scale, tw, st);
    // streaming over planes (one workgroup per column group): 64 + columns
    case 68: return colfft_stream_LC<L, S, 4>(data, planes, nullptr, 0, nx, rows_in, scale, tw, st);
    case 72: return colfft_stream_LC<L, S, 8>(data, planes, nullptr, 0, nx, rows_in, scale, tw, st);
    case 80: return colfft_stream_LC<L, S, 16>(data, planes, nullptr, 0, nx, rows_in, scale, tw, st);
    // wide tilings (L >= 1024), 16 or 32 points per thread, split exchange unless noted:
    // 41: 4x16, 42: 8x16, 43: 8x32, 44: 16x32, 45: 8x32 plain, 46: 4x32, 47: 2x32 plain,
    // 48: 16x64, 49: 8x64
    case 41: return colfft_w_LC<L, S, 4, 16, true>(data, planes, nx, rows_in, scale, tw, st);
    case 42: return colfft_w_LC<L, S, 8, 16, true>(data, planes, nx, rows_in, scale, tw, st);
    case 43: return colfft_w_LC<L, S, 8, 32, true>(data, planes, nx, rows_in, scale, tw, st);
    case 44: return colfft_w_LC<L, S, 16, 32, true>(data, planes, nx, rows_in, scale, tw, st);
    case 45: return colfft_w_LC<L, S, 8, 32, false>(data, planes, nx, rows_in, scale, tw, st);
    case 46: return colfft_w_LC<L, S, 4, 32, true>(data, planes, nx, rows_in, scale, tw, st);
    case 47: return colfft_w_LC<L, S, 2, 32, false>(data, planes, nx, rows_in, scale, tw, st);
    case 48: return colfft_w_LC<L, S, 16, 64, true>(data, planes, nx, rows_in, scale, tw, st);
    case 49: return colfft_w_LC<L, S, 8, 64, true>(data, planes, nx, rows_in, scale, tw, st);
    default: return false;
  }
}

// ── cross spectrum ────────────────────────────────────────────────────────────────────────
// Spectrum of real slot `odd ? 2p+1 : 2p` at bin k from Z = FFT(x_2p + i x_2p+1) at k and −k,
// times 2 (the ½ is folded into the output scale): even: Z + conj Z₋, odd: (Z − conj Z₋)/i.
__device__ __forceinline__ float2 unpack2(float2 z, float2 zm, bool odd) {
  return odd ? make_float2(z.y + zm.y, zm.x - z.x) : make_float2(z.x + zm.x, z.y - zm.y);
}

template <int PID>
__global__ __launch_bounds__(256) void k_cmul(const float2* __restrict__ A, const float2* __restrict__ B,
                                              float2* work, int nx, int ny, float scale) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int NC = P.n_cplx, NCB = P.b_n_cplx;
  const int64_t N = (int64_t)nx * ny;
  // blockIdx.y: the batch entry (its NC spectra in A, its n_out outputs in work; B is shared)
  A += (int64_t)blockIdx.y * NC * N;
  work += (int64_t)blockIdx.y * P.n_out * N;
  const int64_t i = blockIdx.x * 256 + threadIdx.x;
  const int64_t half_rows = ny / 2 + 1;
  if (i >= half_rows * nx) return;
  const int ky = (int)(i / nx), kx = (int)(i - (int64_t)ky * nx);
  const int my = (ny - ky) % ny, mx = (nx - kx) % nx;
  if ((ky == 0 || 2 * ky == ny) && kx > nx / 2) return;   // its mirror bin covers it
  const int64_t k = (int64_t)ky * nx + kx, km = (int64_t)my * nx + mx;
  const bool self = k == km;
  float2 za[NC], zam[NC], zb[NCB], zbm[NCB];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    za[c] = A[c * N + k];
    zam[c] = A[c * N + km];
  }
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    zb[c] = B[c * N + k];
    zbm[c] = B[c * N + km];
  }
  const float sc = 0.25f * scale;
#pragma unroll
  for (int q = 0; q < P.n_out; ++q) {
    const int c1 = 2 * q, c2 = 2 * q + 1;
    const int a1 = P.ca[c1], b1 = P.cb[c1];
    float2 xa = unpack2(za[a1 >> 1], zam[a1 >> 1], a1 & 1), xb = unpack2(zb[b1 >> 1], zbm[b1 >> 1], b1 & 1);
    float2 ya = unpack2(zam[a1 >> 1], za[a1 >> 1], a1 & 1), yb = unpack2(zbm[b1 >> 1], zb[b1 >> 1], b1 & 1);
    float2 p = cmulc(xa, xb), pm = cmulc(ya, yb);
    if (c2 < P.n_corr) {
      const int a2 = P.ca[c2], b2 = P.cb[c2];
      const float2 xa2 = unpack2(za[a2 >> 1], zam[a2 >> 1], a2 & 1), xb2 = unpack2(zb[b2 >> 1], zbm[b2 >> 1], b2 & 1);
      const float2 ya2 = unpack2(zam[a2 >> 1], za[a2 >> 1], a2 & 1), yb2 = unpack2(zbm[b2 >> 1], zb[b2 >> 1], b2 & 1);
      const float2 p2 = cmulc(xa2, xb2), pm2 = cmulc(ya2, yb2);
      p = make_float2(p.x - p2.y, p.y + p2.x);       // p + i p2
      pm = make_float2(pm.x - pm2.y, pm.y + pm2.x);
    }
    work[q * N + k] = make_float2(p.x * sc, p.y * sc);
    if (!self) work[q * N + km] = make_float2(pm.x * sc, pm.y * sc);
  }
}

// ── row IFFT + extract ────────────────────────────────────────────────────────────────────
// Inverse row FFT of complex output q (rows of L) → correlation planes 2q (re) and 2q+1 (im).
// blockIdx.y = b · n_out + q for batch entry b (its planes at b · n_corr).
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
  static constexpr int M = WIDE ? 32 : (V == 2 || (V == 0 && L == 256)) ? 8 : 16;
  static constexpr int TWO = (V == 4 || (V == 0 && L == 1024)) ? 2 : 1;
  static constexpr int MIN_LINES = (64 * M + L - 1) / L;
  static constexpr int LINES =
      RowT<L>::LINES * TWO > MIN_LINES ? RowT<L>::LINES * TWO : MIN_LINES;
  static constexpr int T = LINES * L / M;
  static constexpr int STR = RowT<L>::STR;
};

template <int L, int PID, int V>
__global__ __launch_bounds__((RowX<L, V>::T)) void k_rowifft_extract(const float2* __restrict__ work, float* corr,
                                                                    int ny, const float2* __restrict__ tw) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int T = RowX<L, V>::T, M = RowX<L, V>::M, LINES = RowX<L, V>::LINES, STR = RowX<L, V>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x, g = blockIdx.y;
  const int q = g % P.n_out, b = g / P.n_out;
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * ny;
  const float2* src = work + (int64_t)g * N;
  corr += (int64_t)b * P.n_corr * N;
  float2 v[M];
  rload<L, M, T, LINES, false>(v, tid, [&](int line, int x) {
    const int y = y0 + line;
    return y < ny ? src[(int64_t)y * L + x] : make_float2(0.0f, 0.0f);
  });
  rfft<L, M, T, LINES, false, 1>(v, lds, STR, tw, tid);
  float* o1 = corr + (int64_t)(2 * q) * N;
  float* o2 = corr + (int64_t)(2 * q + 1) * N;
  const bool two = 2 * q + 1 < P.n_corr;
  rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
    const int y = y0 + line;
    if (y < ny) {
      o1[(int64_t)y * L + x] = val.x;
      if (two) o2[(int64_t)y * L + x] = val.y;
    }
  });
}

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
template <int NF>
__device__ __forceinline__ bool chol(double (&g)[4][4], double (&rd)[4]) {
#pragma unroll
  for (int i = 0; i < NF; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double s = g[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= g[i][k] * g[j][k];
      if (i == j) {
        if (!(s > 0.0)) return false;
        const double r = rsqrt(s);   // 1 / L_ii without a division
        g[i][i] = s * r;
        rd[i] = r;
      } else {
        g[i][j] = s * rd[j];
      }
    }
  }
  return true;
}

// The per-lag score (smi.wgsl:825-931 combine_cell) from the correlation planes read through
// C(number, lag): 0 below the overlap floor; FP64 throughout.
template <int PID, typename CF>
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
    if (chol<NF>(ga, ra_) && chol<NF>(gb, rb_)) {
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

template <int PID>
__global__ __launch_bounds__(256) void k_combine(const float* __restrict__ corr, float* map, int64_t nn,
                                                 CombineArgs a, float* arg_part) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  auto C = [&](int num, int64_t t) -> double { return (double)corr[(int64_t)P.kept_of[num] * nn + t]; };
  const double n0 = C(0, 0);
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < nn; t += (int64_t)gridDim.x * 256) {
    const float sv = (float)lag_score<PID>(C, t, floor_n, a, calib);
    map[t] = sv;
    if (better(sv, (uint32_t)t, bv, bi)) {
      bv = sv;
      bi = (uint32_t)t;
    }
  }
  // block argmax
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

// ── rectangle sums ────────────────────────────────────────────────────────────────────────
// When the fixed side's weight plane is the indicator of the rectangle R (the fixed image pasted
// at full resolution), correlation (a, 0) at lag t is the sum of plane a over R + t clipped to the
// canvas: four reads of plane a's summed-area table S(y, x) = sum over rows < y, columns < x
// ((H+1) x (W+1) doubles per plane). The tables are built in three passes over bands of kSatBand
// rows: every table's column sums per band (k_sat_band: one thread per column, all tables from
// one read of the canvas), their exclusive prefix over bands (k_sat_carry), then every
// (band, table) by a workgroup that scans its rows with a running column carry (k_sat_rows).
constexpr int kSatBand = 8;   // rows per band

// The real planes of a canvas stack sample in FP64: w, w f_i, and w f_i f_j = s_i s_j / s_0
// (0 where s_0 <= 1e-6), the last with one reciprocal per sample.
// pl is a compile-time constant at every call (the plan's table → plane map), so every index
// below folds and the sample stays in registers.
__device__ __forceinline__ double real_plane_d(int pl, const float* s, double inv0) {
  if (pl < 5) return (double)s[pl];
  if (s[0] <= 1e-6f) return 0.0;
  constexpr int TI[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
  constexpr int TJ[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
  const int k = pl - 5;
  return (double)s[1 + TI[k]] * (double)s[1 + TJ[k]] * inv0;
}

__device__ __forceinline__ double canvas_inv0(const float* s) { return s[0] > 1e-6f ? 1.0 / (double)s[0] : 0.0; }

__device__ __forceinline__ void canvas_values(const float* __restrict__ canvas, int64_t n, int64_t at, float* s) {
#pragma unroll
  for (int p = 0; p < 5; ++p) s[p] = canvas[p * n + at];
}

// blockIdx.x band, blockIdx.y a chunk of 256 columns: band[(k · nb + b) · W + x] = column sums of
// table k's plane over the band's rows, for every table.
template <int PID>
__global__ __launch_bounds__(256) void k_sat_band(const float* __restrict__ canvas, int W, int H, double* band) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int n_tab = P.n_sat;
  const int b = blockIdx.x, nb = gridDim.x;
  const int x = blockIdx.y * 256 + threadIdx.x;
  if (x >= W) return;
  const int64_t n = (int64_t)W * H;
  const int y0 = b * kSatBand, y1 = min(H, y0 + kSatBand);
  double acc[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) acc[k] = 0.0;
  float sv[kSatBand][5];
#pragma unroll
  for (int i = 0; i < kSatBand; ++i) {   // every row's loads in flight together
    const int y = min(y0 + i, H - 1);
    canvas_values(canvas, n, (int64_t)y * W + x, sv[i]);
  }
#pragma unroll
  for (int i = 0; i < kSatBand; ++i) {
    if (y0 + i < y1) {
      const double inv0 = canvas_inv0(sv[i]);
#pragma unroll
      for (int k = 0; k < 15; ++k)
        if (k < n_tab) acc[k] += real_plane_d(P.sat_plane[k], sv[i], inv0);
    }
  }
#pragma unroll
  for (int k = 0; k < 15; ++k)
    if (k < n_tab) band[((int64_t)k * nb + b) * W + x] = acc[k];
}

// In place: band sums → exclusive prefix over bands, per (table, column). blockIdx.x: table ·
// column chunks of 64; the block's 16 waves take a sixteenth of the bands each.
constexpr int kCarryWaves = 16;
__global__ __launch_bounds__(64 * kCarryWaves) void k_sat_carry(double* band, int W, int nb) {
  __shared__ double seg[kCarryWaves][64];
  const int chunks = (W + 63) / 64;
  const int k = blockIdx.x / chunks, x = (blockIdx.x - k * chunks) * 64 + (threadIdx.x & 63);
  const int q = threadIdx.x >> 6;
  const int per = (nb + kCarryWaves - 1) / kCarryWaves, b0 = q * per, b1 = min(nb, b0 + per);
  double* col = band + (int64_t)k * nb * W + x;
  double tot = 0.0;
  if (x < W)
    for (int b = b0; b < b1; ++b) tot += col[(int64_t)b * W];
  seg[q][threadIdx.x & 63] = tot;
  __syncthreads();
  double acc = 0.0;
  for (int j = 0; j < q; ++j) acc += seg[j][threadIdx.x & 63];
  if (x < W)
    for (int b = b0; b < b1; ++b) {
      const double v = col[(int64_t)b * W];
      col[(int64_t)b * W] = acc;
      acc += v;
    }
}

// Exclusive prefix of G values per thread over the threads of the block, each value its own
// scan (lds: G doubles per wave).
template <int G>
__device__ __forceinline__ void block_excl_scan(double (&v)[G], double* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double inc[G];
#pragma unroll
  for (int g = 0; g < G; ++g) inc[g] = v[g];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const double u = __shfl_up(inc[g], o, 64);
      if (lane >= o) inc[g] += u;
    }
  }
  if (lane == 63)
#pragma unroll
    for (int g = 0; g < G; ++g) lds[w * G + g] = inc[g];
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    double off = 0.0;
    for (int j = 0; j < w; ++j) off += lds[j * G + g];
    v[g] = off + inc[g] - v[g];
  }
  __syncthreads();   // lds is reused by the next scan
}

// blockIdx.x band, blockIdx.y a group of G tables (one read of the canvas for all of them);
// thread t owns columns [t·E, t·E + E) for the scans; each finished row goes out through LDS so
// that lane i stores column x + i (whole cache lines per store instruction).
template <int PID, int NT, int E, int G, int K0>
__device__ __forceinline__ void sat_rows_body(const float* __restrict__ canvas, int W, int H,
                                              const double* __restrict__ band, double* sat, double* lds,
                                              double (*stage)[NT * E]) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int n_tab = P.n_sat, k0 = K0;
  const int b = blockIdx.x, nb = gridDim.x;
  const int xs = threadIdx.x * E;
  const int64_t n = (int64_t)W * H, SW = W + 1, SP = SW * (H + 1);
  constexpr int pl[G] = {P.sat_plane[k0 < n_tab ? k0 : 0], P.sat_plane[k0 + 1 < n_tab ? k0 + 1 : 0],
                         P.sat_plane[k0 + 2 < n_tab ? k0 + 2 : 0]};
  static_assert(G == 3, "three tables per workgroup");
  // column carry cc[g][e] = S_g(y, xs + e + 1): the row prefix of the earlier bands' column sums
  double cc[G][E];
  {
    double tot[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      tot[g] = 0.0;
      const double* crow = band + ((int64_t)min(k0 + g, n_tab - 1) * nb + b) * W;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        tot[g] += xs + e < W ? crow[xs + e] : 0.0;
        cc[g][e] = tot[g];
      }
    }
    block_excl_scan<G>(tot, lds);
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int e = 0; e < E; ++e) cc[g][e] += tot[g];
  }
  if (b == 0)
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (k0 + g < n_tab)
        for (int x = threadIdx.x; x <= W; x += NT) sat[(k0 + g) * SP + x] = 0.0;
  const int y0 = b * kSatBand, y1 = min(H, y0 + kSatBand);
  // the canvas values of the thread's columns, one row ahead of the scans (E <= 4: the registers
  // of a second row fit)
  constexpr bool AHEAD = E <= 4;
  float nxt[E][5];
#define ZC_SAT_LOAD_ROW(YY)                                                                         \
  _Pragma("unroll") for (int e = 0; e < E; ++e) {                                                    \
    _Pragma("unroll") for (int p = 0; p < 5; ++p) nxt[e][p] = 0.0f;                                  \
    if ((YY) < y1 && xs + e < W)                                                                     \
      _Pragma("unroll") for (int p = 0; p < 5; ++p) nxt[e][p] = canvas[p * n + (int64_t)(YY) * W + xs + e]; \
  }
  if (AHEAD) ZC_SAT_LOAD_ROW(y0)
  for (int y = y0; y < y1; ++y) {
    if (!AHEAD) ZC_SAT_LOAD_ROW(y)
    float cur[E][5];
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int p = 0; p < 5; ++p) cur[e][p] = nxt[e][p];
    if (AHEAD) ZC_SAT_LOAD_ROW(y + 1)
    double v[G][E];
    double tot[G];
#pragma unroll
    for (int g = 0; g < G; ++g) tot[g] = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float* sv = cur[e];
      const double inv0 = canvas_inv0(sv);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        tot[g] += real_plane_d(pl[g], sv, inv0);
        v[g][e] = tot[g];
      }
    }
    block_excl_scan<G>(tot, lds);   // its barriers also order the previous row's stage reads
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        cc[g][e] += tot[g] + v[g][e];
        stage[g][xs + e] = cc[g][e];
      }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (k0 + g >= n_tab) continue;
      double* row = sat + (k0 + g) * SP + (int64_t)(y + 1) * SW;
      for (int x = threadIdx.x; x < W; x += NT) row[x + 1] = stage[g][x];
      if (threadIdx.x == 0) row[0] = 0.0;
    }
  }
#undef ZC_SAT_LOAD_ROW
}

// blockIdx.y: the group of G tables (tables G·y .. G·y + G − 1 of the plan); NT threads of E
// columns each cover rows of up to NT·E cells.
template <int PID, int NT, int E, int G>
__global__ __launch_bounds__(NT) void k_sat_rows(const float* __restrict__ canvas, int W, int H,
                                                 const double* __restrict__ band, double* sat) {
  __shared__ double lds[NT / 64 * G];
  __shared__ double stage[G][NT * E];
  static_assert(kPlans[PID].n_sat <= 5 * G, "at most five groups");
  switch (blockIdx.y) {
    case 0: sat_rows_body<PID, NT, E, G, 0>(canvas, W, H, band, sat, lds, stage); break;
    case 1: sat_rows_body<PID, NT, E, G, G>(canvas, W, H, band, sat, lds, stage); break;
    case 2: sat_rows_body<PID, NT, E, G, 2 * G>(canvas, W, H, band, sat, lds, stage); break;
    case 3: sat_rows_body<PID, NT, E, G, 3 * G>(canvas, W, H, band, sat, lds, stage); break;
    default: sat_rows_body<PID, NT, E, G, 4 * G>(canvas, W, H, band, sat, lds, stage); break;
  }
}

int sat_bands(int H) { return (H + kSatBand - 1) / kSatBand; }

template <int PID>
static void sat_build_P(const float* canvas, int W, int H, double* band, double* sat, hipStream_t st) {
  constexpr int n_tab = kPlans[PID].n_sat;
  const int nb = sat_bands(H);
  k_sat_band<PID><<<dim3(nb, (W + 255) / 256), 256, 0, st>>>(canvas, W, H, band);
  k_sat_carry<<<(unsigned)(n_tab * ((W + 63) / 64)), 64 * kCarryWaves, 0, st>>>(band, W, nb);
  constexpr int G = 3;   // tables per workgroup
  const dim3 g(nb, (n_tab + G - 1) / G);
  if (W <= 256 * 4) k_sat_rows<PID, 256, 4, G><<<g, 256, 0, st>>>(canvas, W, H, band, sat);
  else k_sat_rows<PID, 512, 4, G><<<g, 512, 0, st>>>(canvas, W, H, band, sat);
}

bool launch_sat_build(int plan, const float* canvas, int W, int H, double* band, double* sat, hipStream_t st) {
  if (W < 1 || H < 1 || W > 256 * 8) return false;
  switch (plan) {
    case 4: sat_build_P<4>(canvas, W, H, band, sat, st); return true;
    case 5: sat_build_P<5>(canvas, W, H, band, sat, st); return true;
    case 6: sat_build_P<6>(canvas, W, H, band, sat, st); return true;
    case 7: sat_build_P<7>(canvas, W, H, band, sat, st); return true;
    default: return false;
  }
}

// Table offsets of the rectangle R + (dx, dy) clipped to the canvas for lag index t.
struct SatIdx {
  int64_t o00, o01, o10, o11;   // S(y0, x0), S(y0, x1), S(y1, x0), S(y1, x1)
  bool empty;
  bool r0, c0;                  // y0 = 0 / x0 = 0: those corners lie on the table's zero row / column
};

__device__ __forceinline__ SatIdx sat_idx(const SatView& v, int64_t t) {
  const int px = (int)(t & (v.nx - 1)), py = (int)(t >> v.lx);
  const int dx = px <= v.nx / 2 ? px : px - v.nx;
  const int dy = py <= v.ny / 2 ? py : py - v.ny;
  const int x0 = min(max(v.bx0 + dx, 0), v.W), x1 = min(max(v.bx1 + dx, 0), v.W);
  const int y0 = min(max(v.by0 + dy, 0), v.H), y1 = min(max(v.by1 + dy, 0), v.H);
  SatIdx q;
  const int64_t SW = v.W + 1;
  q.empty = x1 <= x0 || y1 <= y0;
  q.r0 = y0 == 0;
  q.c0 = x0 == 0;
  q.o00 = y0 * SW + x0;
  q.o01 = y0 * SW + x1;
  q.o10 = y1 * SW + x0;
  q.o11 = y1 * SW + x1;
  return q;
}

// Corners on the zero row or column are not read (half the lags of a map have y0 = 0 or x0 = 0).
__device__ __forceinline__ double sat_sum(const double* __restrict__ S, const SatIdx& q) {
  if (q.empty) return 0.0;
  const double s11 = S[q.o11];
  const double s01 = q.r0 ? 0.0 : S[q.o01];
  const double s10 = q.c0 ? 0.0 : S[q.o10];
  const double s00 = (q.r0 || q.c0) ? 0.0 : S[q.o00];
  return (s11 - s01) - (s10 - s00);
}

// k_combine for the rectangle plans: correlations with sat_of >= 0 come from the tables.
template <int PID>
__global__ __launch_bounds__(256) void k_combine_rect(const float* __restrict__ corr, float* map, int64_t nn,
                                                      CombineArgs a, SatView sv, float* arg_part) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  const int64_t SP = (int64_t)(sv.W + 1) * (sv.H + 1);
  SatIdx q = sat_idx(sv, 0);
  auto C = [&](int num, int64_t t) -> double {
    if (P.sat_of[num] >= 0) return sat_sum(sv.sat + P.sat_of[num] * SP, q);
    return (double)corr[(int64_t)P.kept_of[num] * nn + t];
  };
  const double n0 = C(0, 0);
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < nn; t += (int64_t)gridDim.x * 256) {
    q = sat_idx(sv, t);
    const float s = (float)lag_score<PID & 3>(C, t, floor_n, a, calib);
    map[t] = s;
    if (better(s, (uint32_t)t, bv, bi)) {
      bv = s;
      bi = (uint32_t)t;
    }
  }
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

// ── roto-scale maps: per-lag score and finish ─────────────────────────────────────────────
// Lag t of the n x n map: row k (signed, lambda), column kx (theta). The correlations were taken
// on a px x py grid: px = n (circular in theta) or a padded length >= 2n, whose two linear lags kx
// and kx - n fold into the circular one; py >= n (linear in lambda). Each correlation is scaled by
// e^(k dlam) dlam dth (extract_pair's area element, smi.wgsl:778-796).
template <int PID>
__global__ __launch_bounds__(256) void k_combine_rs(const float* __restrict__ corr, float* map, RsArgs r,
                                                    CombineArgs a) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  const int64_t nn = (int64_t)r.px * r.py, total = (int64_t)r.n * r.n;
  int64_t base = 0, base2 = -1;
  double sc = 0.0;
  auto at = [&](int64_t t) {
    const int ky = (int)(t / r.n), kx = (int)(t - (int64_t)ky * r.n);
    const int k = ky > r.n / 2 ? ky - r.n : ky;
    const int row = k < 0 ? k + r.py : k;
    base = (int64_t)row * r.px + kx;
    base2 = (r.px != r.n && kx > 0) ? base - r.n + r.px : -1;
    sc = exp((double)k * r.dlam) * r.dlam * r.dth;
  };
  auto C = [&](int num, int64_t) -> double {
    const float* c = corr + (int64_t)P.kept_of[num] * nn;
    double v = (double)c[base];
    if (base2 >= 0) v += (double)c[base2];
    return v * sc;
  };
  at(0);
  const double n0 = C(0, 0);
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    at(t);
    map[t] = (float)lag_score<PID>(C, t, floor_n, a, calib);
  }
}

// fwd averaged with inv at the mirrored lag (-y, -x), or (-y, x) when !mirror; rows whose applied
// scale exp(-k dlam) lies outside [e^lo, e^hi] zeroed (smi.wgsl:1708-1731); block argmax partials.
__global__ __launch_bounds__(256) void k_rs_finish(const float* __restrict__ fwd, const float* __restrict__ inv,
                                                   int n, int mirror, double dlam, double lo, double hi, float* out,
                                                   float* arg_part) {
  const int64_t total = (int64_t)n * n;
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int y = (int)(t / n), x = (int)(t - (int64_t)y * n);
    float v = fwd[t];
    if (inv) {
      const int ym = (n - y) % n, xm = mirror ? (n - x) % n : x;
      v = 0.5f * (v + inv[(int64_t)ym * n + xm]);
    }
    const int k = y > n / 2 ? y - n : y;
    const double sg = -(double)k * dlam;
    if (sg < lo || sg > hi) v = 0.0f;
    out[t] = v;
    if (better(v, (uint32_t)t, bv, bi)) {
      bv = v;
      bi = (uint32_t)t;
    }
  }
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

__global__ __launch_bounds__(256) void k_argmax_finish(const float* __restrict__ part, int nparts,
                                                       const float* __restrict__ map, float* res) {
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int g = threadIdx.x; g < nparts; g += 256) {
    const float v = part[2 * g];
    const uint32_t i = __float_as_uint(part[2 * g + 1]);
    if (i != 0xffffffffu && better(v, i, bv, bi)) {
      bv = v;
      bi = i;
    }
  }
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[0] = sv_[0];
    res[1] = __uint_as_float(si_[0]);
    res[2] = map[0];
  }
}

// ── pose moments ──────────────────────────────────────────────────────────────────────────
// FP32 two-sum (Knuth) accumulation: s + c carries the exact running sum to ~2^-48.
__device__ __forceinline__ void two_sum_add(float& s, float& c, float x) {
  const float t = s + x;
  const float z = t - s;
  c += (s - (t - z)) + (x - z);
  s = t;
}

__device__ __forceinline__ float ramp_w(float v, float lo, float hi) {
  const float a = (v - lo) * 0.5f, b = (hi - v) * 0.5f;
  return fminf(fmaxf(a, 0.0f), 1.0f) * fminf(fmaxf(b, 0.0f), 1.0f);
}

// The 45 overlap sums of a pose (smi.wgsl:980-1057 overlap_moments): moving features fa sampled
// bilinearly at Hi·X for every destination pixel X of fb, tapered over 2 px inside the moving
// image; per-thread FP32 two-sum, then FP64 across the block into part[0..45).
__device__ __forceinline__ void pose_moments_body(const float* __restrict__ fa, int wa, int ha,
                                                  const float* __restrict__ fb, int wb, int hb, const double* Hi,
                                                  double* part, int blk, int nblk) {
  const PoseArgs p = {{Hi[0], Hi[1], Hi[2], Hi[3], Hi[4], Hi[5], Hi[6], Hi[7], Hi[8]}};
  float s[45], c[45];
#pragma unroll
  for (int m = 0; m < 45; ++m) s[m] = c[m] = 0.0f;
  const int64_t na = (int64_t)wa * ha, nb = (int64_t)wb * hb;
  for (int64_t t = (int64_t)blk * 256 + threadIdx.x; t < nb; t += (int64_t)nblk * 256) {
    const int x = (int)(t % wb), y = (int)(t / wb);
    const double X = x + 1.0, Y = y + 1.0;
    const double qz = p.Hi[6] * X + p.Hi[7] * Y + p.Hi[8];
    if (fabs(qz) < 1e-8) continue;
    const double fxd = (p.Hi[0] * X + p.Hi[1] * Y + p.Hi[2]) / qz - 1.0;
    const double fyd = (p.Hi[3] * X + p.Hi[4] * Y + p.Hi[5]) / qz - 1.0;
    if (!(fxd >= 2.0 && fyd >= 2.0 && fxd <= wa - 3.0 && fyd <= ha - 3.0)) continue;   // NaN fails too
    const float fx = (float)fxd, fy = (float)fyd;
    const float wt = ramp_w(fx, 2.0f, (float)wa - 3.0f) * ramp_w(fy, 2.0f, (float)ha - 3.0f);
    if (!(wt > 0.0f)) continue;
    const int x0 = (int)floor(fxd), y0 = (int)floor(fyd);
    const float ax = (float)(fxd - x0), ay = (float)(fyd - y0);
    const int xa = min(max(x0, 0), wa - 1), xb = min(max(x0 + 1, 0), wa - 1);
    const int ya = min(max(y0, 0), ha - 1), yb = min(max(y0 + 1, 0), ha - 1);
    float va[4], vb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* f = fa + q * na;
      const float v00 = f[(int64_t)ya * wa + xa], v10 = f[(int64_t)ya * wa + xb];
      const float v01 = f[(int64_t)yb * wa + xa], v11 = f[(int64_t)yb * wa + xb];
      va[q] = (v00 * (1.0f - ax) + v10 * ax) * (1.0f - ay) + (v01 * (1.0f - ax) + v11 * ax) * ay;
      vb[q] = fb[q * nb + t];
    }
    two_sum_add(s[0], c[0], wt);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      two_sum_add(s[1 + i], c[1 + i], wt * va[i]);
      two_sum_add(s[5 + i], c[5 + i], wt * vb[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) two_sum_add(s[9 + i * 4 + j], c[9 + i * 4 + j], wt * va[i] * vb[j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = i; j < 4; ++j) {
        two_sum_add(s[25 + tri_k(i, j)], c[25 + tri_k(i, j)], wt * va[i] * va[j]);
        two_sum_add(s[35 + tri_k(i, j)], c[35 + tri_k(i, j)], wt * vb[i] * vb[j]);
      }
  }
  double acc[45];
#pragma unroll
  for (int m = 0; m < 45; ++m) acc[m] = (double)s[m] + (double)c[m];
  block_sum<double, 45, 256>(acc, part, threadIdx.x);
}

__global__ __launch_bounds__(256) void k_pose_moments(const float* __restrict__ fa, int wa, int ha,
                                                      const float* __restrict__ fb, int wb, int hb, PoseArgs p,
                                                      double* part) {
  pose_moments_body(fa, wa, ha, fb, wb, hb, p.Hi, part + (int64_t)blockIdx.x * 45, blockIdx.x, gridDim.x);
}

// One pose per blockIdx.y: its partials at part[(y · gridDim.x + x) · 45].
__global__ __launch_bounds__(256) void k_pose_moments_jobs(const PoseJob* __restrict__ jobs, double* part) {
  const PoseJob& j = jobs[blockIdx.y];
  pose_moments_body(j.fa, j.wa, j.ha, j.fb, j.wb, j.hb, j.Hi,
                    part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 45, blockIdx.x, gridDim.x);
}

// The score's tangent gradient and Gauss–Newton matrix at a pose (smi.wgsl:1129-1272
// tangent_grad), one pose per blockIdx.y: per destination pixel, ∂S/∂v_a = 2 (W t_b − M̂ t_a)
// dotted with the sample's gradients gives (cx, cy) (plus the taper's boundary term), times each
// tangent's ∂(sample)/∂ξ_k from M_k = −H⁻¹ G_k; the matrix Σ wt J_aᵀ Q J_a from the four
// scalars g_xᵀ Q g_x, g_xᵀ Q g_y, g_yᵀ Q g_x, g_yᵀ Q g_y. Partials: 8 gradient, then the upper
// triangle row by row (36), FP64 across the block, at part[(y · gridDim.x + x) · 44].
__global__ __launch_bounds__(256) void k_tangent_jobs(const TangentJob* __restrict__ jobs, double* part) {
  const TangentJob& j = jobs[blockIdx.y];
  const int wa = j.wa, ha = j.ha, wb = j.wb;
  const int64_t na = (int64_t)wa * ha, nb = (int64_t)wb * j.hb;
  const float sw = (float)wa, sh = (float)ha;
  float acc[44];
#pragma unroll
  for (int m = 0; m < 44; ++m) acc[m] = 0.0f;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nb; t += (int64_t)gridDim.x * 256) {
    const int x = (int)(t % wb), y = (int)(t / wb);
    const float X0 = (float)x + 1.0f, X1 = (float)y + 1.0f;
    // the sample position in FP64: the bilinear gradients jump at cell edges, so the cell must
    // not depend on the rounding of the position
    const double pxd = j.Hi[0] * (double)X0 + j.Hi[1] * (double)X1 + j.Hi[2];
    const double pyd = j.Hi[3] * (double)X0 + j.Hi[4] * (double)X1 + j.Hi[5];
    const double pzd = j.Hi[6] * (double)X0 + j.Hi[7] * (double)X1 + j.Hi[8];
    if (fabs(pzd) < 1e-8) continue;
    const double fxd = pxd / pzd - 1.0, fyd = pyd / pzd - 1.0;
    if (!(fxd >= 2.0 && fyd >= 2.0 && fxd <= wa - 3.0 && fyd <= ha - 3.0)) continue;   // NaN fails too
    const float phx = (float)pxd, phy = (float)pyd, phz = (float)pzd;
    const float fx = (float)fxd, fy = (float)fyd;
    // taper (smi.wgsl:338-351): value and its derivatives along x and y
    float tvx, dtx, tvy, dty;
    {
      const float a = (fx - 2.0f) * 0.5f, b = (sw - 3.0f - fx) * 0.5f;
      const float ca = fminf(fmaxf(a, 0.0f), 1.0f), cb = fminf(fmaxf(b, 0.0f), 1.0f);
      const float da = (a > 0.0f && a < 1.0f) ? 0.5f : 0.0f, db = (b > 0.0f && b < 1.0f) ? -0.5f : 0.0f;
      tvx = ca * cb;
      dtx = da * cb + ca * db;
    }
    {
      const float a = (fy - 2.0f) * 0.5f, b = (sh - 3.0f - fy) * 0.5f;
      const float ca = fminf(fmaxf(a, 0.0f), 1.0f), cb = fminf(fmaxf(b, 0.0f), 1.0f);
      const float da = (a > 0.0f && a < 1.0f) ? 0.5f : 0.0f, db = (b > 0.0f && b < 1.0f) ? -0.5f : 0.0f;
      tvy = ca * cb;
      dty = da * cb + ca * db;
    }
    const float wt = tvx * tvy, twx = dtx * tvy, twy = tvx * dty;
    const int x0 = (int)fxd, y0 = (int)fyd;   // >= 2: truncation is floor
    const float ax = (float)(fxd - x0), ay = (float)(fyd - y0);
    const int64_t i00 = (int64_t)y0 * wa + x0, i01 = i00 + wa;
    float va[4], gx[4], gy[4], vb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* f = j.fa + q * na;
      const float v00 = f[i00], v10 = f[i00 + 1], v01 = f[i01], v11 = f[i01 + 1];
      va[q] = (v00 * (1.0f - ax) + v10 * ax) * (1.0f - ay) + (v01 * (1.0f - ax) + v11 * ax) * ay;
      gx[q] = (v10 - v00) * (1.0f - ay) + (v11 - v01) * ay;
      gy[q] = (v01 - v00) * (1.0f - ax) + (v11 - v10) * ax;
      vb[q] = j.fb[q * nb + t];
    }
    float ra[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      float s = 0.0f;
#pragma unroll
      for (int b = 0; b < 4; ++b) s += j.W[a * 4 + b] * (vb[b] - j.mb[b]) - j.Mh[a * 4 + b] * (va[b] - j.ma[b]);
      ra[a] = 2.0f * s;
    }
    float bw = 0.0f;
    if (twx != 0.0f || twy != 0.0f) {   // ∂S/∂w: Σ_k g_k φ_k(v_a, v_b) over the 45 monomials
      float s = j.g[0];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        s += j.g[1 + i] * va[i] + j.g[5 + i] * vb[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) s += j.g[9 + i * 4 + k] * va[i] * vb[k];
#pragma unroll
        for (int k = i; k < 4; ++k) s += j.g[25 + tri_k(i, k)] * va[i] * va[k] + j.g[35 + tri_k(i, k)] * vb[i] * vb[k];
      }
      bw = s;
    }
    const float cx = wt * (ra[0] * gx[0] + ra[1] * gx[1] + ra[2] * gx[2] + ra[3] * gx[3]) + bw * twx;
    const float cy = wt * (ra[0] * gy[0] + ra[1] * gy[1] + ra[2] * gy[2] + ra[3] * gy[3]) + bw * twy;
    const float iz2 = 1.0f / fmaxf(phz * phz, 1e-12f);
    float px[8], py[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float* M = j.Mk + 9 * k;
      const float w0 = M[0] * X0 + M[1] * X1 + M[2];
      const float w1 = M[3] * X0 + M[4] * X1 + M[5];
      const float w2 = M[6] * X0 + M[7] * X1 + M[8];
      px[k] = (w0 * phz - phx * w2) * iz2;
      py[k] = (w1 * phz - phy * w2) * iz2;
      acc[k] += cx * px[k] + cy * py[k];
    }
    if (j.hess) {
      float qgx[4], qgy[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        qgx[a] = j.Q[a * 4] * gx[0] + j.Q[a * 4 + 1] * gx[1] + j.Q[a * 4 + 2] * gx[2] + j.Q[a * 4 + 3] * gx[3];
        qgy[a] = j.Q[a * 4] * gy[0] + j.Q[a * 4 + 1] * gy[1] + j.Q[a * 4 + 2] * gy[2] + j.Q[a * 4 + 3] * gy[3];
      }
      const float sxx = gx[0] * qgx[0] + gx[1] * qgx[1] + gx[2] * qgx[2] + gx[3] * qgx[3];
      const float sxy = gx[0] * qgy[0] + gx[1] * qgy[1] + gx[2] * qgy[2] + gx[3] * qgy[3];
      const float syx = gy[0] * qgx[0] + gy[1] * qgx[1] + gy[2] * qgx[2] + gy[3] * qgx[3];
      const float syy = gy[0] * qgy[0] + gy[1] * qgy[1] + gy[2] * qgy[2] + gy[3] * qgy[3];
      float r[8], q[8];
#pragma unroll
      for (int l = 0; l < 8; ++l) {
        r[l] = wt * (sxx * px[l] + sxy * py[l]);
        q[l] = wt * (syx * px[l] + syy * py[l]);
      }
      int h = 8;
#pragma unroll
      for (int k = 0; k < 8; ++k)
#pragma unroll
        for (int l = k; l < 8; ++l) acc[h++] += px[k] * r[l] + py[k] * q[l];
    }
  }
  double accd[44];
#pragma unroll
  for (int m = 0; m < 44; ++m) accd[m] = (double)acc[m];
  block_sum<double, 44, 256>(accd, part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 44, threadIdx.x);
}

// ── rotation × scale sweep (sweep.wgsl:52-371) ─────────────────────────────────────────────
// Features interleaved per pixel (f0..f3 as one float4). The reference's stack carries a fifth
// weight plane w = 1 (smi.zig:175); bilinear taps are all-or-nothing (zero unless every corner
// is inside), so a tap's w contribution is its weight sum and needs no storage.
__global__ __launch_bounds__(256) void k_interleave(const float* __restrict__ feat, int64_t n, float4* il) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    il[i] = make_float4(feat[i], feat[n + i], feat[2 * n + i], feat[3 * n + i]);
}

// One canvas cell's five stack values for a candidate: the mean of k×k bilinear taps of the
// source at H⁻¹ of the cell's tap points (sweep.wgsl:64-85 tap, 120-145 pack_sweep).
__device__ __forceinline__ void sweep_cell(const float4* __restrict__ il, const SweepPack& u, const SweepPose& pz,
                                           int cx, int cy, float* s) {
  const int k = max(pz.taps, 1);
  const float fk = (float)k;
  float a0 = 0.0f;
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float xmax = (float)(u.sw - 1), ymax = (float)(u.sh - 1);
  {
    // A cell whose four corners map (with one sign of the homogeneous coordinate) beyond one
    // edge of the source, by more than rounding, has no valid tap: its values are zero.
    float cxs[4], cys[4];
    bool same = true;
    float zs = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float px = u.ox + (float)(cx + (c & 1)) * u.cell, py = u.oy + (float)(cy + (c >> 1)) * u.cell;
      const float qz = pz.h[6] * px + pz.h[7] * py + pz.h[8];
      if (c == 0) zs = qz;
      same &= (qz > 1e-8f && zs > 1e-8f) || (qz < -1e-8f && zs < -1e-8f);
      cxs[c] = (pz.h[0] * px + pz.h[1] * py + pz.h[2]) / qz - 1.0f;
      cys[c] = (pz.h[3] * px + pz.h[4] * py + pz.h[5]) / qz - 1.0f;
    }
    if (same) {
      const float m = 1e-3f;
      const bool left = cxs[0] < -m && cxs[1] < -m && cxs[2] < -m && cxs[3] < -m;
      const bool right = cxs[0] > xmax + m && cxs[1] > xmax + m && cxs[2] > xmax + m && cxs[3] > xmax + m;
      const bool top = cys[0] < -m && cys[1] < -m && cys[2] < -m && cys[3] < -m;
      const bool bottom = cys[0] > ymax + m && cys[1] > ymax + m && cys[2] > ymax + m && cys[3] > ymax + m;
      if (left || right || top || bottom) {
#pragma unroll
        for (int p = 0; p < 5; ++p) s[p] = 0.0f;
        return;
      }
    }
  }
  for (int ty = 0; ty < k; ++ty) {
    const float py = u.oy + ((float)cy + ((float)ty + 0.5f) / fk) * u.cell;
    for (int tx = 0; tx < k; ++tx) {
      const float px = u.ox + ((float)cx + ((float)tx + 0.5f) / fk) * u.cell;
      const float qx = pz.h[0] * px + pz.h[1] * py + pz.h[2];
      const float qy = pz.h[3] * px + pz.h[4] * py + pz.h[5];
      const float qz = pz.h[6] * px + pz.h[7] * py + pz.h[8];
      if (fabsf(qz) < 1e-8f) continue;
      const float x = qx / qz - 1.0f, y = qy / qz - 1.0f;
      // floor(x) in [0, w - 2] and floor(y) in [0, h - 2]; NaN fails
      if (!(x >= 0.0f && y >= 0.0f && x < xmax && y < ymax)) continue;
      const int x0 = (int)x, y0 = (int)y;
      const float fx = x - (float)x0, fy = y - (float)y0;
      const int64_t i00 = (int64_t)y0 * u.sw + x0;
      const float4 v00 = il[i00], v10 = il[i00 + 1], v01 = il[i00 + u.sw], v11 = il[i00 + u.sw + 1];
      const float w00 = (1.0f - fx) * (1.0f - fy), w10 = fx * (1.0f - fy), w01 = (1.0f - fx) * fy, w11 = fx * fy;
      a0 += w00 + w10 + w01 + w11;
      a.x += w00 * v00.x + w10 * v10.x + w01 * v01.x + w11 * v11.x;
      a.y += w00 * v00.y + w10 * v10.y + w01 * v01.y + w11 * v11.y;
      a.z += w00 * v00.z + w10 * v10.z + w01 * v01.z + w11 * v11.z;
      a.w += w00 * v00.w + w10 * v10.w + w01 * v01.w + w11 * v11.w;
    }
  }
  const float inv = 1.0f / (float)(k * k);
  s[0] = a0 * inv;
  s[1] = a.x * inv;
  s[2] = a.y * inv;
  s[3] = a.z * inv;
  s[4] = a.w * inv;
}

// Candidate windows → real planes → forward row FFTs. blockIdx.y = batch entry b (pose
// poses[b]); plane point (x, y) is canvas cell (gx0 + x, gy0 + y), zero outside the window
// (x < win_x, y < win_y) and the canvas. Rows >= win_y are not written (the column FFT reads
// them as zero). Spectra of entry b at spec + b · n_cplx · N.
template <int L, int PID>
__global__ __launch_bounds__(RowT<L>::T) void k_pack_sweep(const float4* __restrict__ il,
                                                           const SweepPose* __restrict__ poses, SweepPack u,
                                                           float2* spec, const float2* __restrict__ tw) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int T = RowT<L>::T, M = RowT<L>::M, LINES = RowT<L>::LINES, STR = RowT<L>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const SweepPose pz = poses[b];
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * u.ny;
  float s[M][5];
  {
    float2 dummy[M];
    int idx = 0;
    rload<L, M, T, LINES, false>(dummy, tid, [&](int line, int x) {
      const int y = y0 + line;
      const int cx = pz.gx0 + x, cy = pz.gy0 + y;
      if (x < u.win_x && y < u.win_y && cx < u.grid && cy < u.grid) sweep_cell(il, u, pz, cx, cy, s[idx]);
      else {
#pragma unroll
        for (int p = 0; p < 5; ++p) s[idx][p] = 0.0f;
      }
      ++idx;
      return make_float2(0.0f, 0.0f);
    });
  }
  float2* out = spec + (int64_t)b * P.n_cplx * N;
#pragma unroll
  for (int c = 0; c < P.n_cplx; ++c) {
    float2 v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const float re = real_plane(P.slot_plane[2 * c], s[i]);
      const float im = (2 * c + 1 < P.n_real) ? real_plane(P.slot_plane[2 * c + 1], s[i]) : 0.0f;
      v[i] = make_float2(re, im);
    }
    rfft<L, M, T, LINES, false, -1>(v, lds, STR, tw, tid);
    float2* dst = out + c * N;
    rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
      const int y = y0 + line;
      if (y < u.win_y) dst[(int64_t)y * L + x] = val;
    });
    __syncthreads();
  }
}

// Per-lag scores of batch entry blockIdx.y (sweep.wgsl:236-324 combine_sweep): the overlap floor
// is 8% of n at the lag that leaves the moving image where the candidate put it (zero_lag[b]).
template <int PID>
__global__ __launch_bounds__(256) void k_combine_sweep(const float* __restrict__ corr, float* ns, int64_t nn,
                                                       CombineArgs a, const int* __restrict__ zero_lag) {
  constexpr CorrPlan P = kPlans[PID];
  const int b = blockIdx.y;
  const float* cb = corr + (int64_t)b * P.n_corr * nn;
  auto C = [&](int num, int64_t t) -> double { return (double)cb[(int64_t)P.kept_of[num] * nn + t]; };
  const double floor_n = fmax((double)a.min_n, 0.08 * fmax(C(0, zero_lag[b]), 0.0));
  float* out = ns + (int64_t)b * nn;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < nn; t += (int64_t)gridDim.x * 256)
    out[t] = (float)lag_score<PID>(C, t, floor_n, a, false);
}

// Top local maxima of each map (sweep.wgsl:326-371 peaks_sweep). Thread class c (0..255) owns
// the lags t ≡ c (mod 256) and keeps its largest 3×3 local maximum (periodic neighbours; ties:
// the first lag); the n_peak best classes are the peaks. Split over chunks of lags: part holds
// (value, lag) per (entry, chunk, class).
// nx, ny: powers of two (lx = log2 nx), so the periodic neighbours are masks.
__global__ __launch_bounds__(256) void k_peaks_part(const float* __restrict__ ns, int lx, int ny, int64_t chunk,
                                                    float2* part) {
  const int b = blockIdx.y, lid = threadIdx.x;
  const int nx = 1 << lx, mx = nx - 1, my = ny - 1;
  const int64_t nn = (int64_t)nx * ny;
  const float* m = ns + (int64_t)b * nn;
  const int64_t t0 = (int64_t)blockIdx.x * chunk, t1 = min(nn, t0 + chunk);
  float bv = 0.0f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = t0 + lid; t < t1; t += 256) {
    const float v = m[t];
    if (v <= bv) continue;
    const int y = (int)(t >> lx), x = (int)t & mx;
    bool ok = true;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t row = (int64_t)((y + dy) & my) << lx;
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        if (dx == 0 && dy == 0) continue;
        ok &= !(m[row + ((x + dx) & mx)] > v);
      }
    }
    if (ok) {
      bv = v;
      bi = (uint32_t)t;
    }
  }
  part[((int64_t)b * gridDim.x + blockIdx.x) * 256 + lid] = make_float2(bv, __uint_as_float(bi));
}

__global__ __launch_bounds__(256) void k_peaks_final(const float2* __restrict__ part, int chunks, int n_peak,
                                                     float2* peaks) {
  const int b = blockIdx.x, lid = threadIdx.x;
  float bv = 0.0f;
  uint32_t bi = 0xffffffffu;
  for (int c = 0; c < chunks; ++c) {   // chunks in lag order: a later equal value does not replace
    const float2 e = part[((int64_t)b * chunks + c) * 256 + lid];
    if (e.x > bv) {
      bv = e.x;
      bi = __float_as_uint(e.y);
    }
  }
  __shared__ float sv[256];
  __shared__ uint32_t si[256];
  __shared__ float rv[256];
  __shared__ int ri[256];
  sv[lid] = bv;
  si[lid] = bi;
  __syncthreads();
  // n_peak rounds of a block argmax over the classes (descending value, then lower lag)
  for (int k = 0; k < n_peak; ++k) {
    rv[lid] = sv[lid];
    ri[lid] = lid;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (lid < h) {
        const int o = ri[lid + h], c = ri[lid];
        if (rv[lid + h] > rv[lid] || (rv[lid + h] == rv[lid] && si[o] < si[c])) {
          rv[lid] = rv[lid + h];
          ri[lid] = o;
        }
      }
      __syncthreads();
    }
    if (lid == 0) {
      const int w = ri[0];
      peaks[(int64_t)b * n_peak + k] = make_float2(sv[w], __uint_as_float(si[w]));
      sv[w] = -1.0f;   // taken (scores are >= 0)
      si[w] = 0xffffffffu;
    }
    __syncthreads();
  }
}

// ═════════════════════════════════════ launchers ═════════════════════════════════════════
void launch_round_half(float* x, int64_t n, hipStream_t st) {
  k_round_half<<<min(blocks_for(n, 256), 8192u), 256, 0, st>>>(x, n);
}
void launch_maxabs(const float* x, int64_t n, uint32_t* out, hipStream_t st) {
  k_maxabs<<<min(blocks_for(n, 256), 512u), 256, 0, st>>>(x, n, out);
}
void launch_fixsum(const float* x, int64_t n, double scale, float mean, int squared, long long* hi,
                   unsigned long long* lo, hipStream_t st) {
  k_fixsum<<<min(blocks_for(n, 256), 512u), 256, 0, st>>>(x, n, scale, mean, squared, hi, lo);
}
void launch_hist(const float* x, int64_t n, float mean, float spread, uint32_t* hist, hipStream_t st) {
  k_hist<<<min(blocks_for(n, 256), 256u), 256, 0, st>>>(x, n, mean, spread, hist);
}
void launch_rank(const float* x, int64_t n, float mean, float spread, const uint32_t* hist, const uint32_t* cdf,
                 float* rank, hipStream_t st) {
  k_rank<<<min(blocks_for(n, 256), 8192u), 256, 0, st>>>(x, n, mean, spread, hist, cdf, rank);
}
int feats_blocks(int w, int h) { return ((w + 15) / 16) * ((h + 15) / 16); }
int launch_feats(const float* rank, int w, int h, int mode, float* feat, float* mag, double* part, hipStream_t st) {
  dim3 g((w + 15) / 16, (h + 15) / 16);
  k_feats<<<g, 256, 0, st>>>(rank, w, h, mode, feat, mag, part);
  return feats_blocks(w, h);
}
void launch_sum_parts(const double* part, int nparts, int m, double* out, hipStream_t st) {
  k_sum_parts<<<(unsigned)m, 256, 0, st>>>(part, nparts, m, out);
}
void launch_whiten(float* feat, float* stack, int64_t n, const WhitenArgs& a, hipStream_t st) {
  k_whiten<<<min(blocks_for(n, 256), 8192u), 256, 0, st>>>(feat, stack, n, a);
}
int acf_parts(int w, int h) { return ((w + ACF_T - 1) / ACF_T) * ((h + ACF_T - 1) / ACF_T) * ACF_Q; }
void launch_acf(const float* feat, int w, int h, double* part, hipStream_t st) {
  dim3 g((w + ACF_T - 1) / ACF_T, (h + ACF_T - 1) / ACF_T);
  k_acf<<<g, 256, 0, st>>>(feat, w, h, part);
}

void launch_warp_stack(const float* stack, int sw, int sh, const double* Hi, float* canvas, int cw, int ch,
                       hipStream_t st) {
  H9 h;
  for (int i = 0; i < 9; ++i) h.m[i] = Hi[i];
  dim3 g((cw + 15) / 16, (ch + 15) / 16);
  k_warp_stack<<<g, 256, 0, st>>>(stack, sw, sh, h, canvas, cw, ch);
}
void launch_paste_stack(const float* stack, int sw, int sh, int ox, int oy, float* canvas, int cw, int ch,
                        hipStream_t st) {
  dim3 g((cw + 15) / 16, (ch + 15) / 16);
  k_paste_stack<<<g, 256, 0, st>>>(stack, sw, sh, ox, oy, canvas, cw, ch);
}

template <int L>
static void pack_rowfft_L(int plan, int side, const float* canvas, const Resample& r, int ny, float2* spec,
                          const float2* tw, hipStream_t st) {
  dim3 g((r.ch + RowT<L>::LINES - 1) / RowT<L>::LINES);
  // plans without rectangle sums share one table for both sides
  const int ti = plan < 4 ? 2 * plan : 2 * plan + side;
#define K(TAB) k_pack_rowfft<L, TAB><<<g, RowT<L>::T, 0, st>>>(canvas, r, ny, spec, tw)
  switch (ti) {
    case 0: K(0); break;
    case 2: K(2); break;
    case 4: K(4); break;
    case 6: K(6); break;
    case 8: K(8); break;
    case 9: K(9); break;
    case 10: K(10); break;
    case 11: K(11); break;
    case 12: K(12); break;
    case 13: K(13); break;
    case 14: K(14); break;
    default: K(15); break;
  }
#undef K
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

bool launch_pack_rowfft(int plan, int side, const float* canvas, int W, int H, int cw, int ch, int nx, int ny,
                        float2* spec, const float2* tw_x, hipStream_t st) {
  Resample r;
  r.W = W;
  r.H = H;
  r.cw = cw;
  r.ch = ch;
  r.fx = (float)W / (float)(cw > 0 ? cw : 1);
  r.fy = (float)H / (float)(ch > 0 ? ch : 1);
  r.kx = (int)fminf(fmaxf(ceilf(r.fx - 1e-4f), 1.0f), 8.0f);
  r.ky = (int)fminf(fmaxf(ceilf(r.fy - 1e-4f), 1.0f), 8.0f);
  r.identity = (cw == W && ch == H) ? 1 : 0;
#define CALL(LL) pack_rowfft_L<LL>(plan, side, canvas, r, ny, spec, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

template <int L>
static void pack_logpolar_L(int plan, const LogPolarArgs& a, int ny, float2* spec, const float2* tw, hipStream_t st) {
  dim3 g((a.n_lam + RowT<L>::LINES - 1) / RowT<L>::LINES);
  switch (plan) {
    case 0: k_pack_logpolar<L, 0><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
    case 1: k_pack_logpolar<L, 1><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
    case 2: k_pack_logpolar<L, 2><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
    default: k_pack_logpolar<L, 3><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
  }
}

bool launch_pack_logpolar(int plan, const LogPolarArgs& a, int nx, int ny, float2* spec, const float2* tw_x,
                          hipStream_t st) {
  if (plan < 0 || plan > 3 || a.n_th > nx || a.n_lam > ny) return false;
#define CALL(LL) pack_logpolar_L<LL>(plan, a, ny, spec, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

void launch_rs_tables(int n_th, int n_lam, double log0, double span, double* rho, double* cth, double* sth,
                      hipStream_t st) {
  const int m = n_th > n_lam ? n_th : n_lam;
  k_rs_tables<<<(m + 255) / 256, 256, 0, st>>>(n_th, n_lam, log0, span, rho, cth, sth);
}

void launch_combine_rs(int plan, const float* corr, float* map, const RsArgs& r, const CombineArgs& a,
                       hipStream_t st) {
  const int g = combine_blocks((int64_t)r.n * r.n);
  switch (plan) {
    case 0: k_combine_rs<0><<<g, 256, 0, st>>>(corr, map, r, a); break;
    case 1: k_combine_rs<1><<<g, 256, 0, st>>>(corr, map, r, a); break;
    case 2: k_combine_rs<2><<<g, 256, 0, st>>>(corr, map, r, a); break;
    default: k_combine_rs<3><<<g, 256, 0, st>>>(corr, map, r, a); break;
  }
}

void launch_rs_finish(const float* fwd, const float* inv, int n, bool mirror, double dlam, double lo, double hi,
                      float* out, float* arg_part, hipStream_t st) {
  k_rs_finish<<<combine_blocks((int64_t)n * n), 256, 0, st>>>(fwd, inv, n, mirror ? 1 : 0, dlam, lo, hi, out, arg_part);
}

bool launch_colfft(float2* data, int planes, int nx, int ny, int rows_in, bool inverse, const float2* tw_y,
                   hipStream_t st, int lines) {
  // a tiling that cannot launch at this width (nx not a multiple of its columns) falls back to
  // the default tiling
#define CALL(LL)                                                                                   \
  {                                                                                                \
    auto run = [&](int ln) {                                                                       \
      return inverse ? colfft_L<LL, 1>(data, planes, nx, rows_in, 1.0f, tw_y, ln, st)              \
                     : colfft_L<LL, -1>(data, planes, nx, rows_in, 1.0f, tw_y, ln, st);            \
    };                                                                                             \
    return run(lines) || (lines != 0 && run(0));                                                   \
  }
  ZC_POW2_SWITCH(ny, CALL)
#undef CALL
  return true;
}

void launch_cmul(int plan, const float2* spec_a, const float2* spec_b, float2* work, int nx, int ny, hipStream_t st,
                 int batch) {
  const int64_t cnt = (int64_t)(ny / 2 + 1) * nx;
  const float scale = 1.0f / ((float)nx * (float)ny);
  const dim3 g(blocks_for(cnt, 256), batch);
  switch (plan) {
    case 0: k_cmul<0><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    case 1: k_cmul<1><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    case 2: k_cmul<2><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    case 3: k_cmul<3><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    case 4: k_cmul<4><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    case 5: k_cmul<5><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    case 6: k_cmul<6><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
    default: k_cmul<7><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale); break;
  }
}

template <int L, int V>
static void rowifft_LV(int plan, const float2* work, float* corr, int ny, const float2* tw, hipStream_t st, int batch) {
  using X = RowX<L, V>;
  dim3 g((ny + X::LINES - 1) / X::LINES, kPlans[plan].n_out * batch);
  switch (plan) {
    case 0: k_rowifft_extract<L, 0, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    case 1: k_rowifft_extract<L, 1, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    case 2: k_rowifft_extract<L, 2, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    case 3: k_rowifft_extract<L, 3, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    case 4: k_rowifft_extract<L, 4, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    case 5: k_rowifft_extract<L, 5, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    case 6: k_rowifft_extract<L, 6, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
    default: k_rowifft_extract<L, 7, V><<<g, X::T, 0, st>>>(work, corr, ny, tw); break;
  }
}

template <int L>
static void rowifft_L(int plan, const float2* work, float* corr, int ny, const float2* tw, hipStream_t st, int batch) {
  if constexpr (L >= 256) {
    switch (FftRowTiling::v) {
      case 1: rowifft_LV<L, 1>(plan, work, corr, ny, tw, st, batch); return;
      case 2: rowifft_LV<L, 2>(plan, work, corr, ny, tw, st, batch); return;
      case 3: rowifft_LV<L, 3>(plan, work, corr, ny, tw, st, batch); return;
      case 4: rowifft_LV<L, 4>(plan, work, corr, ny, tw, st, batch); return;
      default: break;
    }
  }
  rowifft_LV<L, 0>(plan, work, corr, ny, tw, st, batch);
}

bool launch_rowifft_extract(int plan, const float2* work, float* corr, int nx, int ny, const float2* tw_x,
                            hipStream_t st, int batch) {
#define CALL(LL) rowifft_L<LL>(plan, work, corr, ny, tw_x, st, batch)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

int combine_blocks(int64_t nn) { return (int)std::min<int64_t>((nn + 255) / 256, 4096); }

void launch_combine(int plan, const float* corr, float* map, int64_t nn, const CombineArgs& a, float* arg_part,
                    hipStream_t st, const SatView* sv) {
  const int g = combine_blocks(nn);
  if (plan >= 4 && !sv) throw std::runtime_error("rectangle plan without tables");
  switch (plan) {
    case 0: k_combine<0><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 1: k_combine<1><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 2: k_combine<2><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 3: k_combine<3><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 4: k_combine_rect<4><<<g, 256, 0, st>>>(corr, map, nn, a, *sv, arg_part); break;
    case 5: k_combine_rect<5><<<g, 256, 0, st>>>(corr, map, nn, a, *sv, arg_part); break;
    case 6: k_combine_rect<6><<<g, 256, 0, st>>>(corr, map, nn, a, *sv, arg_part); break;
    default: k_combine_rect<7><<<g, 256, 0, st>>>(corr, map, nn, a, *sv, arg_part); break;
  }
}

void launch_argmax_finish(const float* arg_part, int nparts, const float* map, float* res, hipStream_t st) {
  k_argmax_finish<<<1, 256, 0, st>>>(arg_part, nparts, map, res);
}

bool launch_fft2(float2* data, int planes, int n, bool inverse, const float2* tw, hipStream_t st, int lines) {
  const float sc = inverse ? 1.0f / (float)n : 1.0f;
#define CALL(LL)                                                                                \
  {                                                                                             \
    const int rows = planes * LL;                                                               \
    const unsigned gr = (unsigned)((rows + RowT<LL>::LINES - 1) / RowT<LL>::LINES);              \
    if (inverse) k_rowfft<LL, 1><<<gr, RowT<LL>::T, 0, st>>>(data, rows, sc, tw);               \
    else k_rowfft<LL, -1><<<gr, RowT<LL>::T, 0, st>>>(data, rows, sc, tw);                      \
    return inverse ? colfft_L<LL, 1>(data, planes, n, n, sc, tw, lines, st)                     \
                   : colfft_L<LL, -1>(data, planes, n, n, sc, tw, lines, st);                   \
  }
  ZC_POW2_SWITCH(n, CALL)
#undef CALL
  return true;
}

// Twiddles e^(−2πi t / L), FP64-computed, one table per (device, L) for the life of the process.
static const float2* static_twiddles(int L) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, float2*> tables;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  auto it = tables.find({dev, L});
  if (it != tables.end()) return it->second;
  std::vector<float2> h(L);
  for (int t = 0; t < L; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)L;
    h[t] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * L) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * L, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  tables[{dev, L}] = d;
  return d;
}

bool launch_fft2_twiddled(float2* data, int planes, int n, bool inverse, hipStream_t st) {
  const float2* tw = static_twiddles(n);
  return tw && launch_fft2(data, planes, n, inverse, tw, st, 0);
}

int moments_blocks(int w, int h) {
  const int64_t n = (int64_t)w * h;
  return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 256 * 64 - 1) / (256 * 64)));
}

void launch_pose_moments(const float* feat_a, int wa, int ha, const float* feat_b, int wb, int hb, const PoseArgs& p,
                         double* part, hipStream_t st) {
  k_pose_moments<<<moments_blocks(wb, hb), 256, 0, st>>>(feat_a, wa, ha, feat_b, wb, hb, p, part);
}

void launch_interleave(const float* feat, int64_t n, float4* il, hipStream_t st) {
  k_interleave<<<min(blocks_for(n, 256), 8192u), 256, 0, st>>>(feat, n, il);
}

template <int L>
static void pack_sweep_L(int plan, const float4* il, const SweepPose* poses, int batch, const SweepPack& u,
                         float2* spec, const float2* tw, hipStream_t st) {
  const dim3 g((u.win_y + RowT<L>::LINES - 1) / RowT<L>::LINES, batch);
  switch (plan) {
    case 0: k_pack_sweep<L, 0><<<g, RowT<L>::T, 0, st>>>(il, poses, u, spec, tw); break;
    case 1: k_pack_sweep<L, 1><<<g, RowT<L>::T, 0, st>>>(il, poses, u, spec, tw); break;
    case 2: k_pack_sweep<L, 2><<<g, RowT<L>::T, 0, st>>>(il, poses, u, spec, tw); break;
    default: k_pack_sweep<L, 3><<<g, RowT<L>::T, 0, st>>>(il, poses, u, spec, tw); break;
  }
}

bool launch_pack_sweep(int plan, const float4* il, const SweepPose* poses, int batch, const SweepPack& u, int nx,
                       float2* spec, const float2* tw_x, hipStream_t st) {
  if (u.win_y <= 0) return true;
#define CALL(LL) pack_sweep_L<LL>(plan, il, poses, batch, u, spec, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

void launch_combine_sweep(int plan, const float* corr, float* ns, int64_t nn, int batch, const CombineArgs& a,
                          const int* zero_lag, hipStream_t st) {
  const dim3 g((unsigned)std::min<int64_t>((nn + 255) / 256, 2048), batch);
  switch (plan) {
    case 0: k_combine_sweep<0><<<g, 256, 0, st>>>(corr, ns, nn, a, zero_lag); break;
    case 1: k_combine_sweep<1><<<g, 256, 0, st>>>(corr, ns, nn, a, zero_lag); break;
    case 2: k_combine_sweep<2><<<g, 256, 0, st>>>(corr, ns, nn, a, zero_lag); break;
    default: k_combine_sweep<3><<<g, 256, 0, st>>>(corr, ns, nn, a, zero_lag); break;
  }
}

// lag chunks per map: ~16K lags per workgroup (enough workgroups to fill the chip at any batch)
int peaks_chunks(int64_t nn) { return (int)std::max<int64_t>(1, std::min<int64_t>(1024, nn / 16384)); }

void launch_peaks(const float* ns, int nx, int ny, int batch, int n_peak, float2* part, float2* peaks,
                  hipStream_t st) {
  const int64_t nn = (int64_t)nx * ny;
  const int chunks = peaks_chunks(nn);
  const int64_t chunk = ((nn + chunks - 1) / chunks + 255) / 256 * 256;
  int lx = 0;
  while ((1 << lx) < nx) ++lx;
  k_peaks_part<<<dim3(chunks, batch), 256, 0, st>>>(ns, lx, ny, chunk, part);
  k_peaks_final<<<batch, 256, 0, st>>>(part, chunks, n_peak, peaks);
}

int jobs_blocks(int64_t max_pixels) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(256, (max_pixels + 256 * 32 - 1) / (256 * 32)));
}

void launch_pose_moments_jobs(const PoseJob* jobs, int njobs, int blocks, double* part, double* sums,
                              hipStream_t st) {
  k_pose_moments_jobs<<<dim3(blocks, njobs), 256, 0, st>>>(jobs, part);
  k_sum_parts<<<dim3(45, njobs), 256, 0, st>>>(part, blocks, 45, sums);
}

void launch_tangent_jobs(const TangentJob* jobs, int njobs, int blocks, double* part, double* sums,
                         hipStream_t st) {
  k_tangent_jobs<<<dim3(blocks, njobs), 256, 0, st>>>(jobs, part);
  k_sum_parts<<<dim3(44, njobs), 256, 0, st>>>(part, blocks, 44, sums);
}

}  // namespace zc
