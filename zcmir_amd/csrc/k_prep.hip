// Per-image preparation kernels (rank transform, features, whitening, autocorrelation) and the
// canvas warp / paste of the map.
#include "kdev.hpp"

namespace zc {

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

// one canvas pixel of the warped stack (Hi: canvas → source, 1-based pixels)
__device__ __forceinline__ void warp_stack_px(const float* __restrict__ stack, int sw, int sh, const double* Hi,
                                              float* canvas, int cw, int ch) {
  const int gx = blockIdx.x * 16 + (threadIdx.x & 15), gy = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (gx >= cw || gy >= ch) return;
  const int64_t n = (int64_t)cw * ch, pix = (int64_t)gy * cw + gx, sn = (int64_t)sw * sh;
  const double X = gx + 1.0, Y = gy + 1.0;
  const double qx = Hi[0] * X + Hi[1] * Y + Hi[2];
  const double qy = Hi[3] * X + Hi[4] * Y + Hi[5];
  const double qz = Hi[6] * X + Hi[7] * Y + Hi[8];
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

__global__ __launch_bounds__(256) void k_warp_stack(const float* __restrict__ stack, int sw, int sh, H9 Hi,
                                                    float* canvas, int cw, int ch) {
  warp_stack_px(stack, sw, sh, Hi.m, canvas, cw, ch);
}

// the pose read from device memory (a pose that device code chose: the climb's)
__global__ __launch_bounds__(256) void k_warp_stack_dev(const float* __restrict__ stack, int sw, int sh,
                                                        const double* __restrict__ Hi, float* canvas, int cw, int ch) {
  warp_stack_px(stack, sw, sh, Hi, canvas, cw, ch);
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
void launch_sum_parts(const double* part, int nparts, int m, double* out, hipStream_t st, int batch) {
  k_sum_parts<<<dim3((unsigned)m, (unsigned)batch), 256, 0, st>>>(part, nparts, m, out);
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

void launch_warp_stack_dev(const float* stack, int sw, int sh, const double* Hi_dev, float* canvas, int cw, int ch,
                           hipStream_t st) {
  dim3 g((cw + 15) / 16, (ch + 15) / 16);
  k_warp_stack_dev<<<g, 256, 0, st>>>(stack, sw, sh, Hi_dev, canvas, cw, ch);
}
void launch_paste_stack(const float* stack, int sw, int sh, int ox, int oy, float* canvas, int cw, int ch,
                        hipStream_t st) {
  dim3 g((cw + 15) / 16, (ch + 15) / 16);
  k_paste_stack<<<g, 256, 0, st>>>(stack, sw, sh, ox, oy, canvas, cw, ch);
}


}  // namespace zc
