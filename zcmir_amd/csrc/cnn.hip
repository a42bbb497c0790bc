// The learned rotation × scale prior's kernels (no counterpart in the reference): the network's
// input features (log-polar Fourier magnitudes of downsampled rank-gradient images), 3×3
// convolutions as implicit GEMMs on BF16 MFMA with FP32 accumulation (forward, input gradient,
// weight gradient), the dense head, softmax cross-entropy and Adam.
//
// Layouts: activations NHWC FP32 in HBM (converted to BF16 when they enter an MFMA); weights
// FP32 master copies, conv weights [ky][kx][cin][cout] (= a [9·cin][cout] matrix), dense
// weights [k][j]. The H axis (rotation) is periodic, the W axis (log radius) zero-padded.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.hpp"

namespace zc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 → BF16, round to nearest even (finite inputs)
__device__ __forceinline__ __bf16 to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __builtin_bit_cast(__bf16, (unsigned short)(u >> 16));
}

// One wave's 16 × 16 tile of C = A B over K: v_mfma_f32_16x16x32_bf16, lane l holds A row
// (l mod 16), B column (l mod 16), K elements 8·(l div 16) .. +7 of each 32-wide step; the result
// acc[i] is C row 4·(l div 16) + i, column l mod 16.
template <typename FA, typename FB>
__device__ __forceinline__ f32x4 wave_tile(int m0, int n0, int K, FA&& fa, FB&& fb) {
  const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + g * 8 + j;
      a[j] = to_bf16(k < K ? fa(m0 + r, k) : 0.0f);
      b[j] = to_bf16(k < K ? fb(k, n0 + r) : 0.0f);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  return acc;
}

struct ConvShape {
  int n, h, w, cin, cout, stride, ho, wo;
};

__device__ __forceinline__ int wrap(int y, int h) { return ((y % h) + h) % h; }

// y = relu(conv(x) + b): M = n·ho·wo output pixels, N = cout, K = 9·cin.
__global__ __launch_bounds__(256) void k_conv_fwd(const float* __restrict__ x, const float* __restrict__ wt,
                                                  const float* __restrict__ bias, float* y, ConvShape s, int relu) {
  const int M = s.n * s.ho * s.wo, tn = (s.cout + 15) / 16, tiles = ((M + 15) / 16) * tn;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int m0 = (t / tn) * 16, n0 = (t % tn) * 16;
  const int K = 9 * s.cin;
  auto fa = [&](int m, int k) -> float {
    if (m >= M) return 0.0f;
    const int xo = m % s.wo, yo = (m / s.wo) % s.ho, b = m / (s.wo * s.ho);
    const int tap = k / s.cin, ci = k - tap * s.cin, ky = tap / 3, kx = tap - ky * 3;
    const int yi = wrap(yo * s.stride + ky - 1, s.h), xi = xo * s.stride + kx - 1;
    if (xi < 0 || xi >= s.w) return 0.0f;
    return x[(((int64_t)b * s.h + yi) * s.w + xi) * s.cin + ci];
  };
  auto fb = [&](int k, int c) -> float { return c < s.cout ? wt[(int64_t)k * s.cout + c] : 0.0f; };
  const f32x4 acc = wave_tile(m0, n0, K, fa, fb);
  const int l = threadIdx.x & 63, c = n0 + (l & 15);
  if (c >= s.cout) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + (l >> 4) * 4 + i;
    if (m >= M) continue;
    float v = acc[i] + bias[c];
    if (relu) v = fmaxf(v, 0.0f);
    y[(int64_t)m * s.cout + c] = v;
  }
}

// The output gradient through the activation: dy · [y > 0] (relu) or dy.
__device__ __forceinline__ float dpre(const float* dy, const float* y, int64_t i, int relu) {
  return relu ? (y[i] > 0.0f ? dy[i] : 0.0f) : dy[i];
}

// dx = convᵀ(dy'): M = n·h·w input pixels, N = cin, K = 9·cout; input pixel (yi, xi) receives
// tap (ky, kx) of the output pixel yo·s + ky − 1 ≡ yi (mod h), xo·s + kx − 1 = xi.
__global__ __launch_bounds__(256) void k_conv_bwd_data(const float* __restrict__ dy, const float* __restrict__ y,
                                                       const float* __restrict__ wt, float* dx, ConvShape s, int relu) {
  const int M = s.n * s.h * s.w, tn = (s.cin + 15) / 16, tiles = ((M + 15) / 16) * tn;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int m0 = (t / tn) * 16, n0 = (t % tn) * 16;
  const int K = 9 * s.cout;
  auto fa = [&](int m, int k) -> float {
    if (m >= M) return 0.0f;
    const int xi = m % s.w, yi = (m / s.w) % s.h, b = m / (s.w * s.h);
    const int tap = k / s.cout, co = k - tap * s.cout, ky = tap / 3, kx = tap - ky * 3;
    const int ty = wrap(yi - ky + 1, s.h), tx = xi - kx + 1;
    if (ty % s.stride || tx < 0 || tx % s.stride) return 0.0f;
    const int yo = ty / s.stride, xo = tx / s.stride;
    if (yo >= s.ho || xo >= s.wo) return 0.0f;
    return dpre(dy, y, (((int64_t)b * s.ho + yo) * s.wo + xo) * s.cout + co, relu);
  };
  auto fb = [&](int k, int c) -> float {
    if (c >= s.cin) return 0.0f;
    const int tap = k / s.cout, co = k - tap * s.cout;
    return wt[((int64_t)tap * s.cin + c) * s.cout + co];
  };
  const f32x4 acc = wave_tile(m0, n0, K, fa, fb);
  const int l = threadIdx.x & 63, c = n0 + (l & 15);
  if (c >= s.cin) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + (l >> 4) * 4 + i;
    if (m < M) dx[(int64_t)m * s.cin + c] = acc[i];
  }
}

// dw += Σ_p xcol(p, k) dy'(p, c): rows k of the [9·cin][cout] matrix, the pixel sum split over
// blockIdx.y chunks of `chunk` output pixels, added atomically.
__global__ __launch_bounds__(256) void k_conv_bwd_weight(const float* __restrict__ x, const float* __restrict__ dy,
                                                         const float* __restrict__ y, float* dw, ConvShape s, int relu,
                                                         int chunk) {
  const int K = 9 * s.cin, P = s.n * s.ho * s.wo, tn = (s.cout + 15) / 16, tiles = ((K + 15) / 16) * tn;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int k0 = (t / tn) * 16, n0 = (t % tn) * 16;
  const int p0 = blockIdx.y * chunk, np = min(chunk, P - p0);
  if (np <= 0) return;
  auto fa = [&](int k, int pp) -> float {
    if (k >= K) return 0.0f;
    const int p = p0 + pp;
    const int xo = p % s.wo, yo = (p / s.wo) % s.ho, b = p / (s.wo * s.ho);
    const int tap = k / s.cin, ci = k - tap * s.cin, ky = tap / 3, kx = tap - ky * 3;
    const int yi = wrap(yo * s.stride + ky - 1, s.h), xi = xo * s.stride + kx - 1;
    if (xi < 0 || xi >= s.w) return 0.0f;
    return x[(((int64_t)b * s.h + yi) * s.w + xi) * s.cin + ci];
  };
  auto fb = [&](int pp, int c) -> float {
    return c < s.cout ? dpre(dy, y, (int64_t)(p0 + pp) * s.cout + c, relu) : 0.0f;
  };
  const f32x4 acc = wave_tile(k0, n0, np, fa, fb);
  const int l = threadIdx.x & 63, c = n0 + (l & 15);
  if (c >= s.cout) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = k0 + (l >> 4) * 4 + i;
    if (k < K) atomicAdd(&dw[(int64_t)k * s.cout + c], acc[i]);
  }
}

// db[c] += Σ_p dy'(p, c) (one block per channel).
__global__ __launch_bounds__(256) void k_bias_grad(const float* __restrict__ dy, const float* __restrict__ y,
                                                   float* db, int P, int C, int relu) {
  const int c = blockIdx.x;
  double s = 0.0;
  for (int p = threadIdx.x; p < P; p += 256) s += dpre(dy, y, (int64_t)p * C + c, relu);
  __shared__ double sm[256];
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] += (float)sm[0];
}

// Dense layer: out[n][j] = Σ_k f[n][k] w[k][j] + b[j].
__global__ __launch_bounds__(256) void k_fc_fwd(const float* __restrict__ f, const float* __restrict__ w,
                                                const float* __restrict__ b, float* out, int N, int K, int J) {
  const int tn = (J + 15) / 16, tiles = ((N + 15) / 16) * tn;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int m0 = (t / tn) * 16, n0 = (t % tn) * 16;
  auto fa = [&](int m, int k) -> float { return m < N ? f[(int64_t)m * K + k] : 0.0f; };
  auto fb = [&](int k, int j) -> float { return j < J ? w[(int64_t)k * J + j] : 0.0f; };
  const f32x4 acc = wave_tile(m0, n0, K, fa, fb);
  const int l = threadIdx.x & 63, j = n0 + (l & 15);
  if (j >= J) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + (l >> 4) * 4 + i;
    if (m < N) out[(int64_t)m * J + j] = acc[i] + b[j];
  }
}

// df[n][k] = Σ_j dout[n][j] w[k][j]
__global__ __launch_bounds__(256) void k_fc_bwd_data(const float* __restrict__ dout, const float* __restrict__ w,
                                                     float* df, int N, int K, int J) {
  const int tn = (K + 15) / 16, tiles = ((N + 15) / 16) * tn;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int m0 = (t / tn) * 16, n0 = (t % tn) * 16;
  auto fa = [&](int m, int j) -> float { return m < N ? dout[(int64_t)m * J + j] : 0.0f; };
  auto fb = [&](int j, int k) -> float { return k < K ? w[(int64_t)k * J + j] : 0.0f; };
  const f32x4 acc = wave_tile(m0, n0, J, fa, fb);
  const int l = threadIdx.x & 63, k = n0 + (l & 15);
  if (k >= K) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + (l >> 4) * 4 + i;
    if (m < N) df[(int64_t)m * K + k] = acc[i];
  }
}

// dw[k][j] += Σ_n f[n][k] dout[n][j]
__global__ __launch_bounds__(256) void k_fc_bwd_weight(const float* __restrict__ f, const float* __restrict__ dout,
                                                       float* dw, int N, int K, int J) {
  const int tn = (J + 15) / 16, tiles = ((K + 15) / 16) * tn;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= tiles) return;
  const int k0 = (t / tn) * 16, n0 = (t % tn) * 16;
  auto fa = [&](int k, int n) -> float { return k < K ? f[(int64_t)n * K + k] : 0.0f; };
  auto fb = [&](int n, int j) -> float { return j < J ? dout[(int64_t)n * J + j] : 0.0f; };
  const f32x4 acc = wave_tile(k0, n0, N, fa, fb);
  const int l = threadIdx.x & 63, j = n0 + (l & 15);
  if (j >= J) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = k0 + (l >> 4) * 4 + i;
    if (k < K) dw[(int64_t)k * J + j] += acc[i];
  }
}

// db[j] += Σ_n dout[n][j]
__global__ __launch_bounds__(256) void k_fc_bias_grad(const float* __restrict__ dout, float* db, int N, int J) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= J) return;
  double s = 0.0;
  for (int n = 0; n < N; ++n) s += dout[(int64_t)n * J + j];
  db[j] += (float)s;
}

// Softmax cross-entropy per sample (one wave each): loss[n] = −log p[label], dlogits = (p − 1_label) / N.
__global__ __launch_bounds__(64) void k_xent(const float* __restrict__ logits, const int* __restrict__ labels,
                                             float* loss, float* dlogits, int N, int J) {
  const int n = blockIdx.x, l = threadIdx.x;
  const float* z = logits + (int64_t)n * J;
  float mx = -INFINITY;
  for (int j = l; j < J; j += 64) mx = fmaxf(mx, z[j]);
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  double se = 0.0;
  for (int j = l; j < J; j += 64) se += exp((double)(z[j] - mx));
  for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
  const int lab = labels[n];
  for (int j = l; j < J; j += 64) {
    const double p = exp((double)(z[j] - mx)) / se;
    dlogits[(int64_t)n * J + j] = (float)((p - (j == lab ? 1.0 : 0.0)) / N);
  }
  if (l == 0) loss[n] = (float)(log(se) - (double)(z[lab] - mx));
}

// Adam (bias-corrected) on n parameters.
__global__ __launch_bounds__(256) void k_adam(float* p, const float* __restrict__ g, float* m, float* v, int64_t n,
                                              float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// ── input features ─────────────────────────────────────────────────────────────────────────
// Box downsample of image blockIdx.y (w × h) to S × S.
__global__ __launch_bounds__(256) void k_lp_down(const float* __restrict__ src, int w, int h, float* dst, int S) {
  const int img = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= S * S) return;
  const int x = i % S, y = i / S;
  const int x0 = (int)((int64_t)x * w / S), x1 = max(x0 + 1, (int)((int64_t)(x + 1) * w / S));
  const int y0 = (int)((int64_t)y * h / S), y1 = max(y0 + 1, (int)((int64_t)(y + 1) * h / S));
  const float* s = src + (int64_t)img * w * h;
  float acc = 0.0f;
  for (int yy = y0; yy < y1; ++yy)
    for (int xx = x0; xx < x1; ++xx) acc += s[(int64_t)yy * w + xx];
  dst[(int64_t)img * S * S + i] = acc / (float)((x1 - x0) * (y1 - y0));
}

// Per-image rank (histogram equalization over 256 bins of [0, 1]), in place; one block per image.
__global__ __launch_bounds__(256) void k_lp_rank(float* img, int n) {
  float* a = img + (int64_t)blockIdx.x * n;
  __shared__ uint32_t hist[256];
  __shared__ uint32_t cdf[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) {
    const int b = min(255, max(0, (int)(a[i] * 256.0f)));
    atomicAdd(&hist[b], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int b = 0; b < 256; ++b) {
      cdf[b] = acc;
      acc += hist[b];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) {
    const int b = min(255, max(0, (int)(a[i] * 256.0f)));
    a[i] = ((float)cdf[b] + 0.5f * (float)hist[b]) / (float)n * 2.0f - 1.0f;
  }
}

// Sobel gradient magnitude (clamped borders) times a Hann window → complex plane (re, 0).
__global__ __launch_bounds__(256) void k_lp_grad(const float* __restrict__ r, float2* plane, int S) {
  const int img = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= S * S) return;
  const int x = i % S, y = i / S;
  const float* a = r + (int64_t)img * S * S;
  auto at = [&](int yy, int xx) { return a[min(max(yy, 0), S - 1) * S + min(max(xx, 0), S - 1)]; };
  const float gx = (at(y - 1, x + 1) + 2.0f * at(y, x + 1) + at(y + 1, x + 1)) -
                   (at(y - 1, x - 1) + 2.0f * at(y, x - 1) + at(y + 1, x - 1));
  const float gy = (at(y + 1, x - 1) + 2.0f * at(y + 1, x) + at(y + 1, x + 1)) -
                   (at(y - 1, x - 1) + 2.0f * at(y - 1, x) + at(y - 1, x + 1));
  const float wx = 0.5f - 0.5f * cospif(2.0f * ((float)x + 0.5f) / (float)S);
  const float wy = 0.5f - 0.5f * cospif(2.0f * ((float)y + 0.5f) / (float)S);
  plane[(int64_t)img * S * S + i] = make_float2(sqrtf(gx * gx + gy * gy) * wx * wy, 0.0f);
}

// log(1 + |F|) of spectrum blockIdx.y at T log-polar angles θ = π t / T (t < T) and R radii
// r = r0 · (r1 / r0)^(j / R) (frequency units), bilinear over the periodic spectrum; written to
// channel `ch` of an NHWC [T][R][C] map.
__global__ __launch_bounds__(256) void k_lp_polar(const float2* __restrict__ spec, int S, float* out, int T, int R,
                                                  int C, int ch, float r0, float r1) {
  const int img = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= T * R) return;
  const int t = i / R, j = i % R;
  const float th = 3.14159265358979f * (float)t / (float)T;
  const float rr = r0 * powf(r1 / r0, (float)j / (float)R);
  const float u = rr * cosf(th), v = rr * sinf(th);
  const float fu = floorf(u), fv = floorf(v);
  const float au = u - fu, av = v - fv;
  const int iu = (int)fu, iv = (int)fv;
  const float2* sp = spec + (int64_t)img * S * S;
  auto mag = [&](int vv, int uu) {
    const float2 z = sp[(int64_t)(((vv % S) + S) % S) * S + (((uu % S) + S) % S)];
    return sqrtf(z.x * z.x + z.y * z.y);
  };
  const float m = (mag(iv, iu) * (1.0f - au) + mag(iv, iu + 1) * au) * (1.0f - av) +
                  (mag(iv + 1, iu) * (1.0f - au) + mag(iv + 1, iu + 1) * au) * av;
  out[(((int64_t)img * T + t) * R + j) * C + ch] = logf(1.0f + m);
}

// Zero mean, unit standard deviation per (map, channel) of NHWC [n][T·R][C]; one block each.
__global__ __launch_bounds__(256) void k_lp_norm(float* x, int P, int C) {
  const int img = blockIdx.x / C, ch = blockIdx.x % C;
  float* a = x + (int64_t)img * P * C + ch;
  double s = 0.0, q = 0.0;
  for (int p = threadIdx.x; p < P; p += 256) {
    const double v = a[(int64_t)p * C];
    s += v;
    q += v * v;
  }
  __shared__ double ss[256], qq[256];
  ss[threadIdx.x] = s;
  qq[threadIdx.x] = q;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      ss[threadIdx.x] += ss[threadIdx.x + k];
      qq[threadIdx.x] += qq[threadIdx.x + k];
    }
    __syncthreads();
  }
  const double mean = ss[0] / P, var = fmax(qq[0] / P - mean * mean, 1e-12);
  const float inv = (float)(1.0 / sqrt(var)), mu = (float)mean;
  for (int p = threadIdx.x; p < P; p += 256) a[(int64_t)p * C] = (a[(int64_t)p * C] - mu) * inv;
}

// Bilinear warp of a gray image by Hi (destination → source, 1-based), zero outside, then a
// remap (0 none, 1: 1 − x², 2: √x, 3: sin πx) of the clamped value; one image per blockIdx.y.
__global__ __launch_bounds__(256) void k_warp_gray(const float* __restrict__ src, int w, int h,
                                                   const double* __restrict__ His, const int* __restrict__ remap,
                                                   float* dst) {
  const int img = blockIdx.y;
  const int64_t n = (int64_t)w * h;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double* Hi = His + 9 * img;
  const int x = (int)(i % w), y = (int)(i / w);
  const double X = x + 1.0, Y = y + 1.0;
  const double qz = Hi[6] * X + Hi[7] * Y + Hi[8];
  float v = 0.0f;
  if (fabs(qz) >= 1e-12) {
    const double fx = (Hi[0] * X + Hi[1] * Y + Hi[2]) / qz - 1.0, fy = (Hi[3] * X + Hi[4] * Y + Hi[5]) / qz - 1.0;
    if (fx >= 0.0 && fy >= 0.0 && fx <= w - 1.0 && fy <= h - 1.0) {
      const int x0 = (int)fx, y0 = (int)fy;
      const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
      const float ax = (float)(fx - x0), ay = (float)(fy - y0);
      const float* s = src + (int64_t)img * n;
      v = (s[(int64_t)y0 * w + x0] * (1.0f - ax) + s[(int64_t)y0 * w + x1] * ax) * (1.0f - ay) +
          (s[(int64_t)y1 * w + x0] * (1.0f - ax) + s[(int64_t)y1 * w + x1] * ax) * ay;
      v = fminf(fmaxf(v, 0.0f), 1.0f);
      switch (remap[img]) {
        case 1: v = 1.0f - v * v; break;
        case 2: v = sqrtf(v); break;
        case 3: v = sinpif(v); break;
        default: break;
      }
    }
  }
  dst[(int64_t)img * n + i] = v;
}

static inline unsigned tiles_blocks(int64_t tiles) { return (unsigned)((tiles + 3) / 4); }

}  // namespace zc

// ═══════════════════════════════════ C ABI (device pointers) ═══════════════════════════════════
using zc::ConvShape;

extern "C" {

static hipStream_t S_(void* st) { return static_cast<hipStream_t>(st); }

static ConvShape conv_shape(int n, int h, int w, int cin, int cout, int stride) {
  ConvShape s{n, h, w, cin, cout, stride, (h + stride - 1) / stride, (w + stride - 1) / stride};
  return s;
}

int32_t zc_cnn_conv_fwd(const float* x, const float* w, const float* b, float* y, int n, int h, int wd, int cin,
                        int cout, int stride, int relu, void* st) {
  if (stride < 1 || stride > 2 || h % stride || wd % stride) return -1;
  const ConvShape s = conv_shape(n, h, wd, cin, cout, stride);
  const int64_t tiles = (int64_t)((n * s.ho * s.wo + 15) / 16) * ((cout + 15) / 16);
  zc::k_conv_fwd<<<zc::tiles_blocks(tiles), 256, 0, S_(st)>>>(x, w, b, y, s, relu);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// dx (nullable: first layer), dw and db accumulate (the caller zeroes them per step).
int32_t zc_cnn_conv_bwd(const float* x, const float* w, const float* y, const float* dy, float* dx, float* dw,
                        float* db, int n, int h, int wd, int cin, int cout, int stride, int relu, void* st) {
  if (stride < 1 || stride > 2 || h % stride || wd % stride) return -1;
  const ConvShape s = conv_shape(n, h, wd, cin, cout, stride);
  if (dx) {
    const int64_t tiles = (int64_t)((n * h * wd + 15) / 16) * ((cin + 15) / 16);
    zc::k_conv_bwd_data<<<zc::tiles_blocks(tiles), 256, 0, S_(st)>>>(dy, y, w, dx, s, relu);
  }
  const int P = n * s.ho * s.wo, chunk = 2048;
  const int64_t tiles = (int64_t)((9 * cin + 15) / 16) * ((cout + 15) / 16);
  zc::k_conv_bwd_weight<<<dim3(zc::tiles_blocks(tiles), (P + chunk - 1) / chunk), 256, 0, S_(st)>>>(x, dy, y, dw, s,
                                                                                                       relu, chunk);
  zc::k_bias_grad<<<cout, 256, 0, S_(st)>>>(dy, y, db, P, cout, relu);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t zc_cnn_fc_fwd(const float* f, const float* w, const float* b, float* out, int n, int k, int j, void* st) {
  const int64_t tiles = (int64_t)((n + 15) / 16) * ((j + 15) / 16);
  zc::k_fc_fwd<<<zc::tiles_blocks(tiles), 256, 0, S_(st)>>>(f, w, b, out, n, k, j);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// df (nullable), dw and db accumulate.
int32_t zc_cnn_fc_bwd(const float* f, const float* w, const float* dout, float* df, float* dw, float* db, int n, int k,
                      int j, void* st) {
  if (df) {
    const int64_t tiles = (int64_t)((n + 15) / 16) * ((k + 15) / 16);
    zc::k_fc_bwd_data<<<zc::tiles_blocks(tiles), 256, 0, S_(st)>>>(dout, w, df, n, k, j);
  }
  const int64_t tiles = (int64_t)((k + 15) / 16) * ((j + 15) / 16);
  zc::k_fc_bwd_weight<<<zc::tiles_blocks(tiles), 256, 0, S_(st)>>>(f, dout, dw, n, k, j);
  zc::k_fc_bias_grad<<<(j + 255) / 256, 256, 0, S_(st)>>>(dout, db, n, j);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t zc_cnn_xent(const float* logits, const int* labels, float* loss, float* dlogits, int n, int j, void* st) {
  zc::k_xent<<<n, 64, 0, S_(st)>>>(logits, labels, loss, dlogits, n, j);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int32_t zc_cnn_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
                    int step, void* st) {
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  zc::k_adam<<<(unsigned)std::min<int64_t>((n + 255) / 256, 4096), 256, 0, S_(st)>>>(p, g, m, v, n, lr, b1, b2, eps, c1,
                                                                                        c2);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Input features of `count` gray images (w × h each, device): S × S downsample → rank → Sobel
// magnitude × Hann → 2-D FFT → log(1 + |F|) on T angles × R radii → channel `ch` of the NHWC
// [count][T][R][C] output. Writing the last channel (ch = C − 1) standardizes every channel of
// the output per map (zero mean, unit deviation). work: count · S² · 12 bytes of device scratch.
int32_t zc_lp_features(const float* imgs, int count, int w, int h, int S, int T, int R, int C, int ch, float r0,
                       float r1, float* out, void* work, void* st) {
  if (S < 32 || S > 4096 || (S & (S - 1)) || ch < 0 || ch >= C) return -1;
  hipStream_t s = S_(st);
  float* down = static_cast<float*>(work);
  float2* plane = reinterpret_cast<float2*>(down + (size_t)count * S * S);
  const dim3 g((S * S + 255) / 256, count);
  zc::k_lp_down<<<g, 256, 0, s>>>(imgs, w, h, down, S);
  zc::k_lp_rank<<<count, 256, 0, s>>>(down, S * S);
  zc::k_lp_grad<<<g, 256, 0, s>>>(down, plane, S);
  if (!zc::launch_fft2_twiddled(plane, count, S, false, s)) return -1;
  zc::k_lp_polar<<<dim3((T * R + 255) / 256, count), 256, 0, s>>>(plane, S, out, T, R, C, ch, r0, r1);
  if (ch == C - 1) zc::k_lp_norm<<<count * C, 256, 0, s>>>(out, T * R, C);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Warp `count` gray images (w × h, device) by His (count × 9 doubles, destination → source,
// 1-based) with remaps (count ints) into dst.
int32_t zc_warp_gray(const float* src, int count, int w, int h, const double* His, const int* remap, float* dst,
                     void* st) {
  const int64_t n = (int64_t)w * h;
  zc::k_warp_gray<<<dim3((unsigned)((n + 255) / 256), count), 256, 0, S_(st)>>>(src, w, h, His, remap, dst);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
