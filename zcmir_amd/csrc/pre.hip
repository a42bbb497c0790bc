// Work-image preprocessing on gfx950: RGBA8 to gray, CLAHE, the band pass (Gaussian passes,
// difference, rescale), invert and f16 texel rounding (gls.zig:409-554; gls_mift.wgsl:283-318,
// 1973-2066). Every step is float32 in the order the reference writes it, compiled without
// contraction (no FMA), so the results are reproducible bit for bit.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "pre.hpp"

#pragma clang fp contract(off)

namespace zc {

__device__ __forceinline__ float store_t(float x, int half) { return half ? __half2float(__float2half_rn(x)) : x; }
__device__ __forceinline__ float clamp01(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
// WGSL mix(a, b, t) = a (1 - t) + b t
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// ── RGBA8 → gray: 0.299 R + 0.587 G + 0.114 B of the channels' k / 255 (f32), left to right ──
struct Unorm8 {
  float v[256];
};
constexpr Unorm8 make_unorm8() {
  Unorm8 t{};
  for (int k = 0; k < 256; ++k) t.v[k] = (float)((double)k / 255.0);
  return t;
}
__constant__ Unorm8 kUnorm8 = make_unorm8();

__global__ __launch_bounds__(256) void k_rgba_gray(const uint32_t* __restrict__ rgba, float* gray, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t w = rgba[i];
  const float r = kUnorm8.v[w & 0xffu], g = kUnorm8.v[(w >> 8) & 0xffu], b = kUnorm8.v[(w >> 16) & 0xffu];
  gray[i] = (0.299f * r + 0.587f * g) + 0.114f * b;
}

__global__ __launch_bounds__(256) void k_store(const float* __restrict__ src, float* dst, int64_t n, int half) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = store_t(src[i], half);
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ── CLAHE ──
__device__ __forceinline__ int clahe_reflect(int i, int n, int pad) {
  if (n <= 1) return 0;
  int x = i - pad;
  if (x < 0) x = -x;
  if (x >= n) x = 2 * n - 2 - x;
  return min(max(x, 0), n - 1);
}

// blockIdx.x: tile (ty · hg + tx), blockIdx.y: a slice of the tile's rows; LDS histogram, then
// one global add per nonzero bin.
__global__ __launch_bounds__(256) void k_clahe_hist(const float* __restrict__ src, ClaheArgs a, uint32_t* hist) {
  __shared__ uint32_t lh[256];
  for (int b = threadIdx.x; b < a.nb; b += 256) lh[b] = 0u;
  __syncthreads();
  const int tile = blockIdx.x, ty = tile / a.hg, tx = tile - ty * a.hg;
  const int rows_per = (a.h_tile + gridDim.y - 1) / gridDim.y;
  const int r0 = blockIdx.y * rows_per, r1 = min(a.h_tile, r0 + rows_per);
  const int cnt = (r1 > r0 ? r1 - r0 : 0) * a.w_tile;
  for (int k = threadIdx.x; k < cnt; k += 256) {
    const int gy = ty * a.h_tile + r0 + k / a.w_tile, gx = tx * a.w_tile + k % a.w_tile;
    const int oy = clahe_reflect(gy, a.h, a.h_pad), ox = clahe_reflect(gx, a.w, a.w_pad);
    const float v = clamp01(src[(int64_t)oy * a.w + ox]);
    const float bf = fminf(fmaxf(v * (float)(a.nb - 1), 0.0f), (float)(a.nb - 1));
    atomicAdd(&lh[(int)bf], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < a.nb; b += 256)
    if (lh[b]) atomicAdd(&hist[tile * a.nb + b], lh[b]);
}

// One thread per tile: clip at max(clip · voxels / nb, 1), spread the excess evenly (the
// remainder to the first bins), normalized cumulative sums.
__global__ __launch_bounds__(64) void k_clahe_cdf(const uint32_t* __restrict__ hist, ClaheArgs a, float* cdf) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= a.hg * a.hg) return;
  const int nb = a.nb;
  const uint32_t voxels = (uint32_t)max(a.h_tile * a.w_tile, 1);
  const float limit = fmaxf(a.clip * (float)voxels / (float)nb, 1.0f);
  float sumc = 0.0f;
  for (int i = 0; i < nb; ++i) sumc = sumc + fminf((float)hist[t * nb + i], limit);
  const float clipped = (float)voxels - sumc;
  const float nbf = (float)nb;
  const int residual = (int)(clipped - nbf * floorf(clipped / nbf));
  const float redist = (clipped - (float)residual) / nbf;
  const float den = fmaxf((float)voxels, 1.0f);
  float acc = 0.0f;
  for (int i = 0; i < nb; ++i) {
    float c = fminf((float)hist[t * nb + i], limit) + redist;
    if (i < residual) c = c + 1.0f;
    acc = acc + c;
    cdf[t * nb + i] = clamp01(acc / den);
  }
}

__global__ __launch_bounds__(256) void k_clahe_apply(const float* __restrict__ src, ClaheArgs a,
                                                     const float* __restrict__ cdf, float* dst, int half) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)a.w * a.h) return;
  const int y = (int)(i / a.w), x = (int)(i - (int64_t)y * a.w);
  const int hg = a.hg, nb = a.nb;
  const float v = clamp01(src[i]);
  const float gy = ((float)y + 0.5f) / (float)a.h * (float)(hg - 1);
  const float gx = ((float)x + 0.5f) / (float)a.w * (float)(hg - 1);
  const int y0 = (int)floorf(gy), x0 = (int)floorf(gx);
  const int y1 = min(y0 + 1, hg - 1), x1 = min(x0 + 1, hg - 1);
  const float fy = gy - (float)y0, fx = gx - (float)x0;
  const float bf = v * (float)(nb - 1);
  const int b0 = (int)fminf(fmaxf(floorf(bf), 0.0f), (float)(nb - 2));
  const float fb = bf - (float)b0;
  const int b1 = b0 + 1;
  auto at = [&](int ty, int tx) {
    const float* c = cdf + (ty * hg + tx) * nb;
    return mixf(c[b0], c[b1], fb);
  };
  const float c0 = mixf(at(y0, x0), at(y0, x1), fx);
  const float c1 = mixf(at(y1, x0), at(y1, x1), fx);
  dst[i] = store_t(mixf(c0, c1, fy), half);
}

// ── band pass ──
// sigma 1, radius 3: the float32 roundings of exp(-k^2 / 2), k = -3..3
__constant__ float kGauss7[7] = {0.011108996538242306f, 0.1353352832366127f, 0.6065306597126334f, 1.0f,
                                 0.6065306597126334f, 0.1353352832366127f, 0.011108996538242306f};

__global__ __launch_bounds__(256) void k_gauss(const float* __restrict__ src, float* dst, int w, int h, int axis,
                                               int half) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)w * h) return;
  const int y = (int)(i / w), x = (int)(i - (int64_t)y * w);
  float acc = 0.0f, wsum = 0.0f;
#pragma unroll
  for (int k = -3; k <= 3; ++k) {
    const float ww = kGauss7[k + 3];
    const int xx = axis == 1 ? min(max(x + k, 0), w - 1) : x;
    const int yy = axis == 0 ? min(max(y + k, 0), h - 1) : y;
    acc = acc + src[(int64_t)yy * w + xx] * ww;
    wsum = wsum + ww;
  }
  dst[i] = store_t(acc / wsum, half);
}

// v = clamp(fine) - clamp(coarse); per-block min / max partials
__global__ __launch_bounds__(256) void k_band_diff(const float* __restrict__ fine, const float* __restrict__ coarse,
                                                   float* v, int64_t n, float* part) {
  __shared__ float smin[256], smax[256];
  float lo = INFINITY, hi = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float d = clamp01(fine[i]) - clamp01(coarse[i]);
    v[i] = d;
    lo = fminf(lo, d);
    hi = fmaxf(hi, d);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + k]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = smin[0];
    part[2 * blockIdx.x + 1] = smax[0];
  }
}

// (v - lo) · sc, sc = 1 / (hi - lo) (0 when hi <= lo), lo / hi the min / max over the partials
__global__ __launch_bounds__(256) void k_band_scale(const float* __restrict__ v, const float* __restrict__ part,
                                                    int nparts, float* dst, int64_t n, int half) {
  __shared__ float smin[256], smax[256];
  float lo = INFINITY, hi = -INFINITY;
  for (int k = threadIdx.x; k < nparts; k += 256) {
    lo = fminf(lo, part[2 * k]);
    hi = fmaxf(hi, part[2 * k + 1]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + k]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + k]);
    }
    __syncthreads();
  }
  lo = smin[0];
  hi = smax[0];
  const float sc = hi > lo ? 1.0f / (hi - lo) : 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    dst[i] = store_t((v[i] - lo) * sc, half);
}

__global__ __launch_bounds__(256) void k_invert(float* buf, int64_t n, int half) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) buf[i] = store_t(1.0f - clamp01(buf[i]), half);
}

// ── launchers ──
static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

ClaheArgs clahe_args(int w, int h, double clip, uint32_t n_bins0, uint32_t grid) {
  ClaheArgs a{};
  a.w = w;
  a.h = h;
  a.hg = (int)std::min<uint32_t>(8, std::max<uint32_t>(2, grid));
  a.nb = (int)std::min<uint32_t>(256, std::max<uint32_t>(8, n_bins0));
  a.h_pad = (a.hg - h % a.hg) % a.hg;
  a.w_pad = (a.hg - w % a.hg) % a.hg;
  a.h_tile = std::max(1, (h + a.h_pad) / a.hg);
  a.w_tile = std::max(1, (w + a.w_pad) / a.hg);
  a.clip = (float)clip;
  return a;
}

void launch_rgba_gray(const uint32_t* rgba, float* gray, int64_t n, hipStream_t st) {
  k_rgba_gray<<<blocks(n), 256, 0, st>>>(rgba, gray, n);
}

void launch_store(const float* src, float* dst, int64_t n, bool half, hipStream_t st) {
  k_store<<<blocks(n), 256, 0, st>>>(src, dst, n, half ? 1 : 0);
}

void launch_clahe(const float* src, const ClaheArgs& a, uint32_t* hist, float* cdf, float* dst, bool half,
                  hipStream_t st) {
  const int nt = a.hg * a.hg, nbins = nt * a.nb;
  k_zero_u32<<<blocks(nbins), 256, 0, st>>>(hist, nbins);
  const int slices = std::max(1, std::min(64, a.h_tile / 16));
  k_clahe_hist<<<dim3(nt, slices), 256, 0, st>>>(src, a, hist);
  k_clahe_cdf<<<(nt + 63) / 64, 64, 0, st>>>(hist, a, cdf);
  k_clahe_apply<<<blocks((int64_t)a.w * a.h), 256, 0, st>>>(src, a, cdf, dst, half ? 1 : 0);
}

void launch_gauss(const float* src, float* dst, int w, int h, int axis, bool half, hipStream_t st) {
  k_gauss<<<blocks((int64_t)w * h), 256, 0, st>>>(src, dst, w, h, axis, half ? 1 : 0);
}

int band_parts(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(1, (n + 255) / 256), 1024); }

void launch_band_finish(const float* fine, const float* coarse, float* v, float* part, float* dst, int64_t n, bool half,
                        hipStream_t st) {
  const int np = band_parts(n);
  k_band_diff<<<np, 256, 0, st>>>(fine, coarse, v, n, part);
  k_band_scale<<<np, 256, 0, st>>>(v, part, np, dst, n, half ? 1 : 0);
}

void launch_invert(float* buf, int64_t n, bool half, hipStream_t st) {
  k_invert<<<blocks(n), 256, 0, st>>>(buf, n, half ? 1 : 0);
}

}  // namespace zc
