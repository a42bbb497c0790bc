// Work-image preprocessing launchers (pre.hip): RGBA8 to gray, CLAHE, band pass, invert, texels.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace zc {

// CLAHE geometry (gls.zig:462-489): hg x hg tiles of h_tile x w_tile over the image padded by
// reflection in front (h_pad rows, w_pad columns), nb bins, clip limit
struct ClaheArgs {
  int w, h, hg, nb, h_pad, w_pad, h_tile, w_tile;
  float clip;
};
ClaheArgs clahe_args(int w, int h, double clip, uint32_t n_bins, uint32_t grid);

void launch_rgba_gray(const uint32_t* rgba, float* gray, int64_t n, hipStream_t st);
// dst = src, rounded to f16 when half (the work image's texels)
void launch_store(const float* src, float* dst, int64_t n, bool half, hipStream_t st);
// hist: hg·hg·nb u32, cdf: hg·hg·nb f32 (scratch)
void launch_clahe(const float* src, const ClaheArgs& a, uint32_t* hist, float* cdf, float* dst, bool half,
                  hipStream_t st);
// one Gaussian pass (sigma 1, radius 3, clamped borders) along x (axis 1) or y (axis 0)
void launch_gauss(const float* src, float* dst, int w, int h, int axis, bool half, hipStream_t st);
int band_parts(int64_t n);
// dst = (v - lo) / (hi - lo), v = clamp(fine) - clamp(coarse); v: n floats, part: 2·band_parts(n)
void launch_band_finish(const float* fine, const float* coarse, float* v, float* part, float* dst, int64_t n, bool half,
                        hipStream_t st);
void launch_invert(float* buf, int64_t n, bool half, hipStream_t st);

}  // namespace zc
