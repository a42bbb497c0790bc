// Map kernels, inverse side: row IFFT + extract of the correlation planes (every row tiling).
#include "kdev.hpp"

namespace zc {

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


}  // namespace zc
