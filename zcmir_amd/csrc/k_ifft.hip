// Map kernels, inverse side: row IFFT + extract of the correlation planes (every row tiling).
#include "kdev.hpp"

namespace zc {

// ── row IFFT + extract ────────────────────────────────────────────────────────────────────
// Inverse row FFT of complex output q (rows of L) → correlation planes 2q (re) and 2q+1 (im).
// blockIdx.y = b · n_out + q for batch entry b (its planes at b · n_corr).

template <int L, int PID, int V>
__global__ __launch_bounds__((RowX<L, V>::T)) void k_rowifft_extract(const float2* __restrict__ work, float* corr,
                                                                    int ny, const float2* __restrict__ tw,
                                                                    const int4* __restrict__ dead) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int T = RowX<L, V>::T, M = RowX<L, V>::M, LINES = RowX<L, V>::LINES, STR = RowX<L, V>::STR;
  constexpr int KN = P.kept_of[0];   // the plane of the overlap count n
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x, g = blockIdx.y;
  const int q = g % P.n_out, b = g / P.n_out;
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * ny;
  const float2* src = work + (int64_t)g * N;
  corr += (int64_t)b * P.n_corr * N;
  const int4 d = dead ? dead[b] : make_int4(0, 0, 0, 0);
  if (KN >= 0 && y0 >= d.x && y0 + LINES <= d.y) {
    // every line of the workgroup is a lag row with no overlap: n = 0 there (the exact value of
    // a correlation with nothing in common), and the score reads nothing else below its floor
    if (q == KN / 2) {
      float* on = corr + (int64_t)KN * N + (int64_t)y0 * L;
      for (int i = tid; i < LINES * L; i += T)
        if (y0 + i / L < ny) on[i] = 0.0f;
    }
    return;
  }
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
      if constexpr (L <= 1024 || LINES > 1) {
        // short lines (few registers per lane) or several lines per workgroup: lag columns
        // without overlap store n only; a lag row without overlap in a workgroup with other rows
        // (the column pass may have left its input unstored) stores n = 0
        const bool dead_row = y >= d.x && y < d.y;
        const bool live = !dead_row && (x < d.z || x >= d.w);
        if (2 * q == KN) o1[(int64_t)y * L + x] = dead_row ? 0.0f : val.x;
        else if (live) o1[(int64_t)y * L + x] = val.x;
        if (two) {
          if (2 * q + 1 == KN) o2[(int64_t)y * L + x] = dead_row ? 0.0f : val.y;
          else if (live) o2[(int64_t)y * L + x] = val.y;
        }
      } else {
        o1[(int64_t)y * L + x] = val.x;
        if (two) o2[(int64_t)y * L + x] = val.y;
      }
    }
  });
}


// The rectangle plans' layout: correlation k in component k & 3 of quad plane k >> 2 (float4 per
// lag), so the per-lag score reads four correlations per load. Workgroup (row y, quad p) inverts
// outputs 2p and 2p+1 of row y together (two register groups per thread, one LDS line) and
// stores the four planes' values of each lag as one float4; a missing second output stores 0.
template <int L, int PID>
__global__ __launch_bounds__(L / 16) void k_rowifft_quads(const float2* __restrict__ work, float4* corr4, int ny,
                                                          const float2* __restrict__ tw) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int M = 16, T = L / 16, STR = padded_len(L);
  __shared__ float2 lds[STR];
  const int tid = threadIdx.x, y = blockIdx.x, p = blockIdx.y;
  const int q0 = 2 * p;
  const bool two = q0 + 1 < P.n_out;
  const int64_t N = (int64_t)L * ny;
  const float2* r0 = work + (int64_t)q0 * N + (int64_t)y * L;
  const float2* r1 = r0 + N;
  float2 v[2][M];
  rload<L, M, T, 1, false>(v[0], tid, [&](int, int x) { return r0[x]; });
  rload<L, M, T, 1, false>(v[1], tid, [&](int, int x) { return two ? r1[x] : make_float2(0.0f, 0.0f); });
  rfft_groups<L, M, T, 1, false, 1, 2>(v, lds, STR, tw, tid);
  float4* out = corr4 + (int64_t)p * N + (int64_t)y * L;
  int i = 0;   // rstore visits v[0][0], v[0][1], ... in order: v[1][i] is the same lag of output 2p + 1
  rstore<L, M, T, 1, false>(v[0], tid, [&](int, int x, float2 a) {
    const float2 b = v[1][i++];
    out[x] = make_float4(a.x, a.y, b.x, b.y);
  });
}

template <int L>
static void rowifft_quads_L(int plan, const float2* work, float4* corr4, int ny, const float2* tw, hipStream_t st) {
  {
    const dim3 g(ny, (kPlans[plan].n_out + 1) / 2);
    switch (plan) {
      case 4: k_rowifft_quads<L, 4><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      case 5: k_rowifft_quads<L, 5><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      case 6: k_rowifft_quads<L, 6><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      case 7: k_rowifft_quads<L, 7><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      case 8: k_rowifft_quads<L, 8><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      case 9: k_rowifft_quads<L, 9><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      case 10: k_rowifft_quads<L, 10><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
      default: k_rowifft_quads<L, 11><<<g, L / 16, 0, st>>>(work, corr4, ny, tw); break;
    }
  }
}

bool launch_rowifft_quads(int plan, const float2* work, float4* corr4, int nx, int ny, const float2* tw_x,
                          hipStream_t st) {
  if (plan < 4) return false;
#define CALL(LL) rowifft_quads_L<LL>(plan, work, corr4, ny, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

template <int L, int V>
static void rowifft_LV(int plan, const float2* work, float* corr, int ny, const float2* tw, hipStream_t st, int batch,
                       const int4* dead) {
  using X = RowX<L, V>;
  dim3 g((ny + X::LINES - 1) / X::LINES, kPlans[plan].n_out * batch);
  switch (plan) {
    case 0: k_rowifft_extract<L, 0, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 1: k_rowifft_extract<L, 1, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 2: k_rowifft_extract<L, 2, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 3: k_rowifft_extract<L, 3, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 4: k_rowifft_extract<L, 4, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 5: k_rowifft_extract<L, 5, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 6: k_rowifft_extract<L, 6, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 7: k_rowifft_extract<L, 7, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 8: k_rowifft_extract<L, 8, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 9: k_rowifft_extract<L, 9, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    case 10: k_rowifft_extract<L, 10, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    default: k_rowifft_extract<L, 11, V><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
  }
}

template <int L>
static void rowifft_L(int plan, const float2* work, float* corr, int ny, const float2* tw, hipStream_t st, int batch,
                      const int4* dead) {
  if constexpr (L % 3 == 0) {
    // lines of 2^a·3^b points (the sweep's planes): the default tiling, plans without rectangle sums
    using X = RowX<L, 0>;
    dim3 g((ny + X::LINES - 1) / X::LINES, kPlans[plan & 3].n_out * batch);
    switch (plan & 3) {
      case 0: k_rowifft_extract<L, 0, 0><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
      case 1: k_rowifft_extract<L, 1, 0><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
      case 2: k_rowifft_extract<L, 2, 0><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
      default: k_rowifft_extract<L, 3, 0><<<g, X::T, 0, st>>>(work, corr, ny, tw, dead); break;
    }
    return;
  }
  if constexpr (L >= 256) {
    switch (FftRowTiling::v) {
      case 1: rowifft_LV<L, 1>(plan, work, corr, ny, tw, st, batch, dead); return;
      case 2: rowifft_LV<L, 2>(plan, work, corr, ny, tw, st, batch, dead); return;
      case 3: rowifft_LV<L, 3>(plan, work, corr, ny, tw, st, batch, dead); return;
      case 4: rowifft_LV<L, 4>(plan, work, corr, ny, tw, st, batch, dead); return;
      default: break;
    }
  }
  rowifft_LV<L, 0>(plan, work, corr, ny, tw, st, batch, dead);
}

bool launch_rowifft_extract(int plan, const float2* work, float* corr, int nx, int ny, const float2* tw_x,
                            hipStream_t st, int batch, const int4* dead) {
  if (nx % 3 == 0 && plan > 3) return false;
#define CALL(LL) rowifft_L<LL>(plan, work, corr, ny, tw_x, st, batch, dead)
  ZC_FFT_SWITCH(nx, CALL)
#undef CALL
  return true;
}


}  // namespace zc
