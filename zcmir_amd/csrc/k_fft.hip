// Batched column FFTs (every tiling) and the generic 2-D FFT.
#include "kdev.hpp"

namespace zc {

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
  static constexpr int M = L % 3 == 0 ? 12 : 16;   // 2^a·3^b points: 12 per thread
  static constexpr int T = C * L / M;
  static constexpr int STR = line_stride(L, C);
  static constexpr bool fits = C * STR * 8 <= 160 * 1024 && T <= 1024 && T >= 16;
};

// lines (columns) per workgroup of the column FFT: the default, or a tuned value that fits
// the widest row segments whose LDS fits (measured best at 2048: 8 columns, 1 workgroup per CU)
inline int col_lines_default(int L) {
  if (L % 3 == 0) return L >= 2304 ? 4 : 8;
  return L >= 4096 ? 4 : (L >= 2048 ? 8 : 16);
}
static_assert(ColT<32, 16>::fits && ColT<64, 16>::fits && ColT<128, 16>::fits && ColT<256, 16>::fits &&
                  ColT<512, 16>::fits && ColT<1024, 16>::fits && ColT<1536, 8>::fits && ColT<2048, 8>::fits &&
                  ColT<2304, 4>::fits && ColT<3072, 4>::fits && ColT<4096, 4>::fits,
              "every default column tiling must be launchable");

// skip (nullable; honoured for columns of up to 1024 points and of 2^a·3^b points, where the
// check costs no occupancy): per entry of `per` consecutive planes, output rows [x, y) that no
// later pass reads, left unstored.
template <int L, int S, int C>
__global__ __launch_bounds__((ColT<L, C>::T)) void k_colfft(float2* data, int nx, int planes, int rows_in, float scale,
                                                          const float2* __restrict__ tw, const int4* __restrict__ skip,
                                                          int per) {
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
  if constexpr (L <= 1024 || L % 3 == 0) {
    const int4 sk = skip ? skip[plane / per] : make_int4(0, 0, 0, 0);
    rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
      if (y < sk.x || y >= sk.y) base[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
    });
  } else {
    rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
      base[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
    });
  }
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
      (L % 3 != 0) == (M % 3 != 0) &&   // a line of 2^a·3^b points needs M divisible by 3, and only it
      (M < 32 || T <= 512) &&   // 32 complex points per lane need more than the 128 VGPRs of a 1024-lane block
      (M < 64 || T == 512);     // 64 points: 128 VGPRs of data, two waves per SIMD
};

// skip as in k_colfft, honoured for columns of 2^a·3^b points (the 2048 and 4096 tilings lose a
// wave per SIMD to the check's registers)
template <int L, int S, int C, int M, bool SPLIT>
__global__ __launch_bounds__((ColW<L, C, M, SPLIT>::T), (SPLIT && M == 16 ? 2 : 1)) void k_colfft_w(float2* data, int nx, int planes, int rows_in,
                                                                       float scale, const float2* __restrict__ tw,
                                                                       const int4* __restrict__ skip, int per) {
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
  if constexpr (L % 3 == 0) {
    const int4 sk = skip ? skip[plane / per] : make_int4(0, 0, 0, 0);
    rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
      if (y < sk.x || y >= sk.y) base[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
    });
  } else {
    rstore<L, M, T, C, true>(v, tid, [&](int c, int y, float2 val) {
      base[(unsigned)(y * nx + c)] = make_float2(val.x * scale, val.y * scale);
    });
  }
}

template <int L, int S, int C, int M, bool SPLIT>
static bool colfft_w_LC(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw, hipStream_t st,
                        const int4* skip = nullptr, int per = 1) {
  if constexpr (!ColW<L, C, M, SPLIT>::fits) {
    return false;
  } else {
    if (nx % C) return false;
    const unsigned g = (unsigned)(nx / C * planes);
    k_colfft_w<L, S, C, M, SPLIT><<<g, ColW<L, C, M, SPLIT>::T, 0, st>>>(data, nx, planes, rows_in, scale, tw, skip,
                                                                         per);
    return true;
  }
}

template <int L, int S, int C>
static bool colfft_LC(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw, hipStream_t st,
                      const int4* skip = nullptr, int per = 1) {
  if constexpr (!ColT<L, C>::fits) {
    return false;
  } else {
    if (nx % C) return false;
    const unsigned g = (unsigned)(nx / C * planes);
    k_colfft<L, S, C><<<g, ColT<L, C>::T, 0, st>>>(data, nx, planes, rows_in, scale, tw, skip, per);
    return true;
  }
}

// skip: output rows not to store (k_colfft), honoured by the plain tilings; the others store
// every row
template <int L, int S>
static bool colfft_L(float2* data, int planes, int nx, int rows_in, float scale, const float2* tw, int lines,
                     hipStream_t st, const int4* skip = nullptr, int per = 1) {
  switch (lines > 0 ? lines : col_lines_default(L)) {
    case 1: return colfft_LC<L, S, 1>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 2: return colfft_LC<L, S, 2>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 4: return colfft_LC<L, S, 4>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 8: return colfft_LC<L, S, 8>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 16: return colfft_LC<L, S, 16>(data, planes, nx, rows_in, scale, tw, st, skip, per);
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
    case 41: return colfft_w_LC<L, S, 4, 16, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 42: return colfft_w_LC<L, S, 8, 16, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 43: return colfft_w_LC<L, S, 8, 32, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 44: return colfft_w_LC<L, S, 16, 32, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 45: return colfft_w_LC<L, S, 8, 32, false>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 46: return colfft_w_LC<L, S, 4, 32, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 47: return colfft_w_LC<L, S, 2, 32, false>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 48: return colfft_w_LC<L, S, 16, 64, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 49: return colfft_w_LC<L, S, 8, 64, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    // lines of 2^a·3^b points: 50: 8x48, 51: 4x48, 52: 8x24, 53: 4x24 (split exchange)
    case 50: return colfft_w_LC<L, S, 8, 48, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 51: return colfft_w_LC<L, S, 4, 48, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 52: return colfft_w_LC<L, S, 8, 24, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    case 53: return colfft_w_LC<L, S, 4, 24, true>(data, planes, nx, rows_in, scale, tw, st, skip, per);
    default: return false;
  }
}


bool launch_colfft(float2* data, int planes, int nx, int ny, int rows_in, bool inverse, const float2* tw_y,
                   hipStream_t st, int lines, const int4* skip, int per) {
  // a tiling that cannot launch at this width (nx not a multiple of its columns) falls back to
  // the default tiling
#define CALL(LL)                                                                                   \
  {                                                                                                \
    auto run = [&](int ln) {                                                                       \
      return inverse ? colfft_L<LL, 1>(data, planes, nx, rows_in, 1.0f, tw_y, ln, st, skip, per)   \
                     : colfft_L<LL, -1>(data, planes, nx, rows_in, 1.0f, tw_y, ln, st, skip, per); \
    };                                                                                             \
    return run(lines) || (lines != 0 && run(0));                                                   \
  }
  ZC_FFT_SWITCH(ny, CALL)
#undef CALL
  return true;
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
  ZC_FFT_SWITCH(n, CALL)
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


}  // namespace zc
