// Register-resident batched FFT lines for gfx950.
//
// T threads transform LINES lines of length L (a power of two, 32..4096) together; each thread
// holds M = LINES·L/T points in registers. A line runs as Stockham stages of radix ≤ 16 (each
// radix divides M, so every thread does whole butterflies every stage): the first stage's inputs
// come straight from memory (the caller's loader), the last stage's outputs go straight to
// memory (the caller's storer), and between stages the points move through LDS once (write,
// barrier, read, barrier). A 2048-point line is 16·16·8: two LDS exchanges.
//
// Butterfly b of a stage belongs to thread b mod T (b = tid + u·T). Rows map b to (line, j) with
// j fastest (consecutive threads read consecutive points of a line); columns (LFAST) map line
// fastest, so consecutive threads touch consecutive columns of one row in memory.
#pragma once
#include <hip/hip_runtime.h>

#include "fft.hpp"

namespace zc {

template <int S>
__device__ __forceinline__ void dft16(float2* v) {
  // 16 = 4 x 4: DFT4 over n1 for each n2, twiddle w16^(n2 k1), DFT4 over n2 for each k1
  constexpr float C1 = 0.92387953251128675613f;   // cos(π/8)
  constexpr float S1 = 0.38268343236508977173f;   // sin(π/8)
  constexpr float C2 = 0.70710678118654752440f;
  float2 y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 a0 = v[n2], a1 = v[4 + n2], a2 = v[8 + n2], a3 = v[12 + n2];
    dft4<S>(a0, a1, a2, a3);
    y[n2][0] = a0;
    y[n2][1] = a1;
    y[n2][2] = a2;
    y[n2][3] = a3;
  }
  // w16^m = cos(2πm/16) + S i sin(2πm/16)
  auto tw = [](float2 a, float c, float s) { return make_float2(a.x * c - a.y * s, a.x * s + a.y * c); };
  y[1][1] = tw(y[1][1], C1, S * S1);
  y[1][2] = tw(y[1][2], C2, S * C2);
  y[1][3] = tw(y[1][3], S1, S * C1);
  y[2][1] = tw(y[2][1], C2, S * C2);
  y[2][2] = mul_si<S>(y[2][2]);
  y[2][3] = tw(y[2][3], -C2, S * C2);
  y[3][1] = tw(y[3][1], S1, S * C1);
  y[3][2] = tw(y[3][2], -C2, S * C2);
  y[3][3] = tw(y[3][3], -C1, -S * S1);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b0 = y[0][k1], b1 = y[1][k1], b2 = y[2][k1], b3 = y[3][k1];
    dft4<S>(b0, b1, b2, b3);
    v[k1] = b0;
    v[k1 + 4] = b1;
    v[k1 + 8] = b2;
    v[k1 + 12] = b3;
  }
}

// Mixed radix N = N1 · N2 (Cooley–Tukey, input n2 + N2 n1, output k1 + N1 k2): DFT_N1 over n1 for
// each n2, twiddle w_N^(n2 k1), DFT_N2 over n2 for each k1. 6 = 2 x 3 and 12 = 4 x 3 for lines of
// 2^a·3^b points; w_N^m = cos(2πm/N) + S i sin(2πm/N).
template <int S>
__device__ __forceinline__ void dft6(float2* v) {
  constexpr float H = 0.86602540378443864676f;   // √3/2
  auto tw = [](float2 a, float c, float s) { return make_float2(a.x * c - a.y * s, a.x * s + a.y * c); };
  float2 y[3][2];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    const float2 a = v[n2], b = v[n2 + 3];
    y[n2][0] = cadd(a, b);
    y[n2][1] = csub(a, b);
  }
  y[1][1] = tw(y[1][1], 0.5f, S * H);    // w6^1
  y[2][1] = tw(y[2][1], -0.5f, S * H);   // w6^2
#pragma unroll
  for (int k1 = 0; k1 < 2; ++k1) {
    float2 a = y[0][k1], b = y[1][k1], c = y[2][k1];
    dft3<S>(a, b, c);
    v[k1] = a;
    v[k1 + 2] = b;
    v[k1 + 4] = c;
  }
}

template <int S>
__device__ __forceinline__ void dft12(float2* v) {
  constexpr float H = 0.86602540378443864676f;   // √3/2
  auto tw = [](float2 a, float c, float s) { return make_float2(a.x * c - a.y * s, a.x * s + a.y * c); };
  float2 y[3][4];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    float2 a0 = v[n2], a1 = v[n2 + 3], a2 = v[n2 + 6], a3 = v[n2 + 9];
    dft4<S>(a0, a1, a2, a3);
    y[n2][0] = a0;
    y[n2][1] = a1;
    y[n2][2] = a2;
    y[n2][3] = a3;
  }
  y[1][1] = tw(y[1][1], H, S * 0.5f);       // w12^1
  y[1][2] = tw(y[1][2], 0.5f, S * H);       // w12^2
  y[1][3] = mul_si<S>(y[1][3]);             // w12^3
  y[2][1] = tw(y[2][1], 0.5f, S * H);       // w12^2
  y[2][2] = tw(y[2][2], -0.5f, S * H);      // w12^4
  y[2][3] = make_float2(-y[2][3].x, -y[2][3].y);   // w12^6
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 a = y[0][k1], b = y[1][k1], c = y[2][k1];
    dft3<S>(a, b, c);
    v[k1] = a;
    v[k1 + 4] = b;
    v[k1 + 8] = c;
  }
}

template <int R, int S>
__device__ __forceinline__ void dftr(float2* v) {
  if constexpr (R == 16) dft16<S>(v);
  else if constexpr (R == 12) dft12<S>(v);
  else if constexpr (R == 8) dft8<S>(v);
  else if constexpr (R == 6) dft6<S>(v);
  else if constexpr (R == 4) dft4v<S>(v);
  else if constexpr (R == 3) dft3v<S>(v);
  else dft2<S>(v);
}

// Radices of a length-L line (2^a·3^b) when each thread holds M points: the largest radix of
// {16, 12, 8, 6, 4, 3, 2} dividing both M and what is left, with a trailing (16, 2) / (8, 2)
// rebalanced to (8, 4) / (4, 4). A line with a factor 3 needs M divisible by 3.
template <int L, int M>
struct RPlan {
  struct Arr {
    int r[16];
    int n;
    bool ok;
  };
  static constexpr Arr make() {
    Arr a{};
    constexpr int cand[7] = {16, 12, 8, 6, 4, 3, 2};
    int rem = L;
    a.ok = true;
    while (rem > 1 && a.n < 16) {
      int R = 0;
      for (int i = 0; i < 7 && !R; ++i)
        if (M % cand[i] == 0 && rem % cand[i] == 0) R = cand[i];
      if (!R) {
        a.ok = false;
        break;
      }
      a.r[a.n++] = R;
      rem /= R;
    }
    if (a.n >= 2 && a.r[a.n - 1] == 2 && a.r[a.n - 2] >= 8 && (a.r[a.n - 2] & (a.r[a.n - 2] - 1)) == 0 && M % 4 == 0) {
      a.r[a.n - 2] /= 2;
      a.r[a.n - 1] = 4;
    }
    return a;
  }
  static constexpr Arr A = make();
  static_assert(A.ok, "no radix plan: the points per thread must hold every odd factor of the line");
  static constexpr int stages = A.n;
  static constexpr int radix(int s) { return A.r[s]; }
  static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * A.r[s - 1]; }
};

// Butterfly b → (line, j) of a stage with NB butterflies per line.
template <int LINES, int NB, bool LFAST>
__device__ __forceinline__ void bf_index(int b, int& line, int& j) {
  if constexpr (LFAST) {
    line = b % LINES;
    j = b / LINES;
  } else {
    line = b / NB;
    j = b - line * NB;
  }
}

template <int L, int M, int T, int LINES, bool LFAST, int S, int ST>
__device__ __forceinline__ void rstage_compute(float2 (&v)[M], const float2* __restrict__ tw, int tid) {
  using P = RPlan<L, M>;
  constexpr int R = P::radix(ST), NS = P::ns(ST), NB = L / R, B = M / R;
#pragma unroll
  for (int u = 0; u < B; ++u) {
    if constexpr (NS > 1) {
      int line, j;
      bf_index<LINES, NB, LFAST>(tid + u * T, line, j);
      const int k = j % NS;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[r * k * (L / (NS * R))];
        if constexpr (S > 0) w.y = -w.y;
        v[u * R + r] = cmul(v[u * R + r], w);
      }
    }
    dftr<R, S>(&v[u * R]);
  }
}

// Stage ST's outputs → LDS, then stage ST+1's inputs ← LDS.
template <int L, int M, int T, int LINES, bool LFAST, int ST>
__device__ __forceinline__ void rexchange(float2 (&v)[M], float2* lds, int stride, int tid) {
  using P = RPlan<L, M>;
  constexpr int R = P::radix(ST), NS = P::ns(ST), NB = L / R, B = M / R;
  constexpr int R2 = P::radix(ST + 1), NB2 = L / R2, B2 = M / R2;
#pragma unroll
  for (int u = 0; u < B; ++u) {
    int line, j;
    bf_index<LINES, NB, LFAST>(tid + u * T, line, j);
    const int d = (j / NS) * NS * R + (j % NS);
    float2* dst = lds + line * stride;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[pad(d + r * NS)] = v[u * R + r];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < B2; ++u) {
    int line, j;
    bf_index<LINES, NB2, LFAST>(tid + u * T, line, j);
    const float2* src = lds + line * stride;
#pragma unroll
    for (int r = 0; r < R2; ++r) v[u * R2 + r] = src[pad(j + r * NB2)];
  }
  __syncthreads();
}

template <int L, int M, int T, int LINES, bool LFAST, int S, int ST>
__device__ __forceinline__ void rstages_from(float2 (&v)[M], float2* lds, int stride, const float2* __restrict__ tw,
                                             int tid) {
  rstage_compute<L, M, T, LINES, LFAST, S, ST>(v, tw, tid);
  if constexpr (ST + 1 < RPlan<L, M>::stages) {
    rexchange<L, M, T, LINES, LFAST, ST>(v, lds, stride, tid);
    rstages_from<L, M, T, LINES, LFAST, S, ST + 1>(v, lds, stride, tw, tid);
  }
}

// The whole line FFT on register data v (stage-0 inputs, see load positions below) to the last
// stage's outputs (see store positions). LDS: LINES lines of `stride` complex (>= padded_len(L)).
template <int L, int M, int T, int LINES, bool LFAST, int S>
__device__ __forceinline__ void rfft(float2 (&v)[M], float2* lds, int stride, const float2* __restrict__ tw, int tid) {
  rstages_from<L, M, T, LINES, LFAST, S, 0>(v, lds, stride, tw, tid);
}

// The same exchange through half the LDS: the real parts, then the imaginary parts, each through
// one float buffer of LINES lines of `stride` (four barriers instead of two), so that two
// workgroups of a wide tiling fit on one CU and overlap each other's exchanges with memory.
template <int L, int M, int T, int LINES, bool LFAST, int ST>
__device__ __forceinline__ void rexchange_split(float2 (&v)[M], float* lds, int stride, int tid) {
  using P = RPlan<L, M>;
  constexpr int R = P::radix(ST), NS = P::ns(ST), NB = L / R, B = M / R;
  constexpr int R2 = P::radix(ST + 1), NB2 = L / R2, B2 = M / R2;
#pragma unroll
  for (int part = 0; part < 2; ++part) {
#pragma unroll
    for (int u = 0; u < B; ++u) {
      int line, j;
      bf_index<LINES, NB, LFAST>(tid + u * T, line, j);
      const int d = (j / NS) * NS * R + (j % NS);
      float* dst = lds + line * stride;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[pad(d + r * NS)] = part ? v[u * R + r].y : v[u * R + r].x;
    }
    __syncthreads();
    float w[M];
#pragma unroll
    for (int u = 0; u < B2; ++u) {
      int line, j;
      bf_index<LINES, NB2, LFAST>(tid + u * T, line, j);
      const float* src = lds + line * stride;
#pragma unroll
      for (int r = 0; r < R2; ++r) w[u * R2 + r] = src[pad(j + r * NB2)];
    }
    __syncthreads();
    // the next stage's layout replaces this stage's once both parts are read
    if (part == 0) {
#pragma unroll
      for (int i = 0; i < M; ++i) v[i].x = w[i];
    } else {
#pragma unroll
      for (int i = 0; i < M; ++i) v[i].y = w[i];
    }
  }
}

template <int L, int M, int T, int LINES, bool LFAST, int S, int ST>
__device__ __forceinline__ void rstages_split_from(float2 (&v)[M], float* lds, int stride,
                                                   const float2* __restrict__ tw, int tid) {
  rstage_compute<L, M, T, LINES, LFAST, S, ST>(v, tw, tid);
  if constexpr (ST + 1 < RPlan<L, M>::stages) {
    rexchange_split<L, M, T, LINES, LFAST, ST>(v, lds, stride, tid);
    rstages_split_from<L, M, T, LINES, LFAST, S, ST + 1>(v, lds, stride, tw, tid);
  }
}

template <int L, int M, int T, int LINES, bool LFAST, int S>
__device__ __forceinline__ void rfft_split(float2 (&v)[M], float* lds, int stride, const float2* __restrict__ tw,
                                           int tid) {
  rstages_split_from<L, M, T, LINES, LFAST, S, 0>(v, lds, stride, tw, tid);
}

// G column groups of LINES lines each, time-multiplexing one LDS buffer of LINES lines: every
// stage computes all groups, then exchanges group 0, group 1, ... through the same LDS (two
// barriers per group and exchange). Lets a workgroup own G·LINES adjacent columns (wider row
// segments in memory) with the LDS of LINES columns; its points stay in registers.
template <int L, int M, int T, int LINES, bool LFAST, int S, int G, int ST>
__device__ __forceinline__ void rstages_groups_from(float2 (&v)[G][M], float2* lds, int stride,
                                                    const float2* __restrict__ tw, int tid) {
#pragma unroll
  for (int g = 0; g < G; ++g) rstage_compute<L, M, T, LINES, LFAST, S, ST>(v[g], tw, tid);
  if constexpr (ST + 1 < RPlan<L, M>::stages) {
#pragma unroll
    for (int g = 0; g < G; ++g) rexchange<L, M, T, LINES, LFAST, ST>(v[g], lds, stride, tid);
    rstages_groups_from<L, M, T, LINES, LFAST, S, G, ST + 1>(v, lds, stride, tw, tid);
  }
}

template <int L, int M, int T, int LINES, bool LFAST, int S, int G>
__device__ __forceinline__ void rfft_groups(float2 (&v)[G][M], float2* lds, int stride, const float2* __restrict__ tw,
                                            int tid) {
  rstages_groups_from<L, M, T, LINES, LFAST, S, G, 0>(v, lds, stride, tw, tid);
}

// Stage-0 input positions: for u < M/R0, r < R0: (line, j + r·L/R0) → v[u·R0 + r].
template <int L, int M, int T, int LINES, bool LFAST, typename F>
__device__ __forceinline__ void rload(float2 (&v)[M], int tid, F&& f) {
  using P = RPlan<L, M>;
  constexpr int R = P::radix(0), NB = L / R, B = M / R;
#pragma unroll
  for (int u = 0; u < B; ++u) {
    int line, j;
    bf_index<LINES, NB, LFAST>(tid + u * T, line, j);
#pragma unroll
    for (int r = 0; r < R; ++r) v[u * R + r] = f(line, j + r * NB);
  }
}

// Last-stage output positions: v[u·R + r] is point (line, j + r·L/R) of the transform.
template <int L, int M, int T, int LINES, bool LFAST, typename F>
__device__ __forceinline__ void rstore(const float2 (&v)[M], int tid, F&& f) {
  using P = RPlan<L, M>;
  constexpr int R = P::radix(P::stages - 1), NB = L / R, B = M / R;
#pragma unroll
  for (int u = 0; u < B; ++u) {
    int line, j;
    bf_index<LINES, NB, LFAST>(tid + u * T, line, j);
#pragma unroll
    for (int r = 0; r < R; ++r) f(line, j + r * NB, v[u * R + r]);
  }
}

// LDS stride (complex) for LINES lines: ≥ padded_len(L), ≡ 64 / LINES (mod 16) so that the
// line-fastest accesses of a wave spread over the banks.
__host__ __device__ constexpr int line_stride(int L, int LINES) {
  const int p = padded_len(L);
  const int want = LINES >= 16 ? 4 : (LINES >= 8 ? 8 : 0);
  return LINES == 1 ? p : p + ((want - p % 16) % 16 + 16) % 16;
}

// Block index remapped so that consecutive logical blocks run on the same XCD (the hardware
// deals blocks round-robin over the 8 XCDs): neighbours share their L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  if (nb % 8 != 0) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

}  // namespace zc
