// Batched power-of-two FFT building blocks for gfx950: Stockham autosort stages over one line
// (or several lines side by side) held in LDS, radix 8 with a radix-4 or radix-2·4 tail.
//
// Each stage reads R points at stride L/R, multiplies by the stage twiddles (a per-length
// table e^{-2πi t/L} in global memory, read through the vector L1), performs the radix-R DFT in
// registers and writes the R outputs at stride Ns. One LDS buffer serves all stages: a thread
// holds every point it will write in registers across the barrier that separates the reads
// from the writes. LDS addresses are padded by one complex per 32 to spread the strided
// early-stage writes over the banks.
#pragma once
#include <hip/hip_runtime.h>

namespace zc {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a · conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// a · (s i), s = ±1
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// padded LDS index: one complex of padding per 32
__device__ __forceinline__ constexpr int pad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int padded_len(int L) { return L + (L >> 5); }

// In-register DFTs with sign S (-1 forward, +1 inverse).
template <int S>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S>
__device__ __forceinline__ void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
  float2 a = cadd(x0, x2), b = csub(x0, x2);
  float2 c = cadd(x1, x3), d = mul_si<S>(csub(x1, x3));
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, d);
  x3 = csub(b, d);
}

template <int S>
__device__ __forceinline__ void dft4v(float2* v) { dft4<S>(v[0], v[1], v[2], v[3]); }

// DFT3 with w = e^{S 2πi/3} = −1/2 + S i √3/2: X1 = a − (b + c)/2 + S i (√3/2)(b − c), X2 its
// conjugate-rotated twin.
template <int S>
__device__ __forceinline__ void dft3(float2& a, float2& b, float2& c) {
  constexpr float H = 0.86602540378443864676f;   // √3/2
  const float2 t1 = cadd(b, c);
  const float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
  const float2 d = mul_si<S>(make_float2(H * (b.x - c.x), H * (b.y - c.y)));
  a = cadd(a, t1);
  b = cadd(t2, d);
  c = csub(t2, d);
}

template <int S>
__device__ __forceinline__ void dft3v(float2* v) { dft3<S>(v[0], v[1], v[2]); }

template <int S>
__device__ __forceinline__ void dft8(float2* v) {
  constexpr float C = 0.70710678118654752440f;
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<S>(e0, e1, e2, e3);
  dft4<S>(o0, o1, o2, o3);
  // w8^k = e^{S 2πi k / 8}
  float2 t1 = make_float2(C * (o1.x - S * o1.y), C * (o1.y + S * o1.x));    // o1 (1 + S i)/√2
  float2 t2 = mul_si<S>(o2);                                                // o2 (S i)
  float2 t3 = make_float2(C * (-o3.x - S * o3.y), C * (-o3.y + S * o3.x));  // o3 (-1 + S i)/√2
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1);
  v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2);
  v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3);
  v[7] = csub(e3, t3);
}

template <int R, int S>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 8) dft8<S>(v);
  else if constexpr (R == 4) dft4v<S>(v);
  else dft2<S>(v);
}

// Radix of stage `st` of a length-L plan: radix 8 first, then a tail of one 4 (2^(3a+2)),
// or two 4s for 2^(3a+1) (an 8 and a 2 become 4 · 4), or a lone 2 for L = 2.
template <int L>
struct Plan {
  static constexpr int log2(int n) { return n <= 1 ? 0 : 1 + log2(n >> 1); }
  static constexpr int M = log2(L);
  static constexpr int n8 = (M % 3 == 1 && M > 1) ? M / 3 - 1 : M / 3;
  static constexpr int n4 = (M % 3 == 1 && M > 1) ? 2 : (M % 3 == 2 ? 1 : 0);
  static constexpr int n2 = (M == 1) ? 1 : 0;
  static constexpr int stages = n8 + n4 + n2;
  static constexpr int radix(int st) { return st < n8 ? 8 : (st < n8 + n4 ? 4 : 2); }
  static constexpr int ns(int st) { return st == 0 ? 1 : ns(st - 1) * radix(st - 1); }
};

// One Stockham stage over `lines` lines of length L stored at lds + line * stride (padded
// indices), T threads. tw: e^{-2πi t/L} for t < L.
template <int L, int R, int NS, int S, int T, int LINES>
__device__ __forceinline__ void stage(float2* lds, int stride, const float2* __restrict__ tw, int tid) {
  constexpr int NB = L / R;                          // butterflies per line
  constexpr int TOT = NB * LINES;                    // butterflies per stage
  constexpr int PER = (TOT + T - 1) / T;             // per thread
  float2 v[PER][R];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = tid + u * T;
    if (TOT % T == 0 || b < TOT) {
      const int line = b / NB, j = b - line * NB;
      const float2* src = lds + line * stride;
#pragma unroll
      for (int r = 0; r < R; ++r) v[u][r] = src[pad(j + r * NB)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = tid + u * T;
    if (TOT % T == 0 || b < TOT) {
      const int line = b / NB, j = b - line * NB;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * (L / (NS * R))];
          if constexpr (S > 0) w.y = -w.y;
          v[u][r] = cmul(v[u][r], w);
        }
      }
      dft<R, S>(v[u]);
      const int d = (j / NS) * NS * R + k;
      float2* dst = lds + line * stride;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[pad(d + r * NS)] = v[u][r];
    }
  }
  __syncthreads();
}

template <int L, int S, int T, int LINES, int ST>
__device__ __forceinline__ void stages_from(float2* lds, int stride, const float2* __restrict__ tw, int tid) {
  if constexpr (ST < Plan<L>::stages) {
    stage<L, Plan<L>::radix(ST), Plan<L>::ns(ST), S, T, LINES>(lds, stride, tw, tid);
    stages_from<L, S, T, LINES, ST + 1>(lds, stride, tw, tid);
  }
}

// Full in-LDS FFT of LINES lines of length L (natural order in and out). Callers must
// __syncthreads() after filling the LDS; the routine ends with a barrier.
template <int L, int S, int T, int LINES>
__device__ __forceinline__ void fft_lds(float2* lds, int stride, const float2* __restrict__ tw, int tid) {
  stages_from<L, S, T, LINES, 0>(lds, stride, tw, tid);
}

}  // namespace zc
