// Rotation × scale sweep kernels.
#include "kdev.hpp"

namespace zc {

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
// Two waves per SIMD (at most 256 registers per lane): the five sums of the M cells stay live
// through every spectrum's transform, and at 2304 and 3072 points the plans with 15 real planes
// needed 256 VGPRs + AGPRs, one wave per SIMD. Their few spills cost less than the lost
// occupancy (full-resolution sweep: pack 180 → 129 ms per sweep).
template <int L, int PID>
__global__ __launch_bounds__(RowT<L>::T, 2) void k_pack_sweep(const float4* __restrict__ il,
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
    // the twiddle pointer as an opaque value per spectrum: otherwise the twiddle loads of every
    // transform are merged and held in registers across all of them
    const float2* twc = tw;
    asm volatile("" : "+s"(twc)::"memory");
    float2 v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const float re = real_plane(P.slot_plane[2 * c], s[i]);
      const float im = (2 * c + 1 < P.n_real) ? real_plane(P.slot_plane[2 * c + 1], s[i]) : 0.0f;
      v[i] = make_float2(re, im);
    }
    rfft<L, M, T, LINES, false, -1>(v, lds, STR, twc, tid);
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
// Periodic neighbours wrap at nx and ny (any lengths; maps stay below 2^31 lags).
__global__ __launch_bounds__(256) void k_peaks_part(const float* __restrict__ ns, int nx, int ny, int64_t chunk,
                                                    float2* part) {
  const int b = blockIdx.y, lid = threadIdx.x;
  const int64_t nn = (int64_t)nx * ny;
  const float* m = ns + (int64_t)b * nn;
  const int64_t t0 = (int64_t)blockIdx.x * chunk, t1 = min(nn, t0 + chunk);
  float bv = 0.0f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = t0 + lid; t < t1; t += 256) {
    const float v = m[t];
    if (v <= bv) continue;
    const int y = (int)((uint32_t)t / (uint32_t)nx), x = (int)((uint32_t)t - (uint32_t)y * (uint32_t)nx);
    const int xs[3] = {x == 0 ? nx - 1 : x - 1, x, x == nx - 1 ? 0 : x + 1};
    const int ys[3] = {y == 0 ? ny - 1 : y - 1, y, y == ny - 1 ? 0 : y + 1};
    bool ok = true;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int64_t row = (int64_t)ys[dy] * nx;
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        if (dx == 1 && dy == 1) continue;
        ok &= !(m[row + xs[dx]] > v);
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
  ZC_FFT_SWITCH(nx, CALL)
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
  k_peaks_part<<<dim3(chunks, batch), 256, 0, st>>>(ns, nx, ny, chunk, part);
  k_peaks_final<<<batch, 256, 0, st>>>(part, chunks, n_peak, peaks);
}


}  // namespace zc
