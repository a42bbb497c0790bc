// Map kernels, forward side: fused resample + real-plane pack + row FFT, the log-polar pack and
// its tables, and the Hermitian-unpacking cross spectrum.
#include "kdev.hpp"

namespace zc {

template <int L, int TI>
__global__ __launch_bounds__(RowT<L>::T) void k_pack_rowfft(const float* __restrict__ canvas, Resample r, int ny,
                                                            float2* spec, const float2* __restrict__ tw) {
  constexpr PackTable P = pack_table(TI);
  constexpr int T = RowT<L>::T, M = RowT<L>::M, LINES = RowT<L>::LINES, STR = RowT<L>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * ny;
  float s[M][5];
  {
    float2 dummy[M];
    int idx = 0;
    rload<L, M, T, LINES, false>(dummy, tid, [&](int line, int x) {
      const int y = y0 + line;
      if (x < r.cw && y < r.ch) cell_values(canvas, r, x, y, s[idx]);
      else {
#pragma unroll
        for (int p = 0; p < 5; ++p) s[idx][p] = 0.0f;
      }
      ++idx;
      return make_float2(0.0f, 0.0f);
    });
  }
#pragma unroll
  for (int c = 0; c < P.n_cplx; ++c) {
    float2 v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const float re = real_plane(P.slot_plane[2 * c], s[i]);
      const float im = (2 * c + 1 < P.n_real) ? real_plane(P.slot_plane[2 * c + 1], s[i]) : 0.0f;
      v[i] = make_float2(re, im);
    }
    rfft<L, M, T, LINES, false, -1>(v, lds, STR, tw, tid);
    float2* dst = spec + c * N;
    rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
      const int y = y0 + line;
      if (y < r.ch) dst[(int64_t)y * L + x] = val;
    });
    __syncthreads();   // the next plane's first exchange reuses the LDS
  }
}

// ── roto-scale maps: log-polar pack ───────────────────────────────────────────────────────
// Cell (x, y) of the log-polar grid samples the canvas stack at c + rho_y (cos th_x, sin th_x)
// (pixel i at i; 0 unless all four bilinear taps are inside: smi.wgsl:1418-1452), then every real
// plane is multiplied by rho_y (the area element, smi.wgsl:718-720), and the row FFT follows.
// Rows y < n_lam hold data; cells x >= n_th are 0 (a padded angle axis).
__device__ __forceinline__ void logpolar_values(const LogPolarArgs& a, int x, int y, float* s) {
  const double rho = a.rho[y];
  const double xs = a.cx + rho * a.cth[x], ys = a.cy + rho * a.sth[x];
  const double x0d = floor(xs), y0d = floor(ys);
  const int x0 = (int)x0d, y0 = (int)y0d;
  if (!(xs >= 0.0 && ys >= 0.0) || x0 + 1 >= a.W || y0 + 1 >= a.H) {
#pragma unroll
    for (int p = 0; p < 5; ++p) s[p] = 0.0f;
    return;
  }
  const float fx = (float)(xs - x0d), fy = (float)(ys - y0d);
  const int64_t n = (int64_t)a.W * a.H, o = (int64_t)y0 * a.W + x0;
#pragma unroll
  for (int p = 0; p < 5; ++p) {
    const float* c = a.canvas + p * n + o;
    const float v0 = c[0] * (1.0f - fx) + c[1] * fx;
    const float v1 = c[a.W] * (1.0f - fx) + c[a.W + 1] * fx;
    s[p] = v0 * (1.0f - fy) + v1 * fy;
  }
}

template <int L, int PID>
__global__ __launch_bounds__(RowT<L>::T) void k_pack_logpolar(LogPolarArgs a, int ny, float2* spec,
                                                              const float2* __restrict__ tw) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int T = RowT<L>::T, M = RowT<L>::M, LINES = RowT<L>::LINES, STR = RowT<L>::STR;
  __shared__ float2 lds[LINES * STR];
  const int tid = threadIdx.x;
  const int y0 = blockIdx.x * LINES;
  const int64_t N = (int64_t)L * ny;
  float s[M][5];
  float om[M];
  {
    float2 dummy[M];
    int idx = 0;
    rload<L, M, T, LINES, false>(dummy, tid, [&](int line, int x) {
      const int y = y0 + line;
      if (x < a.n_th && y < a.n_lam) {
        logpolar_values(a, x, y, s[idx]);
        om[idx] = (float)a.rho[y];
      } else {
#pragma unroll
        for (int p = 0; p < 5; ++p) s[idx][p] = 0.0f;
        om[idx] = 0.0f;
      }
      ++idx;
      return make_float2(0.0f, 0.0f);
    });
  }
#pragma unroll
  for (int c = 0; c < P.n_cplx; ++c) {
    float2 v[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const float re = real_plane(P.slot_plane[2 * c], s[i]) * om[i];
      const float im = (2 * c + 1 < P.n_real) ? real_plane(P.slot_plane[2 * c + 1], s[i]) * om[i] : 0.0f;
      v[i] = make_float2(re, im);
    }
    rfft<L, M, T, LINES, false, -1>(v, lds, STR, tw, tid);
    float2* dst = spec + c * N;
    rstore<L, M, T, LINES, false>(v, tid, [&](int line, int x, float2 val) {
      const int y = y0 + line;
      if (y < a.n_lam) dst[(int64_t)y * L + x] = val;
    });
    __syncthreads();
  }
}

// rho_y = exp(log0 + (y + 1/2) / n_lam · span), (cos, sin) of th_x = (x + 1/2) / n_th · 2 pi, FP64.
__global__ __launch_bounds__(256) void k_rs_tables(int n_th, int n_lam, double log0, double span, double* rho,
                                                   double* cth, double* sth) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n_lam) rho[i] = exp(log0 + ((double)i + 0.5) / (double)n_lam * span);
  if (i < n_th) {
    const double th = ((double)i + 0.5) / (double)n_th * 6.283185307179586476925;
    cth[i] = cos(th);
    sth[i] = sin(th);
  }
}

// ── cross spectrum ────────────────────────────────────────────────────────────────────────
// Spectrum of real slot `odd ? 2p+1 : 2p` at bin k from Z = FFT(x_2p + i x_2p+1) at k and −k,
// times 2 (the ½ is folded into the output scale): even: Z + conj Z₋, odd: (Z − conj Z₋)/i.
__device__ __forceinline__ float2 unpack2(float2 z, float2 zm, bool odd) {
  return odd ? make_float2(z.y + zm.y, zm.x - z.x) : make_float2(z.x + zm.x, z.y - zm.y);
}

// Batch entries blockIdx.y · nb .. + nb (at most `batch`) in turn per thread: the fixed spectra B,
// shared by every entry, are read once per thread and kept in registers.
template <int PID>
__global__ __launch_bounds__(256) void k_cmul(const float2* __restrict__ A, const float2* __restrict__ B,
                                              float2* work, int nx, int ny, float scale, int nb, int batch) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr int NC = P.n_cplx, NCB = P.b_n_cplx;
  const int64_t N = (int64_t)nx * ny;
  const int64_t i = blockIdx.x * 256 + threadIdx.x;
  const int64_t half_rows = ny / 2 + 1;
  if (i >= half_rows * nx) return;
  const int ky = (int)(i / nx), kx = (int)(i - (int64_t)ky * nx);
  const int my = (ny - ky) % ny, mx = (nx - kx) % nx;
  if ((ky == 0 || 2 * ky == ny) && kx > nx / 2) return;   // its mirror bin covers it
  const int64_t k = (int64_t)ky * nx + kx, km = (int64_t)my * nx + mx;
  const bool self = k == km;
  float2 zb[NCB], zbm[NCB];
#pragma unroll
  for (int c = 0; c < NCB; ++c) {
    zb[c] = B[c * N + k];
    zbm[c] = B[c * N + km];
  }
  const float sc = 0.25f * scale;
  const int e1 = min(batch, ((int)blockIdx.y + 1) * nb);
  for (int e = (int)blockIdx.y * nb; e < e1; ++e) {
    // entry e: its NC spectra in A, its n_out outputs in work
    const float2* Ae = A + (int64_t)e * NC * N;
    float2* we = work + (int64_t)e * P.n_out * N;
    float2 za[NC], zam[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      za[c] = Ae[c * N + k];
      zam[c] = Ae[c * N + km];
    }
#pragma unroll
    for (int q = 0; q < P.n_out; ++q) {
      const int c1 = 2 * q, c2 = 2 * q + 1;
      const int a1 = P.ca[c1], b1 = P.cb[c1];
      float2 xa = unpack2(za[a1 >> 1], zam[a1 >> 1], a1 & 1), xb = unpack2(zb[b1 >> 1], zbm[b1 >> 1], b1 & 1);
      float2 ya = unpack2(zam[a1 >> 1], za[a1 >> 1], a1 & 1), yb = unpack2(zbm[b1 >> 1], zb[b1 >> 1], b1 & 1);
      float2 p = cmulc(xa, xb), pm = cmulc(ya, yb);
      if (c2 < P.n_corr) {
        const int a2 = P.ca[c2], b2 = P.cb[c2];
        const float2 xa2 = unpack2(za[a2 >> 1], zam[a2 >> 1], a2 & 1), xb2 = unpack2(zb[b2 >> 1], zbm[b2 >> 1], b2 & 1);
        const float2 ya2 = unpack2(zam[a2 >> 1], za[a2 >> 1], a2 & 1), yb2 = unpack2(zbm[b2 >> 1], zb[b2 >> 1], b2 & 1);
        const float2 p2 = cmulc(xa2, xb2), pm2 = cmulc(ya2, yb2);
        p = make_float2(p.x - p2.y, p.y + p2.x);       // p + i p2
        pm = make_float2(pm.x - pm2.y, pm.y + pm2.x);
      }
      we[q * N + k] = make_float2(p.x * sc, p.y * sc);
      if (!self) we[q * N + km] = make_float2(pm.x * sc, pm.y * sc);
    }
  }
}


template <int L>
static void pack_rowfft_L(int plan, int side, const float* canvas, const Resample& r, int ny, float2* spec,
                          const float2* tw, hipStream_t st) {
  dim3 g((r.ch + RowT<L>::LINES - 1) / RowT<L>::LINES);
  // plans without rectangle sums share one table for both sides
  const int ti = plan < 4 ? 2 * plan : 2 * plan + side;
#define K(TAB) k_pack_rowfft<L, TAB><<<g, RowT<L>::T, 0, st>>>(canvas, r, ny, spec, tw)
  switch (ti) {
    case 0: K(0); break;
    case 2: K(2); break;
    case 4: K(4); break;
    case 6: K(6); break;
    case 8: K(8); break;
    case 9: K(9); break;
    case 10: K(10); break;
    case 11: K(11); break;
    case 12: K(12); break;
    case 13: K(13); break;
    case 14: K(14); break;
    case 15: K(15); break;
    case 16: K(16); break;
    case 17: K(17); break;
    case 18: K(18); break;
    case 19: K(19); break;
    case 20: K(20); break;
    case 21: K(21); break;
    case 22: K(22); break;
    default: K(23); break;
  }
#undef K
}


// Resampled maps: every cell's five values (k×k bilinear taps each), one thread per cell, so the
// taps are gathered by the whole chip rather than by a pack launch of ch · nx / 16 threads.
__global__ __launch_bounds__(256) void k_resample_cells(const float* __restrict__ canvas, Resample r, float* cells) {
  const int64_t nc = (int64_t)r.cw * r.ch;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nc) return;
  float s[5];
  cell_values(canvas, r, (int)(t % r.cw), (int)(t / r.cw), s);
#pragma unroll
  for (int p = 0; p < 5; ++p) cells[p * nc + t] = s[p];
}

// The same with a workgroup's 16 × 16 cells reading their canvas footprint from LDS (each canvas
// pixel loaded once per tile instead of by every tap that touches it): identical arithmetic.
constexpr int kResTile = 16, kResTilePx = 2048;
struct ResFoot {
  int x0, y0, w, h;   // the tile's canvas footprint (clamped pixel indices)
};
__host__ __device__ __forceinline__ ResFoot res_footprint(const Resample& r, int cx0, int cy0) {
  const int cx1 = (cx0 + kResTile < r.cw ? cx0 + kResTile : r.cw) - 1;
  const int cy1 = (cy0 + kResTile < r.ch ? cy0 + kResTile : r.ch) - 1;
  const float pxl = ((float)cx0 + 0.5f / (float)r.kx) * r.fx;
  const float pxh = ((float)cx1 + ((float)(r.kx - 1) + 0.5f) / (float)r.kx) * r.fx;
  const float pyl = ((float)cy0 + 0.5f / (float)r.ky) * r.fy;
  const float pyh = ((float)cy1 + ((float)(r.ky - 1) + 0.5f) / (float)r.ky) * r.fy;
  auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const int xl = cl((int)floorf(pxl - 0.5f), r.W - 1), xh = cl((int)floorf(pxh - 0.5f) + 1, r.W - 1);
  const int yl = cl((int)floorf(pyl - 0.5f), r.H - 1), yh = cl((int)floorf(pyh - 0.5f) + 1, r.H - 1);
  return {xl, yl, xh - xl + 1, yh - yl + 1};
}

__global__ __launch_bounds__(kResTile * kResTile) void k_resample_tile(const float* __restrict__ canvas, Resample r,
                                                                       float* cells) {
  __shared__ float tile[5 * kResTilePx];
  const int cx0 = blockIdx.x * kResTile, cy0 = blockIdx.y * kResTile;
  const ResFoot f = res_footprint(r, cx0, cy0);
  const int64_t n = (int64_t)r.W * r.H, fp = (int64_t)f.w * f.h;
  const int x = cx0 + (threadIdx.x % kResTile), y = cy0 + (threadIdx.x / kResTile);
  const int64_t nc = (int64_t)r.cw * r.ch, t = (int64_t)y * r.cw + x;
  float s[5];
  if (fp > kResTilePx) {   // a footprint larger than the tile (not launched so): from the canvas
    if (x >= r.cw || y >= r.ch) return;
    cell_values(canvas, r, x, y, s);
#pragma unroll
    for (int p = 0; p < 5; ++p) cells[p * nc + t] = s[p];
    return;
  }
  for (int q = threadIdx.x; q < 5 * fp; q += kResTile * kResTile) {
    const int p = q / (int)fp, rem = q - p * (int)fp, yy = rem / f.w, xx = rem - yy * f.w;
    tile[q] = canvas[p * n + (int64_t)(f.y0 + yy) * r.W + f.x0 + xx];
  }
  __syncthreads();
  if (x >= r.cw || y >= r.ch) return;
#pragma unroll
  for (int p = 0; p < 5; ++p) s[p] = 0.0f;
  for (int j = 0; j < r.ky; ++j) {
    const float py = ((float)y + ((float)j + 0.5f) / (float)r.ky) * r.fy;
    for (int i = 0; i < r.kx; ++i) {
      const float px = ((float)x + ((float)i + 0.5f) / (float)r.kx) * r.fx;
      // sample_plane's taps, read from the tile
      const float xs = px - 0.5f, ys = py - 0.5f;
      const float x0f = floorf(xs), y0f = floorf(ys);
      const float ax = xs - x0f, ay = ys - y0f;
      const int ix = (int)x0f, iy = (int)y0f;
      const int xa = min(max(ix, 0), r.W - 1) - f.x0, xb = min(max(ix + 1, 0), r.W - 1) - f.x0;
      const int ya = min(max(iy, 0), r.H - 1) - f.y0, yb = min(max(iy + 1, 0), r.H - 1) - f.y0;
#pragma unroll
      for (int p = 0; p < 5; ++p) {
        const float* sp = tile + p * fp;
        const float a = sp[ya * f.w + xa], b = sp[ya * f.w + xb];
        const float c = sp[yb * f.w + xa], d = sp[yb * f.w + xb];
        s[p] += (a * (1.0f - ax) + b * ax) * (1.0f - ay) + (c * (1.0f - ax) + d * ax) * ay;
      }
    }
  }
  const float inv = 1.0f / (float)(r.kx * r.ky);
#pragma unroll
  for (int p = 0; p < 5; ++p) cells[p * nc + t] = s[p] * inv;
}

// Large cells (k×k >= 16 taps): one wave per cell, one lane per tap (coalesced gathers); the tap
// values meet in LDS and lane p adds plane p's in cell_values' order, so the sums are identical.
__global__ __launch_bounds__(256) void k_resample_wave(const float* __restrict__ canvas, Resample r, float* cells) {
  __shared__ float tv[4][5][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t nc = (int64_t)r.cw * r.ch, n = (int64_t)r.W * r.H;
  const int64_t cell = (int64_t)blockIdx.x * 4 + w;
  const int K = r.kx * r.ky;
  const bool live = cell < nc;
  const int x = live ? (int)(cell % r.cw) : 0, y = live ? (int)(cell / r.cw) : 0;
  if (live && lane < K) {
    const int i = lane % r.kx, j = lane / r.kx;
    const float py = ((float)y + ((float)j + 0.5f) / (float)r.ky) * r.fy;
    const float px = ((float)x + ((float)i + 0.5f) / (float)r.kx) * r.fx;
#pragma unroll
    for (int p = 0; p < 5; ++p) tv[w][p][lane] = sample_plane(canvas + p * n, r.W, r.H, px, py);
  }
  __syncthreads();
  if (live && lane < 5) {
    float acc = 0.0f;
    for (int q = 0; q < K; ++q) acc += tv[w][lane][q];
    cells[lane * nc + cell] = acc * (1.0f / (float)K);
  }
}

static void launch_resample(const float* canvas, const Resample& r, float* cells, hipStream_t st) {
  if (r.kx * r.ky >= 16) {
    k_resample_wave<<<(unsigned)(((int64_t)r.cw * r.ch + 3) / 4), 256, 0, st>>>(canvas, r, cells);
    return;
  }
  // the tile form when every tile's footprint fits its LDS (the largest is a full interior tile)
  const ResFoot f = res_footprint(r, 0, 0);
  const int fw = (int)ceilf(kResTile * r.fx) + 4, fh = (int)ceilf(kResTile * r.fy) + 4;
  if ((int64_t)std::max(fw, f.w) * std::max(fh, f.h) <= kResTilePx) {
    const dim3 g((r.cw + kResTile - 1) / kResTile, (r.ch + kResTile - 1) / kResTile);
    k_resample_tile<<<g, kResTile * kResTile, 0, st>>>(canvas, r, cells);
  } else {
    k_resample_cells<<<(unsigned)(((int64_t)r.cw * r.ch + 255) / 256), 256, 0, st>>>(canvas, r, cells);
  }
}

bool launch_pack_rowfft(int plan, int side, const float* canvas, int W, int H, int cw, int ch, int nx, int ny,
                        float2* spec, const float2* tw_x, hipStream_t st, float* cells) {
  Resample r;
  r.W = W;
  r.H = H;
  r.cw = cw;
  r.ch = ch;
  r.fx = (float)W / (float)(cw > 0 ? cw : 1);
  r.fy = (float)H / (float)(ch > 0 ? ch : 1);
  r.kx = (int)fminf(fmaxf(ceilf(r.fx - 1e-4f), 1.0f), 8.0f);
  r.ky = (int)fminf(fmaxf(ceilf(r.fy - 1e-4f), 1.0f), 8.0f);
  r.identity = (cw == W && ch == H) ? 1 : 0;
  r.sx = r.sy = 0;
  if (!r.identity && cells && cw > 0 && ch > 0) {
    launch_resample(canvas, r, cells, st);
    canvas = cells;
    r = Resample{cw, ch, cw, ch, 1, 1, 1, 1.0f, 1.0f, 0, 0};
  }
#define CALL(LL) pack_rowfft_L<LL>(plan, side, canvas, r, ny, spec, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

bool launch_pack_rowfft_shifted(int plan, int side, const float* stack, int sw, int sh, int sx, int sy, int cw,
                                int ch, int nx, int ny, float2* spec, const float2* tw_x, hipStream_t st) {
  const Resample r{sw, sh, cw, ch, 1, 1, 2, 1.0f, 1.0f, sx, sy};
#define CALL(LL) pack_rowfft_L<LL>(plan, side, stack, r, ny, spec, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

template <int L>
static void pack_logpolar_L(int plan, const LogPolarArgs& a, int ny, float2* spec, const float2* tw, hipStream_t st) {
  dim3 g((a.n_lam + RowT<L>::LINES - 1) / RowT<L>::LINES);
  switch (plan) {
    case 0: k_pack_logpolar<L, 0><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
    case 1: k_pack_logpolar<L, 1><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
    case 2: k_pack_logpolar<L, 2><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
    default: k_pack_logpolar<L, 3><<<g, RowT<L>::T, 0, st>>>(a, ny, spec, tw); break;
  }
}

bool launch_pack_logpolar(int plan, const LogPolarArgs& a, int nx, int ny, float2* spec, const float2* tw_x,
                          hipStream_t st) {
  if (plan < 0 || plan > 3 || a.n_th > nx || a.n_lam > ny) return false;
#define CALL(LL) pack_logpolar_L<LL>(plan, a, ny, spec, tw_x, st)
  ZC_POW2_SWITCH(nx, CALL)
#undef CALL
  return true;
}

void launch_rs_tables(int n_th, int n_lam, double log0, double span, double* rho, double* cth, double* sth,
                      hipStream_t st) {
  const int m = n_th > n_lam ? n_th : n_lam;
  k_rs_tables<<<(m + 255) / 256, 256, 0, st>>>(n_th, n_lam, log0, span, rho, cth, sth);
}

void launch_cmul(int plan, const float2* spec_a, const float2* spec_b, float2* work, int nx, int ny, hipStream_t st,
                 int batch) {
  const int64_t cnt = (int64_t)(ny / 2 + 1) * nx;
  const float scale = 1.0f / ((float)nx * (float)ny);
  // entries per thread: as many as leave at least 4096 workgroups (one per entry when the bins
  // alone give fewer)
  const unsigned gx = blocks_for(cnt, 256);
  const int rows = (int)std::min<int64_t>(batch, std::max<int64_t>(1, (4096 + gx - 1) / gx));
  const int nb = (batch + rows - 1) / rows;
  const dim3 g(gx, (unsigned)((batch + nb - 1) / nb));
  switch (plan) {
    case 0: k_cmul<0><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 1: k_cmul<1><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 2: k_cmul<2><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 3: k_cmul<3><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 4: k_cmul<4><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 5: k_cmul<5><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 6: k_cmul<6><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 7: k_cmul<7><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 8: k_cmul<8><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 9: k_cmul<9><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    case 10: k_cmul<10><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
    default: k_cmul<11><<<g, 256, 0, st>>>(spec_a, spec_b, work, nx, ny, scale, nb, batch); break;
  }
}


}  // namespace zc
