// Map kernels, score side: per-lag whitened scores (FFT and rectangle-sum plans, roto-scale),
// summed-area tables and the argmax.
#include "kdev.hpp"

namespace zc {

template <int PID>
__global__ __launch_bounds__(256) void k_combine(const float* __restrict__ corr, float* map, int64_t nn,
                                                 CombineArgs a, float* arg_part) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  auto C = [&](int num, int64_t t) -> double { return (double)corr[(int64_t)P.kept_of[num] * nn + t]; };
  const double n0 = C(0, 0);
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < nn; t += (int64_t)gridDim.x * 256) {
    const float sv = (float)lag_score<PID>(C, t, floor_n, a, calib);
    map[t] = sv;
    if (better(sv, (uint32_t)t, bv, bi)) {
      bv = sv;
      bi = (uint32_t)t;
    }
  }
  // block argmax
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

// ── rectangle sums ────────────────────────────────────────────────────────────────────────
// When the fixed side's weight plane is the indicator of the rectangle R (the fixed image pasted
// at full resolution), correlation (a, 0) at lag t is the sum of plane a over R + t clipped to the
// canvas: four reads of plane a's summed-area table S(y, x) = sum over rows < y, columns < x
// ((H+1) x (W+1) values per plane). Tables 2p and 2p+1 are stored interleaved as one double2 per
// cell (table pair p; a last odd table pairs with zeros), so one 16-byte load at a corner serves
// both. The tables are built in three passes over bands of kSatBand rows: every table's column
// sums per band (k_sat_band: one thread per column, all tables from one read of the canvas), their
// exclusive prefix over bands (k_sat_carry), then every (band, four tables) by a workgroup that
// scans its rows with a running column carry (k_sat_rows).
constexpr int kSatBand = 8;   // rows per band

// The real planes of a canvas stack sample in FP64: w, w f_i, and w f_i f_j = s_i s_j / s_0
// (0 where s_0 <= 1e-6), the last with one reciprocal per sample.
// pl is a compile-time constant at every call (the plan's table → plane map), so every index
// below folds and the sample stays in registers.
__device__ __forceinline__ double real_plane_d(int pl, const float* s, double inv0) {
  if (pl < 5) return (double)s[pl];
  if (s[0] <= 1e-6f) return 0.0;
  constexpr int TI[10] = {0, 0, 0, 0, 1, 1, 1, 2, 2, 3};
  constexpr int TJ[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
  const int k = pl - 5;
  return (double)s[1 + TI[k]] * (double)s[1 + TJ[k]] * inv0;
}

__device__ __forceinline__ double canvas_inv0(const float* s) { return s[0] > 1e-6f ? 1.0 / (double)s[0] : 0.0; }

__device__ __forceinline__ void canvas_values(const float* __restrict__ canvas, int64_t n, int64_t at, float* s) {
#pragma unroll
  for (int p = 0; p < 5; ++p) s[p] = canvas[p * n + at];
}

// blockIdx.x band, blockIdx.y a chunk of 256 columns: band[(k · nb + b) · W + x] = column sums of
// table k's plane over the band's rows, for every table.
template <int SID>
__global__ __launch_bounds__(256) void k_sat_band(const float* __restrict__ canvas, int W, int H, double* band) {
  constexpr SatSet P = sat_set(SID);
  constexpr int n_tab = P.n;
  const int b = blockIdx.x, nb = gridDim.x;
  const int x = blockIdx.y * 256 + threadIdx.x;
  if (x >= W) return;
  const int64_t n = (int64_t)W * H;
  const int y0 = b * kSatBand, y1 = min(H, y0 + kSatBand);
  double acc[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) acc[k] = 0.0;
  float sv[kSatBand][5];
#pragma unroll
  for (int i = 0; i < kSatBand; ++i) {   // every row's loads in flight together
    const int y = min(y0 + i, H - 1);
    canvas_values(canvas, n, (int64_t)y * W + x, sv[i]);
  }
#pragma unroll
  for (int i = 0; i < kSatBand; ++i) {
    if (y0 + i < y1) {
      const double inv0 = canvas_inv0(sv[i]);
#pragma unroll
      for (int k = 0; k < 15; ++k)
        if (k < n_tab) acc[k] += real_plane_d(P.plane[k], sv[i], inv0);
    }
  }
#pragma unroll
  for (int k = 0; k < 15; ++k)
    if (k < n_tab) band[((int64_t)k * nb + b) * W + x] = acc[k];
}

// In place: band sums → exclusive prefix over bands, per (table, column). blockIdx.x: table ·
// column chunks of 64; the block's 16 waves take a sixteenth of the bands each.
constexpr int kCarryWaves = 16;
__global__ __launch_bounds__(64 * kCarryWaves) void k_sat_carry(double* band, int W, int nb) {
  __shared__ double seg[kCarryWaves][64];
  const int chunks = (W + 63) / 64;
  const int k = blockIdx.x / chunks, x = (blockIdx.x - k * chunks) * 64 + (threadIdx.x & 63);
  const int q = threadIdx.x >> 6;
  const int per = (nb + kCarryWaves - 1) / kCarryWaves, b0 = q * per, b1 = min(nb, b0 + per);
  double* col = band + (int64_t)k * nb * W + x;
  double tot = 0.0;
  if (x < W)
    for (int b = b0; b < b1; ++b) tot += col[(int64_t)b * W];
  seg[q][threadIdx.x & 63] = tot;
  __syncthreads();
  double acc = 0.0;
  for (int j = 0; j < q; ++j) acc += seg[j][threadIdx.x & 63];
  if (x < W)
    for (int b = b0; b < b1; ++b) {
      const double v = col[(int64_t)b * W];
      col[(int64_t)b * W] = acc;
      acc += v;
    }
}

// Exclusive prefix of G values per thread over the threads of the block, each value its own
// scan (lds: G doubles per wave).
template <int G>
__device__ __forceinline__ void block_excl_scan(double (&v)[G], double* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double inc[G];
#pragma unroll
  for (int g = 0; g < G; ++g) inc[g] = v[g];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const double u = __shfl_up(inc[g], o, 64);
      if (lane >= o) inc[g] += u;
    }
  }
  if (lane == 63)
#pragma unroll
    for (int g = 0; g < G; ++g) lds[w * G + g] = inc[g];
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G; ++g) {
    double off = 0.0;
    for (int j = 0; j < w; ++j) off += lds[j * G + g];
    v[g] = off + inc[g] - v[g];
  }
  __syncthreads();   // lds is reused by the next scan
}

// blockIdx.x band, blockIdx.y a group of G tables (G / 2 table pairs; one read of the canvas for
// all of them); thread t owns columns [t·E, t·E + E) for the scans; each finished row goes out
// through LDS so that lane i stores column x + i of a pair (whole cache lines per store).
template <int SID, int NT, int E, int G, int K0>
__device__ __forceinline__ void sat_rows_body(const float* __restrict__ canvas, int W, int H,
                                              const double* __restrict__ band, double2* sat, double* lds,
                                              double (*stage)[NT * E]) {
  constexpr SatSet P = sat_set(SID);
  constexpr int n_tab = P.n, k0 = K0;
  const int b = blockIdx.x, nb = gridDim.x;
  const int xs = threadIdx.x * E;
  const int64_t n = (int64_t)W * H, SW = W + 1, SP = SW * (H + 1);
  constexpr int pl[G] = {P.plane[k0 < n_tab ? k0 : 0], P.plane[k0 + 1 < n_tab ? k0 + 1 : 0],
                         P.plane[k0 + 2 < n_tab ? k0 + 2 : 0], P.plane[k0 + 3 < n_tab ? k0 + 3 : 0]};
  static_assert(G == 4 && K0 % 2 == 0, "two table pairs per workgroup");
  // column carry cc[g][e] = S_g(y, xs + e + 1): the row prefix of the earlier bands' column sums
  double cc[G][E];
  {
    double tot[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      tot[g] = 0.0;
      const double* crow = band + ((int64_t)min(k0 + g, n_tab - 1) * nb + b) * W;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        tot[g] += xs + e < W ? crow[xs + e] : 0.0;
        cc[g][e] = tot[g];
      }
    }
    block_excl_scan<G>(tot, lds);
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int e = 0; e < E; ++e) cc[g][e] += tot[g];
  }
  if (b == 0)
#pragma unroll
    for (int h = 0; h < G / 2; ++h)
      if (k0 + 2 * h < n_tab)
        for (int x = threadIdx.x; x <= W; x += NT) sat[(k0 / 2 + h) * SP + x] = make_double2(0.0, 0.0);
  const int y0 = b * kSatBand, y1 = min(H, y0 + kSatBand);
  // the canvas values of the thread's columns, one row ahead of the scans (E <= 4: the registers
  // of a second row fit)
  constexpr bool AHEAD = E <= 4;
  float nxt[E][5];
#define ZC_SAT_LOAD_ROW(YY)                                                                         \
  _Pragma("unroll") for (int e = 0; e < E; ++e) {                                                    \
    _Pragma("unroll") for (int p = 0; p < 5; ++p) nxt[e][p] = 0.0f;                                  \
    if ((YY) < y1 && xs + e < W)                                                                     \
      _Pragma("unroll") for (int p = 0; p < 5; ++p) nxt[e][p] = canvas[p * n + (int64_t)(YY) * W + xs + e]; \
  }
  if (AHEAD) ZC_SAT_LOAD_ROW(y0)
  for (int y = y0; y < y1; ++y) {
    if (!AHEAD) ZC_SAT_LOAD_ROW(y)
    float cur[E][5];
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int p = 0; p < 5; ++p) cur[e][p] = nxt[e][p];
    if (AHEAD) ZC_SAT_LOAD_ROW(y + 1)
    double v[G][E];
    double tot[G];
#pragma unroll
    for (int g = 0; g < G; ++g) tot[g] = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float* sv = cur[e];
      const double inv0 = canvas_inv0(sv);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        tot[g] += real_plane_d(pl[g], sv, inv0);
        v[g][e] = tot[g];
      }
    }
    block_excl_scan<G>(tot, lds);   // its barriers also order the previous row's stage reads
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        cc[g][e] += tot[g] + v[g][e];
        stage[g][xs + e] = cc[g][e];
      }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < G / 2; ++h) {
      if (k0 + 2 * h >= n_tab) continue;
      const bool hi = k0 + 2 * h + 1 < n_tab;   // else the pair's second table is padding
      double2* row = sat + (k0 / 2 + h) * SP + (int64_t)(y + 1) * SW;
      for (int x = threadIdx.x; x < W; x += NT) row[x + 1] = make_double2(stage[2 * h][x], hi ? stage[2 * h + 1][x] : 0.0);
      if (threadIdx.x == 0) row[0] = make_double2(0.0, 0.0);
    }
  }
#undef ZC_SAT_LOAD_ROW
}

// blockIdx.y: the group of G tables (tables G·y .. G·y + G − 1 of the plan); NT threads of E
// columns each cover rows of up to NT·E cells.
template <int SID, int NT, int E, int G>
__global__ __launch_bounds__(NT) void k_sat_rows(const float* __restrict__ canvas, int W, int H,
                                                 const double* __restrict__ band, double2* sat) {
  __shared__ double lds[NT / 64 * G];
  __shared__ double stage[G][NT * E];
  static_assert(sat_set(SID).n <= 4 * G, "at most four groups");
  switch (blockIdx.y) {
    case 0: sat_rows_body<SID, NT, E, G, 0>(canvas, W, H, band, sat, lds, stage); break;
    case 1: sat_rows_body<SID, NT, E, G, G>(canvas, W, H, band, sat, lds, stage); break;
    case 2: sat_rows_body<SID, NT, E, G, 2 * G>(canvas, W, H, band, sat, lds, stage); break;
    default: sat_rows_body<SID, NT, E, G, 3 * G>(canvas, W, H, band, sat, lds, stage); break;
  }
}

int sat_bands(int H) { return (H + kSatBand - 1) / kSatBand; }

template <int SID>
static void sat_build_S(const float* canvas, int W, int H, double* band, double2* sat, hipStream_t st) {
  constexpr int n_tab = sat_set(SID).n;
  const int nb = sat_bands(H);
  k_sat_band<SID><<<dim3(nb, (W + 255) / 256), 256, 0, st>>>(canvas, W, H, band);
  k_sat_carry<<<(unsigned)(n_tab * ((W + 63) / 64)), 64 * kCarryWaves, 0, st>>>(band, W, nb);
  constexpr int G = 4;   // tables per workgroup (two pairs)
  const dim3 g(nb, (n_tab + G - 1) / G);
  if (W <= 256 * 4) k_sat_rows<SID, 256, 4, G><<<g, 256, 0, st>>>(canvas, W, H, band, sat);
  else k_sat_rows<SID, 512, 4, G><<<g, 512, 0, st>>>(canvas, W, H, band, sat);
}

bool launch_sat_build(int plan, int side, const float* canvas, int W, int H, double* band, double2* sat,
                      hipStream_t st) {
  if (W < 1 || H < 1 || W > 256 * 8) return false;
  // side A of a translation plan sums the same planes as its rectangle plan (plan − 4)
  const int sid = side == 2 ? kImageTables : side == 0 ? (plan >= 8 ? plan - 4 : plan) : 16 + plan;
  switch (sid) {
    case kImageTables: sat_build_S<kImageTables>(canvas, W, H, band, sat, st); return true;
    case 4: sat_build_S<4>(canvas, W, H, band, sat, st); return true;
    case 5: sat_build_S<5>(canvas, W, H, band, sat, st); return true;
    case 6: sat_build_S<6>(canvas, W, H, band, sat, st); return true;
    case 7: sat_build_S<7>(canvas, W, H, band, sat, st); return true;
    case 24: sat_build_S<24>(canvas, W, H, band, sat, st); return true;
    case 25: sat_build_S<25>(canvas, W, H, band, sat, st); return true;
    case 26: sat_build_S<26>(canvas, W, H, band, sat, st); return true;
    case 27: sat_build_S<27>(canvas, W, H, band, sat, st); return true;
    default: return false;
  }
}

// Table offsets (elements) of the rectangle R ± (dx, dy) clipped to the canvas for lag index t:
// S(y0, x0), S(y0, x1), S(y1, x0), S(y1, x1). The clamps keep x0 <= x1 and y0 <= y1, so an empty
// rectangle sums to exactly 0, and the tables' row 0 and column 0 hold zeros, so every corner is
// read unconditionally (no branches; the zero row and column stay in the caches).
struct SatIdx {
  uint32_t o00, o01, o10, o11;
  int area;   // (x1 − x0)(y1 − y0): the rectangle's cell count
};

__device__ __forceinline__ SatIdx sat_idx(const SatView& v, int64_t t) {
  const int px = (int)(t & (v.nx - 1)), py = (int)(t >> v.lx);
  const int dx = px <= v.nx / 2 ? px : px - v.nx;
  const int dy = py <= v.ny / 2 ? py : py - v.ny;
  const int sx = v.sign * dx, sy = v.sign * dy;
  const int x0 = min(max(v.bx0 + sx, v.cx0), v.cx1), x1 = min(max(v.bx1 + sx, v.cx0), v.cx1);
  const int y0 = min(max(v.by0 + sy, v.cy0), v.cy1), y1 = min(max(v.by1 + sy, v.cy0), v.cy1);
  const uint32_t SW = (uint32_t)v.stride;
  return {y0 * SW + x0, y0 * SW + x1, y1 * SW + x0, y1 * SW + x1, (x1 - x0) * (y1 - y0)};
}

// The sum over the rectangle from table k (pair k / 2, base = pair · table size in cells; every
// table set of a map stays below 2^32 bytes, so 32-bit cell offsets address it). The two tables of
// a pair load the same four cells, which the compiler merges into one load per corner.
__device__ __forceinline__ double sat_sum(const double2* __restrict__ S, int k, uint32_t SP, const SatIdx& q) {
  const uint32_t base = (uint32_t)(k >> 1) * SP;
  const double2 a = S[base + q.o11], b = S[base + q.o01], c = S[base + q.o10], d = S[base + q.o00];
  return (k & 1) ? (a.y - b.y) - (c.y - d.y) : (a.x - b.x) - (c.x - d.x);
}

// k_combine for the rectangle plans: correlations with sat_of >= 0 come from side A's tables, and
// in translation plans those with bsat_of >= 0 from side B's. A rectangle plan reads canvas A's
// tables (table k: the plan's k-th sum); a translation plan reads each image's own tables (table
// k: plane k), both at an offset (SatView).
//
// Lag order: a thread takes the four lags (px, py), (px + nx/2, py), (px, py + ny/2) and
// (px + nx/2, py + ny/2) one after the other (px < nx/2, py < ny/2). When the rectangles span half
// the plane (images of equal size on a plane twice as large) those four lags read the same table
// cells (a rectangle [d, h) and its complement [0, d) meet at row d, likewise for columns), so the
// cells of one lag are still in the caches for the next three instead of being fetched four times.
template <int PID>
__global__ __launch_bounds__(256) void k_combine_rect(const float4* __restrict__ corr4, float* map, int64_t nn,
                                                      CombineArgs a, SatView sv, SatView svb, float* arg_part) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  constexpr bool TRANS = P.n_bsat > 0;
  const uint32_t SP = sv.cells, SPB = svb.cells;
  SatIdx q = sat_idx(sv, 0), qb = q;
  if constexpr (TRANS) qb = sat_idx(svb, 0);
  // a translation plan's sums come from the image tables (plane p in table p − 1); the moving
  // image's weight plane is 1 on every pixel, so its sum over a rectangle is the area
  auto C = [&](int num, int64_t t) -> double {
    if constexpr (TRANS) {
      if (P.sat_of[num] >= 0) {
        const int pl = P.sat_plane[P.sat_of[num]];
        return pl == 0 ? (double)q.area : sat_sum(sv.sat, pl - 1, SP, q);
      }
      if (P.bsat_of[num] >= 0) return sat_sum(svb.sat, P.bsat_plane[P.bsat_of[num]] - 1, SPB, qb);
    } else {
      if (P.sat_of[num] >= 0) return sat_sum(sv.sat, P.sat_of[num], SP, q);
    }
    // the four correlations of a quad load together (k_rowifft_quads' layout)
    const int k = P.kept_of[num];
    const float4 v = corr4[(int64_t)(k >> 2) * nn + t];
    return (double)((k & 3) == 0 ? v.x : (k & 3) == 1 ? v.y : (k & 3) == 2 ? v.z : v.w);
  };
  const double n0 = C(0, 0);
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  auto one = [&](int64_t t) {
    q = sat_idx(sv, t);
    if constexpr (TRANS) qb = sat_idx(svb, t);
    const float s = (float)lag_score<PID & 3>(C, t, floor_n, a, calib);
    map[t] = s;
    if (better(s, (uint32_t)t, bv, bi)) {
      bv = s;
      bi = (uint32_t)t;
    }
  };
  const int hx = sv.nx >> 1;
  const int64_t nq = nn >> 2, hn = nn >> 1;
  for (int64_t u = blockIdx.x * 256 + threadIdx.x; u < nq; u += (int64_t)gridDim.x * 256) {
    const int64_t t = ((u / hx) << sv.lx) + (u & (hx - 1));
    one(t);
    one(t + hx);
    one(t + hn);
    one(t + hn + hx);
  }
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

// k_combine for translation plans whose two images have the same size and sit in the canvas so
// that every quad of lags (px, py), (px + nx/2, py), (px, py + ny/2), (px + nx/2, py + ny/2)
// splits each image at one cell (R, C): per axis one lag row of the quad sums rows [R, hi) and
// the other rows [0, R) (R = clamp(b + s·p), the axis' fixed rectangle edge moved by the lag),
// likewise for columns (quad_view checks this on the host for every row and column: the moving
// image takes the upper parts at lags >= 0, the fixed image the lower ones). Every rectangle sum
// of the four lags then comes from four cells per table pair: S(R, C) and S(hi, C) (loaded once
// per quad), S(R, hi) (one row per wave) and S(hi, hi) (one per table); the two uniform ones go
// through LDS. That is 2 loads per pair and quad instead of 16. Each lag's sums are formed as
// k_combine_rect forms them ((a − b) − (c − d) with the zero cells as 0), so the map is identical
// bit for bit. Lag row ny/2 is taken at −ny/2 (both are empty: quad_view checks).
__device__ __forceinline__ int quad_r(const QuadAxis& q, int p) { return min(max(q.b + q.s * p, q.c0), q.c1); }

template <int PID>
__global__ __launch_bounds__(256) void k_combine_quad(const float4* __restrict__ corr4, float* map, int64_t nn,
                                                      int lx, CombineArgs a, QuadView va, QuadView vb, double n0,
                                                      float* arg_part) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  static_assert(P.n_bsat > 0, "translation plans only");
  constexpr int NP = (kImageTableCount + 1) / 2;   // table pairs per image
  const int hx = 1 << (lx - 1);
  const int64_t nq = nn >> 2, hn = nn >> 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // per table pair (image A's NP, then image B's): S(hi, hi) − S(R, hi) and S(R, hi), per wave
  __shared__ double2 uni[4][2 * NP][2];
  double2 hh = make_double2(0.0, 0.0);
  if (lane < NP) hh = va.sat[lane * va.cells + va.y.hi * va.stride + va.x.hi];
  else if (lane < 2 * NP) hh = vb.sat[(lane - NP) * vb.cells + vb.y.hi * vb.stride + vb.x.hi];
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int64_t u = blockIdx.x * 256 + threadIdx.x; u < nq; u += (int64_t)gridDim.x * 256) {
    const int py = (int)(u >> (lx - 1)), px = (int)(u & (hx - 1));   // py is the same across the wave
    const int Ra = quad_r(va.y, py), Ca = quad_r(va.x, px), Rb = quad_r(vb.y, py), Cb = quad_r(vb.x, px);
    double2 rc[2 * NP], hc[2 * NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      rc[p] = va.sat[p * va.cells + (uint32_t)Ra * va.stride + Ca];
      hc[p] = va.sat[p * va.cells + (uint32_t)va.y.hi * va.stride + Ca];
      rc[NP + p] = vb.sat[p * vb.cells + (uint32_t)Rb * vb.stride + Cb];
      hc[NP + p] = vb.sat[p * vb.cells + (uint32_t)vb.y.hi * vb.stride + Cb];
    }
    if (lane < 2 * NP) {
      const double2 rh = lane < NP ? va.sat[lane * va.cells + (uint32_t)Ra * va.stride + va.x.hi]
                                   : vb.sat[(lane - NP) * vb.cells + (uint32_t)Rb * vb.stride + vb.x.hi];
      uni[wv][lane][0] = make_double2(hh.x - rh.x, hh.y - rh.y);
      uni[wv][lane][1] = rh;
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t t = ((int64_t)py << lx) + px + ((j & 1) ? hx : 0) + ((j & 2) ? hn : 0);
      // image A takes [R, hi) on an axis where the lag is >= 0 (j < 2 rows, even j columns), image
      // B the complement
      const bool uy = j < 2, ux = (j & 1) == 0;
      // rectangle sum of table `comp` of pair `slot`, upper or lower part per axis
      auto rect = [&](int slot, int comp, bool up_y, bool up_x) -> double {
        const double vrc = comp ? rc[slot].y : rc[slot].x;
        if (up_y && up_x) {
          const double2 w = uni[wv][slot][0];
          return (comp ? w.y : w.x) - ((comp ? hc[slot].y : hc[slot].x) - vrc);
        }
        if (up_y) return (comp ? hc[slot].y : hc[slot].x) - vrc;
        if (up_x) {
          const double2 w = uni[wv][slot][1];
          return (comp ? w.y : w.x) - vrc;
        }
        return vrc;
      };
      auto C = [&](int num, int64_t tt) -> double {
        if (P.sat_of[num] >= 0) {
          const int pl = P.sat_plane[P.sat_of[num]];
          if (pl == 0) {
            const int ay = uy ? va.y.hi - Ra : Ra, ax = ux ? va.x.hi - Ca : Ca;
            return (double)(ax * ay);
          }
          return rect((pl - 1) >> 1, (pl - 1) & 1, uy, ux);
        }
        if (P.bsat_of[num] >= 0) {
          const int pl = P.bsat_plane[P.bsat_of[num]];
          return rect(NP + ((pl - 1) >> 1), (pl - 1) & 1, !uy, !ux);
        }
        const int k = P.kept_of[num];
        const float4 v = corr4[(int64_t)(k >> 2) * nn + tt];
        return (double)((k & 3) == 0 ? v.x : (k & 3) == 1 ? v.y : (k & 3) == 2 ? v.z : v.w);
      };
      const float sc = (float)lag_score<PID & 3, true>(C, t, floor_n, a, calib);
      map[t] = sc;
      if (better(sc, (uint32_t)t, bv, bi)) {
        bv = sc;
        bi = (uint32_t)t;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

// One axis of a SatView in the quad form, or false. Rows (n = ny) or columns (n = nx) p of the
// first half pair with p − n/2; sat_idx clamps each edge b + s·d to [c0, c1].
static bool quad_axis(int n, int b0, int b1, int s, int c0, int c1, QuadAxis* q) {
  auto cl = [&](int v) { return std::min(std::max(v, c0), c1); };
  if (n < 2 || c0 != 0) return false;
  // lag ±n/2: both empty (the quad form takes row n/2 as −n/2, sat_idx as +n/2)
  if (cl(b0 + s * (n / 2)) != cl(b1 + s * (n / 2)) || cl(b0 - s * (n / 2)) != cl(b1 - s * (n / 2))) return false;
  for (int pat = 0; pat < 2; ++pat) {
    const int hi = pat == 0 ? cl(b1) : cl(b1 - s * (n / 2));
    bool ok = true;
    for (int p = 0; p < n / 2 && ok; ++p) {
      const int d = p, e = p - n / 2;
      const int y0d = cl(b0 + s * d), y1d = cl(b1 + s * d), y0e = cl(b0 + s * e), y1e = cl(b1 + s * e);
      ok = pat == 0 ? (y0d == y1e && y0e == 0 && y1d == hi) : (y1d == y0e && y0d == 0 && y1e == hi);
    }
    if (ok) {
      *q = {pat == 0 ? b0 : b1, s, c0, c1, hi, pat};
      return true;
    }
  }
  return false;
}

// The moving image's view (sign +1) must take pattern 0 on both axes, the fixed image's (sign −1)
// pattern 1, as k_combine_quad assumes.
bool quad_view(const SatView& v, QuadView* q) {
  if (v.nx < 128 || v.ny < 2) return false;   // a wave's 64 quads share one row
  q->sat = v.sat;
  q->stride = (uint32_t)v.stride;
  q->cells = v.cells;
  const int want = v.sign > 0 ? 0 : 1;
  return quad_axis(v.ny, v.by0, v.by1, v.sign, v.cy0, v.cy1, &q->y) &&
         quad_axis(v.nx, v.bx0, v.bx1, v.sign, v.cx0, v.cx1, &q->x) && q->y.patB == want && q->x.patB == want;
}

// ── roto-scale maps: per-lag score and finish ─────────────────────────────────────────────
// Lag t of the n x n map: row k (signed, lambda), column kx (theta). The correlations were taken
// on a px x py grid: px = n (circular in theta) or a padded length >= 2n, whose two linear lags kx
// and kx - n fold into the circular one; py >= n (linear in lambda). Each correlation is scaled by
// e^(k dlam) dlam dth (extract_pair's area element, smi.wgsl:778-796).
template <int PID>
__global__ __launch_bounds__(256) void k_combine_rs(const float* __restrict__ corr, float* map, RsArgs r,
                                                    CombineArgs a) {
  constexpr CorrPlan P = kPlans[PID];
  constexpr bool EXACT = PID & 1;
  const int64_t nn = (int64_t)r.px * r.py, total = (int64_t)r.n * r.n;
  int64_t base = 0, base2 = -1;
  double sc = 0.0;
  auto at = [&](int64_t t) {
    const int ky = (int)(t / r.n), kx = (int)(t - (int64_t)ky * r.n);
    const int k = ky > r.n / 2 ? ky - r.n : ky;
    const int row = k < 0 ? k + r.py : k;
    base = (int64_t)row * r.px + kx;
    base2 = (r.px != r.n && kx > 0) ? base - r.n + r.px : -1;
    sc = exp((double)k * r.dlam) * r.dlam * r.dth;
  };
  auto C = [&](int num, int64_t) -> double {
    const float* c = corr + (int64_t)P.kept_of[num] * nn;
    double v = (double)c[base];
    if (base2 >= 0) v += (double)c[base2];
    return v * sc;
  };
  at(0);
  const double n0 = C(0, 0);
  const bool calib = EXACT && a.calib;
  const double floor_n = calib ? (double)a.min_n : fmax((double)a.min_n, 0.08 * fmax(n0, 0.0));
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    at(t);
    map[t] = (float)lag_score<PID>(C, t, floor_n, a, calib);
  }
}

// fwd averaged with inv at the mirrored lag (-y, -x), or (-y, x) when !mirror; rows whose applied
// scale exp(-k dlam) lies outside [e^lo, e^hi] zeroed (smi.wgsl:1708-1731); block argmax partials.
__global__ __launch_bounds__(256) void k_rs_finish(const float* __restrict__ fwd, const float* __restrict__ inv,
                                                   int n, int mirror, double dlam, double lo, double hi, float* out,
                                                   float* arg_part) {
  const int64_t total = (int64_t)n * n;
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int y = (int)(t / n), x = (int)(t - (int64_t)y * n);
    float v = fwd[t];
    if (inv) {
      const int ym = (n - y) % n, xm = mirror ? (n - x) % n : x;
      v = 0.5f * (v + inv[(int64_t)ym * n + xm]);
    }
    const int k = y > n / 2 ? y - n : y;
    const double sg = -(double)k * dlam;
    if (sg < lo || sg > hi) v = 0.0f;
    out[t] = v;
    if (better(v, (uint32_t)t, bv, bi)) {
      bv = v;
      bi = (uint32_t)t;
    }
  }
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    arg_part[2 * blockIdx.x] = sv_[0];
    arg_part[2 * blockIdx.x + 1] = __uint_as_float(si_[0]);
  }
}

__global__ __launch_bounds__(256) void k_argmax_finish(const float* __restrict__ part, int nparts,
                                                       const float* __restrict__ map, float* res) {
  __shared__ float sv_[256];
  __shared__ uint32_t si_[256];
  float bv = -3.4e38f;
  uint32_t bi = 0xffffffffu;
  for (int g = threadIdx.x; g < nparts; g += 256) {
    const float v = part[2 * g];
    const uint32_t i = __float_as_uint(part[2 * g + 1]);
    if (i != 0xffffffffu && better(v, i, bv, bi)) {
      bv = v;
      bi = i;
    }
  }
  sv_[threadIdx.x] = bv;
  si_[threadIdx.x] = bi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k && better(sv_[threadIdx.x + k], si_[threadIdx.x + k], sv_[threadIdx.x], si_[threadIdx.x])) {
      sv_[threadIdx.x] = sv_[threadIdx.x + k];
      si_[threadIdx.x] = si_[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    res[0] = sv_[0];
    res[1] = __uint_as_float(si_[0]);
    res[2] = map[0];
  }
}



int combine_blocks(int64_t nn) { return (int)std::min<int64_t>((nn + 255) / 256, 4096); }

void launch_combine(int plan, const float* corr, float* map, int64_t nn, const CombineArgs& a, float* arg_part,
                    hipStream_t st, const SatView* sv, const SatView* svb) {
  const int g = combine_blocks(nn);
  // every score kernel writes g block partials (launch_argmax_finish reads combine_blocks(nn))
  if (plan >= 8 && sv && svb && MapPath::trans == 1) {
    QuadView qa{}, qb{};
    if (quad_view(*sv, &qa) && quad_view(*svb, &qb) && sv->nx == svb->nx && sv->ny == svb->ny) {
      const double n0 = (double)(std::min(std::max(sv->bx1, sv->cx0), sv->cx1) - std::min(std::max(sv->bx0, sv->cx0), sv->cx1)) *
                        (double)(std::min(std::max(sv->by1, sv->cy0), sv->cy1) - std::min(std::max(sv->by0, sv->cy0), sv->cy1));
      const float4* c4 = reinterpret_cast<const float4*>(corr);
      switch (plan) {
        case 8: k_combine_quad<8><<<g, 256, 0, st>>>(c4, map, nn, sv->lx, a, qa, qb, n0, arg_part); break;
        case 9: k_combine_quad<9><<<g, 256, 0, st>>>(c4, map, nn, sv->lx, a, qa, qb, n0, arg_part); break;
        case 10: k_combine_quad<10><<<g, 256, 0, st>>>(c4, map, nn, sv->lx, a, qa, qb, n0, arg_part); break;
        default: k_combine_quad<11><<<g, 256, 0, st>>>(c4, map, nn, sv->lx, a, qa, qb, n0, arg_part); break;
      }
      return;
    }
  }
  if (plan >= 4 && !sv) throw std::runtime_error("rectangle plan without tables");
  if (plan >= 8 && !svb) throw std::runtime_error("translation plan without the fixed side's tables");
  const SatView b = svb ? *svb : SatView{};
  // rectangle plans: corr holds the correlation quads of launch_rowifft_quads
  const float4* c4 = reinterpret_cast<const float4*>(corr);
  switch (plan) {
    case 0: k_combine<0><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 1: k_combine<1><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 2: k_combine<2><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 3: k_combine<3><<<g, 256, 0, st>>>(corr, map, nn, a, arg_part); break;
    case 4: k_combine_rect<4><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    case 5: k_combine_rect<5><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    case 6: k_combine_rect<6><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    case 7: k_combine_rect<7><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    case 8: k_combine_rect<8><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    case 9: k_combine_rect<9><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    case 10: k_combine_rect<10><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
    default: k_combine_rect<11><<<g, 256, 0, st>>>(c4, map, nn, a, *sv, b, arg_part); break;
  }
}

__global__ __launch_bounds__(256) void k_map_remap(const float* __restrict__ in, int nx, int ny, float* out, int rx,
                                                   int ry) {
  const int64_t n = (int64_t)rx * ry;
  for (int64_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int jy = (int)(t / rx), jx = (int)(t - (int64_t)jy * rx);
    const int ix = 2 * jx <= rx ? jx : jx - rx + nx;
    const int iy = 2 * jy <= ry ? jy : jy - ry + ny;
    out[t] = in[(int64_t)iy * nx + ix];
  }
}

void launch_map_remap(const float* in, int nx, int ny, float* out, int rx, int ry, hipStream_t st) {
  if (rx > nx || ry > ny || rx <= 0 || ry <= 0) throw std::runtime_error("map_remap: the grid to read is smaller");
  k_map_remap<<<(unsigned)std::min<int64_t>(((int64_t)rx * ry + 255) / 256, 8192), 256, 0, st>>>(in, nx, ny, out, rx, ry);
}

void launch_argmax_finish(const float* arg_part, int nparts, const float* map, float* res, hipStream_t st) {
  k_argmax_finish<<<1, 256, 0, st>>>(arg_part, nparts, map, res);
}



void launch_combine_rs(int plan, const float* corr, float* map, const RsArgs& r, const CombineArgs& a,
                       hipStream_t st) {
  const int g = combine_blocks((int64_t)r.n * r.n);
  switch (plan) {
    case 0: k_combine_rs<0><<<g, 256, 0, st>>>(corr, map, r, a); break;
    case 1: k_combine_rs<1><<<g, 256, 0, st>>>(corr, map, r, a); break;
    case 2: k_combine_rs<2><<<g, 256, 0, st>>>(corr, map, r, a); break;
    default: k_combine_rs<3><<<g, 256, 0, st>>>(corr, map, r, a); break;
  }
}

void launch_rs_finish(const float* fwd, const float* inv, int n, bool mirror, double dlam, double lo, double hi,
                      float* out, float* arg_part, hipStream_t st) {
  k_rs_finish<<<combine_blocks((int64_t)n * n), 256, 0, st>>>(fwd, inv, n, mirror ? 1 : 0, dlam, lo, hi, out, arg_part);
}

}  // namespace zc
