// Pose moments and tangent gradients (Gauss–Newton) over pose jobs.
#include "kdev.hpp"

namespace zc {

// ── pose moments ──────────────────────────────────────────────────────────────────────────
__device__ __forceinline__ float ramp_w(float v, float lo, float hi) {
  const float a = (v - lo) * 0.5f, b = (hi - v) * 0.5f;
  return fminf(fmaxf(a, 0.0f), 1.0f) * fminf(fmaxf(b, 0.0f), 1.0f);
}

// Work split of the pose reductions: a wave takes chunks of 64 · kRunPx consecutive destination
// pixels (kRunPx per lane, 64 apart), sums each lane's run in FP32, adds the 64 lanes' runs in six
// pairwise DPP levels and adds the chunk's sum into an FP64 total. A chunk's FP32 sum depends only
// on the chunk, not on the grid, so every launch shape gives the same totals to FP64 rounding,
// and each FP32 rounding sits within 32 sequential + 6 pairwise additions (the error of a total
// stays below 38 · 2^-24 of its sum of |terms|, however many pixels the pose covers).
constexpr int kRunPx = 32;
constexpr int kChunkPx = 64 * kRunPx;

// v + (v of another lane by DPP control CTRL), in the rows of ROWS (other rows add 0)
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_plus(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// The sum of v over the wave, as a uniform value: pairwise within quads (quad_perm 1,0,3,2 and
// 2,3,0,1), across quad pairs (row_half_mirror) and half rows (row_mirror) - every lane then holds
// its 16-lane row's sum - then rows 1 and 3 add the row below (row_bcast:15) and rows 2-3 add
// lane 31 (row_bcast:31), which leaves the total in lane 63.
__device__ __forceinline__ float wave_total(float v) {
  v = dpp_plus<0xB1, 0xF>(v);
  v = dpp_plus<0x4E, 0xF>(v);
  v = dpp_plus<0x141, 0xF>(v);
  v = dpp_plus<0x140, 0xF>(v);
  v = dpp_plus<0x142, 0xA>(v);
  v = dpp_plus<0x143, 0xC>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// A feature plane read through the global address space with a 32-bit element offset (a plane
// holds at most 2^30 pixels, so the byte offset fits 32 bits).
typedef const __attribute__((address_space(1))) char* gbyte_ptr;
__device__ __forceinline__ float plane_at(const float* plane, uint32_t i) {
  return *(const __attribute__((address_space(1))) float*)((gbyte_ptr)plane + (i << 2));
}

// The four features of pixel i: from four planes of n pixels (zc_features' layout) ...
struct PlanarFeat {
  const float* f;
  int64_t n;
  __device__ __forceinline__ float4 at(uint32_t i) const {
    return make_float4(plane_at(f, i), plane_at(f + n, i), plane_at(f + 2 * n, i), plane_at(f + 3 * n, i));
  }
};
// ... or from one float4 per pixel (the interleaved copy): one load per tap instead of four.
typedef float f32x4 __attribute__((ext_vector_type(4)));
struct PackedFeat {
  const float4* f;
  __device__ __forceinline__ float4 at(uint32_t i) const {
    const f32x4 v = ((const __attribute__((address_space(1))) f32x4*)f)[i];
    return make_float4(v.x, v.y, v.z, v.w);
  }
};

__device__ __forceinline__ float comp(const float4& v, int q) { return q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w; }

// The destination pixel index t as (x, y), moved on by one wave step (64 pixels).
__device__ __forceinline__ void step64(int& x, int& y, int wb) {
  x += 64;
  while (x >= wb) {
    x -= wb;
    ++y;
  }
}

// One destination pixel of a pose: its taper weight (0 outside the moving image's inner area,
// where the projection degenerates, or past the last pixel) and the four features of each side
// there. Branch-free, so two pixels' loads can be in flight together: a pixel that does not count
// samples the safe cell (2, 2) of fa (the caller ensures wa, ha >= 5) and fb's last pixel.
struct PosePixel {
  float wt, va[4], vb[4];
};

template <class F>
__device__ __forceinline__ PosePixel pose_pixel(const F& A, const F& B, int wa, int64_t nb, const double* Hi,
                                                double xmax, double ymax, float sw, float sh, int x, int y,
                                                uint32_t t) {
  PosePixel p;
  const bool live = t < (uint64_t)nb;
  const double X = x + 1.0, Y = y + 1.0;
  const double qz = fma(Hi[6], X, fma(Hi[7], Y, Hi[8]));
  const bool proj = fabs(qz) >= 1e-8;
  const double iq = 1.0 / (proj ? qz : 1.0);
  double fxd = fma(Hi[0], X, fma(Hi[1], Y, Hi[2])) * iq - 1.0;
  double fyd = fma(Hi[3], X, fma(Hi[4], Y, Hi[5])) * iq - 1.0;
  const bool in = live && proj && fxd >= 2.0 && fyd >= 2.0 && fxd <= xmax && fyd <= ymax;   // NaN fails too
  fxd = in ? fxd : 2.0;
  fyd = in ? fyd : 2.0;
  const float fx = (float)fxd, fy = (float)fyd;
  p.wt = in ? ramp_w(fx, 2.0f, sw) * ramp_w(fy, 2.0f, sh) : 0.0f;
  const int x0 = (int)fxd, y0 = (int)fyd;   // >= 2: truncation is floor
  const float ax = (float)(fxd - x0), ay = (float)(fyd - y0);
  const uint32_t i00 = (uint32_t)y0 * (uint32_t)wa + (uint32_t)x0, i01 = i00 + (uint32_t)wa;
  const float4 a00 = A.at(i00), a10 = A.at(i00 + 1), a01 = A.at(i01), a11 = A.at(i01 + 1);
  const float4 bv = B.at(live ? t : (uint32_t)(nb - 1));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float v00 = comp(a00, q), v10 = comp(a10, q), v01 = comp(a01, q), v11 = comp(a11, q);
    p.va[q] = (v00 * (1.0f - ax) + v10 * ax) * (1.0f - ay) + (v01 * (1.0f - ax) + v11 * ax) * ay;
    p.vb[q] = comp(bv, q);
  }
  return p;
}

// s += wt · [1, va, vb, va ⊗ vb, upper(va ⊗ va), upper(vb ⊗ vb)] (the 45 monomials' order)
__device__ __forceinline__ void pose_accumulate(float (&s)[45], const PosePixel& p) {
  float wva[4], wvb[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    wva[i] = p.wt * p.va[i];
    wvb[i] = p.wt * p.vb[i];
  }
  s[0] += p.wt;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    s[1 + i] += wva[i];
    s[5 + i] += wvb[i];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) s[9 + i * 4 + j] += wva[i] * p.vb[j];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i; j < 4; ++j) {
      s[25 + tri_k(i, j)] += wva[i] * p.va[j];
      s[35 + tri_k(i, j)] += wvb[i] * p.vb[j];
    }
}

// M FP32 sums of each lane, added over the wave; lane m < M adds sum m into its FP64 total.
template <int M>
__device__ __forceinline__ void wave_flush(const float (&s)[M], double& tot, int lane) {
  float mine = 0.0f;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const float v = wave_total(s[m]);
    mine = lane == m ? v : mine;
  }
  tot += (double)mine;
}

// The workgroup's M totals (lane m of each of its 4 waves holds total m) into out[0..M).
template <int M>
__device__ __forceinline__ void block_flush(double tot, double* out) {
  __shared__ double sm[4][M];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane < M) sm[wv][lane] = tot;
  __syncthreads();
  if (threadIdx.x < M) out[threadIdx.x] = (sm[0][threadIdx.x] + sm[1][threadIdx.x]) + (sm[2][threadIdx.x] + sm[3][threadIdx.x]);
}

// The 45 overlap sums of a pose (smi.wgsl:980-1057 overlap_moments): moving features A (wa x ha)
// sampled bilinearly at Hi·X for every destination pixel X of B (wb x hb), tapered over 2 px
// inside the moving image, into part[0..45) for workgroup blk of nblk (256 threads). Wave w of the
// workgroup takes chunks blk · 4 + w, then every (nblk · 4)-th; two pixels per step.
template <class F>
__device__ __forceinline__ void pose_moments_body(const F& A, int wa, int ha, const F& B, int wb, int hb,
                                                  const double* Hi, double* part, int blk, int nblk) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t nb = (int64_t)wb * hb;
  const double xmax = wa - 3.0, ymax = ha - 3.0;
  const float sw = (float)wa - 3.0f, sh = (float)ha - 3.0f;
  const int64_t nchunk = wa >= 5 && ha >= 5 ? (nb + kChunkPx - 1) / kChunkPx : 0;   // else nothing is inside
  double tot = 0.0;
  for (int64_t c = (int64_t)blk * 4 + wv; c < nchunk; c += (int64_t)nblk * 4) {
    uint32_t t = (uint32_t)(c * kChunkPx) + lane;   // < 2^30 + kChunkPx
    int y = (int)(t / (uint32_t)wb), x = (int)(t - (uint32_t)y * (uint32_t)wb);
    float s[45];
#pragma unroll
    for (int m = 0; m < 45; ++m) s[m] = 0.0f;
    for (int i = 0; i < kRunPx; i += 2) {
      int x1 = x, y1 = y;
      step64(x1, y1, wb);
      const PosePixel p0 = pose_pixel(A, B, wa, nb, Hi, xmax, ymax, sw, sh, x, y, t);
      const PosePixel p1 = pose_pixel(A, B, wa, nb, Hi, xmax, ymax, sw, sh, x1, y1, t + 64);
      pose_accumulate(s, p0);
      pose_accumulate(s, p1);
      x = x1;
      y = y1;
      step64(x, y, wb);
      t += 128;
    }
    wave_flush<45>(s, tot, lane);
  }
  block_flush<45>(tot, part);
}

__global__ __launch_bounds__(256) void k_pose_moments(const float* __restrict__ fa, int wa, int ha,
                                                      const float* __restrict__ fb, int wb, int hb, PoseArgs p,
                                                      double* part) {
  const PlanarFeat A{fa, (int64_t)wa * ha}, B{fb, (int64_t)wb * hb};
  pose_moments_body(A, wa, ha, B, wb, hb, p.Hi, part + (int64_t)blockIdx.x * 45, blockIdx.x, gridDim.x);
}

// One pose per blockIdx.y: its partials at part[(y · gridDim.x + x) · 45].
__global__ __launch_bounds__(256) void k_pose_moments_jobs(const PoseJob* __restrict__ jobs, double* part) {
  const PoseJob& j = jobs[blockIdx.y];
  // an idle job (a slot of a device batch with nothing to sum; images are at least 8 px) leaves its
  // partial sums unwritten: nothing reads them
  if (j.wa < 5 || j.ha < 5) return;
  const PackedFeat A{j.fa}, B{j.fb};
  pose_moments_body(A, j.wa, j.ha, B, j.wb, j.hb, j.Hi, part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 45,
                    blockIdx.x, gridDim.x);
}

// The score's tangent gradient and Gauss–Newton matrix at a pose (smi.wgsl:1129-1272
// tangent_grad), one pose per blockIdx.y: per destination pixel, ∂S/∂v_a = 2 (W t_b − M̂ t_a)
// dotted with the sample's gradients gives (cx, cy) (plus the taper's boundary term), times each
// tangent's ∂(sample)/∂ξ_k from M_k = −H⁻¹ G_k; the matrix Σ wt J_aᵀ Q J_a from the four
// scalars g_xᵀ Q g_x, g_xᵀ Q g_y, g_yᵀ Q g_x, g_yᵀ Q g_y. Partials: 8 gradient, then the upper
// triangle row by row (36), FP64 across the block, at part[(y · gridDim.x + x) · 44].
// One destination pixel (x, y) = pixel t of fb of a tangent job, added into acc (8 gradient
// entries, then the matrix's upper triangle row by row); pixels outside the moving image's inner
// area add nothing.
__device__ __forceinline__ void tangent_pixel(const TangentJob& j, int64_t na, int64_t nb, float sw, float sh, int x,
                                              int y, uint32_t t, float (&acc)[44]) {
  const int wa = j.wa, ha = j.ha;
  const float X0 = (float)x + 1.0f, X1 = (float)y + 1.0f;
  // the sample position in FP64: the bilinear gradients jump at cell edges, so the cell must
  // not depend on the rounding of the position
  const double pxd = j.Hi[0] * (double)X0 + j.Hi[1] * (double)X1 + j.Hi[2];
  const double pyd = j.Hi[3] * (double)X0 + j.Hi[4] * (double)X1 + j.Hi[5];
  const double pzd = j.Hi[6] * (double)X0 + j.Hi[7] * (double)X1 + j.Hi[8];
  if (fabs(pzd) < 1e-8) return;
  const double fxd = pxd / pzd - 1.0, fyd = pyd / pzd - 1.0;
  if (!(fxd >= 2.0 && fyd >= 2.0 && fxd <= wa - 3.0 && fyd <= ha - 3.0)) return;   // NaN fails too
  const float phx = (float)pxd, phy = (float)pyd, phz = (float)pzd;
  const float fx = (float)fxd, fy = (float)fyd;
  // taper (smi.wgsl:338-351): value and its derivatives along x and y
  float tvx, dtx, tvy, dty;
  {
    const float a = (fx - 2.0f) * 0.5f, b = (sw - 3.0f - fx) * 0.5f;
    const float ca = fminf(fmaxf(a, 0.0f), 1.0f), cb = fminf(fmaxf(b, 0.0f), 1.0f);
    const float da = (a > 0.0f && a < 1.0f) ? 0.5f : 0.0f, db = (b > 0.0f && b < 1.0f) ? -0.5f : 0.0f;
    tvx = ca * cb;
    dtx = da * cb + ca * db;
  }
  {
    const float a = (fy - 2.0f) * 0.5f, b = (sh - 3.0f - fy) * 0.5f;
    const float ca = fminf(fmaxf(a, 0.0f), 1.0f), cb = fminf(fmaxf(b, 0.0f), 1.0f);
    const float da = (a > 0.0f && a < 1.0f) ? 0.5f : 0.0f, db = (b > 0.0f && b < 1.0f) ? -0.5f : 0.0f;
    tvy = ca * cb;
    dty = da * cb + ca * db;
  }
  const float wt = tvx * tvy, twx = dtx * tvy, twy = tvx * dty;
  const int x0 = (int)fxd, y0 = (int)fyd;   // >= 2: truncation is floor
  const float ax = (float)(fxd - x0), ay = (float)(fyd - y0);
  const uint32_t i00 = (uint32_t)y0 * (uint32_t)wa + (uint32_t)x0, i01 = i00 + (uint32_t)wa;
  const PackedFeat A{j.fa}, B{j.fb};
  const float4 a00 = A.at(i00), a10 = A.at(i00 + 1), a01 = A.at(i01), a11 = A.at(i01 + 1), bv = B.at(t);
  float va[4], gx[4], gy[4], vb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float v00 = comp(a00, q), v10 = comp(a10, q), v01 = comp(a01, q), v11 = comp(a11, q);
    va[q] = (v00 * (1.0f - ax) + v10 * ax) * (1.0f - ay) + (v01 * (1.0f - ax) + v11 * ax) * ay;
    gx[q] = (v10 - v00) * (1.0f - ay) + (v11 - v01) * ay;
    gy[q] = (v01 - v00) * (1.0f - ax) + (v11 - v10) * ax;
    vb[q] = comp(bv, q);
  }
  float ra[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    float s = 0.0f;
#pragma unroll
    for (int b = 0; b < 4; ++b) s += j.W[a * 4 + b] * (vb[b] - j.mb[b]) - j.Mh[a * 4 + b] * (va[b] - j.ma[b]);
    ra[a] = 2.0f * s;
  }
  float bw = 0.0f;
  if (twx != 0.0f || twy != 0.0f) {   // ∂S/∂w: Σ_k g_k φ_k(v_a, v_b) over the 45 monomials
    float s = j.g[0];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      s += j.g[1 + i] * va[i] + j.g[5 + i] * vb[i];
#pragma unroll
      for (int k = 0; k < 4; ++k) s += j.g[9 + i * 4 + k] * va[i] * vb[k];
#pragma unroll
      for (int k = i; k < 4; ++k) s += j.g[25 + tri_k(i, k)] * va[i] * va[k] + j.g[35 + tri_k(i, k)] * vb[i] * vb[k];
    }
    bw = s;
  }
  const float cx = wt * (ra[0] * gx[0] + ra[1] * gx[1] + ra[2] * gx[2] + ra[3] * gx[3]) + bw * twx;
  const float cy = wt * (ra[0] * gy[0] + ra[1] * gy[1] + ra[2] * gy[2] + ra[3] * gy[3]) + bw * twy;
  const float iz2 = 1.0f / fmaxf(phz * phz, 1e-12f);
  float px[8], py[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const float* M = j.Mk + 9 * k;
    const float w0 = M[0] * X0 + M[1] * X1 + M[2];
    const float w1 = M[3] * X0 + M[4] * X1 + M[5];
    const float w2 = M[6] * X0 + M[7] * X1 + M[8];
    px[k] = (w0 * phz - phx * w2) * iz2;
    py[k] = (w1 * phz - phy * w2) * iz2;
    acc[k] += cx * px[k] + cy * py[k];
  }
  if (j.hess) {
    float qgx[4], qgy[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      qgx[a] = j.Q[a * 4] * gx[0] + j.Q[a * 4 + 1] * gx[1] + j.Q[a * 4 + 2] * gx[2] + j.Q[a * 4 + 3] * gx[3];
      qgy[a] = j.Q[a * 4] * gy[0] + j.Q[a * 4 + 1] * gy[1] + j.Q[a * 4 + 2] * gy[2] + j.Q[a * 4 + 3] * gy[3];
    }
    const float sxx = gx[0] * qgx[0] + gx[1] * qgx[1] + gx[2] * qgx[2] + gx[3] * qgx[3];
    const float sxy = gx[0] * qgy[0] + gx[1] * qgy[1] + gx[2] * qgy[2] + gx[3] * qgy[3];
    const float syx = gy[0] * qgx[0] + gy[1] * qgx[1] + gy[2] * qgx[2] + gy[3] * qgx[3];
    const float syy = gy[0] * qgy[0] + gy[1] * qgy[1] + gy[2] * qgy[2] + gy[3] * qgy[3];
    float r[8], q[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
      r[l] = wt * (sxx * px[l] + sxy * py[l]);
      q[l] = wt * (syx * px[l] + syy * py[l]);
    }
    int h = 8;
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int l = k; l < 8; ++l) acc[h++] += px[k] * r[l] + py[k] * q[l];
  }
}

__global__ __launch_bounds__(256) void k_tangent_jobs(const TangentJob* __restrict__ jobs, double* part) {
  const TangentJob& j = jobs[blockIdx.y];
  if (j.wa < 5 || j.ha < 5) return;   // an idle job: nothing reads its partial sums
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int wa = j.wa, ha = j.ha, wb = j.wb;
  const int64_t na = (int64_t)wa * ha, nb = (int64_t)wb * j.hb;
  const float sw = (float)wa, sh = (float)ha;
  const int64_t nchunk = wa >= 5 && ha >= 5 ? (nb + kChunkPx - 1) / kChunkPx : 0;   // else nothing is inside
  double tot = 0.0;
  // the chunks of the moments kernel: runs of kRunPx pixels per lane in FP32, FP64 per chunk
  for (int64_t c = (int64_t)blockIdx.x * 4 + wv; c < nchunk; c += (int64_t)gridDim.x * 4) {
    uint32_t t = (uint32_t)(c * kChunkPx) + lane;   // < 2^30 + kChunkPx
    int y = (int)(t / (uint32_t)wb), x = (int)(t - (uint32_t)y * (uint32_t)wb);
    float acc[44];
#pragma unroll
    for (int m = 0; m < 44; ++m) acc[m] = 0.0f;
    for (int i = 0; i < kRunPx; ++i) {
      if (t < (uint64_t)nb) tangent_pixel(j, na, nb, sw, sh, x, y, t, acc);
      step64(x, y, wb);
      t += 64;
    }
    wave_flush<44>(acc, tot, lane);
  }
  block_flush<44>(tot, part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 44);
}


// Workgroups of a single pose's moments: about 4 chunks per wave, at most 1024.
int moments_blocks(int w, int h) {
  const int64_t chunks = ((int64_t)w * h + kChunkPx - 1) / kChunkPx;
  return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (chunks + 15) / 16));
}

void launch_pose_moments(const float* feat_a, int wa, int ha, const float* feat_b, int wb, int hb, const PoseArgs& p,
                         double* part, hipStream_t st) {
  k_pose_moments<<<moments_blocks(wb, hb), 256, 0, st>>>(feat_a, wa, ha, feat_b, wb, hb, p, part);
}

// Workgroups per pose: about 4 chunks per wave, at most 256 (more, down to one chunk per wave, for
// launches of few jobs), a multiple of 8 — the hardware deals
// workgroups to the 8 XCDs in turn, so chunk x of consecutive poses (the tries of one slice, which
// read the same features) runs on the same XCD and shares its L2.
int jobs_blocks(int64_t max_pixels, int njobs) {
  const int64_t chunks = (max_pixels + kChunkPx - 1) / kChunkPx;
  int64_t b = std::max<int64_t>(1, std::min<int64_t>(256, (chunks + 15) / 16));
  // few jobs (a single pose and its tries): enough workgroups to fill the chip — about 2048 in
  // all, down to one chunk per wave
  const int64_t fill = (2048 + std::max(njobs, 1) - 1) / std::max(njobs, 1);
  b = std::max(b, std::min<int64_t>(fill, std::min<int64_t>(1024, (chunks + 3) / 4)));
  return (int)((b + 7) / 8 * 8);
}

void launch_pose_moments_jobs(const PoseJob* jobs, int njobs, int blocks, double* part, double* sums,
                              hipStream_t st) {
  k_pose_moments_jobs<<<dim3(blocks, njobs), 256, 0, st>>>(jobs, part);
  launch_sum_parts(part, blocks, 45, sums, st, njobs);
}

void launch_tangent_jobs(const TangentJob* jobs, int njobs, int blocks, double* part, double* sums,
                         hipStream_t st) {
  k_tangent_jobs<<<dim3(blocks, njobs), 256, 0, st>>>(jobs, part);
  launch_sum_parts(part, blocks, 44, sums, st, njobs);
}


}  // namespace zc
