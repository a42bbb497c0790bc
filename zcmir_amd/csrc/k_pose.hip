// Pose moments and tangent gradients (Gauss–Newton) over pose jobs.
#include "kdev.hpp"

namespace zc {

// ── pose moments ──────────────────────────────────────────────────────────────────────────
// FP32 two-sum (Knuth) accumulation: s + c carries the exact running sum to ~2^-48.
__device__ __forceinline__ void two_sum_add(float& s, float& c, float x) {
  const float t = s + x;
  const float z = t - s;
  c += (s - (t - z)) + (x - z);
  s = t;
}

__device__ __forceinline__ float ramp_w(float v, float lo, float hi) {
  const float a = (v - lo) * 0.5f, b = (hi - v) * 0.5f;
  return fminf(fmaxf(a, 0.0f), 1.0f) * fminf(fmaxf(b, 0.0f), 1.0f);
}

// The 45 overlap sums of a pose (smi.wgsl:980-1057 overlap_moments): moving features fa sampled
// bilinearly at Hi·X for every destination pixel X of fb, tapered over 2 px inside the moving
// image; per-thread FP32 two-sum, then FP64 across the block into part[0..45).
__device__ __forceinline__ void pose_moments_body(const float* __restrict__ fa, int wa, int ha,
                                                  const float* __restrict__ fb, int wb, int hb, const double* Hi,
                                                  double* part, int blk, int nblk) {
  const PoseArgs p = {{Hi[0], Hi[1], Hi[2], Hi[3], Hi[4], Hi[5], Hi[6], Hi[7], Hi[8]}};
  float s[45], c[45];
#pragma unroll
  for (int m = 0; m < 45; ++m) s[m] = c[m] = 0.0f;
  const int64_t na = (int64_t)wa * ha, nb = (int64_t)wb * hb;
  for (int64_t t = (int64_t)blk * 256 + threadIdx.x; t < nb; t += (int64_t)nblk * 256) {
    const int x = (int)(t % wb), y = (int)(t / wb);
    const double X = x + 1.0, Y = y + 1.0;
    const double qz = p.Hi[6] * X + p.Hi[7] * Y + p.Hi[8];
    if (fabs(qz) < 1e-8) continue;
    const double fxd = (p.Hi[0] * X + p.Hi[1] * Y + p.Hi[2]) / qz - 1.0;
    const double fyd = (p.Hi[3] * X + p.Hi[4] * Y + p.Hi[5]) / qz - 1.0;
    if (!(fxd >= 2.0 && fyd >= 2.0 && fxd <= wa - 3.0 && fyd <= ha - 3.0)) continue;   // NaN fails too
    const float fx = (float)fxd, fy = (float)fyd;
    const float wt = ramp_w(fx, 2.0f, (float)wa - 3.0f) * ramp_w(fy, 2.0f, (float)ha - 3.0f);
    if (!(wt > 0.0f)) continue;
    const int x0 = (int)floor(fxd), y0 = (int)floor(fyd);
    const float ax = (float)(fxd - x0), ay = (float)(fyd - y0);
    const int xa = min(max(x0, 0), wa - 1), xb = min(max(x0 + 1, 0), wa - 1);
    const int ya = min(max(y0, 0), ha - 1), yb = min(max(y0 + 1, 0), ha - 1);
    float va[4], vb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* f = fa + q * na;
      const float v00 = f[(int64_t)ya * wa + xa], v10 = f[(int64_t)ya * wa + xb];
      const float v01 = f[(int64_t)yb * wa + xa], v11 = f[(int64_t)yb * wa + xb];
      va[q] = (v00 * (1.0f - ax) + v10 * ax) * (1.0f - ay) + (v01 * (1.0f - ax) + v11 * ax) * ay;
      vb[q] = fb[q * nb + t];
    }
    two_sum_add(s[0], c[0], wt);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      two_sum_add(s[1 + i], c[1 + i], wt * va[i]);
      two_sum_add(s[5 + i], c[5 + i], wt * vb[i]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) two_sum_add(s[9 + i * 4 + j], c[9 + i * 4 + j], wt * va[i] * vb[j]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = i; j < 4; ++j) {
        two_sum_add(s[25 + tri_k(i, j)], c[25 + tri_k(i, j)], wt * va[i] * va[j]);
        two_sum_add(s[35 + tri_k(i, j)], c[35 + tri_k(i, j)], wt * vb[i] * vb[j]);
      }
  }
  double acc[45];
#pragma unroll
  for (int m = 0; m < 45; ++m) acc[m] = (double)s[m] + (double)c[m];
  block_sum<double, 45, 256>(acc, part, threadIdx.x);
}

__global__ __launch_bounds__(256) void k_pose_moments(const float* __restrict__ fa, int wa, int ha,
                                                      const float* __restrict__ fb, int wb, int hb, PoseArgs p,
                                                      double* part) {
  pose_moments_body(fa, wa, ha, fb, wb, hb, p.Hi, part + (int64_t)blockIdx.x * 45, blockIdx.x, gridDim.x);
}

// One pose per blockIdx.y: its partials at part[(y · gridDim.x + x) · 45].
__global__ __launch_bounds__(256) void k_pose_moments_jobs(const PoseJob* __restrict__ jobs, double* part) {
  const PoseJob& j = jobs[blockIdx.y];
  pose_moments_body(j.fa, j.wa, j.ha, j.fb, j.wb, j.hb, j.Hi,
                    part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 45, blockIdx.x, gridDim.x);
}

// The score's tangent gradient and Gauss–Newton matrix at a pose (smi.wgsl:1129-1272
// tangent_grad), one pose per blockIdx.y: per destination pixel, ∂S/∂v_a = 2 (W t_b − M̂ t_a)
// dotted with the sample's gradients gives (cx, cy) (plus the taper's boundary term), times each
// tangent's ∂(sample)/∂ξ_k from M_k = −H⁻¹ G_k; the matrix Σ wt J_aᵀ Q J_a from the four
// scalars g_xᵀ Q g_x, g_xᵀ Q g_y, g_yᵀ Q g_x, g_yᵀ Q g_y. Partials: 8 gradient, then the upper
// triangle row by row (36), FP64 across the block, at part[(y · gridDim.x + x) · 44].
__global__ __launch_bounds__(256) void k_tangent_jobs(const TangentJob* __restrict__ jobs, double* part) {
  const TangentJob& j = jobs[blockIdx.y];
  const int wa = j.wa, ha = j.ha, wb = j.wb;
  const int64_t na = (int64_t)wa * ha, nb = (int64_t)wb * j.hb;
  const float sw = (float)wa, sh = (float)ha;
  float acc[44];
#pragma unroll
  for (int m = 0; m < 44; ++m) acc[m] = 0.0f;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < nb; t += (int64_t)gridDim.x * 256) {
    const int x = (int)(t % wb), y = (int)(t / wb);
    const float X0 = (float)x + 1.0f, X1 = (float)y + 1.0f;
    // the sample position in FP64: the bilinear gradients jump at cell edges, so the cell must
    // not depend on the rounding of the position
    const double pxd = j.Hi[0] * (double)X0 + j.Hi[1] * (double)X1 + j.Hi[2];
    const double pyd = j.Hi[3] * (double)X0 + j.Hi[4] * (double)X1 + j.Hi[5];
    const double pzd = j.Hi[6] * (double)X0 + j.Hi[7] * (double)X1 + j.Hi[8];
    if (fabs(pzd) < 1e-8) continue;
    const double fxd = pxd / pzd - 1.0, fyd = pyd / pzd - 1.0;
    if (!(fxd >= 2.0 && fyd >= 2.0 && fxd <= wa - 3.0 && fyd <= ha - 3.0)) continue;   // NaN fails too
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
    const int64_t i00 = (int64_t)y0 * wa + x0, i01 = i00 + wa;
    float va[4], gx[4], gy[4], vb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float* f = j.fa + q * na;
      const float v00 = f[i00], v10 = f[i00 + 1], v01 = f[i01], v11 = f[i01 + 1];
      va[q] = (v00 * (1.0f - ax) + v10 * ax) * (1.0f - ay) + (v01 * (1.0f - ax) + v11 * ax) * ay;
      gx[q] = (v10 - v00) * (1.0f - ay) + (v11 - v01) * ay;
      gy[q] = (v01 - v00) * (1.0f - ax) + (v11 - v10) * ax;
      vb[q] = j.fb[q * nb + t];
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
  double accd[44];
#pragma unroll
  for (int m = 0; m < 44; ++m) accd[m] = (double)acc[m];
  block_sum<double, 44, 256>(accd, part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 44, threadIdx.x);
}


int moments_blocks(int w, int h) {
  const int64_t n = (int64_t)w * h;
  return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 256 * 64 - 1) / (256 * 64)));
}

void launch_pose_moments(const float* feat_a, int wa, int ha, const float* feat_b, int wb, int hb, const PoseArgs& p,
                         double* part, hipStream_t st) {
  k_pose_moments<<<moments_blocks(wb, hb), 256, 0, st>>>(feat_a, wa, ha, feat_b, wb, hb, p, part);
}

int jobs_blocks(int64_t max_pixels) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(256, (max_pixels + 256 * 32 - 1) / (256 * 32)));
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
