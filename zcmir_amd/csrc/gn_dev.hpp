// Device-side FP64 pose algebra and score coefficients: the same arithmetic as the host's
// geometry.hpp / lie.hpp (lie.zig:22-376) and score_coefs (moments.zig:91-172), on plain arrays,
// for the Gauss–Newton batches that run their iterations without the host. Matrices are row-major;
// poses map moving pixels to fixed pixels (1-based).
#pragma once
#include <hip/hip_runtime.h>

namespace zc {
namespace gd {

__device__ __forceinline__ void mul3(const double* a, const double* b, double* o) {
  double t[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) t[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
#pragma unroll
  for (int k = 0; k < 9; ++k) o[k] = t[k];
}

// the adjugate inverse with one Newton refinement step (geometry.hpp inv3)
__device__ __forceinline__ void inv3(const double* m, double* out) {
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  const double A = e * i - f * h, B = f * g - d * i, C = d * h - e * g;
  const double D = c * h - b * i, E = a * i - c * g, F = b * g - a * h;
  const double G = b * f - c * e, Hh = c * d - a * f, I = a * e - b * d;
  const double det = a * A + b * B + c * C;
  const double sg = det > 0 ? 1.0 : (det < 0 ? -1.0 : 0.0);
  const double s = 1.0 / (sg * fmax(fabs(det), 1e-18));
  const double inv[9] = {A * s, D * s, G * s, B * s, E * s, Hh * s, C * s, F * s, I * s};
  double p[9];
  mul3(m, inv, p);
  const double corr[9] = {2 - p[0], -p[1], -p[2], -p[3], 2 - p[4], -p[5], -p[6], -p[7], 2 - p[8]};
  mul3(inv, corr, out);
}

__device__ __forceinline__ void dest_n(double w0, double h0, double* N) {
  const double w = fmax(w0, 1.0), h = fmax(h0, 1.0);
  const double sx = 0.5 * w, sy = 0.5 * h, cx = 0.5 * (w + 1), cy = 0.5 * (h + 1);
  N[0] = 1 / sx; N[1] = 0; N[2] = -cx / sx;
  N[3] = 0; N[4] = 1 / sy; N[5] = -cy / sy;
  N[6] = 0; N[7] = 0; N[8] = 1;
}

__device__ __forceinline__ void hat(const double* x, bool homography, double* A) {
  double px = x[6], py = x[7], iso, z;
  if (!homography) {
    px = py = 0;
    z = 0;
    iso = x[3];
  } else {
    iso = x[3] / 3;
    z = -2 * iso;
  }
  A[0] = iso + x[4]; A[1] = x[5] - x[2]; A[2] = x[0];
  A[3] = x[5] + x[2]; A[4] = iso - x[4]; A[5] = x[1];
  A[6] = px; A[7] = py; A[8] = z;
}

// scaling and squaring with a 13-term Taylor series (lie.hpp expm3)
__device__ __forceinline__ void expm3(const double* A, double* T) {
  double nrm = 0;
  for (int i = 0; i < 9; ++i) nrm = fmax(nrm, fabs(A[i]));
  int s = 0;
  if (nrm > 0.5) s = (int)fmin(10.0, fmax(0.0, ceil(log2(nrm / 0.5))));
  const double sc = pow(0.5, (double)s);
  double B[9], P[9];
  for (int i = 0; i < 9; ++i) B[i] = A[i] * sc;
  for (int i = 0; i < 9; ++i) T[i] = P[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int k = 1; k < 14; ++k) {
    mul3(P, B, P);
    for (int i = 0; i < 9; ++i) P[i] /= (double)k;
    for (int i = 0; i < 9; ++i) T[i] += P[i];
  }
  for (int i = 0; i < s; ++i) mul3(T, T, T);
}

__device__ __forceinline__ void project_group(double* H, bool homography) {
  const double s = fabs(H[8]) > 1e-12 ? H[8] : 1;
  for (int i = 0; i < 9; ++i) H[i] = H[i] / s;
  if (!homography) {
    H[6] = 0;
    H[7] = 0;
    H[8] = 1;
  }
}

// exp(x̂) in normalized coordinates of a dw × dh destination, applied to H0 (lie.hpp compose_n)
__device__ __forceinline__ void compose_n(const double* H0, const double* x, bool homography, double dw, double dh,
                                          double* out) {
  double N[9], Ni[9], A[9], E[9], t[9];
  dest_n(dw, dh, N);
  inv3(N, Ni);
  hat(x, homography, A);
  expm3(A, E);
  mul3(Ni, E, t);
  mul3(t, N, t);
  mul3(t, H0, out);
  project_group(out, homography);
}

__device__ __forceinline__ void vee(const double* A0, bool homography, double* out) {
  double a = A0[0], e = A0[4];
  const double b = A0[1], c = A0[2], d = A0[3], f = A0[5], g = A0[6], h = A0[7], i = A0[8];
  if (homography) {
    const double tr = (a + e + i) / 3;
    a -= tr;
    e -= tr;
    const double iso = (a + e) / 2;
    out[0] = c; out[1] = f; out[2] = (d - b) / 2; out[3] = iso * 3; out[4] = (a - e) / 2; out[5] = (d + b) / 2;
    out[6] = g; out[7] = h;
    return;
  }
  out[0] = c; out[1] = f; out[2] = (d - b) / 2; out[3] = (a + e) / 2; out[4] = (a - e) / 2; out[5] = (d + b) / 2;
  out[6] = 0; out[7] = 0;
}

// generator k (pixel coordinates of a dw × dh destination when dest) (lie.hpp generators)
__device__ __forceinline__ void generator(int k, bool homography, bool dest, double dw, double dh, double* G) {
  double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  x[k] = 1;
  hat(x, homography, G);
  if (dest) {
    double N[9], Ni[9];
    dest_n(dw, dh, N);
    inv3(N, Ni);
    mul3(Ni, G, G);
    mul3(G, N, G);
  }
}

// the adjoint map's matrix: column j = dS/dξ from e_j on H⁻¹ (lie.hpp adjoint_grad, linear in g)
__device__ __forceinline__ void adjoint_cols(const double* H, bool homography, double mw, double mh, double fw,
                                             double fh, double (&PT)[8][8]) {
  double Hi[9], Nm[9], Nf[9], Nfi[9], M[9], Mi[9], t[9];
  inv3(H, Hi);
  dest_n(mw, mh, Nm);
  dest_n(fw, fh, Nf);
  inv3(Nf, Nfi);
  mul3(Nm, Hi, t);
  mul3(t, Nfi, M);
  inv3(M, Mi);
  const int nk = homography ? 8 : 6;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) PT[i][j] = 0;
  for (int i = 0; i < nk; ++i) {
    double G[9], cc[8];
    generator(i, homography, false, 0, 0, G);
    mul3(M, G, t);
    mul3(t, Mi, t);
    vee(t, homography, cc);
    for (int j = 0; j < nk; ++j) PT[i][j] = -cc[j];   // out[i] = Σ_j g[j] PT[i][j]
  }
}

// A x = b (n × n row-major), Gauss–Jordan with partial pivoting (lie.hpp solve_linear)
__device__ __forceinline__ bool solve_linear(const double* A, const double* b, int n, double* x) {
  double M[64];
  for (int i = 0; i < n * n; ++i) M[i] = A[i];
  for (int i = 0; i < n; ++i) x[i] = b[i];
  for (int i = 0; i < n; ++i) {
    int piv = i;
    for (int r = i + 1; r < n; ++r)
      if (fabs(M[r * n + i]) > fabs(M[piv * n + i])) piv = r;
    if (piv != i) {
      for (int j = 0; j < n; ++j) {
        const double t = M[i * n + j];
        M[i * n + j] = M[piv * n + j];
        M[piv * n + j] = t;
      }
      const double t = x[i];
      x[i] = x[piv];
      x[piv] = t;
    }
    const double d = M[i * n + i];
    if (!isfinite(d) || fabs(d) < 1e-300) return false;
    for (int r = 0; r < n; ++r) {
      if (r == i) continue;
      const double f = M[r * n + i] / d;
      if (f == 0) continue;
      for (int j = i; j < n; ++j) M[r * n + j] -= f * M[i * n + j];
      x[r] -= f * x[i];
    }
  }
  for (int i = 0; i < n; ++i) {
    x[i] /= M[i * n + i];
    if (!isfinite(x[i])) return false;
  }
  return true;
}

// solve_linear by a workgroup of at least n · n threads (t: the thread's index) on M (n × n) and
// x (n) in LDS, in place: the same operations on every element as the one-thread form above, with
// the pivot row found by thread 0 and the rows' factors taken before any element changes; the
// result is uniform over the workgroup (piv, f: n + 1 ints and n doubles of LDS scratch).
__device__ __forceinline__ bool solve_linear_wg(double* M, double* x, int n, int t, int* piv, double* f) {
  const int r = t / n, j = t - (t / n) * n;
  for (int i = 0; i < n; ++i) {
    if (t == 0) {
      int p = i;
      for (int q = i + 1; q < n; ++q)
        if (fabs(M[q * n + i]) > fabs(M[p * n + i])) p = q;
      *piv = p;
    }
    __syncthreads();
    const int p = *piv;
    if (p != i) {
      if (t < n) {
        const double v = M[i * n + t];
        M[i * n + t] = M[p * n + t];
        M[p * n + t] = v;
      }
      if (t == 0) {
        const double v = x[i];
        x[i] = x[p];
        x[p] = v;
      }
    }
    __syncthreads();
    const double d = M[i * n + i];
    if (!isfinite(d) || fabs(d) < 1e-300) return false;
    if (t < n) f[t] = t == i ? 0.0 : M[t * n + i] / d;
    __syncthreads();
    if (r < n && r != i && j >= i && f[r] != 0) M[r * n + j] -= f[r] * M[i * n + j];
    if (t < n && t != i && f[t] != 0) x[t] -= f[t] * x[i];
    __syncthreads();
  }
  if (t < n) {
    x[t] /= M[t * n + t];
    piv[1 + t] = isfinite(x[t]) ? 1 : 0;
  }
  __syncthreads();
  bool ok = true;
  for (int q = 0; q < n; ++q) ok = ok && piv[1 + q];
  return ok;
}

__device__ __forceinline__ bool beats(double nw, double old) { return nw > old + fmax(1e-6 * fabs(old), 1e-12); }

__device__ __forceinline__ int tri(int i, int j) {
  const int a = i < j ? i : j, b = i < j ? j : i;
  return a * 4 - (a * (a + 1)) / 2 + b;
}

__device__ __forceinline__ bool chol4(const double* G, double* L) {
  for (int i = 0; i < 16; ++i) L[i] = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = G[i * 4 + j];
      for (int k = 0; k < j; ++k) s -= L[i * 4 + k] * L[j * 4 + k];
      if (i == j) {
        if (!(s > 0)) return false;
        L[i * 5] = sqrt(s);
      } else {
        L[i * 4 + j] = s / L[j * 5];
      }
    }
  return true;
}

__device__ __forceinline__ void chol_inv(const double* L, double* X) {
  for (int c = 0; c < 4; ++c) {
    double y[4];
    for (int i = 0; i < 4; ++i) {
      double s = i == c ? 1 : 0;
      for (int k = 0; k < i; ++k) s -= L[i * 4 + k] * y[k];
      y[i] = s / L[i * 5];
    }
    for (int i = 3; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < 4; ++k) s -= L[k * 4 + i] * X[k * 4 + c];
      X[i * 4 + c] = s / L[i * 5];
    }
  }
}

__device__ __forceinline__ void mm4(const double* A, const double* B, double* o) {
  double t[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 4 + j];
      t[i * 4 + j] = s;
    }
  for (int k = 0; k < 16; ++k) o[k] = t[k];
}

__device__ __forceinline__ void tr4(const double* A, double* o) {
  double t[16];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t[i * 4 + j] = A[j * 4 + i];
  for (int k = 0; k < 16; ++k) o[k] = t[k];
}

// The score's coefficients from the 45 sums (score_coefs in gn.cpp, moments.zig:91-172).
struct DCoefs {
  double score, n;
  double W[16], Mh[16], Q[16], ma[4], mb[4], g[45];
};

__device__ __forceinline__ void score_coefs(const double* st, bool exact, double ridge, int min_n, DCoefs& o) {
  o.score = 0;
  o.n = st[0];
  for (int i = 0; i < 16; ++i) o.W[i] = o.Mh[i] = o.Q[i] = 0;
  for (int i = 0; i < 4; ++i) o.ma[i] = o.mb[i] = 0;
  for (int i = 0; i < 45; ++i) o.g[i] = 0;
  const double n = st[0];
  if (!(n >= min_n)) return;
  double ma[4], mb[4], C[16];
  for (int i = 0; i < 4; ++i) {
    ma[i] = st[1 + i] / n;
    mb[i] = st[5 + i] / n;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) C[i * 4 + j] = st[9 + i * 4 + j] / n - ma[i] * mb[j];
  double W[16], Mh[16], Nh[16];
  for (int i = 0; i < 16; ++i) {
    W[i] = C[i];
    Mh[i] = Nh[i] = 0;
  }
  double phi = 0;
  if (!exact) {
    for (int i = 0; i < 16; ++i) phi += C[i] * C[i];
  } else {
    double GA[16], GB[16];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        GA[i * 4 + j] = st[25 + tri(i, j)] / n - ma[i] * ma[j];
        GB[i * 4 + j] = st[35 + tri(i, j)] / n - mb[i] * mb[j];
      }
    const double ra = ridge * fmax((GA[0] + GA[5] + GA[10] + GA[15]) / 4, 1e-12);
    const double rb = ridge * fmax((GB[0] + GB[5] + GB[10] + GB[15]) / 4, 1e-12);
    for (int i = 0; i < 4; ++i) {
      GA[i * 5] += ra;
      GB[i * 5] += rb;
    }
    double LA[16], LB[16], GAi[16], GBi[16], t[16], Ct[16];
    if (!chol4(GA, LA) || !chol4(GB, LB)) return;
    chol_inv(LA, GAi);
    chol_inv(LB, GBi);
    mm4(GAi, C, t);
    mm4(t, GBi, W);
    for (int i = 0; i < 16; ++i) phi += W[i] * C[i];
    tr4(C, Ct);
    mm4(W, Ct, t);
    mm4(t, GAi, Mh);
    double Wt[16];
    tr4(W, Wt);
    mm4(Wt, C, t);
    mm4(t, GBi, Nh);
    const double tm = (Mh[0] + Mh[5] + Mh[10] + Mh[15]) * ridge / 4;
    const double tn = (Nh[0] + Nh[5] + Nh[10] + Nh[15]) * ridge / 4;
    for (int i = 0; i < 4; ++i) {
      Mh[i * 5] += tm;
      Nh[i * 5] += tn;
    }
  }
  const double score = n * phi;
  if (!isfinite(score)) return;
  o.score = score;
  for (int i = 0; i < 16; ++i) {
    o.W[i] = W[i];
    o.Mh[i] = Mh[i];
  }
  for (int i = 0; i < 4; ++i) {
    o.ma[i] = ma[i];
    o.mb[i] = mb[i];
  }
  double* g = o.g;
  for (int i = 0; i < 4; ++i) {
    double ga = 0, gb = 0;
    for (int j = 0; j < 4; ++j) {
      ga += -2 * W[i * 4 + j] * mb[j] + 2 * Mh[i * 4 + j] * ma[j];
      gb += -2 * W[j * 4 + i] * ma[j] + 2 * Nh[i * 4 + j] * mb[j];
      g[9 + i * 4 + j] = 2 * W[i * 4 + j];
    }
    g[1 + i] = ga;
    g[5 + i] = gb;
    for (int j = i; j < 4; ++j) {
      const double f = i == j ? 1 : 2;
      g[25 + tri(i, j)] = -f * Mh[i * 4 + j];
      g[35 + tri(i, j)] = -f * Nh[i * 4 + j];
    }
  }
  double rest = 0;
  for (int k = 1; k < 45; ++k) rest += g[k] * st[k];
  g[0] = (score - rest) / n;
  if (exact) {
    for (int i = 0; i < 16; ++i) o.Q[i] = Mh[i];
  } else {
    double Ct[16];
    tr4(C, Ct);
    mm4(C, Ct, o.Q);
  }
}

}  // namespace gd
}  // namespace zc
