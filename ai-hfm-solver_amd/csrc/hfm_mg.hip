// Config-mode multigrid kernels for gfx950: weighted-Jacobi sweeps on the
// face-coefficient 7-point operator, fused residual + restriction, dense
// coarsest solve, FP64 flexible-CG passes and per-instance reductions.
//
// Every kernel is driven by a work table (hfm_mg_task[] + per-block task
// index), so one launch covers any mix of instances of different sizes.
// Memory-bound: one thread per cell (per coarse cell for the restriction),
// flattened row-major chunks of kChunk cells per block, all neighbour loads
// from L1/L2, chunks of a task spread contiguously over the 8 XCDs.
#include <math.h>

#include "hfm_common.h"

namespace hfm {

constexpr int kCPT = 4;                 // cells per thread
constexpr int kChunk = kBlock * kCPT;   // cells per block

template <typename C>
__device__ __forceinline__ float ldc(const void* p, int64_t i) {
  return to_f(static_cast<const C*>(p)[i]);
}

// Off-diagonal sum and diagonal of the face operator at cell (k, j, i), index n.
// xat(m, kk, jj, ii) returns x at a neighbour.  Accumulation in type A.
template <typename A, typename C, typename XF>
__device__ __forceinline__ A face_row(const hfm_mg_task& t, int n, int i, int j, int k, XF xat,
                                      A& diag) {
  const int W = t.W, H = t.H, D = t.D, HW = H * W;
  const A cw = (A)ldc<C>(t.c[0], n);
  const A ce = (A)(i < W - 1 ? ldc<C>(t.c[0], n + 1) : ldc<C>(t.c[3], k * H + j));
  const A cs = (A)ldc<C>(t.c[1], n);
  const A cn = (A)(j < H - 1 ? ldc<C>(t.c[1], n + W) : ldc<C>(t.c[4], k * W + i));
  const A cd = (A)ldc<C>(t.c[2], n);
  const A cu = (A)(k < D - 1 ? ldc<C>(t.c[2], n + HW) : ldc<C>(t.c[5], j * W + i));
  diag = cw + ce + cs + cn + cd + cu;
  A off = (A)0;
  if (i > 0) off += cw * (A)xat(n - 1, k, j, i - 1);
  if (i < W - 1) off += ce * (A)xat(n + 1, k, j, i + 1);
  if (j > 0) off += cs * (A)xat(n - W, k, j - 1, i);
  if (j < H - 1) off += cn * (A)xat(n + W, k, j + 1, i);
  if (k > 0) off += cd * (A)xat(n - HW, k - 1, j, i);
  if (k < D - 1) off += cu * (A)xat(n + HW, k + 1, j, i);
  return off;
}

__device__ __forceinline__ void decompose(const hfm_mg_task& t, int n, int& i, int& j, int& k) {
  const uint32_t row = fdiv(uint32_t(n), t.magW);
  i = n - int(row) * t.W;
  k = int(fdiv(row, t.magH));
  j = int(row) - k * t.H;
}

// ------------------------------------------------------------------ sweeps
template <typename T, int MODE, bool DOT>
__global__ __launch_bounds__(kBlock) void k_sweep(const hfm_mg_task* __restrict__ tasks,
                                                  const uint16_t* __restrict__ blk_task,
                                                  const hfm_mg_scalars* __restrict__ sc,
                                                  float omega, float alpha, double* dots) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  if (!sc[t.inst].active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kChunk;
  const T* xin = static_cast<const T*>(t.xin);
  const T* b = static_cast<const T*>(t.b);
  const T* ec = static_cast<const T*>(t.ec);
  T* out = static_cast<T*>(t.xout);
  double acc = 0.0;
  if (base < N) {
    for (int s = 0; s < kCPT; ++s) {
      const int n = base + s * kBlock + threadIdx.x;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      const float bn = to_f(b[n]);
      float r;
      if (MODE == HFM_SWEEP_ZERO) {
        // diagonal only: sum of the six faces
        float diag;
        face_row<float, T>(t, n, i, j, k, [](int, int, int, int) { return 0.0f; }, diag);
        r = omega * bn / diag;
      } else if (MODE == HFM_SWEEP_PLAIN) {
        auto xat = [&](int m, int, int, int) { return to_f(xin[m]); };
        float diag;
        const float off = face_row<float, T>(t, n, i, j, k, xat, diag);
        const float xc = to_f(xin[n]);
        r = xc + omega * (bn - (diag * xc - off)) / diag;
      } else {  // CORRECT: x := xin + alpha * P ec on the fly
        const int Hc = t.Hc, Wc = t.Wc;
        auto xat = [&](int m, int kk, int jj, int ii) {
          return to_f(xin[m]) + alpha * to_f(ec[((kk >> 1) * Hc + (jj >> 1)) * Wc + (ii >> 1)]);
        };
        float diag;
        const float off = face_row<float, T>(t, n, i, j, k, xat, diag);
        const float xc = xat(n, k, j, i);
        r = xc + omega * (bn - (diag * xc - off)) / diag;
      }
      const T rs = from_f<T>(r);
      out[n] = rs;
      if (DOT) acc += double(bn) * double(to_f(rs));
    }
  }
  if (DOT) {
    const double s = block_sum(acc);
    if (threadIdx.x == 0) dots[t.slot0 + chunk] = s;
  }
}

// ------------------------------------------------------------------ residual + restriction
// One thread per coarse cell: bc = sum over children (b - A x) / 8.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_resid_restrict(const hfm_mg_task* __restrict__ tasks,
                                                           const uint16_t* __restrict__ blk_task,
                                                           const hfm_mg_scalars* __restrict__ sc) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  if (!sc[t.inst].active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int Nc = t.Dc * t.Hc * t.Wc;
  const int base = chunk * kChunk;
  if (base >= Nc) return;
  const T* x = static_cast<const T*>(t.xin);
  const T* b = static_cast<const T*>(t.b);
  T* bc = static_cast<T*>(t.xout);
  auto xat = [&](int m, int, int, int) { return to_f(x[m]); };
  const int W = t.W, H = t.H, D = t.D;
  for (int s = 0; s < kCPT; ++s) {
    const int nc = base + s * kBlock + threadIdx.x;
    if (nc >= Nc) break;
    const uint32_t rowc = fdiv(uint32_t(nc), t.magWc);
    const int I = nc - int(rowc) * t.Wc;
    const int K = int(fdiv(rowc, t.magHc));
    const int J = int(rowc) - K * t.Hc;
    float sum = 0.0f;
    for (int a = 0; a < 2; ++a) {
      const int k = 2 * K + a;
      if (k >= D) break;
      for (int bb = 0; bb < 2; ++bb) {
        const int j = 2 * J + bb;
        if (j >= H) break;
        for (int c = 0; c < 2; ++c) {
          const int i = 2 * I + c;
          if (i >= W) break;
          const int n = (k * H + j) * W + i;
          float diag;
          const float off = face_row<float, T>(t, n, i, j, k, xat, diag);
          sum += to_f(b[n]) - (diag * to_f(x[n]) - off);
        }
      }
    }
    bc[nc] = from_f<T>(sum * 0.125f);
  }
}

// ------------------------------------------------------------------ coarsest solve
template <typename T>
__global__ __launch_bounds__(64) void k_coarse_solve(const hfm_mg_task* __restrict__ tasks,
                                                     const hfm_mg_scalars* __restrict__ sc) {
  const hfm_mg_task& t = tasks[blockIdx.x];
  if (!sc[t.inst].active) return;
  const int N = t.D * t.H * t.W;
  if (N > 64) return;  // host rejects such levels; never index past the LDS tile
  const double* ainv = static_cast<const double*>(t.ec);
  const T* b = static_cast<const T*>(t.b);
  __shared__ double bs[64];
  if (threadIdx.x < N) bs[threadIdx.x] = double(to_f(b[threadIdx.x]));
  __syncthreads();
  const int i = threadIdx.x;
  if (i < N) {
    double e = 0.0;
    for (int j = 0; j < N; ++j) e += ainv[j * N + i] * bs[j];
    static_cast<T*>(t.xout)[i] = from_f<T>(float(e));
  }
}

// ------------------------------------------------------------------ FCG passes (fine level)
// Coefficients c are the FP32 fine faces; products accumulated in FP64.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fcg_init(const hfm_mg_task* __restrict__ tasks,
                                                     const uint16_t* __restrict__ blk_task,
                                                     const hfm_mg_scalars* __restrict__ sc,
                                                     double* dots) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  if (!sc[t.inst].active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kChunk;
  double acc = 0.0;
  T* rlow = static_cast<T*>(t.rlow);
  if (base < N) {
    for (int s = 0; s < kCPT; ++s) {
      const int n = base + s * kBlock + threadIdx.x;
      if (n >= N) break;
      const double bv = double(t.b32[n]);
      t.x64[n] = 0.0;
      t.r64[n] = bv;
      rlow[n] = from_f<T>(float(bv));
      acc += bv * bv;
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) dots[t.slot0 + chunk] = s;
}

// p = T(z + beta p_old) (p_old not read when beta == 0); dots = p . (A p)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fcg_dir(const hfm_mg_task* __restrict__ tasks,
                                                    const uint16_t* __restrict__ blk_task,
                                                    const hfm_mg_scalars* __restrict__ sc,
                                                    double* dots) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (!S.active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kChunk;
  const double beta = S.beta;
  const T* z = static_cast<const T*>(t.b);
  const T* pold = static_cast<const T*>(t.xin);
  T* pnew = static_cast<T*>(t.xout);
  auto pat = [&](int m, int, int, int) -> double {
    const double v = beta == 0.0 ? double(to_f(z[m])) : double(to_f(z[m])) + beta * double(to_f(pold[m]));
    return double(round_to<T>(float(v)));
  };
  double acc = 0.0;
  if (base < N) {
    for (int s = 0; s < kCPT; ++s) {
      const int n = base + s * kBlock + threadIdx.x;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      double diag;
      const double off = face_row<double, float>(t, n, i, j, k, pat, diag);
      const double p = pat(n, k, j, i);
      pnew[n] = from_f<T>(float(p));
      acc += p * (diag * p - off);
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) dots[t.slot0 + chunk] = s;
}

// x64 += alpha p; r64 -= alpha A p; rlow = T(r64); dots = r . r
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fcg_update(const hfm_mg_task* __restrict__ tasks,
                                                       const uint16_t* __restrict__ blk_task,
                                                       const hfm_mg_scalars* __restrict__ sc,
                                                       double* dots) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (!S.active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kChunk;
  const double alpha = S.alpha;
  const T* p = static_cast<const T*>(t.xin);
  T* rlow = static_cast<T*>(t.rlow);
  auto pat = [&](int m, int, int, int) -> double { return double(to_f(p[m])); };
  double acc = 0.0;
  if (base < N) {
    for (int s = 0; s < kCPT; ++s) {
      const int n = base + s * kBlock + threadIdx.x;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      double diag;
      const double off = face_row<double, float>(t, n, i, j, k, pat, diag);
      const double pn = pat(n, k, j, i);
      const double q = diag * pn - off;
      t.x64[n] += alpha * pn;
      const double r = t.r64[n] - alpha * q;
      t.r64[n] = r;
      rlow[n] = from_f<T>(float(r));
      acc += r * r;
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) dots[t.slot0 + chunk] = s;
}

// ------------------------------------------------------------------ per-instance scalars
__global__ __launch_bounds__(kBlock) void k_fcg_finalize(hfm_mg_scalars* sc, const double* dots,
                                                         int kind) {
  hfm_mg_scalars& S = sc[blockIdx.x];
  if (!S.active) return;
  double acc = 0.0;
  for (int q = threadIdx.x; q < S.nslot; q += kBlock) acc += dots[S.slot0 + q];
  const double s = block_sum(acc);
  if (threadIdx.x != 0) return;
  switch (kind) {
    case HFM_FIN_INIT:
      S.bnorm2 = s;
      S.rr = s;
      S.rel = 1.0;
      S.iter = 0;
      S.rz = 0.0;
      S.beta = 0.0;
      if (s == 0.0) {  // x = 0 solves b = 0
        S.rel = 0.0;
        S.status = HFM_MG_CONVERGED;
        S.active = 0;
      } else if (!isfinite(s)) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
      }
      break;
    case HFM_FIN_BETA:
      S.beta = (S.iter == 0) ? 0.0 : s / S.rz;
      S.rz = s;
      if (!isfinite(s)) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
      }
      break;
    case HFM_FIN_ALPHA:
      S.pap = s;
      if (!(s > 0.0) || !isfinite(s)) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
      } else {
        S.alpha = S.rz / s;
      }
      break;
    case HFM_FIN_CHECK:
      S.rr = s;
      S.rel = sqrt(s / S.bnorm2);
      S.iter += 1;
      if (S.rel <= S.tol) {
        S.status = HFM_MG_CONVERGED;
        S.active = 0;
      } else if (!isfinite(S.rel)) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
      } else if (S.iter >= S.maxit) {
        S.status = HFM_MG_MAXIT;
        S.active = 0;
      }
      break;
  }
}

// ------------------------------------------------------------------ setup kernels
__global__ __launch_bounds__(kBlock) void k_faces(const float* __restrict__ sx,
                                                  const float* __restrict__ sy,
                                                  const float* __restrict__ sz, float* cx,
                                                  float* cy, float* cz, float* bxh, float* byh,
                                                  float* bzh, int D, int H, int W, float inv_h2) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t N = int64_t(D) * H * W;
  if (n >= N) return;
  const int i = int(n % W);
  const int j = int((n / W) % H);
  const int k = int(n / (int64_t(W) * H));
  const double s2 = double(inv_h2);
  auto hm = [](double a, double b) { return 2.0 * a * b / (a + b); };
  const double ax = sx[n], ay = sy[n], az = sz[n];
  cx[n] = float((i == 0 ? ax : hm(ax, double(sx[n - 1]))) * s2);
  cy[n] = float((j == 0 ? ay : hm(ay, double(sy[n - W]))) * s2);
  cz[n] = float((k == 0 ? az : hm(az, double(sz[n - int64_t(W) * H]))) * s2);
  if (i == W - 1) bxh[int64_t(k) * H + j] = float(ax * s2);
  if (j == H - 1) byh[int64_t(k) * W + i] = float(ay * s2);
  if (k == D - 1) bzh[int64_t(j) * W + i] = float(az * s2);
}

// Coarse faces: (pair over the first listed axis, then the second) / 8, in FP64.
__global__ __launch_bounds__(kBlock) void k_galerkin(const float* __restrict__ fcx,
                                                     const float* __restrict__ fcy,
                                                     const float* __restrict__ fcz,
                                                     const float* __restrict__ fbxh,
                                                     const float* __restrict__ fbyh,
                                                     const float* __restrict__ fbzh, float* cx,
                                                     float* cy, float* cz, float* bxh, float* byh,
                                                     float* bzh, int D, int H, int W) {
  const int Dc = (D + 1) / 2, Hc = (H + 1) / 2, Wc = (W + 1) / 2;
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t Nc = int64_t(Dc) * Hc * Wc;
  if (n >= Nc) return;
  const int I = int(n % Wc);
  const int J = int((n / Wc) % Hc);
  const int K = int(n / (int64_t(Wc) * Hc));
  auto F = [&](const float* a, int k, int j, int i) -> double {
    return (k < D && j < H && i < W) ? double(a[(int64_t(k) * H + j) * W + i]) : 0.0;
  };
  // x faces at fine i = 2I: sum k pairs, then j pairs
  cx[n] = float(((F(fcx, 2 * K, 2 * J, 2 * I) + F(fcx, 2 * K + 1, 2 * J, 2 * I)) +
                 (F(fcx, 2 * K, 2 * J + 1, 2 * I) + F(fcx, 2 * K + 1, 2 * J + 1, 2 * I))) /
                8.0);
  // y faces at fine j = 2J: sum k pairs, then i pairs
  cy[n] = float(((F(fcy, 2 * K, 2 * J, 2 * I) + F(fcy, 2 * K + 1, 2 * J, 2 * I)) +
                 (F(fcy, 2 * K, 2 * J, 2 * I + 1) + F(fcy, 2 * K + 1, 2 * J, 2 * I + 1))) /
                8.0);
  // z faces at fine k = 2K: sum j pairs, then i pairs
  cz[n] = float(((F(fcz, 2 * K, 2 * J, 2 * I) + F(fcz, 2 * K, 2 * J + 1, 2 * I)) +
                 (F(fcz, 2 * K, 2 * J, 2 * I + 1) + F(fcz, 2 * K, 2 * J + 1, 2 * I + 1))) /
                8.0);
  auto P = [](const float* a, int r, int c, int R, int Cn) -> double {
    return (r < R && c < Cn) ? double(a[int64_t(r) * Cn + c]) : 0.0;
  };
  if (I == Wc - 1)  // bxh (D,H): k pairs, then j pairs
    bxh[int64_t(K) * Hc + J] =
        float(((P(fbxh, 2 * K, 2 * J, D, H) + P(fbxh, 2 * K + 1, 2 * J, D, H)) +
               (P(fbxh, 2 * K, 2 * J + 1, D, H) + P(fbxh, 2 * K + 1, 2 * J + 1, D, H))) /
              8.0);
  if (J == Hc - 1)  // byh (D,W): k pairs, then i pairs
    byh[int64_t(K) * Wc + I] =
        float(((P(fbyh, 2 * K, 2 * I, D, W) + P(fbyh, 2 * K + 1, 2 * I, D, W)) +
               (P(fbyh, 2 * K, 2 * I + 1, D, W) + P(fbyh, 2 * K + 1, 2 * I + 1, D, W))) /
              8.0);
  if (K == Dc - 1)  // bzh (H,W): j pairs, then i pairs
    bzh[int64_t(J) * Wc + I] =
        float(((P(fbzh, 2 * J, 2 * I, H, W) + P(fbzh, 2 * J + 1, 2 * I, H, W)) +
               (P(fbzh, 2 * J, 2 * I + 1, H, W) + P(fbzh, 2 * J + 1, 2 * I + 1, H, W))) /
              8.0);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_convert(const float* __restrict__ src, T* dst,
                                                    int64_t n) {
  const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (q < n) dst[q] = from_f<T>(src[q]);
}

// Dense FP64 inverse of the coarsest face operator by Gauss-Jordan elimination
// (symmetric positive definite M-matrix: no pivoting).  One block per task.
__global__ __launch_bounds__(kBlock) void k_coarse_inverse(const hfm_mg_task* __restrict__ tasks) {
  const hfm_mg_task& t = tasks[blockIdx.x];
  const int N = t.D * t.H * t.W;
  if (N > 64) return;  // host rejects such levels; never index past the LDS tile
  __shared__ double M[64][129];
  __shared__ double prow[128];
  __shared__ double fac[64];
  const int N2 = 2 * N;
  for (int e = threadIdx.x; e < N * N2; e += kBlock) {
    const int r = e / N2, c = e % N2;
    M[r][c] = (c >= N) ? (c - N == r ? 1.0 : 0.0) : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int n = threadIdx.x;
    const int i = n % t.W, j = (n / t.W) % t.H, k = n / (t.W * t.H);
    // scatter one row of the operator
    auto xat = [](int, int, int, int) { return 0.0; };
    double diag;
    face_row<double, float>(t, n, i, j, k, xat, diag);
    M[n][n] = diag;
    const int W = t.W, H = t.H, D = t.D, HW = H * W;
    const float* cx = static_cast<const float*>(t.c[0]);
    const float* cy = static_cast<const float*>(t.c[1]);
    const float* cz = static_cast<const float*>(t.c[2]);
    if (i > 0) M[n][n - 1] = -double(cx[n]);
    if (i < W - 1) M[n][n + 1] = -double(cx[n + 1]);
    if (j > 0) M[n][n - W] = -double(cy[n]);
    if (j < H - 1) M[n][n + W] = -double(cy[n + W]);
    if (k > 0) M[n][n - HW] = -double(cz[n]);
    if (k < D - 1) M[n][n + HW] = -double(cz[n + HW]);
  }
  __syncthreads();
  for (int p = 0; p < N; ++p) {
    const double piv = M[p][p];
    for (int c = threadIdx.x; c < N2; c += kBlock) prow[c] = M[p][c] / piv;
    for (int r = threadIdx.x; r < N; r += kBlock) fac[r] = M[r][p];
    __syncthreads();
    for (int e = threadIdx.x; e < N * N2; e += kBlock) {
      const int r = e / N2, c = e % N2;
      M[r][c] = (r == p) ? prow[c] : M[r][c] - fac[r] * prow[c];
    }
    __syncthreads();
  }
  double* out = static_cast<double*>(t.xout);
  for (int e = threadIdx.x; e < N * N; e += kBlock) out[e] = M[e / N][N + e % N];
}

}  // namespace hfm

// ====================================================================== C ABI
using namespace hfm;

namespace {

template <typename F>
int launch_store(int store, const char* what, F&& f) {
  if (store == HFM_STORE_F32) {
    f(float{});
  } else if (store == HFM_STORE_BF16) {
    f(bf16{});
  } else {
    set_error("%s: unknown storage type %d", what, store);
    return HFM_EINVAL;
  }
  return launch_status(what);
}

int check_table(const void* tasks, const void* blk_task, int64_t nblocks, const void* sc,
                const char* what) {
  if (!tasks || !sc || (nblocks > 0 && !blk_task)) {
    set_error("%s: null work table", what);
    return HFM_EINVAL;
  }
  if (nblocks < 0 || nblocks > INT32_MAX) {
    set_error("%s: bad block count %lld", what, (long long)nblocks);
    return HFM_EINVAL;
  }
  return HFM_OK;
}

}  // namespace

extern "C" int hfm_mg_sweep(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                            const hfm_mg_scalars* sc, int store, int mode, float omega,
                            float alpha, double* dots, void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_mg_sweep")) return e;
  if (mode < HFM_SWEEP_ZERO || mode > HFM_SWEEP_CORRECT) {
    set_error("hfm_mg_sweep: bad mode %d", mode);
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  hipStream_t s = as_stream(stream);
  return launch_store(store, "hfm_mg_sweep", [&](auto tag) {
    using T = decltype(tag);
    const dim3 grid{unsigned(nblocks)}, block{unsigned(kBlock)};
    const bool dot = dots != nullptr;
#define HFM_SWEEP_CASE(M)                                                                 \
  if (dot)                                                                                \
    hipLaunchKernelGGL((k_sweep<T, M, true>), grid, block, 0, s, tasks, blk_task, sc, omega, \
                       alpha, dots);                                                      \
  else                                                                                    \
    hipLaunchKernelGGL((k_sweep<T, M, false>), grid, block, 0, s, tasks, blk_task, sc,     \
                       omega, alpha, dots);
    if (mode == HFM_SWEEP_ZERO) {
      HFM_SWEEP_CASE(HFM_SWEEP_ZERO)
    } else if (mode == HFM_SWEEP_PLAIN) {
      HFM_SWEEP_CASE(HFM_SWEEP_PLAIN)
    } else {
      HFM_SWEEP_CASE(HFM_SWEEP_CORRECT)
    }
#undef HFM_SWEEP_CASE
  });
}

extern "C" int hfm_mg_resid_restrict(const hfm_mg_task* tasks, const uint16_t* blk_task,
                                     int64_t nblocks, const hfm_mg_scalars* sc, int store,
                                     void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_mg_resid_restrict")) return e;
  if (nblocks == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_resid_restrict", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_resid_restrict<T>, dim3(unsigned(nblocks)), dim3(kBlock), 0,
                       as_stream(stream), tasks, blk_task, sc);
  });
}

extern "C" int hfm_mg_coarse_solve(const hfm_mg_task* tasks, int64_t ntasks,
                                   const hfm_mg_scalars* sc, int store, void* stream) {
  if (!tasks || !sc || ntasks < 0 || ntasks > INT32_MAX) {
    set_error("hfm_mg_coarse_solve: bad arguments");
    return HFM_EINVAL;
  }
  if (ntasks == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_coarse_solve", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_coarse_solve<T>, dim3(unsigned(ntasks)), dim3(64), 0,
                       as_stream(stream), tasks, sc);
  });
}

#define HFM_FCG_ENTRY(NAME, KERNEL)                                                          \
  extern "C" int NAME(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,     \
                      const hfm_mg_scalars* sc, int store, double* dots, void* stream) {       \
    if (int e = check_table(tasks, blk_task, nblocks, sc, #NAME)) return e;                    \
    if (!dots) {                                                                               \
      set_error(#NAME ": null dot slots");                                                     \
      return HFM_EINVAL;                                                                       \
    }                                                                                          \
    if (nblocks == 0) return HFM_OK;                                                           \
    return launch_store(store, #NAME, [&](auto tag) {                                          \
      using T = decltype(tag);                                                                 \
      hipLaunchKernelGGL(KERNEL<T>, dim3(unsigned(nblocks)), dim3(kBlock), 0,                  \
                         as_stream(stream), tasks, blk_task, sc, dots);                        \
    });                                                                                        \
  }

HFM_FCG_ENTRY(hfm_fcg_init, k_fcg_init)
HFM_FCG_ENTRY(hfm_fcg_dir, k_fcg_dir)
HFM_FCG_ENTRY(hfm_fcg_update, k_fcg_update)
#undef HFM_FCG_ENTRY

extern "C" int hfm_fcg_finalize(hfm_mg_scalars* sc, int64_t ninst, const double* dots, int kind,
                                void* stream) {
  if (!sc || !dots || ninst < 0 || ninst > INT32_MAX || kind < HFM_FIN_INIT ||
      kind > HFM_FIN_CHECK) {
    set_error("hfm_fcg_finalize: bad arguments");
    return HFM_EINVAL;
  }
  if (ninst == 0) return HFM_OK;
  hipLaunchKernelGGL(k_fcg_finalize, dim3(unsigned(ninst)), dim3(kBlock), 0, as_stream(stream),
                     sc, dots, kind);
  return launch_status("hfm_fcg_finalize");
}

extern "C" int hfm_mg_faces(const float* sx, const float* sy, const float* sz,
                            float* const out[6], int64_t D, int64_t H, int64_t W, float inv_h2,
                            void* stream) {
  if (!sx || !sy || !sz || !out || D <= 0 || H <= 0 || W <= 0 || D * H * W > INT32_MAX) {
    set_error("hfm_mg_faces: bad arguments");
    return HFM_EINVAL;
  }
  for (int q = 0; q < 6; ++q)
    if (!out[q]) {
      set_error("hfm_mg_faces: null output %d", q);
      return HFM_EINVAL;
    }
  const int64_t N = D * H * W;
  hipLaunchKernelGGL(k_faces, dim3(unsigned((N + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     as_stream(stream), sx, sy, sz, out[0], out[1], out[2], out[3], out[4],
                     out[5], int(D), int(H), int(W), inv_h2);
  return launch_status("hfm_mg_faces");
}

extern "C" int hfm_mg_galerkin(const float* const fine[6], float* const coarse[6], int64_t D,
                               int64_t H, int64_t W, void* stream) {
  if (!fine || !coarse || D <= 0 || H <= 0 || W <= 0 || D * H * W > INT32_MAX) {
    set_error("hfm_mg_galerkin: bad arguments");
    return HFM_EINVAL;
  }
  for (int q = 0; q < 6; ++q)
    if (!fine[q] || !coarse[q]) {
      set_error("hfm_mg_galerkin: null field %d", q);
      return HFM_EINVAL;
    }
  const int64_t Nc = ((D + 1) / 2) * ((H + 1) / 2) * ((W + 1) / 2);
  hipLaunchKernelGGL(k_galerkin, dim3(unsigned((Nc + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     as_stream(stream), fine[0], fine[1], fine[2], fine[3], fine[4], fine[5],
                     coarse[0], coarse[1], coarse[2], coarse[3], coarse[4], coarse[5], int(D),
                     int(H), int(W));
  return launch_status("hfm_mg_galerkin");
}

extern "C" int hfm_mg_convert(const float* src, void* dst, int64_t n, int store, void* stream) {
  if (!src || !dst || n < 0) {
    set_error("hfm_mg_convert: bad arguments");
    return HFM_EINVAL;
  }
  if (n == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_convert", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_convert<T>, dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       as_stream(stream), src, static_cast<T*>(dst), n);
  });
}

extern "C" int hfm_mg_coarse_inverse(const hfm_mg_task* tasks, int64_t ntasks, void* stream) {
  if (!tasks || ntasks < 0 || ntasks > INT32_MAX) {
    set_error("hfm_mg_coarse_inverse: bad arguments");
    return HFM_EINVAL;
  }
  if (ntasks == 0) return HFM_OK;
  hipLaunchKernelGGL(k_coarse_inverse, dim3(unsigned(ntasks)), dim3(kBlock), 0,
                     as_stream(stream), tasks);
  return launch_status("hfm_mg_coarse_inverse");
}
