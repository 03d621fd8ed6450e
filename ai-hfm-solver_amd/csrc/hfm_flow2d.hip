// 2D periodic incompressible flow on a staggered grid for gfx950: explicit
// two-stage momentum stencils, divergence source, FP64 conjugate gradients
// preconditioned by an FP32 multigrid V-cycle, gradient correction.
//
// The grids are small (256^2 floats = 256 KB per field) and one step is a
// chain of ~40 dependent passes per CG iteration, so launches, not bytes,
// bound a kernel-per-pass design.  Here a team of 1024-thread workgroups owns
// a flow and runs whole steps: passes are separated by team barriers, every
// field stays in the L2 working set, and scalars (dots, CG coefficients, the
// stopping test) are reduced within the team in a fixed order, so every
// workgroup takes the same branches.  A team of one workgroup (barriers are
// workgroup barriers) suits an ensemble, whose flows fill the chip; a single
// flow spreads over a team of up to 64 co-resident workgroups (a cooperative
// launch) that synchronise through a counter in the flow's team scratch.
#include <math.h>

#include "hfm_common.h"

namespace hfm {
namespace {

constexpr int kFT = 1024;      // threads per flow
constexpr int kFWv = kFT / 64;  // waves per flow

// Field pointers are read from the task table, so the compiler cannot tell
// they address global memory and would emit FLAT accesses (which also hold
// the LDS counter that the shuffles and reductions wait on).  Every field is
// accessed through these global-address-space types instead.
#define HFM_GLOBAL __attribute__((address_space(1)))
typedef HFM_GLOBAL float gf;
typedef HFM_GLOBAL double gd;

// Workgroup sums in a fixed order; every thread gets the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ void wg_sum2(double& a, double& b, double* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();  // earlier readers of red are done
  if (lane == 0) {
    red[wid] = a;
    red[kFWv + wid] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
  for (int w = 0; w < kFWv; ++w) {
    sa += red[w];
    sb += red[kFWv + w];
  }
  a = sa;
  b = sb;
}

__device__ __forceinline__ double wg_sum(double a, double* red) {
  double b = 0.0;
  wg_sum2(a, b, red);
  return a;
}

// max that keeps a NaN (fmaxf drops it): a NaN velocity must reach the
// stopping test of the step
__device__ __forceinline__ float max_nan(float a, float b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ float wg_max(float a, float* redf) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a = max_nan(a, __shfl_down(a, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) redf[wid] = a;
  __syncthreads();
  float m = redf[0];
  for (int w = 1; w < kFWv; ++w) m = max_nan(m, redf[w]);
  return m;
}

// ------------------------------------------------------------------ teams
// Team scratch (per flow, zeroed by the host before each launch): the
// arrival counter, a failure word, then two slots of per-workgroup partial
// sums (double pairs) and maxima (float pairs) for the team reductions.
constexpr int kTeamMax = 64;
constexpr double kBarrierTimeoutS = 1.0;  // wall-clock bound of one team barrier wait

struct Team {
  static constexpr bool kTeam = true;
  unsigned* cnt;
  unsigned* fail;
  double* part;   // [2][kTeamMax][2]
  float* partf;   // [2][kTeamMax][2]
  int G, rank;
  unsigned nbar;  // barriers passed (thread 0's count)
  int slot;       // reduction slot to use next
  uint64_t wait_ticks;  // barrier timeout in ticks of the constant-rate wall clock
  __device__ __forceinline__ int gt() const { return rank * kFT + int(threadIdx.x); }
  __device__ __forceinline__ int stride() const { return G * kFT; }
  __device__ __forceinline__ bool lead() const { return rank == 0 && threadIdx.x == 0; }
};

// A flow owned by one workgroup: the same interface with compile-time extents
// (an ensemble's path keeps its constant strides and plain barriers).
struct Solo {
  static constexpr bool kTeam = false;
  static constexpr int G = 1;
  static constexpr int rank = 0;
  __device__ __forceinline__ int gt() const { return int(threadIdx.x); }
  __device__ __forceinline__ int stride() const { return kFT; }
  __device__ __forceinline__ bool lead() const { return threadIdx.x == 0; }
};
__device__ __forceinline__ void team_sync(Solo&) { __syncthreads(); }
__device__ __forceinline__ void team_sum2(Solo&, double& a, double& b, double* red) {
  wg_sum2(a, b, red);
}
__device__ __forceinline__ double team_sum(Solo&, double a, double* red) { return wg_sum(a, red); }
__device__ __forceinline__ void team_max2(Solo&, float& a, float& b, float* redf) {
  a = wg_max(a, redf);
  b = wg_max(b, redf);
}

// Every workgroup of the team has finished the pass before any starts the
// next: workgroup barrier, then (for G > 1) a device-scope release, one
// arrival on the counter and an acquire once all G arrivals of this barrier
// are in.  The wait is bounded: on timeout the failure word is set and the
// team runs on without waiting (its results are then void and reported).
__device__ __forceinline__ void team_sync(Team& t) {
  __syncthreads();
  if (t.G == 1) return;
  if (threadIdx.x == 0) {
    t.nbar += 1;
    const unsigned target = t.nbar * unsigned(t.G);
    __threadfence();
    atomicAdd(t.cnt, 1u);
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(t.cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target &&
           __hip_atomic_load(t.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
      __builtin_amdgcn_s_sleep(1);
      if (wall_clock64() - t0 > t.wait_ticks) {
        atomicExch(t.fail, 1u);
        break;
      }
    }
    __threadfence();
  }
  __syncthreads();
}

// Team sums: workgroup sums, then the G partials summed in rank order by
// every thread (the same value in every workgroup).
__device__ __forceinline__ void team_sum2(Team& t, double& a, double& b, double* red) {
  wg_sum2(a, b, red);
  if (t.G == 1) return;
  double* slot = t.part + t.slot * (2 * kTeamMax);
  t.slot ^= 1;
  if (threadIdx.x == 0) {
    slot[2 * t.rank] = a;
    slot[2 * t.rank + 1] = b;
  }
  team_sync(t);
  double sa = 0.0, sb = 0.0;
  for (int q = 0; q < t.G; ++q) {
    sa += __hip_atomic_load(slot + 2 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sb += __hip_atomic_load(slot + 2 * q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  a = sa;
  b = sb;
}

__device__ __forceinline__ double team_sum(Team& t, double a, double* red) {
  double b = 0.0;
  team_sum2(t, a, b, red);
  return a;
}

__device__ __forceinline__ void team_max2(Team& t, float& a, float& b, float* redf) {
  a = wg_max(a, redf);
  b = wg_max(b, redf);
  if (t.G == 1) return;
  float* slot = t.partf + t.slot * (2 * kTeamMax);
  t.slot ^= 1;
  if (threadIdx.x == 0) {
    slot[2 * t.rank] = a;
    slot[2 * t.rank + 1] = b;
  }
  team_sync(t);
  float ma = 0.0f, mb = 0.0f;
  for (int q = 0; q < t.G; ++q) {
    ma = max_nan(ma, __hip_atomic_load(slot + 2 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    mb = max_nan(mb, __hip_atomic_load(slot + 2 * q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  a = ma;
  b = mb;
}

// Periodic neighbour offsets of cell (j, i) on an n x n level (n = 1 << sh).
struct Nb {
  int c, e, w, n, s;
  __device__ __forceinline__ Nb(int c_, int sh) {
    const int m = (1 << sh) - 1;
    const int j = c_ >> sh, i = c_ & m;
    const int row = j << sh;
    c = c_;
    e = row + ((i + 1) & m);
    w = row + ((i - 1) & m);
    n = (((j + 1) & m) << sh) + i;
    s = (((j - 1) & m) << sh) + i;
  }
};

// Cells c = gt + k G kFT of an nn-cell pass (gt: the thread's index in its
// team), kU at a time: load(c) for each cell of a group before put(c, value)
// for any of them, so a thread has kU cells' loads in flight instead of one.
constexpr int kU = 4;

template <typename V, typename TM, typename Load, typename Put>
__device__ __forceinline__ void cells(const TM& t, int nn, Load&& load, Put&& put) {
  const int st = t.stride();
  for (int c0 = t.gt(); c0 < nn; c0 += kU * st) {
    V val[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + k * st;
      if (c < nn) val[k] = load(c);
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + k * st;
      if (c < nn) put(c, val[k]);
    }
  }
}

struct F2 {
  float a, b;
};

// ------------------------------------------------------------------ momentum
// F_u and F_v at cell (j, i): nu/h^2 lap5 - (velocity-signed upwind advection)/h
// + forcing (u only).  vbar / ubar are the 4-point averages of the other
// component at the u / v point.
__device__ __forceinline__ void momentum_rhs(const gf* __restrict__ u,
                                             const gf* __restrict__ v, int c, int sh,
                                             float nu_h2, float inv_h, float f, float& Fu,
                                             float& Fv) {
  const int m = (1 << sh) - 1;
  const int j = c >> sh, i = c & m;
  const int rc = j << sh, rn = ((j + 1) & m) << sh, rs = ((j - 1) & m) << sh;
  const int ie = (i + 1) & m, iw = (i - 1) & m;
  const float uc = u[c], ue = u[rc + ie], uw = u[rc + iw], un = u[rn + i], us = u[rs + i];
  const float vc = v[c], ve = v[rc + ie], vw = v[rc + iw], vn = v[rn + i], vs = v[rs + i];
  const float vnw = v[rn + iw], use = u[rs + ie];
  const float vbar = 0.25f * ((vw + vc) + (vnw + vn));
  const float ubar = 0.25f * ((us + use) + (uc + ue));
  const float dux = uc > 0.0f ? uc - uw : ue - uc;
  const float duy = vbar > 0.0f ? uc - us : un - uc;
  const float dvx = ubar > 0.0f ? vc - vw : ve - vc;
  const float dvy = vc > 0.0f ? vc - vs : vn - vc;
  const float lap_u = ((ue + uw) + (un + us)) - 4.0f * uc;
  const float lap_v = ((ve + vw) + (vn + vs)) - 4.0f * vc;
  Fu = nu_h2 * lap_u - inv_h * (uc * dux + vbar * duy) + f;
  Fv = nu_h2 * lap_v - inv_h * (ubar * dvx + vc * dvy);
}

// ------------------------------------------------------------------ V-cycle
struct Levels {
  int N, sh, nlev;
  __device__ __forceinline__ int off(int l) const {  // sum_{k<l} (N >> k)^2
    int o = 0;
    for (int k = 0; k < l; ++k) o += (N >> k) * (N >> k);
    return o;
  }
};

// out = x + omega (b - s A x) / (4 s) with x = xin (+ alpha e[parent] when ec).
__device__ __forceinline__ float sweep_cell(const gf* __restrict__ xin,
                                            const gf* __restrict__ ec, float alpha,
                                            const gf* __restrict__ b, int c, int sh, float s,
                                            float omega) {
  const Nb q(c, sh);
  float xc = xin[q.c], xe = xin[q.e], xw = xin[q.w], xn = xin[q.n], xs = xin[q.s];
  if (ec) {
    const int m = (1 << sh) - 1, shc = sh - 1;
    auto par = [&](int k) { return ((k >> sh) >> 1 << shc) + ((k & m) >> 1); };
    xc += alpha * ec[par(q.c)];
    xe += alpha * ec[par(q.e)];
    xw += alpha * ec[par(q.w)];
    xn += alpha * ec[par(q.n)];
    xs += alpha * ec[par(q.s)];
  }
  const float Ax = s * (4.0f * xc - ((xe + xw) + (xn + xs)));
  return xc + omega * (b[c] - Ax) / (4.0f * s);
}

__device__ __forceinline__ float resid_cell(const gf* __restrict__ x,
                                            const gf* __restrict__ b, int c, int sh,
                                            float s) {
  const Nb q(c, sh);
  const float Ax = s * (4.0f * x[q.c] - ((x[q.e] + x[q.w]) + (x[q.n] + x[q.s])));
  return b[c] - Ax;
}

// Passes of one V-cycle level.  CX: the threads that run them (a team or one
// workgroup), sync(): the barrier between dependent passes.
template <typename CX, typename SYNC>
__device__ void vdown_level(const CX& cx, SYNC sync, const hfm_flow2d_task& T,
                            const hfm_flow2d_params& P, const Levels& L, int l) {
  const int nu1 = P.nu1;
  const float omega = P.omega;
  const int sh = L.sh - l, nn = 1 << (2 * sh), o = L.off(l);
  const float s = ldexpf(1.0f, -l), d = 4.0f * s;
  gf* buf[2] = {(gf*)T.mgx + o, (gf*)T.mgy + o};
  const gf* b = (const gf*)T.mgb + o;
  if (l > 0) {
    gf* x0 = buf[0];
    cells<float>(cx, nn, [&](int c) { return omega * b[c] / d; }, [&](int c, float y) { x0[c] = y; });
    sync();
  }
  for (int k = 1; k < nu1; ++k) {
    const gf* xin = buf[(k - 1) & 1];
    gf* out = buf[k & 1];
    cells<float>(cx, nn, [&](int c) { return sweep_cell(xin, nullptr, 0.0f, b, c, sh, s, omega); },
                 [&](int c, float y) { out[c] = y; });
    sync();
  }
  // bc = mean over the 2x2 children of (b - A x)
  const gf* x = buf[(nu1 - 1) & 1];
  const int shc = sh - 1, ncc = 1 << (2 * shc), mc = (1 << shc) - 1;
  gf* bc = (gf*)T.mgb + L.off(l + 1);
  cells<float>(
      cx, ncc,
      [&](int C) {
        const int J = C >> shc, I = C & mc;
        const int f00 = ((2 * J) << sh) + 2 * I, f10 = f00 + (1 << sh);
        const float r00 = resid_cell(x, b, f00, sh, s), r01 = resid_cell(x, b, f00 + 1, sh, s);
        const float r10 = resid_cell(x, b, f10, sh, s), r11 = resid_cell(x, b, f10 + 1, sh, s);
        return 0.25f * ((r00 + r01) + (r10 + r11));
      },
      [&](int C, float y) { bc[C] = y; });
  sync();
}

// coarsest: x = pinv b (FP64)
template <typename CX, typename SYNC>
__device__ void vcoarsest(const CX& cx, SYNC sync, const hfm_flow2d_task& T,
                          const double* __restrict__ pinv, const Levels& L) {
  const int l = L.nlev - 1, nn = 1 << (2 * (L.sh - l)), o = L.off(l);
  const gf* b = (const gf*)T.mgb + o;
  gf* x = (gf*)T.mgx + o;
  for (int r = cx.gt(); r < nn; r += cx.stride()) {
    const gd* row = (const gd*)pinv + int64_t(r) * nn;
    double acc = 0.0;
    for (int c = 0; c < nn; ++c) acc += row[c] * double(b[c]);
    x[r] = float(acc);
  }
  sync();
}

// returns this thread's part of b0 . z (level 0's last sweep writes z)
template <typename CX, typename SYNC>
__device__ double vup_level(const CX& cx, SYNC sync, const hfm_flow2d_task& T,
                            const hfm_flow2d_params& P, const Levels& L, int l) {
  const int nu1 = P.nu1, nu2 = P.nu2;
  const float omega = P.omega, alpha = P.alpha;
  const int sh = L.sh - l, nn = 1 << (2 * sh), o = L.off(l), oc = L.off(l + 1);
  const float s = ldexpf(1.0f, -l);
  gf* buf[2] = {(gf*)T.mgx + o, (gf*)T.mgy + o};
  const gf* b = (const gf*)T.mgb + o;
  const gf* ec = (const gf*)((l + 1 == L.nlev - 1) ? T.mgx + oc
                                                   : ((nu1 + nu2 - 1) & 1 ? T.mgy : T.mgx) + oc);
  double rz = 0.0;
  for (int k = 0; k < nu2; ++k) {
    const int K = nu1 + k;
    const gf* xin = buf[(K - 1) & 1];
    const bool last0 = (l == 0 && k == nu2 - 1);
    gf* out = last0 ? (gf*)T.z : buf[K & 1];
    const gf* e = k == 0 ? ec : nullptr;
    if (last0) {
      cells<F2>(cx, nn, [&](int c) { return F2{sweep_cell(xin, e, alpha, b, c, sh, s, omega), b[c]}; },
                [&](int c, F2 y) {
                  out[c] = y.a;
                  rz += double(y.b) * double(y.a);
                });
    } else {
      cells<float>(cx, nn, [&](int c) { return sweep_cell(xin, e, alpha, b, c, sh, s, omega); },
                   [&](int c, float y) { out[c] = y; });
    }
    sync();
  }
  return rz;
}

// Levels of at most kSoloCells cells run on one workgroup of a team (rank 0)
// between workgroup barriers; the rest of the team waits at one team barrier.
constexpr int kSoloCells = 4096;

// z = V(b0) with b0 = mgb level 0 and the first level-0 sweep (omega b0 / 4)
// already in mgx level 0.  Returns this thread's part of b0 . z.
template <typename TM>
__device__ double vcycle(TM& tm, const hfm_flow2d_task& T, const hfm_flow2d_params& P,
                         const double* __restrict__ pinv, const Levels& L) {
  int ls = 0;  // first level run by one workgroup
  if (TM::kTeam) {
    ls = L.nlev - 1;
    while (ls > 0 && (1 << (2 * (L.sh - ls + 1))) <= kSoloCells) --ls;
  }
  auto tsync = [&] { team_sync(tm); };
  auto wsync = [] { __syncthreads(); };
  for (int l = 0; l < ls; ++l) vdown_level(tm, tsync, T, P, L, l);
  double rz = 0.0;
  if (tm.rank == 0) {
    Solo one;
    for (int l = ls; l < L.nlev - 1; ++l) vdown_level(one, wsync, T, P, L, l);
    vcoarsest(one, wsync, T, pinv, L);
    for (int l = L.nlev - 2; l >= ls; --l) rz += vup_level(one, wsync, T, P, L, l);
  }
  if (TM::kTeam) team_sync(tm);
  for (int l = ls - 1; l >= 0; --l) rz += vup_level(tm, tsync, T, P, L, l);
  return rz;
}

// A x for an FP64 or FP32 field, evaluated in FP64.
template <typename X>
__device__ __forceinline__ double apply64(const HFM_GLOBAL X* __restrict__ x, const Nb& q) {
  return 4.0 * double(x[q.c]) - ((double(x[q.e]) + double(x[q.w])) + (double(x[q.n]) + double(x[q.s])));
}

// ------------------------------------------------------------------ one substep
// Returns the HFM_MG_* status of the projection.
template <typename TM>
__device__ int substep(TM& tm, const hfm_flow2d_task& T, const hfm_flow2d_params& P,
                       const double* __restrict__ pinv, const Levels& L, float dt, bool momentum,
                       double* red) {
  const int sh = L.sh, NN = 1 << (2 * sh), m = (1 << sh) - 1;
  gf* __restrict__ u = (gf*)T.u;
  gf* __restrict__ v = (gf*)T.v;
  const gf* __restrict__ fu = (const gf*)T.fu;
  gf* __restrict__ uh = (gf*)T.uh;
  gf* __restrict__ vh = (gf*)T.vh;
  gd* __restrict__ phi = (gd*)T.phi;
  gd* __restrict__ r = (gd*)T.r;
  gf* __restrict__ mgb = (gf*)T.mgb;
  gf* __restrict__ mgx = (gf*)T.mgx;
  gf* __restrict__ p = (gf*)T.p;
  const gf* __restrict__ z = (const gf*)T.z;
  if (momentum) {
    const float hdt = 0.5f * dt;
    cells<F2>(
        tm, NN,
        [&](int c) {
          F2 F;
          momentum_rhs(u, v, c, sh, P.nu_h2, P.inv_h, fu[c >> sh], F.a, F.b);
          return F2{u[c] + hdt * F.a, v[c] + hdt * F.b};
        },
        [&](int c, F2 y) {
          uh[c] = y.a;
          vh[c] = y.b;
        });
    team_sync(tm);
    // u* in place: a cell reads only its own u, v (before any store of its group)
    cells<F2>(
        tm, NN,
        [&](int c) {
          F2 F;
          momentum_rhs(uh, vh, c, sh, P.nu_h2, P.inv_h, fu[c >> sh], F.a, F.b);
          return F2{u[c] + dt * F.a, v[c] + dt * F.b};
        },
        [&](int c, F2 y) {
          u[c] = y.a;
          v[c] = y.b;
        });
    team_sync(tm);
  }
  // divergence source b (into r), its mean
  double sb = 0.0;
  cells<float>(
      tm, NN,
      [&](int c) {
        const int j = c >> sh, i = c & m;
        return P.h * ((u[c] - u[(j << sh) + ((i + 1) & m)]) + (v[c] - v[(((j + 1) & m) << sh) + i]));
      },
      [&](int c, float bf) {
        r[c] = double(bf);
        sb += double(bf);
      });
  const double mean = team_sum(tm, sb, red) / double(NN);
  // r = (b - mean) - A phi; rl = f32(r); first level-0 sweep
  double bb = 0.0, rr = 0.0;
  const float w4 = P.omega;
  struct D2 {
    double a, b;
  };
  cells<D2>(
      tm, NN,
      [&](int c) {
        const Nb q(c, sh);
        const double bm = r[c] - mean;
        return D2{bm, bm - apply64(phi, q)};
      },
      [&](int c, D2 y) {
        r[c] = y.b;
        const float rl = float(y.b);
        mgb[c] = rl;
        mgx[c] = w4 * rl / 4.0f;
        bb += y.a * y.a;
        rr += y.b * y.b;
      });
  team_sum2(tm, bb, rr, red);
  int status = HFM_MG_RUNNING;
  int it = 0;
  double rel = bb > 0.0 ? sqrt(rr / bb) : 0.0;
  // a non-finite right-hand side or residual (NaN or Inf fields) is a
  // breakdown, never "converged with zero iterations"
  if (!isfinite(bb) || !isfinite(rr)) {
    status = HFM_MG_BREAKDOWN;
    rel = rr / bb;
  } else if (bb == 0.0 || rel <= P.tol) {
    status = HFM_MG_CONVERGED;
  }
  double rz_old = 0.0;
  while (status == HFM_MG_RUNNING) {
    if (it >= P.maxit) {
      status = HFM_MG_MAXIT;
      break;
    }
    const double rz = team_sum(tm, vcycle(tm, T, P, pinv, L), red);
    const double beta = it == 0 ? 0.0 : rz / rz_old;
    rz_old = rz;
    const double bd = double(float(beta));
    if (it == 0) {
      cells<float>(tm, NN, [&](int c) { return z[c]; }, [&](int c, float y) { p[c] = y; });
    } else {
      cells<float>(tm, NN, [&](int c) { return float(bd * double(p[c]) + double(z[c])); },
                   [&](int c, float y) { p[c] = y; });
    }
    team_sync(tm);
    double pq = 0.0;
    cells<double>(tm, NN, [&](int c) { return double(p[c]) * apply64(p, Nb(c, sh)); },
                  [&](int, double y) { pq += y; });
    pq = team_sum(tm, pq, red);
    if (!(pq > 0.0) || !isfinite(pq)) {
      status = HFM_MG_BREAKDOWN;
      break;
    }
    const double a = rz / pq;
    rr = 0.0;
    cells<D2>(
        tm, NN,
        [&](int c) {
          const Nb q(c, sh);
          const double pc = double(p[c]);
          return D2{phi[c] + a * pc, r[c] - a * apply64(p, q)};
        },
        [&](int c, D2 y) {
          phi[c] = y.a;
          r[c] = y.b;
          const float rl = float(y.b);
          mgb[c] = rl;
          mgx[c] = w4 * rl / 4.0f;
          rr += y.b * y.b;
        });
    rr = team_sum(tm, rr, red);
    ++it;
    rel = sqrt(rr / bb);
    if (rel <= P.tol) status = HFM_MG_CONVERGED;
    else if (!isfinite(rel)) status = HFM_MG_BREAKDOWN;
  }
  // gradient correction (phi final: the loop ended on a barrier)
  const double ih = double(P.inv_h);
  cells<F2>(
      tm, NN,
      [&](int c) {
        const Nb q(c, sh);
        const double ph = phi[c];
        return F2{float(double(u[c]) - (ph - phi[q.w]) * ih), float(double(v[c]) - (ph - phi[q.s]) * ih)};
      },
      [&](int c, F2 y) {
        u[c] = y.a;
        v[c] = y.b;
      });
  team_sync(tm);
  if (tm.lead()) {
    double* st = T.stats;
    st[HFM_FLOW_ST_ITERS] = it;
    st[HFM_FLOW_ST_REL] = rel;
    st[HFM_FLOW_ST_ITERS_SUM] += it;
    st[HFM_FLOW_ST_STATUS] = status;
    st[HFM_FLOW_ST_SUBSTEPS] += 1.0;
    st[HFM_FLOW_ST_REL_MAX] = fmax(st[HFM_FLOW_ST_REL_MAX], rel);
  }
  return status;
}

// A team whose barrier timed out reports it (its fields are void).
__device__ __forceinline__ void team_finish(const Solo&, const hfm_flow2d_task&) {}
__device__ __forceinline__ void team_finish(const Team& t, const hfm_flow2d_task& T) {
  if (t.G > 1 && t.lead() &&
      __hip_atomic_load(t.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
    T.stats[HFM_FLOW_ST_STATUS] = HFM_FLOW_TEAM_TIMEOUT;
}

template <typename TM>
__device__ __forceinline__ void flow_steps(TM& tm, const hfm_flow2d_task& T,
                                           const hfm_flow2d_params& P,
                                           const double* __restrict__ pinv, int nsteps, int flags,
                                           double* red, float* redf) {
  Levels L;
  L.N = P.N;
  L.sh = 31 - __clz(P.N);
  L.nlev = P.nlev;
  const int NN = P.N * P.N;
  // a flow whose last projection did not converge stays stopped (read by
  // every workgroup before the first barrier, hence before any stats update)
  const double st0 = T.stats[HFM_FLOW_ST_STATUS];
  if (st0 != double(HFM_MG_CONVERGED) && st0 != double(HFM_MG_RUNNING)) return;
  if (flags & HFM_FLOW_PROJECT_ONLY) {
    substep(tm, T, P, pinv, L, 0.0f, false, red);
    team_finish(tm, T);
    return;
  }
  for (int step = 0; step < nsteps; ++step) {
    float umax = 0.0f, vmax = 0.0f;
    const gf* u = (const gf*)T.u;
    const gf* v = (const gf*)T.v;
    cells<F2>(tm, NN, [&](int c) { return F2{fabsf(u[c]), fabsf(v[c])}; },
              [&](int, F2 a) {
                umax = max_nan(umax, a.a);
                vmax = max_nan(vmax, a.b);
              });
    team_max2(tm, umax, vmax, redf);
    const float q = (umax + vmax) * P.cfl_coef;
    // NaN (a breakdown) or too fast (more substeps than allowed): stop this flow
    if (!(q <= float(P.max_sub))) {
      if (tm.lead()) T.stats[HFM_FLOW_ST_STATUS] = q != q ? HFM_MG_BREAKDOWN : HFM_FLOW_CFL;
      team_finish(tm, T);
      return;
    }
    const int nsub = q > 1.0f ? int(ceilf(q)) : 1;
    const float dts = P.dt / float(nsub);
    bool ok = true;
    for (int k = 0; k < nsub && ok; ++k) ok = substep(tm, T, P, pinv, L, dts, true, red) == HFM_MG_CONVERGED;
    if (!ok) break;
    if (tm.lead()) {
      T.stats[HFM_FLOW_ST_STEPS] += 1.0;
      T.stats[HFM_FLOW_ST_NSUB] = nsub;
    }
  }
  team_finish(tm, T);
}

// One workgroup per flow (an ensemble).
__global__ __launch_bounds__(kFT) void k_flow2d(const hfm_flow2d_task* __restrict__ tasks,
                                               hfm_flow2d_params P,
                                               const double* __restrict__ pinv, int nsteps,
                                               int flags) {
  __shared__ double red[2 * kFWv];
  __shared__ float redf[kFWv];
  Solo tm;
  flow_steps(tm, tasks[blockIdx.x], P, pinv, nsteps, flags, red, redf);
}

// P.team workgroups per flow: workgroup b is rank b % team of flow b / team.
// (Placing a team on one XCD -- every 8th workgroup of an 8x larger launch --
// measured slower than the plain round-robin spread.)
__global__ __launch_bounds__(kFT) void k_flow2d_team(const hfm_flow2d_task* __restrict__ tasks,
                                                    hfm_flow2d_params P,
                                                    const double* __restrict__ pinv, int nsteps,
                                                    int flags, uint64_t wait_ticks) {
  __shared__ double red[2 * kFWv];
  __shared__ float redf[kFWv];
  const int G = P.team;
  const int m = int(blockIdx.x);
  const hfm_flow2d_task& T = tasks[m / G];
  Team tm;
  tm.G = G;
  tm.rank = m % G;
  tm.nbar = 0;
  tm.slot = 0;
  tm.wait_ticks = wait_ticks;
  unsigned* w = (unsigned*)T.team;
  tm.cnt = w;
  tm.fail = w + 1;
  tm.part = (double*)T.team + 1;
  tm.partf = (float*)((double*)T.team + 1 + 4 * kTeamMax);
  flow_steps(tm, T, P, pinv, nsteps, flags, red, redf);
}

}  // namespace
}  // namespace hfm

using namespace hfm;

extern "C" int hfm_flow2d_steps(const hfm_flow2d_task* tasks, int64_t nflows,
                                const hfm_flow2d_params* prm, const double* pinv, int nsteps,
                                int flags, void* stream) {
  if (!tasks || !prm || !pinv || nflows < 0 || nflows > INT32_MAX || nsteps < 0) {
    set_error("hfm_flow2d_steps: bad arguments");
    return HFM_EINVAL;
  }
  const hfm_flow2d_params& P = *prm;
  if (P.nlev < 2 || P.nu1 < 1 || P.nu2 < 1 || P.maxit < 1 || P.max_sub < 1 || !(P.tol >= 0.0) ||
      !(P.dt >= 0.0f) || !(P.h > 0.0f)) {
    set_error("hfm_flow2d_steps: bad parameters");
    return HFM_EINVAL;
  }
  // N a power of two with N^2 < 2^30; coarsest edge >= 2 with at most 1024 cells
  const int64_t nc = P.N >> (P.nlev - 1);
  if (P.N < 4 || P.N > (1 << 15) || (P.N & (P.N - 1)) || P.nlev > 15 || nc < 2 ||
      nc * nc > 1024 || (nc << (P.nlev - 1)) != P.N) {
    set_error("hfm_flow2d_steps: grid %d with %d levels is not supported", P.N, P.nlev);
    return HFM_EINVAL;
  }
  if (P.team < 1 || P.team > kTeamMax || nflows * P.team > INT32_MAX) {
    set_error("hfm_flow2d_steps: team of %d workgroups (1..%d)", P.team, kTeamMax);
    return HFM_EINVAL;
  }
  if (nflows == 0 || (nsteps == 0 && !(flags & HFM_FLOW_PROJECT_ONLY))) return HFM_OK;
  if (P.team == 1) {
    hipLaunchKernelGGL(k_flow2d, dim3(unsigned(nflows)), dim3(kFT), 0, as_stream(stream), tasks,
                       P, pinv, nsteps, flags);
    return launch_status("hfm_flow2d_steps");
  }
  // the barrier timeout in ticks of the device's constant-rate wall clock
  int dev = 0, rate_khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
      rate_khz <= 0) {
    set_error("hfm_flow2d_steps: cannot read the device wall-clock rate");
    return HFM_EDEVICE;
  }
  const hfm_flow2d_task* ta = tasks;
  hfm_flow2d_params pa = P;
  const double* pi = pinv;
  int ns = nsteps, fl = flags;
  uint64_t wt = uint64_t(double(rate_khz) * 1000.0 * kBarrierTimeoutS);
  void* args[] = {(void*)&ta, (void*)&pa, (void*)&pi, (void*)&ns, (void*)&fl, (void*)&wt};
  const hipError_t e =
      hipLaunchCooperativeKernel((const void*)k_flow2d_team, dim3(unsigned(nflows * P.team)),
                                 dim3(kFT),
                                 args, 0, as_stream(stream));
  if (e != hipSuccess) {
    set_error("hfm_flow2d_steps: cooperative launch of %lld workgroups failed: %s",
              (long long)(nflows * P.team), hipGetErrorString(e));
    return HFM_EDEVICE;
  }
  return launch_status("hfm_flow2d_steps");
}

extern "C" int hfm_flow2d_struct_sizes(int64_t* task_size, int64_t* params_size) {
  if (!task_size || !params_size) {
    set_error("hfm_flow2d_struct_sizes: null output");
    return HFM_EINVAL;
  }
  *task_size = int64_t(sizeof(hfm_flow2d_task));
  *params_size = int64_t(sizeof(hfm_flow2d_params));
  return HFM_OK;
}
