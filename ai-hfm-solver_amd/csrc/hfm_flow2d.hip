// 2D periodic incompressible flow on a staggered grid for gfx950: explicit
// two-stage momentum stencils, divergence source, FP64 conjugate gradients
// preconditioned by an FP32 multigrid V-cycle, gradient correction.
//
// The grids are small (256^2 floats = 256 KB per field) and one step is a
// chain of ~40 dependent passes per CG iteration, so launches, not bytes,
// bound a kernel-per-pass design.  Here one 1024-thread workgroup owns a flow
// and runs whole steps: passes are separated by workgroup barriers, every
// field stays in the CU's L1/L2 working set, scalars (dots, CG coefficients,
// the stopping test) never leave the workgroup, and a launch over many flows
// fills the chip with independent workgroups.
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

__device__ __forceinline__ float wg_max(float a, float* redf) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a = fmaxf(a, __shfl_down(a, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) redf[wid] = a;
  __syncthreads();
  float m = redf[0];
  for (int w = 1; w < kFWv; ++w) m = fmaxf(m, redf[w]);
  return m;
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

// Cells c = threadIdx.x + k kFT of an nn-cell pass, kU at a time: load(c)
// for each cell of a group before put(c, value) for any of them, so a
// thread has kU cells' loads in flight instead of one cell's.
constexpr int kU = 4;

template <typename V, typename Load, typename Put>
__device__ __forceinline__ void cells(int nn, Load&& load, Put&& put) {
  for (int c0 = threadIdx.x; c0 < nn; c0 += kU * kFT) {
    V val[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + k * kFT;
      if (c < nn) val[k] = load(c);
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + k * kFT;
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

// z = V(b0) with b0 = mgb level 0 and the first level-0 sweep (omega b0 / 4)
// already in mgx level 0.  Returns this thread's part of b0 . z.
__device__ double vcycle(const hfm_flow2d_task& T, const hfm_flow2d_params& P,
                         const double* __restrict__ pinv, const Levels& L) {
  const int tid = threadIdx.x;
  const int nu1 = P.nu1, nu2 = P.nu2;
  const float omega = P.omega, alpha = P.alpha;
  // down
  for (int l = 0; l < L.nlev - 1; ++l) {
    const int sh = L.sh - l, nn = 1 << (2 * sh), o = L.off(l);
    const float s = ldexpf(1.0f, -l), d = 4.0f * s;
    gf* buf[2] = {(gf*)T.mgx + o, (gf*)T.mgy + o};
    const gf* b = (const gf*)T.mgb + o;
    if (l > 0) {
      gf* x0 = buf[0];
      cells<float>(nn, [&](int c) { return omega * b[c] / d; }, [&](int c, float y) { x0[c] = y; });
      __syncthreads();
    }
    for (int k = 1; k < nu1; ++k) {
      const gf* xin = buf[(k - 1) & 1];
      gf* out = buf[k & 1];
      cells<float>(nn, [&](int c) { return sweep_cell(xin, nullptr, 0.0f, b, c, sh, s, omega); },
                   [&](int c, float y) { out[c] = y; });
      __syncthreads();
    }
    // bc = mean over the 2x2 children of (b - A x)
    const gf* x = buf[(nu1 - 1) & 1];
    const int shc = sh - 1, ncc = 1 << (2 * shc), mc = (1 << shc) - 1;
    gf* bc = (gf*)T.mgb + L.off(l + 1);
    cells<float>(
        ncc,
        [&](int C) {
          const int J = C >> shc, I = C & mc;
          const int f00 = ((2 * J) << sh) + 2 * I, f10 = f00 + (1 << sh);
          const float r00 = resid_cell(x, b, f00, sh, s), r01 = resid_cell(x, b, f00 + 1, sh, s);
          const float r10 = resid_cell(x, b, f10, sh, s), r11 = resid_cell(x, b, f10 + 1, sh, s);
          return 0.25f * ((r00 + r01) + (r10 + r11));
        },
        [&](int C, float y) { bc[C] = y; });
    __syncthreads();
  }
  // coarsest: x = pinv b (FP64)
  {
    const int l = L.nlev - 1, nn = 1 << (2 * (L.sh - l)), o = L.off(l);
    const gf* b = (const gf*)T.mgb + o;
    gf* x = (gf*)T.mgx + o;
    for (int r = tid; r < nn; r += kFT) {
      const gd* row = (const gd*)pinv + int64_t(r) * nn;
      double acc = 0.0;
      for (int c = 0; c < nn; ++c) acc += row[c] * double(b[c]);
      x[r] = float(acc);
    }
    __syncthreads();
  }
  // up
  double rz = 0.0;
  for (int l = L.nlev - 2; l >= 0; --l) {
    const int sh = L.sh - l, nn = 1 << (2 * sh), o = L.off(l), oc = L.off(l + 1);
    const float s = ldexpf(1.0f, -l);
    gf* buf[2] = {(gf*)T.mgx + o, (gf*)T.mgy + o};
    const gf* b = (const gf*)T.mgb + o;
    const gf* ec = (const gf*)((l + 1 == L.nlev - 1) ? T.mgx + oc
                                                     : ((nu1 + nu2 - 1) & 1 ? T.mgy : T.mgx) + oc);
    for (int k = 0; k < nu2; ++k) {
      const int K = nu1 + k;
      const gf* xin = buf[(K - 1) & 1];
      const bool last0 = (l == 0 && k == nu2 - 1);
      gf* out = last0 ? (gf*)T.z : buf[K & 1];
      const gf* e = k == 0 ? ec : nullptr;
      if (last0) {
        cells<F2>(nn, [&](int c) { return F2{sweep_cell(xin, e, alpha, b, c, sh, s, omega), b[c]}; },
                  [&](int c, F2 y) {
                    out[c] = y.a;
                    rz += double(y.b) * double(y.a);
                  });
      } else {
        cells<float>(nn, [&](int c) { return sweep_cell(xin, e, alpha, b, c, sh, s, omega); },
                     [&](int c, float y) { out[c] = y; });
      }
      __syncthreads();
    }
  }
  return rz;
}

// A x for an FP64 or FP32 field, evaluated in FP64.
template <typename X>
__device__ __forceinline__ double apply64(const HFM_GLOBAL X* __restrict__ x, const Nb& q) {
  return 4.0 * double(x[q.c]) - ((double(x[q.e]) + double(x[q.w])) + (double(x[q.n]) + double(x[q.s])));
}

// ------------------------------------------------------------------ one substep
// Returns the HFM_MG_* status of the projection.
__device__ int substep(const hfm_flow2d_task& T, const hfm_flow2d_params& P,
                       const double* __restrict__ pinv, const Levels& L, float dt, bool momentum,
                       double* red) {
  const int tid = threadIdx.x;
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
        NN,
        [&](int c) {
          F2 F;
          momentum_rhs(u, v, c, sh, P.nu_h2, P.inv_h, fu[c >> sh], F.a, F.b);
          return F2{u[c] + hdt * F.a, v[c] + hdt * F.b};
        },
        [&](int c, F2 y) {
          uh[c] = y.a;
          vh[c] = y.b;
        });
    __syncthreads();
    // u* in place: a cell reads only its own u, v (before any store of its group)
    cells<F2>(
        NN,
        [&](int c) {
          F2 F;
          momentum_rhs(uh, vh, c, sh, P.nu_h2, P.inv_h, fu[c >> sh], F.a, F.b);
          return F2{u[c] + dt * F.a, v[c] + dt * F.b};
        },
        [&](int c, F2 y) {
          u[c] = y.a;
          v[c] = y.b;
        });
    __syncthreads();
  }
  // divergence source b (into r), its mean
  double sb = 0.0;
  cells<float>(
      NN,
      [&](int c) {
        const int j = c >> sh, i = c & m;
        return P.h * ((u[c] - u[(j << sh) + ((i + 1) & m)]) + (v[c] - v[(((j + 1) & m) << sh) + i]));
      },
      [&](int c, float bf) {
        r[c] = double(bf);
        sb += double(bf);
      });
  const double mean = wg_sum(sb, red) / double(NN);
  // r = (b - mean) - A phi; rl = f32(r); first level-0 sweep
  double bb = 0.0, rr = 0.0;
  const float w4 = P.omega;
  struct D2 {
    double a, b;
  };
  cells<D2>(
      NN,
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
  wg_sum2(bb, rr, red);
  int status = HFM_MG_RUNNING;
  int it = 0;
  double rel = bb > 0.0 ? sqrt(rr / bb) : 0.0;
  if (bb == 0.0 || rel <= P.tol) status = HFM_MG_CONVERGED;
  double rz_old = 0.0;
  while (status == HFM_MG_RUNNING) {
    if (it >= P.maxit) {
      status = HFM_MG_MAXIT;
      break;
    }
    const double rz = wg_sum(vcycle(T, P, pinv, L), red);
    const double beta = it == 0 ? 0.0 : rz / rz_old;
    rz_old = rz;
    const double bd = double(float(beta));
    if (it == 0) {
      cells<float>(NN, [&](int c) { return z[c]; }, [&](int c, float y) { p[c] = y; });
    } else {
      cells<float>(NN, [&](int c) { return float(bd * double(p[c]) + double(z[c])); },
                   [&](int c, float y) { p[c] = y; });
    }
    __syncthreads();
    double pq = 0.0;
    cells<double>(NN, [&](int c) { return double(p[c]) * apply64(p, Nb(c, sh)); },
                  [&](int, double y) { pq += y; });
    pq = wg_sum(pq, red);
    if (!(pq > 0.0) || !isfinite(pq)) {
      status = HFM_MG_BREAKDOWN;
      break;
    }
    const double a = rz / pq;
    rr = 0.0;
    cells<D2>(
        NN,
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
    rr = wg_sum(rr, red);
    ++it;
    rel = sqrt(rr / bb);
    if (rel <= P.tol) status = HFM_MG_CONVERGED;
    else if (!isfinite(rel)) status = HFM_MG_BREAKDOWN;
  }
  // gradient correction (phi final: the loop ended on a barrier)
  const double ih = double(P.inv_h);
  cells<F2>(
      NN,
      [&](int c) {
        const Nb q(c, sh);
        const double ph = phi[c];
        return F2{float(double(u[c]) - (ph - phi[q.w]) * ih), float(double(v[c]) - (ph - phi[q.s]) * ih)};
      },
      [&](int c, F2 y) {
        u[c] = y.a;
        v[c] = y.b;
      });
  __syncthreads();
  if (tid == 0) {
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

__global__ __launch_bounds__(kFT) void k_flow2d(const hfm_flow2d_task* __restrict__ tasks,
                                               hfm_flow2d_params P,
                                               const double* __restrict__ pinv, int nsteps,
                                               int flags) {
  const hfm_flow2d_task& T = tasks[blockIdx.x];
  __shared__ double red[2 * kFWv];
  __shared__ float redf[kFWv];
  Levels L;
  L.N = P.N;
  L.sh = 31 - __clz(P.N);
  L.nlev = P.nlev;
  const int NN = P.N * P.N;
  // a flow whose last projection did not converge stays stopped
  const double st0 = T.stats[HFM_FLOW_ST_STATUS];
  if (st0 != double(HFM_MG_CONVERGED) && st0 != double(HFM_MG_RUNNING)) return;
  if (flags & HFM_FLOW_PROJECT_ONLY) {
    substep(T, P, pinv, L, 0.0f, false, red);
    return;
  }
  for (int step = 0; step < nsteps; ++step) {
    float umax = 0.0f, vmax = 0.0f;
    const gf* u = (const gf*)T.u;
    const gf* v = (const gf*)T.v;
    cells<F2>(NN, [&](int c) { return F2{fabsf(u[c]), fabsf(v[c])}; },
              [&](int, F2 a) {
                umax = fmaxf(umax, a.a);
                vmax = fmaxf(vmax, a.b);
              });
    umax = wg_max(umax, redf);
    vmax = wg_max(vmax, redf);
    const float q = (umax + vmax) * P.cfl_coef;
    // NaN or too fast: stop this flow
    if (!(q <= float(P.max_sub))) {
      if (threadIdx.x == 0) T.stats[HFM_FLOW_ST_STATUS] = HFM_FLOW_CFL;
      return;
    }
    const int nsub = q > 1.0f ? int(ceilf(q)) : 1;
    const float dts = P.dt / float(nsub);
    for (int k = 0; k < nsub; ++k) {
      if (substep(T, P, pinv, L, dts, true, red) != HFM_MG_CONVERGED) return;
    }
    if (threadIdx.x == 0) {
      T.stats[HFM_FLOW_ST_STEPS] += 1.0;
      T.stats[HFM_FLOW_ST_NSUB] = nsub;
    }
  }
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
  if (nflows == 0 || (nsteps == 0 && !(flags & HFM_FLOW_PROJECT_ONLY))) return HFM_OK;
  hipLaunchKernelGGL(k_flow2d, dim3(unsigned(nflows)), dim3(kFT), 0, as_stream(stream), tasks, P,
                     pinv, nsteps, flags);
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
