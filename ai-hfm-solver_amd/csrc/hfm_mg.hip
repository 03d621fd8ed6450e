// Config-mode multigrid kernels for gfx950: weighted-Jacobi sweeps on the
// face-coefficient 7-point operator, fused residual + restriction, dense
// coarsest solve, FP64 flexible-CG passes and per-instance reductions.
//
// Every kernel is driven by a work table (hfm_mg_task[] + per-block task
// index), so one launch covers any mix of instances of different sizes.
// Memory-bound: flattened row-major chunks per block, 16-byte row loads,
// neighbour rows and planes from L1/L2, chunks of a task spread contiguously
// over the 8 XCDs.
#include <math.h>

#include <type_traits>

#include "hfm_common.h"

namespace hfm {

// Cells per block: every cell kernel covers kCellChunk flattened cells per
// block; the restriction covers kCoarseChunk coarse cells per block.
// Vector path (row width W % 8 == 0, all offsets 8-element aligned by the
// host): a thread owns a run of 8 consecutive x-cells, loads 16-byte rows of
// every field and takes its x-neighbours from the adjacent lanes.  Scalar
// path (other widths, small coarse levels): 8 cells per thread, stride 256.
constexpr int kRun = 8;
constexpr int kCellChunk = kBlock * kRun;
constexpr int kCoarseChunk = kBlock * 4;

// FP64 solver vectors (x64, r64) of a vector-path task are stored
// lane-interleaved: in every 512-cell segment (the runs of one wave), cells
// 8l + 2q and 8l + 2q + 1 of lane l's run sit at segment offsets 128 q + 2 l
// and 128 q + 2 l + 1, so each of the wave's four 16-byte accesses per vector
// covers 1 KiB contiguously instead of striding 64 bytes per lane.  The host
// pads these arrays to a multiple of kIlvSeg cells.  Scalar-path tasks keep
// the natural order.
constexpr int kIlvSeg = 64 * kRun;
__device__ __forceinline__ int ilv_run(int n0) {  // position of run n0's first pair
  return (n0 & ~(kIlvSeg - 1)) + ((n0 & (kIlvSeg - 1)) >> 2);
}
__device__ __forceinline__ int ilv_cell(int n) {
  const int o = n & (kIlvSeg - 1);
  return (n & ~(kIlvSeg - 1)) + 128 * ((o & 7) >> 1) + 2 * (o >> 3) + (o & 1);
}

template <typename C>
__device__ __forceinline__ float ldc(const void* p, int64_t i) {
  return ldg(gin<C>(p) + i);
}
template <typename C>
__device__ __forceinline__ float ldc(const HFM_GLOBAL C* p, int64_t i) {
  return ldg(p + i);
}

// ------------------------------------------------------------------ 8-wide row access
// Native vector types (16- and 8-byte accesses through global pointers).
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld8(const HFM_GLOBAL float* p, float v[8]) {
  const f4v a = ((const HFM_GLOBAL f4v*)p)[0];
  const f4v b = ((const HFM_GLOBAL f4v*)p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void ld8(const HFM_GLOBAL bf16* p, float v[8]) {
  const u4v u = *(const HFM_GLOBAL u4v*)p;
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = __uint_as_float(w[q] << 16);
    v[2 * q + 1] = __uint_as_float(w[q] & 0xFFFF0000u);
  }
}
__device__ __forceinline__ void ld4(const HFM_GLOBAL float* p, float v[4]) {
  const f4v a = *(const HFM_GLOBAL f4v*)p;
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void ld4(const HFM_GLOBAL bf16* p, float v[4]) {
  const u2v u = *(const HFM_GLOBAL u2v*)p;
  v[0] = __uint_as_float(u.x << 16); v[1] = __uint_as_float(u.x & 0xFFFF0000u);
  v[2] = __uint_as_float(u.y << 16); v[3] = __uint_as_float(u.y & 0xFFFF0000u);
}
__device__ __forceinline__ void st8(HFM_GLOBAL float* p, const float v[8]) {
  ((HFM_GLOBAL f4v*)p)[0] = f4v{v[0], v[1], v[2], v[3]};
  ((HFM_GLOBAL f4v*)p)[1] = f4v{v[4], v[5], v[6], v[7]};
}
__device__ __forceinline__ uint32_t pack2(float a, float b) { return pack_bf16x2(a, b); }
__device__ __forceinline__ void st8(HFM_GLOBAL bf16* p, const float v[8]) {
  *(HFM_GLOBAL u4v*)p = u4v{pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7])};
}
__device__ __forceinline__ void st4(HFM_GLOBAL float* p, const float v[4]) {
  *(HFM_GLOBAL f4v*)p = f4v{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void st4(HFM_GLOBAL bf16* p, const float v[4]) {
  *(HFM_GLOBAL u2v*)p = u2v{pack2(v[0], v[1]), pack2(v[2], v[3])};
}
// FP16 storage: the same accesses, converted by the hardware's f16 <-> f32 moves
__device__ __forceinline__ uint32_t pack2h(float a, float b) { return pack_f16x2(a, b); }
__device__ __forceinline__ void ld8(const HFM_GLOBAL f16* p, float v[8]) {
  const u4v u = *(const HFM_GLOBAL u4v*)p;
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = f16_lo(w[q]);
    v[2 * q + 1] = f16_hi(w[q]);
  }
}
__device__ __forceinline__ void ld4(const HFM_GLOBAL f16* p, float v[4]) {
  const u2v u = *(const HFM_GLOBAL u2v*)p;
  v[0] = f16_lo(u.x); v[1] = f16_hi(u.x);
  v[2] = f16_lo(u.y); v[3] = f16_hi(u.y);
}
__device__ __forceinline__ void st8(HFM_GLOBAL f16* p, const float v[8]) {
  *(HFM_GLOBAL u4v*)p = u4v{pack2h(v[0], v[1]), pack2h(v[2], v[3]), pack2h(v[4], v[5]),
                            pack2h(v[6], v[7])};
}
__device__ __forceinline__ void st4(HFM_GLOBAL f16* p, const float v[4]) {
  *(HFM_GLOBAL u2v*)p = u2v{pack2h(v[0], v[1]), pack2h(v[2], v[3])};
}
// The run at n0 of a lane-interleaved FP64 vector (see kIlvSeg).  These
// vectors are read once per launch: non-temporal loads keep them from
// displacing the stencil rows other lanes still re-read from the caches
// (3.8 % faster update launches).
__device__ __forceinline__ void ld8d_ilv(const HFM_GLOBAL double* p, int n0, double v[8]) {
  const HFM_GLOBAL double* b = p + ilv_run(n0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const d2v a = __builtin_nontemporal_load((const HFM_GLOBAL d2v*)(b + 128 * q));
    v[2 * q] = a.x;
    v[2 * q + 1] = a.y;
  }
}
__device__ __forceinline__ void st8d_ilv(HFM_GLOBAL double* p, int n0, const double v[8]) {
  HFM_GLOBAL double* b = p + ilv_run(n0);
#pragma unroll
  for (int q = 0; q < 4; ++q) *(HFM_GLOBAL d2v*)(b + 128 * q) = d2v{v[2 * q], v[2 * q + 1]};
}

// Streaming (non-temporal) variants for arrays read or written exactly once
// per launch: they are not kept in the caches that serve the stencil reuse.

__device__ __forceinline__ void ld8_nt(const HFM_GLOBAL float* p, float v[8]) {
  const f4v a = __builtin_nontemporal_load((const HFM_GLOBAL f4v*)p);
  const f4v b = __builtin_nontemporal_load((const HFM_GLOBAL f4v*)p + 1);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void ld8_nt(const HFM_GLOBAL bf16* p, float v[8]) {
  const u4v u = __builtin_nontemporal_load((const HFM_GLOBAL u4v*)p);
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = __uint_as_float(w[q] << 16);
    v[2 * q + 1] = __uint_as_float(w[q] & 0xFFFF0000u);
  }
}
__device__ __forceinline__ void st8_nt(HFM_GLOBAL float* p, const float v[8]) {
  __builtin_nontemporal_store(f4v{v[0], v[1], v[2], v[3]}, (HFM_GLOBAL f4v*)p);
  __builtin_nontemporal_store(f4v{v[4], v[5], v[6], v[7]}, (HFM_GLOBAL f4v*)p + 1);
}
__device__ __forceinline__ void st8_nt(HFM_GLOBAL bf16* p, const float v[8]) {
  __builtin_nontemporal_store(u4v{pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]),
                                  pack2(v[6], v[7])},
                              (HFM_GLOBAL u4v*)p);
}
__device__ __forceinline__ void ld8_nt(const HFM_GLOBAL f16* p, float v[8]) {
  const u4v u = __builtin_nontemporal_load((const HFM_GLOBAL u4v*)p);
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = f16_lo(w[q]);
    v[2 * q + 1] = f16_hi(w[q]);
  }
}
__device__ __forceinline__ void st8_nt(HFM_GLOBAL f16* p, const float v[8]) {
  __builtin_nontemporal_store(u4v{pack2h(v[0], v[1]), pack2h(v[2], v[3]), pack2h(v[4], v[5]),
                                  pack2h(v[6], v[7])},
                              (HFM_GLOBAL u4v*)p);
}

// ------------------------------------------------------------------ stencil arithmetic
// The products that feed a sum are written as explicit fused multiply-adds:
// a cell's value then rounds the same in every kernel and path that computes
// it (a row of a run, a single neighbour, a fused pass), whatever the
// compiler would contract in each context.
__device__ __forceinline__ float fma_a(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_a(double a, double b, double c) { return __builtin_fma(a, b, c); }

// x + alpha * e: the prolonged coarse correction added to x
__device__ __forceinline__ float correct_val(float x, float e, float alpha) {
  return __builtin_fmaf(alpha, e, x);
}

// ------------------------------------------------------------------ scalar stencil row
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
  if (i > 0) off = fma_a(cw, (A)xat(n - 1, k, j, i - 1), off);
  if (i < W - 1) off = fma_a(ce, (A)xat(n + 1, k, j, i + 1), off);
  if (j > 0) off = fma_a(cs, (A)xat(n - W, k, j - 1, i), off);
  if (j < H - 1) off = fma_a(cn, (A)xat(n + W, k, j + 1, i), off);
  if (k > 0) off = fma_a(cd, (A)xat(n - HW, k - 1, j, i), off);
  if (k < D - 1) off = fma_a(cu, (A)xat(n + HW, k + 1, j, i), off);
  return off;
}

__device__ __forceinline__ void decompose(const hfm_mg_task& t, int n, int& i, int& j, int& k) {
  const uint32_t row = fdiv(uint32_t(n), t.magW);
  i = n - int(row) * t.W;
  k = int(fdiv(row, t.magH));
  j = int(row) - k * t.H;
}

// ------------------------------------------------------------------ vector stencil run
struct Run {
  int n0, i0, j, k;
};

// Face coefficients of a run: west face cw[m], east ce[m], south cs, north cn,
// down cd, up cu.  Every address is in bounds (lanes past the end of a task
// carry run 0), so all rows are loaded without branches and waited for
// together; a boundary selects the boundary-face array instead.
template <typename C>
struct RunCoef {
  float cw[8], ce[8], cs[8], cn[8], cd[8], cu[8];
  __device__ __forceinline__ void load(const hfm_mg_task& t, const Run& r) {
    const HFM_GLOBAL C* cx = gin<C>(t.c[0]);
    const HFM_GLOBAL C* cy = gin<C>(t.c[1]);
    const HFM_GLOBAL C* cz = gin<C>(t.c[2]);
    const HFM_GLOBAL C* pn =
        r.j + 1 < t.H ? cy + r.n0 + t.W : gin<C>(t.c[4]) + r.k * t.W + r.i0;
    const HFM_GLOBAL C* pu =
        r.k + 1 < t.D ? cz + r.n0 + t.H * t.W : gin<C>(t.c[5]) + r.j * t.W + r.i0;
    const HFM_GLOBAL C* pe =
        r.i0 + kRun == t.W ? gin<C>(t.c[3]) + r.k * t.H + r.j : cx + r.n0 + kRun;
    ld8(cx + r.n0, cw);
    ld8(cy + r.n0, cs);
    ld8(pn, cn);
    ld8(cz + r.n0, cd);
    ld8(pu, cu);
    const float eface = ldg(pe);
#pragma unroll
    for (int m = 0; m < 7; ++m) ce[m] = cw[m + 1];
    ce[7] = eface;
  }
  template <typename A>
  __device__ __forceinline__ void diag(A d[8]) const {
#pragma unroll
    for (int m = 0; m < 8; ++m)
      d[m] = (A)cw[m] + (A)ce[m] + (A)cs[m] + (A)cn[m] + (A)cd[m] + (A)cu[m];
  }
};

// x sources for the run stencil: run(n, k, j, i0, v) loads 8 values at a
// row-aligned position, one(n, k, j, i) a single value.
template <typename T>
struct SrcPlain {
  const HFM_GLOBAL T* x;
  __device__ __forceinline__ void run(int n, int, int, int, float v[8]) const { ld8(x + n, v); }
  __device__ __forceinline__ float one(int n, int, int, int) const { return ldg(x + n); }
};

// x := xin + alpha * P ec (nearest prolongation of the coarse correction)
template <typename T>
struct SrcCorrect {
  const HFM_GLOBAL T* x;
  const HFM_GLOBAL T* ec;
  int Hc, Wc;
  float alpha;
  __device__ __forceinline__ void run(int n, int k, int j, int i0, float v[8]) const {
    float f[8], e[4];
    ld8(x + n, f);
    ld4(ec + ((k >> 1) * Hc + (j >> 1)) * Wc + (i0 >> 1), e);
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = correct_val(f[m], e[m >> 1], alpha);
  }
  __device__ __forceinline__ float one(int n, int k, int j, int i) const {
    return correct_val(ldg(x + n), ldg(ec + ((k >> 1) * Hc + (j >> 1)) * Wc + (i >> 1)), alpha);
  }
};

// p := T(z + beta p_old) (one fused multiply-add in FP32; p_old unread when beta == 0)
template <typename T>
struct SrcDir {
  const HFM_GLOBAL T* z;
  const HFM_GLOBAL T* pold;
  float beta;
  __device__ __forceinline__ float make(float zv, float pv) const {
    return round_to<T>(beta == 0.0f ? zv : fmaf(beta, pv, zv));
  }
  // p_old is loaded unconditionally (a buffer of the right size; its value
  // is discarded when beta == 0) so the loads issue together
  __device__ __forceinline__ void run(int n, int, int, int, float v[8]) const {
    float zf[8], pf[8];
    ld8(z + n, zf);
    ld8(pold + n, pf);
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = make(zf[m], pf[m]);
  }
  __device__ __forceinline__ float one(int n, int, int, int) const {
    return make(ldg(z + n), ldg(pold + n));
  }
};

// Diagonal of the storage-rounded operator at cell (k, j, i): the six faces
// rounded to T and summed in the order of RunCoef::diag / face_row.
template <typename T>
__device__ __forceinline__ float diag_rounded(const hfm_mg_task& t, int n, int i, int j, int k) {
  const int W = t.W, H = t.H, D = t.D, HW = H * W;
  const float cw = round_to<T>(ldc<float>(t.c[0], n));
  const float ce = round_to<T>(i < W - 1 ? ldc<float>(t.c[0], n + 1) : ldc<float>(t.c[3], k * H + j));
  const float cs = round_to<T>(ldc<float>(t.c[1], n));
  const float cn = round_to<T>(j < H - 1 ? ldc<float>(t.c[1], n + W) : ldc<float>(t.c[4], k * W + i));
  const float cd = round_to<T>(ldc<float>(t.c[2], n));
  const float cu = round_to<T>(k < D - 1 ? ldc<float>(t.c[2], n + HW) : ldc<float>(t.c[5], j * W + i));
  return cw + ce + cs + cn + cd + cu;
}

// off = sum of face * neighbour over a run in w, e, s, n, d, u order; wv / ev
// are the x-neighbours of the run's ends (0 when missing), a missing s/n/d/u
// neighbour contributes face * 0.  Every kernel forms the sum through this
// one expression, so fused and separate passes round identically.
template <typename A, typename C>
__device__ __forceinline__ void offdiag_sum(const RunCoef<C>& q, const A xc[8], A wv, A ev,
                                            const float vs[8], const float vn[8],
                                            const float vd[8], const float vu[8], bool south,
                                            bool north, bool down, bool up, A off[8]) {
#pragma unroll
  for (int m = 0; m < 8; ++m) off[m] = (A)q.cw[m] * (m ? xc[m - 1] : wv);
#pragma unroll
  for (int m = 0; m < 8; ++m) off[m] = fma_a((A)q.ce[m], m < 7 ? xc[m + 1] : ev, off[m]);
#pragma unroll
  for (int m = 0; m < 8; ++m) off[m] = fma_a((A)q.cs[m], south ? (A)vs[m] : (A)0, off[m]);
#pragma unroll
  for (int m = 0; m < 8; ++m) off[m] = fma_a((A)q.cn[m], north ? (A)vn[m] : (A)0, off[m]);
#pragma unroll
  for (int m = 0; m < 8; ++m) off[m] = fma_a((A)q.cd[m], down ? (A)vd[m] : (A)0, off[m]);
#pragma unroll
  for (int m = 0; m < 8; ++m) off[m] = fma_a((A)q.cu[m], up ? (A)vu[m] : (A)0, off[m]);
}

// omega v / d of the weighted-Jacobi sweeps, formed as (omega v) times the
// hardware reciprocal of d (v_rcp_f32, within 1 ulp) instead of the IEEE
// division sequence (about ten instructions).  Every sweep of every kernel
// forms its quotient through this helper, so fused and separate passes and
// the one-launch coarse cycle stay bitwise equal.
__device__ __forceinline__ float wdiv(float omega, float v, float d) {
  return (omega * v) * __builtin_amdgcn_rcpf(d);
}

// The weighted-Jacobi value and the residual of one cell from its diagonal,
// off-diagonal sum, right-hand side and current value.
__device__ __forceinline__ float resid_val(float xc, float d, float bv, float off) {
  return bv - __builtin_fmaf(d, xc, -off);
}
__device__ __forceinline__ float jacobi_val(float xc, float d, float bv, float off, float omega) {
  return xc + wdiv(omega, resid_val(xc, d, bv, off), d);
}

// xc = x on the run, off = sum of face * neighbour (w, e, s, n, d, u order).
// The five rows and the two row-end values are loaded first, from in-bounds
// addresses (a missing neighbour re-reads the run and is zeroed after).
template <typename A, typename C, typename SRC>
__device__ __forceinline__ void run_offdiag(const hfm_mg_task& t, const Run& r,
                                            const RunCoef<C>& q, const SRC& src, A xc[8],
                                            A off[8]) {
  const int W = t.W, HW = t.H * t.W;
  const bool west = r.i0 > 0, east = r.i0 + kRun < W;
  const bool south = r.j > 0, north = r.j + 1 < t.H, down = r.k > 0, up = r.k + 1 < t.D;
  float xf[8], vs[8], vn[8], vd[8], vu[8];
  src.run(r.n0, r.k, r.j, r.i0, xf);
  src.run(south ? r.n0 - W : r.n0, r.k, south ? r.j - 1 : r.j, r.i0, vs);
  src.run(north ? r.n0 + W : r.n0, r.k, north ? r.j + 1 : r.j, r.i0, vn);
  src.run(down ? r.n0 - HW : r.n0, down ? r.k - 1 : r.k, r.j, r.i0, vd);
  src.run(up ? r.n0 + HW : r.n0, up ? r.k + 1 : r.k, r.j, r.i0, vu);
  // the x-neighbours of the run's ends, read directly by every lane (the
  // adjacent lanes' rows are in the same cache lines)
  const float wl = src.one(west ? r.n0 - 1 : r.n0, r.k, r.j, west ? r.i0 - 1 : r.i0);
  const float el = src.one(east ? r.n0 + kRun : r.n0, r.k, r.j, east ? r.i0 + kRun : r.i0);
#pragma unroll
  for (int m = 0; m < 8; ++m) xc[m] = (A)xf[m];
  const A wv = west ? (A)wl : (A)0;
  const A ev = east ? (A)el : (A)0;
  offdiag_sum<A>(q, xc, wv, ev, vs, vn, vd, vu, south, north, down, up, off);
}

__device__ __forceinline__ bool vec_task(const hfm_mg_task& t) { return (t.W & (kRun - 1)) == 0; }

// The run starting at n0, or run 0 for a lane past the end of its task (its
// loads stay in bounds; its results are not stored).
__device__ __forceinline__ Run run_at(const hfm_mg_task& t, int n0, bool valid) {
  Run r;
  r.n0 = valid ? n0 : 0;
  decompose(t, r.n0, r.i0, r.j, r.k);
  return r;
}

// ------------------------------------------------------------------ per-cell forms
// The sweep of cell n (index (k, j, i)) before rounding: ZERO omega b/d, PLAIN
// x + omega (b - A x)/d, CORRECT the same with x := xin + alpha P ec.
template <typename T, int MODE>
__device__ __forceinline__ float sweep_one(const hfm_mg_task& t, int n, int i, int j, int k,
                                           const HFM_GLOBAL T* xin, const HFM_GLOBAL T* ec,
                                           float bn, float omega, float alpha) {
  float diag;
  if (MODE == HFM_SWEEP_ZERO) {
    face_row<float, T>(t, n, i, j, k, [](int, int, int, int) { return 0.0f; }, diag);
    return wdiv(omega, bn, diag);
  } else if (MODE == HFM_SWEEP_PLAIN) {
    auto xat = [&](int m, int, int, int) { return ldg(xin + m); };
    const float off = face_row<float, T>(t, n, i, j, k, xat, diag);
    const float xc = ldg(xin + n);
    return jacobi_val(xc, diag, bn, off, omega);
  } else {
    const int Hc = t.Hc, Wc = t.Wc;
    auto xat = [&](int m, int kk, int jj, int ii) {
      return correct_val(ldg(xin + m), ldg(ec + ((kk >> 1) * Hc + (jj >> 1)) * Wc + (ii >> 1)),
                         alpha);
    };
    const float off = face_row<float, T>(t, n, i, j, k, xat, diag);
    const float xc = xat(n, k, j, i);
    return jacobi_val(xc, diag, bn, off, omega);
  }
}

// Coarse cell nc of R (b - A x): the sum over its (up to 2x2x2) children, / 8.
template <typename T>
__device__ __forceinline__ float restrict_one(const hfm_mg_task& t, int nc,
                                              const HFM_GLOBAL T* x, const HFM_GLOBAL T* b) {
  const int W = t.W, H = t.H, D = t.D;
  const uint32_t rowc = fdiv(uint32_t(nc), t.magWc);
  const int I = nc - int(rowc) * t.Wc;
  const int K = int(fdiv(rowc, t.magHc));
  const int J = int(rowc) - K * t.Hc;
  auto xat = [&](int m, int, int, int) { return ldg(x + m); };
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
        sum += resid_val(ldg(x + n), diag, ldg(b + n), off);
      }
    }
  }
  return sum * 0.125f;
}

// ------------------------------------------------------------------ sweeps
// T: storage of the level's vectors; O: storage of this sweep's output (the
// fine level's last post-sweep may write the preconditioned residual z in a
// wider type than the cycle's internals).
// B: threads per block (a block covers 8 B cells); DOT launches use kBlock.
// PB: the body runs as half of a persistent workgroup, whose other half
// shares its barriers: an early exit still takes part in them.
template <typename T, int MODE, bool DOT, typename O, int B, bool PB = false>
__device__ __forceinline__ void sweep_body(const hfm_mg_task* __restrict__ tasks,
                                           const uint16_t* __restrict__ blk_task,
                                           const hfm_mg_scalars* __restrict__ sc, float omega,
                                           float alpha, double* dots, const Blk& bk) {
  const int g = bk.g;
  const hfm_mg_task& t = tasks[blk_task[g]];
  if (!sc[t.inst].active) {
    if (DOT && PB) block_sum(0.0, bk.half);  // the workgroup's barriers
    return;
  }
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * (B * kRun);
  const HFM_GLOBAL T* xin = gin<T>(t.xin);
  const HFM_GLOBAL T* b = gin<T>(t.b);
  const HFM_GLOBAL T* ec = gin<T>(t.ec);
  HFM_GLOBAL O* out = gout<O>(t.xout);
  double acc = 0.0;
  if (vec_task(t)) {
    const int n0 = base + bk.tid * kRun;
    const bool valid = n0 < N;
    const Run r = run_at(t, n0, valid);
    RunCoef<T> q;
    q.load(t, r);
    float d[8], bv[8], o[8];
    ld8_nt(b + r.n0, bv);
    q.diag(d);
    if (MODE == HFM_SWEEP_ZERO) {
#pragma unroll
      for (int m = 0; m < 8; ++m) o[m] = wdiv(omega, bv[m], d[m]);
    } else {
      float xc[8], off[8];
      if (MODE == HFM_SWEEP_PLAIN) {
        run_offdiag<float>(t, r, q, SrcPlain<T>{xin}, xc, off);
      } else {
        run_offdiag<float>(t, r, q, SrcCorrect<T>{xin, ec, t.Hc, t.Wc, alpha}, xc, off);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) o[m] = jacobi_val(xc[m], d[m], bv[m], off[m], omega);
    }
    if (valid) {
      st8_nt(out + n0, o);
      if (DOT) {
#pragma unroll
        for (int m = 0; m < 8; ++m) acc += double(bv[m]) * double(round_to<O>(o[m]));
      }
    }
  } else if (base < N) {
    for (int s = 0; s < kRun; ++s) {
      const int n = base + s * B + bk.tid;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      const float bn = ldg(b + n);
      const float r = sweep_one<T, MODE>(t, n, i, j, k, xin, ec, bn, omega, alpha);
      const O rs = from_f<O>(r);
      put(out + n, rs);
      if (DOT) acc += double(bn) * double(to_f(rs));
    }
  }
  if (DOT) {
    const double s = block_sum(acc, bk.half);
    if (bk.tid == 0) dots[t.slot0 + chunk] = s;
  }
}

template <typename T, int MODE, bool DOT, typename O, int B>
__global__ __launch_bounds__(B) void k_sweep(const hfm_mg_task* __restrict__ tasks,
                                             const uint16_t* __restrict__ blk_task,
                                             const hfm_mg_scalars* __restrict__ sc, float omega,
                                             float alpha, double* dots) {
  sweep_body<T, MODE, DOT, O, B>(tasks, blk_task, sc, omega, alpha, dots, launch_blk());
}

// ------------------------------------------------------------------ residual + restriction
// bc = sum over the (up to 2x2x2) children of (b - A x), over 8.
// Vector path: a thread owns 4 consecutive coarse cells = a run of 8 fine
// cells in each of its 2 fine rows x 2 fine planes.
template <typename T, int B>
__device__ __forceinline__ void resid_restrict_body(const hfm_mg_task* __restrict__ tasks,
                                                    const uint16_t* __restrict__ blk_task,
                                                    const hfm_mg_scalars* __restrict__ sc,
                                                    const Blk& bk) {
  const int g = bk.g;
  const hfm_mg_task& t = tasks[blk_task[g]];
  if (!sc[t.inst].active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int Nc = t.Dc * t.Hc * t.Wc;
  const int base = chunk * (B * 4);
  const HFM_GLOBAL T* x = gin<T>(t.xin);
  const HFM_GLOBAL T* b = gin<T>(t.b);
  HFM_GLOBAL T* bc = gout<T>(t.xout);
  const int W = t.W, H = t.H, D = t.D;
  if (vec_task(t)) {
    const int nc0 = base + bk.tid * 4;
    const bool cvalid = nc0 < Nc;
    int I0 = 0, J = 0, K = 0;
    if (cvalid) {
      const uint32_t rowc = fdiv(uint32_t(nc0), t.magWc);
      I0 = nc0 - int(rowc) * t.Wc;
      K = int(fdiv(rowc, t.magHc));
      J = int(rowc) - K * t.Hc;
    }
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
      for (int bb = 0; bb < 2; ++bb) {
        const int k = 2 * K + a, j = 2 * J + bb;
        const bool valid = cvalid && k < D && j < H;
        Run r;  // a missing child row (odd extent, tail lane) reads run 0, adds nothing
        r.i0 = valid ? 2 * I0 : 0;
        r.j = valid ? j : 0;
        r.k = valid ? k : 0;
        r.n0 = (r.k * H + r.j) * W + r.i0;
        RunCoef<T> q;
        q.load(t, r);
        float d[8], xc[8], off[8], bv[8];
        ld8(b + r.n0, bv);
        q.diag(d);
        run_offdiag<float>(t, r, q, SrcPlain<T>{x}, xc, off);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const float r2 = resid_val(xc[2 * m], d[2 * m], bv[2 * m], off[2 * m]);
          const float r3 = resid_val(xc[2 * m + 1], d[2 * m + 1], bv[2 * m + 1], off[2 * m + 1]);
          if (valid) {
            s[m] += r2;
            s[m] += r3;
          }
        }
      }
    }
    if (cvalid) {
      float o[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) o[m] = s[m] * 0.125f;
      st4(bc + nc0, o);
    }
    return;
  }
  if (base >= Nc) return;
  for (int s = 0; s < 4; ++s) {
    const int nc = base + s * B + bk.tid;
    if (nc >= Nc) break;
    put(bc + nc, from_f<T>(restrict_one<T>(t, nc, x, b)));
  }
}

template <typename T, int B>
__global__ __launch_bounds__(B) void k_resid_restrict(const hfm_mg_task* __restrict__ tasks,
                                                      const uint16_t* __restrict__ blk_task,
                                                      const hfm_mg_scalars* __restrict__ sc) {
  resid_restrict_body<T, B>(tasks, blk_task, sc, launch_blk());
}

// ------------------------------------------------------------------ fused pass pairs
// Two consecutive passes of a level in one launch; the first pass's result
// never leaves the chip.  A block owns a tile of one task: `tile_rows` rows
// of `tile_planes` planes, plus one halo row on each side when the tile does
// not span the plane and one halo plane on each side.  It walks the tile
// plane by plane: pass 1 (a Jacobi sweep) on plane kk, then pass 2 on plane
// kk - 1, which needs pass 1 on planes kk - 2 .. kk.  A thread owns a run of
// kPairRun x-cells of one tile row.  The plane's pass-1 input, its south
// faces and pass 1's result go through LDS (in-plane neighbours); z-neighbours
// stay in registers, and each plane's rows are loaded a plane ahead (a
// deeper prefetch costs registers, and the walk is bound by its instructions
// more than by latency).  The halo rows and planes are computed by pass 1 only.
//   SR (!S2): pass 1 PLAIN sweep xin -> xout, pass 2 the residual b - A xout
//     summed over each coarse cell's (up to 2x2x2) children / 8 -> rlow.
//   S2:       pass 1 CORRECT sweep (xin + alpha P ec), pass 2 PLAIN sweep -> xout
//     (type O), with the r.z partials when DOT.
// Pass 1's result is rounded to T as a stored sweep is, and both passes use
// the arithmetic of k_sweep / k_resid_restrict, so the outputs equal the
// separate launches bitwise (the r.z partials are summed per tile instead).
// Short runs and large blocks keep the per-thread state small: the walk is
// latency-bound, so it needs waves more than it needs wide rows.
constexpr int kPairRun = 4;
constexpr int kPairBlock = 512;
constexpr int kPairPlane = kPairRun * kPairBlock;  // cells of a tile plane, halo rows included

// kPairRun values as loaded (storage type T), unpacked to float where used,
// so the rows a thread keeps across planes cost T's width in registers.
template <typename T>
struct Raw4;
template <>
struct Raw4<float> {
  f4v a;
  __device__ __forceinline__ void load(const HFM_GLOBAL float* p) { a = *(const HFM_GLOBAL f4v*)p; }
  __device__ __forceinline__ void load_nt(const HFM_GLOBAL float* p) {
    a = __builtin_nontemporal_load((const HFM_GLOBAL f4v*)p);
  }
  __device__ __forceinline__ void lds_load(const float* p) { a = *(const f4v*)p; }
  __device__ __forceinline__ void lds_store(float* p) const { *(f4v*)p = a; }
  __device__ __forceinline__ void get(float v[4]) const {
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  }
  __device__ __forceinline__ void set(const float v[4]) { a = f4v{v[0], v[1], v[2], v[3]}; }
  __device__ __forceinline__ void zero() { a = f4v{0.0f, 0.0f, 0.0f, 0.0f}; }
  __device__ __forceinline__ void store_nt(HFM_GLOBAL float* p) const {
    __builtin_nontemporal_store(a, (HFM_GLOBAL f4v*)p);
  }
};
template <>
struct Raw4<bf16> {
  u2v u;
  __device__ __forceinline__ void load(const HFM_GLOBAL bf16* p) { u = *(const HFM_GLOBAL u2v*)p; }
  __device__ __forceinline__ void load_nt(const HFM_GLOBAL bf16* p) {
    u = __builtin_nontemporal_load((const HFM_GLOBAL u2v*)p);
  }
  __device__ __forceinline__ void lds_load(const bf16* p) { u = *(const u2v*)p; }
  __device__ __forceinline__ void lds_store(bf16* p) const { *(u2v*)p = u; }
  __device__ __forceinline__ void get(float v[4]) const {
    v[0] = __uint_as_float(u.x << 16); v[1] = __uint_as_float(u.x & 0xFFFF0000u);
    v[2] = __uint_as_float(u.y << 16); v[3] = __uint_as_float(u.y & 0xFFFF0000u);
  }
  __device__ __forceinline__ void set(const float v[4]) {  // rounded to BF16 here
    u = u2v{pack2(v[0], v[1]), pack2(v[2], v[3])};
  }
  __device__ __forceinline__ void zero() { u = u2v{0u, 0u}; }
  __device__ __forceinline__ void store_nt(HFM_GLOBAL bf16* p) const {
    __builtin_nontemporal_store(u, (HFM_GLOBAL u2v*)p);
  }
};
template <>
struct Raw4<f16> {
  u2v u;
  __device__ __forceinline__ void load(const HFM_GLOBAL f16* p) { u = *(const HFM_GLOBAL u2v*)p; }
  __device__ __forceinline__ void load_nt(const HFM_GLOBAL f16* p) {
    u = __builtin_nontemporal_load((const HFM_GLOBAL u2v*)p);
  }
  __device__ __forceinline__ void lds_load(const f16* p) { u = *(const u2v*)p; }
  __device__ __forceinline__ void lds_store(f16* p) const { *(u2v*)p = u; }
  __device__ __forceinline__ void get(float v[4]) const {
    v[0] = f16_lo(u.x); v[1] = f16_hi(u.x);
    v[2] = f16_lo(u.y); v[3] = f16_hi(u.y);
  }
  __device__ __forceinline__ void set(const float v[4]) {  // rounded to FP16 here
    u = u2v{pack2h(v[0], v[1]), pack2h(v[2], v[3])};
  }
  __device__ __forceinline__ void zero() { u = u2v{0u, 0u}; }
  __device__ __forceinline__ void store_nt(HFM_GLOBAL f16* p) const {
    __builtin_nontemporal_store(u, (HFM_GLOBAL u2v*)p);
  }
};
__device__ __forceinline__ void lds_ld4(const float* p, float v[4]) {
  const f4v a = *(const f4v*)p;
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
}
__device__ __forceinline__ void lds_st4(float* p, const float v[4]) {
  *(f4v*)p = f4v{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void lds_ld4(const bf16* p, float v[4]) {
  Raw4<bf16> r;
  r.lds_load(p);
  r.get(v);
}
__device__ __forceinline__ float lds_ld1(const float* p) { return *p; }
__device__ __forceinline__ float lds_ld1(const bf16* p) { return to_f(*p); }
__device__ __forceinline__ void lds_ld4(const f16* p, float v[4]) {
  Raw4<f16> r;
  r.lds_load(p);
  r.get(v);
}
__device__ __forceinline__ float lds_ld1(const f16* p) { return to_f(*p); }
__device__ __forceinline__ void ld2(const HFM_GLOBAL float* p, float v[2]) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v a = *(const HFM_GLOBAL f2v*)p;
  v[0] = a.x; v[1] = a.y;
}
__device__ __forceinline__ void ld2(const HFM_GLOBAL bf16* p, float v[2]) {
  const uint32_t u = *(const HFM_GLOBAL uint32_t*)p;
  v[0] = __uint_as_float(u << 16);
  v[1] = __uint_as_float(u & 0xFFFF0000u);
}
__device__ __forceinline__ void st2(HFM_GLOBAL float* p, const float v[2]) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  *(HFM_GLOBAL f2v*)p = f2v{v[0], v[1]};
}
__device__ __forceinline__ void st2(HFM_GLOBAL bf16* p, const float v[2]) {
  *(HFM_GLOBAL uint32_t*)p = pack2(v[0], v[1]);
}
__device__ __forceinline__ void ld2(const HFM_GLOBAL f16* p, float v[2]) {
  const uint32_t u = *(const HFM_GLOBAL uint32_t*)p;
  v[0] = f16_lo(u);
  v[1] = f16_hi(u);
}
__device__ __forceinline__ void st2(HFM_GLOBAL f16* p, const float v[2]) {
  *(HFM_GLOBAL uint32_t*)p = pack2h(v[0], v[1]);
}
__device__ __forceinline__ void st4_nt(HFM_GLOBAL float* p, const float v[4]) {
  __builtin_nontemporal_store(f4v{v[0], v[1], v[2], v[3]}, (HFM_GLOBAL f4v*)p);
}
__device__ __forceinline__ void st4_nt(HFM_GLOBAL bf16* p, const float v[4]) {
  __builtin_nontemporal_store(u2v{pack2(v[0], v[1]), pack2(v[2], v[3])}, (HFM_GLOBAL u2v*)p);
}
__device__ __forceinline__ void st4_nt(HFM_GLOBAL f16* p, const float v[4]) {
  __builtin_nontemporal_store(u2v{pack2h(v[0], v[1]), pack2h(v[2], v[3])}, (HFM_GLOBAL u2v*)p);
}

// Element n of a field: for 2-byte types a uniform base with a 32-bit byte
// offset, which the loads take as their SGPR base and VGPR offset (no 64-bit
// address arithmetic per load; a task holds fewer than 2^31 cells, so the
// offset fits); a 4-byte field may exceed 4 GiB and takes 64-bit addresses.
// n is taken modulo 2^32: an index computed with a negative step is exact as
// long as the element it names exists.
template <typename T>
__device__ __forceinline__ const HFM_GLOBAL T* elem_at(const HFM_GLOBAL T* base, uint32_t n) {
  if constexpr (sizeof(T) <= 2)
    return (const HFM_GLOBAL T*)((const HFM_GLOBAL char*)base + n * uint32_t(sizeof(T)));
  else
    return base + n;
}

// Two values as loaded (the coarse correction under a run), unpacked where used.
template <typename T>
struct Raw2;
template <>
struct Raw2<float> {
  typedef float f2v __attribute__((ext_vector_type(2)));
  f2v a;
  __device__ __forceinline__ void load(const HFM_GLOBAL float* p) { a = *(const HFM_GLOBAL f2v*)p; }
  __device__ __forceinline__ void get(float v[2]) const { v[0] = a.x; v[1] = a.y; }
  __device__ __forceinline__ void zero() { a = f2v{0.0f, 0.0f}; }
};
template <>
struct Raw2<bf16> {
  uint32_t u;
  __device__ __forceinline__ void load(const HFM_GLOBAL bf16* p) { u = *(const HFM_GLOBAL uint32_t*)p; }
  __device__ __forceinline__ void get(float v[2]) const {
    v[0] = __uint_as_float(u << 16);
    v[1] = __uint_as_float(u & 0xFFFF0000u);
  }
  __device__ __forceinline__ void zero() { u = 0u; }
};
template <>
struct Raw2<f16> {
  uint32_t u;
  __device__ __forceinline__ void load(const HFM_GLOBAL f16* p) { u = *(const HFM_GLOBAL uint32_t*)p; }
  __device__ __forceinline__ void get(float v[2]) const { v[0] = f16_lo(u); v[1] = f16_hi(u); }
  __device__ __forceinline__ void zero() { u = 0u; }
};

// A run's pass-1 input as loaded: x, or x and the coarse correction below it
// (x + alpha P ec is formed where it is used, by correct_val as SrcCorrect).
template <typename T, bool CORR>
struct XRun {
  Raw4<T> x;
  Raw2<T> e;
  __device__ __forceinline__ void load(const HFM_GLOBAL T* xp, const HFM_GLOBAL T* ec, int n, int k,
                                       int j, int i0, int Hc, int Wc) {
    x.load(xp + n);
    if (CORR) e.load(ec + ((k >> 1) * Hc + (j >> 1)) * Wc + (i0 >> 1));
  }
  // from the run's and the correction's addresses
  __device__ __forceinline__ void load_at(const HFM_GLOBAL T* xp, const HFM_GLOBAL T* ep) {
    x.load(xp);
    if (CORR) e.load(ep);
  }
  // a plane past the grid's ends: get() then yields +0 (x + alpha * 0 = +0)
  __device__ __forceinline__ void zero() {
    x.zero();
    e.zero();
  }
  __device__ __forceinline__ void get(float v[4], float alpha) const {
    x.get(v);
    if (CORR) {
      float ev[2];
      e.get(ev);
#pragma unroll
      for (int m = 0; m < 4; ++m) v[m] = correct_val(v[m], ev[m >> 1], alpha);
    }
  }
};

// Diagonal and off-diagonal sum of a run from its face rows (cx: west faces,
// ef: the east face of the run's last cell, cy / cn: south / north, cd / cu:
// down / up) and neighbour values, in the order and with the rounding of
// RunCoef::diag and offdiag_sum; each face row is unpacked where it is used.
// CALC_D = false: d holds the diagonal already (the same faces' sum).
template <typename T, bool CALC_D = true>
__device__ __forceinline__ void stencil_run(const Raw4<T>& cx, float ef, const Raw4<T>& cy,
                                            const Raw4<T>& cn, const Raw4<T>& cd,
                                            const Raw4<T>& cu, const float xc[4], float wv,
                                            float ev, const float vs[4], const float vn[4],
                                            const float vd[4], const float vu[4], bool south,
                                            bool north, bool down, bool up, float d[4],
                                            float off[4]) {
  constexpr int N = kPairRun;
  float c[N], ce[N];
  cx.get(c);
#pragma unroll
  for (int m = 0; m < N - 1; ++m) ce[m] = c[m + 1];
  ce[N - 1] = ef;
#pragma unroll
  for (int m = 0; m < N; ++m) {
    if (CALC_D) d[m] = c[m] + ce[m];
    off[m] = c[m] * (m ? xc[m - 1] : wv);
    off[m] = fma_a(ce[m], m < N - 1 ? xc[m + 1] : ev, off[m]);
  }
  cy.get(c);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    if (CALC_D) d[m] += c[m];
    off[m] = fma_a(c[m], south ? vs[m] : 0.0f, off[m]);
  }
  cn.get(c);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    if (CALC_D) d[m] += c[m];
    off[m] = fma_a(c[m], north ? vn[m] : 0.0f, off[m]);
  }
  cd.get(c);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    if (CALC_D) d[m] += c[m];
    off[m] = fma_a(c[m], down ? vd[m] : 0.0f, off[m]);
  }
  cu.get(c);
#pragma unroll
  for (int m = 0; m < N; ++m) {
    if (CALC_D) d[m] += c[m];
    off[m] = fma_a(c[m], up ? vu[m] : 0.0f, off[m]);
  }
}

// Loads of one plane kk for pass 1, issued a plane ahead of their use.
template <typename T, bool CORR>
struct PlaneIn {
  XRun<T, CORR> xn;    // pass-1 input on plane kk + 1 (own run)
  XRun<T, CORR> edge;  // pass-1 input of the row past the tile's edge (edge rows only)
  Raw4<T> cx, cy, cu, b, cn;  // faces of plane kk; cn: the north faces (from memory
                              // for edge and boundary rows, else from LDS)
  T ef;                       // the east face of the run's last cell, as loaded
  Raw4<T> y;                  // pass 1's result on plane kk
};

// Sum of v over a kPairBlock-thread block (valid in thread 0).
__device__ __forceinline__ double pair_block_sum(double v) {
  __shared__ double part[kPairBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kPairBlock / 64; ++w) s += part[w];
  }
  return s;
}

// The pair's LDS.  A plane's rows are laid out from row 0, or, in a tile
// with halo rows, from row 1: rows 0 and NR + 1 then hold the rows just past
// the tile's first and last rows, which the threads of those rows load and
// put there.  A tile with halo rows has at most kPairBlock / 4 threads per
// row, so W <= kPairBlock and the NR + 2 rows take at most kPairPlane +
// kPairEdge cells.  lx / ly end in a run of zeros (at kPairZero), which a
// thread reads in place of a neighbour past the grid's edge: the masked
// form's +0, without a select per cell.
constexpr int kPairEdge = 2 * kPairBlock;
constexpr int kPairZero = kPairPlane + kPairEdge;
template <typename T, bool S2>
struct PairLds {
  float lx[2][kPairZero + kPairRun];  // pass-1 input of the plane
  float ly[2][kPairZero + kPairRun];  // pass-1 result of the plane, rounded to T
  float ld[2][kPairPlane + kPairEdge / 2];  // the plane's diagonal, from pass 1 to pass 2:
                                      // each thread's run at its own lx offset (no barrier,
                                      // no registers; rows 0 .. NR of the layout)
  float lr[2][S2 ? 1 : kPairPlane / 2];  // odd rows' residuals (restriction)
  double acc[S2 ? kPairBlock : 1];       // each thread's r.z partial (its own slot: no
                                         // barrier; in LDS, it costs the walk no registers)
};

template <typename T, bool S2, typename O, bool DOT>
__device__ __forceinline__ void smooth_pair_body(const hfm_mg_task* __restrict__ tasks,
                                                 const uint16_t* __restrict__ blk_task,
                                                 const hfm_mg_scalars* __restrict__ sc,
                                                 float omega, float alpha, double* dots,
                                                 PairLds<T, S2>& lds, const Blk& bk) {
  constexpr int N = kPairRun;
  constexpr int kZero = kPairZero;
  auto& lx = lds.lx;
  auto& ly = lds.ly;
  auto& ld = lds.ld;
  auto& lr = lds.lr;
  // Tiles are numbered contiguously over the launch's tasks (a task's tiles
  // are blk0 .. blk0 + nblk - 1); each XCD takes a contiguous range of them,
  // so a task's neighbouring tiles (which re-read each other's halo) share an
  // L2 and tasks with one tile still spread over the XCDs.  The table is
  // padded to a multiple of 8 entries with 0xFFFF.
  const int tile = xcd_chunk(bk.g, bk.nblocks);
  const int task = blk_task[tile];
  if (task == 0xFFFF) return;
  const hfm_mg_task& t = tasks[task];
  if (!sc[t.inst].active) return;
  const int W = t.W, H = t.H, D = t.D, HW = H * W;
  const int R = t.tile_rows, Z = t.tile_planes;
  const int ny = (H + R - 1) / R;
  const int lb = tile - t.blk0;
  if (DOT && bk.tid == 0) {
    // the finalize sums every dot slot of the instance: the ones past the
    // tiles hold zero
    const int nslot = sc[t.inst].nslot;
    for (int s = t.nblk + lb; s < nslot; s += t.nblk) dots[t.slot0 + s] = 0.0;
  }
  const int yb = lb % ny, zb = lb / ny;
  const int Q = W / N;
  const int jh = R < H ? 1 : 0;  // halo rows
  const int Rb = min(R, H - yb * R);
  const int NR = Rb + 2 * jh;
  const int r = int(bk.tid) / Q;
  const int i0 = N * (int(bk.tid) - r * Q);
  const int j = yb * R - jh + r;
  // live: the thread's row is a row of the plane (tile or halo); rows past
  // the plane's edges and threads past the tile do no work
  const bool live = r < NR && j >= 0 && j < H;
  const bool outrow = live && r >= jh && r < jh + Rb;
  const int jc = min(max(j, 0), H - 1);
  const int li = live ? (r + jh) * W + i0 : 0;
  const bool south = j > 0, north = j + 1 < H, west = i0 > 0, east = i0 + N < W;
  // the rows next to the tile's first and last rows are not in LDS: the
  // threads of those rows load them (every other thread loads its own row
  // in their place, so that every thread issues the same loads)
  const bool edge_s = r == 0 && south, edge_n = r + 1 == NR && north;
  const int je = edge_s ? jc - 1 : (edge_n ? jc + 1 : jc);
  // LDS offsets of the in-plane neighbours (rows r - 1 / r + 1, the cells
  // before / after the run), or of the zeros where the grid ends there or
  // the neighbour row comes from memory
  const int os = south ? li - W : kZero, on = north ? li + W : kZero;
  const int ow = west ? li - 1 : kZero, oe = east ? li + N : kZero;
  if (bk.tid < N) {  // read after the walk's first barrier
    lx[0][kZero + bk.tid] = 0.0f;
    lx[1][kZero + bk.tid] = 0.0f;
    ly[0][kZero + bk.tid] = 0.0f;
    ly[1][kZero + bk.tid] = 0.0f;
  }
  const int k0 = zb * Z, k1 = min(k0 + Z, D);  // output planes [k0, k1)
  const int ks = max(k0 - 1, 0);
  const int kmax = min(k1, D - 1);  // the last plane the walk loads
  const HFM_GLOBAL T* xin = gin<T>(t.xin);
  const HFM_GLOBAL T* ec = gin<T>(t.ec);
  const HFM_GLOBAL T* cx = gin<T>(t.c[0]);
  const HFM_GLOBAL T* cy = gin<T>(t.c[1]);
  const HFM_GLOBAL T* cz = gin<T>(t.c[2]);
  const HFM_GLOBAL T* b = gin<T>(t.b);
  const int Hc = t.Hc, Wc = t.Wc;
  auto cell = [&](int k) { return (k * H + jc) * W + i0; };
  const HFM_GLOBAL T* bxh = gin<T>(t.c[3]);
  const HFM_GLOBAL T* byh = gin<T>(t.c[4]);
  const HFM_GLOBAL T* bzh = gin<T>(t.c[5]);
  // The loads of a plane: every thread issues all of them, unconditionally,
  // from in-bounds addresses (rows outside the tile read the clamped row jc),
  // and every value goes into its record as loaded.  So the compiler counts
  // the loads in flight exactly and waits for a plane's loads only where
  // they are first used, a plane later; a load under a branch, or a
  // conversion or select at the load, makes it wait for every load at once.
  // On the last plane the x-row above is read from the plane itself and
  // taken as +0 where it is used.
  // element offsets: of the thread's run in a plane, of a plane, of the edge
  // row from the thread's row (modulo 2^32: it may step back a row), of the
  // correction under the thread's and the edge row's run in a coarse plane,
  // of a coarse plane
  const uint32_t row = uint32_t(jc * W + i0), uHW = uint32_t(HW);
  const uint32_t edge = uint32_t((je - jc) * W);
  const uint32_t ecown = uint32_t((jc >> 1) * Wc + (i0 >> 1));
  const uint32_t ecedge = uint32_t((je >> 1) * Wc + (i0 >> 1));
  const uint32_t HWc = uint32_t(Hc * Wc);
  auto load_plane = [&](PlaneIn<T, S2>& pi, int kk) {
    const int n = cell(kk);
    const bool up1 = kk + 1 < D;
    const uint32_t nu = uint32_t(kk) * uHW + row;
    const int ku = up1 ? kk + 1 : kk;
    pi.xn.load_at(elem_at(xin, up1 ? nu + uHW : nu), elem_at(ec, uint32_t(ku >> 1) * HWc + ecown));
    pi.cx.load(elem_at(cx, nu));
    pi.cy.load(elem_at(cy, nu));
    pi.cu.load(elem_at(up1 ? cz : bzh, up1 ? nu + uHW : row));
    pi.b.load_nt(elem_at(b, nu));
    pi.ef = ldraw(east ? cx + n + N : bxh + kk * H + jc);
    pi.edge.load_at(elem_at(xin, nu + edge), elem_at(ec, uint32_t(kk >> 1) * HWc + ecedge));
    pi.cn.load(north ? cy + n + W : byh + kk * W + i0);
  };

  // The walk keeps three plane records in rotation (previous, current, the
  // one being loaded); the step is written once and unrolled three times so
  // that the rotation costs no register moves.  Pass 1 takes its down
  // neighbours from the other LDS buffer, which holds its own run of the
  // plane below.  The z-neighbours past the grid's ends are +0 (the x-row
  // below plane 0 in LDS, the x-row above plane D - 1 and the pass-1 results
  // of planes -1 and D in the records), so both passes take them unmasked:
  // the +0 that the masked form substitutes.
  PlaneIn<T, S2> pa, pb, pc;
  float ev0[N];             // an even row's residuals on an even plane
#pragma unroll
  for (int m = 0; m < N; ++m) ev0[m] = 0.0f;
  pa.y.set(ev0);  // zeros
  pc.y.set(ev0);
  {
    // pa stands for plane ks - 1 (its x-row is x(ks), its up faces cz(ks));
    // pb is plane ks.  The pass-1 input on plane ks - 1 goes to the LDS
    // buffer that step ks reads its down neighbours from (a thread reads
    // only its own run there, so no barrier orders it).
    float xd[N];
    if (ks > 0) {
      XRun<T, S2> xr;
      xr.load(xin, ec, cell(ks - 1), ks - 1, jc, i0, Hc, Wc);
      xr.get(xd, alpha);
    } else {
#pragma unroll
      for (int m = 0; m < N; ++m) xd[m] = 0.0f;
    }
    if (live) lds_st4(&lx[(ks & 1) ^ 1][li], xd);
    pa.xn.load(xin, ec, cell(ks), ks, jc, i0, Hc, Wc);
    pa.cu.load(cz + cell(ks));
    load_plane(pb, ks);
  }
  if (DOT) lds.acc[bk.tid] = 0.0;

  auto step = [&](int kk, PlaneIn<T, S2>& prv, PlaneIn<T, S2>& cur, PlaneIn<T, S2>& spare) {
    const bool p1 = kk < D;  // pass 1 runs on plane kk (uniform)
    const int buf = kk & 1;
    // what the pass pair still needs of plane kk - 2 before its record is reused
    const Raw4<T> pcd = spare.cu;      // down faces of plane kk - 1
    const Raw4<T> ym2 = spare.y;       // pass 1 on plane kk - 2
    float xc[N], yu[N];
    if (p1 && live) {
      prv.xn.get(xc, alpha);
      lds_st4(&lx[buf][li], xc);
      if (edge_s || edge_n) {  // the row past the tile's edge, from memory
        float ve[N];
        cur.edge.get(ve, alpha);
        lds_st4(&lx[buf][edge_s ? os : on], ve);
      }
    }
    // the next plane (past the tile's or the grid's last one: the last one
    // again, whose values are not used), after this plane's edge row is used,
    // so that the two records' edge rows need no registers at the same time
    load_plane(spare, min(kk + 1, kmax));
    __syncthreads();
    if (p1 && live) {
      float vs[N], vn[N], vd[N], vu[N], d[N], off[N];
      lds_ld4(&lx[buf][os], vs);
      lds_ld4(&lx[buf][on], vn);
      lds_ld4(&lx[buf ^ 1][li], vd);  // this run on plane kk - 1
      cur.xn.get(vu, alpha);
      if (kk + 1 >= D) {  // past the grid's last plane (uniform): +0
#pragma unroll
        for (int m = 0; m < N; ++m) vu[m] = 0.0f;
      }
      const float wv = lx[buf][ow];
      const float ev = lx[buf][oe];
      stencil_run<T>(cur.cx, to_f(cur.ef), cur.cy, cur.cn, prv.cu, cur.cu, xc, wv, ev, vs, vn, vd,
                     vu, true, true, true, true, d, off);
      lds_st4(&ld[buf][li], d);  // for pass 2 on this plane, a step later
      float bv[N], y[N];
      cur.b.get(bv);
#pragma unroll
      for (int m = 0; m < N; ++m) y[m] = jacobi_val(xc[m], d[m], bv[m], off[m], omega);
      // y is rounded to T where it is packed (the record, the store): one
      // conversion per two cells; pass 2 takes the rounded values as floats
      // (from LDS, or yu for this run)
      cur.y.set(y);
      if (!S2 && outrow && kk >= k0 && kk < k1) cur.y.store_nt(gout<T>(t.xout) + cell(kk));
      cur.y.get(yu);
      lds_st4(&ly[buf][li], yu);
    } else {
#pragma unroll
      for (int m = 0; m < N; ++m) yu[m] = 0.0f;
      cur.y.set(yu);  // plane kk is past the last (its pass-1 value is unused)
    }
    if (kk - 1 >= k0) {  // pass 2 on plane kp (uniform)
      const int kp = kk - 1;
      const int pb2 = kp & 1;
      float e[N];
      if (outrow) {
        float yc[N], vs[N], vn[N], vd[N], d[N], off[N], bv[N];
        lds_ld4(&ly[pb2][li], yc);
        ym2.get(vd);
        lds_ld4(&ly[pb2][os], vs);
        lds_ld4(&ly[pb2][on], vn);
        lds_ld4(&ld[pb2][li], d);
        const float wv = ly[pb2][ow];
        const float ev = ly[pb2][oe];
        stencil_run<T, false>(prv.cx, to_f(prv.ef), prv.cy, prv.cn, pcd, prv.cu, yc, wv, ev, vs,
                              vn, vd, yu, true, true, true, true, d, off);
        prv.b.get(bv);
        if constexpr (S2) {
          float o[N];
#pragma unroll
          for (int m = 0; m < N; ++m) o[m] = jacobi_val(yc[m], d[m], bv[m], off[m], omega);
          Raw4<O> ro;
          ro.set(o);  // z as stored; the dot takes the stored values
          ro.store_nt(gout<O>(t.xout) + cell(kp));
          if (DOT) {
            ro.get(o);
            double acc = lds.acc[bk.tid];
#pragma unroll
            for (int m = 0; m < N; ++m) acc += double(bv[m]) * double(o[m]);
            lds.acc[bk.tid] = acc;
          }
        } else {
#pragma unroll
          for (int m = 0; m < N; ++m) e[m] = resid_val(yc[m], d[m], bv[m], off[m]);
        }
      }
      if constexpr (!S2) {
        // rows 2J and 2J+1 of planes 2K and 2K+1 make coarse row (K, J): odd
        // rows hand their residuals to the even row, which sums the children
        // in k_resid_restrict's order (plane, row, cell)
        const int pidx = ((r - jh) >> 1) * W + i0;
        const bool odd = (j & 1) != 0;
        if (outrow && odd) lds_st4(&lr[pb2][pidx], e);
        if (outrow && !odd && !(kp & 1)) {
#pragma unroll
          for (int m = 0; m < N; ++m) ev0[m] = e[m];
        }
        if ((kp & 1) || kp == D - 1) {
          __syncthreads();
          if (outrow && !odd) {
            const bool pair = north;        // row 2J+1 exists
            const bool second = (kp & 1);   // plane 2K+1 exists
            float s2v[N / 2], o0[N];
#pragma unroll
            for (int m = 0; m < N / 2; ++m) {
              s2v[m] = 0.0f;
              s2v[m] += ev0[2 * m];
              s2v[m] += ev0[2 * m + 1];
            }
            if (pair) {
              lds_ld4(&lr[0][pidx], o0);
#pragma unroll
              for (int m = 0; m < N / 2; ++m) {
                s2v[m] += o0[2 * m];
                s2v[m] += o0[2 * m + 1];
              }
            }
            if (second) {
#pragma unroll
              for (int m = 0; m < N / 2; ++m) {
                s2v[m] += e[2 * m];
                s2v[m] += e[2 * m + 1];
              }
              if (pair) {
                lds_ld4(&lr[1][pidx], o0);
#pragma unroll
                for (int m = 0; m < N / 2; ++m) {
                  s2v[m] += o0[2 * m];
                  s2v[m] += o0[2 * m + 1];
                }
              }
            }
            float o[N / 2];
#pragma unroll
            for (int m = 0; m < N / 2; ++m) o[m] = s2v[m] * 0.125f;
            st2(gout<T>(t.rlow) + ((kp >> 1) * Hc + (j >> 1)) * Wc + (i0 >> 1), o);
          }
        }
      }
    }
  };
  for (int kk = ks; kk <= k1; kk += 3) {
    step(kk, pa, pb, pc);
    if (kk + 1 <= k1) step(kk + 1, pb, pc, pa);
    if (kk + 2 <= k1) step(kk + 2, pc, pa, pb);
  }
  if (DOT) {
    const double s = pair_block_sum(lds.acc[bk.tid]);
    if (bk.tid == 0) dots[t.slot0 + lb] = s;
  }
}

template <typename T, bool S2, typename O, bool DOT>
__global__ __launch_bounds__(kPairBlock, sizeof(T) == 2 ? 4 : 2) void k_smooth_pair(
    const hfm_mg_task* __restrict__ tasks, const uint16_t* __restrict__ blk_task,
    const hfm_mg_scalars* __restrict__ sc, float omega, float alpha, double* dots) {
  __shared__ PairLds<T, S2> lds;
  smooth_pair_body<T, S2, O, DOT>(tasks, blk_task, sc, omega, alpha, dots, lds, launch_blk());
}

// ------------------------------------------------------------------ coarsest solve
template <typename T>
__global__ __launch_bounds__(64) void k_coarse_solve(const hfm_mg_task* __restrict__ tasks,
                                                     const hfm_mg_scalars* __restrict__ sc) {
  const hfm_mg_task& t = tasks[blockIdx.x];
  if (!sc[t.inst].active) return;
  const int N = t.D * t.H * t.W;
  if (N > 64) return;  // host rejects such levels; never index past the LDS tile
  const HFM_GLOBAL double* ainv = gin<double>(t.ec);
  const HFM_GLOBAL T* b = gin<T>(t.b);
  __shared__ double bs[64];
  if (threadIdx.x < N) bs[threadIdx.x] = double(ldg(b + threadIdx.x));
  __syncthreads();
  const int i = threadIdx.x;
  if (i < N) {
    double e = 0.0;
    for (int j = 0; j < N; ++j) e += ainv[j * N + i] * bs[j];
    put(gout<T>(t.xout) + i, from_f<T>(float(e)));
  }
}

// ------------------------------------------------------------------ coarse-level cycle
// One workgroup per instance runs the V-cycle's tail: every level from the
// chain's first (its rhs B already restricted into place) to the coarsest and
// back, with barriers between passes, so the small levels cost one launch
// instead of ~5 per level.  Record q of an instance's chain describes level
// lc + q: D/H/W, Dc/Hc/Wc + magics of the next level, c = storage faces,
// xin = X, xout = Y (the sweep ping-pong), b = B, ec = the next level's B (the
// restriction target), or, for the coarsest, the dense FP64 inverse.  The
// buffers and the per-cell arithmetic are those of the per-level launches
// (sweep k writes (X, Y)[k % 2]; the level's result ends in
// (X, Y)[(nu1 + nu2 - 1) % 2], the coarsest's in X).
constexpr int kCycleThreads = 1024;

// NT threads (a whole workgroup) run instance bk.g's tail; the per-cell loops
// stride by NT, so the result does not depend on it.
template <typename T, int NT>
__device__ __forceinline__ void coarse_cycle_body(const hfm_mg_task* __restrict__ recs,
                                                  const int* __restrict__ chain,
                                                  const hfm_mg_scalars* __restrict__ sc, int nu1,
                                                  int nu2, float omega, float alpha,
                                                  const Blk& bk) {
  const int first = chain[2 * bk.g], count = chain[2 * bk.g + 1];
  if (count < 1 || !sc[recs[first].inst].active) return;
  const int tid = bk.tid;
  // xin and ec are read-only for the per-level kernels; here X and the next
  // level's B are written as well
  auto X = [&](int q) { return gout<T>(const_cast<void*>(recs[first + q].xin)); };
  auto Y = [&](int q) { return gout<T>(recs[first + q].xout); };
  auto buf = [&](int q, int kk) { return (kk & 1) ? Y(q) : X(q); };
  const int last = count - 1;
  // down
  for (int q = 0; q < last; ++q) {
    const hfm_mg_task& t = recs[first + q];
    const int N = t.D * t.H * t.W;
    const HFM_GLOBAL T* b = gin<T>(t.b);
    for (int kk = 0; kk < nu1; ++kk) {
      const HFM_GLOBAL T* xin = buf(q, kk - 1);
      HFM_GLOBAL T* out = buf(q, kk);
      for (int n = tid; n < N; n += NT) {
        int i, j, k;
        decompose(t, n, i, j, k);
        const float bn = ldg(b + n);
        const float r = kk == 0 ? sweep_one<T, HFM_SWEEP_ZERO>(t, n, i, j, k, xin, nullptr, bn,
                                                               omega, alpha)
                                : sweep_one<T, HFM_SWEEP_PLAIN>(t, n, i, j, k, xin, nullptr, bn,
                                                                omega, alpha);
        put(out + n, from_f<T>(r));
      }
      __syncthreads();
    }
    const HFM_GLOBAL T* x = buf(q, nu1 - 1);
    HFM_GLOBAL T* bc = gout<T>(const_cast<void*>(t.ec));
    const int Nc = t.Dc * t.Hc * t.Wc;
    for (int nc = tid; nc < Nc; nc += NT) put(bc + nc, from_f<T>(restrict_one<T>(t, nc, x, b)));
    __syncthreads();
  }
  // coarsest: x = Ainv b
  {
    const hfm_mg_task& t = recs[first + last];
    const int N = t.D * t.H * t.W;
    if (N <= 64) {
      const HFM_GLOBAL double* ainv = gin<double>(t.ec);
      const HFM_GLOBAL T* b = gin<T>(t.b);
      __shared__ double bs[64];
      if (tid < N) bs[tid] = double(ldg(b + tid));
      __syncthreads();
      if (tid < N) {
        double e = 0.0;
        for (int jj = 0; jj < N; ++jj) e += ainv[jj * N + tid] * bs[jj];
        put(X(last) + tid, from_f<T>(float(e)));
      }
    }
    __syncthreads();
  }
  // up
  for (int q = last - 1; q >= 0; --q) {
    const hfm_mg_task& t = recs[first + q];
    const int N = t.D * t.H * t.W;
    const HFM_GLOBAL T* b = gin<T>(t.b);
    const HFM_GLOBAL T* ec = q + 1 == last ? X(last) : buf(q + 1, nu1 + nu2 - 1);
    for (int s2 = 0; s2 < nu2; ++s2) {
      const int kk = nu1 + s2;
      const HFM_GLOBAL T* xin = buf(q, kk - 1);
      HFM_GLOBAL T* out = buf(q, kk);
      for (int n = tid; n < N; n += NT) {
        int i, j, k;
        decompose(t, n, i, j, k);
        const float bn = ldg(b + n);
        const float r = s2 == 0 ? sweep_one<T, HFM_SWEEP_CORRECT>(t, n, i, j, k, xin, ec, bn,
                                                                  omega, alpha)
                                : sweep_one<T, HFM_SWEEP_PLAIN>(t, n, i, j, k, xin, nullptr, bn,
                                                                omega, alpha);
        put(out + n, from_f<T>(r));
      }
      __syncthreads();
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kCycleThreads) void k_coarse_cycle(
    const hfm_mg_task* __restrict__ recs, const int* __restrict__ chain,
    const hfm_mg_scalars* __restrict__ sc, int nu1, int nu2, float omega, float alpha) {
  coarse_cycle_body<T, kCycleThreads>(recs, chain, sc, nu1, nu2, omega, alpha, launch_blk());
}

// ------------------------------------------------------------------ FCG passes (fine level)
// Coefficients c are the FP32 fine faces; products accumulated in FP64.
template <typename T, bool SWEEP>
__global__ __launch_bounds__(kBlock) void k_fcg_init(const hfm_mg_task* __restrict__ tasks,
                                                     const uint16_t* __restrict__ blk_task,
                                                     const hfm_mg_scalars* __restrict__ sc,
                                                     float omega, double* dots) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (!S.active || S.iter != -1) return;  // staged slots only
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kCellChunk;
  double acc = 0.0;
  HFM_GLOBAL T* rlow = gout<T>(t.rlow);
  HFM_GLOBAL T* x0 = gout<T>(t.xout);
  const HFM_GLOBAL float* b32 = gin<float>(t.b32);
  HFM_GLOBAL double* x64 = gout<double>(t.x64);
  HFM_GLOBAL double* r64 = gout<double>(t.r64);
  const bool ilv = vec_task(t);
  for (int s = 0; s < kRun; ++s) {
    const int n = base + s * kBlock + threadIdx.x;
    if (n >= N) break;
    const double bv = double(b32[n]);
    const int n64 = ilv ? ilv_cell(n) : n;
    x64[n64] = 0.0;
    r64[n64] = bv;
    const T rt = from_f<T>(rscale<T>(float(bv), S.sexp));
    put(rlow + n, rt);
    acc += bv * bv;
    if (SWEEP) {
      int i, j, k;
      decompose(t, n, i, j, k);
      put(x0 + n, from_f<T>(wdiv(omega, to_f(rt), diag_rounded<T>(t, n, i, j, k))));
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) dots[t.slot0 + chunk] = s;
}

// p = T(z + beta p_old); dots = p . (A p) in FP32 per cell (it only sets the
// step length; the residual update recomputes A p in FP64).  The vector path
// evaluates p . A p as the operator's sum over faces,
//   sum_cells cx (p - p_w)^2 + cy (p - p_s)^2 + cz (p - p_d)^2
//             + [high boundary] (bxh | byh | bzh) p^2      (zero ghosts),
// which needs only the west, south and down neighbours.
template <typename T, bool PB = false>
__device__ __forceinline__ void fcg_dir_body(const hfm_mg_task* __restrict__ tasks,
                                             const uint16_t* __restrict__ blk_task,
                                             const hfm_mg_scalars* __restrict__ sc, double* dots,
                                             const Blk& bk) {
  const int g = bk.g;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (!S.active) {
    if (PB) block_sum(0.0, bk.half);  // the workgroup's barriers
    return;
  }
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kCellChunk;
  const SrcDir<T> src{gin<T>(t.b), gin<T>(t.xin), float(S.beta)};
  HFM_GLOBAL T* pnew = gout<T>(t.xout);
  double acc = 0.0;
  if (vec_task(t)) {
    const int n0 = base + bk.tid * kRun;
    const bool valid = n0 < N;
    const Run r = run_at(t, n0, valid);
    const int W = t.W, HW = t.H * t.W;
    const bool west = r.i0 > 0, south = r.j > 0, down = r.k > 0;
    const bool east_b = r.i0 + kRun == W, north_b = r.j + 1 == t.H, up_b = r.k + 1 == t.D;
    float p[8], ps[8], pd[8], cx[8], cy[8], cz[8];
    src.run(r.n0, r.k, r.j, r.i0, p);
    src.run(south ? r.n0 - W : r.n0, r.k, r.j, r.i0, ps);
    src.run(down ? r.n0 - HW : r.n0, r.k, r.j, r.i0, pd);
    const float pw = src.one(west ? r.n0 - 1 : r.n0, r.k, r.j, r.i0);
    ld8(gin<float>(t.c[0]) + r.n0, cx);
    ld8(gin<float>(t.c[1]) + r.n0, cy);
    ld8(gin<float>(t.c[2]) + r.n0, cz);
    if (valid) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const float dw = p[m] - (m ? p[m - 1] : (west ? pw : 0.0f));
        const float ds = p[m] - (south ? ps[m] : 0.0f);
        const float dd = p[m] - (down ? pd[m] : 0.0f);
        acc += double(cx[m] * dw * dw + cy[m] * ds * ds + cz[m] * dd * dd);
      }
    }
    // the high-boundary faces (b p^2 on the last x, y and z layers): only
    // waves that hold such a run load them
    const bool edge = valid && (east_b || north_b || up_b);
    if (__any(edge)) {
      float by[8], bz[8];
      ld8(gin<float>(t.c[4]) + r.k * W + r.i0, by);
      ld8(gin<float>(t.c[5]) + r.j * W + r.i0, bz);
      const float bx = ldg(gin<float>(t.c[3]) + r.k * t.H + r.j);
      if (edge) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          float e = 0.0f;
          if (m == 7 && east_b) e += bx;
          if (north_b) e += by[m];
          if (up_b) e += bz[m];
          acc += double(e * (p[m] * p[m]));
        }
      }
    }
    if (valid) {
      st8_nt(pnew + n0, p);
    }
  } else if (base < N) {
    auto pat = [&](int m, int, int, int) -> float { return src.one(m, 0, 0, 0); };
    for (int s = 0; s < kRun; ++s) {
      const int n = base + s * kBlock + bk.tid;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      float diag;
      const float off = face_row<float, float>(t, n, i, j, k, pat, diag);
      const float p = pat(n, k, j, i);
      put(pnew + n, from_f<T>(p));
      acc += double(p * (diag * p - off));
    }
  }
  const double s = block_sum(acc, bk.half);
  if (bk.tid == 0) dots[t.slot0 + chunk] = s;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fcg_dir(const hfm_mg_task* __restrict__ tasks,
                                                    const uint16_t* __restrict__ blk_task,
                                                    const hfm_mg_scalars* __restrict__ sc,
                                                    double* dots) {
  fcg_dir_body<T>(tasks, blk_task, sc, dots, launch_blk());
}

// r64 -= alpha A p; rlow = T(r64); dots = r . r; x64 every other iteration.
// p is stored as P (the search-direction type), rlow and xout as T.
// The solution update is deferred: with S.xlag == 0 x64 is not touched (the
// finalize CHECK then records alpha_prev and xlag = 1); with S.xlag == 1
// x64 = fma(alpha, p, fma(alpha_prev, p_prev, x64)) with p_prev in ec (the
// other search-direction buffer), the two updates in their own order, so the
// result is bitwise that of updating x64 every iteration while x64 moves
// through HBM on half of the iterations.  k_fcg_flush applies a pending
// update once an instance stops.
// SWEEP: also the next preconditioner's first (zero initial guess) fine-level
// sweep, xout = omega rlow / d with d the diagonal of the storage-rounded faces
// (identical to a separate HFM_SWEEP_ZERO launch).
template <typename T, bool SWEEP, typename P, bool PB = false>
__device__ __forceinline__ void fcg_update_body(const hfm_mg_task* __restrict__ tasks,
                                                const uint16_t* __restrict__ blk_task,
                                                const hfm_mg_scalars* __restrict__ sc,
                                                float omega, double* dots, const Blk& bk) {
  const int g = bk.g;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (!S.active) {
    if (PB) block_sum(0.0, bk.half);  // the workgroup's barriers
    return;
  }
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kCellChunk;
  const double alpha = S.alpha;
  const bool lag = S.xlag != 0;
  const double alpha_prev = S.alpha_prev;
  const int sexp = S.sexp;  // the fine residual's scale (FP16 storage)
  const SrcPlain<P> src{gin<P>(t.xin)};
  const SrcPlain<P> prev{gin<P>(t.ec)};
  HFM_GLOBAL T* rlow = gout<T>(t.rlow);
  HFM_GLOBAL T* x0 = gout<T>(t.xout);
  HFM_GLOBAL double* x64 = gout<double>(t.x64);
  HFM_GLOBAL double* r64 = gout<double>(t.r64);
  double acc = 0.0;
  if (vec_task(t)) {
    const int n0 = base + bk.tid * kRun;
    const bool valid = n0 < N;
    const Run r = run_at(t, n0, valid);
    RunCoef<float> q;
    q.load(t, r);
    double rv[8], xv[8];
    float pp[8];
    ld8d_ilv(r64, r.n0, rv);
    if (lag) {
      ld8d_ilv(x64, r.n0, xv);
      prev.run(r.n0, 0, 0, 0, pp);
    }
    double d[8], p[8], off[8];
    q.diag(d);
    run_offdiag<double>(t, r, q, src, p, off);
    if (valid) {
      float rl[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        rv[m] = rv[m] - alpha * (d[m] * p[m] - off[m]);
        rl[m] = rscale<T>(float(rv[m]), sexp);
        acc += rv[m] * rv[m];
      }
      st8d_ilv(r64, n0, rv);
      st8(rlow + n0, rl);
      if (lag) {
#pragma unroll
        for (int m = 0; m < 8; ++m) xv[m] = fma(alpha, p[m], fma(alpha_prev, double(pp[m]), xv[m]));
        st8d_ilv(x64, n0, xv);
      }
      if (SWEEP) {
        float o[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const float dT = round_to<T>(q.cw[m]) + round_to<T>(q.ce[m]) + round_to<T>(q.cs[m]) +
                           round_to<T>(q.cn[m]) + round_to<T>(q.cd[m]) + round_to<T>(q.cu[m]);
          o[m] = wdiv(omega, round_to<T>(rl[m]), dT);
        }
        st8_nt(x0 + n0, o);
      }
    }
  } else if (base < N) {
    auto pat = [&](int m, int, int, int) -> double { return double(src.one(m, 0, 0, 0)); };
    for (int s = 0; s < kRun; ++s) {
      const int n = base + s * kBlock + bk.tid;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      double diag;
      const double off = face_row<double, float>(t, n, i, j, k, pat, diag);
      const double pn = pat(n, k, j, i);
      const double q = diag * pn - off;
      if (lag) x64[n] = fma(alpha, pn, fma(alpha_prev, double(prev.one(n, 0, 0, 0)), x64[n]));
      const double rr = r64[n] - alpha * q;
      r64[n] = rr;
      const T rt = from_f<T>(rscale<T>(float(rr), sexp));
      put(rlow + n, rt);
      acc += rr * rr;
      if (SWEEP) put(x0 + n, from_f<T>(wdiv(omega, to_f(rt), diag_rounded<T>(t, n, i, j, k))));
    }
  }
  const double s = block_sum(acc, bk.half);
  if (bk.tid == 0) dots[t.slot0 + chunk] = s;
}

// The launch keeps the body written out in the kernel itself (a copy of
// fcg_update_body with PB = false): compiled through the shared inline body
// it takes 194 VGPRs instead of 152, one wave per SIMD less.
template <typename T, bool SWEEP, typename P>
__global__ __launch_bounds__(kBlock) void k_fcg_update(const hfm_mg_task* __restrict__ tasks,
                                                       const uint16_t* __restrict__ blk_task,
                                                       const hfm_mg_scalars* __restrict__ sc,
                                                       float omega, double* dots) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (!S.active) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kCellChunk;
  const double alpha = S.alpha;
  const bool lag = S.xlag != 0;
  const double alpha_prev = S.alpha_prev;
  const int sexp = S.sexp;  // the fine residual's scale (FP16 storage)
  const SrcPlain<P> src{gin<P>(t.xin)};
  const SrcPlain<P> prev{gin<P>(t.ec)};
  HFM_GLOBAL T* rlow = gout<T>(t.rlow);
  HFM_GLOBAL T* x0 = gout<T>(t.xout);
  HFM_GLOBAL double* x64 = gout<double>(t.x64);
  HFM_GLOBAL double* r64 = gout<double>(t.r64);
  double acc = 0.0;
  if (vec_task(t)) {
    const int n0 = base + threadIdx.x * kRun;
    const bool valid = n0 < N;
    const Run r = run_at(t, n0, valid);
    RunCoef<float> q;
    q.load(t, r);
    double rv[8], xv[8];
    float pp[8];
    ld8d_ilv(r64, r.n0, rv);
    if (lag) {
      ld8d_ilv(x64, r.n0, xv);
      prev.run(r.n0, 0, 0, 0, pp);
    }
    double d[8], p[8], off[8];
    q.diag(d);
    run_offdiag<double>(t, r, q, src, p, off);
    if (valid) {
      float rl[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        rv[m] = rv[m] - alpha * (d[m] * p[m] - off[m]);
        rl[m] = rscale<T>(float(rv[m]), sexp);
        acc += rv[m] * rv[m];
      }
      st8d_ilv(r64, n0, rv);
      st8(rlow + n0, rl);
      if (lag) {
#pragma unroll
        for (int m = 0; m < 8; ++m) xv[m] = fma(alpha, p[m], fma(alpha_prev, double(pp[m]), xv[m]));
        st8d_ilv(x64, n0, xv);
      }
      if (SWEEP) {
        float o[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const float dT = round_to<T>(q.cw[m]) + round_to<T>(q.ce[m]) + round_to<T>(q.cs[m]) +
                           round_to<T>(q.cn[m]) + round_to<T>(q.cd[m]) + round_to<T>(q.cu[m]);
          o[m] = wdiv(omega, round_to<T>(rl[m]), dT);
        }
        st8_nt(x0 + n0, o);
      }
    }
  } else if (base < N) {
    auto pat = [&](int m, int, int, int) -> double { return double(src.one(m, 0, 0, 0)); };
    for (int s = 0; s < kRun; ++s) {
      const int n = base + s * kBlock + threadIdx.x;
      if (n >= N) break;
      int i, j, k;
      decompose(t, n, i, j, k);
      double diag;
      const double off = face_row<double, float>(t, n, i, j, k, pat, diag);
      const double pn = pat(n, k, j, i);
      const double q = diag * pn - off;
      if (lag) x64[n] = fma(alpha, pn, fma(alpha_prev, double(prev.one(n, 0, 0, 0)), x64[n]));
      const double rr = r64[n] - alpha * q;
      r64[n] = rr;
      const T rt = from_f<T>(rscale<T>(float(rr), sexp));
      put(rlow + n, rt);
      acc += rr * rr;
      if (SWEEP) put(x0 + n, from_f<T>(wdiv(omega, to_f(rt), diag_rounded<T>(t, n, i, j, k))));
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) dots[t.slot0 + chunk] = s;
}

// The pending half of the deferred solution update of instances that have
// stopped: x64 = fma(alpha_prev, p_prev, x64), p_prev in xin (xlag 1: the
// direction of the last update) or ec (xlag 2: the one before, when the
// instance stopped between its direction and its update).  Same tasks as
// k_fcg_update of the launch that preceded the stopping CHECK.
template <typename P>
__device__ __forceinline__ void fcg_flush_body(const hfm_mg_task* __restrict__ tasks,
                                               const uint16_t* __restrict__ blk_task,
                                               const hfm_mg_scalars* __restrict__ sc,
                                               const Blk& bk) {
  const int g = bk.g;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const hfm_mg_scalars& S = sc[t.inst];
  if (S.active || S.xlag == 0) return;
  const int chunk = xcd_chunk(g - t.blk0, t.nblk);
  const int N = t.D * t.H * t.W;
  const int base = chunk * kCellChunk;
  const double ap = S.alpha_prev;
  const SrcPlain<P> prev{gin<P>(S.xlag == 1 ? t.xin : t.ec)};
  HFM_GLOBAL double* x64 = gout<double>(t.x64);
  if (vec_task(t)) {
    const int n0 = base + bk.tid * kRun;
    if (n0 < N) {
      double xv[8];
      float pp[8];
      ld8d_ilv(x64, n0, xv);
      prev.run(n0, 0, 0, 0, pp);
#pragma unroll
      for (int m = 0; m < 8; ++m) xv[m] = fma(ap, double(pp[m]), xv[m]);
      st8d_ilv(x64, n0, xv);
    }
  } else {
    for (int s = 0; s < kRun; ++s) {
      const int n = base + s * kBlock + bk.tid;
      if (n >= N) break;
      x64[n] = fma(ap, double(prev.one(n, 0, 0, 0)), x64[n]);
    }
  }
}

template <typename P>
__global__ __launch_bounds__(kBlock) void k_fcg_flush(const hfm_mg_task* __restrict__ tasks,
                                                      const uint16_t* __restrict__ blk_task,
                                                      const hfm_mg_scalars* __restrict__ sc) {
  fcg_flush_body<P>(tasks, blk_task, sc, launch_blk());
}

// ------------------------------------------------------------------ per-instance scalars
// After a residual of squared norm rr has been written (INIT, CHECK): the z
// to come was made with 2^sexp; the next residual gets 2^(sbase - ilogb ||r||)
// (kept when rr is zero or not finite).
__device__ __forceinline__ void rescale(hfm_mg_scalars& S, double rr) {
  S.zexp = S.sexp;
  if (S.scaled && rr > 0.0 && isfinite(rr)) S.sexp = S.sbase - ilogb(sqrt(rr));
}

template <bool PB = false>
__device__ __forceinline__ void fcg_finalize_body(hfm_mg_scalars* sc, const double* dots,
                                                  int kind, const Blk& bk) {
  hfm_mg_scalars& S = sc[bk.g];
  if (!S.active) {
    // a stopped instance's pending solution update was applied by the flush
    // after its last CHECK; BETA opens the next iteration
    if (kind == HFM_FIN_BETA && bk.tid == 0) S.xlag = 0;
    if (PB) block_sum(0.0, bk.half);  // the workgroup's barriers
    return;
  }
  // INIT applies to slots staged by the host (iter == -1), the other steps to
  // initialised slots; a pool can then admit instances while others iterate.
  if ((kind == HFM_FIN_INIT) != (S.iter == -1)) {
    if (PB) block_sum(0.0, bk.half);
    return;
  }
  double acc = 0.0;
  for (int q = bk.tid; q < S.nslot; q += kBlock) acc += dots[S.slot0 + q];
  const double s = block_sum(acc, bk.half);
  if (bk.tid != 0) return;
  switch (kind) {
    case HFM_FIN_INIT:
      S.bnorm2 = s;
      S.rr = s;
      S.rel = 1.0;
      S.iter = 0;
      S.rz = 0.0;
      S.beta = 0.0;
      S.xlag = 0;
      S.alpha_prev = 0.0;
      rescale(S, s);
      if (s == 0.0) {  // x = 0 solves b = 0
        S.rel = 0.0;
        S.status = HFM_MG_CONVERGED;
        S.active = 0;
      } else if (!isfinite(s)) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
      }
      break;
    case HFM_FIN_BETA: {
      // r.z of the rescaled z (2^zexp z from 2^zexp r): divided by 2^zexp it
      // is the r.z of z scaled by 2^zexp once, so that beta, alpha and p all
      // follow the rescaled z (exact: a power of 2)
      const double rz = S.zexp ? ldexp(s, -S.zexp) : s;
      S.beta = (S.iter == 0 || S.method == HFM_METHOD_REFINE) ? 0.0 : rz / S.rz;
      S.rz = rz;
      if (!isfinite(s)) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
        if (S.xlag) S.xlag = 2;  // pending p is now the update's previous direction
      }
      break;
    }
    case HFM_FIN_ALPHA:
      S.pap = s;
      if (S.method == HFM_METHOD_REFINE ? !isfinite(s) : (!(s > 0.0) || !isfinite(s))) {
        S.status = HFM_MG_BREAKDOWN;
        S.active = 0;
        if (S.xlag) S.xlag = 2;
      } else {
        S.alpha = S.method == HFM_METHOD_REFINE ? (S.zexp ? ldexp(1.0, -S.zexp) : 1.0)
                                                : S.rz / s;
      }
      break;
    case HFM_FIN_CHECK:
      S.rr = s;
      S.rel = sqrt(s / S.bnorm2);
      S.iter += 1;
      rescale(S, s);
      // the update just run applied x's pending half and its own (xlag 1)
      // or left its own pending (xlag 0)
      S.xlag = S.xlag ? 0 : 1;
      if (S.xlag) S.alpha_prev = S.alpha;
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

__global__ __launch_bounds__(kBlock) void k_fcg_finalize(hfm_mg_scalars* sc, const double* dots,
                                                         int kind) {
  fcg_finalize_body(sc, dots, kind, launch_blk());
}

// ------------------------------------------------------------------ setup kernels
// Harmonic face coefficients of cell n (FP64 arithmetic, rounded to FP32).
__device__ __forceinline__ void faces_cell(int64_t n, const float* __restrict__ sx,
                                           const float* __restrict__ sy,
                                           const float* __restrict__ sz, float* const* o, int D,
                                           int H, int W, float inv_h2) {
  const int i = int(n % W);
  const int j = int((n / W) % H);
  const int k = int(n / (int64_t(W) * H));
  const double s2 = double(inv_h2);
  auto hm = [](double a, double b) { return 2.0 * a * b / (a + b); };
  const double ax = sx[n], ay = sy[n], az = sz[n];
  o[0][n] = float((i == 0 ? ax : hm(ax, double(sx[n - 1]))) * s2);
  o[1][n] = float((j == 0 ? ay : hm(ay, double(sy[n - W]))) * s2);
  o[2][n] = float((k == 0 ? az : hm(az, double(sz[n - int64_t(W) * H]))) * s2);
  if (i == W - 1) o[3][int64_t(k) * H + j] = float(ax * s2);
  if (j == H - 1) o[4][int64_t(k) * W + i] = float(ay * s2);
  if (k == D - 1) o[5][int64_t(j) * W + i] = float(az * s2);
}

__global__ __launch_bounds__(kBlock) void k_faces(const float* __restrict__ sx,
                                                  const float* __restrict__ sy,
                                                  const float* __restrict__ sz, float* cx,
                                                  float* cy, float* cz, float* bxh, float* byh,
                                                  float* bzh, int D, int H, int W, float inv_h2) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (n >= int64_t(D) * H * W) return;
  float* const o[6] = {cx, cy, cz, bxh, byh, bzh};
  faces_cell(n, sx, sy, sz, o, D, H, W, inv_h2);
}

// Batched: task slots xin = sx, b = sy, ec = sz, o[0..5] = output faces.
__global__ __launch_bounds__(kBlock) void k_faces_batched(const hfm_mg_task* __restrict__ tasks,
                                                          const uint16_t* __restrict__ blk_task,
                                                          float inv_h2) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const int D = t.D, H = t.H, W = t.W;
  const int64_t N = int64_t(D) * H * W, HW = int64_t(H) * W;
  const int64_t base = int64_t(g - t.blk0) * kCellChunk;
  float* const o[6] = {static_cast<float*>(t.o[0]), static_cast<float*>(t.o[1]),
                       static_cast<float*>(t.o[2]), static_cast<float*>(t.o[3]),
                       static_cast<float*>(t.o[4]), static_cast<float*>(t.o[5])};
  const float* sx = static_cast<const float*>(t.xin);
  const float* sy = static_cast<const float*>(t.b);
  const float* sz = static_cast<const float*>(t.ec);
  if (W % kRun == 0) {
    // a run of 8 cells of one row: 16-byte rows of every array, the
    // arithmetic of faces_cell per cell
    const int64_t n0 = base + threadIdx.x * kRun;
    if (n0 >= N) return;
    const int i0 = int(n0 % W);
    const int j = int((n0 / W) % H);
    const int k = int(n0 / HW);
    const double s2 = double(inv_h2);
    auto hm = [](double a, double b) { return 2.0 * a * b / (a + b); };
    float ax[kRun], ay[kRun], az[kRun], py[kRun], pz[kRun];
    ld8(gin<float>(sx) + n0, ax);
    ld8(gin<float>(sy) + n0, ay);
    ld8(gin<float>(sz) + n0, az);
    ld8(gin<float>(sy) + (j > 0 ? n0 - W : n0), py);
    ld8(gin<float>(sz) + (k > 0 ? n0 - HW : n0), pz);
    const float pw = ldg(gin<float>(sx) + (i0 > 0 ? n0 - 1 : n0));
    float ox[kRun], oy[kRun], oz[kRun];
#pragma unroll
    for (int m = 0; m < kRun; ++m) {
      const double a = ax[m];
      ox[m] = float((i0 + m == 0 ? a : hm(a, double(m ? ax[m - 1] : pw))) * s2);
      oy[m] = float((j == 0 ? double(ay[m]) : hm(double(ay[m]), double(py[m]))) * s2);
      oz[m] = float((k == 0 ? double(az[m]) : hm(double(az[m]), double(pz[m]))) * s2);
    }
    st8(gout<float>(o[0]) + n0, ox);
    st8(gout<float>(o[1]) + n0, oy);
    st8(gout<float>(o[2]) + n0, oz);
    if (i0 + kRun == W) o[3][int64_t(k) * H + j] = float(double(ax[kRun - 1]) * s2);
    if (j == H - 1) {
#pragma unroll
      for (int m = 0; m < kRun; ++m) ay[m] = float(double(ay[m]) * s2);
      st8(gout<float>(o[4]) + int64_t(k) * W + i0, ay);
    }
    if (k == D - 1) {
#pragma unroll
      for (int m = 0; m < kRun; ++m) az[m] = float(double(az[m]) * s2);
      st8(gout<float>(o[5]) + int64_t(j) * W + i0, az);
    }
    return;
  }
  for (int q = 0; q < kRun; ++q) {
    const int64_t n = base + q * kBlock + threadIdx.x;
    if (n >= N) return;
    faces_cell(n, sx, sy, sz, o, D, H, W, inv_h2);
  }
}

// Coarse faces of coarse cell n: (pair over the first listed axis, then the
// second) / 8, in FP64, rounded to FP32.
__device__ __forceinline__ void galerkin_cell(int64_t n, const float* const* f, float* const* o,
                                              int D, int H, int W) {
  const int Dc = (D + 1) / 2, Hc = (H + 1) / 2, Wc = (W + 1) / 2;
  const int I = int(n % Wc);
  const int J = int((n / Wc) % Hc);
  const int K = int(n / (int64_t(Wc) * Hc));
  auto F = [&](const float* a, int k, int j, int i) -> double {
    return (k < D && j < H && i < W) ? double(a[(int64_t(k) * H + j) * W + i]) : 0.0;
  };
  // x faces at fine i = 2I: sum k pairs, then j pairs
  o[0][n] = float(((F(f[0], 2 * K, 2 * J, 2 * I) + F(f[0], 2 * K + 1, 2 * J, 2 * I)) +
                   (F(f[0], 2 * K, 2 * J + 1, 2 * I) + F(f[0], 2 * K + 1, 2 * J + 1, 2 * I))) /
                  8.0);
  // y faces at fine j = 2J: sum k pairs, then i pairs
  o[1][n] = float(((F(f[1], 2 * K, 2 * J, 2 * I) + F(f[1], 2 * K + 1, 2 * J, 2 * I)) +
                   (F(f[1], 2 * K, 2 * J, 2 * I + 1) + F(f[1], 2 * K + 1, 2 * J, 2 * I + 1))) /
                  8.0);
  // z faces at fine k = 2K: sum j pairs, then i pairs
  o[2][n] = float(((F(f[2], 2 * K, 2 * J, 2 * I) + F(f[2], 2 * K, 2 * J + 1, 2 * I)) +
                   (F(f[2], 2 * K, 2 * J, 2 * I + 1) + F(f[2], 2 * K, 2 * J + 1, 2 * I + 1))) /
                  8.0);
  auto P = [](const float* a, int r, int c, int R, int Cn) -> double {
    return (r < R && c < Cn) ? double(a[int64_t(r) * Cn + c]) : 0.0;
  };
  if (I == Wc - 1)  // bxh (D,H): k pairs, then j pairs
    o[3][int64_t(K) * Hc + J] =
        float(((P(f[3], 2 * K, 2 * J, D, H) + P(f[3], 2 * K + 1, 2 * J, D, H)) +
               (P(f[3], 2 * K, 2 * J + 1, D, H) + P(f[3], 2 * K + 1, 2 * J + 1, D, H))) /
              8.0);
  if (J == Hc - 1)  // byh (D,W): k pairs, then i pairs
    o[4][int64_t(K) * Wc + I] =
        float(((P(f[4], 2 * K, 2 * I, D, W) + P(f[4], 2 * K + 1, 2 * I, D, W)) +
               (P(f[4], 2 * K, 2 * I + 1, D, W) + P(f[4], 2 * K + 1, 2 * I + 1, D, W))) /
              8.0);
  if (K == Dc - 1)  // bzh (H,W): j pairs, then i pairs
    o[5][int64_t(J) * Wc + I] =
        float(((P(f[5], 2 * J, 2 * I, H, W) + P(f[5], 2 * J + 1, 2 * I, H, W)) +
               (P(f[5], 2 * J, 2 * I + 1, H, W) + P(f[5], 2 * J + 1, 2 * I + 1, H, W))) /
              8.0);
}

__global__ __launch_bounds__(kBlock) void k_galerkin(const float* __restrict__ fcx,
                                                     const float* __restrict__ fcy,
                                                     const float* __restrict__ fcz,
                                                     const float* __restrict__ fbxh,
                                                     const float* __restrict__ fbyh,
                                                     const float* __restrict__ fbzh, float* cx,
                                                     float* cy, float* cz, float* bxh, float* byh,
                                                     float* bzh, int D, int H, int W) {
  const int64_t n = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (n >= int64_t((D + 1) / 2) * ((H + 1) / 2) * ((W + 1) / 2)) return;
  const float* const f[6] = {fcx, fcy, fcz, fbxh, fbyh, fbzh};
  float* const o[6] = {cx, cy, cz, bxh, byh, bzh};
  galerkin_cell(n, f, o, D, H, W);
}

// Batched: task extents D/H/W are the fine level's, c[0..5] the fine FP32
// faces, o[0..5] the coarse FP32 faces; blocks cover the coarse cells.
__global__ __launch_bounds__(kBlock) void k_galerkin_batched(const hfm_mg_task* __restrict__ tasks,
                                                             const uint16_t* __restrict__ blk_task) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const int64_t n = int64_t(g - t.blk0) * kBlock + threadIdx.x;
  if (n >= int64_t((t.D + 1) / 2) * ((t.H + 1) / 2) * ((t.W + 1) / 2)) return;
  const float* const f[6] = {static_cast<const float*>(t.c[0]), static_cast<const float*>(t.c[1]),
                             static_cast<const float*>(t.c[2]), static_cast<const float*>(t.c[3]),
                             static_cast<const float*>(t.c[4]), static_cast<const float*>(t.c[5])};
  float* const o[6] = {static_cast<float*>(t.o[0]), static_cast<float*>(t.o[1]),
                       static_cast<float*>(t.o[2]), static_cast<float*>(t.o[3]),
                       static_cast<float*>(t.o[4]), static_cast<float*>(t.o[5])};
  galerkin_cell(n, f, o, t.D, t.H, t.W);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_convert(const float* __restrict__ src, T* dst,
                                                    int64_t n) {
  const int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (q < n) dst[q] = from_f<T>(src[q]);
}

// Batched: task xin = FP32 source, xout = destination (T), D = element count.
// A thread converts 8 consecutive elements (16-byte rows when both arrays
// are 16-byte aligned, else element by element); a block covers kCellChunk.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_convert_batched(const hfm_mg_task* __restrict__ tasks,
                                                            const uint16_t* __restrict__ blk_task) {
  const int g = blockIdx.x;
  const hfm_mg_task& t = tasks[blk_task[g]];
  const int64_t n = t.D;
  const int64_t q0 = (int64_t(g - t.blk0) * kBlock + threadIdx.x) * kRun;
  if (q0 >= n) return;
  const HFM_GLOBAL float* src = gin<float>(t.xin) + q0;
  HFM_GLOBAL T* dst = gout<T>(t.xout) + q0;
  if (q0 + kRun <= n && ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) &
                         15) == 0) {
    float v[kRun];
    ld8_nt(src, v);
    st8(dst, v);
  } else {
    for (int64_t q = 0; q < kRun && q0 + q < n; ++q) put(dst + q, from_f<T>(ldg(src + q)));
  }
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


// ------------------------------------------------------------------ persistent iteration
// One launch runs a program of phases (hfm_phase): each phase is one of the
// launches above -- its task table, block map and arguments -- and its work
// items are that launch's blocks (for the finalize and the coarse cycle: its
// instances).  A workgroup of 512 threads runs an item of a 512-thread body
// (the fused pass pairs, the coarse cycle) or two items of a 256-thread body
// side by side, each half with its own block index; the bodies are the
// launches' own, so every value and every partial sum is the launch's.
//
// Order: for each instance the phases run in program order.  Item i of
// phase p for instance s waits until every item of phase prev[p][s] (the
// instance's previous phase) for s has finished: a counter per (phase,
// instance), raised by the item's workgroup after the item (release at agent
// scope), read by the waiting one (acquire); there is no grid barrier, so one
// instance's passes overlap another's.  Every workgroup takes its items in
// program order and the grid is co-resident (a cooperative launch), so the
// earliest unfinished item can always proceed.  Every wait is bounded by
// wall-clock time: on a timeout the failure word is set, every workgroup
// stops at its next item, and the host reports the program as failed.
//
// Items on XCD x: workgroup w runs on XCD w % 8 (round-robin dispatch), and
// the j-th of an XCD's workgroups takes a contiguous range of the phase's
// items g with g % 8 == x -- the items a launch would have sent to that XCD,
// so neighbouring chunks of a level stay in one L2.
constexpr int kPersistThreads = 512;
constexpr int kPairLdsBytes = sizeof(PairLds<float, false>) > sizeof(PairLds<float, true>)
                                  ? int(sizeof(PairLds<float, false>))
                                  : int(sizeof(PairLds<float, true>));

// the item's instance slot, or -1 for a padding item
__device__ __forceinline__ int item_slot(const hfm_phase& ph, int g) {
  if (ph.kind == HFM_PH_FINALIZE) return ph.slot0 + g;
  if (ph.kind == HFM_PH_COARSE_CYCLE) {
    const int first = ph.chain[2 * g], count = ph.chain[2 * g + 1];
    return count > 0 ? ph.tasks[first].inst : -1;
  }
  // a fused pair's block g runs tile xcd_chunk(g) (smooth_pair_body)
  const bool tiles = ph.kind == HFM_PH_SMOOTH_RESTRICT || ph.kind == HFM_PH_SMOOTH2;
  const int q = ph.blk_task[tiles ? xcd_chunk(g, ph.nitems) : g];
  return q == 0xFFFF ? -1 : ph.tasks[q].inst;
}

// thread 0 of a (half-)block: wait for the item's previous phase; false on failure
__device__ __forceinline__ bool persist_wait(const hfm_persist_ctl& ctl, int p, int slot,
                                             uint64_t ticks) {
  if (slot < 0) return true;
  const int q = ctl.prev[p * ctl.nslot + slot];
  if (q < 0) return true;
  const uint32_t target = uint32_t(ctl.need[q * ctl.nslot + slot]);
  const uint32_t* c = ctl.cnt + q * ctl.nslot + slot;
  // poll with relaxed loads; one acquire fence (a cache invalidation) once met
  if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    const uint64_t t0 = wall_clock64();
    for (;;) {
      __builtin_amdgcn_s_sleep(2);
      if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
      if (__hip_atomic_load(ctl.fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
        return false;
      if (wall_clock64() - t0 > ticks) {
        atomicExch(ctl.fail, 1u);
        return false;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

__device__ __forceinline__ void persist_signal(const hfm_persist_ctl& ctl, int p, int slot,
                                               int count) {
  if (slot < 0 || count <= 0) return;
  // the release writes back this XCD's L2: the items' stores (ordered before
  // it by the workgroup barrier) reach the other XCDs
  __hip_atomic_fetch_add(ctl.cnt + p * ctl.nslot + slot, uint32_t(count), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// the 256-thread bodies; `g < 0`: this half has no item (it takes part in
// the body's barriers only)
template <typename T, typename P>
__device__ __forceinline__ void persist_half(const hfm_phase& ph, int g, int half,
                                             hfm_mg_scalars* sc) {
  const Blk bk{g, ph.nitems, int(threadIdx.x) & (kBlock - 1), half};
  const bool none = g < 0;
  switch (ph.kind) {
    case HFM_PH_SWEEP:
      if (ph.dot) {
        if (none) { block_sum(0.0, half); break; }
        switch (ph.mode) {
          case HFM_SWEEP_ZERO:
            sweep_body<T, HFM_SWEEP_ZERO, true, P, kBlock, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha, ph.dots, bk); break;
          case HFM_SWEEP_PLAIN:
            sweep_body<T, HFM_SWEEP_PLAIN, true, P, kBlock, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha, ph.dots, bk); break;
          default:
            sweep_body<T, HFM_SWEEP_CORRECT, true, P, kBlock, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha, ph.dots, bk); break;
        }
      } else if (!none) {
        switch (ph.mode) {
          case HFM_SWEEP_ZERO:
            sweep_body<T, HFM_SWEEP_ZERO, false, T, kBlock, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha, ph.dots, bk); break;
          case HFM_SWEEP_PLAIN:
            sweep_body<T, HFM_SWEEP_PLAIN, false, T, kBlock, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha, ph.dots, bk); break;
          default:
            sweep_body<T, HFM_SWEEP_CORRECT, false, T, kBlock, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha, ph.dots, bk); break;
        }
      }
      break;
    case HFM_PH_RESTRICT:
      if (!none) resid_restrict_body<T, kBlock>(ph.tasks, ph.blk_task, sc, bk);
      break;
    case HFM_PH_DIR:
      if (none) block_sum(0.0, half);
      else fcg_dir_body<P, true>(ph.tasks, ph.blk_task, sc, ph.dots, bk);
      break;
    case HFM_PH_UPDATE:
      if (none) block_sum(0.0, half);
      else fcg_update_body<T, true, P, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.dots, bk);
      break;
    case HFM_PH_FLUSH:
      if (!none) fcg_flush_body<P>(ph.tasks, ph.blk_task, sc, bk);
      break;
    case HFM_PH_FINALIZE:
      if (none) block_sum(0.0, half);
      else fcg_finalize_body<true>(sc + ph.slot0, ph.dots, ph.mode, bk);
      break;
    default:
      break;
  }
}

template <typename T, typename P>
__global__ __launch_bounds__(kPersistThreads) void k_persistent(const hfm_phase* __restrict__ prog,
                                                                int nphase, hfm_persist_ctl ctl,
                                                                hfm_mg_scalars* sc,
                                                                uint64_t ticks) {
  __shared__ __align__(16) unsigned char lds[kPairLdsBytes];
  __shared__ int s_go[2];
  const int G = int(gridDim.x);
  const int w = int(blockIdx.x);
  const int x = w & 7, j = w >> 3, per = G >> 3;
  const int tid = int(threadIdx.x), half = tid >> 8;
  const bool lead = (tid & (kBlock - 1)) == 0;  // thread 0 of a half
  for (int p = 0; p < nphase; ++p) {
    const hfm_phase ph = prog[p];
    const int n8 = ph.nitems > x ? (ph.nitems - x + 7) / 8 : 0;  // items with g % 8 == x
    const int e0 = int(int64_t(n8) * j / per), e1 = int(int64_t(n8) * (j + 1) / per);
    const bool whole = ph.kind == HFM_PH_SMOOTH_RESTRICT || ph.kind == HFM_PH_SMOOTH2 ||
                       ph.kind == HFM_PH_COARSE_CYCLE;
    // Per half (the same values in every thread): the slot whose finished
    // items this workgroup has not signalled yet and their count, and the
    // slot whose dependency is known to be met.  Consecutive items mostly
    // belong to one instance: they then take no wait and no signal of their
    // own (the count is raised once, when the slot changes or the phase ends).
    int pend[2] = {-1, -1}, npend[2] = {0, 0}, okslot[2] = {-1, -1};
    for (int e = e0; e < e1; e += whole ? 1 : 2) {
      const int g0 = x + 8 * e;
      const int g1 = whole ? -1 : (e + 1 < e1 ? g0 + 8 : -1);
      const int sl[2] = {item_slot(ph, g0), g1 >= 0 ? item_slot(ph, g1) : -1};
      const int nh = whole ? 1 : 2;
      bool fast = true;
      for (int h = 0; h < nh; ++h)
        fast = fast && (sl[h] < 0 || (sl[h] == pend[h] && sl[h] == okslot[h]));
      if (!fast || whole) {
        // the previous items are done by every thread before their signal and
        // before this item touches the LDS
        __syncthreads();
        if (lead && (whole ? tid == 0 : true) && half < nh) {
          bool ok = true;
          if (sl[half] != pend[half] && npend[half] > 0) {
            persist_signal(ctl, p, pend[half], npend[half]);
          }
          if (sl[half] >= 0 && sl[half] != okslot[half])
            ok = persist_wait(ctl, p, sl[half], ticks);
          s_go[half] = ok ? 1 : 0;
        }
        __syncthreads();
        const bool go = s_go[0] != 0 && (nh == 1 || s_go[1] != 0);
        if (!go) return;  // a wait failed: every workgroup leaves (uniformly here)
        for (int h = 0; h < nh; ++h) {
          if (sl[h] != pend[h]) {
            pend[h] = sl[h];
            npend[h] = 0;
          }
          okslot[h] = sl[h];
        }
      }
      if (whole) {
        const Blk bk{g0, ph.nitems, tid, 0};
        if (ph.kind == HFM_PH_COARSE_CYCLE) {
          coarse_cycle_body<T, kPersistThreads>(ph.tasks, ph.chain, sc, ph.nu1, ph.nu2,
                                                ph.omega, ph.alpha, bk);
        } else if (ph.kind == HFM_PH_SMOOTH_RESTRICT) {
          smooth_pair_body<T, false, T, false>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha,
                                               ph.dots,
                                               *reinterpret_cast<PairLds<T, false>*>(lds), bk);
        } else if (ph.dot) {
          smooth_pair_body<T, true, P, true>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha,
                                             ph.dots,
                                             *reinterpret_cast<PairLds<T, true>*>(lds), bk);
        } else {
          smooth_pair_body<T, true, T, false>(ph.tasks, ph.blk_task, sc, ph.omega, ph.alpha,
                                              ph.dots,
                                              *reinterpret_cast<PairLds<T, true>*>(lds), bk);
        }
      } else {
        persist_half<T, P>(ph, half == 0 ? g0 : g1, half, sc);
      }
      for (int h = 0; h < nh; ++h)
        if (sl[h] >= 0) npend[h] += 1;
    }
    // the phase's last signals, once every thread is done
    __syncthreads();
    if (lead && npend[half] > 0 && (whole ? tid == 0 : true))
      persist_signal(ctl, p, pend[half], npend[half]);
  }
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
  } else if (store == HFM_STORE_F16) {
    f(f16{});
  } else {
    set_error("%s: unknown storage type %d", what, store);
    return HFM_EINVAL;
  }
  return launch_status(what);
}

// Two storage types (inner, outer): f(T{}, O{}).  The outer type is FP32 or
// the inner 16-bit type's kind (BF16 with FP32 or BF16, FP16 with FP16).
template <typename F>
int launch_store2(int store, int outer, const char* what, F&& f) {
  const bool ok = outer == HFM_STORE_F32 ||
                  (outer == HFM_STORE_BF16 && store != HFM_STORE_F16) ||
                  (outer == HFM_STORE_F16 && store == HFM_STORE_F16);
  if (!ok) {
    if (outer < HFM_STORE_F32 || outer > HFM_STORE_F16)
      set_error("%s: unknown storage type %d", what, outer);
    else
      set_error("%s: storage %d with outer storage %d is not built", what, store, outer);
    return HFM_EINVAL;
  }
  return launch_store(store, what, [&](auto tag) {
    using T = decltype(tag);
    if (outer == HFM_STORE_F32) {
      f(tag, float{});
    } else if constexpr (is_f16<T>::value) {
      f(tag, f16{});
    } else {
      f(tag, bf16{});
    }
  });
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

namespace {
// Threads per block of the tunable cell kernels: 128, 256 or 512.
bool block_ok(int block) { return block == 128 || block == 256 || block == 512; }

template <int B, typename T, typename O>
void launch_sweep(int mode, bool dot, dim3 grid, hipStream_t s, const hfm_mg_task* tasks,
                  const uint16_t* blk_task, const hfm_mg_scalars* sc, float omega, float alpha,
                  double* dots) {
  const dim3 block{unsigned(B)};
#define HFM_SWEEP_CASE(M)                                                                       \
  if (dot)                                                                                      \
    hipLaunchKernelGGL((k_sweep<T, M, true, O, kBlock>), grid, dim3(kBlock), 0, s, tasks,       \
                       blk_task, sc, omega, alpha, dots);                                       \
  else                                                                                          \
    hipLaunchKernelGGL((k_sweep<T, M, false, O, B>), grid, block, 0, s, tasks, blk_task, sc,     \
                       omega, alpha, dots);
  if (mode == HFM_SWEEP_ZERO) {
    HFM_SWEEP_CASE(HFM_SWEEP_ZERO)
  } else if (mode == HFM_SWEEP_PLAIN) {
    HFM_SWEEP_CASE(HFM_SWEEP_PLAIN)
  } else {
    HFM_SWEEP_CASE(HFM_SWEEP_CORRECT)
  }
#undef HFM_SWEEP_CASE
}
}  // namespace

extern "C" int hfm_mg_sweep(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                            const hfm_mg_scalars* sc, int store, int out_store, int mode,
                            float omega, float alpha, int block, double* dots, void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_mg_sweep")) return e;
  if (mode < HFM_SWEEP_ZERO || mode > HFM_SWEEP_CORRECT) {
    set_error("hfm_mg_sweep: bad mode %d", mode);
    return HFM_EINVAL;
  }
  if (!block_ok(block) || (dots && block != kBlock)) {
    set_error("hfm_mg_sweep: block of %d threads not supported%s", block,
              dots ? " with dot partials (256 only)" : "");
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  hipStream_t s = as_stream(stream);
  return launch_store2(store, out_store, "hfm_mg_sweep", [&](auto tag, auto otag) {
    using T = decltype(tag);
    using O = decltype(otag);
    const dim3 grid{unsigned(nblocks)};
    const bool dot = dots != nullptr;
    if (block == 128)
      launch_sweep<128, T, O>(mode, dot, grid, s, tasks, blk_task, sc, omega, alpha, dots);
    else if (block == 256)
      launch_sweep<256, T, O>(mode, dot, grid, s, tasks, blk_task, sc, omega, alpha, dots);
    else
      launch_sweep<512, T, O>(mode, dot, grid, s, tasks, blk_task, sc, omega, alpha, dots);
  });
}

extern "C" int hfm_mg_resid_restrict(const hfm_mg_task* tasks, const uint16_t* blk_task,
                                     int64_t nblocks, const hfm_mg_scalars* sc, int store,
                                     int block, void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_mg_resid_restrict")) return e;
  if (!block_ok(block)) {
    set_error("hfm_mg_resid_restrict: block of %d threads not supported", block);
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_resid_restrict", [&](auto tag) {
    using T = decltype(tag);
    const dim3 grid{unsigned(nblocks)};
    hipStream_t s = as_stream(stream);
    if (block == 128)
      hipLaunchKernelGGL((k_resid_restrict<T, 128>), grid, dim3(128), 0, s, tasks, blk_task, sc);
    else if (block == 256)
      hipLaunchKernelGGL((k_resid_restrict<T, 256>), grid, dim3(256), 0, s, tasks, blk_task, sc);
    else
      hipLaunchKernelGGL((k_resid_restrict<T, 512>), grid, dim3(512), 0, s, tasks, blk_task, sc);
  });
}

namespace {
// Planes per tile of the fused pass pairs: a tile computes pass 1 on two
// halo planes, so longer walks waste less; 32 keeps a few tiles per task.
constexpr int kTilePlanes = 32;
}  // namespace

extern "C" int hfm_mg_tile(int64_t D, int64_t H, int64_t W, int32_t max_planes,
                           int32_t* tile_rows, int32_t* tile_planes, int64_t* ntiles) {
  if (!tile_rows || !tile_planes || !ntiles || D < 1 || H < 1 || W < 1 ||
      D * H * W > INT32_MAX || max_planes == 1 || max_planes < 0 || max_planes > 4096) {
    set_error("hfm_mg_tile: bad arguments");
    return HFM_EINVAL;
  }
  const int64_t zmax = max_planes ? max_planes : kTilePlanes;
  if (W % kRun) {
    set_error("hfm_mg_tile: W = %lld is not a multiple of %d", (long long)W, kRun);
    return HFM_EINVAL;
  }
  const int64_t Q = W / kPairRun;  // runs per row = threads per row
  int64_t R, ny;
  if (H * Q <= kPairBlock) {       // the whole plane fits: no halo rows
    R = H;
    ny = 1;
  } else {
    const int64_t rmax = (kPairBlock / Q - 2) & ~int64_t(1);  // tile rows + 2 halo rows fit
    if (rmax < 2) {
      set_error("hfm_mg_tile: W = %lld is too wide for a tile", (long long)W);
      return HFM_EINVAL;
    }
    ny = (H + rmax - 1) / rmax;
    R = 2 * ((H + 2 * ny - 1) / (2 * ny));  // balanced, even (restriction pairs)
  }
  const int64_t nz = (D + zmax - 1) / zmax;
  const int64_t Z = 2 * ((D + 2 * nz - 1) / (2 * nz));
  *tile_rows = int32_t(R);
  *tile_planes = int32_t(Z);
  *ntiles = ny * ((D + Z - 1) / Z);
  return HFM_OK;
}

extern "C" int hfm_mg_smooth_restrict(const hfm_mg_task* tasks, const uint16_t* blk_task,
                                      int64_t nblocks, const hfm_mg_scalars* sc, int store,
                                      float omega, void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_mg_smooth_restrict")) return e;
  if (nblocks % 8) {  // the tile -> XCD map is a permutation of 8-block groups
    set_error("hfm_mg_smooth_restrict: %lld blocks is not a multiple of 8", (long long)nblocks);
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_smooth_restrict", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL((k_smooth_pair<T, false, T, false>), dim3(unsigned(nblocks)),
                       dim3(kPairBlock), 0, as_stream(stream), tasks, blk_task, sc, omega, 0.0f,
                       nullptr);
  });
}

extern "C" int hfm_mg_smooth2(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                              const hfm_mg_scalars* sc, int store, int out_store, float omega,
                              float alpha, double* dots, void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_mg_smooth2")) return e;
  if (nblocks % 8) {  // the tile -> XCD map is a permutation of 8-block groups
    set_error("hfm_mg_smooth2: %lld blocks is not a multiple of 8", (long long)nblocks);
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  return launch_store2(store, out_store, "hfm_mg_smooth2", [&](auto tag, auto otag) {
    using T = decltype(tag);
    using O = decltype(otag);
    const dim3 grid{unsigned(nblocks)};
    hipStream_t s = as_stream(stream);
    if (dots)
      hipLaunchKernelGGL((k_smooth_pair<T, true, O, true>), grid, dim3(kPairBlock), 0, s, tasks,
                         blk_task, sc, omega, alpha, dots);
    else
      hipLaunchKernelGGL((k_smooth_pair<T, true, O, false>), grid, dim3(kPairBlock), 0, s,
                         tasks, blk_task, sc, omega, alpha, dots);
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

extern "C" int hfm_mg_coarse_cycle(const hfm_mg_task* recs, const int32_t* chain,
                                   int64_t ninst, const hfm_mg_scalars* sc, int store, int nu1,
                                   int nu2, float omega, float alpha, void* stream) {
  if (!recs || !chain || !sc || ninst < 0 || ninst > INT32_MAX || nu1 < 1 || nu2 < 1) {
    set_error("hfm_mg_coarse_cycle: bad arguments");
    return HFM_EINVAL;
  }
  if (ninst == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_coarse_cycle", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_coarse_cycle<T>, dim3(unsigned(ninst)), dim3(kCycleThreads), 0,
                       as_stream(stream), recs, chain, sc, nu1, nu2, omega, alpha);
  });
}

namespace {
int check_fcg(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
              const hfm_mg_scalars* sc, double* dots, const char* what) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, what)) return e;
  if (!dots) {
    set_error("%s: null dot slots", what);
    return HFM_EINVAL;
  }
  return HFM_OK;
}
}  // namespace

extern "C" int hfm_fcg_dir(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                           const hfm_mg_scalars* sc, int store, double* dots, void* stream) {
  if (int e = check_fcg(tasks, blk_task, nblocks, sc, dots, "hfm_fcg_dir")) return e;
  if (nblocks == 0) return HFM_OK;
  return launch_store(store, "hfm_fcg_dir", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_fcg_dir<T>, dim3(unsigned(nblocks)), dim3(kBlock), 0, as_stream(stream),
                       tasks, blk_task, sc, dots);
  });
}

#define HFM_FCG_INIT_ENTRY(NAME, SWEEP)                                                         \
  extern "C" int NAME(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,      \
                      const hfm_mg_scalars* sc, int store, HFM_OMEGA_ARG double* dots,          \
                      void* stream) {                                                            \
    if (int e = check_fcg(tasks, blk_task, nblocks, sc, dots, #NAME)) return e;                 \
    if (nblocks == 0) return HFM_OK;                                                            \
    return launch_store(store, #NAME, [&](auto tag) {                                           \
      using T = decltype(tag);                                                                  \
      hipLaunchKernelGGL((k_fcg_init<T, SWEEP>), dim3(unsigned(nblocks)), dim3(kBlock), 0,      \
                         as_stream(stream), tasks, blk_task, sc, HFM_OMEGA_VAL, dots);          \
    });                                                                                         \
  }
#define HFM_FCG_UPDATE_ENTRY(NAME, SWEEP)                                                       \
  extern "C" int NAME(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,      \
                      const hfm_mg_scalars* sc, int store, int pstore, HFM_OMEGA_ARG            \
                      double* dots, void* stream) {                                             \
    if (int e = check_fcg(tasks, blk_task, nblocks, sc, dots, #NAME)) return e;                 \
    if (nblocks == 0) return HFM_OK;                                                            \
    return launch_store2(store, pstore, #NAME, [&](auto tag, auto ptag) {                       \
      using T = decltype(tag);                                                                  \
      using P = decltype(ptag);                                                                 \
      hipLaunchKernelGGL((k_fcg_update<T, SWEEP, P>), dim3(unsigned(nblocks)), dim3(kBlock), 0, \
                         as_stream(stream), tasks, blk_task, sc, HFM_OMEGA_VAL, dots);          \
    });                                                                                         \
  }

extern "C" int hfm_fcg_flush(const hfm_mg_task* tasks, const uint16_t* blk_task, int64_t nblocks,
                             const hfm_mg_scalars* sc, int pstore, void* stream) {
  if (int e = check_table(tasks, blk_task, nblocks, sc, "hfm_fcg_flush")) return e;
  if (nblocks == 0) return HFM_OK;
  return launch_store(pstore, "hfm_fcg_flush", [&](auto tag) {
    using P = decltype(tag);
    hipLaunchKernelGGL((k_fcg_flush<P>), dim3(unsigned(nblocks)), dim3(kBlock), 0,
                       as_stream(stream), tasks, blk_task, sc);
  });
}

#define HFM_OMEGA_ARG
#define HFM_OMEGA_VAL 1.0f
HFM_FCG_INIT_ENTRY(hfm_fcg_init, false)
HFM_FCG_UPDATE_ENTRY(hfm_fcg_update, false)
#undef HFM_OMEGA_ARG
#undef HFM_OMEGA_VAL
#define HFM_OMEGA_ARG float omega,
#define HFM_OMEGA_VAL omega
HFM_FCG_INIT_ENTRY(hfm_fcg_init_sweep, true)
HFM_FCG_UPDATE_ENTRY(hfm_fcg_update_sweep, true)
#undef HFM_OMEGA_ARG
#undef HFM_OMEGA_VAL
#undef HFM_FCG_INIT_ENTRY
#undef HFM_FCG_UPDATE_ENTRY

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

extern "C" int hfm_mg_faces_batched(const hfm_mg_task* tasks, const uint16_t* blk_task,
                                    int64_t nblocks, float inv_h2, void* stream) {
  if (!tasks || (nblocks > 0 && !blk_task) || nblocks < 0 || nblocks > INT32_MAX) {
    set_error("hfm_mg_faces_batched: bad arguments");
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  hipLaunchKernelGGL(k_faces_batched, dim3(unsigned(nblocks)), dim3(kBlock), 0,
                     as_stream(stream), tasks, blk_task, inv_h2);
  return launch_status("hfm_mg_faces_batched");
}

extern "C" int hfm_mg_galerkin_batched(const hfm_mg_task* tasks, const uint16_t* blk_task,
                                       int64_t nblocks, void* stream) {
  if (!tasks || (nblocks > 0 && !blk_task) || nblocks < 0 || nblocks > INT32_MAX) {
    set_error("hfm_mg_galerkin_batched: bad arguments");
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  hipLaunchKernelGGL(k_galerkin_batched, dim3(unsigned(nblocks)), dim3(kBlock), 0,
                     as_stream(stream), tasks, blk_task);
  return launch_status("hfm_mg_galerkin_batched");
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

extern "C" int hfm_mg_convert_batched(const hfm_mg_task* tasks, const uint16_t* blk_task,
                                      int64_t nblocks, int store, void* stream) {
  if (!tasks || (nblocks > 0 && !blk_task) || nblocks < 0 || nblocks > INT32_MAX) {
    set_error("hfm_mg_convert_batched: bad arguments");
    return HFM_EINVAL;
  }
  if (nblocks == 0) return HFM_OK;
  return launch_store(store, "hfm_mg_convert_batched", [&](auto tag) {
    using T = decltype(tag);
    hipLaunchKernelGGL(k_convert_batched<T>, dim3(unsigned(nblocks)), dim3(kBlock), 0,
                       as_stream(stream), tasks, blk_task);
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

namespace {
template <typename T, typename P>
const void* persistent_kernel() {
  return reinterpret_cast<const void*>(&k_persistent<T, P>);
}

int persistent_fn(int store, int pstore, const void** fn) {
  if (store == HFM_STORE_F32 && pstore == HFM_STORE_F32) *fn = persistent_kernel<float, float>();
  else if (store == HFM_STORE_BF16 && pstore == HFM_STORE_BF16) *fn = persistent_kernel<bf16, bf16>();
  else if (store == HFM_STORE_BF16 && pstore == HFM_STORE_F32) *fn = persistent_kernel<bf16, float>();
  else if (store == HFM_STORE_F32 && pstore == HFM_STORE_BF16) *fn = persistent_kernel<float, bf16>();
  else if (store == HFM_STORE_F16 && pstore == HFM_STORE_F16) *fn = persistent_kernel<f16, f16>();
  else if (store == HFM_STORE_F16 && pstore == HFM_STORE_F32) *fn = persistent_kernel<f16, float>();
  else return HFM_EINVAL;
  return HFM_OK;
}
}  // namespace

extern "C" int hfm_mg_persistent_workgroups(int store, int pstore, int32_t* n) {
  const void* fn = nullptr;
  if (!n || persistent_fn(store, pstore, &fn) != HFM_OK) {
    set_error("hfm_mg_persistent_workgroups: bad arguments");
    return HFM_EINVAL;
  }
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, kPersistThreads, 0) != hipSuccess ||
      per < 1) {
    set_error("hfm_mg_persistent_workgroups: occupancy query failed");
    return HFM_EDEVICE;
  }
  *n = int32_t((int64_t(cus) * per) / 8 * 8);
  return HFM_OK;
}

extern "C" int hfm_mg_persistent(const hfm_phase* prog, int32_t nphase, const hfm_persist_ctl* ctl,
                                 hfm_mg_scalars* sc, int store, int pstore, int32_t workgroups,
                                 double timeout_s, void* stream) {
  const void* fn = nullptr;
  if (!prog || !ctl || !sc || nphase < 0 || !ctl->cnt || !ctl->need || !ctl->prev ||
      !ctl->fail || ctl->nslot < 1 || workgroups < 8 || workgroups % 8 ||
      !(timeout_s > 0.0) || persistent_fn(store, pstore, &fn) != HFM_OK) {
    set_error("hfm_mg_persistent: bad arguments");
    return HFM_EINVAL;
  }
  if (nphase == 0) return HFM_OK;
  int dev = 0, rate_khz = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
      rate_khz <= 0) {
    set_error("hfm_mg_persistent: cannot read the device wall-clock rate");
    return HFM_EDEVICE;
  }
  const hfm_phase* pr = prog;
  int np = nphase;
  hfm_persist_ctl c = *ctl;
  hfm_mg_scalars* s = sc;
  uint64_t ticks = uint64_t(double(rate_khz) * 1000.0 * timeout_s);
  void* args[] = {(void*)&pr, (void*)&np, (void*)&c, (void*)&s, (void*)&ticks};
  const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(unsigned(workgroups)),
                                                  dim3(kPersistThreads), args, 0,
                                                  as_stream(stream));
  if (e != hipSuccess) {
    set_error("hfm_mg_persistent: cooperative launch of %d workgroups failed: %s", workgroups,
              hipGetErrorString(e));
    return HFM_EDEVICE;
  }
  return launch_status("hfm_mg_persistent");
}

extern "C" int hfm_mg_persist_struct_sizes(int64_t* phase_size, int64_t* ctl_size) {
  if (!phase_size || !ctl_size) {
    set_error("hfm_mg_persist_struct_sizes: null output");
    return HFM_EINVAL;
  }
  *phase_size = int64_t(sizeof(hfm_phase));
  *ctl_size = int64_t(sizeof(hfm_persist_ctl));
  return HFM_OK;
}
