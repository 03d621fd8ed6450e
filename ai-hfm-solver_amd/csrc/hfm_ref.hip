// Reference-mode kernels for gfx950: the reference's fixed-weight Keras
// layers (3x3x3 SAME convolutions, the six-direction variable-coefficient
// operator, 2^d mean restriction, nearest upsampling) and their fused
// smoother/residual forms, plus the boundary-plane fills.  FP32 throughout.
//
// 64 x 4 thread blocks, so a wavefront reads 64 consecutive floats of a row.
// The cell kernels take one output cell per thread and one z plane per grid
// row; the 3x3x3 stencils on large grids are column walkers (one thread walks
// a z column of 18 planes with a register window, see below).
#include <stdarg.h>
#include <stdio.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "hfm_common.h"

namespace hfm {

// ------------------------------------------------------------------ errors
namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int launch_status(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: launch failed: %s", what, hipGetErrorString(e));
    return HFM_EDEVICE;
  }
  return HFM_OK;
}

// ------------------------------------------------------------------ helpers
struct Weights {
  float w[27];
};

constexpr int kBX = 64, kBY = 4;

struct Grid3 {
  int D, H, W;
  __device__ __forceinline__ int64_t idx(int k, int j, int i) const {
    return (int64_t(k) * H + j) * W + i;
  }
  // &x(k, j, i) as plane k's base (formed by the scalar unit where k is
  // wave-uniform, as in the column walkers) plus a 32-bit byte offset in the
  // plane: a plane holds fewer than 2^30 cells (dims_ok)
  template <typename T>
  __device__ __forceinline__ T* cell(T* x, int k, int j, int i) const {
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    T* plane = x + int64_t(k) * (int64_t(H) * W);
    return reinterpret_cast<T*>(reinterpret_cast<B*>(plane) +
                                uint32_t(j * W + i) * uint32_t(sizeof(T)));
  }
  // cell() for a wave-uniform plane k (the column walkers' planes): the
  // in-plane offset is tied to k by an empty asm statement, so it is formed
  // next to the access and the access takes the plane's base in scalar
  // registers plus the offset (global_load ... saddr), instead of a
  // loop-invariant 64-bit offset hoisted out of the walk and added to the
  // base with a VALU instruction per access.  That frees the offsets' vector
  // registers but holds a base per field in scalar registers, which pays
  // where it lifts occupancy (k_pg_coef 3 -> 4 waves per SIMD, k_grad3) and
  // costs elsewhere.  k must be the same in every lane (the asm takes it in
  // a scalar register).
  template <typename T>
  __device__ __forceinline__ T* cellu(T* x, int k, int j, int i) const {
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    T* plane = x + int64_t(k) * (int64_t(H) * W);
    uint32_t o = uint32_t(j * W + i) * uint32_t(sizeof(T));
    asm("" : "+v"(o) : "s"(k));
    return reinterpret_cast<T*>(reinterpret_cast<B*>(plane) + o);
  }
  // without short-circuit evaluation (no branches between a cell's loads)
  __device__ __forceinline__ bool in(int k, int j, int i) const {
    return (unsigned(k) < unsigned(D)) & (unsigned(j) < unsigned(H)) & (unsigned(i) < unsigned(W));
  }
};

// Tap sets of the 3x3x3 stencils (row-major t = 9a+3b+c): the 7-point cross
// or all 27.  A kernel is instantiated per set and the host picks the cross
// when every non-zero weight lies on it (the other taps have zero weight, so
// skipping them is exact).
constexpr uint32_t kCross = (1u << 4) | (1u << 10) | (1u << 12) | (1u << 13) | (1u << 14) |
                            (1u << 16) | (1u << 22);
// Also the 18-tap sets of the first-derivative stencils (the advection and
// divergence layers: the taps of the centre x column, y row or z plane are
// zero), TAPS = kTapsX / kTapsY / kTapsZ.
constexpr int kTapsX = 180, kTapsY = 181, kTapsZ = 182;
constexpr uint32_t tap_mask(int taps) {
  uint32_t m = 0;
  for (int t = 0; t < 27; ++t) {
    const bool on = taps == 27 ? true
                    : taps == 7 ? ((kCross >> t) & 1u) != 0
                    : taps == kTapsX ? t % 3 != 1
                    : taps == kTapsY ? (t / 3) % 3 != 1
                                     : t / 9 != 1;
    if (on) m |= 1u << t;
  }
  return m;
}
template <int TAPS>
__device__ constexpr bool tap_on(int t) {
  return (tap_mask(TAPS) >> t) & 1u;
}
// the smallest tap set holding every non-zero weight
inline int taps_of(uint32_t mask) {
  for (int taps : {7, kTapsX, kTapsY, kTapsZ})
    if (!(mask & ~tap_mask(taps))) return taps;
  return 27;
}

// The taps of the 3x3x3 neighbourhood of (k, j, i) with zero ghosts.  All are
// loaded from in-bounds addresses (a tap outside the grid re-reads the centre
// and is zeroed) without branches, so a cell's loads are in flight together.
template <int TAPS, typename XF>
__device__ __forceinline__ void load_nbhd(const Grid3& g, int k, int j, int i, XF xat,
                                          float v[27]) {
#pragma unroll
  for (int t = 0; t < 27; ++t) {
    if (!tap_on<TAPS>(t)) {
      v[t] = 0.0f;
      continue;
    }
    const int kk = k + t / 9 - 1, jj = j + (t / 3) % 3 - 1, ii = i + t % 3 - 1;
    const bool in = g.in(kk, jj, ii);
    const float x = xat(in ? kk : k, in ? jj : j, in ? ii : i);
    v[t] = in ? x : 0.0f;
  }
}

// sum_t K[t] v[t] over the tap set, in tap order
template <int TAPS>
__device__ __forceinline__ float wsum(const Weights& K, const float v[27]) {
  float acc = 0.0f;
#pragma unroll
  for (int t = 0; t < 27; ++t)
    if (tap_on<TAPS>(t)) acc += K.w[t] * v[t];
  return acc;
}

// The weights of NK stencils in LDS, for a walker whose weights and field
// pointers exceed the scalar registers (the compiler would keep part of them
// in vector-register lanes and read each back with a v_readlane, a VALU
// instruction, at every use).  Staged once per workgroup (every wave takes
// part in the barrier, before any returns), read per step as broadcasts (up
// to four contiguous weights per ds_read).  The index offset zero4() is
// redefined at every step by an empty asm statement, so the reads stay at
// their use instead of being hoisted out of the walk into NK * 27 live
// registers.
template <int NK>
struct WeightLds {
  float w[NK][28];  // rows of 28 floats: 16-byte aligned
  __device__ __forceinline__ void stage(const Weights* k) {
    const int t = int(threadIdx.x) + kBX * int(threadIdx.y);
    for (int q = t; q < NK * 27; q += kBX * kBY) w[q / 27][q % 27] = k[q / 27].w[q % 27];
    __syncthreads();
  }
};
__device__ __forceinline__ int zero4() {
  int z = 0;
  asm volatile("" : "+v"(z));
  return z << 2;
}
// sum_t K[t] v[t] over the tap set in tap order (as wsum), with K from LDS,
// for two tap vectors (two output rows): each weight read once for both
template <int TAPS>
__device__ __forceinline__ void wsum2_lds(const float* w, const float v0[27], const float v1[27],
                                          float& a0, float& a1) {
  a0 = 0.0f;
  a1 = 0.0f;
#pragma unroll
  for (int t = 0; t < 27; ++t)
    if (tap_on<TAPS>(t)) {
      const float wt = w[t];
      a0 += wt * v0[t];
      a1 += wt * v1[t];
    }
}

// sum_t K[t] x(k+a-1, j+b-1, i+c-1) over the tap set (zero ghosts)
template <int TAPS, typename XF>
__device__ __forceinline__ float conv_at(const Weights& K, const Grid3& g, int k, int j, int i,
                                         XF xat) {
  float v[27];
  load_nbhd<TAPS>(g, k, j, i, xat, v);
  return wsum<TAPS>(K, v);
}

// six-direction operator sum_dir m_dir (x - x_nb), order x-, x+, y-, y+, z-, z+
struct Var6 {
  const float* m[6];
  const float* tot;
};

template <typename XF>
__device__ __forceinline__ float var6_at(const Var6& V, const Grid3& g, int k, int j, int i,
                                         XF xat) {
  const int64_t n = g.idx(k, j, i);
  const float xc = xat(k, j, i);
  // a neighbour outside the grid re-reads the centre (in bounds) and is zeroed
  auto nb = [&](int kk, int jj, int ii) {
    const bool in = g.in(kk, jj, ii);
    const float x = xat(in ? kk : k, in ? jj : j, in ? ii : i);
    return in ? x : 0.0f;
  };
  float acc = 0.0f;
  acc += (xc - nb(k, j, i - 1)) * V.m[0][n];
  acc += (xc - nb(k, j, i + 1)) * V.m[1][n];
  acc += (xc - nb(k, j - 1, i)) * V.m[2][n];
  acc += (xc - nb(k, j + 1, i)) * V.m[3][n];
  acc += (xc - nb(k - 1, j, i)) * V.m[4][n];
  acc += (xc - nb(k + 1, j, i)) * V.m[5][n];
  return acc;
}

__device__ __forceinline__ bool cell_of_thread(const Grid3& g, int& k, int& j, int& i) {
  i = blockIdx.x * kBX + threadIdx.x;
  j = blockIdx.y * kBY + threadIdx.y;
  k = blockIdx.z;
  return i < g.W && j < g.H;
}

// ------------------------------------------------------------------ column walkers
// The 3x3x3 kernels on large grids walk z: a thread owns the column (j, i)
// over kKZ planes and keeps the 3 x 3 (plane, row) window of its column in
// registers, loading one new plane's three rows per output cell (the loads
// of the plane after next are in flight while a cell is computed).  The x
// neighbours come from the adjacent lanes: a wave holds 64 consecutive x
// cells of one row and outputs the inner 62 (its two end lanes only load
// the halo).  Per output cell a field costs 3 row loads instead of 27 tap
// loads, and the taps are the same values in the same order as load_nbhd's.
constexpr int kKZ = 18;  // a multiple of three (walk3)
constexpr int kOutX = kBX - 2;
// Output rows per thread: RY = 2 for the one-field kernels (a window of 4
// rows serves two outputs), 1 where several fields' windows are live.

struct Col {
  int i, j, lane, k0, k1;  // j: the first of the thread's RY rows
  bool in;   // 0 <= i < W: the lane loads its column
  bool out;  // an inner lane of a column inside the grid: it writes results
};

// The walkers' tiles, XCD-aware.  The dispatcher deals consecutive workgroup
// ids (x fastest) round-robin to the 8 XCDs, each with its own L2, so the
// linear id L is taken as tile T = (the range of XCD L % 8) + L / 8, and T
// enumerates the tiles slab by slab: the y rows of tiles are cut into eight
// slabs, and a slab's tiles go x fastest, then y, then z.  An XCD thus walks
// one slab over every z chunk (the same work as the others, the short last
// chunk included), and the tiles it runs at once are a patch of adjacent
// rows, so the rows a tile shares with its y neighbours (the windows' halo
// rows) are fetched once into that XCD's L2.
constexpr unsigned kXCDs = 8;
__device__ __forceinline__ void xcd_tile(unsigned& bx, unsigned& by, unsigned& bz) {
  const unsigned nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
  const unsigned n = nx * ny * nz;
  const unsigned L = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned x = L % kXCDs, per = n / kXCDs, rem = n % kXCDs;
  const unsigned T = x * per + min(x, rem) + L / kXCDs;
  const unsigned sy = (ny + kXCDs - 1) / kXCDs;  // rows of tiles per slab
  const unsigned q = T / (nx * nz * sy), t = T % (nx * nz * sy);
  const unsigned rows = min(sy, ny - q * sy);     // the last slab may be shorter
  bx = t % nx;
  by = q * sy + (t / nx) % rows;
  bz = (t / nx) / rows;
}

// false: the whole wave lies past the last row
template <int RY>
__device__ __forceinline__ bool col_begin(const Grid3& g, Col& c) {
  unsigned bx, by, bz;
  xcd_tile(bx, by, bz);
  c.lane = threadIdx.x;
  c.i = int(bx) * kOutX + c.lane - 1;
  c.j = (int(by) * kBY + int(threadIdx.y)) * RY;
  c.k0 = int(bz) * kKZ;
  c.k1 = min(c.k0 + kKZ, g.D);
  c.in = unsigned(c.i) < unsigned(g.W);
  c.out = c.in && c.lane >= 1 && c.lane <= kOutX;
  return c.j < g.H;
}

template <int RY>
dim3 grid_col(const Grid3& g) {
  return dim3(unsigned((g.W + kOutX - 1) / kOutX), unsigned((g.H + kBY * RY - 1) / (kBY * RY)),
              unsigned((g.D + kKZ - 1) / kKZ));
}

// What a window fetches per cell: the loaded values, kept as loaded until the
// plane is committed a plane later (a conversion, select or sum at the load
// would make the walk wait for its loads at once), and whether the cell
// counts (else it is taken as +0).  Fetch1: the value a; Fetch2: a - b.
struct Fetch1 {
  float a;
  bool keep;
  __device__ __forceinline__ float value() const { return a; }
};
struct Fetch2 {
  float a, b;
  bool keep;
  __device__ __forceinline__ float value() const { return a - b; }
};

template <int RY, typename FT = Fetch1, int PF = 1>
struct Win27 {
  static_assert(PF == 1 || PF == 2, "prefetch distance of one or two steps");
  static constexpr int R = RY + 2;
  // three plane slots [slot][row], x offsets -1, 0, +1 (zero outside): at a
  // step of phase P (walk3) slot (a + P) % 3 holds plane k - 1 + a, so the
  // planes rotate through the slots without copies
  float l[3][R], m[3][R], r[3][R];
  // fetched planes' rows at this lane's column: a plane is fetched PF steps
  // before the step that commits it (into sv[phase of the fetching step]), so
  // PF planes' loads are in flight while a cell is computed
  FT sv[3][R];

  // issue the loads of plane kk of the field xat(kk, jj, ii) into sv[B] (no
  // waiting): every load is issued, from in-bounds addresses, and taken as +0
  // where the cell is outside the grid
  template <int B, typename XF>
  __device__ __forceinline__ void fetch(const Grid3& g, const Col& c, int kk, XF xat) {
    const bool kin = unsigned(kk) < unsigned(g.D);
    const int kc = kin ? kk : c.k0;
    const int ic = c.in ? c.i : 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int jj = c.j - 1 + q;
      const bool rin = unsigned(jj) < unsigned(g.H);
      sv[B][q] = xat(kc, rin ? jj : c.j, ic);  // the row offset is the walk's constant
      sv[B][q].keep = sv[B][q].keep & (kin & rin & c.in);
    }
  }
  // slot S <- the plane fetched into sv[B] (x neighbours from the adjacent lanes)
  template <int S, int B>
  __device__ __forceinline__ void commit(const Col&) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const float x = sv[B][q].keep ? sv[B][q].value() : 0.0f;
      l[S][q] = __shfl_up(x, 1, 64);    // the end lanes' outer values are unused
      m[S][q] = x;
      r[S][q] = __shfl_down(x, 1, 64);
    }
  }
  // the taps of output row j + ry at the centre plane, at phase P (load_nbhd
  // order, t = 9a + 3b + c)
  template <int TAPS, int P>
  __device__ __forceinline__ void taps(float v[27], int ry) const {
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int a = (t / 9 + P) % 3, b = (t / 3) % 3 + ry, cc = t % 3;
      v[t] = !tap_on<TAPS>(t) ? 0.0f : cc == 0 ? l[a][b] : cc == 1 ? m[a][b] : r[a][b];
    }
  }
  template <int P>
  __device__ __forceinline__ float centre(int ry) const {
    return m[(1 + P) % 3][1 + ry];
  }
  // planes k0 - 1 and k0 in slots 0 and 1; planes k0 + 1 .. k0 + PF fetched
  // (plane k0 + 1 + d is committed at step d, from the buffer of step d - PF)
  template <typename XF>
  __device__ __forceinline__ void start(const Grid3& g, const Col& c, XF xat) {
    fetch<0>(g, c, c.k0 - 1, xat);
    commit<0, 0>(c);
    fetch<0>(g, c, c.k0, xat);
    commit<1, 0>(c);
    fetch<(3 - PF) % 3>(g, c, c.k0 + 1, xat);
    if constexpr (PF == 2) fetch<2>(g, c, c.k0 + 2, xat);
  }
  // at output plane k (phase P): plane k + 1 into its slot, plane k + 1 + PF's
  // loads issued
  template <int P, typename XF>
  __device__ __forceinline__ void advance(const Grid3& g, const Col& c, int k, XF xat) {
    commit<(P + 2) % 3, (P + 3 - PF) % 3>(c);
    fetch<P>(g, c, k + 1 + PF, xat);
  }
};

template <int P>
using Phase = std::integral_constant<int, P>;

// Prefetch distance (steps) of the walkers.  Two steps ahead only pays in
// k_pg_flux (whose two windows leave registers for a second plane in flight);
// elsewhere the extra registers cost occupancy and the walk is no faster.
constexpr int kWalkPF = 1;

// The walk over a column's planes k0 .. k1 - 1, unrolled by three: step(P, k)
// at phase P = (k - k0) % 3, so the window slots and a walker's per-plane
// records (indexed by phase) rotate without copies, and no value loaded for
// a later step is moved (which would wait for it) at the end of a step.  With
// k1 - k0 not a multiple of three (only the last chunk of a column, kKZ being
// one) the walk ends with one or two steps past k1: they load from clamped
// in-bounds addresses and store nothing (a step stores only for k < k1).
template <typename Step>
__device__ __forceinline__ void walk3(const Col& c, Step&& step) {
  for (int k = c.k0; k < c.k1; k += 3) {
    step(Phase<0>{}, k);
    step(Phase<1>{}, k + 1);
    step(Phase<2>{}, k + 2);
  }
}

// The thread's own cells of rows c.j .. c.j + RY - 1 at plane k of field p,
// from in-bounds addresses (a row or plane past the grid re-reads the last
// one).  The walkers load them a plane ahead, as the window's planes: a value
// loaded in one step and used in the same step makes the walk wait there for
// every load in flight, the window's next plane included.
// (U: through Grid3::cellu)
template <int RY, bool U = false>
__device__ __forceinline__ void own_cells(const Grid3& g, const Col& c, int k, const float* p,
                                          float v[RY]) {
  const int ic = c.in ? c.i : 0;
  const int kc = min(k, g.D - 1);
#pragma unroll
  for (int ry = 0; ry < RY; ++ry) {
    const int jr = min(c.j + ry, g.H - 1);
    v[ry] = U ? *g.cellu(p, kc, jr, ic) : *g.cell(p, kc, jr, ic);
  }
}

// (x - A/d) + b/d, the reference's evaluation order
__device__ __forceinline__ float jacobi_form(float x, float ax, float b, float d) {
  return (x - ax / d) + b / d;
}

// (x - c A) + c b: the smoother whose 1/diag is a layer weight c
// (ARTn/Multigrid_cnn.ipynb:126-398, "w - bias(A(w)) + bias(r)" with the bias
// layer's centre weight 1/107), each product rounded as the layer's output is
__device__ __forceinline__ float jacobi_mul_form(float x, float ax, float b, float c) {
  return __fadd_rn(__fsub_rn(x, __fmul_rn(c, ax)), __fmul_rn(c, b));
}

// The obstacle damping f / (1 + dt sigma) (bluff_body, UM/AI_HFM_v3.ipynb cell
// 19): one expression for every kernel that applies it, so fused and separate
// passes round alike.
__device__ __forceinline__ float bluff_apply(float f, float s, float dt) {
  return f / __builtin_fmaf(dt, s, 1.0f);
}

// ------------------------------------------------------------------ kernels
template <int TAPS>
__global__ __launch_bounds__(kBX* kBY) void k_conv27(const float* __restrict__ x,
                                                     float* __restrict__ y,
                                                     const float* __restrict__ b, Weights K,
                                                     float d, Grid3 g, int form) {
  Col c;
  if (!col_begin<2>(g, c)) return;
  auto xat = [&](int kk, int jj, int ii) { return Fetch1{*g.cell(x, kk, jj, ii), true}; };
  const int f = form & HFM_FORM_MASK;
  const float* bp = f == HFM_FORM_APPLY ? x : b;  // APPLY reads no b (x stands in)
  constexpr int PF = kWalkPF;
  Win27<2, Fetch1, PF> win;
  win.start(g, c, xat);
  float bv[3][2];  // b of the plane of phase P in bv[P], loaded PF steps ahead
  for (int d = 0; d < PF; ++d) own_cells<2>(g, c, c.k0 + d, bp, bv[d]);
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
    win.template advance<P>(g, c, k, xat);
    own_cells<2>(g, c, k + PF, bp, bv[(P + PF) % 3]);
#pragma unroll
    for (int ry = 0; ry < 2; ++ry) {
      float v[27];
      win.template taps<TAPS, P>(v, ry);
      const float ax = wsum<TAPS>(K, v);
      if (c.out && k < c.k1 && c.j + ry < g.H) {
        float* const yp = g.cell(y, k, c.j + ry, c.i);
        const float bc = bv[P][ry];
        const float xc = win.template centre<P>(ry);
        float out = ax;
        if (f == HFM_FORM_RESID) out = ax - bc;
        else if (f == HFM_FORM_JACOBI) out = jacobi_form(xc, ax, bc, d);
        else if (f == HFM_FORM_SUBTRACT) out = bc - ax;
        else if (f == HFM_FORM_AXPY) out = bc + d * ax;
        else if (f == HFM_FORM_JACOBI_MUL) out = jacobi_mul_form(xc, ax, bc, d);
        if ((form & HFM_FORM_ZERO_OUTLET) && c.i == g.W - 1) out = 0.0f;
        *yp = out;
      }
    }
  });
}

// The fine residual r = A x - b (k_conv27's RESID form: the same taps and
// sums; with HFM_FORM_ZERO_OUTLET r = 0 on the x = W-1 plane) restricted in
// the same walk to the next level, rc = k_restrict_mean(r) with factors
// (2, 2, 2) and weight 1/8; r itself is not written.  A thread's two rows are
// a row pair of the restriction, the x pairs are adjacent lanes (an output
// tile starts on an even x) and the z pairs consecutive planes (chunks start
// on even planes), and the eight children are summed in k_restrict_mean's
// order (plane, row, cell), so rc is bitwise the two separate launches'.
// D, H and W are even.
template <int TAPS>
__global__ __launch_bounds__(kBX* kBY) void k_resid_restrict27(const float* __restrict__ x,
                                                               const float* __restrict__ b,
                                                               float* __restrict__ rc, Weights K,
                                                               Grid3 g, int form) {
  Col c;
  if (!col_begin<2>(g, c)) return;
  auto xat = [&](int kk, int jj, int ii) { return Fetch1{*g.cell(x, kk, jj, ii), true}; };
  constexpr int PF = kWalkPF;
  Win27<2, Fetch1, PF> win;
  win.start(g, c, xat);
  float bv[3][2];  // b of the plane of phase P in bv[P], loaded PF steps ahead
  for (int d = 0; d < PF; ++d) own_cells<2>(g, c, c.k0 + d, b, bv[d]);
  const bool zo = (form & HFM_FORM_ZERO_OUTLET) != 0;
  const Grid3 gc{g.D / 2, g.H / 2, g.W / 2};
  const bool owner = c.out && (c.i & 1) == 0;  // the lane of a pair's even x
  constexpr float wgt = 1.0f / 8.0f;
  float acc = 0.0f;  // the coarse cell's sum over a plane pair
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
    win.template advance<P>(g, c, k, xat);
    own_cells<2>(g, c, k + PF, b, bv[(P + PF) % 3]);
    float r[2];
#pragma unroll
    for (int ry = 0; ry < 2; ++ry) {
      float v[27];
      win.template taps<TAPS, P>(v, ry);
      const float ax = wsum<TAPS>(K, v);
      r[ry] = ax - bv[P][ry];
      if (zo && c.i == g.W - 1) r[ry] = 0.0f;
    }
    const float r0x = __shfl_down(r[0], 1, 64), r1x = __shfl_down(r[1], 1, 64);
    if (!(k & 1)) acc = 0.0f;
    acc += wgt * r[0];
    acc += wgt * r0x;
    acc += wgt * r[1];
    acc += wgt * r1x;
    if ((k & 1) && k < c.k1 && owner) *gc.cell(rc, k >> 1, c.j >> 1, c.i >> 1) = acc;
  });
}

__global__ __launch_bounds__(kBX* kBY) void k_var6(const float* __restrict__ x,
                                                   float* __restrict__ y,
                                                   const float* __restrict__ b, Var6 V, Grid3 g,
                                                   int form) {
  int k, j, i;
  if (!cell_of_thread(g, k, j, i)) return;
  const int64_t n = g.idx(k, j, i);
  const float ax = var6_at(V, g, k, j, i, [&](int kk, int jj, int ii) { return *g.cell(x, kk, jj, ii); });
  float out = ax;
  if (form == HFM_FORM_RESID) out = ax - b[n];
  else if (form == HFM_FORM_JACOBI) out = jacobi_form(x[n], ax, b[n], V.tot[n]);
  y[n] = out;
}

__global__ __launch_bounds__(kBX* kBY) void k_restrict_mean(const float* __restrict__ x,
                                                            float* __restrict__ y, Grid3 gf,
                                                            Grid3 gc, int fz, int fy, int fx,
                                                            float wgt) {
  int K, J, I;
  if (!cell_of_thread(gc, K, J, I)) return;
  float acc = 0.0f;
  for (int a = 0; a < fz; ++a)
    for (int b = 0; b < fy; ++b)
      for (int c = 0; c < fx; ++c) acc += wgt * x[gf.idx(K * fz + a, J * fy + b, I * fx + c)];
  y[gc.idx(K, J, I)] = acc;
}

__global__ __launch_bounds__(kBX* kBY) void k_prolong_nearest(const float* __restrict__ xc,
                                                              float* __restrict__ y, Grid3 gf,
                                                              Grid3 gc, int fz, int fy, int fx) {
  int k, j, i;
  if (!cell_of_thread(gf, k, j, i)) return;
  y[gf.idx(k, j, i)] = xc[gc.idx(k / fz, j / fy, i / fx)];
}

// UpSampling by (fz, fy, fx) then the all-ones convolution of window size 2
// on every upsampled axis with 'SAME' padding, which for an even window pads
// one zero after the end only: y(k, j, i) = sum over k..k+fz-1, j..j+fy-1,
// i..i+fx-1 (inside the grid) of the upsampled field
// (ARTn/Multigrid_cnn.ipynb:342-398).  Terms are summed in window order.
__global__ __launch_bounds__(kBX* kBY) void k_prolong_upsum(const float* __restrict__ xc,
                                                            float* __restrict__ y, Grid3 gf,
                                                            Grid3 gc, int fz, int fy, int fx) {
  int k, j, i;
  if (!cell_of_thread(gf, k, j, i)) return;
  float acc = 0.0f;
  for (int a = 0; a < fz; ++a) {
    const int kk = k + a;
    if (kk >= gf.D) break;
    for (int b = 0; b < fy; ++b) {
      const int jj = j + b;
      if (jj >= gf.H) break;
      for (int c = 0; c < fx; ++c) {
        const int ii = i + c;
        if (ii >= gf.W) break;
        acc += xc[gc.idx(kk / fz, jj / fy, ii / fx)];
      }
    }
  }
  y[gf.idx(k, j, i)] = acc;
}

// Fixed-order FP64 sum of n floats by one workgroup: thread t sums t, t +
// 1024, ... in order, then the 1024 partials are added pairwise in a fixed
// tree.
__global__ __launch_bounds__(1024) void k_sum_f64(const float* __restrict__ x, int64_t n,
                                                  double* __restrict__ out) {
  __shared__ double part[1024];
  double acc = 0.0;
  for (int64_t q = threadIdx.x; q < n; q += 1024) acc += double(x[q]);
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (int(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

template <bool VAR6, int TAPS>
__global__ __launch_bounds__(kBX* kBY) void k_prolong_jacobi(const float* __restrict__ wc,
                                                             const float* __restrict__ r,
                                                             float* __restrict__ y, Weights K,
                                                             float d, Var6 V,
                                                             Grid3 gf, Grid3 gc, int fz, int fy,
                                                             int fx) {
  int k, j, i;
  if (!cell_of_thread(gf, k, j, i)) return;
  const int64_t n = gf.idx(k, j, i);
  auto wat = [&](int kk, int jj, int ii) { return wc[gc.idx(kk / fz, jj / fy, ii / fx)]; };
  const float wp = wat(k, j, i);
  if (VAR6) {
    y[n] = jacobi_form(wp, var6_at(V, gf, k, j, i, wat), r[n], V.tot[n]);
  } else {
    y[n] = jacobi_form(wp, conv_at<TAPS>(K, gf, k, j, i, wat), r[n], d);
  }
}

// The same two kernels for the constant 3x3x3 stencil, as column walkers.
template <int TAPS>
__global__ __launch_bounds__(kBX* kBY) void k_prolong_jacobi_col(const float* __restrict__ wc,
                                                                 const float* __restrict__ r,
                                                                 float* __restrict__ y, Weights K,
                                                                 float d, Grid3 gf, Grid3 gc,
                                                                 int fz, int fy, int fx) {
  Col c;
  if (!col_begin<2>(gf, c)) return;
  auto wat = [&](int kk, int jj, int ii) { return Fetch1{*gc.cell(wc, kk / fz, jj / fy, ii / fx), true}; };
  constexpr int PF = kWalkPF;
  Win27<2, Fetch1, PF> win;
  win.start(gf, c, wat);
  float rv[3][2];  // r of the plane of phase P in rv[P], loaded PF steps ahead
  for (int d = 0; d < PF; ++d) own_cells<2>(gf, c, c.k0 + d, r, rv[d]);
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
    win.template advance<P>(gf, c, k, wat);
    own_cells<2>(gf, c, k + PF, r, rv[(P + PF) % 3]);
#pragma unroll
    for (int ry = 0; ry < 2; ++ry) {
      float v[27];
      win.template taps<TAPS, P>(v, ry);
      const float ax = wsum<TAPS>(K, v);
      if (c.out && k < c.k1 && c.j + ry < gf.H)
        *gf.cell(y, k, c.j + ry, c.i) =
            jacobi_form(win.template centre<P>(ry), ax, rv[P][ry], d);
    }
  });
}

// The correction w is a field of the fine grid, or (PROL) the coarse field
// wc read through the nearest prolongation: w(k, j, i) = wc(k >> sz, j >> sy,
// i >> sx), which is bitwise the prolonged field without writing it.
struct CorrW {
  const float* w;
  Grid3 gc;
  int sz, sy, sx;
  template <bool PROL>
  __device__ __forceinline__ float at(const Grid3& g, int kk, int jj, int ii) const {
    return PROL ? *gc.cell(w, kk >> sz, jj >> sy, ii >> sx) : *g.cell(w, kk, jj, ii);
  }
};

template <int TAPS, bool PROL>
__global__ __launch_bounds__(kBX* kBY) void k_correct_jacobi_col(const float* __restrict__ x,
                                                                 CorrW cw,
                                                                 const float* __restrict__ b,
                                                                 float* __restrict__ y, Weights K,
                                                                 float d, Grid3 g, int flags) {
  Col c;
  if (!col_begin<2>(g, c)) return;
  const bool zo = flags & HFM_CORRECT_ZERO_OUTLET;
  // x - w, or 0 at the outlet (HFM_CORRECT_ZERO_OUTLET)
  auto vat = [&](int kk, int jj, int ii) {
    return Fetch2{*g.cell(x, kk, jj, ii), cw.template at<PROL>(g, kk, jj, ii),
                  !(zo && ii == g.W - 1)};
  };
  constexpr int PF = kWalkPF;
  Win27<2, Fetch2, PF> win;
  win.start(g, c, vat);
  float bv[3][2];  // b of the plane of phase P in bv[P], loaded PF steps ahead
  for (int d = 0; d < PF; ++d) own_cells<2>(g, c, c.k0 + d, b, bv[d]);
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
    win.template advance<P>(g, c, k, vat);
    own_cells<2>(g, c, k + PF, b, bv[(P + PF) % 3]);
#pragma unroll
    for (int ry = 0; ry < 2; ++ry) {
      float v[27];
      win.template taps<TAPS, P>(v, ry);
      const float ax = wsum<TAPS>(K, v);
      if (c.out && k < c.k1 && c.j + ry < g.H) {
        const float xc = win.template centre<P>(ry);
        *g.cell(y, k, c.j + ry, c.i) = (flags & HFM_CORRECT_RECIPROCAL)
                                           ? jacobi_mul_form(xc, ax, bv[P][ry], d)
                                           : jacobi_form(xc, ax, bv[P][ry], d);
      }
    }
  });
}

template <bool VAR6, int TAPS, bool PROL>
__global__ __launch_bounds__(kBX* kBY) void k_correct_jacobi(const float* __restrict__ x,
                                                             CorrW cw,
                                                             const float* __restrict__ b,
                                                             float* __restrict__ y, Weights K,
                                                             float d, Var6 V,
                                                             Grid3 g, int flags) {
  int k, j, i;
  if (!cell_of_thread(g, k, j, i)) return;
  const int64_t n = g.idx(k, j, i);
  const bool zo = flags & HFM_CORRECT_ZERO_OUTLET;
  auto vat = [&](int kk, int jj, int ii) {
    const int64_t m = g.idx(kk, jj, ii);
    return (zo && ii == g.W - 1) ? 0.0f : x[m] - cw.template at<PROL>(g, kk, jj, ii);
  };
  const float v = vat(k, j, i);
  if (VAR6) {
    y[n] = jacobi_form(v, var6_at(V, g, k, j, i, vat), b[n], V.tot[n]);
  } else {
    y[n] = jacobi_form(v, conv_at<TAPS>(K, g, k, j, i, vat), b[n], d);
  }
}

// Explicit advection-diffusion of NF fields phi by the velocity (u, v, w):
//   out_f = f * (((dif(phi) - u*ax(phi)) - v*ay(phi)) - w*az(phi)) + base_f
// (UM/AI_HFM_v3.ipynb cell 25: a_u = dif(u) - u*xadv(u) - v*yadv(u) - w*zadv(u);
// b_u = 0.5*a_u + u; values_u = values_u + c_u; the scalar alpha the same way).
struct Kern4 {
  Weights k[4];       // dif, ax, ay, az
  uint32_t mask[4];
};

// Petrov-Galerkin dissipation (UM/INHALE_demo.ipynb cell 20), in the
// reference's evaluation order.  Inputs are the damped velocities
// f~ = f / (1 + dt sigma).  Coefficients, per component f~:
//   k = min(8*0.25 |1/3 (|u~| + |v~| + |w~|) dif(f~)| / (eps + (|ax| + |ay| + |az|)/3), 0.5)
//       / (1 + dt sigma)
// with one 27-point neighbourhood load per component serving the four stencils.
template <int TX, int TY, int TZ>
__global__ __launch_bounds__(kBX* kBY) void k_pg_coef(const float* __restrict__ u,
                                                      const float* __restrict__ v,
                                                      const float* __restrict__ w,
                                                      const float* __restrict__ sigma, float* k0,
                                                      float* k1, float* k2, Kern4 K, float dt,
                                                      float eps, Grid3 g) {
  Col c;
  if (!col_begin<1>(g, c)) return;
  const float* f[3] = {u, v, w};
  float* out[3] = {k0, k1, k2};
  constexpr int PF = kWalkPF;
  Win27<1, Fetch1, PF> win[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float* x = f[q];
    win[q].start(g, c, [&](int kk, int jj, int ii) { return Fetch1{*g.cellu(x, kk, jj, ii), true}; });
  }
  float sv[3][1];  // sigma of the plane of phase P in sv[P], loaded PF steps ahead
  for (int d = 0; d < PF; ++d) own_cells<1, true>(g, c, c.k0 + d, sigma, sv[d]);
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float* x = f[q];
      win[q].template advance<P>(
          g, c, k, [&](int kk, int jj, int ii) { return Fetch1{*g.cellu(x, kk, jj, ii), true}; });
    }
    own_cells<1, true>(g, c, k + PF, sigma, sv[(P + PF) % 3]);
    const float speed = (1.0f / 3.0f) * ((fabsf(win[0].template centre<P>(0)) +
                                          fabsf(win[1].template centre<P>(0))) +
                                         fabsf(win[2].template centre<P>(0)));
    const float damp = 1.0f + dt * sv[P][0];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float nb[27];
      win[q].template taps<27, P>(nb, 0);
      const float dif = wsum<27>(K.k[0], nb);
      const float ax = wsum<TX>(K.k[1], nb);
      const float ay = wsum<TY>(K.k[2], nb);
      const float az = wsum<TZ>(K.k[3], nb);
      const float num = 2.0f * fabsf(speed * dif);
      const float den = eps + ((fabsf(ax) + fabsf(ay)) + fabsf(az)) / 3.0f;
      if (c.out && k < c.k1) *g.cellu(out[q], k, c.j, c.i) = fminf(num / den, 0.5f) / damp;
    }
  });
}

// The dissipation flux 3*0.5 (k dif(f~) + dif(f~ k) - f~ dif(k)) per
// component: the neighbourhoods of f~ and k are loaded once and the product
// field f~ k is formed tap by tap (zero outside the grid, as its SAME padding).
__global__ __launch_bounds__(kBX* kBY) void k_pg_flux(const float* f0, const float* f1,
                                                      const float* f2, const float* c0,
                                                      const float* c1, const float* c2, float* o0,
                                                      float* o1, float* o2, Weights Kd, Grid3 g) {
  Col c;
  if (!col_begin<1>(g, c)) return;
  const float* f[3] = {f0, f1, f2};
  const float* kf[3] = {c0, c1, c2};
  float* out[3] = {o0, o1, o2};
  // one component at a time over the column (two windows each)
#pragma unroll 1
  for (int q = 0; q < 3; ++q) {
    const float* x = f[q];
    const float* y = kf[q];
    auto xat = [&](int kk, int jj, int ii) { return Fetch1{*g.cell(x, kk, jj, ii), true}; };
    auto yat = [&](int kk, int jj, int ii) { return Fetch1{*g.cell(y, kk, jj, ii), true}; };
    Win27<1, Fetch1, 2> wf, wk;
    wf.start(g, c, xat);
    wk.start(g, c, yat);
    walk3(c, [&](auto ph, int k) {
      constexpr int P = decltype(ph)::value;
      wf.template advance<P>(g, c, k, xat);
      wk.template advance<P>(g, c, k, yat);
      float fv[27], kv[27], fk[27];
      wf.template taps<27, P>(fv, 0);
      wk.template taps<27, P>(kv, 0);
#pragma unroll
      for (int t = 0; t < 27; ++t) fk[t] = fv[t] * kv[t];
      const float dif_f = wsum<27>(Kd, fv);
      const float dif_fk = wsum<27>(Kd, fk);
      const float dif_k = wsum<27>(Kd, kv);
      if (c.out && k < c.k1)
        *g.cell(out[q], k, c.j, c.i) = 1.5f * ((kv[13] * dif_f + dif_fk) - fv[13] * dif_k);
    });
  }
}

// out_q = f * (dif(phi_q) - u ax(phi_q) - v ay(phi_q) - w az(phi_q)) + base_q
// for NF fields by one column walker.  The four stencils are folded into one
// weight set per cell, dif - u ax - v ay - w az (the velocities are the
// cell's), applied to each field: 27 + 81 / NF FMAs per field instead of 108.
// This sums the same products in another order than four separate
// convolutions, so results agree with the layer-by-layer form to rounding.
// TX / TY / TZ: the tap sets of the three advection stencils; the fold skips
// their structurally zero taps (54 FMAs instead of 81 for the first-derivative
// sets), which leaves every finite result unchanged.
// SH (three fields): 1 when the velocities are the advected fields (u, v, w =
// phi, the momentum stages), 2 when the bases are as well (the predictor): the
// cell's velocities (and bases) are then the windows' centres, not loads of
// their own, which spares their loads, registers and pointers.
template <int NF, bool DIS, int SH = 0, int TX = 27, int TY = 27, int TZ = 27>
__global__ __launch_bounds__(kBX* kBY) void k_advect27_v(
    const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
    const float* p0, const float* p1, const float* p2, const float* b0, const float* b1,
    const float* b2, float* o0, float* o1, float* o2, Kern4 K, float f, Grid3 g,
    const float* d0, const float* d1, const float* d2, const float* __restrict__ sigma, float dt,
    int post) {
  Col c;
  if (!col_begin<1>(g, c)) return;
  const float* dis[3] = {d0, d1, d2};
  const float* phi[3] = {p0, p1, p2};
  const float* base[3] = {b0, b1, b2};
  float* out[3] = {o0, o1, o2};
  constexpr int PF = kWalkPF;
  Win27<1, Fetch1, PF> win[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const float* x = phi[q];
    win[q].start(g, c, [&](int kk, int jj, int ii) { return Fetch1{*g.cell(x, kk, jj, ii), true}; });
  }
  // the cell's own values (velocity, bases, dissipation, sigma), loaded one
  // plane ahead (the last plane's successor re-reads the last plane) into the
  // record of the next phase
  struct Cells {
    float u, v, w, b[NF], d[NF], s;
  };
  const bool has_s = post && sigma;
  const float* sg = has_s ? sigma : u;
  static_assert(SH == 0 || NF == 3, "shared velocities need three fields");
  auto load_cells = [&](int kk, Cells& cv) {
    const int kc = min(kk, g.D - 1), ic = c.in ? c.i : 0;
    if (SH == 0) {
      cv.u = *g.cell(u, kc, c.j, ic);
      cv.v = *g.cell(v, kc, c.j, ic);
      cv.w = *g.cell(w, kc, c.j, ic);
    }
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      if (SH < 2) cv.b[q] = *g.cell(base[q], kc, c.j, ic);
      cv.d[q] = DIS ? *g.cell(dis[q], kc, c.j, ic) : 0.0f;
    }
    cv.s = *g.cell(sg, kc, c.j, ic);  // sigma (or u, unused): every load is issued
  };
  Cells rec[3];
  for (int d = 0; d < PF; ++d) load_cells(c.k0 + d, rec[d]);
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
    const Cells& cur = rec[P];
    // the cell's velocities: the windows' centre plane (committed a step
    // earlier) when they are the advected fields
    const float uc = SH ? win[0].template centre<P>(0) : cur.u;
    const float vc = SH ? win[NF > 1 ? 1 : 0].template centre<P>(0) : cur.v;
    const float wc = SH ? win[NF > 2 ? 2 : 0].template centre<P>(0) : cur.w;
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      const float* x = phi[q];
      win[q].template advance<P>(
          g, c, k, [&](int kk, int jj, int ii) { return Fetch1{*g.cell(x, kk, jj, ii), true}; });
    }
    load_cells(k + PF, rec[(P + PF) % 3]);
    // the cell's combined weights dif - u ax - v ay - w az, shared by the
    // NF fields: half the FMAs of four stencils per field
    float cw[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      float a = DIS ? 0.0f : K.k[0].w[t];
      if (tap_on<TX>(t)) a = __builtin_fmaf(-uc, K.k[1].w[t], a);
      if (tap_on<TY>(t)) a = __builtin_fmaf(-vc, K.k[2].w[t], a);
      if (tap_on<TZ>(t)) a = __builtin_fmaf(-wc, K.k[3].w[t], a);
      cw[t] = a;
    }
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      float t[27];
      win[q].template taps<27, P>(t, 0);
      const float bq = SH == 2 ? win[q].template centre<P>(0) : cur.b[q];
      // DIS: the dissipation term is given (dis0..2 in place of the phi's
      // diffusion stencil, UM/INHALE_demo.ipynb cell 24)
      float a = DIS ? cur.d[q] : 0.0f;
#pragma unroll
      for (int s = 0; s < 27; ++s) a = __builtin_fmaf(cw[s], t[s], a);
      if (c.out && k < c.k1) {
        float o = f * a + bq;
        if (post) {
          // the scalar's fill (every boundary cell 0) and nbluff dampings
          const bool face = c.i == 0 || c.i == g.W - 1 || c.j == 0 || c.j == g.H - 1 || k == 0 ||
                            k == g.D - 1;
          if ((post & HFM_ADV_INDICATOR) && face) o = 0.0f;
          const int nb = (post >> HFM_ADV_BLUFF_SHIFT) & 3;
          const float s = has_s ? cur.s : 0.0f;
          for (int r = 0; r < nb; ++r) o = bluff_apply(o, s, dt);
        }
        *g.cell(out[q], k, c.j, c.i) = o;
      }
    }
  });
}

// y = scale * sum_q K_q (*) x_q (q < n <= 3); SOURCE_BC: the x = 0 plane takes
// the value computed at x = 1 (boundary_condition_source, UM/AI_HFM_v3.ipynb cell 19).
struct Kern3 {
  Weights k[3];
  uint32_t mask[3];
};


// RY: output rows per thread (a window of RY + 2 rows per field).
template <int NF, int T0 = 27, int T1 = 27, int T2 = 27, int RY = 1>
__global__ __launch_bounds__(kBX* kBY) void k_conv27_sum_v(const float* x0, const float* x1,
                                                           const float* x2, float* __restrict__ y,
                                                           Kern3 K, float scale, Grid3 g,
                                                           int flags) {
  Col c;
  if (!col_begin<RY>(g, c)) return;
  const float* xs[3] = {x0, x1, x2};
  Win27<RY, Fetch1, kWalkPF> win[NF];
#pragma unroll
  for (int q = 0; q < NF; ++q) {
    const float* xq = xs[q];
    win[q].start(g, c, [&](int kk, int jj, int ii) { return Fetch1{*g.cell(xq, kk, jj, ii), true}; });
  }
  // boundary_condition_source: the x = 0 cell takes the x = 1 cell's sum
  const bool src = (flags & HFM_SUM_SOURCE_BC) && g.W > 1;
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      const float* xq = xs[q];
      win[q].template advance<P>(
          g, c, k, [&](int kk, int jj, int ii) { return Fetch1{*g.cell(xq, kk, jj, ii), true}; });
    }
#pragma unroll
    for (int ry = 0; ry < RY; ++ry) {
      float acc = 0.0f;
#pragma unroll
      for (int q = 0; q < NF; ++q) {
        float t[27];
        win[q].template taps<27, P>(t, ry);
        acc += q == 0 ? wsum<T0>(K.k[0], t) : q == 1 ? wsum<T1>(K.k[1], t) : wsum<T2>(K.k[2], t);
      }
      const float from_right = __shfl_down(acc, 1, 64);
      if (src && c.i == 0) acc = from_right;
      if (c.out && k < c.k1 && c.j + ry < g.H) *g.cell(y, k, c.j + ry, c.i) = scale * acc;
    }
  });
}

// The urban step's pressure gradient for the three velocity components in one
// walk over p: o_q = f_q' - K_q (*) p' (the SUBTRACT form of k_conv27 per
// component, the same taps and sums), where
//   p' = boundary_condition_pressure(p) (HFM_GRAD_P_BC): each tap reads p at the
//        cell the fill's last assignment copies from (0 at the outlet); pout,
//        when given, receives p' itself;
//   f' = bluff(f) (HFM_GRAD_PRE_BLUFF);
// and with HFM_GRAD_POST_BLUFF o_q is damped as well, except at the cells the
// velocity fill copies from (u at j = 1, j = H-2 and k = D-2; v at k = D-2):
// hfm_bc_velocity_bluff_inplace_f32 gathers those undamped and damps them.
template <int T0, int T1, int T2>
__global__ __launch_bounds__(kBX* kBY) void k_grad3(const float* __restrict__ p,
                                                    const float* f0, const float* f1,
                                                    const float* f2, float* o0, float* o1,
                                                    float* o2, float* __restrict__ pout, Kern3 K,
                                                    const float* __restrict__ sigma, float dt,
                                                    Grid3 g, int flags) {
  __shared__ __align__(16) WeightLds<3> kw;
  kw.stage(K.k);
  Col c;
  if (!col_begin<2>(g, c)) return;
  const bool pbc = flags & HFM_GRAD_P_BC;
  auto pat = [&](int kk, int jj, int ii) {
    // boundary_condition_pressure: the last assignment covering the cell wins
    // (x = W-1 -> 0, x = 0, y = 0, y = H-1, z = 0, z = D-1 copy their
    // neighbour), as selects: a branch between the walk's loads would make it
    // wait for all of them
    const bool kt = pbc & (kk == g.D - 1), kb = pbc & (kk == 0);
    const bool kf = !(kt | kb);
    const bool jt = kf & pbc & (jj == g.H - 1), jb = kf & pbc & (jj == 0);
    const bool jf = kf & !(jt | jb);
    const bool ib = jf & pbc & (ii == 0);
    const bool zero = jf & pbc & !ib & (ii == g.W - 1);
    const int k2 = kt ? g.D - 2 : (kb ? 1 : kk);
    const int j2 = jt ? g.H - 2 : (jb ? 1 : jj);
    const int i2 = ib ? 1 : ii;
    return Fetch1{*g.cellu(p, k2, j2, i2), !zero};
  };
  const float* fin[3] = {f0, f1, f2};
  float* out[3] = {o0, o1, o2};
  const bool damp = flags & (HFM_GRAD_PRE_BLUFF | HFM_GRAD_POST_BLUFF);
  const float* sg = damp ? sigma : p;  // sigma, or p (unused): every load is issued
  // the cells' own values (f, sigma) of a plane, loaded one plane ahead from
  // in-bounds addresses (a row past the grid re-reads the last row)
  struct Cells {
    float f[2][3], s[2];
  };
  const int ic = c.in ? c.i : 0;
  auto load_cells = [&](int kk, Cells& cv) {
    const int kc = min(kk, g.D - 1);
#pragma unroll
    for (int ry = 0; ry < 2; ++ry) {
      const int jr = min(c.j + ry, g.H - 1);
#pragma unroll
      for (int q = 0; q < 3; ++q) cv.f[ry][q] = *g.cellu(fin[q], kc, jr, ic);
      cv.s[ry] = *g.cellu(sg, kc, jr, ic);
    }
  };
  constexpr int PF = kWalkPF;
  Win27<2, Fetch1, PF> win;
  win.start(g, c, pat);
  Cells rec[3];  // the cells of the plane of phase P in rec[P], loaded PF steps ahead
  for (int d = 0; d < PF; ++d) load_cells(c.k0 + d, rec[d]);
  walk3(c, [&](auto ph, int k) {
    constexpr int P = decltype(ph)::value;
    const Cells& cur = rec[P];
    win.template advance<P>(g, c, k, pat);
    load_cells(k + PF, rec[(P + PF) % 3]);
    // the two rows' gradient sums, each weight read once for both
    float ax[2][3];
    {
      float v0[27], v1[27];
      win.template taps<27, P>(v0, 0);
      win.template taps<27, P>(v1, 1);
      wsum2_lds<T0>(&kw.w[0][zero4()], v0, v1, ax[0][0], ax[1][0]);
      wsum2_lds<T1>(&kw.w[1][zero4()], v0, v1, ax[0][1], ax[1][1]);
      wsum2_lds<T2>(&kw.w[2][zero4()], v0, v1, ax[0][2], ax[1][2]);
    }
#pragma unroll
    for (int ry = 0; ry < 2; ++ry) {
      const int j = c.j + ry;
      if (c.out && k < c.k1 && j < g.H) {
        const float s = cur.s[ry];
        const bool src_u = j == 1 || j == g.H - 2 || k == g.D - 2, src_v = k == g.D - 2;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          float fv = cur.f[ry][q];
          if (flags & HFM_GRAD_PRE_BLUFF) fv = bluff_apply(fv, s, dt);
          float o = fv - ax[ry][q];
          const bool src = q == 0 ? src_u : q == 1 ? src_v : false;
          if ((flags & HFM_GRAD_POST_BLUFF) && !src) o = bluff_apply(o, s, dt);
          *g.cellu(out[q], k, j, c.i) = o;
        }
        if (pout) *g.cellu(pout, k, j, c.i) = win.template centre<P>(ry);
      }
    }
  });
}

// Boundary fills: the value of the last reference assignment covering a cell.
__global__ __launch_bounds__(kBX* kBY) void k_bc(const float* __restrict__ in,
                                                 float* __restrict__ out, Grid3 g, int kind) {
  int k, j, i;
  if (!cell_of_thread(g, k, j, i)) return;
  const int D = g.D, H = g.H, W = g.W;
  float v = in[g.idx(k, j, i)];
  switch (kind) {
    case HFM_BC_TOP_COPY:
      if (k == 0) v = in[g.idx(1, j, i)];
      break;
    case HFM_BC_PRESSURE:
      if (i == W - 1) v = 0.0f;
      if (i == 0) v = in[g.idx(k, j, 1)];
      if (j == 0) v = in[g.idx(k, 1, i)];
      if (j == H - 1) v = in[g.idx(k, H - 2, i)];
      if (k == 0) v = in[g.idx(1, j, i)];
      if (k == D - 1) v = in[g.idx(D - 2, j, i)];
      break;
    case HFM_BC_SOURCE:
      if (i == 0) v = in[g.idx(k, j, 1)];
      break;
    case HFM_BC_INDICATOR:
      if (i == 0 || i == W - 1 || j == 0 || j == H - 1 || k == 0 || k == D - 1) v = 0.0f;
      break;
    case HFM_BC_ZERO_OUTLET:
      if (i == W - 1) v = 0.0f;
      break;
  }
  out[g.idx(k, j, i)] = v;
}

__global__ __launch_bounds__(kBX* kBY) void k_bc_velocity(const float* __restrict__ u,
                                                          const float* __restrict__ v,
                                                          const float* __restrict__ w,
                                                          float* __restrict__ uo,
                                                          float* __restrict__ vo,
                                                          float* __restrict__ wo, Grid3 g,
                                                          float ub,
                                                          const float* __restrict__ inflow) {
  int k, j, i;
  if (!cell_of_thread(g, k, j, i)) return;
  const int D = g.D, H = g.H, W = g.W;
  const int64_t n = g.idx(k, j, i);
  float a = u[n], b = v[n], c = w[n];
  if (i == 0 || i == W - 1) {
    a = inflow ? inflow[k] : ub;
    b = 0.0f;
    c = 0.0f;
  }
  if (j == 0) {
    a = u[g.idx(k, 1, i)];
    b = 0.0f;
    c = 0.0f;
  }
  if (j == H - 1) {
    a = u[g.idx(k, H - 2, i)];
    b = 0.0f;
    c = 0.0f;
  }
  if (k == 0) {
    a = 0.0f;
    b = 0.0f;
    c = 0.0f;
  }
  if (k == D - 1) {
    a = u[g.idx(D - 2, j, i)];
    b = v[g.idx(D - 2, j, i)];
    c = 0.0f;
  }
  uo[n] = a;
  vo[n] = b;
  wo[n] = c;
}

// In-place velocity fill.  The values boundary cells copy from the rows and
// plane next to them (u at j = 1 and j = H-2, u and v at k = D-2) are
// gathered before any cell is written; then each boundary cell is written
// once, by one thread, with k_bc_velocity's rule, so the result is bitwise
// the out-of-place fill's.  scratch: [u(k,1,i) | u(k,H-2,i) | u(D-2,j,i) |
// v(D-2,j,i)], 2 D W + 2 H W floats.
__global__ __launch_bounds__(256) void k_bc_velocity_gather(const float* __restrict__ u,
                                                            const float* __restrict__ v,
                                                            float* __restrict__ scratch, Grid3 g) {
  const int64_t DW = int64_t(g.D) * g.W, HW = int64_t(g.H) * g.W;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < DW) {
    const int k = int(t / g.W), i = int(t % g.W);
    scratch[t] = u[g.idx(k, 1, i)];
    scratch[DW + t] = u[g.idx(k, g.H - 2, i)];
  }
  if (t < HW) {
    const int j = int(t / g.W), i = int(t % g.W);
    scratch[2 * DW + t] = u[g.idx(g.D - 2, j, i)];
    scratch[2 * DW + HW + t] = v[g.idx(g.D - 2, j, i)];
  }
}

// Boundary cells, each once: the two x faces (all k, j), then the two y
// faces without their x-edge cells, then the two z faces without edges.
__global__ __launch_bounds__(256) void k_bc_velocity_inplace(float* u, float* v, float* w,
                                                             const float* __restrict__ scratch,
                                                             Grid3 g, float ub,
                                                             const float* __restrict__ inflow,
                                                             const float* __restrict__ sigma,
                                                             float dt) {
  const int D = g.D, H = g.H, W = g.W;
  const int64_t nx = 2 * int64_t(D) * H, ny = 2 * int64_t(D) * (W - 2);
  const int64_t nz = 2 * int64_t(H - 2) * (W - 2);
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int k, j, i;
  if (t < nx) {
    i = (t & 1) ? W - 1 : 0;
    t >>= 1;
    k = int(t / H);
    j = int(t % H);
  } else if ((t -= nx) < ny) {
    j = (t & 1) ? H - 1 : 0;
    t >>= 1;
    k = int(t / (W - 2));
    i = 1 + int(t % (W - 2));
  } else if ((t -= ny) < nz) {
    k = (t & 1) ? D - 1 : 0;
    t >>= 1;
    j = 1 + int(t / (W - 2));
    i = 1 + int(t % (W - 2));
  } else {
    return;
  }
  const int64_t DW = int64_t(D) * W, HW = int64_t(H) * W;
  const int64_t n = g.idx(k, j, i);
  float a = u[n], b = v[n], c = w[n];
  if (i == 0 || i == W - 1) {
    a = inflow ? inflow[k] : ub;
    b = 0.0f;
    c = 0.0f;
  }
  if (j == 0) {
    a = scratch[int64_t(k) * W + i];
    b = 0.0f;
    c = 0.0f;
  }
  if (j == H - 1) {
    a = scratch[DW + int64_t(k) * W + i];
    b = 0.0f;
    c = 0.0f;
  }
  if (k == 0) {
    a = 0.0f;
    b = 0.0f;
    c = 0.0f;
  }
  if (k == D - 1) {
    a = scratch[2 * DW + int64_t(j) * W + i];
    b = scratch[2 * DW + HW + int64_t(j) * W + i];
    c = 0.0f;
  }
  if (sigma) {  // the bluff body after the fill
    const float s = sigma[n];
    a = bluff_apply(a, s, dt);
    b = bluff_apply(b, s, dt);
    c = bluff_apply(c, s, dt);
  }
  u[n] = a;
  v[n] = b;
  w[n] = c;
}

// The bluff body at the cells the fill copies from that are not boundary
// cells themselves, from their gathered (undamped) values, each once: u at
// j = 1 and j = H-2 for 1 <= k <= D-3, then u and v on the plane k = D-2.
__global__ __launch_bounds__(256) void k_bc_velocity_src_bluff(float* u, float* v,
                                                               const float* __restrict__ scratch,
                                                               const float* __restrict__ sigma,
                                                               float dt, Grid3 g) {
  const int D = g.D, H = g.H, W = g.W;
  const int nrow = H - 2 > 1 ? 2 : 1;  // rows j = 1 and j = H-2 (one row when H = 3)
  const int64_t ny = int64_t(nrow) * (D - 3) * (W - 2), nz = int64_t(H - 2) * (W - 2);
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t DW = int64_t(D) * W, HW = int64_t(H) * W;
  if (t < ny) {
    const int r = int(t % nrow);
    t /= nrow;
    const int k = 1 + int(t / (W - 2)), i = 1 + int(t % (W - 2));
    const int j = r ? H - 2 : 1;
    const int64_t n = g.idx(k, j, i);
    u[n] = bluff_apply(scratch[(r ? DW : 0) + int64_t(k) * W + i], sigma[n], dt);
  } else if ((t -= ny) < nz) {
    const int j = 1 + int(t / (W - 2)), i = 1 + int(t % (W - 2));
    const int64_t n = g.idx(D - 2, j, i);
    u[n] = bluff_apply(scratch[2 * DW + int64_t(j) * W + i], sigma[n], dt);
    v[n] = bluff_apply(scratch[2 * DW + HW + int64_t(j) * W + i], sigma[n], dt);
  }
}

// bluff_body on NF (1 to 3) fields in one pass: sigma is read once per cell
// for all of them (an output may be its input)
struct BluffFields {
  const float* in[3];
  float* out[3];
};
template <int NF>
__global__ __launch_bounds__(256) void k_bluff(BluffFields f, const float* __restrict__ sigma,
                                               float dt, int64_t n) {
  const int64_t q = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (q >= n) return;
  const float s = sigma[q];
  float v[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) v[k] = f.in[k][q];
#pragma unroll
  for (int k = 0; k < NF; ++k) f.out[k][q] = bluff_apply(v[k], s, dt);
}

}  // namespace hfm

// ====================================================================== C ABI
using namespace hfm;

namespace {

// a plane of fewer than 2^30 cells: the kernels address a cell as its plane's
// base plus a 32-bit byte offset
bool dims_ok(int64_t D, int64_t H, int64_t W) {
  return D > 0 && H > 0 && W > 0 && D <= 65535 && D * H * W <= INT32_MAX &&
         H * W < (int64_t(1) << 30);
}

dim3 grid_for(const Grid3& g) {
  return dim3(unsigned((g.W + kBX - 1) / kBX), unsigned((g.H + kBY - 1) / kBY), unsigned(g.D));
}

// Tap sets of a three-kernel walker: the x / y / z first-derivative sets in
// that order, all 7-point, or all 27 (any other mix).
enum { kSetXYZ, kSet7, kSet27 };
int tapset3(const uint32_t m[3]) {
  const int a = taps_of(m[0]), b = taps_of(m[1]), c = taps_of(m[2]);
  if (a == kTapsX && b == kTapsY && c == kTapsZ) return kSetXYZ;
  if (a == 7 && b == 7 && c == 7) return kSet7;
  return kSet27;
}

bool make_weights(const float* K, Weights& w, uint32_t& mask) {
  if (!K) return false;
  mask = 0;
  for (int t = 0; t < 27; ++t) {
    w.w[t] = K[t];
    if (K[t] != 0.0f) mask |= 1u << t;
  }
  return true;
}

bool make_var6(const float* const m[6], const float* tot, Var6& V, bool need_tot) {
  if (!m) return false;
  for (int q = 0; q < 6; ++q) {
    if (!m[q]) return false;
    V.m[q] = m[q];
  }
  V.tot = tot;
  return !(need_tot && !tot);
}

}  // namespace

extern "C" const char* hfm_last_error(void) { return g_err; }

extern "C" int hfm_version(void) { return 1; }

extern "C" int hfm_struct_sizes(int64_t* task_size, int64_t* scalars_size) {
  if (!task_size || !scalars_size) {
    set_error("hfm_struct_sizes: null output");
    return HFM_EINVAL;
  }
  *task_size = int64_t(sizeof(hfm_mg_task));
  *scalars_size = int64_t(sizeof(hfm_mg_scalars));
  return HFM_OK;
}

extern "C" int hfm_conv27_f32(const float* x, float* y, const float* b, const float K[27],
                              float d, int64_t D, int64_t H, int64_t W, int form, void* stream) {
  Weights w;
  uint32_t mask;
  const int fm = form & HFM_FORM_MASK;
  if (!x || !y || !dims_ok(D, H, W) || !make_weights(K, w, mask) || x == y ||
      fm < HFM_FORM_APPLY || fm > HFM_FORM_JACOBI_MUL ||
      (form & ~(HFM_FORM_MASK | HFM_FORM_ZERO_OUTLET)) ||
      (fm != HFM_FORM_APPLY && !b)) {
    set_error("hfm_conv27_f32: bad arguments");
    return HFM_EINVAL;
  }
  const Grid3 g{int(D), int(H), int(W)};
  auto go = [&](auto taps) {
    hipLaunchKernelGGL(k_conv27<decltype(taps)::value>, grid_col<2>(g), dim3(kBX, kBY), 0,
                       as_stream(stream), x, y, b, w, d, g, form);
  };
  switch (taps_of(mask)) {
    case 7: go(std::integral_constant<int, 7>{}); break;
    case kTapsX: go(std::integral_constant<int, kTapsX>{}); break;
    case kTapsY: go(std::integral_constant<int, kTapsY>{}); break;
    case kTapsZ: go(std::integral_constant<int, kTapsZ>{}); break;
    default: go(std::integral_constant<int, 27>{}); break;
  }
  return launch_status("hfm_conv27_f32");
}

extern "C" int hfm_var6_f32(const float* x, float* y, const float* b, const float* const m[6],
                            const float* tot, int64_t D, int64_t H, int64_t W, int form,
                            void* stream) {
  Var6 V;
  if (!x || !y || x == y || !dims_ok(D, H, W) ||
      !make_var6(m, tot, V, form == HFM_FORM_JACOBI) || form < HFM_FORM_APPLY ||
      form > HFM_FORM_JACOBI || (form != HFM_FORM_APPLY && !b)) {
    set_error("hfm_var6_f32: bad arguments");
    return HFM_EINVAL;
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipLaunchKernelGGL(k_var6, grid_for(g), dim3(kBX, kBY), 0, as_stream(stream), x, y, b, V, g,
                     form);
  return launch_status("hfm_var6_f32");
}

static bool factors_ok(int fz, int fy, int fx) {
  return (fz == 1 || fz == 2) && (fy == 1 || fy == 2) && (fx == 1 || fx == 2);
}

extern "C" int hfm_restrict_mean_f32(const float* x, float* y, int64_t D, int64_t H, int64_t W,
                                     int fz, int fy, int fx, void* stream) {
  if (!x || !y || !dims_ok(D, H, W) || !factors_ok(fz, fy, fx) || D % fz || H % fy || W % fx) {
    set_error("hfm_restrict_mean_f32: bad arguments (D,H,W)=(%lld,%lld,%lld) f=(%d,%d,%d)",
              (long long)D, (long long)H, (long long)W, fz, fy, fx);
    return HFM_EINVAL;
  }
  const Grid3 gf{int(D), int(H), int(W)};
  const Grid3 gc{int(D / fz), int(H / fy), int(W / fx)};
  hipLaunchKernelGGL(k_restrict_mean, grid_for(gc), dim3(kBX, kBY), 0, as_stream(stream), x, y,
                     gf, gc, fz, fy, fx, 1.0f / float(fz * fy * fx));
  return launch_status("hfm_restrict_mean_f32");
}

extern "C" int hfm_resid_restrict_f32(const float* x, const float* b, float* rc,
                                      const float K[27], int64_t D, int64_t H, int64_t W,
                                      int flags, void* stream) {
  Weights w;
  uint32_t mask;
  if (!x || !b || !rc || rc == x || rc == b || !dims_ok(D, H, W) || !make_weights(K, w, mask) ||
      D % 2 || H % 2 || W % 2 || (flags & ~HFM_FORM_ZERO_OUTLET)) {
    set_error("hfm_resid_restrict_f32: bad arguments (D,H,W)=(%lld,%lld,%lld) flags=%d",
              (long long)D, (long long)H, (long long)W, flags);
    return HFM_EINVAL;
  }
  const Grid3 g{int(D), int(H), int(W)};
  auto go = [&](auto taps) {
    hipLaunchKernelGGL(k_resid_restrict27<decltype(taps)::value>, grid_col<2>(g),
                       dim3(kBX, kBY), 0, as_stream(stream), x, b, rc, w, g, flags);
  };
  if (taps_of(mask) == 7)
    go(std::integral_constant<int, 7>{});
  else
    go(std::integral_constant<int, 27>{});
  return launch_status("hfm_resid_restrict_f32");
}

extern "C" int hfm_prolong_nearest_f32(const float* xc, float* y, int64_t Dc, int64_t Hc,
                                       int64_t Wc, int fz, int fy, int fx, void* stream) {
  if (!xc || !y || !factors_ok(fz, fy, fx) || !dims_ok(Dc * fz, Hc * fy, Wc * fx)) {
    set_error("hfm_prolong_nearest_f32: bad arguments");
    return HFM_EINVAL;
  }
  const Grid3 gc{int(Dc), int(Hc), int(Wc)};
  const Grid3 gf{int(Dc * fz), int(Hc * fy), int(Wc * fx)};
  hipLaunchKernelGGL(k_prolong_nearest, grid_for(gf), dim3(kBX, kBY), 0, as_stream(stream), xc,
                     y, gf, gc, fz, fy, fx);
  return launch_status("hfm_prolong_nearest_f32");
}

extern "C" int hfm_prolong_upsum_f32(const float* xc, float* y, int64_t Dc, int64_t Hc,
                                     int64_t Wc, int fz, int fy, int fx, void* stream) {
  if (!xc || !y || xc == y || !factors_ok(fz, fy, fx) || !dims_ok(Dc * fz, Hc * fy, Wc * fx)) {
    set_error("hfm_prolong_upsum_f32: bad arguments");
    return HFM_EINVAL;
  }
  const Grid3 gc{int(Dc), int(Hc), int(Wc)};
  const Grid3 gf{int(Dc * fz), int(Hc * fy), int(Wc * fx)};
  hipLaunchKernelGGL(k_prolong_upsum, grid_for(gf), dim3(kBX, kBY), 0, as_stream(stream), xc, y,
                     gf, gc, fz, fy, fx);
  return launch_status("hfm_prolong_upsum_f32");
}

extern "C" int hfm_sum_f64(const float* x, int64_t n, double* out, void* stream) {
  if (!x || !out || n < 0) {
    set_error("hfm_sum_f64: bad arguments");
    return HFM_EINVAL;
  }
  hipLaunchKernelGGL(k_sum_f64, dim3(1), dim3(1024), 0, as_stream(stream), x, n, out);
  return launch_status("hfm_sum_f64");
}

extern "C" int hfm_prolong_jacobi_f32(const float* wc, const float* r, float* y,
                                      const float K[27], float d, const float* const m[6],
                                      const float* tot, int64_t D, int64_t H, int64_t W, int fz,
                                      int fy, int fx, void* stream) {
  Weights w{};
  uint32_t mask = 0;
  Var6 V{};
  const bool var6 = m != nullptr;
  if (!wc || !r || !y || !dims_ok(D, H, W) || !factors_ok(fz, fy, fx) || D % fz || H % fy ||
      W % fx || (var6 ? !make_var6(m, tot, V, true) : !make_weights(K, w, mask))) {
    set_error("hfm_prolong_jacobi_f32: bad arguments");
    return HFM_EINVAL;
  }
  const Grid3 gf{int(D), int(H), int(W)};
  const Grid3 gc{int(D / fz), int(H / fy), int(W / fx)};
  hipStream_t s = as_stream(stream);
  if (var6)
    hipLaunchKernelGGL((k_prolong_jacobi<true, 7>), grid_for(gf), dim3(kBX, kBY), 0, s, wc, r, y,
                       w, d, V, gf, gc, fz, fy, fx);
  else if (taps_of(mask) == 7)
    hipLaunchKernelGGL((k_prolong_jacobi_col<7>), grid_col<2>(gf), dim3(kBX, kBY), 0, s, wc, r, y,
                       w, d, gf, gc, fz, fy, fx);
  else
    hipLaunchKernelGGL((k_prolong_jacobi_col<27>), grid_col<2>(gf), dim3(kBX, kBY), 0, s, wc, r, y,
                       w, d, gf, gc, fz, fy, fx);
  return launch_status("hfm_prolong_jacobi_f32");
}

namespace {
int correct_jacobi(const char* what, const float* x, const CorrW& cw, bool prol, const float* b,
                   float* y, const float K[27], float d, const float* const m[6],
                   const float* tot, int64_t D, int64_t H, int64_t W, int flags, void* stream) {
  Weights w{};
  uint32_t mask = 0;
  Var6 V{};
  const bool var6 = m != nullptr;
  if (!x || !cw.w || !b || !y || y == x || y == cw.w || !dims_ok(D, H, W) ||
      (flags & ~(HFM_CORRECT_ZERO_OUTLET | HFM_CORRECT_RECIPROCAL)) ||
      (var6 && (flags & HFM_CORRECT_RECIPROCAL)) ||
      (var6 ? !make_var6(m, tot, V, true) : !make_weights(K, w, mask))) {
    set_error("%s: bad arguments", what);
    return HFM_EINVAL;
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipStream_t s = as_stream(stream);
  if (prol) {
    if (var6)
      hipLaunchKernelGGL((k_correct_jacobi<true, 7, true>), grid_for(g), dim3(kBX, kBY), 0, s, x,
                         cw, b, y, w, d, V, g, flags);
    else if (taps_of(mask) == 7)
      hipLaunchKernelGGL((k_correct_jacobi_col<7, true>), grid_col<2>(g), dim3(kBX, kBY), 0, s, x,
                         cw, b, y, w, d, g, flags);
    else
      hipLaunchKernelGGL((k_correct_jacobi_col<27, true>), grid_col<2>(g), dim3(kBX, kBY), 0, s,
                         x, cw, b, y, w, d, g, flags);
  } else {
    if (var6)
      hipLaunchKernelGGL((k_correct_jacobi<true, 7, false>), grid_for(g), dim3(kBX, kBY), 0, s, x,
                         cw, b, y, w, d, V, g, flags);
    else if (taps_of(mask) == 7)
      hipLaunchKernelGGL((k_correct_jacobi_col<7, false>), grid_col<2>(g), dim3(kBX, kBY), 0, s,
                         x, cw, b, y, w, d, g, flags);
    else
      hipLaunchKernelGGL((k_correct_jacobi_col<27, false>), grid_col<2>(g), dim3(kBX, kBY), 0, s,
                         x, cw, b, y, w, d, g, flags);
  }
  return launch_status(what);
}
}  // namespace

extern "C" int hfm_correct_jacobi_f32(const float* x, const float* wv, const float* b, float* y,
                                      const float K[27], float d, const float* const m[6],
                                      const float* tot, int64_t D, int64_t H, int64_t W, int flags,
                                      void* stream) {
  CorrW cw{wv, Grid3{int(D), int(H), int(W)}, 0, 0, 0};
  return correct_jacobi("hfm_correct_jacobi_f32", x, cw, false, b, y, K, d, m, tot, D, H, W,
                        flags, stream);
}

extern "C" int hfm_prolong_correct_jacobi_f32(const float* x, const float* wc, const float* b,
                                              float* y, const float K[27], float d,
                                              const float* const m[6], const float* tot,
                                              int64_t D, int64_t H, int64_t W, int fz, int fy,
                                              int fx, int flags, void* stream) {
  if (!factors_ok(fz, fy, fx) || !dims_ok(D, H, W) || D % fz || H % fy || W % fx) {
    set_error("hfm_prolong_correct_jacobi_f32: bad arguments (D,H,W)=(%lld,%lld,%lld) f=(%d,%d,%d)",
              (long long)D, (long long)H, (long long)W, fz, fy, fx);
    return HFM_EINVAL;
  }
  CorrW cw{wc, Grid3{int(D / fz), int(H / fy), int(W / fx)}, fz - 1, fy - 1, fx - 1};
  return correct_jacobi("hfm_prolong_correct_jacobi_f32", x, cw, true, b, y, K, d, m, tot, D, H,
                        W, flags, stream);
}

// k_advect27_v with the first-derivative tap sets when the three advection
// stencils hold them (mask[1..3] of K), else over all 27 taps
template <int NF, bool DIS, int SH, typename... A>
void launch_advect27_sh(const Kern4& K, const Grid3& g, hipStream_t s, A... args) {
  if (tapset3(K.mask + 1) == kSetXYZ)
    hipLaunchKernelGGL((k_advect27_v<NF, DIS, SH, kTapsX, kTapsY, kTapsZ>), grid_col<1>(g),
                       dim3(kBX, kBY), 0, s, args...);
  else
    hipLaunchKernelGGL((k_advect27_v<NF, DIS, SH>), grid_col<1>(g), dim3(kBX, kBY), 0, s,
                       args...);
}
// sh: 1 when the three velocities are the three advected fields, 2 when the
// bases are too (three fields only)
template <int NF, bool DIS, typename... A>
void launch_advect27(int sh, const Kern4& K, const Grid3& g, hipStream_t s, A... args) {
  if (NF == 3 && sh == 2)
    launch_advect27_sh<NF, DIS, NF == 3 ? 2 : 0>(K, g, s, args...);
  else if (NF == 3 && sh == 1)
    launch_advect27_sh<NF, DIS, NF == 3 ? 1 : 0>(K, g, s, args...);
  else
    launch_advect27_sh<NF, DIS, 0>(K, g, s, args...);
}
// the shared-velocity form of a launch (see k_advect27_v's SH)
inline int advect_sharing(int nf, const float* const vel[3], const float* const* phi,
                          const float* const* base) {
  if (nf != 3) return 0;
  for (int q = 0; q < 3; ++q)
    if (vel[q] != phi[q]) return 0;
  for (int q = 0; q < 3; ++q)
    if (base[q] != phi[q]) return 1;
  return 2;
}

static int advect27(const char* what, const float* const vel[3], const float* const* phi,
                    const float* const* base, float* const* out, int nf, const float K[4][27],
                    float f, const float* sigma, float dt, int post, int64_t D, int64_t H,
                    int64_t W, void* stream) {
  Kern4 kk;
  if (!vel || !vel[0] || !vel[1] || !vel[2] || !phi || !base || !out || !K ||
      (nf != 1 && nf != 3) || !dims_ok(D, H, W) ||
      (post & ~(HFM_ADV_INDICATOR | (3 << HFM_ADV_BLUFF_SHIFT))) ||
      (((post >> HFM_ADV_BLUFF_SHIFT) & 3) && !sigma)) {
    set_error("%s: bad arguments", what);
    return HFM_EINVAL;
  }
  for (int q = 0; q < 4; ++q) make_weights(K[q], kk.k[q], kk.mask[q]);
  struct P3 {
    const float* a[3];
  } pp{}, bb{};
  struct O3 {
    float* a[3];
  } oo{};
  for (int q = 0; q < nf; ++q) {
    if (!phi[q] || !base[q] || !out[q]) {
      set_error("%s: null field %d", what, q);
      return HFM_EINVAL;
    }
    for (int r = 0; r < 3; ++r)
      if (out[q] == vel[r] || out[q] == phi[r < nf ? r : 0] || (sigma && out[q] == sigma)) {
        set_error("%s: output %d aliases an input", what, q);
        return HFM_EINVAL;
      }
    pp.a[q] = phi[q];
    bb.a[q] = base[q];
    oo.a[q] = out[q];
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipStream_t s = as_stream(stream);
  const float* const none = nullptr;
  const int sh = advect_sharing(nf, vel, phi, base);
  if (nf == 3)
    launch_advect27<3, false>(sh, kk, g, s, vel[0], vel[1], vel[2], pp.a[0], pp.a[1], pp.a[2],
                              bb.a[0], bb.a[1], bb.a[2], oo.a[0], oo.a[1], oo.a[2], kk, f, g,
                              none, none, none, sigma, dt, post);
  else
    launch_advect27<1, false>(0, kk, g, s, vel[0], vel[1], vel[2], pp.a[0], pp.a[1], pp.a[2],
                              bb.a[0], bb.a[1], bb.a[2], oo.a[0], oo.a[1], oo.a[2], kk, f, g,
                              none, none, none, sigma, dt, post);
  return launch_status(what);
}

extern "C" int hfm_advect27_f32(const float* const vel[3], const float* const* phi,
                                const float* const* base, float* const* out, int nf,
                                const float K[4][27], float f, int64_t D, int64_t H, int64_t W,
                                void* stream) {
  return advect27("hfm_advect27_f32", vel, phi, base, out, nf, K, f, nullptr, 0.0f, 0, D, H, W,
                  stream);
}

extern "C" int hfm_advect27_post_f32(const float* const vel[3], const float* const* phi,
                                     const float* const* base, float* const* out, int nf,
                                     const float K[4][27], float f, const float* sigma, float dt,
                                     int post, int64_t D, int64_t H, int64_t W, void* stream) {
  return advect27("hfm_advect27_post_f32", vel, phi, base, out, nf, K, f, sigma, dt, post, D, H,
                  W, stream);
}

extern "C" int hfm_advect27_dis_f32(const float* const vel[3], const float* const* phi,
                                    const float* const* base, const float* const* dis,
                                    float* const* out, int nf, const float K[3][27], float f,
                                    int64_t D, int64_t H, int64_t W, void* stream) {
  Kern4 kk{};
  if (!vel || !vel[0] || !vel[1] || !vel[2] || !phi || !base || !dis || !out || !K ||
      (nf != 1 && nf != 3) || !dims_ok(D, H, W)) {
    set_error("hfm_advect27_dis_f32: bad arguments");
    return HFM_EINVAL;
  }
  for (int q = 0; q < 3; ++q) make_weights(K[q], kk.k[q + 1], kk.mask[q + 1]);
  const float* pp[3] = {};
  const float* bb[3] = {};
  const float* dd[3] = {};
  float* oo[3] = {};
  for (int q = 0; q < nf; ++q) {
    if (!phi[q] || !base[q] || !dis[q] || !out[q]) {
      set_error("hfm_advect27_dis_f32: null field %d", q);
      return HFM_EINVAL;
    }
    for (int r = 0; r < 3; ++r)
      if (out[q] == vel[r] || out[q] == phi[r < nf ? r : 0] || out[q] == dis[r < nf ? r : 0]) {
        set_error("hfm_advect27_dis_f32: output %d aliases an input", q);
        return HFM_EINVAL;
      }
    pp[q] = phi[q];
    bb[q] = base[q];
    dd[q] = dis[q];
    oo[q] = out[q];
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipStream_t s = as_stream(stream);
  const float* const none = nullptr;
  const int sh = advect_sharing(nf, vel, phi, base);
  if (nf == 3)
    launch_advect27<3, true>(sh, kk, g, s, vel[0], vel[1], vel[2], pp[0], pp[1], pp[2], bb[0],
                             bb[1], bb[2], oo[0], oo[1], oo[2], kk, f, g, dd[0], dd[1], dd[2],
                             none, 0.0f, 0);
  else
    launch_advect27<1, true>(0, kk, g, s, vel[0], vel[1], vel[2], pp[0], pp[1], pp[2], bb[0],
                             bb[1], bb[2], oo[0], oo[1], oo[2], kk, f, g, dd[0], dd[1], dd[2],
                             none, 0.0f, 0);
  return launch_status("hfm_advect27_dis_f32");
}

extern "C" int hfm_pg_coef_f32(const float* const vt[3], const float* sigma, float* const kout[3],
                               const float K[4][27], float dt, float eps, int64_t D, int64_t H,
                               int64_t W, void* stream) {
  Kern4 kk{};
  if (!vt || !sigma || !kout || !K || !dims_ok(D, H, W)) {
    set_error("hfm_pg_coef_f32: bad arguments");
    return HFM_EINVAL;
  }
  for (int q = 0; q < 3; ++q)
    if (!vt[q] || !kout[q] || kout[q] == sigma) {
      set_error("hfm_pg_coef_f32: null or aliased field %d", q);
      return HFM_EINVAL;
    }
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r)
      if (kout[q] == vt[r]) {
        set_error("hfm_pg_coef_f32: output %d aliases an input", q);
        return HFM_EINVAL;
      }
  for (int q = 0; q < 4; ++q) make_weights(K[q], kk.k[q], kk.mask[q]);
  const Grid3 g{int(D), int(H), int(W)};
  const uint32_t adv[3] = {kk.mask[1], kk.mask[2], kk.mask[3]};
  if (tapset3(adv) == kSetXYZ)
    hipLaunchKernelGGL((k_pg_coef<kTapsX, kTapsY, kTapsZ>), grid_col<1>(g), dim3(kBX, kBY), 0,
                       as_stream(stream), vt[0], vt[1], vt[2], sigma, kout[0], kout[1], kout[2], kk,
                       dt, eps, g);
  else
    hipLaunchKernelGGL((k_pg_coef<27, 27, 27>), grid_col<1>(g), dim3(kBX, kBY), 0,
                       as_stream(stream), vt[0], vt[1], vt[2], sigma, kout[0], kout[1], kout[2], kk,
                       dt, eps, g);
  return launch_status("hfm_pg_coef_f32");
}

extern "C" int hfm_pg_flux_f32(const float* const vt[3], const float* const kin[3],
                               float* const out[3], const float Kdif[27], int64_t D, int64_t H,
                               int64_t W, void* stream) {
  Weights w{};
  uint32_t mask = 0;
  if (!vt || !kin || !out || !Kdif || !dims_ok(D, H, W) || !make_weights(Kdif, w, mask)) {
    set_error("hfm_pg_flux_f32: bad arguments");
    return HFM_EINVAL;
  }
  for (int q = 0; q < 3; ++q) {
    if (!vt[q] || !kin[q] || !out[q]) {
      set_error("hfm_pg_flux_f32: null field %d", q);
      return HFM_EINVAL;
    }
    for (int r = 0; r < 3; ++r)
      if (out[q] == vt[r] || out[q] == kin[r]) {
        set_error("hfm_pg_flux_f32: output %d aliases an input", q);
        return HFM_EINVAL;
      }
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipLaunchKernelGGL(k_pg_flux, grid_col<1>(g), dim3(kBX, kBY), 0, as_stream(stream), vt[0], vt[1],
                     vt[2], kin[0], kin[1], kin[2], out[0], out[1], out[2], w, g);
  return launch_status("hfm_pg_flux_f32");
}

extern "C" int hfm_conv27_sum_f32(const float* const* x, float* y, int nf, const float K[3][27],
                                  float scale, int64_t D, int64_t H, int64_t W, int flags,
                                  void* stream) {
  Kern3 kk;
  if (!x || !y || !K || nf < 1 || nf > 3 || !dims_ok(D, H, W) || (flags & ~HFM_SUM_SOURCE_BC)) {
    set_error("hfm_conv27_sum_f32: bad arguments");
    return HFM_EINVAL;
  }
  const float* xs[3] = {nullptr, nullptr, nullptr};
  for (int q = 0; q < nf; ++q) {
    if (!x[q] || x[q] == y) {
      set_error("hfm_conv27_sum_f32: bad field %d", q);
      return HFM_EINVAL;
    }
    xs[q] = x[q];
    make_weights(K[q], kk.k[q], kk.mask[q]);
  }
  for (int q = nf; q < 3; ++q) {
    kk.mask[q] = 0;
    xs[q] = xs[0];
  }
  const Grid3 g{int(D), int(H), int(W)};
  constexpr int kRY = 2;  // rows per thread of the three-field divergence
  if (nf == 3 && tapset3(kk.mask) == kSetXYZ)
    hipLaunchKernelGGL((k_conv27_sum_v<3, kTapsX, kTapsY, kTapsZ, kRY>), grid_col<kRY>(g),
                       dim3(kBX, kBY), 0, as_stream(stream), xs[0], xs[1], xs[2], y, kk, scale, g,
                       flags);
  else if (nf == 3)
    hipLaunchKernelGGL(k_conv27_sum_v<3>, grid_col<1>(g), dim3(kBX, kBY), 0, as_stream(stream),
                       xs[0], xs[1], xs[2], y, kk, scale, g, flags);
  else if (nf == 2)
    hipLaunchKernelGGL(k_conv27_sum_v<2>, grid_col<1>(g), dim3(kBX, kBY), 0, as_stream(stream),
                       xs[0], xs[1], xs[2], y, kk, scale, g, flags);
  else
    hipLaunchKernelGGL(k_conv27_sum_v<1>, grid_col<1>(g), dim3(kBX, kBY), 0, as_stream(stream),
                       xs[0], xs[1], xs[2], y, kk, scale, g, flags);
  return launch_status("hfm_conv27_sum_f32");
}

extern "C" int hfm_grad3_f32(const float* p, const float* const f[3], float* const o[3],
                             float* pout, const float K[3][27], const float* sigma, float dt,
                             int flags, int64_t D, int64_t H, int64_t W, void* stream) {
  const char* what = "hfm_grad3_f32";
  Kern3 kk;
  const int known = HFM_GRAD_P_BC | HFM_GRAD_PRE_BLUFF | HFM_GRAD_POST_BLUFF;
  if (!p || !f || !o || !K || !dims_ok(D, H, W) || (flags & ~known) ||
      ((flags & (HFM_GRAD_PRE_BLUFF | HFM_GRAD_POST_BLUFF)) && !sigma) ||
      ((flags & HFM_GRAD_P_BC) && (D < 2 || H < 2 || W < 2)) || pout == p) {
    set_error("%s: bad arguments", what);
    return HFM_EINVAL;
  }
  for (int q = 0; q < 3; ++q) {
    if (!f[q] || !o[q] || o[q] == p || o[q] == pout || (sigma && o[q] == sigma)) {
      set_error("%s: null or aliased field %d", what, q);
      return HFM_EINVAL;
    }
    if (!make_weights(K[q], kk.k[q], kk.mask[q])) {
      set_error("%s: bad weights %d", what, q);
      return HFM_EINVAL;
    }
  }
  // each output may overwrite its own f (a cell reads only its own f value)
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r)
      if (r != q && o[q] == f[r]) {
        set_error("%s: output %d aliases field %d", what, q, r);
        return HFM_EINVAL;
      }
  const Grid3 g{int(D), int(H), int(W)};
  switch (tapset3(kk.mask)) {
    case kSetXYZ:
      hipLaunchKernelGGL((k_grad3<kTapsX, kTapsY, kTapsZ>), grid_col<2>(g), dim3(kBX, kBY), 0,
                         as_stream(stream), p, f[0], f[1], f[2], o[0], o[1], o[2], pout, kk, sigma,
                         dt, g, flags);
      break;
    case kSet7:
      hipLaunchKernelGGL((k_grad3<7, 7, 7>), grid_col<2>(g), dim3(kBX, kBY), 0,
                         as_stream(stream), p, f[0], f[1], f[2], o[0], o[1], o[2], pout, kk, sigma,
                         dt, g, flags);
      break;
    default:
      hipLaunchKernelGGL((k_grad3<27, 27, 27>), grid_col<2>(g), dim3(kBX, kBY), 0,
                         as_stream(stream), p, f[0], f[1], f[2], o[0], o[1], o[2], pout, kk, sigma,
                         dt, g, flags);
  }
  return launch_status(what);
}

extern "C" int hfm_bc_f32(const float* in, float* out, int64_t D, int64_t H, int64_t W, int kind,
                          void* stream) {
  const bool needs2 = kind == HFM_BC_TOP_COPY || kind == HFM_BC_PRESSURE || kind == HFM_BC_SOURCE;
  if (!in || !out || in == out || !dims_ok(D, H, W) || kind < HFM_BC_TOP_COPY ||
      kind > HFM_BC_ZERO_OUTLET ||
      (needs2 && ((kind != HFM_BC_SOURCE && D < 2) || (kind == HFM_BC_PRESSURE && H < 2) ||
                  (kind != HFM_BC_TOP_COPY && W < 2)))) {
    set_error("hfm_bc_f32: bad arguments (kind %d)", kind);
    return HFM_EINVAL;
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipLaunchKernelGGL(k_bc, grid_for(g), dim3(kBX, kBY), 0, as_stream(stream), in, out, g, kind);
  return launch_status("hfm_bc_f32");
}

// The no-allocation fill's scratch: one copy of this array per device, held
// until the process ends.  Host code never passes the symbol itself to a
// kernel; it passes the device address hipGetSymbolAddress returns, looked
// up once per device (with that device current) and kept per device.
static __device__ __align__(256) float scratch[HFM_BC_NOALLOC_SCRATCH_FLOATS];

namespace {
int bc_velocity_inplace(const char* what, float* u, float* v, float* w, float* scratch, int64_t D,
                        int64_t H, int64_t W, float ub, const float* inflow, void* stream,
                        const float* sigma = nullptr, float dt = 0.0f) {
  const Grid3 g{int(D), int(H), int(W)};
  hipStream_t s = as_stream(stream);
  const int64_t ng = std::max(D * W, H * W);
  hipLaunchKernelGGL(k_bc_velocity_gather, dim3(unsigned((ng + 255) / 256)), dim3(256), 0, s, u, v,
                     scratch, g);
  const int64_t nb = 2 * D * H + 2 * D * (W - 2) + 2 * (H - 2) * (W - 2);
  hipLaunchKernelGGL(k_bc_velocity_inplace, dim3(unsigned((nb + 255) / 256)), dim3(256), 0, s, u,
                     v, w, scratch, g, ub, inflow, sigma, dt);
  if (sigma && D >= 3 && H >= 3 && W >= 3) {
    const int64_t ns = (H - 2 > 1 ? 2 : 1) * (D - 3) * (W - 2) + (H - 2) * (W - 2);
    hipLaunchKernelGGL(k_bc_velocity_src_bluff, dim3(unsigned((ns + 255) / 256)), dim3(256), 0, s,
                       u, v, scratch, sigma, dt, g);
  }
  return launch_status(what);
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_dev[kMaxDevices] = {};

int scratch_of_current_device(const char* what, float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("%s: no usable current device (%s)", what, hipGetErrorString(e));
    return HFM_EDEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error("%s: hipGetSymbolAddress of the scratch array failed on device %d: %s", what,
                dev, hipGetErrorString(e));
      return HFM_EDEVICE;
    }
    g_scratch_dev[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return HFM_OK;
}
}  // namespace

extern "C" int hfm_bc_velocity_inplace_f32(float* u, float* v, float* w, float* scratch,
                                           int64_t D, int64_t H, int64_t W, float ub,
                                           const float* inflow, void* stream) {
  if (!u || !v || !w || !scratch || !dims_ok(D, H, W) || D < 2 || H < 2 || W < 2) {
    set_error("hfm_bc_velocity_inplace_f32: bad arguments");
    return HFM_EINVAL;
  }
  return bc_velocity_inplace("hfm_bc_velocity_inplace_f32", u, v, w, scratch, D, H, W, ub, inflow,
                             stream);
}

extern "C" int hfm_bc_velocity_bluff_inplace_f32(float* u, float* v, float* w, float* scratch,
                                                 int64_t D, int64_t H, int64_t W, float ub,
                                                 const float* inflow, const float* sigma,
                                                 float dt, void* stream) {
  if (!u || !v || !w || !scratch || !sigma || !dims_ok(D, H, W) || D < 2 || H < 2 || W < 2) {
    set_error("hfm_bc_velocity_bluff_inplace_f32: bad arguments");
    return HFM_EINVAL;
  }
  return bc_velocity_inplace("hfm_bc_velocity_bluff_inplace_f32", u, v, w, scratch, D, H, W, ub,
                             inflow, stream, sigma, dt);
}

extern "C" int hfm_bc_velocity_inplace_noalloc_f32(float* u, float* v, float* w, int64_t D,
                                                   int64_t H, int64_t W, float ub,
                                                   const float* inflow, void* stream) {
  const char* what = "hfm_bc_velocity_inplace_noalloc_f32";
  if (!u || !v || !w || !dims_ok(D, H, W) || D < 2 || H < 2 || W < 2) {
    set_error("%s: bad arguments", what);
    return HFM_EINVAL;
  }
  const int64_t need = 2 * D * W + 2 * H * W;
  if (need > HFM_BC_NOALLOC_SCRATCH_FLOATS) {
    set_error("%s: a (%lld, %lld, %lld) grid needs %lld scratch floats, more than the %lld this "
              "entry point holds; use hfm_bc_velocity_inplace_f32 with a workspace",
              what, (long long)D, (long long)H, (long long)W, (long long)need,
              (long long)HFM_BC_NOALLOC_SCRATCH_FLOATS);
    return HFM_EINVAL;
  }
  float* ptr = nullptr;
  const int rc = scratch_of_current_device(what, &ptr);
  if (rc != HFM_OK) return rc;
  return bc_velocity_inplace(what, u, v, w, ptr, D, H, W, ub, inflow, stream);
}

extern "C" int hfm_bc_velocity_f32(const float* u, const float* v, const float* w, float* uo,
                                   float* vo, float* wo, int64_t D, int64_t H, int64_t W,
                                   float ub, const float* inflow, void* stream) {
  if (!u || !v || !w || !uo || !vo || !wo || uo == u || vo == v || wo == w ||
      !dims_ok(D, H, W) || D < 2 || H < 2 || W < 2) {
    set_error("hfm_bc_velocity_f32: bad arguments");
    return HFM_EINVAL;
  }
  const Grid3 g{int(D), int(H), int(W)};
  hipLaunchKernelGGL(k_bc_velocity, grid_for(g), dim3(kBX, kBY), 0, as_stream(stream), u, v, w,
                     uo, vo, wo, g, ub, inflow);
  return launch_status("hfm_bc_velocity_f32");
}

extern "C" int hfm_bluff_body_f32(const float* const* in, float* const* out, int nfields,
                                  const float* sigma, float dt, int64_t ncells, void* stream) {
  if (!in || !out || !sigma || nfields < 0 || ncells < 0) {
    set_error("hfm_bluff_body_f32: bad arguments");
    return HFM_EINVAL;
  }
  for (int f = 0; f < nfields; ++f) {
    if (!in[f] || !out[f]) {
      set_error("hfm_bluff_body_f32: null field %d", f);
      return HFM_EINVAL;
    }
  }
  // up to three fields per launch, which reads every input of a cell before
  // writing its outputs; an output that is another field's input keeps the
  // field-by-field order of one launch per field
  bool cross = false;
  for (int a = 0; a < nfields; ++a)
    for (int b2 = 0; b2 < nfields; ++b2) cross |= a != b2 && out[a] == in[b2];
  const int per = cross ? 1 : 3;
  const dim3 grid(unsigned((ncells + 255) / 256));
  for (int f0 = 0; f0 < nfields && ncells > 0; f0 += per) {
    const int nf = std::min(per, nfields - f0);
    BluffFields bf{};
    for (int k = 0; k < nf; ++k) {
      bf.in[k] = in[f0 + k];
      bf.out[k] = out[f0 + k];
    }
    if (nf == 3)
      hipLaunchKernelGGL(k_bluff<3>, grid, dim3(256), 0, as_stream(stream), bf, sigma, dt, ncells);
    else if (nf == 2)
      hipLaunchKernelGGL(k_bluff<2>, grid, dim3(256), 0, as_stream(stream), bf, sigma, dt, ncells);
    else
      hipLaunchKernelGGL(k_bluff<1>, grid, dim3(256), 0, as_stream(stream), bf, sigma, dt, ncells);
    if (int e = launch_status("hfm_bluff_body_f32")) return e;
  }
  return HFM_OK;
}
