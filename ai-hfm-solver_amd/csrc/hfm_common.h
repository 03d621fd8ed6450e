// Shared device helpers for libhfm (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hfm.h"

namespace hfm {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
int launch_status(const char* what);  // HFM_OK or HFM_EDEVICE after a launch

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- storage types
struct bf16 {
  uint16_t v;
};

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(bf16 v) { return __uint_as_float(uint32_t(v.v) << 16); }

template <typename T>
__device__ __forceinline__ T from_f(float f);
template <>
__device__ __forceinline__ float from_f<float>(float f) {
  return f;
}
// Round to nearest even (gfx950's v_cvt_pk_bf16_f32; NaN stays NaN).
template <>
__device__ __forceinline__ bf16 from_f<bf16>(float f) {
  return bf16{__builtin_bit_cast(uint16_t, (__bf16)f)};
}
// Two values rounded to BF16 and packed (low half = a): one conversion.
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __bf16 h2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, h2));
}
// IEEE half storage (FP16): conversions round to nearest even; values past
// the range become +-inf, below it subnormal or zero.
struct f16 {
  uint16_t v;
};
__device__ __forceinline__ float to_f(f16 v) { return float(__builtin_bit_cast(_Float16, v.v)); }
template <>
__device__ __forceinline__ f16 from_f<f16>(float f) {
  return f16{__builtin_bit_cast(uint16_t, (_Float16)f)};
}
// Two values rounded to FP16 and packed (low half = a).
__device__ __forceinline__ uint32_t pack_f16x2(float a, float b) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef _Float16 h2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, h2));
}
__device__ __forceinline__ float f16_lo(uint32_t w) {
  return float(__builtin_bit_cast(_Float16, uint16_t(w & 0xFFFFu)));
}
__device__ __forceinline__ float f16_hi(uint32_t w) {
  return float(__builtin_bit_cast(_Float16, uint16_t(w >> 16)));
}

template <typename T>
struct is_f16 {
  static constexpr bool value = false;
};
template <>
struct is_f16<f16> {
  static constexpr bool value = true;
};
// The fine residual as the V-cycle gets it: FP16 storage scales it by 2^e
// (exact), the other types take it as is.
template <typename T>
__device__ __forceinline__ float rscale(float v, int e) {
  if constexpr (is_f16<T>::value) return __builtin_ldexpf(v, e);
  return v;
}

// Value after a round trip through storage type T.
template <typename T>
__device__ __forceinline__ float round_to(float f) {
  return to_f(from_f<T>(f));
}

// ---------------------------------------------------------------- global pointers
// Pointers read from work tables are generic to the compiler, which then
// emits FLAT accesses (they also count against the LDS counter that shuffles
// and reductions wait on).  Kernels cast them once to the global address space.
#define HFM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const HFM_GLOBAL T* gin(const void* p) {
  return (const HFM_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ HFM_GLOBAL T* gout(void* p) {
  return (HFM_GLOBAL T*)p;
}
// Element access through global pointers (bf16 as its 16-bit pattern).
__device__ __forceinline__ float ldg(const HFM_GLOBAL float* p) { return *p; }
__device__ __forceinline__ float ldg(const HFM_GLOBAL bf16* p) {
  return __uint_as_float(uint32_t(*(const HFM_GLOBAL uint16_t*)p) << 16);
}
__device__ __forceinline__ void put(HFM_GLOBAL float* p, float v) { *p = v; }
__device__ __forceinline__ void put(HFM_GLOBAL float* p, bf16 v) { *p = to_f(v); }
__device__ __forceinline__ void put(HFM_GLOBAL bf16* p, bf16 v) { *(HFM_GLOBAL uint16_t*)p = v.v; }
__device__ __forceinline__ void put(HFM_GLOBAL bf16* p, float v) {
  *(HFM_GLOBAL uint16_t*)p = from_f<bf16>(v).v;
}
__device__ __forceinline__ float ldg(const HFM_GLOBAL f16* p) {
  return to_f(f16{*(const HFM_GLOBAL uint16_t*)p});
}
// One stored value as loaded, not converted (for values used well after their load).
__device__ __forceinline__ float ldraw(const HFM_GLOBAL float* p) { return *p; }
__device__ __forceinline__ bf16 ldraw(const HFM_GLOBAL bf16* p) {
  return bf16{*(const HFM_GLOBAL uint16_t*)p};
}
__device__ __forceinline__ f16 ldraw(const HFM_GLOBAL f16* p) {
  return f16{*(const HFM_GLOBAL uint16_t*)p};
}
__device__ __forceinline__ void put(HFM_GLOBAL float* p, f16 v) { *p = to_f(v); }
__device__ __forceinline__ void put(HFM_GLOBAL f16* p, f16 v) { *(HFM_GLOBAL uint16_t*)p = v.v; }
__device__ __forceinline__ void put(HFM_GLOBAL f16* p, float v) {
  *(HFM_GLOBAL uint16_t*)p = from_f<f16>(v).v;
}

// ---------------------------------------------------------------- division
// q = n / d for n < 2^32 with mag = floor(2^64 / d) + 1 (mag = 0 means d = 1).
__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint64_t mag) {
  return mag ? uint32_t(__umul64hi(uint64_t(n), mag)) : n;
}

// ---------------------------------------------------------------- reductions
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// Deterministic block sum: fixed shuffle tree per wave, waves summed in order.
// `half` selects the kBlock-thread half of a 2 kBlock-thread workgroup that
// runs two blocks' work side by side (the persistent kernel); both halves
// call it together (its barriers are the workgroup's).
__device__ __forceinline__ double block_sum(double v, int half = 0) {
  __shared__ double part[2][kWaves];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = (threadIdx.x >> 6) & (kWaves - 1);
  if (lane == 0) part[half][wid] = v;
  __syncthreads();
  double s = 0.0;
  if ((threadIdx.x & (kBlock - 1)) == 0) {
    for (int w = 0; w < kWaves; ++w) s += part[half][w];
  }
  __syncthreads();
  return s;  // valid in the half's thread 0
}

// A block's coordinates: virtual block g of a launch of nblocks blocks and the
// thread's index in it.  A kernel launch passes blockIdx / gridDim /
// threadIdx; the persistent kernel passes the work item it runs (half = which
// kBlock-thread half of its workgroup runs a kBlock-thread block).
struct Blk {
  int g, nblocks, tid, half;
};
__device__ __forceinline__ Blk launch_blk() {
  return Blk{int(blockIdx.x), int(gridDim.x), int(threadIdx.x), 0};
}

// ---------------------------------------------------------------- work table
// Block g of a launch belongs to task blk_task[g]; tasks start at a multiple
// of 8 blocks and own a multiple of 8 blocks, so (g - blk0) % 8 is the XCD the
// block is dispatched to.  Each XCD gets a contiguous range of chunks.
__device__ __forceinline__ int xcd_chunk(int local, int nblk) {
  const int per = nblk >> 3;
  return (local & 7) * per + (local >> 3);
}

}  // namespace hfm
