// Device helpers shared by the HIP kernels (included by .hip sources only).
#pragma once

#include <hip/hip_runtime.h>

namespace psx {

// Work-group -> tile mapping that gives each of the 8 XCDs a contiguous run
// of tiles (consecutive work-groups are dispatched round-robin over XCDs),
// so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_tile() {
  const int per = gridDim.x >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

// A work-group-uniform record read through the scalar unit (constant
// address space), so it needs no vector load and no lane broadcast.
__device__ __forceinline__ int4 load_uniform(const int4* p, int i) {
  const auto* q = (const __attribute__((address_space(4))) int*)(p + i);
  return make_int4(q[0], q[1], q[2], q[3]);
}

__device__ __forceinline__ int2 load_uniform(const int2* p, int i) {
  const auto* q = (const __attribute__((address_space(4))) int*)(p + i);
  return make_int2(q[0], q[1]);
}

// Lane exchange inside groups of four lanes (DPP quad permutation; 0xB1
// swaps neighbours, 0x4E swaps pairs), for a double.
// Lanes whose source falls outside the row read 0 (bound_ctrl).
template <int CTRL>
__device__ __forceinline__ double quad_dpp(double x) {
  const unsigned long long v = static_cast<unsigned long long>(__double_as_longlong(x));
  const unsigned lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(v)), CTRL, 0xF, 0xF, true));
  const unsigned hi = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(v >> 32)), CTRL, 0xF, 0xF, true));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Sum of one value over the four lanes of a quad, (x0 + x1) + (x2 + x3),
// the same bits in every lane of the quad.
__device__ __forceinline__ double quad_sum(double x) {
  x += quad_dpp<0xB1>(x);
  return x + quad_dpp<0x4E>(x);
}

// The same for a float (all four lanes of the quad active).
__device__ __forceinline__ float quad_sum(float x) {
  x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));
  return x + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, true));
}

// Node owning unit u: bit b of starts[] marks the first unit of a node with
// units, nz[] lists those nodes in order.
template <int NW>
__device__ __forceinline__ int unit_owner(const unsigned long long* starts, const uint8_t* nz, int u) {
  const int w = u >> 6;
  const unsigned long long below = (2ull << (u & 63)) - 1ull;   // bits 0..u%64
  int c = 0;
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const unsigned long long m = starts[i];
    c += __popcll(i < w ? m : i == w ? (m & below) : 0ull);
  }
  return nz[c - 1];
}

// Index of the lowest set bit of a nonzero mask.
__device__ __forceinline__ int low_bit(uint32_t m) { return __ffs(int(m)) - 1; }
__device__ __forceinline__ int low_bit(unsigned long long m) { return __ffsll(static_cast<long long>(m)) - 1; }

// Bits [b, e) of a 64-bit word, e - b < 64.
__device__ __forceinline__ unsigned long long bit_range(int b, int e) {
  return ((1ull << (e - b)) - 1ull) << b;
}

}  // namespace psx
