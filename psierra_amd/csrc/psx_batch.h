// Batched launches over the tiles of many plans (device side; included by
// the .hip sources): uniform scalar loads, the instance that owns a global
// tile or node, and the tile loop of one launch.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "psx_device.h"
#include "psx_internal.h"

namespace psx {

// A work-group-uniform value (a plan view, a pointer, an index) read through
// the scalar unit, dword by dword.
template <typename T>
__device__ __forceinline__ T load_uniform_val(const T* p) {
  static_assert(sizeof(T) % 4 == 0 && std::is_trivially_copyable<T>::value, "dword-copyable type");
  T out;
  const auto* q = (const __attribute__((address_space(4))) uint32_t*)p;
  uint32_t* o = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) o[i] = q[i];
  return out;
}

// A plan view in the constant address space: its fields are read through
// the scalar unit where they are used, as a kernel argument's are.
using ConstDevPlan = __attribute__((address_space(4))) DevPlan;

// Instance owning global item g: the largest i with start[i] <= g
// (start[0] = 0; an empty instance has the next one's start and is skipped).
template <bool UNIFORM>
__device__ __forceinline__ int batch_owner(const int32_t* start, int n, int g) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    const int v = UNIFORM ? load_uniform_val(start + mid) : start[mid];
    if (v <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Tiles of a batched launch.  PERSISTENT: work-group w of XCD x takes tiles
// x * per + w, x * per + w + P, ... of the global list (P work-groups per
// XCD, as many as stay resident); the thread index is passed through an
// empty asm so that the compiler recomputes what derives from it in each
// tile instead of keeping it live across the loop (registers, occupancy).
// Otherwise one tile per work-group, XCD-contiguous as in a single-plan
// launch.
template <bool PERSISTENT, typename Body>
__device__ __forceinline__ void batch_tiles(const BatchDev& b, Body body) {
  if constexpr (PERSISTENT) {
    const int per = (b.n_tiles + 7) >> 3;
    const int x = int(blockIdx.x & 7u), w0 = int(blockIdx.x >> 3), P = int(gridDim.x >> 3);
    for (int w = w0; w < per; w += P) {
      const int g = x * per + w;
      if (g >= b.n_tiles) break;
      const int inst = batch_owner<true>(b.tile_start, b.n, g);
      if (w != w0) __syncthreads();   // the previous tile's LDS reads are done
      int t = int(threadIdx.x);
      __asm__ volatile("" : "+v"(t));
      body(inst, g - load_uniform_val(b.tile_start + inst), t);
    }
  } else {
    const int g = xcd_tile();
    if (g >= b.n_tiles) return;
    const int inst = batch_owner<true>(b.tile_start, b.n, g);
    body(inst, g - load_uniform_val(b.tile_start + inst), int(threadIdx.x));
  }
}

// Grid of a batched launch: persistent, as many work-groups as stay
// resident on the device (never more than there are tiles); otherwise one per
// tile.  A multiple of 8 either way (one share per XCD).
static inline int batch_grid(const void* kernel, int threads, int cus, int n_tiles, bool persistent) {
  int64_t g = n_tiles;
  if (persistent) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
      per_cu = 1;
    g = std::min<int64_t>(int64_t(cus) * per_cu, n_tiles);
  }
  return int(8 * ((g + 7) / 8));
}

}  // namespace psx
