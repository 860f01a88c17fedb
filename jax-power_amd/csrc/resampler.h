// Device helpers of the mass-assignment kernels (jaxpower/resamplers.py:64-131): the 1-D
// kernels K_p(s) and their derivatives, the stencil origin rule and grid coordinates.
#pragma once

#include "paint.h"

namespace mp {

template <int ORDER>
__host__ __device__ __forceinline__ float kern1d(float s) {
  if constexpr (ORDER == 1) {
    return 1.f;
  } else if constexpr (ORDER == 2) {
    return 1.f - s;
  } else if constexpr (ORDER == 3) {
    return s <= 0.5f ? 0.75f - s * s : 0.5f * (1.5f - s) * (1.5f - s);
  } else {
    return s <= 1.f ? (4.f - 6.f * s * s + 3.f * s * s * s) * (1.f / 6.f) : (2.f - s) * (2.f - s) * (2.f - s) * (1.f / 6.f);
  }
}

// dK_p(s)/ds (one-sided value at the kinks, where K_p is continuous).
template <int ORDER>
__host__ __device__ __forceinline__ float kern1d_ds(float s) {
  if constexpr (ORDER == 1) {
    return 0.f;
  } else if constexpr (ORDER == 2) {
    return -1.f;
  } else if constexpr (ORDER == 3) {
    return s <= 0.5f ? -2.f * s : -(1.5f - s);
  } else {
    return s <= 1.f ? -2.f * s + 1.5f * s * s : -0.5f * (2.f - s) * (2.f - s);
  }
}

__host__ __device__ __forceinline__ int pos_mod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ T round_or_floor(T u, bool odd) {
  return odd ? rint(u) : floor(u);
}

template <typename T>
__device__ __forceinline__ void to_grid(const T* __restrict__ pos, int64_t i, const GridMap& g, T xg[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) xg[a] = (pos[3 * i + a] + (T)g.shift[a]) / (T)g.cell[a];
}

}  // namespace mp
