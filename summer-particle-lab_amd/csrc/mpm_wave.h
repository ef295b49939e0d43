// Wave-wide reductions by DPP lane moves for gfx950 (GFX9 DPP controls).
//
// Lane i of each 16-lane row folds in lanes i-1, i-2..i-3, i-4..i-7 and
// i-8..i-15 of its row (a source outside the row reads as the identity), so
// lane 15 of a row holds the row's result; the broadcast of lane 15 of each
// row to the next row (rows 1 and 3) and of lane 31 to rows 2 and 3 fold the
// rows into lane 63, whose value is returned to every lane. The folds never
// count a lane twice, so sums are exact reductions as well as min/max.
// Every lane of the wave must be active.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143, DPP_WAVE_SHR1 = 0x138;

// x moved by a DPP control; lanes whose source is invalid or whose row is
// masked off take `identity`. 4- and 8-byte types.
template <int CTRL, int ROWS, typename T>
__device__ __forceinline__ T dpp_move(T x, T identity) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "4- or 8-byte values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, identity),
                                                             __builtin_bit_cast(int, x), CTRL, ROWS, 0xf, false));
  } else {
    const uint64_t xi = __builtin_bit_cast(uint64_t, x), id = __builtin_bit_cast(uint64_t, identity);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)id, (int)(uint32_t)xi, CTRL, ROWS, 0xf,
                                                              false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp((int)(uint32_t)(id >> 32), (int)(uint32_t)(xi >> 32),
                                                              CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  }
}

template <typename T>
__device__ __forceinline__ T lane63(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
  } else {
    const uint64_t xi = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)xi, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(xi >> 32), 63);
    return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
  }
}

// Fold of x over the wave with `fold` (associative and commutative) and its
// identity; the result is wave-uniform.
template <typename T, typename Fold>
__device__ __forceinline__ T wave_reduce(T x, T identity, Fold fold) {
  x = fold(x, dpp_move<DPP_ROW_SHR1, 0xf>(x, identity));
  x = fold(x, dpp_move<DPP_ROW_SHR2, 0xf>(x, identity));
  x = fold(x, dpp_move<DPP_ROW_SHR4, 0xf>(x, identity));
  x = fold(x, dpp_move<DPP_ROW_SHR8, 0xf>(x, identity));
  x = fold(x, dpp_move<DPP_ROW_BCAST15, 0xa>(x, identity));
  x = fold(x, dpp_move<DPP_ROW_BCAST31, 0xc>(x, identity));
  return lane63(x);
}

// A wave-uniform value moved into scalar registers, four bytes at a time.
template <typename T>
__device__ __forceinline__ T to_scalar(const T& v) {
  static_assert(sizeof(T) % 4 == 0, "whole words only");
  T out;
  const int* src = reinterpret_cast<const int*>(&v);
  int* dst = reinterpret_cast<int*>(&out);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
  return out;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
  return wave_reduce(x, T(0), [](T a, T b) { return a + b; });
}

}  // namespace mpm
