// Shared device/host helpers for the lime MI355X kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lime {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// One 16x16x4 FP32 MFMA: D = A(16x4) * B(4x16) + C.
// Operand layout (wave64): lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// the result's register v of lane l is D[4 * (l / 16) + v][l % 16].
__device__ __forceinline__ floatx4 mfma_f32(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// DPP lane moves (no LDS round trip).  CTRL: 0xB1 = quad_perm(1,0,3,2), lane l reads
// l ^ 1; 0x4E = quad_perm(2,3,0,1), l ^ 2; 0x141 = row_half_mirror, l reads the lane at
// the mirrored position of its 8-lane half-row; 0x140 = row_mirror, of its 16-lane row.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the lanes of each aligned group of G lanes (G = 1, 2 or 4); every lane of the
// group receives the group's sum.
template <int G>
__device__ __forceinline__ float group_sum(float x) {
  static_assert(G == 1 || G == 2 || G == 4, "groups of 1, 2 or 4 lanes");
  if constexpr (G >= 2) x += dpp_f32<0xB1>(x);
  if constexpr (G >= 4) x += dpp_f32<0x4E>(x);
  return x;
}

// Sum over the 64 lanes of the wave, returned to every lane: within each 16-lane row on
// DPP (after the quad steps every lane of a quad holds the quad's sum, so the mirrored
// lane of the other half-row / half of the row holds the other partial), then across rows.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xB1>(x);
  x += dpp_f64<0x4E>(x);
  x += dpp_f64<0x141>(x);
  x += dpp_f64<0x140>(x);
  x += __shfl_xor(x, 16, 64);
  x += __shfl_xor(x, 32, 64);
  return x;
}

__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx b runs on XCD b % 8),
// each with its own L2.  Renumber them so that XCD x runs a contiguous range of logical
// ids: workgroups with consecutive logical ids (the tiles of one operator, the operators
// sharing one generator) then share one L2 instead of every XCD fetching the same data.
// A bijection for any grid size.
__device__ __forceinline__ int xcd_block() {
  constexpr int X = 8;
  const int g = gridDim.x, b = blockIdx.x;
  const int per = g / X, rem = g % X, x = b % X;
  return x * per + (x < rem ? x : rem) + b / X;
}

}  // namespace lime
