// MFMA operand traits for gfx950 and 16-byte LDS vector access.
#pragma once
#include <hip/hip_runtime.h>

namespace rf {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

// v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16], B[l/16][l%16]; output
// register r of lane l is D[l/16 + 4r][l%16] (checked on gfx950).
template <>
struct Mfma<double> {
  static constexpr int KT = 16;
  using acc_t = v4d;
  using lds16_t = v2d;  // elements of one 16-byte LDS access
  __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int orow(int lane, int r) { return lane / 16 + 4 * r; }
};

// v_mfma_f32_16x16x4f32: same operand layout; output register r of lane l is
// D[4(l/16) + r][l%16] (checked on gfx950).
template <>
struct Mfma<float> {
  static constexpr int KT = 32;
  using acc_t = v4f;
  using lds16_t = v4f;
  __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int orow(int lane, int r) { return 4 * (lane / 16) + r; }
};

// four consecutive elements to a 16-byte aligned LDS address
__device__ inline void lds_put4(float* p, float a, float b, float c, float d) {
  *reinterpret_cast<v4f*>(p) = v4f{a, b, c, d};
}
__device__ inline void lds_put4(double* p, double a, double b, double c, double d) {
  *reinterpret_cast<v2d*>(p) = v2d{a, b};
  *reinterpret_cast<v2d*>(p + 2) = v2d{c, d};
}

}  // namespace rf
