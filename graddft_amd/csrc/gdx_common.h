// Shared definitions for the gdx HIP library: status codes, error capture, MFMA vector
// types, layout constants and small device helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/gdx.h"

namespace gdx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// AO planes stored per grid point: value, d/dx, d/dy, d/dz, laplacian
constexpr int AO_PLANES = 5;
// per-point, per-spin density record: rho, drho/dx, drho/dy, drho/dz, tau, lapl
constexpr int DENS_STRIDE = 6;
// per-point, per-spin potential record consumed by the V_xc build:
// v_rho, 2 v_grad (x, y, z), 2 v_lapl, u = v_tau / 2 + 2 v_lapl
constexpr int VPT_STRIDE = 6;

void set_error(const std::string& msg);

#define GDX_HIP_CHECK(expr)                                                          \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::gdx::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));           \
      return GDX_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

#define GDX_REQUIRE(cond, msg)                                                       \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      ::gdx::set_error(msg);                                                         \
      return GDX_ERR_INVALID;                                                        \
    }                                                                                \
  } while (0)

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Accumulator layout of v_mfma_f64_16x16x4f64: element q of lane l holds row 4q + l/16,
// column l%16 (measured on gfx950; the f32/bf16 16x16 variants use row 4(l/16) + q).
__device__ __forceinline__ int mfma64_row(int lane, int q) { return 4 * q + (lane >> 4); }
__device__ __forceinline__ int mfma32_row(int lane, int q) { return 4 * (lane >> 4) + q; }

__device__ __forceinline__ double warp_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// sum across the 16 lanes that share (lane / 16)
__device__ __forceinline__ double sum16(double v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// MLP head: c_o(z) for the pre-activation z = (H W_last + b)_o.
//  mode 0 (NeuralFunctional head, grad_dft/functional.py:407-419): c = 2 sigmoid(z / 2)
//  mode 1 (B3LYP plus a neural correction): c = c_B3LYP,o + z
__device__ __forceinline__ double b3lyp_coef(int o) {
  return o == 0 ? 1.0 - 0.2 : (o == 1 ? 0.72 : (o == 2 ? 1.0 - 0.81 : 0.81));
}
__device__ __forceinline__ double head_coef(double z, int mode, int o) {
  return mode == 1 ? b3lyp_coef(o) + z : 2.0 / (1.0 + exp(-z / 2.0));
}
// dc/dz and d2c/dz2 from c
__device__ __forceinline__ double head_d1(double c, int mode) { return mode == 1 ? 1.0 : 0.5 * c * (1.0 - 0.5 * c); }
__device__ __forceinline__ double head_d2(double c, int mode) {
  return mode == 1 ? 0.0 : 0.5 * head_d1(c, 0) * (1.0 - c);
}

}  // namespace gdx
