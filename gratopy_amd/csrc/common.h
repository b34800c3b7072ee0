// Shared helpers of the gfx950 projection library: error reporting for the
// C ABI, array-order strides, and per-type arithmetic primitives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/gratopy_hip.h"

namespace gt {

// ---- error state of the C ABI (one message per calling thread) ------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// Returns GT_EHIP (and records the message) when the last launch failed.
int check_launch(const char *what);

inline size_t align_up(size_t n, size_t a = 256) { return (n + a - 1) / a * a; }

// ---- launch-shape knobs set by the autotuner (gt_tune_set) -----------------
struct Tuning {
  int fwd_group = 4;  // detector bins per lane in the forward kernels (1, 2, 4, 8, 16, 32)
  int back_rows = 1;  // pixels per lane (along y) in the backprojection (1, 2, 4, 8)
  int fwd_bands = 1;  // row bands of the forward strip walk (1, 2, 4, 8, 16, 32; fwd_group > 1)
  int back_bands = 1;  // angle bands of the backprojection (1, 2, 4, 8)
  int back_stage = 0;  // angles per LDS staging chunk of the backprojection (0 = unstaged, 8, 16)
  int back_tile = 16;  // lanes along x of a staged backprojection tile (16, 32)
  int fwd_chunk = 1;   // 4: bin-group strip rows read 16 bytes per load (1 = per pixel)
  int fwd_compact = 1;  // 1: banded forward launches only the bin groups that meet the band
  int automatic = 1;   // 1: each call picks its launch shape from its size; 0: the knobs above
};
Tuning &tuning();

// ---- array orders (gratopy/radon.cl:28-31) --------------------------------
struct Strides3 {
  int64_t s0, s1, s2;  // stride of index 0 (x or s), index 1 (y or a), index 2 (z)
};

// Fortran order: index 0 fastest; C order: index 2 fastest.
inline Strides3 order_strides(int order, int n0, int n1, int n2) {
  if (order == GT_ORDER_F) return {1, (int64_t)n0, (int64_t)n0 * n1};
  return {(int64_t)n2 * n1, (int64_t)n2, 1};
}

// ---- per-type arithmetic ----------------------------------------------------
template <typename real>
struct Arith;

template <>
struct Arith<float> {
  // v_rcp_f32: 1 ulp; the projections tolerate it (reference OpenCL allows
  // 2.5 ulp for single precision division).
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float clamp(float x, float lo, float hi) {
    return __builtin_amdgcn_fmed3f(x, lo, hi);
  }
  // x - floor(x), exact (v_fract_f32)
  static __device__ __forceinline__ float fract(float x) { return __builtin_amdgcn_fractf(x); }
};

template <>
struct Arith<double> {
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double clamp(double x, double lo, double hi) {
    return fmin(fmax(x, lo), hi);
  }
  static __device__ __forceinline__ double fract(double x) { return __builtin_amdgcn_fract(x); }
};

}  // namespace gt

#define GT_REQUIRE(cond, msg)                   \
  do {                                          \
    if (!(cond)) return gt::fail(GT_EINVAL, msg); \
  } while (0)
