// Device-side helpers for operator kernels compiled at run time (plugin
// sources of HipKernelSpec).  The loader puts this directory on the include
// path, so a plugin source starts with `#include "gratopy_plugin.h"`.
//
// Kernel ABI (the naming and argument order of the reference's operator
// kernels, gratopy/operator/opencl.py:55-101 and 444-454):
//   extern "C" __global__ void {base}[_ad]_{float|double}_{out_order}{in_order}(
//       real *out, const real *in, extra..., gt_range range)
// `range` is the output shape, the reference's OpenCL global size (absent axes
// are 1).  The launcher covers it with a grid of blocks; threads beyond it must
// return (gt_global_id tells them apart).
#pragma once

struct gt_range {
  int n0, n1, n2;
};

// Coordinates of this thread in the output shape; false for the threads of the
// last blocks that lie beyond it.
__device__ inline bool gt_global_id(const gt_range &r, int &i0, int &i1, int &i2) {
  i0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  i1 = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  i2 = (int)(blockIdx.z * blockDim.z + threadIdx.z);
  return i0 < r.n0 && i1 < r.n1 && i2 < r.n2;
}

// Element offset of (i0, i1, i2) in a dense (n0, n1, n2) array stored in
// Fortran order ('f': i0 fastest) or C order ('c': i2 fastest), as the
// reference's pos_img_f / pos_img_c (gratopy/radon.cl:28-31).
template <char ORDER>
__device__ inline long long gt_offset(int i0, int i1, int i2, int n0, int n1, int n2) {
  if (ORDER == 'f') return i0 + (long long)n0 * (i1 + (long long)n1 * i2);
  return i2 + (long long)n2 * (i1 + (long long)n1 * i0);
}

constexpr char gt_order_f = 'f';
constexpr char gt_order_c = 'c';

// GT_EXPORT_VARIANTS(name, body, params, args) defines the eight ABI kernels
// name_{float|double}_{f|c}{f|c}.  `body` is a device function template
//   template <typename real, char OUT, char IN> __device__ void body(...);
// `params` is the parenthesised parameter list, written with the type name
// `real`, and `args` the parenthesised argument list passed on to `body`.
#define GT_VARIANT_(name, body, params, args, R, O, I)                 \
  namespace gt_variant_##name##_##R##_##O##I {                        \
    using real = R;                                                   \
    extern "C" __global__ void name##_##R##_##O##I params {           \
      body<real, gt_order_##O, gt_order_##I> args;                    \
    }                                                                 \
  }

#define GT_EXPORT_TYPE_(name, body, params, args, R) \
  GT_VARIANT_(name, body, params, args, R, f, f)     \
  GT_VARIANT_(name, body, params, args, R, f, c)     \
  GT_VARIANT_(name, body, params, args, R, c, f)     \
  GT_VARIANT_(name, body, params, args, R, c, c)

#define GT_EXPORT_VARIANTS(name, body, params, args)  \
  GT_EXPORT_TYPE_(name, body, params, args, float)    \
  GT_EXPORT_TYPE_(name, body, params, args, double)
