// Test plugin: forward out = 2*in + 1, adjoint out = 3*in - 1, elementwise over
// the output shape, for every element type and pair of array orders.
#include "gratopy_plugin.h"

template <typename real, char OUT, char IN>
__device__ void affine_forward(real *out, const real *in, gt_range r) {
  int x, y, z;
  if (!gt_global_id(r, x, y, z)) return;
  out[gt_offset<OUT>(x, y, z, r.n0, r.n1, r.n2)] =
      (real)2 * in[gt_offset<IN>(x, y, z, r.n0, r.n1, r.n2)] + (real)1;
}

template <typename real, char OUT, char IN>
__device__ void affine_adjoint(real *out, const real *in, gt_range r) {
  int x, y, z;
  if (!gt_global_id(r, x, y, z)) return;
  out[gt_offset<OUT>(x, y, z, r.n0, r.n1, r.n2)] =
      (real)3 * in[gt_offset<IN>(x, y, z, r.n0, r.n1, r.n2)] - (real)1;
}

GT_EXPORT_VARIANTS(affine, affine_forward, (real *out, const real *in, gt_range r), (out, in, r))
GT_EXPORT_VARIANTS(affine_ad, affine_adjoint, (real *out, const real *in, gt_range r),
                   (out, in, r))
