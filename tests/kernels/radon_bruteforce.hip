// Test plugin: a brute-force parallel-beam Radon transform with the
// reference's kernel ABI and geometry tables (gratopy/radon.cl:66-161 and
// 183-241).  The forward visits every pixel of the image for every detector
// bin instead of walking the ray's strip; pixels outside the strip get
// non-positive weights and are skipped, so the terms added are the strip's,
// in the same order.  The adjoint is the reference's two-tap interpolation.
#include "gratopy_plugin.h"

template <typename real, char OUT, char IN>
__device__ void radon_bruteforce(real *sino, const real *img, const real *ofs, const real *geo,
                                 gt_range r) {
  int s, a, z;
  if (!gt_global_id(r, s, a, z)) return;
  const int Nx = (int)geo[2], Ny = (int)geo[3];
  const float delta_x = (float)geo[0], delta_xi = (float)geo[1];
  const real *o = ofs + 8 * a;
  const bool swap = o[5] != (real)0;  // walk columns instead of rows
  const real X = swap ? o[1] : o[0], Y = swap ? o[0] : o[1];
  const int Nxx = swap ? Ny : Nx, Nyy = swap ? Nx : Ny;
  real acc = 0;
  for (int y = 0; y < Nyy; ++y) {
    const real d = (real)y * Y + o[2] - (real)s;
    for (int x = 0; x < Nxx; ++x) {
      const real w = (real)1 - fabs((real)x * X + d);
      if (w > (real)0) {
        const int px = swap ? y : x, py = swap ? x : y;
        acc += w * img[gt_offset<IN>(px, py, z, Nx, Ny, r.n2)];
      }
    }
  }
  sino[gt_offset<OUT>(s, a, z, r.n0, r.n1, r.n2)] =
      acc * (real)delta_x * (real)delta_x / (real)delta_xi;
}

template <typename real, char OUT, char IN>
__device__ void radon_bruteforce_ad(real *img, const real *sino, const real *ofs,
                                    const real *geo, gt_range r) {
  int x, y, z;
  if (!gt_global_id(r, x, y, z)) return;
  const int Ns = (int)geo[4], Na = (int)geo[5];
  real acc = 0;
  for (int a = 0; a < Na; ++a) {
    const real *o = ofs + 8 * a;
    const real s = (real)x * o[0] + (real)y * o[1] + o[2];
    int sm = (int)floor(s);
    int sp = sm + 1;
    real wp = (real)1 - ((real)sp - s);
    real wm = (real)1 - (s - (real)sm);
    if (sm < 0 || sm >= Ns) {
      wm = 0;
      sm = 0;
    }
    if (sp < 0 || sp >= Ns) {
      wp = 0;
      sp = 0;
    }
    acc += o[4] * (wm * sino[gt_offset<IN>(sm, a, z, Ns, Na, r.n2)] +
                   wp * sino[gt_offset<IN>(sp, a, z, Ns, Na, r.n2)]);
  }
  img[gt_offset<OUT>(x, y, z, r.n0, r.n1, r.n2)] = acc;
}

GT_EXPORT_VARIANTS(radon, radon_bruteforce,
                   (real *out, const real *in, const real *ofs, const real *geo, gt_range r),
                   (out, in, ofs, geo, r))
GT_EXPORT_VARIANTS(radon_ad, radon_bruteforce_ad,
                   (real *out, const real *in, const real *ofs, const real *geo, gt_range r),
                   (out, in, ofs, geo, r))
