// Device bodies of the projection kernels, shared by the single-problem
// launches (projection.hip) and the batched persistent launches (batch.hip).
//
// Forward (gratopy/radon.cl:66-161, gratopy/fanbeam.cl:72-256):
//   sino[s,a,z] = sum over the pixels of the strip between the rays through
//   detector positions s-1 and s+1 of hat-weight * img.  Rays closer to the
//   x axis walk a y-fastest copy of the image (made by a tiled transpose), so
//   a wave always reads along the fastest axis; rows that cannot touch the
//   image are skipped (closed-form row range).
//   * per-ray body: one lane per detector bin, the reference's loop order;
//   * bin-group body: a lane owns K adjacent bins [s0, s0+K) and walks the
//     union strip between the rays through s0-1 and s0+K once per row; each
//     strip pixel is projected once and its two hat weights (1-f, f) go to
//     bins floor(p), floor(p)+1, accumulated in lane-private LDS slots.  Per
//     bin the summation order is the reference's (rows ascending, then along
//     the row), so results agree to rounding.
// Backward (gratopy/radon.cl:183-241, gratopy/fanbeam.cl:290-391):
//   img[x,y,z] = sum_a lerp(S_a, p_a(x,y)) [ * 1/t_a(x,y) for fanbeam ] over a
//   zero-padded sinogram S already multiplied by the angle weight (and
//   sdpd[s]/R for fanbeam) and a per-angle affine table read with scalar loads.
#pragma once

#include "common.h"

namespace gt {

template <typename real>
struct alignas(sizeof(real) * 8) AngleRec {
  real c[8];
};

// Everything one projection needs.  Forward jobs use sino/so/imgX/imgY (an
// x-fastest and a y-fastest copy of the image: row strides rowX/rowY, slice
// strides sliceX/sliceY; rows may be padded with zeros), backward jobs
// img/io/S/tab.  Batched jobs also name src, the x-fastest image the copies
// are made from (and the TV dual step reads).  Epilogues (data shares the sinogram's strides):
//   mode 0: store the projection;
//   mode 1: forward stores value - data[k] (Landweber residual), backward
//           stores img[k] + scale * value (Landweber update);
//   mode 2: one step of the TV primal-dual iteration of
//           gratopy/gratopy.py:2554-2762 for a 2-D slice: forward updates
//           lambda = sino in place, (lambda + sigma_p (value - data)) * c
//           (total_variation.cl:95-106); backward writes
//           img = max(0, u + tau (div v - norming value)) and
//           u <- 2 img - u (total_variation.cl:108-148 and the extragradient
//           step gratopy.py:2708-2717).
template <typename real>
struct Job {
  int fan;
  int Nx, Ny, Ns, Na, Nz;
  const real *ofs, *sdpd, *geo;
  real *sino;
  Strides3 so;
  const real *imgX, *imgY, *src;
  int64_t rowX, rowY, sliceX, sliceY;
  real *part;  // batched jobs: the slice's band partials of a row-banded forward
  real *img;
  Strides3 io;
  const real *S;
  const AngleRec<real> *tab;
  int mode;
  const real *data;
  real scale;
  real *u;
  const real *v;
  float sigma_p, c, tau, norming;
};

template <int K>
struct Slots {
  static constexpr int kCount = K + 4;
};

// One strip pixel's contribution to its lane's slots (rows of 256 entries,
// one entry per lane; bin s0 + k is slot k + 2): value v at detector
// position p in slot units, clamped to [0.5, K + 2.5] so that truncation is
// floor; hat weights slot[floor p] += (1 - f) v, slot[floor p + 1] += f v,
// f = p - floor p.
template <typename real>
__device__ __forceinline__ void deposit(real *acc, real p, real v) {
  real *slot = acc + (int)p * 256;
  const real f = Arith<real>::fract(p);
  slot[0] += v - f * v;
  slot[256] += f * v;
}

// ---------------------------------------------------------------------------
// epilogues
// ---------------------------------------------------------------------------
// The empty asm pins the projection value in a register: the epilogue then
// starts from the rounded value, as when a separate kernel reads it back,
// instead of contracting the projection's last multiply into an FMA.
template <typename real>
__device__ __forceinline__ void round_barrier(real &v) {
  asm volatile("" : "+v"(v));
}

template <typename real>
__device__ __forceinline__ void store_fwd(const Job<real> &J, int s, int a, int z, real v) {
  const int64_t k = s * J.so.s0 + a * J.so.s1 + z * J.so.s2;
  if (J.mode != 0) round_barrier(v);
  if (J.mode == 1) v -= J.data[k];
  if (J.mode == 2) v = (J.sino[k] + (real)J.sigma_p * (v - J.data[k])) * (real)J.c;
  J.sino[k] = v;
}

// Primal step at pixel (x, y) of a 2-D slice; kstar = (B lambda)(x, y).
// The divergence is the negative adjoint of the forward-difference gradient
// with a zero last difference (total_variation.cl:108-148).
template <typename real>
__device__ __forceinline__ void tv_primal_store(const Job<real> &J, int x, int y, int64_t i,
                                                real kstar) {
  const real *vp = J.v + 4 * i;
  real d0 = vp[0], d1 = vp[1];
  if (x == J.Nx - 1) d0 = 0;
  if (x > 0) d0 -= J.v[4 * (i - J.io.s0) + 0];
  if (y == J.Ny - 1) d1 = 0;
  if (y > 0) d1 -= J.v[4 * (i - J.io.s1) + 1];
  const real d2 = 0;
  const real old = J.u[i];
  const real un = fmax((real)0, old + (real)J.tau * (d0 + d1 + d2 - (real)J.norming * kstar));
  J.img[i] = un;
  J.u[i] = (real)2 * un - old;
}

template <typename real>
__device__ __forceinline__ void store_back(const Job<real> &J, int x, int y, int z, real v) {
  const int64_t k = x * J.io.s0 + y * J.io.s1 + z * J.io.s2;
  if (J.mode != 0) round_barrier(v);
  if (J.mode == 2) {
    tv_primal_store(J, x, y, k, v);
    return;
  }
  if (J.mode == 1) v = J.img[k] + J.scale * v;
  J.img[k] = v;
}

// Dual step at pixel (x, y) of a 2-D slice (total_variation.cl:54-93 and
// the projection onto the unit ball, 150-167): v += sigma grad(ubar), then
// v /= max(1, |v|).  ubar is read through src (x fastest), v has 4 entries
// per pixel at 4 * (x + Nx y).
template <typename real>
__device__ __forceinline__ void tv_dual_pixel(const Job<real> &J, real *__restrict__ v, float sigma,
                                              int x, int y) {
  const int64_t i = x + (int64_t)J.Nx * y;
  const real *u = J.src;
  const real c = u[i];
  real g0 = -c, g1 = -c, g2 = -c;
  g0 = x < J.Nx - 1 ? g0 + u[i + 1] : (real)0;
  g1 = y < J.Ny - 1 ? g1 + u[i + J.Nx] : (real)0;
  g2 = (real)0;
  real *vp = v + 4 * i;
  const real sg = (real)sigma;
  real v0 = vp[0] + sg * g0;
  real v1 = vp[1] + sg * g1;
  real v2 = vp[2] + sg * g2;
  const real norm = hypot(hypot(v0, v1), v2);
  if (norm > (real)1) {
    v0 /= norm;
    v1 /= norm;
    v2 /= norm;
  }
  vp[0] = v0;
  vp[1] = v1;
  vp[2] = v2;
}

// ---------------------------------------------------------------------------
// 32x32 tile of a copy between two orders (LDS transpose).
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ void relayout_tile(real (*tile)[33], real *__restrict__ dst, Strides3 ds,
                                              const real *__restrict__ src, Strides3 ss, int N0,
                                              int N1, int64_t z, int b0, int b1) {
  const int read0fast = ss.s0 <= ss.s1;
  const int write0fast = ds.s0 <= ds.s1;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  for (int k = ly; k < 32; k += 8) {
    const int i0 = read0fast ? b0 + lx : b0 + k;
    const int i1 = read0fast ? b1 + k : b1 + lx;
    if (i0 < N0 && i1 < N1) tile[i0 - b0][i1 - b1] = src[i0 * ss.s0 + i1 * ss.s1 + z * ss.s2];
  }
  __syncthreads();
  for (int k = ly; k < 32; k += 8) {
    const int i0 = write0fast ? b0 + lx : b0 + k;
    const int i1 = write0fast ? b1 + k : b1 + lx;
    if (i0 < N0 && i1 < N1) dst[i0 * ds.s0 + i1 * ds.s1 + z * ds.s2] = tile[i0 - b0][i1 - b1];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Row ranges.  Conservative [ylo, yhi]: rows outside have no pixel in the
// ray's strip (one row and one column of margin).
// ---------------------------------------------------------------------------
__device__ inline void clip_rows(double lo, double hi, int Nyy, int &ylo, int &yhi) {
  if (lo != lo || hi != hi) {  // NaN: fall back to every row
    ylo = 0;
    yhi = Nyy - 1;
    return;
  }
  // clamp in floating point first: infinities must not reach the int conversion
  lo = fmin(fmax(lo - 1.0, -1.0), (double)Nyy);
  hi = fmax(fmin(hi + 1.0, (double)Nyy), -1.0);
  ylo = (int)floor(lo);
  yhi = (int)ceil(hi);
  ylo = ylo < 0 ? 0 : ylo;
  yhi = yhi > Nyy - 1 ? Nyy - 1 : yhi;
}

// m*u >= L restricts u to a half line (or to nothing / everything).
__device__ inline void restrict_ge(double m, double L, double &lo, double &hi) {
  if (m > 0)
    lo = fmax(lo, L / m);
  else if (m < 0)
    hi = fmin(hi, L / m);
  else if (L > 0) {
    lo = 1.0;
    hi = 0.0;
  }
}

// Parallel strip of half width `half` detector units around detector
// position c (axes already swapped): pixels with |x*X + y*Y + off - c| < half.
// The strip centre in x is g(y) = -(off - c + y*Y)/X, half width half/|X|.
__device__ inline void radon_rows(double X, double Y, double off_minus_c, double half, int Nxx,
                                  int Nyy, int &ylo, int &yhi) {
  const double h = half / fabs(X);
  const double g0 = -off_minus_c / X, m = -Y / X;
  double lo = -INFINITY, hi = INFINITY;
  restrict_ge(m, -1.0 - h - g0, lo, hi);           // g(y) >= -1 - h
  restrict_ge(-m, g0 - ((double)Nxx + h), lo, hi);  // g(y) <= Nxx + h
  if (lo > hi) {
    ylo = 0;
    yhi = -1;
    return;
  }
  clip_rows(lo, hi, Nyy, ylo, yhi);
}

// Fan strip (axes already swapped): row y is visited on [xlow(y), xhigh(y)]
// with xhigh = qx + dpx*u, xlow = qx + dmx*u, u = y - midy - qy.
__device__ inline void fan_rows(double qx, double qy, double dpx, double dmx, double midx,
                                double midy, int Nxx, int Nyy, int &ylo, int &yhi) {
  double lo = -INFINITY, hi = INFINITY;
  restrict_ge(dpx, -midx - 1.0 - qx, lo, hi);              // xhigh >= -midx - 1
  restrict_ge(-dmx, -((double)Nxx - midx - qx), lo, hi);  // xlow <= Nxx - midx
  if (lo > hi) {
    ylo = 0;
    yhi = -1;
    return;
  }
  clip_rows(lo + midy + qy, hi + midy + qy, Nyy, ylo, yhi);
}

// CH consecutive elements of an image row in one load (CH * sizeof(real) <= 16
// bytes; element alignment suffices).
template <typename real, int CH>
struct RowChunk {
  real v[CH];
};

// A pointer into global memory.  Loads through it stay global_load: through
// a generic pointer they would be flat loads, which also count against the
// LDS wait counter and so would stall every LDS access of the strip walk.
template <typename real>
using GPtr = const __attribute__((address_space(1))) real *;

template <typename real, int CH>
__device__ __forceinline__ RowChunk<real, CH> load_chunk(GPtr<real> p) {
  RowChunk<real, CH> c;
  __builtin_memcpy(&c, (const __attribute__((address_space(1))) RowChunk<real, CH> *)p, sizeof(c));
  return c;
}

// px += CH for the strip walks; the empty asm keeps px itself as the loop's
// induction variable (otherwise the compiler rebuilds the address from the
// row start, a step counter and the first pixel's offset on every chunk, two
// VALU instructions more)
template <typename real, int CH>
__device__ __forceinline__ void advance(GPtr<real> &px) {
  px += CH;
  asm volatile("" : "+v"(px));
}

// Strip walks over a buffer descriptor (BUFW): the two row-padded image
// copies lie in one workspace block starting at imgX, and a lane walks a row
// by a 32-bit byte offset into it, so the per-chunk address step and loop
// test are 32-bit operations instead of 64-bit pointer arithmetic.  The
// launch takes this form only where the block is under 2^31 bytes
// (strip_buffer_fits); offsets past its end would read zeros.
template <typename real>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t strip_buffer(const Job<real> &J) {
  const int64_t bytes = (int64_t)((const char *)(J.imgY + (int64_t)J.Nz * J.sliceY) -
                                  (const char *)J.imgX);
  return __builtin_amdgcn_make_buffer_rsrc((void *)J.imgX, (short)0, (int)bytes, 0x00020000);
}

inline bool strip_buffer_fits(size_t bytes) { return bytes < ((size_t)1 << 31); }

template <typename real, int CH>
__device__ __forceinline__ RowChunk<real, CH> load_chunk_buf(__amdgpu_buffer_rsrc_t rs, int off) {
  static_assert(CH * sizeof(real) == 16, "buffer strip walks read 16 bytes per load");
  return __builtin_bit_cast(RowChunk<real, CH>, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
}

// ---------------------------------------------------------------------------
// Fan ray geometry of a strip bounded by the rays through detector positions
// s_lo and s_hi (fanbeam.cl:100-184 with s-1, s+1 replaced by s_lo, s_hi).
// ---------------------------------------------------------------------------
template <typename real>
struct FanStrip {
  real ddx, ddy, dx, dy, qx, qy, dpx, dmx, midx, midy, RE;
  int Nxx, Nyy;
  const real *base;
  int64_t row_stride;
  // both bounding rays advance the same way along the row axis, so each row
  // meets the strip in one interval [xlow(y), xhigh(y)].  The rays of a
  // single bin always do on the reference's geometries; the union strip of a
  // bin group can fail it when bins are very wide (a few detector bins over
  // a large angle), and the group then projects its bins one by one.
  bool regular;
};

template <typename real>
__device__ __forceinline__ FanStrip<real> fan_strip(const Job<real> &J, int a, int z, real s_lo,
                                                    real s_hi) {
  FanStrip<real> F;
  const real *geo = J.geo;
  const real R = geo[0];
  const real delta_xi = geo[2];
  const real mid_det = geo[5];
  F.midx = geo[3];
  F.midy = geo[4];
  F.RE = geo[1];
  const real *o = J.ofs + 8 * (int64_t)a;
  real dx = o[0], dy = o[1], qx = o[2], qy = o[3];
  const real d0x = o[4], d0y = o[5];
  const real sp = -mid_det + s_hi;
  const real sm = -mid_det + s_lo;
  real dpx = d0x + dx * sp - qx, dpy = d0y + dy * sp - qy;
  real dmx = d0x + dx * sm - qx, dmy = d0y + dy * sm - qy;
  real ddx = d0x - qx, ddy = d0y - qy;
  const real RR = R * R;
  ddx /= RR;
  ddy /= RR;
  const real dxi2 = delta_xi * delta_xi;
  dx /= dxi2;
  dy /= dxi2;
  F.RE *= (real)1 / R;
  F.Nxx = J.Nx;
  F.Nyy = J.Ny;
  if (fabs(dpx) < fabs(dpy) && fabs(dmx) < fabs(dmy)) {
    F.base = J.imgX + z * J.sliceX;
    F.row_stride = J.rowX;
  } else {
    real t;
    t = ddx; ddx = ddy; ddy = t;
    t = dx; dx = dy; dy = t;
    t = qx; qx = qy; qy = t;
    t = dpx; dpx = dpy; dpy = t;
    t = dmx; dmx = dmy; dmy = t;
    t = F.midx; F.midx = F.midy; F.midy = t;
    F.Nyy = J.Nx;
    F.Nxx = J.Ny;
    F.base = J.imgY + z * J.sliceY;
    F.row_stride = J.rowY;
  }
  F.regular = dpy * dmy > (real)0;
  dpx /= dpy;
  dmx /= dmy;
  if (qy * (dmx - dpx) < (real)0) {
    const real t = dpx;
    dpx = dmx;
    dmx = t;
  }
  F.ddx = ddx; F.ddy = ddy; F.dx = dx; F.dy = dy; F.qx = qx; F.qy = qy;
  F.dpx = dpx; F.dmx = dmx;
  return F;
}

// ---------------------------------------------------------------------------
// Forward bodies.
// ---------------------------------------------------------------------------
template <typename real>
__device__ __forceinline__ void radon_fwd_ray(const Job<real> &J, int s, int a, int z) {
  // the reference reads the scales as float even in double kernels (radon.cl:82-83)
  const float delta_x = (float)J.geo[0];
  const float delta_xi = (float)J.geo[1];
  const real *o = J.ofs + 8 * (int64_t)a;
  real X = o[0], Y = o[1];
  const real off = o[2], inv = o[3];
  int Nxx = J.Nx, Nyy = J.Ny;
  const real *base;
  int64_t row_stride;
  if (o[5] != (real)0) {
    const real t = X;
    X = Y;
    Y = t;
    Nxx = J.Ny;
    Nyy = J.Nx;
    base = J.imgY + z * J.sliceY;
    row_stride = J.rowY;
  } else {
    base = J.imgX + z * J.sliceX;
    row_stride = J.rowX;
  }
  int ylo, yhi;
  radon_rows((double)X, (double)Y, (double)off - (double)s, 1.0, Nxx, Nyy, ylo, yhi);
  real acc = 0;
  for (int y = ylo; y <= yhi; ++y) {
    const real d = (real)y * Y + off - (real)s;
    int xl = (int)((-1 - d) * inv);
    int xh = (int)((1 - d) * inv);
    if (inv < (real)0) {
      const int t = xl;
      xl = xh;
      xh = t;
    }
    xl = xl < 0 ? 0 : xl;
    xh = xh > Nxx - 1 ? Nxx - 1 : xh;
    const real *row = base + (int64_t)y * row_stride;
    for (int x = xl; x <= xh; ++x) {
      const real w = (real)1 - fabs((real)x * X + d);
      const real v = row[x];
      acc = w > (real)0 ? acc + w * v : acc;
    }
  }
  store_fwd(J, s, a, z, acc * (real)delta_x * (real)delta_x / (real)delta_xi);
}

__device__ __forceinline__ void band_rows(int Nyy, int index, int count, int &ylo, int &yhi);

// Unscaled sum of one fan ray (fanbeam.cl:113-250); with band_count > 1 only
// over the rows of band band_index of its frame.
template <typename real>
__device__ __forceinline__ real fan_ray_sum(const Job<real> &J, int s, int a, int z,
                                            int band_index = 0, int band_count = 1) {
  using A = Arith<real>;
  const FanStrip<real> F = fan_strip(J, a, z, (real)s - (real)1, (real)s + (real)1);
  const real smd = (real)s - J.geo[5];
  int ylo, yhi;
  fan_rows((double)F.qx, (double)F.qy, (double)F.dpx, (double)F.dmx, (double)F.midx,
           (double)F.midy, F.Nxx, F.Nyy, ylo, yhi);
  if (band_count > 1) band_rows(F.Nyy, band_index, band_count, ylo, yhi);
  const real xcap = (real)(F.Nxx - 1) - F.midx;
  real acc = 0;
  for (int y = ylo; y <= yhi; ++y) {
    const real u = (real)y - F.midy;
    const real xhigh = F.qx + F.dpx * (u - F.qy);
    const real xlow = F.qx + F.dmx * (u - F.qy);
    // bounds of the strip in this row; the outer fmax/fmin only keep
    // infinities (rays parallel to the rows) away from the int conversion
    const int xh = (int)floor(fmax(fmin(xcap, xhigh) + F.midx, (real)-1));
    const int xl = (int)ceil(fmin(fmax(-F.midx, xlow) + F.midx, (real)F.Nxx));
    const real xr = (real)xl - F.midx;
    real t = F.ddx * xr + F.ddy * u + F.RE;
    real ss = F.dx * xr + F.dy * u;
    const real *row = F.base + (int64_t)y * F.row_stride;
    for (int x = xl; x <= xh; ++x) {
      const real r = A::rcp(t);
      const real xi = ss * r;
      const real w = ((real)1 - fabs(smd - xi)) * r;
      acc += w * row[x];
      t += F.ddx;
      ss += F.dx;
    }
  }
  return acc;
}

// sino[s] = acc * sdpd[s] / delta_xi * delta_x / R (fanbeam.cl:254-255)
template <typename real>
__device__ __forceinline__ void fan_store(const Job<real> &J, int s, int a, int z, real acc) {
  const real R = J.geo[0], delta_xi = J.geo[2], delta_x = J.geo[10];
  store_fwd(J, s, a, z, acc * J.sdpd[s] / delta_xi * delta_x / R);
}

template <typename real>
__device__ __forceinline__ void fan_fwd_ray(const Job<real> &J, int s, int a, int z) {
  fan_store(J, s, a, z, fan_ray_sum(J, s, a, z));
}

// Row band of a banded forward launch: the strip walk is restricted to rows
// [index*Nyy/count, (index+1)*Nyy/count) of its frame and the raw per-bin
// sums go to part[((index*Nz + z)*Na + a)*Ns + s]; a second pass adds the
// bands in order and applies the scale and epilogue (band_reduce_entry).
template <typename real>
struct Band {
  int index, count;
  real *part;
};

template <typename real>
__device__ __forceinline__ real *band_slot(const Job<real> &J, const Band<real> &B, int s, int a,
                                           int z) {
  return B.part + (((int64_t)B.index * J.Nz + z) * J.Na + a) * J.Ns + s;
}

__device__ __forceinline__ void band_rows(int Nyy, int index, int count, int &ylo, int &yhi) {
  const int r0 = (int)((int64_t)Nyy * index / count);
  const int r1 = (int)((int64_t)Nyy * (index + 1) / count) - 1;
  ylo = ylo > r0 ? ylo : r0;
  yhi = yhi < r1 ? yhi : r1;
}

// Whole-column blocks of the bin-group forward cover kColumnGroups adjacent
// groups of kColumnAngles adjacent angles (256 lanes).
constexpr int kColumnAngles = 16;
constexpr int kColumnGroups = 256 / kColumnAngles;

// Lane of a whole-column block: wave y takes groups [4 y, 4 y + 4) of the
// column for all 16 angles, lane x the angle x / 4 and group 4 y + x % 4.
// Returns the angle's index in the block; *goff the group's offset in the
// column.  A wave then spans a few groups at 16 angles 0.25 degrees apart
// (on the 1440-angle stacks): its lanes' strips have nearly equal widths
// per row, so a row costs the wave little more than its lanes' own 16-byte
// chunks (64 groups of one angle span ~40 % of a fan, whose strip widths
// differ by a chunk or more), and its loads span a narrow part of the row.
__device__ __forceinline__ int column_lane(int &goff) {
  constexpr int LW = 64 / kColumnAngles;
  goff = LW * threadIdx.y + (threadIdx.x % LW);
  return threadIdx.x / LW;
}

// Lane mapping of the bin-group forward launches (256-lane blocks of 64 x 4;
// groups = ceil(Ns / K) bin groups per angle).  The whole columns of
// kColumnGroups groups come first, one block per (column, kColumnAngles
// angles), lanes by column_lane: its waves walk nearly the same pixels,
// which the CU's cache then serves once.  The groups left over
// (groups % kColumnGroups of them; all of a small slice's) are packed over
// consecutive angles, lane (a * tail + j) taking group cols * kColumnGroups
// + j of angle a, so their waves are full instead of mostly idle and a block
// holds the same detector edge for a run of adjacent angles.  Returns a (>=
// Na for lanes past the end).
template <int K>
__device__ __forceinline__ int group_lane(int block, int Ns, int Na, int &g) {
  constexpr int A = kColumnAngles, C = kColumnGroups;
  const int groups = (Ns + K - 1) / K;
  const int cols = groups / C;
  const int nfb = cols * ((Na + A - 1) / A);
  if (block < nfb) {
    const int ab = block / cols;
    int goff;
    const int j = column_lane(goff);
    g = (block - ab * cols) * C + goff;
    return ab * A + j;
  }
  const int tail = groups - cols * C;
  const int lin = (block - nfb) * 256 + threadIdx.y * 64 + threadIdx.x;
  const int a = lin / tail;
  g = cols * C + (lin - a * tail);
  return a;
}

// Blocks of a bin-group forward launch over one slice (group_lane).
inline int64_t group_blocks(int Ns, int Na, int K) {
  const int64_t groups = (Ns + K - 1) / K, cols = groups / kColumnGroups;
  const int64_t tail = groups - cols * kColumnGroups;
  return cols * ((Na + kColumnAngles - 1) / kColumnAngles) + (Na * tail + 255) / 256;
}

// acc: this lane's slot 0 in an LDS array of Slots<K>::kCount rows of 256.
// CH > 1: the strip row is read CH pixels per load.  The last load of a row
// may reach up to CH - 1 pixels past the strip: inside the row those lie
// outside the union strip, so their hat weights land in the discarded slots
// (0, 1, K + 2, K + 3); past the end of the row they are the copy's zero
// padding (the single-problem launches pad every row by 4 elements).
template <typename real, int K, bool BANDED = false, int CH = 1, bool BUFW = false>
__device__ __forceinline__ void fan_fwd_group(const Job<real> &J, int s0, int a, int z, real *acc,
                                              Band<real> B = {}) {
  using A = Arith<real>;
#pragma unroll
  for (int k = 0; k < Slots<K>::kCount; ++k) acc[k * 256] = 0;
  const FanStrip<real> F = fan_strip(J, a, z, (real)(s0 - 1), (real)(s0 + K));
  if (!F.regular) {
    for (int k = 0; k < K; ++k) {
      const int s = s0 + k;
      if (s >= J.Ns) break;
      if (!BANDED)
        fan_store(J, s, a, z, fan_ray_sum(J, s, a, z));
      else  // the ray's rows in this band, as for the regular groups
        *band_slot(J, B, s, a, z) = fan_ray_sum(J, s, a, z, B.index, B.count);
    }
    return;
  }
  int ylo, yhi;
  fan_rows((double)F.qx, (double)F.qy, (double)F.dpx, (double)F.dmx, (double)F.midx,
           (double)F.midy, F.Nxx, F.Nyy, ylo, yhi);
  if (BANDED) band_rows(F.Nyy, B.index, B.count, ylo, yhi);
  // Strip bounds in pixel columns: the bounding rays cross row y at
  // qx + midx + dpx (y - midy - qy) and qx + midx + dmx (y - midy - qy),
  // clamped to the image (the fmax/fmin also keep infinities, rays parallel
  // to the rows, away from the int conversion).
  const real qxm = F.qx + F.midx, yq = F.midy + F.qy;
  const real xmax = (real)(F.Nxx - 1);
  // The detector position relative to bin s0 - 2 (the slot index) is
  // p = q/t with q = ss + (mid_det - s0 + 2)*t.  q is carried instead of ss
  // (|ss/t| is up to Ns/2, q/t lies in [1, K+2]), so single precision does
  // not lose the bin fraction to cancellation; each row starts from q at
  // pixel column 0 of row ylo (q00) plus the row and column terms, in double,
  // one FMA each (no running sum: q at column 0 is up to ~Ns in size, so
  // summing row steps would drift).  t is of order one with terms no larger,
  // so its row start is formed in the working precision.  p is positive, so
  // truncation is floor.
  const double c0 = (double)J.geo[5] - (double)s0 + 2.0;
  const double tdx = F.ddx, tdy = F.ddy;
  const double qdx = (double)F.dx + c0 * tdx, qdy = (double)F.dy + c0 * tdy;
  const real dq = (real)qdx;
  const real lo = (real)0.5, hi = (real)K + (real)2.5;
  const double ud0 = (double)ylo - (double)F.midy;
  const real t00 = (real)(tdy * ud0 + (double)F.RE - tdx * (double)F.midx);
  const double q00 = qdy * ud0 + c0 * (double)F.RE - qdx * (double)F.midx;
  // row y starts at F.base + y * F.row_stride, walked by one add per row
  const real *row = F.base + (int64_t)ylo * F.row_stride;
  // BUFW: the same walk by byte offsets from J.imgX
  [[maybe_unused]] __amdgpu_buffer_rsrc_t rs;
  [[maybe_unused]] int roff = 0, rstep = 0;
  if constexpr (BUFW) {
    rs = strip_buffer(J);
    roff = (int)((row - J.imgX) * (int64_t)sizeof(real));
    rstep = (int)(F.row_stride * (int64_t)sizeof(real));
  }
  for (int y = ylo; y <= yhi; ++y, row += F.row_stride, roff += rstep) {
    const real w = (real)y - yq;
    const int xh = (int)floor(fmax(fmin(xmax, fma(F.dpx, w, qxm)), (real)-1));
    const int xl = (int)ceil(fmin(fmax((real)0, fma(F.dmx, w, qxm)), (real)F.Nxx));
    if (xl > xh) continue;
    const int dy = y - ylo;
    real t = fma(F.ddx, (real)xl, fma(F.ddy, (real)dy, t00));
    real q = (real)fma(qdx, (double)xl, fma(qdy, (double)dy, q00));
    auto chunk = [&](const RowChunk<real, CH> &c) {
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const real r = A::rcp(t);
        const real p = A::clamp(q * r, lo, hi);
        deposit(acc, p, c.v[i] * r);
        t += F.ddx;
        q += dq;
      }
    };
    if constexpr (BUFW) {
      const int end = roff + xh * (int)sizeof(real);
      // two chunks per turn, each loaded one chunk ahead of its use; the
      // last load may read past the row (the next row, or zeros past the
      // buffer end) and is not used.  The scheduling barriers keep the
      // compiler from sinking each load below the work it should overlap.
      int off = roff + xl * (int)sizeof(real);
      RowChunk<real, CH> c0 = load_chunk_buf<real, CH>(rs, off);
      for (;;) {
        const RowChunk<real, CH> c1 = load_chunk_buf<real, CH>(rs, off + 16);
        __builtin_amdgcn_sched_barrier(0);
        chunk(c0);
        off += 16;
        if (off > end) break;
        c0 = load_chunk_buf<real, CH>(rs, off + 16);
        __builtin_amdgcn_sched_barrier(0);
        chunk(c1);
        off += 16;
        if (off > end) break;
      }
    } else {
      // walked by address (xl <= xh here): the end pointer is the loop bound
      const GPtr<real> end = (GPtr<real>)(row + xh);
      GPtr<real> px = (GPtr<real>)(row + xl);
      do {
        chunk(load_chunk<real, CH>(px));
        advance<real, CH>(px);
      } while (px <= end);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = s0 + k;
    if (s >= J.Ns) continue;
    const real sum = acc[(k + 2) * 256];
    if (BANDED)
      *band_slot(J, B, s, a, z) = sum;
    else
      fan_store(J, s, a, z, sum);
  }
}

// float(dx)^2 / float(ds), the reference's Radon scale (radon.cl:82-83, 159-160)
template <typename real>
__device__ __forceinline__ real radon_scale(const Job<real> &J) {
  const float delta_x = (float)J.geo[0];
  const float delta_xi = (float)J.geo[1];
  return (real)delta_x * (real)delta_x / (real)delta_xi;
}

// CH: pixels per load along the strip row, as in fan_fwd_group.
template <typename real, int K, bool BANDED = false, int CH = 1, bool BUFW = false>
__device__ __forceinline__ void radon_fwd_group(const Job<real> &J, int s0, int a, int z,
                                                real *acc, Band<real> B = {}) {
  using A = Arith<real>;
#pragma unroll
  for (int k = 0; k < Slots<K>::kCount; ++k) acc[k * 256] = 0;
  const real *o = J.ofs + 8 * (int64_t)a;
  real X = o[0], Y = o[1];
  const real off = o[2], inv = o[3];
  int Nxx = J.Nx, Nyy = J.Ny;
  const real *base;
  int64_t row_stride;
  if (o[5] != (real)0) {
    const real t = X;
    X = Y;
    Y = t;
    Nxx = J.Ny;
    Nyy = J.Nx;
    base = J.imgY + z * J.sliceY;
    row_stride = J.rowY;
  } else {
    base = J.imgX + z * J.sliceX;
    row_stride = J.rowX;
  }
  // union strip of the bins [s0, s0+K): centre s0 + (K-1)/2, half width (K+1)/2
  int ylo, yhi;
  radon_rows((double)X, (double)Y, (double)off - ((double)s0 + 0.5 * (K - 1)), 0.5 * (K + 1), Nxx,
             Nyy, ylo, yhi);
  if (BANDED) band_rows(Nyy, B.index, B.count, ylo, yhi);
  const real lo = (real)0.5, hi = (real)K + (real)2.5;
  // the strip's rays through s0 - 1 and s0 + K, ordered by the column they
  // reach first along a row (inv = 1/X carries the direction)
  const real s_first = inv < (real)0 ? (real)(s0 + K) : (real)(s0 - 1);
  const real s_last = inv < (real)0 ? (real)(s0 - 1) : (real)(s0 + K);
  const double p00 = (double)off - (double)s0 + 2.0;  // pixel (0, 0) relative to bin s0 - 2
  // row y starts at base + y * row_stride, walked by one add per row
  const real *row = base + (int64_t)ylo * row_stride;
  // BUFW: the same walk by byte offsets from J.imgX
  [[maybe_unused]] __amdgpu_buffer_rsrc_t rs;
  [[maybe_unused]] int roff = 0, rstep = 0;
  if constexpr (BUFW) {
    rs = strip_buffer(J);
    roff = (int)((row - J.imgX) * (int64_t)sizeof(real));
    rstep = (int)(row_stride * (int64_t)sizeof(real));
  }
  for (int y = ylo; y <= yhi; ++y, row += row_stride, roff += rstep) {
    const real d = (real)y * Y + off;  // detector position of (0, y)
    int xl = (int)((s_first - d) * inv);
    int xh = (int)((s_last - d) * inv);
    xl = xl < 0 ? 0 : xl;
    xh = xh > Nxx - 1 ? Nxx - 1 : xh;
    if (xl > xh) continue;
    // detector position of (xl, y) relative to bin s0 - 2 (the slot index),
    // evaluated in double: x*X and y*Y + off are up to Ns in size and nearly
    // cancel; inside the strip it lies in [1, K+2], so truncation is floor
    real pr = (real)fma((double)X, (double)xl, fma((double)Y, (double)y, p00));
    auto chunk = [&](const RowChunk<real, CH> &c) {
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        deposit(acc, A::clamp(pr, lo, hi), c.v[i]);
        pr += X;
      }
    };
    if constexpr (BUFW) {
      const int end = roff + xh * (int)sizeof(real);
      // two chunks per turn, each loaded one chunk ahead of its use; the
      // last load may read past the row (the next row, or zeros past the
      // buffer end) and is not used.  The scheduling barriers keep the
      // compiler from sinking each load below the work it should overlap.
      int o = roff + xl * (int)sizeof(real);
      RowChunk<real, CH> c0 = load_chunk_buf<real, CH>(rs, o);
      for (;;) {
        const RowChunk<real, CH> c1 = load_chunk_buf<real, CH>(rs, o + 16);
        __builtin_amdgcn_sched_barrier(0);
        chunk(c0);
        o += 16;
        if (o > end) break;
        c0 = load_chunk_buf<real, CH>(rs, o + 16);
        __builtin_amdgcn_sched_barrier(0);
        chunk(c1);
        o += 16;
        if (o > end) break;
      }
    } else {
      const GPtr<real> end = (GPtr<real>)(row + xh);
      GPtr<real> px = (GPtr<real>)(row + xl);
      do {
        chunk(load_chunk<real, CH>(px));
        advance<real, CH>(px);
      } while (px <= end);
    }
  }
  const real scale = radon_scale(J);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = s0 + k;
    if (s >= J.Ns) continue;
    const real sum = acc[(k + 2) * 256];
    if (BANDED)
      *band_slot(J, B, s, a, z) = sum;
    else
      store_fwd(J, s, a, z, sum * scale);
  }
}

// Sum of the row bands of bin (s, a, z) in band order, then the per-ray scale
// and epilogue of the unbanded bodies.
template <typename real>
__device__ __forceinline__ void band_reduce_entry(const Job<real> &J, const real *__restrict__ part,
                                                  int count, int s, int a, int z,
                                                  const int2 *__restrict__ range = nullptr,
                                                  int K = 1) {
  const int64_t stride = (int64_t)J.Nz * J.Na * J.Ns;
  const real *p = part + ((int64_t)z * J.Na + a) * J.Ns + s;
  if (range) {
    // compacted launch: band b wrote the groups range[b * Na + a] only; the
    // others are exact zeros there (no strip pixel in the band), added as
    // such so that the sum is the uncompacted one bit for bit
    const int g = s / K;
    real acc = 0;
    for (int b = 0; b < count; ++b) {
      const int2 r = range[(int64_t)b * J.Na + a];
      const real v = (g >= r.x && g < r.y) ? p[b * stride] : (real)0;
      acc = b == 0 ? v : acc + v;
    }
    if (J.fan)
      fan_store(J, s, a, z, acc);
    else
      store_fwd(J, s, a, z, acc * radon_scale(J));
    return;
  }
  real acc = p[0];
  for (int b = 1; b < count; ++b) acc += p[b * stride];
  if (J.fan)
    fan_store(J, s, a, z, acc);
  else
    store_fwd(J, s, a, z, acc * radon_scale(J));
}

// ---------------------------------------------------------------------------
// Backprojection prologue and body.
// ---------------------------------------------------------------------------
// Per-angle affine table: radon p = c0 x + c1 y + c2 (with c2 = off + 2, the
// detector coordinate in padded-sinogram indices); fan t = c0 x + c1 y + c2,
// ss = c3 x + c4 y + c5 (t = dd.(P - q), ss = d.(P - q), P = (x,y) - mid,
// dd = (d0 - q)/R^2, d = (xd, yd)/dxi^2; fanbeam.cl:339-355).
template <typename real>
__device__ __forceinline__ void back_table_entry(AngleRec<real> *__restrict__ tab,
                                                 const real *__restrict__ ofs,
                                                 const real *__restrict__ geo, int fan, int a) {
  const real *o = ofs + 8 * (int64_t)a;
  AngleRec<real> rec;
  for (int k = 0; k < 8; ++k) rec.c[k] = 0;
  if (!fan) {
    rec.c[0] = o[0];
    rec.c[1] = o[1];
    rec.c[2] = o[2] + (real)2;
  } else {
    const double R = geo[0], dxi = geo[2], midx = geo[3], midy = geo[4];
    const double qx = o[2], qy = o[3];
    const double ddx = ((double)o[4] - qx) / (R * R), ddy = ((double)o[5] - qy) / (R * R);
    const double sx = (double)o[0] / (dxi * dxi), sy = (double)o[1] / (dxi * dxi);
    const double cx = midx + qx, cy = midy + qy;
    rec.c[0] = (real)ddx;
    rec.c[1] = (real)ddy;
    rec.c[2] = (real)(-(ddx * cx + ddy * cy));
    rec.c[3] = (real)sx;
    rec.c[4] = (real)sy;
    rec.c[5] = (real)(-(sx * cx + sy * cy));
  }
  tab[a] = rec;
}

// S[(z*Na + a)*(Ns+4) + j], j = s + 2: sino[s,a,z] * weight[a] (* sdpd[s]/R); zeros at both ends.
template <typename real>
__device__ __forceinline__ void stage_entry(real *__restrict__ S, const real *__restrict__ sino,
                                            Strides3 si, const real *__restrict__ ofs,
                                            const real *__restrict__ sdpd,
                                            const real *__restrict__ geo, int fan, int Ns, int Na,
                                            int j, int a, int z) {
  const int Nsp = Ns + 4;
  real v = 0;
  const int s = j - 2;
  if (s >= 0 && s < Ns) {
    const real x = sino[s * si.s0 + a * si.s1 + z * si.s2];
    if (!fan)
      v = x * ofs[8 * (int64_t)a + 4];
    else
      v = x * sdpd[s] * (ofs[8 * (int64_t)a + 6] / geo[0]);
  }
  S[((int64_t)z * Na + a) * Nsp + j] = v;
}

// RY pixels (x, y0 .. y0+RY-1) of slice z.  The detector coordinate p is
// taken in padded-sinogram indices (bin s at s + 2) and clamped to
// [0.5, Ns + 2.5], where the two taps floor(p), floor(p) + 1 stay inside the
// padded row and truncation equals floor.
//
// BUF: the two taps are read with one buffer load through a descriptor of
// the slice's padded sinogram (per-angle row offset in a scalar register,
// 32-bit per-lane offset), which saves the 64-bit address arithmetic; needs
// Na * (Ns + 4) * sizeof(real) < 2^31 (buffer_fits).  Offsets outside the
// descriptor would read 0 rather than fault; the clamp keeps them inside.
template <typename real>
struct Taps {
  real v0, v1;
};

template <typename real>
__device__ __forceinline__ Taps<real> buffer_taps(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  Taps<real> t;
  if constexpr (sizeof(real) == 4) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
    t.v0 = v.x;
    t.v1 = v.y;
  } else {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0));
    t.v0 = v.x;
    t.v1 = v.y;
  }
  return t;
}

// One element at byte offset voff of a raw buffer (zero outside it).
template <typename real>
__device__ __forceinline__ real buffer_elem(__amdgpu_buffer_rsrc_t rs, int voff) {
  if constexpr (sizeof(real) == 4)
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0));
  else
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 0, 0));
}

inline bool buffer_fits(size_t elem, int Ns, int Na) {
  return (size_t)Na * (size_t)(Ns + 4) * elem < ((size_t)1 << 31);
}

// Angles [a0, a1) only; with part != nullptr the raw sums go to
// part[((band*Nz + z)*Ny + y)*Nx + x] (band_back_reduce_entry adds the bands
// in order and applies the epilogue), otherwise they are stored.
template <typename real, bool FAN, int RY, bool BUF = false>
__device__ __forceinline__ void back_pixels(const Job<real> &J, int x, int y0, int z, int a0 = 0,
                                            int a1 = -1, real *part = nullptr, int band = 0) {
  using A = Arith<real>;
  if (a1 < 0) a1 = J.Na;
  const real xr = (real)x, yr = (real)y0;
  const int Ns = J.Ns, Na = J.Na;
  const int Nsp = Ns + 4;
  const real *Sz = J.S + (int64_t)z * Na * Nsp;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)Sz, (short)0, BUF ? (int)((size_t)Na * Nsp * sizeof(real)) : 0, 0x00020000);
  const AngleRec<real> *__restrict__ tab = J.tab;
  const real lo = (real)0.5, hi = (real)Ns + (real)2.5;
  const real mid_det = FAN ? J.geo[5] + (real)2 : (real)0;
  real acc[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) acc[r] = 0;
#pragma unroll 2
  for (int a = a0; a < a1; ++a) {
    const AngleRec<real> &c = tab[a];
    const real *row = Sz + (int64_t)a * Nsp;
    // row r of the lane: one FMA from row 0 (no running sum, which would add
    // a rounding of a value ~Ns in size per row)
    const real t0 = FAN ? fma(c.c[0], xr, fma(c.c[1], yr, c.c[2])) : (real)1;
    const real q0 = FAN ? fma(c.c[3], xr, fma(c.c[4], yr, c.c[5])) : fma(c.c[0], xr, fma(c.c[1], yr, c.c[2]));
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const real t = FAN ? (r == 0 ? t0 : fma(c.c[1], (real)r, t0)) : (real)1;
      const real q = r == 0 ? q0 : fma(FAN ? c.c[4] : c.c[1], (real)r, q0);
      real p, w = 1;
      if (FAN) {
        w = A::rcp(t);
        p = fma(q, w, mid_det);
      } else {
        p = q;
      }
      p = A::clamp(p, lo, hi);
      const real f = A::fract(p);
      real v0, v1;
      if (BUF) {
        const Taps<real> tp = buffer_taps<real>(rs, (int)p * (int)sizeof(real),
                                                a * Nsp * (int)sizeof(real));
        v0 = tp.v0;
        v1 = tp.v1;
      } else {
        const real *tap = row + (int)p;
        v0 = tap[0];
        v1 = tap[1];
      }
      const real v = v0 + f * (v1 - v0);
      acc[r] = FAN ? acc[r] + v * w : acc[r] + v;
    }
  }
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int y = y0 + r;
    if (y >= J.Ny) break;
    if (part)
      part[(((int64_t)band * J.Nz + z) * J.Ny + y) * J.Nx + x] = acc[r];
    else
      store_back(J, x, y, z, acc[r]);
  }
}

// Sum of the angle bands of pixel (x, y, z) in band order, then the epilogue.
template <typename real>
__device__ __forceinline__ void band_back_reduce_entry(const Job<real> &J,
                                                       const real *__restrict__ part, int count,
                                                       int x, int y, int z) {
  const int64_t stride = (int64_t)J.Nz * J.Ny * J.Nx;
  const real *p = part + ((int64_t)z * J.Ny + y) * J.Nx + x;
  real acc = p[0];
  for (int b = 1; b < count; ++b) acc += p[b * stride];
  store_back(J, x, y, z, acc);
}

// ---------------------------------------------------------------------------
// LDS-staged backprojection.  A block's pixel tile reaches, for each angle,
// only a short window of the padded sinogram row; the block copies the
// windows of a chunk of angles into LDS with coalesced loads and reads both
// taps of every pixel from there, so the per-pixel gathers leave the vector
// memory pipeline (whose address unit limits the unstaged kernel).
// ---------------------------------------------------------------------------
// LDS elements per staged window: 16 windows of 304 floats and their
// records take 20032 bytes, so 8 blocks fit a CU's 160 KB
constexpr int kStageWidth = 304;

// Window of one angle for one tile: padded-sinogram indices [w0, w0 + n),
// zero outside [0, Ns + 4).  Coordinates are taken relative to the tile's
// corner pixel (xa, ya), with the corner's values formed in double: for a
// lane at (xa + dx, ya + dy)
//   radon  p = c0 dx + c1 dy + k0              (k0 = p(xa, ya) - w0)
//   fan    t = c0 dx + c1 dy + k0,  q = k3 dx + k4 dy + k1,  p = q / t + k2
// where, for a staged window, q = ss + (mid_det + 2 - w0) t is the window
// coordinate times t (k0 = t(xa, ya), k1 = q(xa, ya), k3 = c3 + (mid_det + 2
// - w0) c0, k4 likewise, k2 = 0).  Every term is then of the size of the
// window (a few hundred at most) instead of the detector (up to Ns), so
// single precision keeps ~2^-24 of the window, not of Ns.  Every pixel of
// the tile has 1 <= p < n - 2 (to rounding), so the taps floor(p), floor(p)
// + 1 need no clamp: where the unstaged kernel clamps p into [0.5, Ns + 2.5],
// both windows hold zeros.  n = 0 marks an angle read from memory instead
// (window wider than kStageWidth, or a fan tile reaching t <= 0); its
// constants give p in padded-sinogram indices (fan: q = ss, k2 = mid_det +
// 2), clamped to [lo, hi].
template <typename real>
struct StageWindow {
  real k0, k1, k2, k3, k4, lo, hi;
  int w0, n;
};

// The detector coordinate is a linear (radon) or linear-fractional (fan, with
// t > 0 on the tile) function of the pixel, so its extremes over the tile
// [xa, xb] x [ya, yb] (the pixels its lanes evaluate) lie at the corners.  One
// index of margin on each side covers the rounding of the per-pixel evaluation.
template <typename real, bool FAN>
__device__ __forceinline__ StageWindow<real> stage_window(const Job<real> &J, int a, int xa, int xb,
                                                          int ya, int yb) {
  const AngleRec<real> &c = J.tab[a];
  const double mid = FAN ? (double)J.geo[5] + 2.0 : 0.0;
  double pmin = 0, pmax = 0;
  bool ok = true;
  for (int k = 0; k < 4; ++k) {
    const double x = (k & 1) ? xb : xa, y = (k & 2) ? yb : ya;
    double p;
    if (FAN) {
      const double t = (double)c.c[0] * x + (double)c.c[1] * y + (double)c.c[2];
      const double ss = (double)c.c[3] * x + (double)c.c[4] * y + (double)c.c[5];
      ok = ok && t > 0;
      p = ss / t + mid;
    } else {
      p = (double)c.c[0] * x + (double)c.c[1] * y + (double)c.c[2];
    }
    pmin = k == 0 ? p : fmin(pmin, p);
    pmax = k == 0 ? p : fmax(pmax, p);
  }
  StageWindow<real> W;
  const double span = pmax - pmin;
  // |pmin| < 1e8 keeps the staging copy's byte offsets sizeof(real) * (w0 + j)
  // within int
  ok = ok && span == span && span < kStageWidth - 4 && fabs(pmin) < 1e8;
  const int w0 = ok ? (int)floor(pmin) - 1 : 0;
  W.w0 = w0;
  W.n = ok ? (int)floor(pmax) + 3 - w0 : 0;
  const double xc = xa, yc = ya;
  if (FAN) {
    const double tc = (double)c.c[0] * xc + (double)c.c[1] * yc + (double)c.c[2];
    const double sc = (double)c.c[3] * xc + (double)c.c[4] * yc + (double)c.c[5];
    const double k = ok ? mid - w0 : 0.0;  // window shift folded into q
    W.k0 = (real)tc;
    W.k1 = (real)(sc + k * tc);
    W.k3 = (real)((double)c.c[3] + k * (double)c.c[0]);
    W.k4 = (real)((double)c.c[4] + k * (double)c.c[1]);
    W.k2 = (real)(ok ? 0.0 : mid);
  } else {
    const double pc = (double)c.c[0] * xc + (double)c.c[1] * yc + (double)c.c[2];
    W.k0 = (real)(ok ? pc - w0 : pc);
    W.k1 = W.k2 = W.k3 = W.k4 = (real)0;
  }
  W.lo = (real)0.5;
  W.hi = (real)J.Ns + (real)2.5;
  return W;
}

// Adds one angle's contribution to the RY pixels (xa + dx, ya + dy + r) of
// this lane.  STAGED: `row` is the angle's LDS window (no clamp); otherwise
// the padded sinogram row in memory (clamped as in the unstaged kernel).
// Row r of the lane is one FMA from row 0, as in back_pixels (row 0 is t0, q0
// themselves: fma(c, 0, t0) is not folded, IEEE zeros and NaNs).
template <typename real, bool FAN, int RY, bool STAGED>
__device__ __forceinline__ void staged_angle(const AngleRec<real> &c, const StageWindow<real> &W,
                                             const real *row, real dx, real dy, real *acc) {
  using A = Arith<real>;
  const real t0 = FAN ? fma(c.c[0], dx, fma(c.c[1], dy, W.k0)) : (real)1;
  const real q0 = FAN ? fma(W.k3, dx, fma(W.k4, dy, W.k1)) : fma(c.c[0], dx, fma(c.c[1], dy, W.k0));
  if constexpr (STAGED) {
    // In three passes: every row's tap index, then all 2 RY LDS taps, then
    // the interpolation.  The taps of all rows are then in flight together
    // instead of one row's at a time.  The LDS-qualified pointer keeps the
    // compiler from merging these taps with the memory path's global reads
    // into generic (flat) loads, which would wait on global memory.
    real wr[RY], fr[RY], v0[RY], v1[RY];
    int ir[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const real t = FAN ? (r == 0 ? t0 : fma(c.c[1], (real)r, t0)) : (real)1;
      const real q = r == 0 ? q0 : fma(FAN ? W.k4 : c.c[1], (real)r, q0);
      real p = q;
      wr[r] = 1;
      if (FAN) {
        wr[r] = A::rcp(t);
        p = q * wr[r];
      }
      fr[r] = A::fract(p);
      ir[r] = (int)p;
    }
    const __attribute__((address_space(3))) real *win =
        (const __attribute__((address_space(3))) real *)row;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      v0[r] = win[ir[r]];
      v1[r] = win[ir[r] + 1];
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const real v = v0[r] + fr[r] * (v1[r] - v0[r]);
      acc[r] = FAN ? acc[r] + v * wr[r] : acc[r] + v;
    }
  } else {
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const real t = FAN ? (r == 0 ? t0 : fma(c.c[1], (real)r, t0)) : (real)1;
      const real q = r == 0 ? q0 : fma(FAN ? W.k4 : c.c[1], (real)r, q0);
      real p = q, w = 1;
      if (FAN) {
        w = A::rcp(t);
        p = fma(q, w, W.k2);
      }
      p = A::clamp(p, W.lo, W.hi);
      const real f = A::fract(p);
      const real *tap = row + (int)p;
      const real v = tap[0] + f * (tap[1] - tap[0]);
      acc[r] = FAN ? acc[r] + v * w : acc[r] + v;
    }
  }
}

// Occupancy target of the whole-grid kernels built on back_staged_tile
// (back_staged_kernel, batch_back_staged_kernel; not the angle-banded form,
// whose grids are too small to fill it): one block's 16 angle windows
// (kStageWidth) take 20 KB of LDS, room for 8 blocks (8 waves per SIMD) per
// CU.  Left to itself the compiler gives the fanbeam tile 94
// VGPRs, which allow 5; held to 8 waves per SIMD (__launch_bounds__(256, 8))
// it fits in 64, with a few registers spilled outside the angle loop.  On the
// 32 x 2048^2 stacks the backprojection goes from 51.9 to 48.4 ms (fanbeam,
// 7 waves with 320-element windows) and 48.6 (8 waves); parallel beam 32.7
// -> 31.9 ms.
constexpr int kStagedWavesPerSimd = 8;

// LDS-staged backprojection of one tile: the block (256 threads) owns the TX x
// (256/TX)*RY pixels starting at (bx*TX, by*(256/TX)*RY) of slice z; per chunk
// of CA angles it computes each angle's window (stage_window), copies it into
// `rows` (CA * kStageWidth elements of LDS; zeros outside the padded sinogram),
// then every lane adds the chunk's angles to its RY pixels (staged_angle).
// Lanes outside the image take part in the copies and barriers only.  Angles
// [a_begin, a_end) only; with part != nullptr the raw sums go to
// part[((band*Nz + z)*Ny + y)*Nx + x] (band_back_reduce_entry adds the bands in
// order and applies the epilogue), otherwise they are stored.
template <typename real, bool FAN, int RY, int CA, int TX>
__device__ __forceinline__ void back_staged_tile(const Job<real> &J, int bx, int by, int z,
                                                 real *rows, StageWindow<real> *win, int a_begin = 0,
                                                 int a_end = -1, real *part = nullptr,
                                                 int band = 0) {
  static_assert(CA % 4 == 0 && CA <= 64, "each of the 4 waves copies CA / 4 windows");
  if (a_end < 0) a_end = J.Na;
  constexpr int TY = 256 / TX;  // lane rows
  const int tid = threadIdx.x;
  const int x = bx * TX + (tid % TX);
  const int y0 = (by * TY + tid / TX) * RY;
  const bool active = x < J.Nx && y0 < J.Ny;
  const int xa = bx * TX, xb = min(xa + TX - 1, J.Nx - 1);
  // rows evaluated: a lane whose first row is inside the image evaluates all RY
  const int ya = by * TY * RY, yb = min(ya + TY * RY - 1, J.Ny + RY - 2);
  const int Nsp = J.Ns + 4;
  const real *Sz = J.S + (int64_t)z * J.Na * Nsp;
  // the lane's first pixel relative to the tile corner (xa, ya)
  const real dx = (real)(tid % TX), dy = (real)((tid / TX) * RY);
  real acc[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) acc[r] = 0;
  for (int a0 = a_begin; a0 < a_end; a0 += CA) {
    const int na = min(CA, a_end - a0);
    if (tid < na) win[tid] = stage_window<real, FAN>(J, a0 + tid, xa, xb, ya, yb);
    __syncthreads();
    // every angle of the chunk staged (a wave-uniform vote): the accumulation
    // loop below then runs without the per-angle branch to the memory path
    const int lane = tid & 63;
    const bool all_staged = __ballot(lane < na && win[lane < na ? lane : 0].n == 0) == 0;
    // Wave w copies windows w, w + 4, ...  A buffer descriptor over each
    // padded row makes every index outside [0, Nsp) read zero (a negative
    // offset is a large unsigned one), with no test per element.  Each step
    // loads one element of each of the wave's windows before storing any, so
    // those loads are in flight together.  The wave index goes through
    // readfirstlane so the compiler knows it is uniform and keeps the
    // descriptors in scalar registers.  The range check takes the register
    // offset plus the instruction's offset field without 32-bit wrap-around:
    // the empty asm on each offset keeps the compiler from moving a constant
    // part of it into that field (-4 + 256 would then read as out of range
    // instead of as element 63).  The empty asm on each loaded value keeps a
    // load from being sunk into its conditional store.  Neither asm is
    // volatile: a volatile asm counts as a possible memory write, and the
    // angle table's loads after it would no longer be scalar.
    {
      constexpr int Q = CA / 4;  // windows per wave
      const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
      __amdgpu_buffer_rsrc_t rs[Q];
      int n[Q], w0[Q], nmax = 0;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int k = wv + 4 * q;
        const bool has = k < na;
        n[q] = has ? win[k].n : 0;
        w0[q] = has ? win[k].w0 : 0;
        rs[q] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(Sz + (int64_t)(a0 + (has ? k : 0)) * Nsp), (short)0, Nsp * (int)sizeof(real),
            0x00020000);
        nmax = max(nmax, n[q]);
      }
      for (int base = 0; base < nmax; base += 64) {
        const int j = base + lane;
        real v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          int off = (w0[q] + j) * (int)sizeof(real);
          asm("" : "+v"(off));
          v[q] = buffer_elem<real>(rs[q], off);
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          asm("" : "+v"(v[q]));
          if (j < n[q]) rows[(wv + 4 * q) * kStageWidth + j] = v[q];
        }
      }
    }
    __syncthreads();
    if (active && all_staged) {
      for (int k = 0; k < na; ++k)
        staged_angle<real, FAN, RY, true>(J.tab[a0 + k], win[k], rows + k * kStageWidth, dx, dy,
                                          acc);
    } else if (active) {
      for (int k = 0; k < na; ++k) {
        const StageWindow<real> W = win[k];
        const AngleRec<real> &c = J.tab[a0 + k];
        if (W.n)
          staged_angle<real, FAN, RY, true>(c, W, rows + k * kStageWidth, dx, dy, acc);
        else
          staged_angle<real, FAN, RY, false>(c, W, Sz + (int64_t)(a0 + k) * Nsp, dx, dy, acc);
      }
    }
    __syncthreads();
  }
  if (!active) return;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int y = y0 + r;
    if (y >= J.Ny) break;
    if (part)
      part[(((int64_t)band * J.Nz + z) * J.Ny + y) * J.Nx + x] = acc[r];
    else
      store_back(J, x, y, z, acc[r]);
  }
}

}  // namespace gt
