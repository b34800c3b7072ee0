// Parallel-beam and fanbeam forward/backprojection for gfx950.
//
// Forward (ray-driven gather, one lane per detector bin s of angle a):
//   sino[s,a,z] = sum over the pixels of the strip between the rays through
//   detector positions s-1 and s+1 of hat-weight * img, exactly the
//   quadrature of gratopy/radon.cl:66-161 and gratopy/fanbeam.cl:72-256.
//   A wave holds 64 consecutive bins of one angle, so each image row it
//   visits is read as a short contiguous run.  Rays that are closer to the
//   x axis walk the image by columns; for them the kernel reads a y-fastest
//   copy of the image made by a tiled transpose, so every wave always reads
//   along the fastest axis.  Rows that cannot touch the image are skipped
//   (closed-form row range per ray).
//
// Backward (pixel-driven gather, one lane per pixel):
//   img[x,y,z] = sum_a lerp(S_a, xi_a(x,y)) [ * 1/t_a(x,y) for fan ]
//   where S is a zero-padded copy of the sinogram already multiplied by the
//   angle weight (and by sdpd[s]/R for fanbeam), so the inner loop has no
//   bounds tests and no divergent table loads (gratopy/radon.cl:183-241,
//   gratopy/fanbeam.cl:290-391).  Per-angle affine coefficients come from a
//   table built once per call and are read with scalar loads.

#include "common.h"

namespace gt {

template <typename real>
struct alignas(sizeof(real) * 8) AngleRec {
  real c[8];
};

// ---------------------------------------------------------------------------
// Tiled copy between two orders of an N0 x N1 x N2 array (32x32 tiles in LDS).
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void relayout_kernel(real *__restrict__ dst, Strides3 ds,
                                                       const real *__restrict__ src, Strides3 ss,
                                                       int N0, int N1, int read0fast,
                                                       int write0fast) {
  __shared__ real tile[32][33];
  const int b0 = blockIdx.x * 32, b1 = blockIdx.y * 32;
  const int64_t z = blockIdx.z;
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
}

template <typename real>
static void launch_relayout(hipStream_t st, real *dst, Strides3 ds, const real *src, Strides3 ss,
                            int N0, int N1, int N2) {
  const int read0fast = ss.s0 <= ss.s1;
  const int write0fast = ds.s0 <= ds.s1;
  dim3 grid((N0 + 31) / 32, (N1 + 31) / 32, N2);
  hipLaunchKernelGGL(relayout_kernel<real>, grid, dim3(256), 0, st, dst, ds, src, ss, N0, N1,
                     read0fast, write0fast);
}

// ---------------------------------------------------------------------------
// Row ranges.  Both return a conservative [ylo, yhi]: rows outside it have no
// pixel inside the ray's strip (one row and one column of margin are added).
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

// Parallel ray (axes already swapped): pixels with |x*X + y*Y + c| < 1 and
// 0 <= x <= Nxx-1.  The strip centre is g(y) = -(c + y*Y)/X with half width 1/|X|.
__device__ inline void radon_rows(double X, double Y, double c, int Nxx, int Nyy, int &ylo,
                                  int &yhi) {
  const double h = 1.0 / fabs(X);
  const double g0 = -c / X, m = -Y / X;
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

// Fan ray (axes already swapped): row y is visited on [xlow(y), xhigh(y)] with
// xhigh = qx + dpx*u, xlow = qx + dmx*u, u = y - midy - qy.
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

// ---------------------------------------------------------------------------
// Forward projections.  Block = 64 bins x 4 angles; grid z = slice.
// imgX[x + Nx*(y + Ny*z)], imgY[y + Ny*(x + Nx*z)].
// ---------------------------------------------------------------------------
template <typename real>
__global__ __launch_bounds__(256) void radon_fwd_kernel(real *__restrict__ sino, Strides3 so,
                                                        const real *__restrict__ imgX,
                                                        const real *__restrict__ imgY,
                                                        const real *__restrict__ ofs,
                                                        const real *__restrict__ geo, int Nx,
                                                        int Ny, int Ns, int Na) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (s >= Ns || a >= Na) return;
  // the reference reads the scales as float even in double kernels (radon.cl:82-83)
  const float delta_x = (float)geo[0];
  const float delta_xi = (float)geo[1];
  const real *o = ofs + 8 * (int64_t)a;
  real X = o[0], Y = o[1];
  const real off = o[2], inv = o[3];
  int Nxx = Nx, Nyy = Ny;
  const real *base;
  int64_t row_stride;
  const int64_t slice = (int64_t)z * Nx * Ny;
  if (o[5] != (real)0) {
    const real t = X;
    X = Y;
    Y = t;
    Nxx = Ny;
    Nyy = Nx;
    base = imgY + slice;
    row_stride = Ny;
  } else {
    base = imgX + slice;
    row_stride = Nx;
  }
  int ylo, yhi;
  radon_rows((double)X, (double)Y, (double)off - (double)s, Nxx, Nyy, ylo, yhi);
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
  sino[s * so.s0 + a * so.s1 + z * so.s2] =
      acc * (real)delta_x * (real)delta_x / (real)delta_xi;
}

template <typename real>
__global__ __launch_bounds__(256) void fan_fwd_kernel(real *__restrict__ sino, Strides3 so,
                                                      const real *__restrict__ imgX,
                                                      const real *__restrict__ imgY,
                                                      const real *__restrict__ ofs,
                                                      const real *__restrict__ sdpd,
                                                      const real *__restrict__ geo, int Nx,
                                                      int Ny, int Ns, int Na) {
  using A = Arith<real>;
  const int s = blockIdx.x * 64 + threadIdx.x;
  const int a = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (s >= Ns || a >= Na) return;
  const real R = geo[0];
  real RE = geo[1];
  const real delta_xi = geo[2];
  real midx = geo[3], midy = geo[4];
  const real mid_det = geo[5];
  const real delta_x = geo[10];
  const real *o = ofs + 8 * (int64_t)a;
  real dx = o[0], dy = o[1], qx = o[2], qy = o[3];
  const real d0x = o[4], d0y = o[5];
  // directions from the source to detector positions s+1 and s-1
  const real sp = -mid_det + (real)s + (real)1;
  const real sm = -mid_det + (real)s - (real)1;
  real dpx = d0x + dx * sp - qx, dpy = d0y + dy * sp - qy;
  real dmx = d0x + dx * sm - qx, dmy = d0y + dy * sm - qy;
  real ddx = d0x - qx, ddy = d0y - qy;
  const real RR = R * R;
  ddx /= RR;
  ddy /= RR;
  const real dxi2 = delta_xi * delta_xi;
  dx /= dxi2;
  dy /= dxi2;
  RE *= (real)1 / R;
  int Nxx = Nx, Nyy = Ny;
  const real *base;
  int64_t row_stride;
  const int64_t slice = (int64_t)z * Nx * Ny;
  if (fabs(dpx) < fabs(dpy) && fabs(dmx) < fabs(dmy)) {
    base = imgX + slice;
    row_stride = Nx;
  } else {
    real t;
    t = ddx; ddx = ddy; ddy = t;
    t = dx; dx = dy; dy = t;
    t = qx; qx = qy; qy = t;
    t = dpx; dpx = dpy; dpy = t;
    t = dmx; dmx = dmy; dmy = t;
    t = midx; midx = midy; midy = t;
    Nyy = Nx;
    Nxx = Ny;
    base = imgY + slice;
    row_stride = Ny;
  }
  dpx /= dpy;
  dmx /= dmy;
  if (qy * (dmx - dpx) < (real)0) {
    const real t = dpx;
    dpx = dmx;
    dmx = t;
  }
  const real smd = (real)s - mid_det;
  int ylo, yhi;
  fan_rows((double)qx, (double)qy, (double)dpx, (double)dmx, (double)midx, (double)midy, Nxx,
           Nyy, ylo, yhi);
  const real xcap = (real)(Nxx - 1) - midx;
  real acc = 0;
  for (int y = ylo; y <= yhi; ++y) {
    const real u = (real)y - midy;
    const real xhigh = qx + dpx * (u - qy);
    const real xlow = qx + dmx * (u - qy);
    // bounds of the strip in this row; the outer fmax/fmin only keep
    // infinities (rays parallel to the rows) away from the int conversion
    const int xh = (int)floor(fmax(fmin(xcap, xhigh) + midx, (real)-1));
    const int xl = (int)ceil(fmin(fmax(-midx, xlow) + midx, (real)Nxx));
    const real xr = (real)xl - midx;
    real t = ddx * xr + ddy * u + RE;
    real ss = dx * xr + dy * u;
    const real *row = base + (int64_t)y * row_stride;
    for (int x = xl; x <= xh; ++x) {
      const real r = A::rcp(t);
      const real xi = ss * r;
      const real w = ((real)1 - fabs(smd - xi)) * r;
      acc += w * row[x];
      t += ddx;
      ss += dx;
    }
  }
  sino[s * so.s0 + a * so.s1 + z * so.s2] = acc * sdpd[s] / delta_xi * delta_x / R;
}

// ---------------------------------------------------------------------------
// Forward projections, bin-group form.  A lane owns K adjacent detector bins
// [s0, s0+K) of one angle and walks, row by row, the union strip between the
// rays through detector positions s0-1 and s0+K.  Each pixel of the strip is
// projected once; its two hat weights (1-f, f) go to the bins floor(p) and
// floor(p)+1 it touches, accumulated in lane-private LDS slots (slot k+2 holds
// bin s0+k; slots outside [2, K+2) collect the neighbours' share and are
// dropped).  Per bin, the summation order is the reference's (rows ascending,
// then along the row), so results agree with the per-ray loops to rounding;
// the per-row bound computation is shared by K rays and each pixel is
// projected once instead of twice.
// ---------------------------------------------------------------------------
template <int K>
struct Slots {
  static constexpr int kCount = K + 4;
};

template <typename real, int K>
__global__ __launch_bounds__(256) void fan_fwd_group_kernel(
    real *__restrict__ sino, Strides3 so, const real *__restrict__ imgX,
    const real *__restrict__ imgY, const real *__restrict__ ofs, const real *__restrict__ sdpd,
    const real *__restrict__ geo, int Nx, int Ny, int Ns, int Na) {
  using A = Arith<real>;
  __shared__ real slots[Slots<K>::kCount * 256];
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int s0 = (blockIdx.x * 64 + threadIdx.x) * K;
  const int a = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (s0 >= Ns || a >= Na) return;
  real *acc = slots + tid;
#pragma unroll
  for (int k = 0; k < Slots<K>::kCount; ++k) acc[k * 256] = 0;

  const real R = geo[0];
  real RE = geo[1];
  const real delta_xi = geo[2];
  real midx = geo[3], midy = geo[4];
  const real mid_det = geo[5];
  const real delta_x = geo[10];
  const real *o = ofs + 8 * (int64_t)a;
  real dx = o[0], dy = o[1], qx = o[2], qy = o[3];
  const real d0x = o[4], d0y = o[5];
  // rays through detector positions s0 + K (upper) and s0 - 1 (lower)
  const real sp = -mid_det + (real)(s0 + K);
  const real sm = -mid_det + (real)(s0 - 1);
  real dpx = d0x + dx * sp - qx, dpy = d0y + dy * sp - qy;
  real dmx = d0x + dx * sm - qx, dmy = d0y + dy * sm - qy;
  real ddx = d0x - qx, ddy = d0y - qy;
  const real RR = R * R;
  ddx /= RR;
  ddy /= RR;
  const real dxi2 = delta_xi * delta_xi;
  dx /= dxi2;
  dy /= dxi2;
  RE *= (real)1 / R;
  int Nxx = Nx, Nyy = Ny;
  const real *base;
  int64_t row_stride;
  const int64_t slice = (int64_t)z * Nx * Ny;
  if (fabs(dpx) < fabs(dpy) && fabs(dmx) < fabs(dmy)) {
    base = imgX + slice;
    row_stride = Nx;
  } else {
    real t;
    t = ddx; ddx = ddy; ddy = t;
    t = dx; dx = dy; dy = t;
    t = qx; qx = qy; qy = t;
    t = dpx; dpx = dpy; dpy = t;
    t = dmx; dmx = dmy; dmy = t;
    t = midx; midx = midy; midy = t;
    Nyy = Nx;
    Nxx = Ny;
    base = imgY + slice;
    row_stride = Ny;
  }
  dpx /= dpy;
  dmx /= dmy;
  if (qy * (dmx - dpx) < (real)0) {
    const real t = dpx;
    dpx = dmx;
    dmx = t;
  }
  int ylo, yhi;
  fan_rows((double)qx, (double)qy, (double)dpx, (double)dmx, (double)midx, (double)midy, Nxx,
           Nyy, ylo, yhi);
  const real xcap = (real)(Nxx - 1) - midx;
  const real c0 = mid_det - (real)s0;  // detector position relative to s0
  const real lo = (real)-1.5, hi = (real)K + (real)0.5;
  for (int y = ylo; y <= yhi; ++y) {
    const real u = (real)y - midy;
    const real xhigh = qx + dpx * (u - qy);
    const real xlow = qx + dmx * (u - qy);
    const int xh = (int)floor(fmax(fmin(xcap, xhigh) + midx, (real)-1));
    const int xl = (int)ceil(fmin(fmax(-midx, xlow) + midx, (real)Nxx));
    const real xr = (real)xl - midx;
    real t = ddx * xr + ddy * u + RE;
    real ss = dx * xr + dy * u;
    const real *row = base + (int64_t)y * row_stride;
    for (int x = xl; x <= xh; ++x) {
      const real r = A::rcp(t);
      const real p = A::clamp(ss * r + c0, lo, hi);
      const real fl = floor(p);
      const int j = (int)fl;
      const real f = p - fl;
      const real v = row[x] * r;
      real *slot = acc + (j + 2) * 256;
      slot[0] += v - f * v;
      slot[256] += f * v;
      t += ddx;
      ss += dx;
    }
  }
  const real scale_tail = delta_x / R;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = s0 + k;
    if (s < Ns) sino[s * so.s0 + a * so.s1 + z * so.s2] = acc[(k + 2) * 256] * sdpd[s] / delta_xi * scale_tail;
  }
}

template <typename real, int K>
__global__ __launch_bounds__(256) void radon_fwd_group_kernel(
    real *__restrict__ sino, Strides3 so, const real *__restrict__ imgX,
    const real *__restrict__ imgY, const real *__restrict__ ofs, const real *__restrict__ geo,
    int Nx, int Ny, int Ns, int Na) {
  using A = Arith<real>;
  __shared__ real slots[Slots<K>::kCount * 256];
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int s0 = (blockIdx.x * 64 + threadIdx.x) * K;
  const int a = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (s0 >= Ns || a >= Na) return;
  real *acc = slots + tid;
#pragma unroll
  for (int k = 0; k < Slots<K>::kCount; ++k) acc[k * 256] = 0;
  const float delta_x = (float)geo[0];
  const float delta_xi = (float)geo[1];
  const real *o = ofs + 8 * (int64_t)a;
  real X = o[0], Y = o[1];
  const real off = o[2], inv = o[3];
  int Nxx = Nx, Nyy = Ny;
  const real *base;
  int64_t row_stride;
  const int64_t slice = (int64_t)z * Nx * Ny;
  if (o[5] != (real)0) {
    const real t = X;
    X = Y;
    Y = t;
    Nxx = Ny;
    Nyy = Nx;
    base = imgY + slice;
    row_stride = Ny;
  } else {
    base = imgX + slice;
    row_stride = Nx;
  }
  // union strip of the bins [s0, s0+K): centre bin s0 + (K-1)/2, half width (K+1)/2
  int ylo, yhi;
  {
    const double half = 0.5 * (K + 1);
    const double Xd = X, Yd = Y;
    const double h = half / fabs(Xd);
    const double g0 = -((double)off - ((double)s0 + 0.5 * (K - 1))) / Xd, m = -Yd / Xd;
    double rlo = -INFINITY, rhi = INFINITY;
    restrict_ge(m, -1.0 - h - g0, rlo, rhi);
    restrict_ge(-m, g0 - ((double)Nxx + h), rlo, rhi);
    if (rlo > rhi) {
      ylo = 0;
      yhi = -1;
    } else {
      clip_rows(rlo, rhi, Nyy, ylo, yhi);
    }
  }
  const real lo = (real)-1.5, hi = (real)K + (real)0.5;
  const real sl = (real)(s0 - 1), sh = (real)(s0 + K);
  for (int y = ylo; y <= yhi; ++y) {
    const real d = (real)y * Y + off;  // detector position of (0, y)
    int xl = (int)((sl - d) * inv);
    int xh = (int)((sh - d) * inv);
    if (inv < (real)0) {
      const int t = xl;
      xl = xh;
      xh = t;
    }
    xl = xl < 0 ? 0 : xl;
    xh = xh > Nxx - 1 ? Nxx - 1 : xh;
    const real c = d - (real)s0;
    const real *row = base + (int64_t)y * row_stride;
    for (int x = xl; x <= xh; ++x) {
      const real p = A::clamp((real)x * X + c, lo, hi);
      const real fl = floor(p);
      const int j = (int)fl;
      const real f = p - fl;
      const real v = row[x];
      real *slot = acc + (j + 2) * 256;
      slot[0] += v - f * v;
      slot[256] += f * v;
    }
  }
  const real scale = (real)delta_x * (real)delta_x / (real)delta_xi;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = s0 + k;
    if (s < Ns) sino[s * so.s0 + a * so.s1 + z * so.s2] = acc[(k + 2) * 256] * scale;
  }
}

// ---------------------------------------------------------------------------
// Backprojection prologue: per-angle affine table and padded weighted sinogram.
// ---------------------------------------------------------------------------
template <typename real, bool FAN>
__global__ void back_table_kernel(AngleRec<real> *__restrict__ tab, const real *__restrict__ ofs,
                                  const real *__restrict__ geo, int Na) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= Na) return;
  const real *o = ofs + 8 * (int64_t)a;
  AngleRec<real> rec;
  for (int k = 0; k < 8; ++k) rec.c[k] = 0;
  if (!FAN) {
    rec.c[0] = o[0];
    rec.c[1] = o[1];
    rec.c[2] = o[2];
  } else {
    // t(x,y) = dd.(P - q) and ss(x,y) = d.(P - q), P = (x,y) - mid, with
    // dd = (d0 - q)/R^2 and d = (xd, yd)/dxi^2 (gratopy/fanbeam.cl:339-355)
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

// S[(z*Na + a)*(Ns+4) + 2 + s] = sino[s,a,z] * weight[a] (* sdpd[s]); two zeros on each side.
template <typename real, bool FAN>
__global__ __launch_bounds__(256) void stage_sino_kernel(real *__restrict__ S,
                                                         const real *__restrict__ sino,
                                                         Strides3 si, const real *__restrict__ ofs,
                                                         const real *__restrict__ sdpd,
                                                         const real *__restrict__ geo, int Ns,
                                                         int Na) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int a = blockIdx.y;
  const int z = blockIdx.z;
  const int Nsp = Ns + 4;
  if (j >= Nsp) return;
  real v = 0;
  const int s = j - 2;
  if (s >= 0 && s < Ns) {
    const real x = sino[s * si.s0 + a * si.s1 + z * si.s2];
    if (!FAN) {
      v = x * ofs[8 * (int64_t)a + 4];
    } else {
      v = x * sdpd[s] * (ofs[8 * (int64_t)a + 6] / geo[0]);
    }
  }
  S[((int64_t)z * Na + a) * Nsp + j] = v;
}

// One lane per RY pixels (x, y0..y0+RY-1); a block covers 16 x 16*RY pixels.
// Per angle the lane evaluates the affine coordinates once and advances them
// by one row step per pixel; the RY accumulators are independent chains.
template <typename real, bool FAN, int RY>
__global__ __launch_bounds__(256) void back_kernel(real *__restrict__ img, Strides3 io,
                                                   const real *__restrict__ S,
                                                   const AngleRec<real> *__restrict__ tab,
                                                   const real *__restrict__ geo, int Nx, int Ny,
                                                   int Ns, int Na) {
  using A = Arith<real>;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y0 = (blockIdx.y * 16 + (threadIdx.x >> 4)) * RY;
  const int z = blockIdx.z;
  if (x >= Nx || y0 >= Ny) return;
  const real xr = (real)x, yr = (real)y0;
  const int Nsp = Ns + 4;
  const real *Sz = S + (int64_t)z * Na * Nsp + 2;
  const real lo = (real)-1.5, hi = (real)Ns + (real)0.5;
  const real mid_det = FAN ? geo[5] : (real)0;
  real acc[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) acc[r] = 0;
#pragma unroll 2
  for (int a = 0; a < Na; ++a) {
    const AngleRec<real> &c = tab[a];
    const real *row = Sz + (int64_t)a * Nsp;
    real t = FAN ? c.c[0] * xr + c.c[1] * yr + c.c[2] : (real)1;
    real q = FAN ? c.c[3] * xr + c.c[4] * yr + c.c[5] : c.c[0] * xr + c.c[1] * yr + c.c[2];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      real p, w = 1;
      if (FAN) {
        w = A::rcp(t);
        p = q * w + mid_det;
      } else {
        p = q;
      }
      p = A::clamp(p, lo, hi);
      const real fl = floor(p);
      const int i = (int)fl;
      const real f = p - fl;
      const real v0 = row[i], v1 = row[i + 1];
      const real v = v0 + f * (v1 - v0);
      acc[r] = FAN ? acc[r] + v * w : acc[r] + v;
      if (FAN) {
        t += c.c[1];
        q += c.c[4];
      } else {
        q += c.c[1];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int y = y0 + r;
    if (y < Ny) img[x * io.s0 + y * io.s1 + z * io.s2] = acc[r];
  }
}

// ---------------------------------------------------------------------------
// Single-line (impulse response) kernels for the sparse system matrix.
// ---------------------------------------------------------------------------
template <typename real>
__global__ void single_line_radon_kernel(real *__restrict__ sino, Strides3 so, int px, int py,
                                         const real *__restrict__ ofs,
                                         const real *__restrict__ geo, int Nx, int Ny, int Ns,
                                         int Na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (s >= Ns) return;
  const float delta_x = (float)geo[0];
  const float delta_xi = (float)geo[1];
  const real *o = ofs + 8 * (int64_t)a;
  real X = o[0], Y = o[1];
  const real off = o[2], inv = o[3];
  int Nxx = Nx, x = px, y = py;
  if (o[5] != (real)0) {
    const real t = X;
    X = Y;
    Y = t;
    Nxx = Ny;
    x = py;
    y = px;
  }
  const real d = (real)y * Y + off;
  int xl = (int)(((real)s - 1 - d) * inv);
  int xh = (int)(((real)s + 1 - d) * inv);
  if (inv < (real)0) {
    const int t = xl;
    xl = xh;
    xh = t;
  }
  xl = xl < 0 ? 0 : xl;
  xh = xh > Nxx - 1 ? Nxx - 1 : xh;
  real acc = 0;
  if (xl <= x && x <= xh) {
    const real w = (real)1 - fabs((real)x * X + d - (real)s);
    if (w > (real)0) acc = w;
  }
  sino[s * so.s0 + a * so.s1] = acc * (real)delta_x * (real)delta_x / (real)delta_xi;
}

template <typename real>
__global__ void single_line_fan_kernel(real *__restrict__ sino, Strides3 so, int px, int py,
                                       const real *__restrict__ ofs, const real *__restrict__ sdpd,
                                       const real *__restrict__ geo, int Nx, int Ny, int Ns,
                                       int Na) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (s >= Ns) return;
  const real R = geo[0];
  real RE = geo[1];
  const real delta_xi = geo[2];
  real midx = geo[3], midy = geo[4];
  const real mid_det = geo[5];
  const real delta_x = geo[10];
  const real *o = ofs + 8 * (int64_t)a;
  real dx = o[0], dy = o[1], qx = o[2], qy = o[3];
  const real d0x = o[4], d0y = o[5];
  const real sp = -mid_det + (real)s + (real)1;
  const real sm = -mid_det + (real)s - (real)1;
  real dpx = d0x + dx * sp - qx, dpy = d0y + dy * sp - qy;
  real dmx = d0x + dx * sm - qx, dmy = d0y + dy * sm - qy;
  real ddx = d0x - qx, ddy = d0y - qy;
  const real RR = R * R;
  ddx /= RR;
  ddy /= RR;
  const real dxi2 = delta_xi * delta_xi;
  dx /= dxi2;
  dy /= dxi2;
  RE *= (real)1 / R;
  int Nxx = Nx;
  int x = px, y = py;
  if (!(fabs(dpx) < fabs(dpy) && fabs(dmx) < fabs(dmy))) {
    real t;
    t = ddx; ddx = ddy; ddy = t;
    t = dx; dx = dy; dy = t;
    t = qx; qx = qy; qy = t;
    t = dpx; dpx = dpy; dpy = t;
    t = dmx; dmx = dmy; dmy = t;
    t = midx; midx = midy; midy = t;
    Nxx = Ny;
    x = py;
    y = px;
  }
  dpx /= dpy;
  dmx /= dmy;
  if (qy * (dmx - dpx) < (real)0) {
    const real t = dpx;
    dpx = dmx;
    dmx = t;
  }
  const real u = (real)y - midy;
  const real xhigh = qx + dpx * (u - qy);
  const real xlow = qx + dmx * (u - qy);
  const int xh = (int)floor(fmax(fmin((real)(Nxx - 1) - midx, xhigh) + midx, (real)-1));
  const int xl = (int)ceil(fmin(fmax(-midx, xlow) + midx, (real)Nxx));
  real acc = 0;
  if (xl <= x && x <= xh) {
    const real t = ddx * ((real)x - midx) + ddy * u + RE;
    const real ss = dx * ((real)x - midx) + dy * u;
    const real xi = ss / t;
    acc = ((real)1 - fabs((real)s - xi - mid_det)) / (R * t);
  }
  sino[s * so.s0 + a * so.s1] = acc * sdpd[s] / delta_xi * delta_x;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
struct Layout {
  int Nx, Ny, Ns, Na, Nz;
};

static int check_sizes(const Layout &L) {
  GT_REQUIRE(L.Nx > 0 && L.Ny > 0 && L.Ns > 0 && L.Na > 0 && L.Nz > 0,
             "all dimensions must be positive");
  GT_REQUIRE(L.Nz <= 65535, "Nz must not exceed 65535");
  GT_REQUIRE(L.Na <= 65535, "Na must not exceed 65535");
  GT_REQUIRE((int64_t)L.Nx * L.Ny * L.Nz < ((int64_t)1 << 40), "image too large");
  return GT_OK;
}

static size_t elem_size(int dtype) { return dtype == GT_FLOAT64 ? 8 : 4; }

static size_t fwd_ws(int dtype, const Layout &L, int img_order) {
  const size_t img = align_up(elem_size(dtype) * (size_t)L.Nx * L.Ny * L.Nz);
  size_t n = 0;
  if (img_order != GT_ORDER_F) n += img;                      // x-fastest copy
  if (!(img_order == GT_ORDER_C && L.Nz == 1)) n += img;      // y-fastest copy
  return n;
}

static size_t back_ws(int dtype, const Layout &L) {
  return align_up(elem_size(dtype) * 8 * (size_t)L.Na) +
         align_up(elem_size(dtype) * (size_t)(L.Ns + 4) * L.Na * L.Nz);
}

template <typename real, bool FAN>
static int fwd_impl(hipStream_t st, real *sino, const real *img, const real *ofs,
                    const real *sdpd, const real *geo, const Layout &L, int sino_order,
                    int img_order, void *ws, size_t ws_bytes) {
  const size_t need = fwd_ws(sizeof(real) == 8 ? GT_FLOAT64 : GT_FLOAT32, L, img_order);
  if (ws_bytes < need) return fail(GT_ENOSPC, "projection workspace too small");
  GT_REQUIRE(need == 0 || ws != nullptr, "workspace pointer is NULL");
  const Strides3 is = order_strides(img_order, L.Nx, L.Ny, L.Nz);
  const Strides3 xs = {1, L.Nx, (int64_t)L.Nx * L.Ny};  // x fastest
  const Strides3 ys = {L.Ny, 1, (int64_t)L.Nx * L.Ny};  // y fastest
  char *w = static_cast<char *>(ws);
  const size_t img_bytes = align_up(sizeof(real) * (size_t)L.Nx * L.Ny * L.Nz);
  const real *imgX = img, *imgY = img;
  if (img_order != GT_ORDER_F) {
    real *cp = reinterpret_cast<real *>(w);
    w += img_bytes;
    launch_relayout<real>(st, cp, xs, img, is, L.Nx, L.Ny, L.Nz);
    imgX = cp;
  }
  if (!(img_order == GT_ORDER_C && L.Nz == 1)) {
    real *cp = reinterpret_cast<real *>(w);
    launch_relayout<real>(st, cp, ys, img, is, L.Nx, L.Ny, L.Nz);
    imgY = cp;
  }
  const Strides3 so = order_strides(sino_order, L.Ns, L.Na, L.Nz);
  const int K = tuning().fwd_group;
  dim3 block(64, 4);
  dim3 grid((L.Ns + 64 * K - 1) / (64 * K), (L.Na + 3) / 4, L.Nz);
  if (FAN) {
    if (K == 8)
      hipLaunchKernelGGL((fan_fwd_group_kernel<real, 8>), grid, block, 0, st, sino, so, imgX, imgY,
                         ofs, sdpd, geo, L.Nx, L.Ny, L.Ns, L.Na);
    else if (K == 4)
      hipLaunchKernelGGL((fan_fwd_group_kernel<real, 4>), grid, block, 0, st, sino, so, imgX, imgY,
                         ofs, sdpd, geo, L.Nx, L.Ny, L.Ns, L.Na);
    else if (K == 2)
      hipLaunchKernelGGL((fan_fwd_group_kernel<real, 2>), grid, block, 0, st, sino, so, imgX, imgY,
                         ofs, sdpd, geo, L.Nx, L.Ny, L.Ns, L.Na);
    else
      hipLaunchKernelGGL(fan_fwd_kernel<real>, grid, block, 0, st, sino, so, imgX, imgY, ofs, sdpd,
                         geo, L.Nx, L.Ny, L.Ns, L.Na);
  } else {
    if (K == 8)
      hipLaunchKernelGGL((radon_fwd_group_kernel<real, 8>), grid, block, 0, st, sino, so, imgX,
                         imgY, ofs, geo, L.Nx, L.Ny, L.Ns, L.Na);
    else if (K == 4)
      hipLaunchKernelGGL((radon_fwd_group_kernel<real, 4>), grid, block, 0, st, sino, so, imgX,
                         imgY, ofs, geo, L.Nx, L.Ny, L.Ns, L.Na);
    else if (K == 2)
      hipLaunchKernelGGL((radon_fwd_group_kernel<real, 2>), grid, block, 0, st, sino, so, imgX,
                         imgY, ofs, geo, L.Nx, L.Ny, L.Ns, L.Na);
    else
      hipLaunchKernelGGL(radon_fwd_kernel<real>, grid, block, 0, st, sino, so, imgX, imgY, ofs,
                         geo, L.Nx, L.Ny, L.Ns, L.Na);
  }
  return check_launch(FAN ? "fan forward projection" : "radon forward projection");
}

template <typename real, bool FAN>
static int back_impl(hipStream_t st, real *img, const real *sino, const real *ofs,
                     const real *sdpd, const real *geo, const Layout &L, int img_order,
                     int sino_order, void *ws, size_t ws_bytes) {
  const size_t need = back_ws(sizeof(real) == 8 ? GT_FLOAT64 : GT_FLOAT32, L);
  if (ws_bytes < need) return fail(GT_ENOSPC, "projection workspace too small");
  GT_REQUIRE(ws != nullptr, "workspace pointer is NULL");
  char *w = static_cast<char *>(ws);
  AngleRec<real> *tab = reinterpret_cast<AngleRec<real> *>(w);
  real *S = reinterpret_cast<real *>(w + align_up(sizeof(real) * 8 * (size_t)L.Na));
  hipLaunchKernelGGL((back_table_kernel<real, FAN>), dim3((L.Na + 63) / 64), dim3(64), 0, st, tab,
                     ofs, geo, L.Na);
  const Strides3 si = order_strides(sino_order, L.Ns, L.Na, L.Nz);
  hipLaunchKernelGGL((stage_sino_kernel<real, FAN>), dim3((L.Ns + 4 + 255) / 256, L.Na, L.Nz),
                     dim3(256), 0, st, S, sino, si, ofs, sdpd, geo, L.Ns, L.Na);
  const Strides3 io = order_strides(img_order, L.Nx, L.Ny, L.Nz);
  const int RY = tuning().back_rows;
  dim3 grid((L.Nx + 15) / 16, (L.Ny + 16 * RY - 1) / (16 * RY), L.Nz);
  if (RY == 4)
    hipLaunchKernelGGL((back_kernel<real, FAN, 4>), grid, dim3(256), 0, st, img, io, S, tab, geo,
                       L.Nx, L.Ny, L.Ns, L.Na);
  else if (RY == 2)
    hipLaunchKernelGGL((back_kernel<real, FAN, 2>), grid, dim3(256), 0, st, img, io, S, tab, geo,
                       L.Nx, L.Ny, L.Ns, L.Na);
  else
    hipLaunchKernelGGL((back_kernel<real, FAN, 1>), grid, dim3(256), 0, st, img, io, S, tab, geo,
                       L.Nx, L.Ny, L.Ns, L.Na);
  return check_launch(FAN ? "fan backprojection" : "radon backprojection");
}

static int check_common(int dtype, const void *out, const void *in, const void *ofs,
                        const void *geo, int out_order, int in_order) {
  GT_REQUIRE(dtype == GT_FLOAT32 || dtype == GT_FLOAT64, "dtype must be GT_FLOAT32 or GT_FLOAT64");
  GT_REQUIRE(out && in && ofs && geo, "NULL array argument");
  GT_REQUIRE(out != in, "output must not alias the input");
  GT_REQUIRE((out_order == GT_ORDER_F || out_order == GT_ORDER_C) &&
                 (in_order == GT_ORDER_F || in_order == GT_ORDER_C),
             "order must be GT_ORDER_F or GT_ORDER_C");
  return GT_OK;
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_projection_workspace_bytes(int geometry, int adjoint, int dtype, int Nx,
                                                int Ny, int Ns, int Na, int Nz, int in_order) {
  (void)geometry;
  if (Nx <= 0 || Ny <= 0 || Ns <= 0 || Na <= 0 || Nz <= 0) return 0;
  const Layout L{Nx, Ny, Ns, Na, Nz};
  return adjoint ? back_ws(dtype, L) : fwd_ws(dtype, L, in_order);
}

#define GT_DISPATCH(dtype, CALL_F32, CALL_F64) \
  ((dtype) == GT_FLOAT64 ? (CALL_F64) : (CALL_F32))

extern "C" int gt_radon_fwd(gt_stream_t stream, int dtype, void *sino, const void *img,
                            const void *ofs, const void *geo, int Nx, int Ny, int Ns, int Na,
                            int Nz, int sino_order, int img_order, void *workspace,
                            size_t workspace_bytes) {
  int rc = check_common(dtype, sino, img, ofs, geo, sino_order, img_order);
  if (rc) return rc;
  const Layout L{Nx, Ny, Ns, Na, Nz};
  if ((rc = check_sizes(L))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return GT_DISPATCH(
      dtype,
      (fwd_impl<float, false>(st, (float *)sino, (const float *)img, (const float *)ofs, nullptr,
                              (const float *)geo, L, sino_order, img_order, workspace,
                              workspace_bytes)),
      (fwd_impl<double, false>(st, (double *)sino, (const double *)img, (const double *)ofs,
                               nullptr, (const double *)geo, L, sino_order, img_order, workspace,
                               workspace_bytes)));
}

extern "C" int gt_fan_fwd(gt_stream_t stream, int dtype, void *sino, const void *img,
                          const void *ofs, const void *sdpd, const void *geo, int Nx, int Ny,
                          int Ns, int Na, int Nz, int sino_order, int img_order, void *workspace,
                          size_t workspace_bytes) {
  int rc = check_common(dtype, sino, img, ofs, geo, sino_order, img_order);
  if (rc) return rc;
  GT_REQUIRE(sdpd != nullptr, "NULL sdpd");
  const Layout L{Nx, Ny, Ns, Na, Nz};
  if ((rc = check_sizes(L))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return GT_DISPATCH(
      dtype,
      (fwd_impl<float, true>(st, (float *)sino, (const float *)img, (const float *)ofs,
                             (const float *)sdpd, (const float *)geo, L, sino_order, img_order,
                             workspace, workspace_bytes)),
      (fwd_impl<double, true>(st, (double *)sino, (const double *)img, (const double *)ofs,
                              (const double *)sdpd, (const double *)geo, L, sino_order, img_order,
                              workspace, workspace_bytes)));
}

extern "C" int gt_radon_back(gt_stream_t stream, int dtype, void *img, const void *sino,
                             const void *ofs, const void *geo, int Nx, int Ny, int Ns, int Na,
                             int Nz, int img_order, int sino_order, void *workspace,
                             size_t workspace_bytes) {
  int rc = check_common(dtype, img, sino, ofs, geo, img_order, sino_order);
  if (rc) return rc;
  const Layout L{Nx, Ny, Ns, Na, Nz};
  if ((rc = check_sizes(L))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return GT_DISPATCH(
      dtype,
      (back_impl<float, false>(st, (float *)img, (const float *)sino, (const float *)ofs, nullptr,
                               (const float *)geo, L, img_order, sino_order, workspace,
                               workspace_bytes)),
      (back_impl<double, false>(st, (double *)img, (const double *)sino, (const double *)ofs,
                                nullptr, (const double *)geo, L, img_order, sino_order, workspace,
                                workspace_bytes)));
}

extern "C" int gt_fan_back(gt_stream_t stream, int dtype, void *img, const void *sino,
                           const void *ofs, const void *sdpd, const void *geo, int Nx, int Ny,
                           int Ns, int Na, int Nz, int img_order, int sino_order, void *workspace,
                           size_t workspace_bytes) {
  int rc = check_common(dtype, img, sino, ofs, geo, img_order, sino_order);
  if (rc) return rc;
  GT_REQUIRE(sdpd != nullptr, "NULL sdpd");
  const Layout L{Nx, Ny, Ns, Na, Nz};
  if ((rc = check_sizes(L))) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return GT_DISPATCH(
      dtype,
      (back_impl<float, true>(st, (float *)img, (const float *)sino, (const float *)ofs,
                              (const float *)sdpd, (const float *)geo, L, img_order, sino_order,
                              workspace, workspace_bytes)),
      (back_impl<double, true>(st, (double *)img, (const double *)sino, (const double *)ofs,
                               (const double *)sdpd, (const double *)geo, L, img_order,
                               sino_order, workspace, workspace_bytes)));
}

extern "C" int gt_single_line(gt_stream_t stream, int geometry, int dtype, void *sino, int x,
                              int y, const void *ofs, const void *sdpd, const void *geo, int Nx,
                              int Ny, int Ns, int Na, int sino_order) {
  GT_REQUIRE(dtype == GT_FLOAT32 || dtype == GT_FLOAT64, "dtype must be GT_FLOAT32 or GT_FLOAT64");
  GT_REQUIRE(sino && ofs && geo, "NULL array argument");
  GT_REQUIRE(geometry == GT_GEOMETRY_RADON || geometry == GT_GEOMETRY_FAN, "unknown geometry");
  GT_REQUIRE(geometry == GT_GEOMETRY_RADON || sdpd != nullptr, "NULL sdpd");
  GT_REQUIRE(Nx > 0 && Ny > 0 && Ns > 0 && Na > 0 && Na <= 65535, "invalid dimensions");
  GT_REQUIRE(x >= 0 && x < Nx && y >= 0 && y < Ny, "pixel outside the image");
  GT_REQUIRE(sino_order == GT_ORDER_F || sino_order == GT_ORDER_C, "invalid order");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Strides3 so = order_strides(sino_order, Ns, Na, 1);
  dim3 grid((Ns + 255) / 256, Na);
  if (geometry == GT_GEOMETRY_RADON) {
    if (dtype == GT_FLOAT64)
      hipLaunchKernelGGL(single_line_radon_kernel<double>, grid, dim3(256), 0, st, (double *)sino,
                         so, x, y, (const double *)ofs, (const double *)geo, Nx, Ny, Ns, Na);
    else
      hipLaunchKernelGGL(single_line_radon_kernel<float>, grid, dim3(256), 0, st, (float *)sino,
                         so, x, y, (const float *)ofs, (const float *)geo, Nx, Ny, Ns, Na);
  } else {
    if (dtype == GT_FLOAT64)
      hipLaunchKernelGGL(single_line_fan_kernel<double>, grid, dim3(256), 0, st, (double *)sino,
                         so, x, y, (const double *)ofs, (const double *)sdpd, (const double *)geo,
                         Nx, Ny, Ns, Na);
    else
      hipLaunchKernelGGL(single_line_fan_kernel<float>, grid, dim3(256), 0, st, (float *)sino, so,
                         x, y, (const float *)ofs, (const float *)sdpd, (const float *)geo, Nx, Ny,
                         Ns, Na);
  }
  return check_launch("single line projection");
}
