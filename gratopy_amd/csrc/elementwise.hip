// Streaming kernels of the solvers: angle weighting, axpy variants,
// deterministic reductions, and the TV primal-dual pointwise updates
// (gratopy/utilities.cl, gratopy/total_variation.cl).  All are HBM bound:
// grid-stride loops over 256-thread blocks, one element per lane per step.

#include "common.h"

namespace gt {

static inline dim3 stream_grid(size_t n) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;  // 256 CUs x 32 blocks; the loop covers the rest
  if (blocks == 0) blocks = 1;
  return dim3((unsigned)blocks);
}

#define GRID_STRIDE(i, n)                                                      \
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); \
       i += (size_t)gridDim.x * blockDim.x)

// ---- sinogram angle weighting (utilities.cl:38-67) -------------------------
template <typename real>
__global__ void weight_sino_kernel(real *out, const real *in, const real *__restrict__ w, int Ns,
                                   int Na, int Nz, Strides3 so, int divide) {
  const size_t n = (size_t)Ns * Na * Nz;
  GRID_STRIDE(i, n) {
    // i enumerates (s, a, z) with s fastest; the element lives at its order's offset
    const int s = (int)(i % Ns);
    const size_t r = i / Ns;
    const int a = (int)(r % Na);
    const int z = (int)(r / Na);
    const int64_t k = s * so.s0 + a * so.s1 + z * so.s2;
    out[k] = divide ? in[k] / w[a] : in[k] * w[a];
  }
}

// ---- axpy family (utilities.cl:69-80) --------------------------------------
template <typename real>
__global__ void axpy_kernel(real *rhs, real a, const real *x, size_t n) {
  GRID_STRIDE(i, n) rhs[i] += a * x[i];
}

template <typename real>
__global__ void axpby_kernel(real *rhs, real a, const real *x, const real *y, size_t n) {
  GRID_STRIDE(i, n) rhs[i] = a * x[i] + y[i];
}

template <typename real>
__global__ void axpy_dev_kernel(real *rhs, const double *a_dev, double mult, const real *x,
                                size_t n) {
  const real a = (real)(mult * a_dev[0]);
  GRID_STRIDE(i, n) rhs[i] += a * x[i];
}

template <typename real>
__global__ void axpby_dev_kernel(real *rhs, const double *a_dev, double mult, const real *x,
                                 const real *y, size_t n) {
  const real a = (real)(mult * a_dev[0]);
  GRID_STRIDE(i, n) rhs[i] = a * x[i] + y[i];
}

template <typename real>
__global__ void scale_dev_kernel(real *x, const double *s_dev, int divide, size_t n) {
  const real s = (real)s_dev[0];
  GRID_STRIDE(i, n) x[i] = divide ? x[i] / s : x[i] * s;
}

template <typename real>
__global__ void scale_kernel(real *x, real a, size_t n) {
  GRID_STRIDE(i, n) x[i] *= a;
}

template <typename real>
__global__ void fill_kernel(real *x, real v, size_t n) {
  GRID_STRIDE(i, n) x[i] = v;
}

template <typename real>
__global__ void sub_kernel(real *out, const real *a, const real *b, size_t n) {
  GRID_STRIDE(i, n) out[i] = a[i] - b[i];
}

// ---- deterministic reductions ----------------------------------------------
constexpr int kReduceBlocks = 1024;

__device__ inline double block_sum(double v) {
  __shared__ double part[256 / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < (int)(blockDim.x / 64); ++k) t += part[k];
  }
  return t;
}

template <typename real, bool DOT>
__global__ __launch_bounds__(256) void reduce_stage1(double *partial, const real *x,
                                                     const real *y, size_t n) {
  double acc = 0;
  GRID_STRIDE(i, n) acc += DOT ? (double)x[i] * (double)y[i] : (double)x[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void reduce_stage2(double *out, const double *partial,
                                                     int nparts) {
  double acc = 0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += partial[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) out[0] = t;
}

static int reduce_blocks(size_t n) {
  size_t b = (n + 2047) / 2048;
  if (b < 1) b = 1;
  if (b > (size_t)kReduceBlocks) b = kReduceBlocks;
  return (int)b;
}

template <typename real, bool DOT>
static int reduce_impl(hipStream_t st, double *out, const real *x, const real *y, size_t n,
                       void *ws, size_t ws_bytes) {
  if (ws_bytes < sizeof(double) * kReduceBlocks) return fail(GT_ENOSPC, "reduction workspace too small");
  double *partial = static_cast<double *>(ws);
  const int nb = reduce_blocks(n);
  hipLaunchKernelGGL((reduce_stage1<real, DOT>), dim3(nb), dim3(256), 0, st, partial, x, y, n);
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(256), 0, st, out, partial, nb);
  return check_launch("reduction");
}

__global__ void cg_scalars_kernel(double *d, double scale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (scale > 0) {
    d[4] = scale * (d[0] / d[1]);
  } else {
    d[5] = d[2] / d[0];
    d[0] = d[2];
  }
}

// ---- TV primal-dual updates (total_variation.cl) ----------------------------
// Pixel (x, y, z) of an Nx x Ny x Nz image in order `o` lives at
// x*st.s0 + y*st.s1 + z*st.s2; the dual variable holds 4 reals per pixel.
struct ImgIdx {
  int x, y, z;
  int64_t i;
};

__device__ inline ImgIdx img_index(size_t lin, int Nx, int Ny, int order, Strides3 st) {
  // lin enumerates pixels in memory order so that lanes touch consecutive addresses
  ImgIdx r;
  if (order == GT_ORDER_F) {
    r.x = (int)(lin % Nx);
    const size_t q = lin / Nx;
    r.y = (int)(q % Ny);
    r.z = (int)(q / Ny);
  } else {
    const int Nz = (int)(st.s1);  // C order: y stride is Nz
    r.z = (int)(lin % Nz);
    const size_t q = lin / Nz;
    r.y = (int)(q % Ny);
    r.x = (int)(q / Ny);
  }
  r.i = r.x * st.s0 + r.y * st.s1 + r.z * st.s2;
  return r;
}

template <typename real>
__device__ inline void dual_update(real *__restrict__ v, const real *__restrict__ u,
                                   const ImgIdx &p, int Nx, int Ny, int Nz, Strides3 st,
                                   float sigma, float zdist, bool project) {
  const real c = u[p.i];
  real g0 = -c, g1 = -c, g2 = -c;
  g0 = p.x < Nx - 1 ? g0 + u[p.i + st.s0] : (real)0;
  g1 = p.y < Ny - 1 ? g1 + u[p.i + st.s1] : (real)0;
  g2 = p.z < Nz - 1 ? g2 + u[p.i + st.s2] : (real)0;
  g2 = zdist > 0.0f ? g2 / (real)zdist : (real)0;
  real *vp = v + 4 * p.i;
  const real sg = (real)sigma;
  real v0 = vp[0] + sg * g0;
  real v1 = vp[1] + sg * g1;
  real v2 = vp[2] + sg * g2;
  if (project) {
    const real norm = hypot(hypot(v0, v1), v2);
    if (norm > (real)1) {
      v0 /= norm;
      v1 /= norm;
      v2 /= norm;
    }
  }
  vp[0] = v0;
  vp[1] = v1;
  vp[2] = v2;
}

template <typename real>
__global__ __launch_bounds__(256) void tv_v_kernel(real *v, const real *u, float sigma,
                                                   float zdist, int Nx, int Ny, int Nz, int order,
                                                   Strides3 st, int project) {
  const size_t n = (size_t)Nx * Ny * Nz;
  GRID_STRIDE(lin, n) {
    const ImgIdx p = img_index(lin, Nx, Ny, order, st);
    dual_update<real>(v, u, p, Nx, Ny, Nz, st, sigma, zdist, project != 0);
  }
}

template <typename real>
__global__ __launch_bounds__(256) void tv_normv_kernel(real *v, int Nx, int Ny, int Nz, int order,
                                                       Strides3 st) {
  const size_t n = (size_t)Nx * Ny * Nz;
  GRID_STRIDE(lin, n) {
    const ImgIdx p = img_index(lin, Nx, Ny, order, st);
    real *vp = v + 4 * p.i;
    const real norm = hypot(hypot(vp[0], vp[1]), vp[2]);
    if (norm > (real)1) {
      vp[0] /= norm;
      vp[1] /= norm;
      vp[2] /= norm;
    }
  }
}

template <typename real>
__global__ __launch_bounds__(256) void tv_lambda_kernel(real *lambda, const real *ku,
                                                        const real *f, float sigma,
                                                        float sigmap1inv, size_t n) {
  const real sg = (real)sigma, c = (real)sigmap1inv;
  GRID_STRIDE(i, n) lambda[i] = (lambda[i] + sg * (ku[i] - f[i])) * c;
}

// u_out = max(0, u_in + tau*(div v - norming*kstar)); with extra != 0 also
// u_in <- 2*u_out - u_in (the extragradient step).
template <typename real>
__global__ __launch_bounds__(256) void tv_u_kernel(real *u_out, real *u_in, const real *v,
                                                   const real *kstar, float tau, float norming,
                                                   float zdist, int Nx, int Ny, int Nz, int order,
                                                   Strides3 st, int extra) {
  const size_t n = (size_t)Nx * Ny * Nz;
  GRID_STRIDE(lin, n) {
    const ImgIdx p = img_index(lin, Nx, Ny, order, st);
    const real *vp = v + 4 * p.i;
    real d0 = vp[0], d1 = vp[1], d2 = vp[2];
    if (p.x == Nx - 1) d0 = 0;
    if (p.x > 0) d0 -= v[4 * (p.i - st.s0) + 0];
    if (p.y == Ny - 1) d1 = 0;
    if (p.y > 0) d1 -= v[4 * (p.i - st.s1) + 1];
    if (p.z == Nz - 1) d2 = 0;
    if (p.z > 0) d2 -= v[4 * (p.i - st.s2) + 2];
    d2 = zdist > 0.0f ? d2 / (real)zdist : (real)0;
    const real old = u_in[p.i];
    const real un = fmax((real)0, old + (real)tau * (d0 + d1 + d2 - (real)norming * kstar[p.i]));
    u_out[p.i] = un;
    if (extra) u_in[p.i] = (real)2 * un - old;
  }
}

template <typename real>
__global__ __launch_bounds__(256) void tv_extra_kernel(const real *u_, real *u, size_t n) {
  GRID_STRIDE(i, n) u[i] = (real)2 * u_[i] - u[i];
}

// ---- adjoint (vector-Jacobian product) of one TV primal-dual step ------------
// Forward step (as the kernels above implement it):
//   a   = v + sigma grad(ubar)          v'  = a / max(1, |a|)
//   lam'= (lam + sigma_p (K ubar - f)) c
//   u'  = max(0, u + tau (div v' - n B lam'))        ubar' = 2 u' - u
// The adjoint kernels below take the cotangents of (u', ubar', v', lam').
// grad has forward differences with a zero last difference (Neumann), and
// div = -grad^T for this pair (total_variation.cl:54-148).

// wb = [u' > 0] (ub' + 2 ubarb');  ub = wb - ubarb'
template <typename real>
__global__ __launch_bounds__(256) void tv_adj_primal_kernel(real *wb, real *ub, const real *ub_next,
                                                            const real *ubarb_next,
                                                            const real *u_next, size_t n) {
  GRID_STRIDE(i, n) {
    const real g = ub_next[i] + (real)2 * ubarb_next[i];
    const real w = u_next[i] > (real)0 ? g : (real)0;
    const real ubar = ubarb_next[i];
    wb[i] = w;
    ub[i] = w - ubar;
  }
}

template <typename real>
__device__ inline void forward_grad(const real *__restrict__ f, const ImgIdx &p, int Nx, int Ny,
                                    int Nz, Strides3 st, float zdist, real &g0, real &g1, real &g2) {
  const real c = f[p.i];
  g0 = p.x < Nx - 1 ? -c + f[p.i + st.s0] : (real)0;
  g1 = p.y < Ny - 1 ? -c + f[p.i + st.s1] : (real)0;
  g2 = p.z < Nz - 1 ? -c + f[p.i + st.s2] : (real)0;
  g2 = zdist > 0.0f ? g2 / (real)zdist : (real)0;
}

// ab = J_P(a) (vb' - tau grad(wb)) with a = v + sigma grad(ubar) recomputed;
// J_P = I where |a| <= 1, (I - a a^T/|a|^2)/|a| where the ball projection was active.
template <typename real>
__global__ __launch_bounds__(256) void tv_adj_dual_kernel(real *ab, const real *vb_next,
                                                          const real *wb, const real *v,
                                                          const real *ubar, float sigma, float tau,
                                                          float zdist, int Nx, int Ny, int Nz,
                                                          int order, Strides3 st) {
  const size_t n = (size_t)Nx * Ny * Nz;
  GRID_STRIDE(lin, n) {
    const ImgIdx p = img_index(lin, Nx, Ny, order, st);
    real g0, g1, g2, w0, w1, w2;
    forward_grad(ubar, p, Nx, Ny, Nz, st, zdist, g0, g1, g2);
    forward_grad(wb, p, Nx, Ny, Nz, st, zdist, w0, w1, w2);
    const real *vp = v + 4 * p.i;
    const real sg = (real)sigma, ta = (real)tau;
    const real a0 = vp[0] + sg * g0, a1 = vp[1] + sg * g1, a2 = vp[2] + sg * g2;
    const real *bp = vb_next + 4 * p.i;
    real x0 = bp[0] - ta * w0, x1 = bp[1] - ta * w1, x2 = bp[2] - ta * w2;
    const real norm = hypot(hypot(a0, a1), a2);
    if (norm > (real)1) {
      const real h0 = a0 / norm, h1 = a1 / norm, h2 = a2 / norm;
      const real d = h0 * x0 + h1 * x1 + h2 * x2;
      x0 = (x0 - h0 * d) / norm;
      x1 = (x1 - h1 * d) / norm;
      x2 = (x2 - h2 * d) / norm;
    }
    real *out = ab + 4 * p.i;
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = 0;
  }
}

// ubarb = coef * bterm - sigma div(ab)
template <typename real>
__global__ __launch_bounds__(256) void tv_adj_div_kernel(real *ubarb, const real *bterm, real coef,
                                                         const real *ab, float sigma, float zdist,
                                                         int Nx, int Ny, int Nz, int order,
                                                         Strides3 st) {
  const size_t n = (size_t)Nx * Ny * Nz;
  GRID_STRIDE(lin, n) {
    const ImgIdx p = img_index(lin, Nx, Ny, order, st);
    const real *vp = ab + 4 * p.i;
    real d0 = vp[0], d1 = vp[1], d2 = vp[2];
    if (p.x == Nx - 1) d0 = 0;
    if (p.x > 0) d0 -= ab[4 * (p.i - st.s0) + 0];
    if (p.y == Ny - 1) d1 = 0;
    if (p.y > 0) d1 -= ab[4 * (p.i - st.s1) + 1];
    if (p.z == Nz - 1) d2 = 0;
    if (p.z > 0) d2 -= ab[4 * (p.i - st.s2) + 2];
    d2 = zdist > 0.0f ? d2 / (real)zdist : (real)0;
    ubarb[p.i] = coef * bterm[p.i] - (real)sigma * (d0 + d1 + d2);
  }
}

// bb = c (lamb' - coefR * w_a * Rwb);  lamb = bb;  fb -= sigma_p * bb;  winv = bb / w_a
template <typename real>
__global__ __launch_bounds__(256) void tv_adj_lambda_kernel(real *lamb, real *fb, real *winv,
                                                            const real *lamb_next, const real *rwb,
                                                            const real *__restrict__ w, real coefR,
                                                            float c, float sigma_p, int Ns, int Na,
                                                            int Nz, Strides3 so) {
  const size_t n = (size_t)Ns * Na * Nz;
  GRID_STRIDE(i, n) {
    const int s = (int)(i % Ns);
    const size_t r = i / Ns;
    const int a = (int)(r % Na);
    const int z = (int)(r / Na);
    const int64_t k = s * so.s0 + a * so.s1 + z * so.s2;
    const real bb = (real)c * (lamb_next[k] - coefR * w[a] * rwb[k]);
    lamb[k] = bb;
    fb[k] -= (real)sigma_p * bb;
    winv[k] = bb / w[a];
  }
}

static int check_dtype(int dtype) {
  GT_REQUIRE(dtype == GT_FLOAT32 || dtype == GT_FLOAT64, "dtype must be GT_FLOAT32 or GT_FLOAT64");
  return GT_OK;
}

static int check_img(int Nx, int Ny, int Nz, int order) {
  GT_REQUIRE(Nx > 0 && Ny > 0 && Nz > 0, "all dimensions must be positive");
  GT_REQUIRE(order == GT_ORDER_F || order == GT_ORDER_C, "invalid order");
  return GT_OK;
}

}  // namespace gt

using namespace gt;

#define AS(T, p) reinterpret_cast<T>(p)

extern "C" int gt_weight_sino(gt_stream_t stream, int dtype, void *out, const void *in,
                              const void *angle_weights, int Ns, int Na, int Nz, int order,
                              int divide) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(out && in && angle_weights, "NULL array argument");
  GT_REQUIRE(Ns > 0 && Na > 0 && Nz > 0, "all dimensions must be positive");
  GT_REQUIRE(order == GT_ORDER_F || order == GT_ORDER_C, "invalid order");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t n = (size_t)Ns * Na * Nz;
  const Strides3 so = order_strides(order, Ns, Na, Nz);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(weight_sino_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, out), AS(const double *, in),
                       AS(const double *, angle_weights), Ns, Na, Nz, so, divide);
  else
    hipLaunchKernelGGL(weight_sino_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, out), AS(const float *, in), AS(const float *, angle_weights),
                       Ns, Na, Nz, so, divide);
  return check_launch("weight_sino");
}

extern "C" int gt_axpy(gt_stream_t stream, int dtype, void *rhs, double a, const void *x,
                       size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(rhs && x, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(axpy_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, rhs),
                       a, AS(const double *, x), n);
  else
    hipLaunchKernelGGL(axpy_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, rhs),
                       (float)a, AS(const float *, x), n);
  return check_launch("axpy");
}

extern "C" int gt_axpby(gt_stream_t stream, int dtype, void *rhs, double a, const void *x,
                        const void *y, size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(rhs && x && y, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(axpby_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, rhs),
                       a, AS(const double *, x), AS(const double *, y), n);
  else
    hipLaunchKernelGGL(axpby_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, rhs),
                       (float)a, AS(const float *, x), AS(const float *, y), n);
  return check_launch("axpby");
}

extern "C" int gt_axpy_dev(gt_stream_t stream, int dtype, void *rhs, const void *a_dev,
                           double mult, const void *x, size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(rhs && x && a_dev, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(axpy_dev_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, rhs), AS(const double *, a_dev), mult, AS(const double *, x),
                       n);
  else
    hipLaunchKernelGGL(axpy_dev_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, rhs),
                       AS(const double *, a_dev), mult, AS(const float *, x), n);
  return check_launch("axpy_dev");
}

extern "C" int gt_axpby_dev(gt_stream_t stream, int dtype, void *rhs, const void *a_dev,
                            double mult, const void *x, const void *y, size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(rhs && x && y && a_dev, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(axpby_dev_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, rhs), AS(const double *, a_dev), mult,
                       AS(const double *, x), AS(const double *, y), n);
  else
    hipLaunchKernelGGL(axpby_dev_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, rhs), AS(const double *, a_dev), mult, AS(const float *, x),
                       AS(const float *, y), n);
  return check_launch("axpby_dev");
}

extern "C" int gt_scale_dev(gt_stream_t stream, int dtype, void *x, const void *s_dev, int divide,
                            size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(x && s_dev, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(scale_dev_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, x), AS(const double *, s_dev), divide, n);
  else
    hipLaunchKernelGGL(scale_dev_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, x),
                       AS(const double *, s_dev), divide, n);
  return check_launch("scale_dev");
}

extern "C" int gt_scale(gt_stream_t stream, int dtype, void *x, double a, size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(x, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(scale_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, x), a,
                       n);
  else
    hipLaunchKernelGGL(scale_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, x),
                       (float)a, n);
  return check_launch("scale");
}

extern "C" int gt_fill(gt_stream_t stream, int dtype, void *x, double value, size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(x, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(fill_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, x),
                       value, n);
  else
    hipLaunchKernelGGL(fill_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, x),
                       (float)value, n);
  return check_launch("fill");
}

extern "C" int gt_sub(gt_stream_t stream, int dtype, void *out, const void *a, const void *b,
                      size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(out && a && b, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(sub_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, out),
                       AS(const double *, a), AS(const double *, b), n);
  else
    hipLaunchKernelGGL(sub_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, out),
                       AS(const float *, a), AS(const float *, b), n);
  return check_launch("sub");
}

extern "C" size_t gt_reduce_workspace_bytes(size_t n) {
  (void)n;
  return align_up(sizeof(double) * kReduceBlocks);
}

extern "C" int gt_dot(gt_stream_t stream, int dtype, void *out_dev, const void *x, const void *y,
                      size_t n, void *workspace, size_t workspace_bytes) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(out_dev && x && y && workspace, "NULL argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    return reduce_impl<double, true>(st, AS(double *, out_dev), AS(const double *, x),
                                     AS(const double *, y), n, workspace, workspace_bytes);
  return reduce_impl<float, true>(st, AS(double *, out_dev), AS(const float *, x),
                                  AS(const float *, y), n, workspace, workspace_bytes);
}

extern "C" int gt_sum(gt_stream_t stream, int dtype, void *out_dev, const void *x, size_t n,
                      void *workspace, size_t workspace_bytes) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(out_dev && x && workspace, "NULL argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    return reduce_impl<double, false>(st, AS(double *, out_dev), AS(const double *, x), nullptr, n,
                                      workspace, workspace_bytes);
  return reduce_impl<float, false>(st, AS(double *, out_dev), AS(const float *, x), nullptr, n,
                                   workspace, workspace_bytes);
}

extern "C" int gt_cg_scalars(gt_stream_t stream, void *d, double scale) {
  GT_REQUIRE(d, "NULL argument");
  hipLaunchKernelGGL(cg_scalars_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                     AS(double *, d), scale);
  return check_launch("cg_scalars");
}

extern "C" int gt_tv_update_v(gt_stream_t stream, int dtype, void *v, const void *u, float sigma,
                              float z_distance, int Nx, int Ny, int Nz, int order) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Nx, Ny, Nz, order))) return rc;
  GT_REQUIRE(v && u, "NULL array argument");
  const Strides3 st3 = order_strides(order, Nx, Ny, Nz);
  const size_t n = (size_t)Nx * Ny * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_v_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, v),
                       AS(const double *, u), sigma, z_distance, Nx, Ny, Nz, order, st3, 0);
  else
    hipLaunchKernelGGL(tv_v_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, v),
                       AS(const float *, u), sigma, z_distance, Nx, Ny, Nz, order, st3, 0);
  return check_launch("tv_update_v");
}

extern "C" int gt_tv_dual(gt_stream_t stream, int dtype, void *v, const void *u, float sigma,
                          float z_distance, int Nx, int Ny, int Nz, int order) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Nx, Ny, Nz, order))) return rc;
  GT_REQUIRE(v && u, "NULL array argument");
  const Strides3 st3 = order_strides(order, Nx, Ny, Nz);
  const size_t n = (size_t)Nx * Ny * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_v_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, v),
                       AS(const double *, u), sigma, z_distance, Nx, Ny, Nz, order, st3, 1);
  else
    hipLaunchKernelGGL(tv_v_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, v),
                       AS(const float *, u), sigma, z_distance, Nx, Ny, Nz, order, st3, 1);
  return check_launch("tv_dual");
}

extern "C" int gt_tv_update_normv(gt_stream_t stream, int dtype, void *v, void *normv, int Nx,
                                  int Ny, int Nz, int order) {
  (void)normv;  // the reference allocates it but never reads or writes it
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Nx, Ny, Nz, order))) return rc;
  GT_REQUIRE(v, "NULL array argument");
  const Strides3 st3 = order_strides(order, Nx, Ny, Nz);
  const size_t n = (size_t)Nx * Ny * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_normv_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, v), Nx, Ny, Nz, order, st3);
  else
    hipLaunchKernelGGL(tv_normv_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, v),
                       Nx, Ny, Nz, order, st3);
  return check_launch("tv_update_normv");
}

extern "C" int gt_tv_update_lambda(gt_stream_t stream, int dtype, void *lambda, const void *ku,
                                   const void *f, float sigma, float sigmap1inv, int Ns, int Na,
                                   int Nz, int order) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Ns, Na, Nz, order))) return rc;
  GT_REQUIRE(lambda && ku && f, "NULL array argument");
  const size_t n = (size_t)Ns * Na * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_lambda_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, lambda), AS(const double *, ku), AS(const double *, f), sigma,
                       sigmap1inv, n);
  else
    hipLaunchKernelGGL(tv_lambda_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, lambda), AS(const float *, ku), AS(const float *, f), sigma,
                       sigmap1inv, n);
  return check_launch("tv_update_lambda");
}

static int tv_u_common(gt_stream_t stream, int dtype, void *u_out, void *u_in, const void *v,
                       const void *kstar, float tau, float norming, float zdist, int Nx, int Ny,
                       int Nz, int order, int extra) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Nx, Ny, Nz, order))) return rc;
  GT_REQUIRE(u_out && u_in && v && kstar, "NULL array argument");
  GT_REQUIRE(u_out != u_in, "u_new must not alias u_old");
  const Strides3 st3 = order_strides(order, Nx, Ny, Nz);
  const size_t n = (size_t)Nx * Ny * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_u_kernel<double>, stream_grid(n), dim3(256), 0, st, AS(double *, u_out),
                       AS(double *, u_in), AS(const double *, v), AS(const double *, kstar), tau,
                       norming, zdist, Nx, Ny, Nz, order, st3, extra);
  else
    hipLaunchKernelGGL(tv_u_kernel<float>, stream_grid(n), dim3(256), 0, st, AS(float *, u_out),
                       AS(float *, u_in), AS(const float *, v), AS(const float *, kstar), tau,
                       norming, zdist, Nx, Ny, Nz, order, st3, extra);
  return check_launch(extra ? "tv_primal" : "tv_update_u");
}

extern "C" int gt_tv_update_u(gt_stream_t stream, int dtype, void *u, const void *u_old,
                              const void *v, const void *kstarlambda, float tau, float norming,
                              float z_distance, int Nx, int Ny, int Nz, int order) {
  return tv_u_common(stream, dtype, u, const_cast<void *>(u_old), v, kstarlambda, tau, norming,
                     z_distance, Nx, Ny, Nz, order, 0);
}

extern "C" int gt_tv_primal(gt_stream_t stream, int dtype, void *u_new, void *u_old,
                            const void *v, const void *kstarlambda, float tau, float norming,
                            float z_distance, int Nx, int Ny, int Nz, int order) {
  return tv_u_common(stream, dtype, u_new, u_old, v, kstarlambda, tau, norming, z_distance, Nx,
                     Ny, Nz, order, 1);
}

extern "C" int gt_tv_update_extra(gt_stream_t stream, int dtype, void *u_, void *u, size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(u_ && u, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_extra_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(const double *, u_), AS(double *, u), n);
  else
    hipLaunchKernelGGL(tv_extra_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(const float *, u_), AS(float *, u), n);
  return check_launch("tv_update_extra");
}

// ---- C ABI of the TV step adjoint -------------------------------------------
extern "C" int gt_tv_adj_primal(gt_stream_t stream, int dtype, void *wb, void *ub,
                                const void *ub_next, const void *ubarb_next, const void *u_next,
                                size_t n) {
  int rc = check_dtype(dtype);
  if (rc) return rc;
  GT_REQUIRE(wb && ub && ub_next && ubarb_next && u_next, "NULL array argument");
  if (n == 0) return GT_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_adj_primal_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, wb), AS(double *, ub), AS(const double *, ub_next),
                       AS(const double *, ubarb_next), AS(const double *, u_next), n);
  else
    hipLaunchKernelGGL(tv_adj_primal_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, wb), AS(float *, ub), AS(const float *, ub_next),
                       AS(const float *, ubarb_next), AS(const float *, u_next), n);
  return check_launch("tv_adj_primal");
}

extern "C" int gt_tv_adj_dual(gt_stream_t stream, int dtype, void *ab, const void *vb_next,
                              const void *wb, const void *v, const void *ubar, float sigma,
                              float tau, float z_distance, int Nx, int Ny, int Nz, int order) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Nx, Ny, Nz, order))) return rc;
  GT_REQUIRE(ab && vb_next && wb && v && ubar, "NULL array argument");
  GT_REQUIRE(ab != vb_next && ab != v, "ab must not alias its inputs");
  const Strides3 st3 = order_strides(order, Nx, Ny, Nz);
  const size_t n = (size_t)Nx * Ny * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_adj_dual_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, ab), AS(const double *, vb_next), AS(const double *, wb),
                       AS(const double *, v), AS(const double *, ubar), sigma, tau, z_distance, Nx,
                       Ny, Nz, order, st3);
  else
    hipLaunchKernelGGL(tv_adj_dual_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, ab), AS(const float *, vb_next), AS(const float *, wb),
                       AS(const float *, v), AS(const float *, ubar), sigma, tau, z_distance, Nx,
                       Ny, Nz, order, st3);
  return check_launch("tv_adj_dual");
}

extern "C" int gt_tv_adj_div(gt_stream_t stream, int dtype, void *ubarb, const void *bterm,
                             double coef, const void *ab, float sigma, float z_distance, int Nx,
                             int Ny, int Nz, int order) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Nx, Ny, Nz, order))) return rc;
  GT_REQUIRE(ubarb && bterm && ab, "NULL array argument");
  const Strides3 st3 = order_strides(order, Nx, Ny, Nz);
  const size_t n = (size_t)Nx * Ny * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_adj_div_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, ubarb), AS(const double *, bterm), coef, AS(const double *, ab),
                       sigma, z_distance, Nx, Ny, Nz, order, st3);
  else
    hipLaunchKernelGGL(tv_adj_div_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, ubarb), AS(const float *, bterm), (float)coef,
                       AS(const float *, ab), sigma, z_distance, Nx, Ny, Nz, order, st3);
  return check_launch("tv_adj_div");
}

extern "C" int gt_tv_adj_lambda(gt_stream_t stream, int dtype, void *lamb, void *fb, void *winv,
                                const void *lamb_next, const void *rwb, const void *angle_weights,
                                double coef_r, float c, float sigma_p, int Ns, int Na, int Nz,
                                int order) {
  int rc = check_dtype(dtype);
  if (rc || (rc = check_img(Ns, Na, Nz, order))) return rc;
  GT_REQUIRE(lamb && fb && winv && lamb_next && rwb && angle_weights, "NULL array argument");
  const Strides3 so = order_strides(order, Ns, Na, Nz);
  const size_t n = (size_t)Ns * Na * Nz;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == GT_FLOAT64)
    hipLaunchKernelGGL(tv_adj_lambda_kernel<double>, stream_grid(n), dim3(256), 0, st,
                       AS(double *, lamb), AS(double *, fb), AS(double *, winv),
                       AS(const double *, lamb_next), AS(const double *, rwb),
                       AS(const double *, angle_weights), coef_r, c, sigma_p, Ns, Na, Nz, so);
  else
    hipLaunchKernelGGL(tv_adj_lambda_kernel<float>, stream_grid(n), dim3(256), 0, st,
                       AS(float *, lamb), AS(float *, fb), AS(float *, winv),
                       AS(const float *, lamb_next), AS(const float *, rwb),
                       AS(const float *, angle_weights), (float)coef_r, c, sigma_p, Ns, Na, Nz, so);
  return check_launch("tv_adj_lambda");
}
