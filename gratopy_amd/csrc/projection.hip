// Parallel-beam and fanbeam forward/backprojection for gfx950: the
// single-problem launches of the C ABI.  The device bodies (quadrature,
// row ranges, bin groups, backprojection gather) live in kernels.h and are
// shared with the batched launches.
//
// Forward: one lane per detector bin (64 bins x 4 angles per block) or per
// bin group (256 consecutive (angle, group) pairs per block); the slice is
// the last grid axis.  Rays closer to the x axis read an x-fastest copy of the
// image, the others a y-fastest copy, so a wave always reads along the
// fastest axis (gratopy/radon.cl:66-161, gratopy/fanbeam.cl:72-256).
// Backward: a block is 16 x 16*RY pixels; a prologue builds the per-angle
// affine table and the zero-padded, pre-weighted sinogram
// (gratopy/radon.cl:183-241, gratopy/fanbeam.cl:290-391).

#include <algorithm>
#include <mutex>
#include <string>

#include "kernels.h"

namespace gt {

// pad > 0: the copy's rows along its fast axis are pad elements longer than
// the image; the blocks at the end of the fast axis also zero those elements
// of their 32 rows (so no separate fill of the padding is needed).
template <typename real>
__global__ __launch_bounds__(256) void relayout_kernel(real *__restrict__ dst, Strides3 ds,
                                                       const real *__restrict__ src, Strides3 ss,
                                                       int N0, int N1, int pad) {
  __shared__ real tile[32][33];
  const int64_t z = blockIdx.z;
  relayout_tile(tile, dst, ds, src, ss, N0, N1, z, blockIdx.x * 32, blockIdx.y * 32);
  if (pad) {
    const bool fast0 = ds.s0 <= ds.s1;
    const int nf = fast0 ? N0 : N1, ns = fast0 ? N1 : N0;
    const int bf = (fast0 ? blockIdx.x : blockIdx.y) * 32, bs = (fast0 ? blockIdx.y : blockIdx.x) * 32;
    if (bf + 32 >= nf) {
      for (int t = threadIdx.x; t < 32 * pad; t += 256) {
        const int r = bs + t / pad, f = nf + t % pad;
        if (r < ns) dst[(fast0 ? f : r) * ds.s0 + (fast0 ? r : f) * ds.s1 + z * ds.s2] = 0;
      }
    }
  }
}

template <typename real>
static void launch_relayout(hipStream_t st, real *dst, Strides3 ds, const real *src, Strides3 ss,
                            int N0, int N1, int N2, int pad = 0) {
  dim3 grid((N0 + 31) / 32, (N1 + 31) / 32, N2);
  hipLaunchKernelGGL(relayout_kernel<real>, grid, dim3(256), 0, st, dst, ds, src, ss, N0, N1, pad);
}

// One lane per detector bin (the reference's loops): 64 bins x 4 angles.
template <typename real, bool FAN>
__global__ __launch_bounds__(256) void fwd_ray_kernel(Job<real> J) {
  const int a = blockIdx.y * 4 + threadIdx.y;
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= J.Ns || a >= J.Na) return;
  if (FAN)
    fan_fwd_ray(J, s, a, blockIdx.z);
  else
    radon_fwd_ray(J, s, a, blockIdx.z);
}

// Bin groups: a lane owns bins [K*g, K*g + K) of angle a (group_lane) and
// reads its strip rows CH pixels per load.
template <typename real, bool FAN, int K, int CH, bool BUFW = false>
__global__ __launch_bounds__(256) void fwd_group_kernel(Job<real> J) {
  __shared__ real slots[Slots<K>::kCount * 256];
  int g;
  const int a = group_lane<K>(blockIdx.x, J.Ns, J.Na, g);
  if (a >= J.Na) return;
  real *acc = slots + threadIdx.y * 64 + threadIdx.x;
  if (FAN)
    fan_fwd_group<real, K, false, CH, BUFW>(J, g * K, a, blockIdx.y, acc);
  else
    radon_fwd_group<real, K, false, CH, BUFW>(J, g * K, a, blockIdx.y, acc);
}

// Banded form of the bin-group forward: blockIdx.x = band + nbands * block,
// so consecutive blocks, which the dispatcher hands to the 8 XCDs in turn,
// walk different row bands and each XCD's L2 serves one band of the image.
template <typename real, bool FAN, int K, int CH, bool BUFW = false>
__global__ __launch_bounds__(256) void fwd_band_kernel(Job<real> J, real *part, int nbands) {
  __shared__ real slots[Slots<K>::kCount * 256];
  const int band = blockIdx.x % nbands;
  int g;
  const int a = group_lane<K>(blockIdx.x / nbands, J.Ns, J.Na, g);
  if (a >= J.Na) return;
  real *acc = slots + threadIdx.y * 64 + threadIdx.x;
  const Band<real> B{band, nbands, part};
  if (FAN)
    fan_fwd_group<real, K, true, CH, BUFW>(J, g * K, a, blockIdx.y, acc, B);
  else
    radon_fwd_group<real, K, true, CH, BUFW>(J, g * K, a, blockIdx.y, acc, B);
}

// ---- compacted banded forward ----------------------------------------------
// In a row band only some bin groups of an angle have strip pixels: with the
// parallel beam's detector spanning the image diagonal a third of the groups
// miss the image, and at oblique angles a band's shadow on the detector is
// narrower still.  A prologue lists, per band, the groups that can meet it,
// and the projection launch runs only those, packed into full waves.
struct FwdPlan {
  int2 *range;   // [band][angle]: active groups [x, y)
  int2 *counts;  // [band]: x = column entries used (a multiple of kPlanAngles), y = packed lanes used
  int2 *waves;   // [band][wcap]: (angle, first group) of one angle's column, or (-1, 0)
  int2 *tail;    // [band][tcap]: (angle, group) of one lane
  int64_t wcap, tcap;
};

// A block of the compacted launch walks one column of kPlanCols groups of
// kPlanAngles adjacent angles (column_lane), as a whole-column block of the
// full launch does.  Angles per block measured on the 32 x 2048^2 stacks
// (fanbeam / parallel-beam forward, ms): 4: 78.1 / 72.0, 8: 75.6 / 70.2,
// 16: 74.4 / 69.8, 32: 75.0 / 69.3, 64: 80.4 / 71.5.
constexpr int kPlanAngles = kColumnAngles;
constexpr int kPlanCols = kColumnGroups;

// Capacities of a plan: whole columns of active groups per angle, in sets of
// kPlanAngles angles, and up to kPlanCols - 1 leftover groups per angle.
static void plan_caps(int Ns, int Na, int K, int64_t &wcap, int64_t &tcap) {
  const int64_t groups = (Ns + K - 1) / K;
  wcap = kPlanAngles * (int64_t)((Na + kPlanAngles - 1) / kPlanAngles) * (groups / kPlanCols);
  tcap = (int64_t)Na * std::min<int64_t>(kPlanCols - 1, groups);
}

static size_t plan_bytes(int Ns, int Na, int K, int nb) {
  int64_t wcap, tcap;
  plan_caps(Ns, Na, K, wcap, tcap);
  return align_up(sizeof(int2) * (size_t)nb * ((size_t)Na + 1 + wcap + tcap));
}

// Groups of angle a whose union strip (bins gK - 1 .. gK + K) can meet a
// pixel of row band b.  The detector coordinate of a pixel centre is linear
// (parallel beam) or linear-fractional with a denominator of one sign over
// the image (fanbeam: the source lies outside it), so over the band's
// rectangle it takes its extremes at the corners.  Fanbeam groups pick their
// frame (rows along y or along x) from their own rays, so both frames' bands
// are covered.  Two bins of margin cover the projection kernels' rounding;
// non-finite coordinates give every group.
template <typename real>
__device__ int2 band_group_range(const Job<real> &J, int K, int groups, int b, int nb, int a) {
  const real *o = J.ofs + 8 * (int64_t)a;
  double lo = INFINITY, hi = -INFINITY;
  bool bad = false, any = false;
  auto corner = [&](double x, double y) {
    double s;
    if (J.fan) {
      const double ux = x - (double)J.geo[3] - (double)o[2], uy = y - (double)J.geo[4] - (double)o[3];
      const double wx = (double)o[4] - (double)o[2], wy = (double)o[5] - (double)o[3];
      s = (double)J.geo[5] + (wx * uy - wy * ux) / ((double)o[1] * ux - (double)o[0] * uy);
    } else {
      s = (double)o[0] * x + (double)o[1] * y + (double)o[2];
    }
    bad = bad || !(fabs(s) < 1e300);
    lo = fmin(lo, s);
    hi = fmax(hi, s);
  };
  // frame with rows along y (band over y) and, where it applies, along x
  const bool rows_y = J.fan || o[5] == (real)0, rows_x = J.fan || o[5] != (real)0;
  if (rows_y) {
    const int r0 = (int)((int64_t)J.Ny * b / nb), r1 = (int)((int64_t)J.Ny * (b + 1) / nb) - 1;
    if (r0 <= r1) {
      any = true;
      corner(0, r0), corner(J.Nx - 1, r0), corner(0, r1), corner(J.Nx - 1, r1);
    }
  }
  if (rows_x) {
    const int r0 = (int)((int64_t)J.Nx * b / nb), r1 = (int)((int64_t)J.Nx * (b + 1) / nb) - 1;
    if (r0 <= r1) {
      any = true;
      corner(r0, 0), corner(r1, 0), corner(r0, J.Ny - 1), corner(r1, J.Ny - 1);
    }
  }
  if (!any) return make_int2(0, 0);
  if (bad) return make_int2(0, groups);
  const double m = 2.0;
  const double g0 = fmax(floor((lo - m) / K) - 1.0, 0.0);
  const double g1 = fmin(floor((hi + m) / K) + 2.0, (double)groups);
  return g0 < g1 ? make_int2((int)g0, (int)g1) : make_int2(0, 0);
}

// Exclusive prefix sum of v over a 256-thread block; *total = the block's sum.
__device__ inline int block_scan_256(int v, int *total) {
  __shared__ int wt[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wt[w] = x;
  __syncthreads();
  int off = 0;
  for (int i = 0; i < w; ++i) off += wt[i];
  *total = wt[0] + wt[1] + wt[2] + wt[3];
  __syncthreads();
  return off + x - v;
}

// One block per row band: the active range of every angle, then the wave
// table -- per set of kPlanAngles angles, column c of each of its angles that
// has it (entry kPlanAngles (set offset + c) + j), so a block walks the same
// column of adjacent angles -- and the leftover groups (count mod kPlanCols)
// of all angles packed in angle order.
template <typename real>
__global__ __launch_bounds__(256) void fwd_plan_kernel(Job<real> J, int K, int nb, FwdPlan P) {
  constexpr int A = kPlanAngles, C = kPlanCols;
  const int b = blockIdx.x, Na = J.Na;
  const int groups = (J.Ns + K - 1) / K;
  int2 *range = P.range + (int64_t)b * Na;
  int2 *waves = P.waves + (int64_t)b * P.wcap;
  int2 *tail = P.tail + (int64_t)b * P.tcap;
  for (int a = threadIdx.x; a < Na; a += 256) range[a] = band_group_range(J, K, groups, b, nb, a);
  __syncthreads();
  const int nq = (Na + A - 1) / A;
  int used = 0, total;
  for (int base = 0; base < nq; base += 256) {
    const int q = base + threadIdx.x;
    int cols[A] = {}, mc = 0;
    if (q < nq)
      for (int j = 0; j < A; ++j)
        if (A * q + j < Na) {
          const int2 r = range[A * q + j];
          cols[j] = (r.y - r.x) / C;
          mc = cols[j] > mc ? cols[j] : mc;
        }
    const int off = used + block_scan_256(mc, &total);
    for (int c = 0; c < mc; ++c)
      for (int j = 0; j < A; ++j) {
        const int a = A * q + j;
        waves[A * (int64_t)(off + c) + j] =
            c < cols[j] ? make_int2(a, range[a].x + C * c) : make_int2(-1, 0);
      }
    used += total;
  }
  const int entries = A * used;
  used = 0;
  for (int base = 0; base < Na; base += 256) {
    const int a = base + threadIdx.x;
    int n = 0, g0 = 0;
    if (a < Na) {
      const int2 r = range[a];
      n = (r.y - r.x) % C;
      g0 = r.y - n;
    }
    const int off = used + block_scan_256(n, &total);
    for (int j = 0; j < n; ++j) tail[(int64_t)off + j] = make_int2(a, g0 + j);
    used += total;
  }
  if (threadIdx.x == 0) P.counts[b] = make_int2(entries, used);
}

// Banded forward over a plan: block blk of band b is a column block while
// blk < entries / kPlanAngles (its angles: entries kPlanAngles blk + j, lanes
// by column_lane), then a block of 256 packed lanes.
template <typename real, bool FAN, int K, int CH, bool BUFW>
__global__ __launch_bounds__(256) void fwd_band_compact_kernel(Job<real> J, real *part, int nbands,
                                                               FwdPlan P) {
  __shared__ real slots[Slots<K>::kCount * 256];
  const int band = blockIdx.x % nbands;
  const int blk = blockIdx.x / nbands;
  const int2 cnt = P.counts[band];
  const int fb = cnt.x / kPlanAngles;
  int a, g;
  if (blk < fb) {
    int goff;
    const int j = column_lane(goff);
    const int2 e = P.waves[(int64_t)band * P.wcap + kPlanAngles * blk + j];
    if (e.x < 0) return;
    a = e.x;
    g = e.y + goff;
  } else {
    const int64_t k = (int64_t)(blk - fb) * 256 + threadIdx.y * 64 + threadIdx.x;
    if (k >= cnt.y) return;
    const int2 e = P.tail[(int64_t)band * P.tcap + k];
    a = e.x;
    g = e.y;
  }
  real *acc = slots + threadIdx.y * 64 + threadIdx.x;
  const Band<real> B{band, nbands, part};
  if (FAN)
    fan_fwd_group<real, K, true, CH, BUFW>(J, g * K, a, blockIdx.y, acc, B);
  else
    radon_fwd_group<real, K, true, CH, BUFW>(J, g * K, a, blockIdx.y, acc, B);
}

template <typename real>
__global__ __launch_bounds__(256) void band_reduce_kernel(Job<real> J, const real *part,
                                                          int nbands, const int2 *range, int K) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s < J.Ns) band_reduce_entry(J, part, nbands, s, blockIdx.y, blockIdx.z, range, K);
}

// Backprojection prologue, block (j block, angle, slice): the zero-padded,
// pre-weighted sinogram row, and (first block of the angle's row in slice 0)
// the angle's affine table entry.
template <typename real>
__global__ __launch_bounds__(256) void stage_sino_kernel(real *__restrict__ S,
                                                         AngleRec<real> *__restrict__ tab,
                                                         const real *__restrict__ sino,
                                                         Strides3 si, const real *__restrict__ ofs,
                                                         const real *__restrict__ sdpd,
                                                         const real *__restrict__ geo, int fan,
                                                         int Ns, int Na) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j < Ns + 4) stage_entry(S, sino, si, ofs, sdpd, geo, fan, Ns, Na, j, blockIdx.y, blockIdx.z);
  if (j == 0 && blockIdx.z == 0) back_table_entry(tab, ofs, geo, fan, blockIdx.y);
}

template <typename real, bool FAN, int RY, bool BUF>
__global__ __launch_bounds__(256) void back_kernel(Job<real> J) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y0 = (blockIdx.y * 16 + (threadIdx.x >> 4)) * RY;
  if (x >= J.Nx || y0 >= J.Ny) return;
  back_pixels<real, FAN, RY, BUF>(J, x, y0, blockIdx.z);
}

// Banded form: blockIdx.x = band + nbands * column block; band b sums the
// angles [b*Na/nbands, (b+1)*Na/nbands), so each XCD's L2 serves the padded
// sinogram rows of one angle band.
template <typename real, bool FAN, int RY, bool BUF>
__global__ __launch_bounds__(256) void back_band_kernel(Job<real> J, real *part, int nbands) {
  const int band = blockIdx.x % nbands;
  const int x = (blockIdx.x / nbands) * 16 + (threadIdx.x & 15);
  const int y0 = (blockIdx.y * 16 + (threadIdx.x >> 4)) * RY;
  if (x >= J.Nx || y0 >= J.Ny) return;
  const int a0 = (int)((int64_t)J.Na * band / nbands);
  const int a1 = (int)((int64_t)J.Na * (band + 1) / nbands);
  back_pixels<real, FAN, RY, BUF>(J, x, y0, blockIdx.z, a0, a1, part, band);
}

template <typename real>
__global__ __launch_bounds__(256) void back_reduce_kernel(Job<real> J, const real *part,
                                                          int nbands) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x < J.Nx && y < J.Ny) band_back_reduce_entry(J, part, nbands, x, y, blockIdx.z);
}

// LDS-staged form (back_staged_tile): one tile of TX x (256/TX)*RY pixels per block.
template <typename real, bool FAN, int RY, int CA, int TX>
__global__ __launch_bounds__(256, kStagedWavesPerSimd) void back_staged_kernel(Job<real> J) {
  __shared__ real rows[CA * kStageWidth];
  __shared__ StageWindow<real> win[CA];
  back_staged_tile<real, FAN, RY, CA, TX>(J, blockIdx.x, blockIdx.y, blockIdx.z, rows, win);
}

// LDS-staged and angle-banded: blockIdx.x = band + nbands * tile column (as
// back_band_kernel); each block stages and sums the angles of its band into
// the band partials, which back_reduce_kernel adds in band order.
template <typename real, bool FAN, int RY, int CA, int TX>
__device__ __forceinline__ void back_staged_band_body(const Job<real> &J, real *part, int nbands) {
  __shared__ real rows[CA * kStageWidth];
  __shared__ StageWindow<real> win[CA];
  const int band = blockIdx.x % nbands;
  const int a0 = (int)((int64_t)J.Na * band / nbands);
  const int a1 = (int)((int64_t)J.Na * (band + 1) / nbands);
  back_staged_tile<real, FAN, RY, CA, TX>(J, blockIdx.x / nbands, blockIdx.y, blockIdx.z, rows, win,
                                          a0, a1, part, band);
}

// Grids of at least kStagedCapBlocks blocks: held to kStagedWavesPerSimd, as
// back_staged_kernel (a 2048^2 fan slice, 16384 blocks: 1.50 against 1.58 ms
// uncapped).
template <typename real, bool FAN, int RY, int CA, int TX>
__global__ __launch_bounds__(256, kStagedWavesPerSimd) void back_staged_band_kernel(Job<real> J, real *part,
                                                                                    int nbands) {
  back_staged_band_body<real, FAN, RY, CA, TX>(J, part, nbands);
}

// Smaller grids leave CUs short of blocks, so the cap would only cost
// registers: uncapped (a 512^2 fan slice, 1024 blocks: 61 against 78 us).
template <typename real, bool FAN, int RY, int CA, int TX>
__global__ __launch_bounds__(256) void back_staged_band_small_kernel(Job<real> J, real *part, int nbands) {
  back_staged_band_body<real, FAN, RY, CA, TX>(J, part, nbands);
}

// 8 blocks per CU of the 256
constexpr int64_t kStagedCapBlocks = 2048;

template <typename real, bool FAN, int RY, int CA, int TX>
static void launch_staged_tx(hipStream_t st, const Job<real> &J, int nbands, real *part) {
  const dim3 grid(((J.Nx + TX - 1) / TX) * nbands, (J.Ny + (256 / TX) * RY - 1) / ((256 / TX) * RY),
                  J.Nz);
  if (nbands > 1) {
    if ((int64_t)grid.x * grid.y * grid.z >= kStagedCapBlocks)
      hipLaunchKernelGGL((back_staged_band_kernel<real, FAN, RY, CA, TX>), grid, dim3(256), 0, st, J,
                         part, nbands);
    else
      hipLaunchKernelGGL((back_staged_band_small_kernel<real, FAN, RY, CA, TX>), grid, dim3(256), 0,
                         st, J, part, nbands);
    hipLaunchKernelGGL(back_reduce_kernel<real>, dim3((J.Nx + 15) / 16, (J.Ny + 15) / 16, J.Nz),
                       dim3(256), 0, st, J, part, nbands);
  } else {
    hipLaunchKernelGGL((back_staged_kernel<real, FAN, RY, CA, TX>), grid, dim3(256), 0, st, J);
  }
}

template <typename real, bool FAN, int RY, int CA>
static void launch_staged(hipStream_t st, const Job<real> &J, int tx, int nbands, real *part) {
  if (tx == 32)
    launch_staged_tx<real, FAN, RY, CA, 32>(st, J, nbands, part);
  else
    launch_staged_tx<real, FAN, RY, CA, 16>(st, J, nbands, part);
}


template <typename real, bool FAN, int RY>
static void launch_back(hipStream_t st, const Job<real> &J, int nbands, int stage, int tx,
                        real *part) {
  const dim3 grid(((J.Nx + 15) / 16) * nbands, (J.Ny + 16 * RY - 1) / (16 * RY), J.Nz);
  const bool buf = buffer_fits(sizeof(real), J.Ns, J.Na);
  if (stage == 8) {
    launch_staged<real, FAN, RY, 8>(st, J, tx, nbands, part);
  } else if (stage == 16) {
    launch_staged<real, FAN, RY, 16>(st, J, tx, nbands, part);
  } else if (nbands > 1) {
    if (buf)
      hipLaunchKernelGGL((back_band_kernel<real, FAN, RY, true>), grid, dim3(256), 0, st, J, part,
                         nbands);
    else
      hipLaunchKernelGGL((back_band_kernel<real, FAN, RY, false>), grid, dim3(256), 0, st, J, part,
                         nbands);
    hipLaunchKernelGGL(back_reduce_kernel<real>, dim3((J.Nx + 15) / 16, (J.Ny + 15) / 16, J.Nz),
                       dim3(256), 0, st, J, part, nbands);
  } else if (buf) {
    hipLaunchKernelGGL((back_kernel<real, FAN, RY, true>), grid, dim3(256), 0, st, J);
  } else {
    hipLaunchKernelGGL((back_kernel<real, FAN, RY, false>), grid, dim3(256), 0, st, J);
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
  bool fan;  // fanbeam geometry (the automatic launch shapes depend on it)
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

// Launch shape of one call: the tuning knobs, or (GT_TUNE_AUTO, the default)
// a choice by the geometry and the work of the call, Nx*Ny*Na*Nz pixel-angles,
// from autotuner trials on MI355X (256^2 to 32 x 2048^2, parallel and fan
// beam): small problems need row or angle bands to fill the GPU; larger ones
// run the chunked forward and the LDS-staged backprojection.
struct FwdShape {
  int K, bands, chunk;  // bands and chunked reads need bin groups (K > 1)
  bool compact;         // banded: launch only the groups that meet each band (FwdPlan)
};
struct BackShape {
  int RY, bands, stage, tile;  // stage > 0: LDS-staged, with or without angle bands
};

static double work_of(const Layout &L) { return (double)L.Nx * L.Ny * L.Na * L.Nz; }

static FwdShape fwd_shape(const Layout &L) {
  const Tuning &t = tuning();
  if (!t.automatic) {
    const bool g = t.fwd_group > 1;
    // K = 32 exists with chunked reads only
    const int chunk = g ? (t.fwd_group == 32 ? 4 : t.fwd_chunk) : 1;
    const int bands = g ? t.fwd_bands : 1;
    return {t.fwd_group, bands, chunk, bands > 1 && t.fwd_compact != 0};
  }
  const double w = work_of(L);
  // tiny calls: no row-padded copies (one image copy fewer to launch); the
  // row bands stay, without them a tiny launch has a handful of blocks
  if (w < 2e6 || (L.fan && w < 5e7)) return {4, 8, 1, false};
  const int K = w < 2e9 ? 8 : 16;
  // Row bands: the fewest of 4, 8, 16, 32 that give the launch at least
  // 12288 blocks (48 per CU); more than 8 only with at least 16 rows per
  // band.  Bands
  // multiply the blocks of a launch that alone would leave the GPU part idle,
  // but each costs partial sums and strip set-up.  Forward ms over 4 / 8 / 16
  // / 32 bands (16-angle blocks, compacted): one 2048^2 slice, 1440 angles:
  // fanbeam 2.96 / 2.71 / 2.61 / 2.78, parallel beam 2.86 / 2.49 / 2.32 / 2.37;
  // 32 such slices: fanbeam 71.0 / 72.0 / 74.5 / -, parallel beam 70.1 /
  // 69.5 / 70.0 / -; one 512^2 fan slice, 360 angles: - / 0.158 / 0.142 / 0.144.
  const int64_t gb = group_blocks(L.Ns, L.Na, K) * L.Nz;
  int bands = 4;
  while (bands < 32 && gb * bands < 12288 && (bands < 8 || std::min(L.Nx, L.Ny) >= 32 * bands))
    bands *= 2;
  // compaction pays once the prologue's few microseconds are small against
  // the projection
  return {K, bands, 4, w >= 1e9};
}

static BackShape back_shape(const Layout &L) {
  const Tuning &t = tuning();
  if (!t.automatic) return {t.back_rows, t.back_bands, t.back_stage, t.back_tile};
  if (work_of(L) < 2e6) return {4, 1, 0, 16};  // launch-bound: no band reduction
  if (work_of(L) < 5e7) return {4, 8, 0, 16};
  // Staged tiles hold 2048 pixels at either width; angle bands while the
  // tiles of the call fill the GPU fewer than ~3 times (4096 tiles).
  const int bands = (double)L.Nx * L.Ny * L.Nz < 4096.0 * 2048.0 ? 8 : 1;
  return {8, bands, 16, L.fan ? 32 : 16};
}

// Chunked strip reads (chunk = 4) read 16 bytes per load and need both image
// copies with every row padded by kRowPad zeros; otherwise an F-ordered input
// is read in place.
constexpr int kRowPad = 4;

static size_t fwd_ws(int dtype, const Layout &L, int img_order) {
  const size_t e = elem_size(dtype);
  const FwdShape sh = fwd_shape(L);
  size_t n = 0;
  if (sh.chunk > 1) {
    n += align_up(e * (size_t)(L.Nx + kRowPad) * L.Ny * L.Nz);  // x-fastest padded copy
    n += align_up(e * (size_t)(L.Ny + kRowPad) * L.Nx * L.Nz);  // y-fastest padded copy
  } else {
    const size_t img = align_up(e * (size_t)L.Nx * L.Ny * L.Nz);
    if (img_order != GT_ORDER_F) n += img;                  // x-fastest copy
    if (!(img_order == GT_ORDER_C && L.Nz == 1)) n += img;  // y-fastest copy
  }
  if (sh.compact) n += plan_bytes(L.Ns, L.Na, sh.K, sh.bands);               // group plan
  if (sh.bands > 1) n += align_up(e * (size_t)sh.bands * L.Ns * L.Na * L.Nz);  // band partials
  return n;
}

static size_t back_partials(int dtype, const Layout &L) {
  const int nb = back_shape(L).bands;
  return nb > 1 ? align_up(elem_size(dtype) * (size_t)nb * L.Nx * L.Ny * L.Nz) : 0;
}

static size_t back_ws(int dtype, const Layout &L) {
  return align_up(elem_size(dtype) * 8 * (size_t)L.Na) +
         align_up(elem_size(dtype) * (size_t)(L.Ns + 4) * L.Na * L.Nz) + back_partials(dtype, L);
}

template <typename real>
static Job<real> make_job(bool fan, const Layout &L, const real *ofs, const real *sdpd,
                          const real *geo) {
  Job<real> J{};
  J.fan = fan;
  J.Nx = L.Nx;
  J.Ny = L.Ny;
  J.Ns = L.Ns;
  J.Na = L.Na;
  J.Nz = L.Nz;
  J.ofs = ofs;
  J.sdpd = sdpd;
  J.geo = geo;
  return J;
}

// K = 32 comes with chunked reads (16-byte strip loads) only.
// bufw: the chunked strip walks go by 32-bit offsets into one buffer
// descriptor over both image copies (strip_buffer_fits).
template <typename real, bool FAN, int K>
static int launch_fwd_group(hipStream_t st, const Job<real> &J, int nb, bool chunk, real *part,
                            const FwdPlan *plan, bool bufw) {
  constexpr int CH = 16 / sizeof(real);
  const dim3 block(64, 4);
  if (plan) {
    const int64_t per_band = plan->wcap / kPlanAngles + (plan->tcap + 255) / 256;
    GT_REQUIRE(per_band * nb < ((int64_t)1 << 31), "too many detector groups");
    hipLaunchKernelGGL(fwd_plan_kernel<real>, dim3(nb), dim3(256), 0, st, J, K, nb, *plan);
    const dim3 grid((unsigned)(per_band * nb), J.Nz);
    if (chunk && bufw)
      hipLaunchKernelGGL((fwd_band_compact_kernel<real, FAN, K, CH, true>), grid, block, 0, st, J,
                         part, nb, *plan);
    else if (chunk)
      hipLaunchKernelGGL((fwd_band_compact_kernel<real, FAN, K, CH, false>), grid, block, 0, st, J,
                         part, nb, *plan);
    else if constexpr (K <= 16)
      hipLaunchKernelGGL((fwd_band_compact_kernel<real, FAN, K, 1, false>), grid, block, 0, st, J,
                         part, nb, *plan);
    return GT_OK;
  }
  const int64_t blocks = group_blocks(J.Ns, J.Na, K);
  GT_REQUIRE(blocks * 256 < ((int64_t)1 << 31) && blocks * nb < ((int64_t)1 << 31),
             "too many detector groups");
  const dim3 grid((unsigned)(blocks * nb), J.Nz);
  if (chunk && bufw) {
    if (nb > 1)
      hipLaunchKernelGGL((fwd_band_kernel<real, FAN, K, CH, true>), grid, block, 0, st, J, part, nb);
    else
      hipLaunchKernelGGL((fwd_group_kernel<real, FAN, K, CH, true>), grid, block, 0, st, J);
  } else if (chunk) {
    if (nb > 1)
      hipLaunchKernelGGL((fwd_band_kernel<real, FAN, K, CH>), grid, block, 0, st, J, part, nb);
    else
      hipLaunchKernelGGL((fwd_group_kernel<real, FAN, K, CH>), grid, block, 0, st, J);
  } else if constexpr (K <= 16) {
    if (nb > 1)
      hipLaunchKernelGGL((fwd_band_kernel<real, FAN, K, 1>), grid, block, 0, st, J, part, nb);
    else
      hipLaunchKernelGGL((fwd_group_kernel<real, FAN, K, 1>), grid, block, 0, st, J);
  }
  return GT_OK;
}

template <typename real, bool FAN>
static int fwd_impl(hipStream_t st, real *sino, const real *img, const real *ofs,
                    const real *sdpd, const real *geo, const Layout &L, int sino_order,
                    int img_order, void *ws, size_t ws_bytes) {
  const size_t need = fwd_ws(sizeof(real) == 8 ? GT_FLOAT64 : GT_FLOAT32, L, img_order);
  if (ws_bytes < need) return fail(GT_ENOSPC, "projection workspace too small");
  GT_REQUIRE(need == 0 || ws != nullptr, "workspace pointer is NULL");
  const Strides3 is = order_strides(img_order, L.Nx, L.Ny, L.Nz);
  char *w = static_cast<char *>(ws);
  Job<real> J = make_job<real>(FAN, L, ofs, sdpd, geo);
  J.imgX = img;
  J.imgY = img;
  J.rowX = L.Nx;
  J.rowY = L.Ny;
  J.sliceX = J.sliceY = (int64_t)L.Nx * L.Ny;
  const FwdShape sh = fwd_shape(L);
  if (sh.chunk > 1) {
    const int64_t rx = L.Nx + kRowPad, ry = L.Ny + kRowPad;
    real *cx = reinterpret_cast<real *>(w);
    w += align_up(sizeof(real) * (size_t)rx * L.Ny * L.Nz);
    real *cy = reinterpret_cast<real *>(w);
    launch_relayout<real>(st, cx, Strides3{1, rx, rx * L.Ny}, img, is, L.Nx, L.Ny, L.Nz, kRowPad);
    launch_relayout<real>(st, cy, Strides3{ry, 1, ry * L.Nx}, img, is, L.Nx, L.Ny, L.Nz, kRowPad);
    J.imgX = cx;
    J.imgY = cy;
    J.rowX = rx;
    J.rowY = ry;
    J.sliceX = rx * L.Ny;
    J.sliceY = ry * L.Nx;
  } else {
    const Strides3 xs = {1, L.Nx, (int64_t)L.Nx * L.Ny};  // x fastest
    const Strides3 ys = {L.Ny, 1, (int64_t)L.Nx * L.Ny};  // y fastest
    const size_t img_bytes = align_up(sizeof(real) * (size_t)L.Nx * L.Ny * L.Nz);
    if (img_order != GT_ORDER_F) {
      real *cp = reinterpret_cast<real *>(w);
      w += img_bytes;
      launch_relayout<real>(st, cp, xs, img, is, L.Nx, L.Ny, L.Nz);
      J.imgX = cp;
    }
    if (!(img_order == GT_ORDER_C && L.Nz == 1)) {
      real *cp = reinterpret_cast<real *>(w);
      launch_relayout<real>(st, cp, ys, img, is, L.Nx, L.Ny, L.Nz);
      J.imgY = cp;
    }
  }
  J.sino = sino;
  J.so = order_strides(sino_order, L.Ns, L.Na, L.Nz);
  const int K = sh.K, nb = sh.bands;
  const bool chunk = sh.chunk > 1;
  // both padded copies (imgX first, then imgY) in one descriptor's reach
  const bool bufw = sh.chunk > 1 &&
                    strip_buffer_fits((size_t)((const char *)(J.imgY + (int64_t)L.Nz * J.sliceY) -
                                               (const char *)J.imgX));
  real *part = nullptr;
  FwdPlan plan{}, *pp = nullptr;
  if (nb > 1) {
    // the band partials are the last part of the workspace, the group plan
    // (compacted launches) comes before them, after the image copies
    const size_t part_bytes = align_up(sizeof(real) * (size_t)nb * L.Ns * L.Na * L.Nz);
    part = reinterpret_cast<real *>(static_cast<char *>(ws) + (need - part_bytes));
    if (sh.compact) {
      int2 *q = reinterpret_cast<int2 *>(static_cast<char *>(ws) + (need - part_bytes -
                                                                     plan_bytes(L.Ns, L.Na, K, nb)));
      plan_caps(L.Ns, L.Na, K, plan.wcap, plan.tcap);
      plan.range = q;
      plan.counts = q + (int64_t)nb * L.Na;
      plan.waves = plan.counts + nb;
      plan.tail = plan.waves + nb * plan.wcap;
      pp = &plan;
    }
  }
  if (K == 1) {
    const dim3 grid((L.Ns + 63) / 64, (L.Na + 3) / 4, L.Nz);
    hipLaunchKernelGGL((fwd_ray_kernel<real, FAN>), grid, dim3(64, 4), 0, st, J);
  } else {
    int rc;
    if (K == 32)
      rc = launch_fwd_group<real, FAN, 32>(st, J, nb, chunk, part, pp, bufw);
    else if (K == 16)
      rc = launch_fwd_group<real, FAN, 16>(st, J, nb, chunk, part, pp, bufw);
    else if (K == 8)
      rc = launch_fwd_group<real, FAN, 8>(st, J, nb, chunk, part, pp, bufw);
    else if (K == 4)
      rc = launch_fwd_group<real, FAN, 4>(st, J, nb, chunk, part, pp, bufw);
    else
      rc = launch_fwd_group<real, FAN, 2>(st, J, nb, chunk, part, pp, bufw);
    if (rc) return rc;
  }
  if (nb > 1)
    hipLaunchKernelGGL(band_reduce_kernel<real>, dim3((L.Ns + 255) / 256, L.Na, L.Nz), dim3(256),
                       0, st, J, part, nb, pp ? (const int2 *)plan.range : nullptr, K);
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
  const Strides3 si = order_strides(sino_order, L.Ns, L.Na, L.Nz);
  hipLaunchKernelGGL(stage_sino_kernel<real>, dim3((L.Ns + 4 + 255) / 256, L.Na, L.Nz), dim3(256),
                     0, st, S, tab, sino, si, ofs, sdpd, geo, (int)FAN, L.Ns, L.Na);
  Job<real> J = make_job<real>(FAN, L, ofs, sdpd, geo);
  J.img = img;
  J.io = order_strides(img_order, L.Nx, L.Ny, L.Nz);
  J.S = S;
  J.tab = tab;
  const BackShape sh = back_shape(L);
  const int RY = sh.RY, nb = sh.bands;
  GT_REQUIRE((int64_t)((L.Nx + 15) / 16) * nb < ((int64_t)1 << 31), "image too wide");
  // the band partials are the last part of the workspace
  const size_t part_bytes = back_partials(sizeof(real) == 8 ? GT_FLOAT64 : GT_FLOAT32, L);
  real *part = reinterpret_cast<real *>(w + (need - part_bytes));
  if (RY == 8)
    launch_back<real, FAN, 8>(st, J, nb, sh.stage, sh.tile, part);
  else if (RY == 4)
    launch_back<real, FAN, 4>(st, J, nb, sh.stage, sh.tile, part);
  else if (RY == 2)
    launch_back<real, FAN, 2>(st, J, nb, sh.stage, sh.tile, part);
  else
    launch_back<real, FAN, 1>(st, J, nb, sh.stage, sh.tile, part);
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
  if (Nx <= 0 || Ny <= 0 || Ns <= 0 || Na <= 0 || Nz <= 0) return 0;
  const Layout L{Nx, Ny, Ns, Na, Nz, geometry == GT_GEOMETRY_FAN};
  return adjoint ? back_ws(dtype, L) : fwd_ws(dtype, L, in_order);
}

extern "C" int gt_projection_shape(int geometry, int adjoint, int Nx, int Ny, int Ns, int Na,
                                   int Nz, int *shape) {
  if (!shape) return fail(GT_EINVAL, "shape pointer is NULL");
  const Layout L{Nx, Ny, Ns, Na, Nz, geometry == GT_GEOMETRY_FAN};
  int rc = check_sizes(L);
  if (rc) return rc;
  if (adjoint) {
    const BackShape b = back_shape(L);
    shape[0] = b.RY, shape[1] = b.bands, shape[2] = b.stage, shape[3] = b.tile;
  } else {
    const FwdShape f = fwd_shape(L);
    shape[0] = f.K, shape[1] = f.bands, shape[2] = f.chunk, shape[3] = f.compact ? 1 : 0;
  }
  return GT_OK;
}

#define GT_DISPATCH(dtype, CALL_F32, CALL_F64) \
  ((dtype) == GT_FLOAT64 ? (CALL_F64) : (CALL_F32))

extern "C" int gt_radon_fwd(gt_stream_t stream, int dtype, void *sino, const void *img,
                            const void *ofs, const void *geo, int Nx, int Ny, int Ns, int Na,
                            int Nz, int sino_order, int img_order, void *workspace,
                            size_t workspace_bytes) {
  int rc = check_common(dtype, sino, img, ofs, geo, sino_order, img_order);
  if (rc) return rc;
  const Layout L{Nx, Ny, Ns, Na, Nz, false};
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
  const Layout L{Nx, Ny, Ns, Na, Nz, true};
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
  const Layout L{Nx, Ny, Ns, Na, Nz, false};
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
  const Layout L{Nx, Ny, Ns, Na, Nz, true};
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

// ---------------------------------------------------------------------------
// Projections without allocation: the workspace is a device array of the
// library (one per GPU, held until the process ends).  The array's name is
// a host-side stand-in in host code; its device address on each GPU comes
// from hipGetSymbolAddress with that GPU current, once, and is kept here.
// ---------------------------------------------------------------------------
namespace gt {

static __device__ __align__(256) unsigned char scratch[GT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the stream's device, which must be current.
static int scratch_address(hipStream_t st, void **out) {
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return fail(GT_EHIP, "hipGetDevice failed");
  hipDevice_t sdev = cur;
  if (st != nullptr && hipStreamGetDevice(st, &sdev) != hipSuccess)
    return fail(GT_EHIP, "hipStreamGetDevice failed");
  if (sdev != cur)
    return fail(GT_EINVAL, "the stream belongs to device " + std::to_string(sdev) + " but device " +
                               std::to_string(cur) + " is current");
  GT_REQUIRE(cur >= 0 && cur < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[cur] == nullptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(GT_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[cur] = p;
  }
  *out = g_scratch_ptr[cur];
  return GT_OK;
}

// One projection (any geometry and direction) on a given workspace.
template <typename real>
static int project_with(hipStream_t st, bool fan, int adjoint, void *out, const void *in,
                        const void *ofs, const void *sdpd, const void *geo, const Layout &L,
                        int out_order, int in_order, void *ws, size_t wb) {
  real *o = static_cast<real *>(out);
  const real *i = static_cast<const real *>(in), *t = static_cast<const real *>(ofs);
  const real *d = static_cast<const real *>(sdpd), *g = static_cast<const real *>(geo);
  if (adjoint)
    return fan ? back_impl<real, true>(st, o, i, t, d, g, L, out_order, in_order, ws, wb)
               : back_impl<real, false>(st, o, i, t, nullptr, g, L, out_order, in_order, ws, wb);
  return fan ? fwd_impl<real, true>(st, o, i, t, d, g, L, out_order, in_order, ws, wb)
             : fwd_impl<real, false>(st, o, i, t, nullptr, g, L, out_order, in_order, ws, wb);
}

}  // namespace gt

extern "C" int gt_projection_noalloc(gt_stream_t stream, int geometry, int adjoint, int dtype,
                                     void *out, const void *in, const void *ofs, const void *sdpd,
                                     const void *geo, int Nx, int Ny, int Ns, int Na, int Nz,
                                     int out_order, int in_order) {
  GT_REQUIRE(geometry == GT_GEOMETRY_RADON || geometry == GT_GEOMETRY_FAN, "unknown geometry");
  GT_REQUIRE(adjoint == 0 || adjoint == 1, "adjoint must be 0 or 1");
  int rc = check_common(dtype, out, in, ofs, geo, out_order, in_order);
  if (rc) return rc;
  const bool fan = geometry == GT_GEOMETRY_FAN;
  GT_REQUIRE(!fan || sdpd != nullptr, "NULL sdpd");
  const Layout L{Nx, Ny, Ns, Na, Nz, fan};
  if ((rc = check_sizes(L))) return rc;
  const size_t need = adjoint ? back_ws(dtype, L) : fwd_ws(dtype, L, in_order);
  if (need > GT_NOALLOC_SCRATCH_BYTES)
    return fail(GT_ENOSPC, "projection needs " + std::to_string(need) +
                               " bytes of workspace; the no-allocation entry point holds " +
                               std::to_string((size_t)GT_NOALLOC_SCRATCH_BYTES) +
                               " (use the entry points with a caller workspace)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  void *ws = nullptr;
  if ((rc = scratch_address(st, &ws))) return rc;
  const size_t wb = GT_NOALLOC_SCRATCH_BYTES;
  if (dtype == GT_FLOAT64)
    return project_with<double>(st, fan, adjoint, out, in, ofs, sdpd, geo, L, out_order, in_order,
                                ws, wb);
  return project_with<float>(st, fan, adjoint, out, in, ofs, sdpd, geo, L, out_order, in_order, ws,
                             wb);
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
