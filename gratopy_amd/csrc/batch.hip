// Batched projections over many 2-D slices of uneven geometry (any mix of
// parallel-beam and fanbeam, image sizes, angle and detector counts) and
// the Landweber / TV primal-dual iterations over them, a handful of launches
// per iteration for the whole batch.
//
// A plan (gt_batch_plan_create) holds, in one device allocation:
//   * job records (struct Job of kernels.h), one per slice and table kind
//     (plain projections, Landweber, TV even / odd iteration);
//   * the per-angle backprojection tables of every slice (built once);
//   * per-slice scratch: a y-fastest copy of the image (and, for chunked
//     strip reads, a row-padded x-fastest copy; rows of both then end in
//     zeros) and the padded, pre-weighted sinogram of the backprojection;
//   * tile lists: one entry (slice, tile coordinates) per thread block of
//     each launch kind, sorted by estimated cost (largest first) so the
//     hardware dispatcher, which hands out blocks in order as slots free,
//     runs a longest-processing-time-first schedule over the uneven slices.
// Every launch is a 1-D grid over one tile list; a block reads its tile and
// its slice's job record with scalar loads and runs the same device bodies
// as the single-problem kernels, so per slice the results are bitwise those
// of projection.hip and of the elementwise solver kernels.

#include <algorithm>
#include <vector>

#include "kernels.h"

struct gt_batch_plan {
  int dtype;
  int n_items;
  int fwd_group, back_rows;
  int fwd_bands;              // > 1: row-banded forward (per-slice partials, ordered reduction)
  int fwd_chunk;              // 4: forward strip rows read from row-padded copies, 16 bytes per load
  int back_stage, back_tile;  // > 0: LDS-staged backprojection (angles per chunk, tile width)
  bool has_data;  // every item has data (Landweber, TV)
  bool has_tv;    // every item has data, img2 and field (TV)
  void *mem;  // one hipMalloc: jobs, tiles, tables, scratch
  size_t bytes;
  const void *jobs[4];  // per table kind: Job<real>[n_items]
  const int4 *tiles[5];
  int n_tiles[5];
};

namespace gt {

enum TileKind { kRelayout = 0, kForward = 1, kStage = 2, kBackward = 3, kDual = 4 };
enum JobKind { kPlain = 0, kLandweber = 1, kTvEven = 2, kTvOdd = 3 };

// The empty asm hands back the (uniform) pointer typed as global; accesses
// through it then compile to global and scalar loads.  A plain cast to global
// and back would be folded away.
template <typename T>
__device__ __forceinline__ T *global_ptr(T *p) {
  __attribute__((address_space(1))) T *g;
  asm("" : "=s"(g) : "0"(p));
  return (T *)g;
}

// A batched kernel's job comes from a table in global memory, so the compiler
// cannot tell where its pointers point and would access them with generic
// (flat) instructions, which also count against the LDS wait counter.  The
// copy marks every pointer as global (all of them are device allocations), as
// the pointers of a by-value kernel argument are.
template <typename real>
__device__ __forceinline__ Job<real> load_job(const Job<real> *__restrict__ jobs, int i) {
  Job<real> J = jobs[i];
  J.ofs = global_ptr(J.ofs);
  J.sdpd = global_ptr(J.sdpd);
  J.geo = global_ptr(J.geo);
  J.sino = global_ptr(J.sino);
  J.imgX = global_ptr(J.imgX);
  J.imgY = global_ptr(J.imgY);
  J.src = global_ptr(J.src);
  J.part = global_ptr(J.part);
  J.img = global_ptr(J.img);
  J.S = global_ptr(J.S);
  J.tab = global_ptr(J.tab);
  J.data = global_ptr(J.data);
  J.u = global_ptr(J.u);
  J.v = global_ptr(J.v);
  return J;
}

// tile (slice, 32x32 block): the forward's image copies of src (the y-fastest
// one, and the row-padded x-fastest one when imgX is not src itself)
template <typename real>
__global__ __launch_bounds__(256) void batch_relayout_kernel(const Job<real> *__restrict__ jobs,
                                                             const int4 *__restrict__ tiles) {
  __shared__ real tile[32][33];
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  const Strides3 ss = {1, J.Nx, (int64_t)J.Nx * J.Ny};
  const Strides3 ys = {J.rowY, 1, J.sliceY};
  relayout_tile(tile, const_cast<real *>(J.imgY), ys, J.src, ss, J.Nx, J.Ny, 0, t.y, t.z);
  if (J.imgX != J.src) {
    const Strides3 xs = {1, J.rowX, J.sliceX};
    relayout_tile(tile, const_cast<real *>(J.imgX), xs, J.src, ss, J.Nx, J.Ny, 0, t.y, t.z);
  }
}

// K = 1: tile (slice, bin block, angle block), 64 bins x 4 angles.  K > 1:
// tile (slice, block of the slice's groups as in the single-problem
// launches, group_lane).
template <typename real, int K, int CH = 1>
__global__ __launch_bounds__(256) void batch_fwd_kernel(const Job<real> *__restrict__ jobs,
                                                        const int4 *__restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  if constexpr (K == 1) {
    const int a = t.z * 4 + threadIdx.y;
    const int s = t.y * 64 + threadIdx.x;
    if (s >= J.Ns || a >= J.Na) return;
    if (J.fan)
      fan_fwd_ray(J, s, a, 0);
    else
      radon_fwd_ray(J, s, a, 0);
  } else {
    __shared__ real slots[Slots<K>::kCount * 256];
    int g;
    const int a = group_lane<K>(t.y, J.Ns, J.Na, g);
    if (a >= J.Na) return;
    real *acc = slots + threadIdx.y * 64 + threadIdx.x;
    if (J.fan)
      fan_fwd_group<real, K, false, CH>(J, g * K, a, 0, acc);
    else
      radon_fwd_group<real, K, false, CH>(J, g * K, a, 0, acc);
  }
}

// tile (slice, block of the slice's groups (group_lane), 0, band): one row
// band of their strip walks; raw sums go to the slice's band partials
template <typename real, int K, int CH, int NB>
__global__ __launch_bounds__(256) void batch_fwd_band_kernel(const Job<real> *__restrict__ jobs,
                                                             const int4 *__restrict__ tiles) {
  __shared__ real slots[Slots<K>::kCount * 256];
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  int g;
  const int a = group_lane<K>(t.y, J.Ns, J.Na, g);
  if (a >= J.Na) return;
  real *acc = slots + threadIdx.y * 64 + threadIdx.x;
  const Band<real> B{t.w, NB, J.part};
  if (J.fan)
    fan_fwd_group<real, K, true, CH>(J, g * K, a, 0, acc, B);
  else
    radon_fwd_group<real, K, true, CH>(J, g * K, a, 0, acc, B);
}

// tile (slice, angle): the bands of one sinogram row added in band order, then
// the forward's scale and epilogue
template <typename real, int NB>
__global__ __launch_bounds__(256) void batch_band_reduce_kernel(const Job<real> *__restrict__ jobs,
                                                                const int4 *__restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  for (int s = threadIdx.x; s < J.Ns; s += 256) band_reduce_entry(J, J.part, NB, s, t.y, 0);
}

// tile (slice, angle): one padded sinogram row
template <typename real>
__global__ __launch_bounds__(256) void batch_stage_kernel(const Job<real> *__restrict__ jobs,
                                                          const int4 *__restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  for (int j = threadIdx.x; j < J.Ns + 4; j += 256)
    stage_entry(const_cast<real *>(J.S), J.sino, J.so, J.ofs, J.sdpd, J.geo, J.fan, J.Ns, J.Na, j,
                t.y, 0);
}

template <typename real, int RY>
__global__ __launch_bounds__(256) void batch_back_kernel(const Job<real> *__restrict__ jobs,
                                                         const int4 *__restrict__ tiles) {
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  const int x = t.y * 16 + (threadIdx.x & 15);
  const int y0 = (t.z * 16 + (threadIdx.x >> 4)) * RY;
  if (x >= J.Nx || y0 >= J.Ny) return;
  if (J.fan)
    back_pixels<real, true, RY>(J, x, y0, 0);
  else
    back_pixels<real, false, RY>(J, x, y0, 0);
}

// tile (slice, tile x, tile y): LDS-staged backprojection of TX x (256/TX)*RY pixels
template <typename real, int RY, int CA, int TX>
__global__ __launch_bounds__(256, kStagedWavesPerSimd) void batch_back_staged_kernel(
    const Job<real> *__restrict__ jobs, const int4 *__restrict__ tiles) {
  __shared__ real rows[CA * kStageWidth];
  __shared__ StageWindow<real> win[CA];
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  if (J.fan)
    back_staged_tile<real, true, RY, CA, TX>(J, t.y, t.z, 0, rows, win);
  else
    back_staged_tile<real, false, RY, CA, TX>(J, t.y, t.z, 0, rows, win);
}

// tile (slice, first pixel): 256 pixels of the dual update
template <typename real>
__global__ __launch_bounds__(256) void batch_dual_kernel(const Job<real> *__restrict__ jobs,
                                                         const int4 *__restrict__ tiles,
                                                         float sigma) {
  const int4 t = tiles[blockIdx.x];
  const Job<real> J = load_job(jobs, t.x);
  const int64_t p = (int64_t)t.y * 256 + threadIdx.x;
  if (p >= (int64_t)J.Nx * J.Ny) return;
  tv_dual_pixel(J, const_cast<real *>(J.v), sigma, (int)(p % J.Nx), (int)(p / J.Nx));
}

template <typename real>
__global__ void batch_table_kernel(const Job<real> *__restrict__ jobs, int n_items) {
  const int i = blockIdx.y;
  if (i >= n_items) return;
  const Job<real> J = load_job(jobs, i);
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < J.Na; a += gridDim.x * blockDim.x)
    back_table_entry(const_cast<AngleRec<real> *>(J.tab), J.ofs, J.geo, J.fan, a);
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
static int check_item(const gt_batch_item &it) {
  GT_REQUIRE(it.geometry == GT_GEOMETRY_RADON || it.geometry == GT_GEOMETRY_FAN,
             "batch item: unknown geometry");
  GT_REQUIRE(it.Nx > 0 && it.Ny > 0 && it.Ns > 0 && it.Na > 0, "batch item: sizes must be positive");
  // sizes index with 32-bit lanes: pixels, and (angle, bin) pairs of a slice
  GT_REQUIRE(it.Na <= 65535 && (int64_t)it.Nx * it.Ny < ((int64_t)1 << 31) &&
                 (int64_t)it.Na * it.Ns < ((int64_t)1 << 31) - 256,
             "batch item: too large");
  GT_REQUIRE(it.ofs && it.geo, "batch item: NULL table");
  GT_REQUIRE(it.geometry == GT_GEOMETRY_RADON || it.sdpd, "batch item: NULL sdpd");
  GT_REQUIRE(it.img && it.sino, "batch item: NULL img or sino");
  return GT_OK;
}

struct TileSet {
  std::vector<int4> t;
  std::vector<double> cost;
  void add(int item, int a, int b, double c, int w = 0) {
    t.push_back(make_int4(item, a, b, w));
    cost.push_back(c);
  }
  // largest first; ties keep slice order
  std::vector<int4> sorted() const {
    std::vector<size_t> idx(t.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<int4> out(t.size());
    for (size_t i = 0; i < idx.size(); ++i) out[i] = t[idx[i]];
    return out;
  }
};

// Rows of the forward's padded image copies end in this many zeros (chunked reads).
constexpr int kBatchRowPad = 4;

template <typename real>
static int plan_create(gt_batch_plan **out, int dtype, int n, const gt_batch_item *items, int K,
                       int RY, int nb, int chunk, int stage, int tx) {
  const bool chunked = K > 1 && chunk > 1;
  nb = K > 1 && nb > 1 ? nb : 1;
  const int pad = chunked ? kBatchRowPad : 0;
  const int ty = 256 / tx;  // lane rows of a staged tile
  // ---- layout of the single allocation ----
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align_up(bytes);
    return at;
  };
  const size_t jobs_at = take(4 * sizeof(Job<real>) * (size_t)n);
  std::vector<size_t> tab_at(n), S_at(n), Y_at(n), X_at(n), part_at(n);
  const size_t copies_from = off;
  for (int i = 0; i < n; ++i) {
    const gt_batch_item &it = items[i];
    tab_at[i] = take(sizeof(AngleRec<real>) * (size_t)it.Na);
    S_at[i] = take(sizeof(real) * (size_t)(it.Ns + 4) * it.Na);
    Y_at[i] = take(sizeof(real) * (size_t)(it.Ny + pad) * it.Nx);
    if (chunked) X_at[i] = take(sizeof(real) * (size_t)(it.Nx + pad) * it.Ny);
  }
  const size_t copies_to = off;
  if (nb > 1)
    for (int i = 0; i < n; ++i) part_at[i] = take(sizeof(real) * (size_t)nb * items[i].Ns * items[i].Na);
  TileSet sets[5];
  for (int i = 0; i < n; ++i) {
    const gt_batch_item &it = items[i];
    // forward tile: ~ rows crossed x strip width per ray group
    const double cf =
        (double)std::max(it.Nx, it.Ny) * (K + 1) * (it.geometry == GT_GEOMETRY_FAN ? 1.5 : 1.0);
    // bands innermost: the 8 bands of a tile are dispatched to the 8 XCDs in turn
    if (K > 1) {  // blocks of 256 lanes over (angle, group) pairs (group_lane)
      const int64_t nblk = group_blocks(it.Ns, it.Na, K);
      for (int64_t b = 0; b < nblk; ++b)
        for (int w = 0; w < nb; ++w) sets[kForward].add(i, (int)b, 0, cf / nb, w);
    } else {  // 64 bins x 4 angles
      const int nsb = (it.Ns + 63) / 64, nab = (it.Na + 3) / 4;
      for (int b = 0; b < nsb; ++b)
        for (int a = 0; a < nab; ++a) sets[kForward].add(i, b, a, cf);
    }
    for (int bx = 0; bx < (it.Nx + 31) / 32; ++bx)
      for (int by = 0; by < (it.Ny + 31) / 32; ++by) sets[kRelayout].add(i, bx * 32, by * 32, 1.0);
    for (int a = 0; a < it.Na; ++a) sets[kStage].add(i, a, 0, (double)it.Ns);
    const double cb = (double)it.Na * RY * (it.geometry == GT_GEOMETRY_FAN ? 1.5 : 1.0);
    const int tw = stage ? tx : 16, th = (stage ? ty : 16) * RY;  // back tile in pixels
    for (int bx = 0; bx < (it.Nx + tw - 1) / tw; ++bx)
      for (int by = 0; by < (it.Ny + th - 1) / th; ++by) sets[kBackward].add(i, bx, by, cb);
    const int64_t npix = (int64_t)it.Nx * it.Ny;
    for (int64_t p = 0; p < (npix + 255) / 256; ++p) sets[kDual].add(i, (int)p, 0, 1.0);
  }
  size_t tiles_at[5];
  std::vector<int4> sorted[5];
  for (int k = 0; k < 5; ++k) {
    sorted[k] = sets[k].sorted();
    tiles_at[k] = take(sizeof(int4) * std::max<size_t>(sorted[k].size(), 1));
  }
  void *mem = nullptr;
  if (hipMalloc(&mem, off) != hipSuccess) {
    (void)hipGetLastError();
    return fail(GT_EHIP, "batch plan: device allocation failed");
  }
  char *base = static_cast<char *>(mem);
  // the image copies start as zeros: their row padding is never written again
  if (hipMemset(base + copies_from, 0, copies_to - copies_from) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(mem);
    return fail(GT_EHIP, "batch plan: clearing the image copies failed");
  }

  // ---- job records ----
  std::vector<Job<real>> jobs(4 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const gt_batch_item &it = items[i];
    Job<real> J{};
    J.fan = it.geometry == GT_GEOMETRY_FAN;
    J.Nx = it.Nx;
    J.Ny = it.Ny;
    J.Ns = it.Ns;
    J.Na = it.Na;
    J.Nz = 1;
    J.ofs = static_cast<const real *>(it.ofs);
    J.sdpd = static_cast<const real *>(it.sdpd);
    J.geo = static_cast<const real *>(it.geo);
    J.sino = static_cast<real *>(it.sino);
    J.so = order_strides(GT_ORDER_F, it.Ns, it.Na, 1);
    J.imgY = reinterpret_cast<const real *>(base + Y_at[i]);
    const real *xcopy = chunked ? reinterpret_cast<const real *>(base + X_at[i]) : nullptr;
    J.rowX = it.Nx + pad;
    J.rowY = it.Ny + pad;
    J.sliceX = J.rowX * it.Ny;
    J.sliceY = J.rowY * it.Nx;
    J.part = nb > 1 ? reinterpret_cast<real *>(base + part_at[i]) : nullptr;
    J.io = order_strides(GT_ORDER_F, it.Nx, it.Ny, 1);
    J.S = reinterpret_cast<const real *>(base + S_at[i]);
    J.tab = reinterpret_cast<const AngleRec<real> *>(base + tab_at[i]);
    J.data = static_cast<const real *>(it.data);
    J.v = static_cast<const real *>(it.field);
    J.sigma_p = (float)it.p[1];
    J.c = (float)it.p[2];
    J.tau = (float)it.p[3];
    J.norming = (float)it.p[4];
    real *img = static_cast<real *>(it.img);
    real *img2 = static_cast<real *>(it.img2);

    Job<real> P = J;  // plain: sino = R img, img = B sino
    P.mode = 0;
    P.src = img;
    P.imgX = chunked ? xcopy : img;
    P.img = img;
    Job<real> L = P;  // Landweber: sino = R img - data; img += p0 * B sino
    L.mode = 1;
    L.scale = (real)it.p[0];
    Job<real> T0 = J;  // TV, even step: u_bar in img2, u in img; u' -> img2, u_bar' -> img
    T0.mode = 2;
    T0.src = img2;
    T0.imgX = chunked ? xcopy : img2;
    T0.img = img2;
    T0.u = img;
    Job<real> T1 = T0;  // TV, odd step: roles swapped
    T1.src = img;
    T1.imgX = chunked ? xcopy : img;
    T1.img = img;
    T1.u = img2;
    jobs[kPlain * (size_t)n + i] = P;
    jobs[kLandweber * (size_t)n + i] = L;
    jobs[kTvEven * (size_t)n + i] = T0;
    jobs[kTvOdd * (size_t)n + i] = T1;
  }
  hipError_t e = hipMemcpy(base + jobs_at, jobs.data(), sizeof(Job<real>) * jobs.size(),
                           hipMemcpyHostToDevice);
  for (int k = 0; k < 5 && e == hipSuccess; ++k)
    if (!sorted[k].empty())
      e = hipMemcpy(base + tiles_at[k], sorted[k].data(), sizeof(int4) * sorted[k].size(),
                    hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const Job<real> *dj = reinterpret_cast<const Job<real> *>(base + jobs_at);
    int maxNa = 1;
    for (int i = 0; i < n; ++i) maxNa = std::max(maxNa, items[i].Na);
    hipLaunchKernelGGL(batch_table_kernel<real>, dim3((maxNa + 255) / 256, n), dim3(256), 0, 0, dj,
                       n);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) {
    (void)hipFree(mem);
    return fail(GT_EHIP, std::string("batch plan: ") + hipGetErrorString(e));
  }
  gt_batch_plan *p = new gt_batch_plan{};
  p->dtype = dtype;
  p->n_items = n;
  p->fwd_group = K;
  p->back_rows = RY;
  p->fwd_bands = nb;
  p->fwd_chunk = chunked ? chunk : 1;
  p->back_stage = stage;
  p->back_tile = tx;
  p->has_data = true;
  p->has_tv = true;
  for (int i = 0; i < n; ++i) {
    p->has_data = p->has_data && items[i].data != nullptr;
    p->has_tv = p->has_tv && items[i].data && items[i].img2 && items[i].field;
  }
  p->mem = mem;
  p->bytes = off;
  for (int k = 0; k < 4; ++k) p->jobs[k] = base + jobs_at + sizeof(Job<real>) * (size_t)n * k;
  for (int k = 0; k < 5; ++k) {
    p->tiles[k] = reinterpret_cast<const int4 *>(base + tiles_at[k]);
    p->n_tiles[k] = (int)sorted[k].size();
  }
  *out = p;
  return GT_OK;
}

template <typename real, int K, int CH>
static void launch_fwd_banded(const gt_batch_plan *p, hipStream_t st, const Job<real> *jobs,
                              dim3 grid, dim3 block) {
  switch (p->fwd_bands) {
    case 8:
      hipLaunchKernelGGL((batch_fwd_band_kernel<real, K, CH, 8>), grid, block, 0, st, jobs,
                         p->tiles[kForward]);
      hipLaunchKernelGGL((batch_band_reduce_kernel<real, 8>), dim3(p->n_tiles[kStage]), dim3(256), 0,
                         st, jobs, p->tiles[kStage]);
      break;
    case 4:
      hipLaunchKernelGGL((batch_fwd_band_kernel<real, K, CH, 4>), grid, block, 0, st, jobs,
                         p->tiles[kForward]);
      hipLaunchKernelGGL((batch_band_reduce_kernel<real, 4>), dim3(p->n_tiles[kStage]), dim3(256), 0,
                         st, jobs, p->tiles[kStage]);
      break;
    default:
      hipLaunchKernelGGL((batch_fwd_band_kernel<real, K, CH, 2>), grid, block, 0, st, jobs,
                         p->tiles[kForward]);
      hipLaunchKernelGGL((batch_band_reduce_kernel<real, 2>), dim3(p->n_tiles[kStage]), dim3(256), 0,
                         st, jobs, p->tiles[kStage]);
  }
}

template <typename real, int K>
static void launch_fwd_k(const gt_batch_plan *p, hipStream_t st, const Job<real> *jobs, dim3 grid,
                         dim3 block) {
  constexpr int CH = 16 / sizeof(real);
  if (p->fwd_bands > 1) {
    if (p->fwd_chunk > 1)
      launch_fwd_banded<real, K, CH>(p, st, jobs, grid, block);
    else
      launch_fwd_banded<real, K, 1>(p, st, jobs, grid, block);
    return;
  }
  if (p->fwd_chunk > 1)
    hipLaunchKernelGGL((batch_fwd_kernel<real, K, CH>), grid, block, 0, st, jobs, p->tiles[kForward]);
  else
    hipLaunchKernelGGL((batch_fwd_kernel<real, K, 1>), grid, block, 0, st, jobs, p->tiles[kForward]);
}

template <typename real, int RY, int CA>
static void launch_back_staged(const gt_batch_plan *p, hipStream_t st, const Job<real> *jobs) {
  const dim3 grid(p->n_tiles[kBackward]);
  if (p->back_tile == 32)
    hipLaunchKernelGGL((batch_back_staged_kernel<real, RY, CA, 32>), grid, dim3(256), 0, st, jobs,
                       p->tiles[kBackward]);
  else
    hipLaunchKernelGGL((batch_back_staged_kernel<real, RY, CA, 16>), grid, dim3(256), 0, st, jobs,
                       p->tiles[kBackward]);
}

template <typename real, int RY>
static void launch_back_ry(const gt_batch_plan *p, hipStream_t st, const Job<real> *jobs) {
  if (p->back_stage == 16)
    launch_back_staged<real, RY, 16>(p, st, jobs);
  else if (p->back_stage == 8)
    launch_back_staged<real, RY, 8>(p, st, jobs);
  else
    hipLaunchKernelGGL((batch_back_kernel<real, RY>), dim3(p->n_tiles[kBackward]), dim3(256), 0, st,
                       jobs, p->tiles[kBackward]);
}

template <typename real>
static void launch_fwd(const gt_batch_plan *p, hipStream_t st, const Job<real> *jobs) {
  if (p->n_tiles[kRelayout])
    hipLaunchKernelGGL(batch_relayout_kernel<real>, dim3(p->n_tiles[kRelayout]), dim3(256), 0, st,
                       jobs, p->tiles[kRelayout]);
  const dim3 grid(p->n_tiles[kForward]), block(64, 4);
  switch (p->fwd_group) {
    case 16:
      launch_fwd_k<real, 16>(p, st, jobs, grid, block);
      break;
    case 8:
      launch_fwd_k<real, 8>(p, st, jobs, grid, block);
      break;
    case 4:
      launch_fwd_k<real, 4>(p, st, jobs, grid, block);
      break;
    case 2:
      launch_fwd_k<real, 2>(p, st, jobs, grid, block);
      break;
    default:
      hipLaunchKernelGGL((batch_fwd_kernel<real, 1>), grid, block, 0, st, jobs, p->tiles[kForward]);
  }
}

template <typename real>
static void launch_back(const gt_batch_plan *p, hipStream_t st, const Job<real> *jobs) {
  hipLaunchKernelGGL(batch_stage_kernel<real>, dim3(p->n_tiles[kStage]), dim3(256), 0, st, jobs,
                     p->tiles[kStage]);
  switch (p->back_rows) {
    case 8:
      launch_back_ry<real, 8>(p, st, jobs);
      break;
    case 4:
      launch_back_ry<real, 4>(p, st, jobs);
      break;
    case 2:
      launch_back_ry<real, 2>(p, st, jobs);
      break;
    default:
      launch_back_ry<real, 1>(p, st, jobs);
  }
}

template <typename real>
static int launch(const gt_batch_plan *p, hipStream_t st, int op, double sigma) {
  const Job<real> *J[4];
  for (int k = 0; k < 4; ++k) J[k] = static_cast<const Job<real> *>(p->jobs[k]);
  switch (op) {
    case GT_BATCH_FORWARD:
      launch_fwd<real>(p, st, J[kPlain]);
      break;
    case GT_BATCH_BACKWARD:
      launch_back<real>(p, st, J[kPlain]);
      break;
    case GT_BATCH_LANDWEBER:
      launch_fwd<real>(p, st, J[kLandweber]);
      launch_back<real>(p, st, J[kLandweber]);
      break;
    case GT_BATCH_TV_EVEN:
    case GT_BATCH_TV_ODD: {
      const Job<real> *T = J[op == GT_BATCH_TV_EVEN ? kTvEven : kTvOdd];
      hipLaunchKernelGGL(batch_dual_kernel<real>, dim3(p->n_tiles[kDual]), dim3(256), 0, st, T,
                         p->tiles[kDual], (float)sigma);
      launch_fwd<real>(p, st, T);
      launch_back<real>(p, st, T);
      break;
    }
    default:
      return fail(GT_EINVAL, "gt_batch_launch: unknown operation");
  }
  return check_launch("batched projection");
}

}  // namespace gt

using namespace gt;

extern "C" int gt_batch_plan_create(gt_batch_plan **plan, int dtype, int n_items,
                                    const gt_batch_item *items) {
  GT_REQUIRE(plan != nullptr, "NULL plan pointer");
  *plan = nullptr;
  GT_REQUIRE(dtype == GT_FLOAT32 || dtype == GT_FLOAT64, "dtype must be GT_FLOAT32 or GT_FLOAT64");
  GT_REQUIRE(n_items > 0 && items != nullptr, "a batch needs at least one item");
  GT_REQUIRE(n_items <= 65535, "a batch holds at most 65535 items");
  for (int i = 0; i < n_items; ++i) {
    const int rc = check_item(items[i]);
    if (rc) return rc;
  }
  // automatic launch shapes: the batch autotuner's choice on 64 uneven slices
  // (128^2..1024^2, 60..720 angles); batches have no backprojection angle bands
  Tuning t = tuning();
  t.fwd_bands = std::min(t.fwd_bands, 8);  // batch forwards take up to 8 row bands
  t.fwd_group = std::min(t.fwd_group, 16);  // and up to 16 bins per lane
  if (t.automatic) {
    t.fwd_group = 16;
    t.fwd_bands = 8;
    t.fwd_chunk = 4;
    t.back_rows = 8;
    t.back_stage = 16;
    t.back_tile = 32;
  }
  return dtype == GT_FLOAT64
             ? plan_create<double>(plan, dtype, n_items, items, t.fwd_group, t.back_rows, t.fwd_bands,
                                   t.fwd_chunk, t.back_stage, t.back_tile)
             : plan_create<float>(plan, dtype, n_items, items, t.fwd_group, t.back_rows, t.fwd_bands,
                                  t.fwd_chunk, t.back_stage, t.back_tile);
}

extern "C" int gt_batch_plan_destroy(gt_batch_plan *plan) {
  if (!plan) return GT_OK;
  const hipError_t e = hipFree(plan->mem);
  delete plan;
  if (e != hipSuccess) return fail(GT_EHIP, "gt_batch_plan_destroy: hipFree failed");
  return GT_OK;
}

extern "C" size_t gt_batch_plan_bytes(const gt_batch_plan *plan) { return plan ? plan->bytes : 0; }

extern "C" int gt_batch_launch(gt_batch_plan *plan, gt_stream_t stream, int op, double sigma) {
  GT_REQUIRE(plan != nullptr, "NULL plan");
  GT_REQUIRE(op != GT_BATCH_LANDWEBER || plan->has_data,
             "gt_batch_launch: Landweber needs data on every item");
  GT_REQUIRE(!(op == GT_BATCH_TV_EVEN || op == GT_BATCH_TV_ODD) || plan->has_tv,
             "gt_batch_launch: TV needs data, img2 and field on every item");
  hipStream_t st = static_cast<hipStream_t>(stream);
  return plan->dtype == GT_FLOAT64 ? launch<double>(plan, st, op, sigma)
                                   : launch<float>(plan, st, op, sigma);
}
