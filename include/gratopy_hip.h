/*
 * gratopy_hip.h -- C ABI of the MI355X (gfx950) projection library.
 *
 * Every compute entry point is asynchronous on the HIP stream it is given
 * (pass NULL for the null stream), takes raw device pointers and plain sizes,
 * allocates nothing (only gt_batch_plan_create allocates, and
 * gt_batch_plan_destroy frees), and returns a status code (GT_OK on success).  The text
 * of the last failure of the calling thread is available from
 * gt_last_error().  Arrays are dense and either Fortran ordered (GT_ORDER_F:
 * x (or s) fastest, z slowest) or C ordered (GT_ORDER_C: z fastest), exactly
 * as in the reference's pos_img_f / pos_img_c / pos_sino_f / pos_sino_c
 * macros (gratopy/radon.cl:28-31).
 *
 * Geometry tables (ofs, sdpd, geo, angle_weights) are the reference's device
 * buffers, byte for byte: the layouts written by radon_struct
 * (gratopy/gratopy.py:553-569) and fanbeam_struct (gratopy/gratopy.py:950-988),
 * converted to the element type of the call.
 *
 * Each projection entry point needs a device workspace of at least
 * gt_projection_workspace_bytes(...) bytes (a transposed image copy for the
 * forward direction; a zero-padded, pre-weighted sinogram copy and a per-angle
 * table for the backward direction).
 */
#ifndef GRATOPY_HIP_H
#define GRATOPY_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *gt_stream_t; /* a hipStream_t */

enum { GT_FLOAT32 = 0, GT_FLOAT64 = 1 };
enum { GT_ORDER_F = 0, GT_ORDER_C = 1 };
enum { GT_GEOMETRY_RADON = 0, GT_GEOMETRY_FAN = 1 };
enum { GT_OK = 0, GT_EINVAL = 1, GT_ENOSPC = 2, GT_EHIP = 3 };

/* Library version (major*10000 + minor*100 + patch). */
int gt_version(void);

/* Launch-shape knobs (process-wide; set by the autotuner, gratopy_amd/autotune.py).
 * No reference counterpart: the reference leaves local sizes to the OpenCL
 * driver (gratopy/gratopy.py:304-313).  Results do not depend on them beyond
 * rounding.
 *   GT_TUNE_FWD_GROUP: detector bins per lane in the forward projection (1, 2, 4, 8, 16,
 *                      32; 32 always reads 16 bytes per load)
 *   GT_TUNE_BACK_ROWS: pixels per lane in the backprojection (1, 2, 4, 8)
 *   GT_TUNE_FWD_BANDS: row bands of the forward strip walk (1, 2, 4, 8, 16, 32; applies
 *                      when GT_TUNE_FWD_GROUP > 1).  With B > 1 bands, consecutive thread
 *                      blocks take different bands (with 8, each XCD's L2 holds one band
 *                      of the image; more bands give a small image more blocks; batch
 *                      plans take at most 8), and the workspace grows by B * Ns * Na * Nz
 *                      elements for the per-band partial sums.
 *   GT_TUNE_BACK_BANDS: angle bands of the backprojection (1, 2, 4, 8), on the same
 *                      pattern (each XCD's L2 holds the padded sinogram rows of one band;
 *                      the workspace grows by B * Nx * Ny * Nz elements).
 *   GT_TUNE_BACK_STAGE: angles per chunk of the LDS-staged backprojection (0 = off, 8,
 *                      16): each block copies the sinogram window its pixel tile
 *                      reaches, per angle, into LDS and reads the interpolation taps from
 *                      there; with GT_TUNE_BACK_BANDS > 1 each block sums one band of
 *                      the angles.
 *   GT_TUNE_BACK_TILE: lanes along x of a staged backprojection tile (16 or 32; the
 *                      tile is that many pixels wide and 256 / width * rows-per-lane
 *                      pixels tall).
 *   GT_TUNE_FWD_CHUNK: 4 = the forward bin groups read their strip rows 16 bytes per
 *                      load (4 float32 or 2 float64 pixels) from row-padded copies of
 *                      the image (the workspace then holds two padded copies); 1 = one
 *                      pixel per load.
 *   GT_TUNE_FWD_COMPACT: 1 (the default) = a banded forward launches, per row band and
 *                      angle, only the bin groups whose strips can meet the band (a
 *                      prologue kernel lists them in the workspace); 0 = every group in
 *                      every band.  The results are the same bit for bit.  Batch plans do
 *                      not compact.
 *   GT_TUNE_AUTO:      1 (the default) = every call (and batch plan) picks its launch
 *                      shape from its geometry and sizes and the knobs above are
 *                      ignored; 0 = the knobs above apply.  Workspace sizes follow the
 *                      same choice. */
enum {
  GT_TUNE_FWD_GROUP = 0,
  GT_TUNE_BACK_ROWS = 1,
  GT_TUNE_FWD_BANDS = 2,
  GT_TUNE_BACK_BANDS = 3,
  GT_TUNE_BACK_STAGE = 4,
  GT_TUNE_BACK_TILE = 5,
  GT_TUNE_FWD_CHUNK = 6,
  GT_TUNE_AUTO = 7,
  GT_TUNE_FWD_COMPACT = 8
};
int gt_tune_set(int key, int value);
int gt_tune_get(int key);

/* Copies the calling thread's last error message into buf (NUL terminated,
 * truncated to n bytes); returns the full message length. */
size_t gt_last_error(char *buf, size_t n);

/* Workspace bytes needed by gt_{radon,fan}_{fwd,back} for these sizes.
 * geometry is that of the call (GT_GEOMETRY_RADON or GT_GEOMETRY_FAN: the
 * automatic launch shape, and with it the workspace, depends on it);
 * adjoint = 0 for the forward projection, 1 for the backprojection;
 * in_order is the order of the input array of that call.  A call given less
 * workspace than it needs returns GT_ENOSPC. */
size_t gt_projection_workspace_bytes(int geometry, int adjoint, int dtype, int Nx, int Ny,
                                     int Ns, int Na, int Nz, int in_order);

/* The launch shape a projection call with these sizes takes under the current
 * knobs (the automatic choice when GT_TUNE_AUTO = 1): forward: shape[0..3] =
 * detector bins per lane, row bands, pixels per strip load, compacted (0/1);
 * backprojection: shape[0..3] = pixels per lane, angle bands, angles per LDS
 * staging chunk (0 = unstaged), staged tile width.  No reference counterpart
 * (it describes this library's launches).  GT_OK, or GT_EINVAL for invalid sizes. */
int gt_projection_shape(int geometry, int adjoint, int Nx, int Ny, int Ns, int Na, int Nz,
                        int *shape);

/* Parallel-beam forward projection sino = R img.
 * Replaces kernel radon_<t>_<sino_order><img_order> (gratopy/radon.cl:66-161),
 * launched by gratopy.radon (gratopy/gratopy.py:258-316). */
int gt_radon_fwd(gt_stream_t stream, int dtype, void *sino, const void *img, const void *ofs,
                 const void *geo, int Nx, int Ny, int Ns, int Na, int Nz, int sino_order,
                 int img_order, void *workspace, size_t workspace_bytes);

/* Parallel-beam backprojection img = R^# sino.
 * Replaces kernel radon_ad_<t>_<img_order><sino_order> (gratopy/radon.cl:183-241),
 * launched by gratopy.radon_ad (gratopy/gratopy.py:319-377). */
int gt_radon_back(gt_stream_t stream, int dtype, void *img, const void *sino, const void *ofs,
                  const void *geo, int Nx, int Ny, int Ns, int Na, int Nz, int img_order,
                  int sino_order, void *workspace, size_t workspace_bytes);

/* Fanbeam forward projection.
 * Replaces kernel fanbeam_<t>_<sino_order><img_order> (gratopy/fanbeam.cl:72-256),
 * launched by gratopy.fanbeam (gratopy/gratopy.py:581-640). */
int gt_fan_fwd(gt_stream_t stream, int dtype, void *sino, const void *img, const void *ofs,
               const void *sdpd, const void *geo, int Nx, int Ny, int Ns, int Na, int Nz,
               int sino_order, int img_order, void *workspace, size_t workspace_bytes);

/* Fanbeam backprojection.
 * Replaces kernel fanbeam_ad_<t>_<img_order><sino_order> (gratopy/fanbeam.cl:290-391),
 * launched by gratopy.fanbeam_ad (gratopy/gratopy.py:643-703). */
int gt_fan_back(gt_stream_t stream, int dtype, void *img, const void *sino, const void *ofs,
                const void *sdpd, const void *geo, int Nx, int Ny, int Ns, int Na, int Nz,
                int img_order, int sino_order, void *workspace, size_t workspace_bytes);

/* Projection without allocation: gt_{radon,fan}_{fwd,back} (geometry GT_GEOMETRY_RADON
 * or GT_GEOMETRY_FAN; adjoint 0 = forward, out = sino, in = img; adjoint 1 =
 * backprojection, out = img, in = sino; sdpd may be NULL for Radon), with the
 * workspace taken from a device array that the library holds on every GPU it runs
 * on, GT_NOALLOC_SCRATCH_BYTES (128 MiB) each, until the process ends.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Allocates nothing.  The stream must belong to the current device.  A call whose
 * workspace (gt_projection_workspace_bytes) exceeds the array returns GT_ENOSPC.
 * Every call on one GPU uses the same array: calls that overlap on one GPU (on
 * different streams, or from different threads) race.  Results are those of the
 * entry points above. */
#define GT_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int gt_projection_noalloc(gt_stream_t stream, int geometry, int adjoint, int dtype, void *out,
                          const void *in, const void *ofs, const void *sdpd, const void *geo,
                          int Nx, int Ny, int Ns, int Na, int Nz, int out_order, int in_order);

/* Sinogram of a unit impulse at pixel (x, y): one column of the system matrix.
 * sdpd is ignored (may be NULL) for the parallel geometry.
 * Replaces single_line_radon_* (gratopy/radon.cl:263-341) and
 * single_line_fanbeam_* (gratopy/fanbeam.cl:422-568), used by
 * ProjectionSettings.create_sparse_matrix (gratopy/gratopy.py:2014-2155). */
int gt_single_line(gt_stream_t stream, int geometry, int dtype, void *sino, int x, int y,
                   const void *ofs, const void *sdpd, const void *geo, int Nx, int Ny, int Ns,
                   int Na, int sino_order);

/* out[s,a,z] = in[s,a,z] * w[a]  (divide != 0: / w[a]); out may alias in.
 * Replaces multiply_* / divide_* (gratopy/utilities.cl:38-67) as launched by
 * gratopy.weight_sinogram (gratopy/gratopy.py:2188-2266). */
int gt_weight_sino(gt_stream_t stream, int dtype, void *out, const void *in,
                   const void *angle_weights, int Ns, int Na, int Nz, int order, int divide);

/* rhs += a*x over n elements.  Replaces equ_mul_add_* (gratopy/utilities.cl:69-73),
 * gratopy.equ_mul_add (gratopy/gratopy.py:2269-2295).  a is rounded to the
 * element type before use, as the reference passes it as a `real` argument. */
int gt_axpy(gt_stream_t stream, int dtype, void *rhs, double a, const void *x, size_t n);

/* rhs = a*x + y.  Replaces mul_add_add_* (gratopy/utilities.cl:75-80),
 * gratopy.mul_add_add (gratopy/gratopy.py:2298-2326). */
int gt_axpby(gt_stream_t stream, int dtype, void *rhs, double a, const void *x, const void *y,
             size_t n);

/* As gt_axpy / gt_axpby, with the scalar read on the device: a = mult * (*a_dev)
 * (a_dev points to one double).  No reference counterpart: lets the
 * conjugate-gradient loop (gratopy/gratopy.py:2519-2548) keep alpha and beta on
 * the device so that its iterations can be captured in a HIP graph. */
int gt_axpy_dev(gt_stream_t stream, int dtype, void *rhs, const void *a_dev, double mult,
                const void *x, size_t n);
int gt_axpby_dev(gt_stream_t stream, int dtype, void *rhs, const void *a_dev, double mult,
                 const void *x, const void *y, size_t n);

/* x /= (*s_dev) (divide != 0) or x *= (*s_dev), s_dev a device double.
 * Replaces the host-synchronised `img /= normsqr` of gratopy.normest
 * (gratopy/gratopy.py:2370-2376). */
int gt_scale_dev(gt_stream_t stream, int dtype, void *x, const void *s_dev, int divide, size_t n);

/* x *= a over n elements.  Replaces the pyopencl `output *= scalar` of
 * _OpenCLOperator.apply_to (gratopy/operator/opencl.py:436-437). */
int gt_scale(gt_stream_t stream, int dtype, void *x, double a, size_t n);

/* x[:] = value (a fill kernel; never a memset node). */
int gt_fill(gt_stream_t stream, int dtype, void *x, double value, size_t n);

/* out = a - b elementwise (out may alias a).  Replaces the pyopencl
 * elementwise `sinonew -= sino` of gratopy.landweber (gratopy/gratopy.py:2436). */
int gt_sub(gt_stream_t stream, int dtype, void *out, const void *a, const void *b, size_t n);

/* Workspace bytes for gt_dot / gt_sum over n elements. */
size_t gt_reduce_workspace_bytes(size_t n);

/* *out_dev (a device double) = sum_i x[i]*y[i], accumulated in double with a
 * fixed reduction tree (deterministic).  Replaces pyopencl.array.vdot / dot as
 * used by gratopy.conjugate_gradients (gratopy/gratopy.py:2527-2543). */
int gt_dot(gt_stream_t stream, int dtype, void *out_dev, const void *x, const void *y, size_t n,
           void *workspace, size_t workspace_bytes);

/* *out_dev = sum_i x[i].  Replaces pyopencl.array.sum in gratopy.normest
 * (gratopy/gratopy.py:2372). */
int gt_sum(gt_stream_t stream, int dtype, void *out_dev, const void *x, size_t n, void *workspace,
           size_t workspace_bytes);

/* Conjugate-gradient scalar update on device doubles d[0..7]:
 * mode 0: d[4] = scale * d[0] / d[1]   (alpha = dx^2/ds <sold,sold>/<Wq,q>)
 * mode 1: d[5] = d[2] / d[0]           (beta = <snew,snew>/<sold,sold>), then d[0] = d[2]
 * The call performs mode 0 when scale > 0 and mode 1 when scale <= 0.
 * Mirrors the host arithmetic of gratopy.conjugate_gradients (gratopy/gratopy.py:2523-2538). */
int gt_cg_scalars(gt_stream_t stream, void *d, double scale);

/* TV primal-dual pointwise updates, one per reference kernel
 * (gratopy/total_variation.cl).  v is the dual variable, 4 elements per pixel
 * (a padded 3-vector), indexed by the pixel's position in the image order.
 * sigma, tau, ... are float as in the reference (gratopy/gratopy.py:2647-2685). */
int gt_tv_update_v(gt_stream_t stream, int dtype, void *v, const void *u, float sigma,
                   float z_distance, int Nx, int Ny, int Nz, int order); /* total_variation.cl:54-93 */
int gt_tv_update_normv(gt_stream_t stream, int dtype, void *v, void *normv, int Nx, int Ny,
                       int Nz, int order); /* total_variation.cl:150-167 */
int gt_tv_update_lambda(gt_stream_t stream, int dtype, void *lambda, const void *ku,
                        const void *f, float sigma, float sigmap1inv, int Ns, int Na, int Nz,
                        int order); /* total_variation.cl:95-106 */
int gt_tv_update_u(gt_stream_t stream, int dtype, void *u, const void *u_old, const void *v,
                   const void *kstarlambda, float tau, float norming, float z_distance, int Nx,
                   int Ny, int Nz, int order); /* total_variation.cl:108-148 */
int gt_tv_update_extra(gt_stream_t stream, int dtype, void *u_, void *u,
                       size_t n); /* u = 2 u_ - u; the ElementwiseKernel of gratopy.py:2708-2717 */

/* Fused TV updates (same arithmetic, fewer passes over HBM):
 * gt_tv_dual   = update_v followed by update_NormV_unchor;
 * gt_tv_primal = update_u (into u_new) followed by the extragradient step
 *                u_old <- 2 u_new - u_old.  */
int gt_tv_dual(gt_stream_t stream, int dtype, void *v, const void *u, float sigma,
               float z_distance, int Nx, int Ny, int Nz, int order);
int gt_tv_primal(gt_stream_t stream, int dtype, void *u_new, void *u_old, const void *v,
                 const void *kstarlambda, float tau, float norming, float z_distance, int Nx,
                 int Ny, int Nz, int order);

/* Vector-Jacobian product of one TV primal-dual step (the step of
 * gt_tv_dual / gt_tv_update_lambda / gt_tv_primal with K the forward and B the
 * backprojection).  No reference counterpart: used by the differentiable
 * unrolled TV reconstruction (gratopy_amd/tv_unrolled.py).  Cotangents carry a
 * "_next" suffix for the step's outputs.
 *   gt_tv_adj_primal: wb = [u_next > 0](ub_next + 2 ubarb_next); ub = wb - ubarb_next
 *   gt_tv_adj_dual:   ab = J(a)(vb_next - tau grad wb), a = v + sigma grad ubar,
 *                     J the Jacobian of the pointwise unit-ball projection
 *   gt_tv_adj_div:    ubarb = coef*bterm - sigma div(ab)
 *   gt_tv_adj_lambda: bb = c(lamb_next - coef_r w_a rwb); lamb = bb; fb -= sigma_p bb;
 *                     winv = bb / w_a  */
int gt_tv_adj_primal(gt_stream_t stream, int dtype, void *wb, void *ub, const void *ub_next,
                     const void *ubarb_next, const void *u_next, size_t n);
int gt_tv_adj_dual(gt_stream_t stream, int dtype, void *ab, const void *vb_next, const void *wb,
                   const void *v, const void *ubar, float sigma, float tau, float z_distance,
                   int Nx, int Ny, int Nz, int order);
int gt_tv_adj_div(gt_stream_t stream, int dtype, void *ubarb, const void *bterm, double coef,
                  const void *ab, float sigma, float z_distance, int Nx, int Ny, int Nz,
                  int order);
int gt_tv_adj_lambda(gt_stream_t stream, int dtype, void *lamb, void *fb, void *winv,
                     const void *lamb_next, const void *rwb, const void *angle_weights,
                     double coef_r, float c, float sigma_p, int Ns, int Na, int Nz, int order);

/* ---- batched slices of uneven geometry -------------------------------------
 * No single reference entry point: replaces a host loop that calls
 * forwardprojection / backprojection (gratopy/gratopy.py:126-255), landweber
 * (gratopy/gratopy.py:2380-2446) or total_variation
 * (gratopy/gratopy.py:2554-2762) once per slice, each slice with its own
 * ProjectionSettings.  One item is one 2-D slice; img, img2 (Nx x Ny) and
 * sino, data (Ns x Na) are Fortran ordered; field holds 4 entries per pixel
 * (the TV dual variable, laid out as total_variation's V buffer).
 *
 * Operations (one gt_batch_launch = one step for every item):
 *   GT_BATCH_FORWARD   sino = R img
 *   GT_BATCH_BACKWARD  img = B sino
 *   GT_BATCH_LANDWEBER sino = R img - data;  img += p[0] * B sino
 *                      (p[0] = -w / ||R||^2)
 *   GT_BATCH_TV_EVEN / GT_BATCH_TV_ODD: one Chambolle-Pock iteration of
 *     total_variation with slice_thickness 0; p = {unused, sigma/||R||,
 *     mu'/(sigma+mu'), tau, 1/||R||} and `sigma` the dual step.  Even steps
 *     read u_bar from img2 and u from img and leave the new u in img2 and
 *     u_bar in img; odd steps swap the roles.  Start from zero img, img2,
 *     field and sino (lambda); after an even number of steps u is in img.
 * The plan keeps the items' pointers; results per item are bitwise those of
 * the single-problem entry points and elementwise TV kernels under the same
 * tuning knobs.  The forward bin group, row bands and chunked strip reads,
 * the backprojection rows per lane and LDS staging (chunk, tile width) are
 * taken from the tuning knobs (gt_tune_set) when the plan is created, with at
 * most 16 bins per lane and 8 row bands; angle
 * bands of the backprojection do not apply.
 */
typedef struct gt_batch_item {
  int geometry; /* GT_GEOMETRY_RADON or GT_GEOMETRY_FAN */
  int Nx, Ny, Ns, Na;
  const void *ofs, *sdpd, *geo; /* the item's tables (sdpd NULL for Radon) */
  void *img, *sino;
  const void *data; /* measured sinogram (Landweber, TV) */
  void *img2;       /* second image (TV) */
  void *field;      /* TV dual variable, 4*Nx*Ny */
  double p[6];      /* per-item scalars, see above */
} gt_batch_item;

typedef struct gt_batch_plan gt_batch_plan;

enum {
  GT_BATCH_FORWARD = 0,
  GT_BATCH_BACKWARD = 1,
  GT_BATCH_LANDWEBER = 2,
  GT_BATCH_TV_EVEN = 3,
  GT_BATCH_TV_ODD = 4
};

/* Allocates one device block (job records, tile lists, per-item tables and
 * scratch) and builds the per-angle tables; synchronizes the device. */
int gt_batch_plan_create(gt_batch_plan **plan, int dtype, int n_items, const gt_batch_item *items);
int gt_batch_plan_destroy(gt_batch_plan *plan);
size_t gt_batch_plan_bytes(const gt_batch_plan *plan);
/* Asynchronous on `stream`; capturable in a HIP graph (kernel launches only). */
int gt_batch_launch(gt_batch_plan *plan, gt_stream_t stream, int op, double sigma);

/* ---- operator kernels compiled at run time ---------------------------------
 * Replace the OpenCL program build, kernel lookup and launch of the operator
 * interface (gratopy/operator/opencl.py:262-305 _get_program /
 * _get_projection_kernel, 444-454 _invoke_kernel).  A plugin source is HIP C++
 * whose kernels have C linkage and the names
 * {base}[_ad]_{float|double}_{out_order}{in_order}; their arguments are
 * (out*, in*, extra..., gt_range) where gt_range (gratopy_amd/csrc/gratopy_plugin.h)
 * is the output shape, the reference's global size.
 *
 * gt_plugin_compile: hiprtc compile of `source` for `arch` (e.g. "gfx950"),
 * with -O3 and packed FP32 instructions off ahead of `options`; needs no GPU.
 * The code object goes to `image` if `capacity` suffices; `*image_size` always
 * receives its size.  Compile errors carry the compiler log in gt_last_error. */
int gt_plugin_compile(const char *source, const char *name, const char *arch,
                      const char *const *options, int n_options, void *image, size_t capacity,
                      size_t *image_size);
/* Architecture name of the current device, e.g. "gfx950:sramecc+:xnack-". */
int gt_plugin_device_arch(char *buf, size_t n);
/* Load / unload a code object on the current device (unload synchronizes it). */
int gt_plugin_module_load(const void *image, void **module);
int gt_plugin_module_unload(void *module);
/* Kernel of a loaded module by name. */
int gt_plugin_function(void *module, const char *name, void **function);
/* Launch grid[0..2] blocks of block[0..2] threads on `stream`; args[i] points
 * to the value of kernel argument i. */
int gt_plugin_launch(gt_stream_t stream, void *function, const unsigned *grid,
                     const unsigned *block, void **args);

#ifdef __cplusplus
}
#endif

#endif /* GRATOPY_HIP_H */
