/*
 * Body of the CPU oracle, included once per element type by oracle.c with
 *   REAL  = float | double
 *   SFX   = f32   | f64
 *   FLOOR/CEIL/FABS/FMIN/FMAX/HYPOT = the libm functions of that type.
 *
 * TEST INFRASTRUCTURE ONLY.  Each function restates one OpenCL kernel of the
 * reference with the same loop order, the same per-element arithmetic and
 * the same (int)/floor/ceil bound computations; the citation above each
 * function names the reference lines it follows.  Order arguments are 0 (F)
 * or 1 (C), as in gratopy/radon.cl:28-31.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* gratopy/radon.cl:66-161 (kernel radon_<t>_<o1><o2>) */
void FN(oracle_radon_fwd)(REAL *sino, const REAL *img, const REAL *ofs, const REAL *geo, int Nx,
                          int Ny, int Ns, int Na, int Nz, int sino_order, int img_order) {
  const float delta_x = (float)geo[0];
  const float delta_xi = (float)geo[1];
#pragma omp parallel for collapse(3) schedule(dynamic, 16)
  for (int z = 0; z < Nz; ++z)
    for (int a = 0; a < Na; ++a)
      for (int s = 0; s < Ns; ++s) {
        REAL ox = ofs[8 * a + 0], oy = ofs[8 * a + 1];
        const REAL oz = ofs[8 * a + 2], ow = ofs[8 * a + 3];
        const REAL reverse_mask = ofs[8 * a + 5];
        int Nxx = Nx, Nyy = Ny, horizontal = 1;
        if (reverse_mask != (REAL)0) {
          horizontal = 0;
          const REAL t = ox;
          ox = oy;
          oy = t;
          Nxx = Ny;
          Nyy = Nx;
        }
        REAL acc = 0;
        for (int y = 0; y < Nyy; ++y) {
          const REAL d = (REAL)y * oy + oz - (REAL)s;
          int x_low = (int)((-1 - d) * ow);
          int x_high = (int)((1 - d) * ow);
          if (ow < (REAL)0) {
            const int t = x_low;
            x_low = x_high;
            x_high = t;
          }
          if (x_low < 0) x_low = 0;
          if (x_high > Nxx - 1) x_high = Nxx - 1;
          for (int x = x_low; x <= x_high; ++x) {
            const REAL weight = (REAL)1 - FABS((REAL)x * ox + d);
            if (weight > (REAL)0) {
              const int ix = horizontal ? x : y, iy = horizontal ? y : x;
              acc += weight * img[img_pos(ix, iy, z, Nx, Ny, Nz, img_order)];
            }
          }
        }
        sino[sino_pos(s, a, z, Ns, Na, Nz, sino_order)] =
            acc * (REAL)delta_x * (REAL)delta_x / (REAL)delta_xi;
      }
}

/* gratopy/radon.cl:183-241 (kernel radon_ad_<t>_<o1><o2>) */
void FN(oracle_radon_back)(REAL *img, const REAL *sino, const REAL *ofs, const REAL *geo, int Nx,
                           int Ny, int Ns, int Na, int Nz, int img_order, int sino_order) {
  (void)geo;
#pragma omp parallel for collapse(3) schedule(static)
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y)
      for (int x = 0; x < Nx; ++x) {
        REAL acc = 0;
        for (int a = 0; a < Na; ++a) {
          const REAL *o = ofs + 8 * a;
          const REAL Delta_phi = o[4];
          const REAL s = (REAL)x * o[0] + (REAL)y * o[1] + o[2];
          int sm = (int)FLOOR(s);
          int sp = sm + 1;
          REAL weightp = 1 - ((REAL)sp - s);
          REAL weightm = 1 - (s - (REAL)sm);
          if (sm < 0 || sm >= Ns) {
            weightm = 0;
            sm = 0;
          }
          if (sp < 0 || sp >= Ns) {
            weightp = 0;
            sp = 0;
          }
          acc += Delta_phi * (weightm * sino[sino_pos(sm, a, z, Ns, Na, Nz, sino_order)] +
                              weightp * sino[sino_pos(sp, a, z, Ns, Na, Nz, sino_order)]);
        }
        img[img_pos(x, y, z, Nx, Ny, Nz, img_order)] = acc;
      }
}

/* Per-ray setup shared by the fanbeam forward and single-line kernels,
 * gratopy/fanbeam.cl:100-184 (and 446-523). */
typedef struct {
  REAL ddx, ddy, dx, dy, qx, qy, dpx, dmx, midx, midy, RE;
  int horizontal, Nxx, Nyy;
} FN(fan_ray);

static FN(fan_ray) FN(fan_ray_setup)(const REAL *ofs, const REAL *geo, int s, int a, int Nx,
                                     int Ny) {
  FN(fan_ray) r;
  const REAL R = geo[0];
  REAL RE = geo[1];
  const REAL delta_xi = geo[2];
  REAL midx = geo[3], midy = geo[4];
  const REAL midpoint_det = geo[5];
  const REAL *o = ofs + 8 * a;
  REAL dx = o[0], dy = o[1], qx = o[2], qy = o[3];
  const REAL d0x = o[4], d0y = o[5];
  const REAL kp = -midpoint_det + (REAL)s + 1;
  const REAL km = -midpoint_det + (REAL)s - 1;
  REAL dpx = d0x + dx * kp - qx, dpy = d0y + dy * kp - qy;
  REAL dmx = d0x + dx * km - qx, dmy = d0y + dy * km - qy;
  REAL ddx = d0x - qx, ddy = d0y - qy;
  ddx /= R * R;
  ddy /= R * R;
  dx /= delta_xi * delta_xi;
  dy /= delta_xi * delta_xi;
  RE *= 1 / R;
  int Nyy = Ny, Nxx = Nx, horizontal;
  if (FABS(dpx) < FABS(dpy) && FABS(dmx) < FABS(dmy)) {
    horizontal = 1;
  } else {
    REAL t;
    horizontal = 0;
    t = ddx; ddx = ddy; ddy = t;
    t = dx; dx = dy; dy = t;
    t = qx; qx = qy; qy = t;
    t = dpx; dpx = dpy; dpy = t;
    t = dmx; dmx = dmy; dmy = t;
    t = midx; midx = midy; midy = t;
    Nyy = Nx;
    Nxx = Ny;
  }
  dpx /= dpy;
  dmx /= dmy;
  if (qy * (dmx - dpx) < (REAL)0) {
    const REAL t = dpx;
    dpx = dmx;
    dmx = t;
  }
  r.ddx = ddx; r.ddy = ddy; r.dx = dx; r.dy = dy; r.qx = qx; r.qy = qy;
  r.dpx = dpx; r.dmx = dmx; r.midx = midx; r.midy = midy; r.RE = RE;
  r.horizontal = horizontal; r.Nxx = Nxx; r.Nyy = Nyy;
  return r;
}

/* gratopy/fanbeam.cl:72-256 (kernel fanbeam_<t>_<o1><o2>) */
void FN(oracle_fan_fwd)(REAL *sino, const REAL *img, const REAL *ofs, const REAL *sdpd,
                        const REAL *geo, int Nx, int Ny, int Ns, int Na, int Nz, int sino_order,
                        int img_order) {
  const REAL R = geo[0], delta_xi = geo[2], midpoint_det = geo[5], delta_x = geo[10];
#pragma omp parallel for collapse(3) schedule(dynamic, 16)
  for (int z = 0; z < Nz; ++z)
    for (int a = 0; a < Na; ++a)
      for (int s = 0; s < Ns; ++s) {
        const FN(fan_ray) r = FN(fan_ray_setup)(ofs, geo, s, a, Nx, Ny);
        const REAL s_midpoint_det = (REAL)s - midpoint_det;
        REAL acc = 0;
        for (int y = 0; y < r.Nyy; ++y) {
          const REAL xhigh = r.qx + r.dpx * ((REAL)y - r.midy - r.qy);
          const REAL xlow = r.qx + r.dmx * ((REAL)y - r.midy - r.qy);
          const int xhighint = (int)FLOOR(FMIN((REAL)(r.Nxx - 1) - r.midx, xhigh) + r.midx);
          const int xlowint = (int)CEIL(FMAX(-r.midx, xlow) + r.midx);
          REAL t = r.ddx * ((REAL)xlowint - r.midx) + r.ddy * ((REAL)y - r.midy) + r.RE;
          REAL ss = r.dx * ((REAL)xlowint - r.midx) + r.dy * ((REAL)y - r.midy);
          for (int x = xlowint; x <= xhighint; ++x) {
            const REAL xi = ss / t;
            const REAL Weight = ((REAL)1 - FABS(s_midpoint_det - xi)) / t;
            const int ix = r.horizontal ? x : y, iy = r.horizontal ? y : x;
            acc += Weight * img[img_pos(ix, iy, z, Nx, Ny, Nz, img_order)];
            t += r.ddx;
            ss += r.dx;
          }
        }
        sino[sino_pos(s, a, z, Ns, Na, Nz, sino_order)] = acc * sdpd[s] / delta_xi * delta_x / R;
      }
}

/* gratopy/fanbeam.cl:290-391 (kernel fanbeam_ad_<t>_<o1><o2>) */
void FN(oracle_fan_back)(REAL *img, const REAL *sino, const REAL *ofs, const REAL *sdpd,
                         const REAL *geo, int Nx, int Ny, int Ns, int Na, int Nz, int img_order,
                         int sino_order) {
  const REAL R = geo[0], delta_xi = geo[2], midx = geo[3], midy = geo[4], midpoint_det = geo[5];
  const REAL R_sqr_inv = (REAL)1 / (R * R);
  const REAL delta_xi_sqr_inv = (REAL)1 / (delta_xi * delta_xi);
#pragma omp parallel for collapse(3) schedule(static)
  for (int z = 0; z < Nz; ++z)
    for (int yy = 0; yy < Ny; ++yy)
      for (int xx = 0; xx < Nx; ++xx) {
        const REAL Px = (REAL)xx - midx, Py = (REAL)yy - midy;
        REAL acc = 0;
        for (int a = 0; a < Na; ++a) {
          const REAL *o = ofs + 8 * a;
          REAL dx = o[0], dy = o[1];
          const REAL qx = o[2], qy = o[3];
          REAL ddx = o[4] - qx, ddy = o[5] - qy;
          const REAL Delta_Phi = o[6];
          ddx *= R_sqr_inv;
          ddy *= R_sqr_inv;
          dx *= delta_xi_sqr_inv;
          dy *= delta_xi_sqr_inv;
          const REAL t = ddx * (Px - qx) + ddy * (Py - qy);
          const REAL ss = dx * (Px - qx) + dy * (Py - qy);
          const REAL xi = ss / t + midpoint_det;
          int xim = (int)FLOOR(xi);
          int xip = xim + 1;
          REAL Weightp = (REAL)1 - ((REAL)xip - xi);
          REAL Weightm = (REAL)1 - (xi - (REAL)xim);
          if (xim < 0 || xim >= Ns) {
            Weightm = 0;
            xim = 0;
          } else {
            Weightm *= sdpd[xim];
          }
          if (xip < 0 || xip >= Ns) {
            Weightp = 0;
            xip = 0;
          } else {
            Weightp *= sdpd[xip];
          }
          acc += Delta_Phi *
                 (Weightm * sino[sino_pos(xim, a, z, Ns, Na, Nz, sino_order)] +
                  Weightp * sino[sino_pos(xip, a, z, Ns, Na, Nz, sino_order)]) /
                 (R * t);
        }
        img[img_pos(xx, yy, z, Nx, Ny, Nz, img_order)] = acc;
      }
}

/* gratopy/radon.cl:263-341 (kernel single_line_radon_<t>_<o1><o2>), Nz = 1 */
void FN(oracle_single_line_radon)(REAL *sino, int px, int py, const REAL *ofs, const REAL *geo,
                                  int Nx, int Ny, int Ns, int Na, int sino_order) {
  const float delta_x = (float)geo[0];
  const float delta_xi = (float)geo[1];
  for (int a = 0; a < Na; ++a)
    for (int s = 0; s < Ns; ++s) {
      REAL ox = ofs[8 * a + 0], oy = ofs[8 * a + 1];
      const REAL oz = ofs[8 * a + 2], ow = ofs[8 * a + 3];
      int Nxx = Nx, x = px, y = py;
      if (ofs[8 * a + 5] != (REAL)0) {
        const REAL t = ox;
        ox = oy;
        oy = t;
        Nxx = Ny;
        x = py;
        y = px;
      }
      REAL acc = 0;
      const REAL d = (REAL)y * oy + oz;
      int x_low = (int)(((REAL)s - 1 - d) * ow);
      int x_high = (int)(((REAL)s + 1 - d) * ow);
      if (ow < (REAL)0) {
        const int t = x_low;
        x_low = x_high;
        x_high = t;
      }
      if (x_low < 0) x_low = 0;
      if (x_high > Nxx - 1) x_high = Nxx - 1;
      if (x_low <= x && x <= x_high) {
        const REAL weight = (REAL)1 - FABS((REAL)x * ox + d - (REAL)s);
        if (weight > (REAL)0) acc = weight;
      }
      sino[sino_pos(s, a, 0, Ns, Na, 1, sino_order)] =
          acc * (REAL)delta_x * (REAL)delta_x / (REAL)delta_xi;
    }
}

/* gratopy/fanbeam.cl:422-568 (kernel single_line_fanbeam_<t>_<o1><o2>), Nz = 1 */
void FN(oracle_single_line_fan)(REAL *sino, int px, int py, const REAL *ofs, const REAL *sdpd,
                                const REAL *geo, int Nx, int Ny, int Ns, int Na, int sino_order) {
  const REAL R = geo[0], delta_xi = geo[2], midpoint_det = geo[5], delta_x = geo[10];
  for (int a = 0; a < Na; ++a)
    for (int s = 0; s < Ns; ++s) {
      const FN(fan_ray) r = FN(fan_ray_setup)(ofs, geo, s, a, Nx, Ny);
      const int x = r.horizontal ? px : py, y = r.horizontal ? py : px;
      const REAL xhigh = r.qx + r.dpx * ((REAL)y - r.midy - r.qy);
      const REAL xlow = r.qx + r.dmx * ((REAL)y - r.midy - r.qy);
      const int xhighint = (int)FLOOR(FMIN((REAL)(r.Nxx - 1) - r.midx, xhigh) + r.midx);
      const int xlowint = (int)CEIL(FMAX(-r.midx, xlow) + r.midx);
      REAL acc = 0;
      if (xlowint <= x && x <= xhighint) {
        const REAL t = r.ddx * ((REAL)x - r.midx) + r.ddy * ((REAL)y - r.midy) + r.RE;
        const REAL ss = r.dx * ((REAL)x - r.midx) + r.dy * ((REAL)y - r.midy);
        const REAL xi = ss / t;
        acc = ((REAL)1 - FABS((REAL)s - xi - midpoint_det)) / (R * t);
      }
      sino[sino_pos(s, a, 0, Ns, Na, 1, sino_order)] = acc * sdpd[s] / delta_xi * delta_x;
    }
}

/* gratopy/utilities.cl:38-67 (multiply_* / divide_*) */
void FN(oracle_weight_sino)(REAL *out, const REAL *in, const REAL *w, int Ns, int Na, int Nz,
                            int order, int divide) {
  for (int z = 0; z < Nz; ++z)
    for (int a = 0; a < Na; ++a)
      for (int s = 0; s < Ns; ++s) {
        const size_t i = sino_pos(s, a, z, Ns, Na, Nz, order);
        out[i] = divide ? in[i] / w[a] : in[i] * w[a];
      }
}

/* gratopy/utilities.cl:69-73 (equ_mul_add_*) */
void FN(oracle_equ_mul_add)(REAL *rhs, REAL a, const REAL *x, size_t n) {
  for (size_t i = 0; i < n; ++i) rhs[i] += a * x[i];
}

/* gratopy/utilities.cl:75-80 (mul_add_add_*) */
void FN(oracle_mul_add_add)(REAL *rhs, REAL a, const REAL *x, const REAL *y, size_t n) {
  for (size_t i = 0; i < n; ++i) rhs[i] = a * x[i] + y[i];
}

/* gratopy/total_variation.cl:54-93 (update_v_*); v holds 4 reals per pixel */
void FN(oracle_tv_update_v)(REAL *v, const REAL *u, float sigma, float z_distance, int Nx, int Ny,
                            int Nz, int order) {
  const size_t sx = img_pos(1, 0, 0, Nx, Ny, Nz, order), sy = img_pos(0, 1, 0, Nx, Ny, Nz, order),
               sz = img_pos(0, 0, 1, Nx, Ny, Nz, order);
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y)
      for (int x = 0; x < Nx; ++x) {
        const size_t i = img_pos(x, y, z, Nx, Ny, Nz, order);
        REAL v0 = -u[i], v1 = -u[i], v2 = -u[i];
        if (x < Nx - 1) v0 += u[i + sx]; else v0 = 0;
        if (y < Ny - 1) v1 += u[i + sy]; else v1 = 0;
        if (z < Nz - 1) v2 += u[i + sz]; else v2 = 0;
        if (z_distance > 0) v2 /= z_distance; else v2 = 0;
        v[4 * i + 0] += sigma * v0;
        v[4 * i + 1] += sigma * v1;
        v[4 * i + 2] += sigma * v2;
      }
}

/* gratopy/total_variation.cl:150-167 (update_NormV_unchor_*) */
void FN(oracle_tv_update_normv)(REAL *v, int Nx, int Ny, int Nz, int order) {
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y)
      for (int x = 0; x < Nx; ++x) {
        const size_t i = img_pos(x, y, z, Nx, Ny, Nz, order);
        const REAL norm = HYPOT(HYPOT(v[4 * i], v[4 * i + 1]), v[4 * i + 2]);
        if (norm > (REAL)1) {
          v[4 * i] /= norm;
          v[4 * i + 1] /= norm;
          v[4 * i + 2] /= norm;
        }
      }
}

/* gratopy/total_variation.cl:95-106 (update_lambda_L2_*) */
void FN(oracle_tv_update_lambda)(REAL *lambda, const REAL *ku, const REAL *f, float sigma,
                                 float sigmap1inv, size_t n) {
  for (size_t i = 0; i < n; ++i) lambda[i] = (lambda[i] + sigma * (ku[i] - f[i])) * sigmap1inv;
}

/* gratopy/total_variation.cl:108-148 (update_u_*): u = max(0, u_ + tau(div v - n K*l)) */
void FN(oracle_tv_update_u)(REAL *u, const REAL *u_, const REAL *v, const REAL *kstar, float tau,
                            float norming, float z_distance, int Nx, int Ny, int Nz, int order) {
  const size_t sx = img_pos(1, 0, 0, Nx, Ny, Nz, order), sy = img_pos(0, 1, 0, Nx, Ny, Nz, order),
               sz = img_pos(0, 0, 1, Nx, Ny, Nz, order);
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y)
      for (int x = 0; x < Nx; ++x) {
        const size_t i = img_pos(x, y, z, Nx, Ny, Nz, order);
        REAL v0 = v[4 * i], v1 = v[4 * i + 1], v2 = v[4 * i + 2];
        if (x == Nx - 1) v0 = 0;
        if (x > 0) v0 -= v[4 * (i - sx)];
        if (y == Ny - 1) v1 = 0;
        if (y > 0) v1 -= v[4 * (i - sy) + 1];
        if (z == Nz - 1) v2 = 0;
        if (z > 0) v2 -= v[4 * (i - sz) + 2];
        if (z_distance > 0) v2 /= z_distance; else v2 = 0;
        u[i] = FMAX((REAL)0, u_[i] + tau * (v0 + v1 + v2 - norming * kstar[i]));
      }
}

/* ElementwiseKernel update_extra_ of gratopy/gratopy.py:2708-2717: u = 2 u_ - u */
void FN(oracle_tv_update_extra)(const REAL *u_, REAL *u, size_t n) {
  for (size_t i = 0; i < n; ++i) u[i] = (REAL)2 * u_[i] - u[i];
}

#undef FN
#undef CAT
#undef CAT2
