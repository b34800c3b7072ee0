/*
 * CPU oracle of the gratopy projection hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference OpenCL kernels (gratopy/radon.cl,
 * gratopy/fanbeam.cl, gratopy/total_variation.cl, gratopy/utilities.cl),
 * compiled with -ffp-contract=off so that every product and sum is rounded
 * separately, in the element type of the call, like the reference source
 * reads.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker.
 *
 * How each function is pinned (see tests/test_oracle.py): against the
 * reference's own control numbers (tests/test_radon.py, tests/test_fanbeam.py
 * "evaluate_control_numbers"), against an independent dense closed-form
 * evaluation (oracle/closed_form.py) on small grids, and against adjointness
 * and mass-preservation invariants.
 */
#include <math.h>
#include <stddef.h>

/* gratopy/radon.cl:28-31 */
static inline size_t img_pos(int x, int y, int z, int Nx, int Ny, int Nz, int order) {
  return order == 0 ? (size_t)x + (size_t)Nx * ((size_t)y + (size_t)Ny * z)
                    : (size_t)z + (size_t)Nz * ((size_t)y + (size_t)Ny * x);
}
static inline size_t sino_pos(int s, int a, int z, int Ns, int Na, int Nz, int order) {
  return order == 0 ? (size_t)s + (size_t)Ns * ((size_t)a + (size_t)Na * z)
                    : (size_t)z + (size_t)Nz * ((size_t)a + (size_t)Na * s);
}

#define REAL float
#define SFX f32
#define FLOOR floorf
#define CEIL ceilf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#define HYPOT hypotf
#include "oracle_kernels.h"
#undef REAL
#undef SFX
#undef FLOOR
#undef CEIL
#undef FABS
#undef FMIN
#undef FMAX
#undef HYPOT

#define REAL double
#define SFX f64
#define FLOOR floor
#define CEIL ceil
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#define HYPOT hypot
#include "oracle_kernels.h"
