/*
 * mtr_synth.h -- seeded synthetic tilt-series generator (C ABI).
 *
 * The reference ships no data (its only run reads cluster paths,
 * run_multiaxis.sh:6-41), so benchmarks and tests build their inputs from:
 *   - an additive phantom of random ellipsoids,
 *   - per-view affine geometry in the reference's coefficient convention
 *     (u = xc0 + xc1*X + xc2*Y + xc3*Z - 1, FunctionOnGPU.cu:17-18),
 *   - analytic chord-length ray sums through the ellipsoids,
 *   - i.i.d. Gaussian noise scaled to a fraction of the largest ray sum.
 * Random numbers come from xoshiro256** seeded through splitmix64, with
 * U(0,1) = (x >> 11) * 2^-53 and Box-Muller normals, so every stream is
 * reproducible bit for bit on any host.
 */
#ifndef MTR_SYNTH_H_
#define MTR_SYNTH_H_

#include <stdint.h>
#include "mtr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Doubles per ellipsoid record: centre[3], semi-axes[3], rotation[9]
 * (rows are the principal axes in volume coordinates), density. */
#define MTR_ELLIPSOID_DOUBLES 16

/* Geometry conventions. */
enum {
    /* beam along Z at zero tilt; tilt axis in the XY plane at angle phi
     * from X: u = cu + cos(t)*(cos(phi) dX + sin(phi) dY) + sin(t) dZ,
     *         v = cv - sin(phi) dX + cos(phi) dY                          */
    MTR_GEOM_SLAB = 0,
    /* tilt axis along Z: u = cu + cos(t) dX + sin(t) dY, v = cv + dZ      */
    MTR_GEOM_AXIS_Z = 1
};

/* Per-view coefficients xc[4*AngN], yc[4*AngN] for tilts theta_deg[AngN]. */
int mtr_synth_geometry(const mtr_volume_size *vol, const mtr_projection_size *prj, int kind,
                       double phi_deg, const double *theta_deg, double *xc, double *yc);
/* n ellipsoids: centres uniform in the box, semi-axes uniform [amin,amax],
 * uniform random orientation (Shoemake quaternion), density uniform
 * [dmin,dmax]. */
int mtr_synth_phantom(uint64_t seed, const mtr_volume_size *vol, int n, double amin,
                      double amax, double dmin, double dmax, double *ell);
/* Analytic ray sums p[AngN][Y][X]: for each detector cell the line of
 * voxel-space points that the geometry maps onto it, integrated through
 * every ellipsoid.  threads <= 0 uses all host cores. */
int mtr_synth_project(const mtr_volume_size *vol, const mtr_projection_size *prj,
                      const double *xc, const double *yc, int n, const double *ell, float *p,
                      int threads);
/* p += sigma * N(0,1) with sigma = frac * max(p).  Returns sigma in
 * *sigma_out (nullable). */
int mtr_synth_noise(uint64_t seed, int64_t count, float *p, double frac, double *sigma_out);
/* Voxelised phantom (density summed over ellipsoids containing each voxel
 * centre), [Z][Y][X]. */
int mtr_synth_voxelize(const mtr_volume_size *vol, int n, const double *ell, float *x,
                       int threads);
/* Raw generator access for tests: n doubles U(0,1) and n normals. */
int mtr_synth_uniform(uint64_t seed, int64_t n, double *out);
int mtr_synth_normal(uint64_t seed, int64_t n, double *out);

#ifdef __cplusplus
}
#endif

#endif /* MTR_SYNTH_H_ */
