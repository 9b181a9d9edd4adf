/*
 * sw2d_gen.h -- seeded synthetic inputs for the benchmark configurations
 * (no counterpart in the reference, which ships only measured case data).
 * All generators are deterministic for a given seed (splitmix64-seeded
 * xoshiro256** stream, fixed traversal order) and return 0 or a negative
 * SW2D_E_* code.
 */
#ifndef SW2D_GEN_H
#define SW2D_GEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Diamond-square fractal terrain on the smallest (2^k+1)^2 lattice that
 * covers rows x cols, cropped to the top-left rows x cols and rescaled to
 * [zmin, zmax].  The displacement amplitude is multiplied by `roughness`
 * at every halving of the step. */
int sw2d_gen_fractal_dem(int rows, int cols, uint64_t seed, double roughness,
                         double zmin, double zmax, double *out);
/* Adds a plane falling by `slope` per metre toward the given edge
 * (SW2D_EDGE_* bit): z += slope * dx * (distance in cells to that edge). */
int sw2d_gen_tilt(int rows, int cols, double dx, double slope, int edge,
                  double *z);
int sw2d_gen_uniform(size_t n, uint64_t seed, double lo, double hi,
                     double *out);
/* exp(N(mu, sigma^2)) */
int sw2d_gen_lognormal(size_t n, uint64_t seed, double mu, double sigma,
                       double *out);
/* Gamma(shape k, scale theta) (Marsaglia-Tsang, with the k < 1 boost) */
int sw2d_gen_gamma(size_t n, uint64_t seed, double k, double theta,
                   double *out);

#ifdef __cplusplus
}
#endif
#endif
