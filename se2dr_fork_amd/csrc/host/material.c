/*
 * material.c -- seeded cell-wise material models for the benchmark
 * workloads (random per cell, layered).  The reference's own perturbed model
 * (SpecFECtxSetPerturbedMaterialProperties_Velocity, se2wave.c:1177-1236)
 * draws from PETSc's RAND48; these models use splitmix64 instead so that the
 * same arrays can be regenerated anywhere from the seed.
 */
#include <math.h>
#include <stdint.h>

#include "host_internal.h"

/* splitmix64 (Steele, Lea, Flood, OOPSLA 2014). */
uint64_t se2w_rng_next(uint64_t *state)
{
  uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double se2w_rng_uniform(uint64_t *state) { return (double)(se2w_rng_next(state) >> 11) * 0x1.0p-53; }

int se2w_material_random_cells(int mx, int my, uint64_t seed, double vp_lo, double vp_hi, double rho_lo, double rho_hi,
                               double *vp, double *vs, double *rho)
{
  uint64_t st = seed;
  int e;
  if (mx < 1 || my < 1 || !vp || !vs || !rho) return host_fail(SE2W_HOST_ERR_ARG, "bad material arguments");
  for (e = 0; e < mx * my; e++) {
    vp[e] = vp_lo + (vp_hi - vp_lo) * se2w_rng_uniform(&st);
    rho[e] = rho_lo + (rho_hi - rho_lo) * se2w_rng_uniform(&st);
    vs[e] = vp[e] / sqrt(3.0);
  }
  return 0;
}

int se2w_material_layered(int mx, int my, int nlayers, uint64_t seed, double vp_lo, double vp_hi, double rho_lo,
                          double rho_hi, double *vp, double *vs, double *rho)
{
  uint64_t st = seed;
  double lvp[4096], lrho[4096];
  int l, ei, ej;
  if (mx < 1 || my < 1 || nlayers < 1 || nlayers > 4096 || !vp || !vs || !rho)
    return host_fail(SE2W_HOST_ERR_ARG, "bad material arguments");
  for (l = 0; l < nlayers; l++) {
    lvp[l] = vp_lo + (vp_hi - vp_lo) * se2w_rng_uniform(&st);
    lrho[l] = rho_lo + (rho_hi - rho_lo) * se2w_rng_uniform(&st);
  }
  for (ej = 0; ej < my; ej++) {
    int layer = (int)(((long)ej * nlayers) / my);
    if (layer >= nlayers) layer = nlayers - 1;
    for (ei = 0; ei < mx; ei++) {
      const int e = ei + ej * mx;
      vp[e] = lvp[layer];
      rho[e] = lrho[layer];
      vs[e] = vp[e] / sqrt(3.0);
    }
  }
  return 0;
}
