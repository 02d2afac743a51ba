/*
 * material.c -- cell-wise material models: the reference's perturbed model
 * (SpecFECtxSetPerturbedMaterialProperties_Velocity, se2wave.c:1177-1236),
 * whose draws follow PETSc's PETSCRAND48 generator (srand48/drand48), and the
 * seeded models of the benchmark workloads (random per cell, layered), which
 * use splitmix64 so that the same arrays can be regenerated anywhere.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

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

/* The 48-bit linear congruential generator of POSIX drand48:
 * X <- (0x5DEECE66D X + 0xB) mod 2^48, value X / 2^48; srand48(s) sets the
 * high 32 bits of X to s and the low 16 bits to 0x330E. */
void se2w_rand48_seed(uint64_t *state, long seed)
{
  *state = (((uint64_t)(uint32_t)seed) << 16) | 0x330EULL;
}

double se2w_rand48_next(uint64_t *state)
{
  *state = (0x5DEECE66DULL * *state + 0xBULL) & 0xFFFFFFFFFFFFULL;
  return ldexp((double)*state, -48);
}

/* Draws of VecSetRandom with a PETSCRAND48 PetscRandom on one process:
 * seeded with `seed`, value = (high - low) * drand48() + low per entry. */
static void rand48_fill(double *x, int n, long seed, double low, double high)
{
  uint64_t st;
  const double width = high - low;
  int i;
  se2w_rand48_seed(&st, seed);
  for (i = 0; i < n; i++) {
    const double d = se2w_rand48_next(&st);
    const double t = width * d;
    x[i] = t + low;
  }
}

int se2w_spec_set_perturbed_material_velocity(SE2WSpec spec, double vp0, double delta_vp, double vs0, double delta_vs,
                                              double rho0, double delta_rho)
{
  double *vp, *vs, *rho;
  int ne, ierr;
  if (!spec) return host_fail(SE2W_HOST_ERR_ARG, "null spec");
  ne = spec->mx * spec->my;
  vp = (double *)malloc(sizeof(double) * ne);
  vs = (double *)malloc(sizeof(double) * ne);
  rho = (double *)malloc(sizeof(double) * ne);
  if (!vp || !vs || !rho) {
    free(vp); free(vs); free(rho);
    return host_fail(SE2W_HOST_ERR_MEM, "out of memory");
  }
  /* seeds 1, 2, 3 for Vp, Vs, rho (se2wave.c:1194-1208) */
  rand48_fill(vp, ne, 1, vp0 - delta_vp, vp0 + delta_vp);
  rand48_fill(vs, ne, 2, vs0 - delta_vs, vs0 + delta_vs);
  rand48_fill(rho, ne, 3, rho0 - delta_rho, rho0 + delta_rho);
  ierr = se2w_spec_set_cell_material_velocity(spec, vp, vs, rho);
  free(vp); free(vs); free(rho);
  if (ierr) return ierr;
  printf("  [material]     Vp0 = %1.8e : delta = %+1.8e\n", vp0, delta_vp);
  printf("  [material]     Vs0 = %1.8e : delta = %+1.8e\n", vs0, delta_vs);
  printf("  [material]    rho0 = %1.8e : delta = %+1.8e\n", rho0, delta_rho);
  return 0;
}
