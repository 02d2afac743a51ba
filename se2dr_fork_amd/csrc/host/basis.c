/*
 * basis.c -- GLL quadrature and the 1D Lagrange basis of se2wave.
 *
 * The basis is built the way the reference builds it: monomial Vandermonde
 * V[k][i] = xi_k^i, basis coefficients from LU solves V c = e_k
 * (se2wave.c:279-334, 413-468), then evaluation of the monomial expansion
 * with entries below 1e-12 set to zero (352-354, 494-496).  The LU follows
 * reference LAPACK dgetrf/dgetrs (the factorisation PETSc's PCLU applies to
 * a SeqDense matrix): partial pivoting on the first largest |a|, multipliers
 * scaled by the reciprocal pivot, rank-1 updates a + l*(-u), column-oriented
 * triangular solves.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "host_internal.h"

#define GLL_PI 3.14159265358979323846264338327950288

/* CreateGLLCoordsWeights (se2wave.c:103-209): Newton iteration on the
 * Legendre recurrence from the Chebyshev-Gauss-Lobatto guess, until the
 * 2-norm of the update is <= 1e-12.  Weights 2/(N(N+1)P_N^2) are formed on
 * the descending nodes and not reordered when the nodes are flipped. */
int se2w_gll(int p, double *xi, double *w)
{
  const int n1 = p + 1;
  double *x, *xprev, *leg, res;
  int i, k;
  if (p < 1) return host_fail(SE2W_HOST_ERR_ARG, "basis degree must be >= 1 (got %d)", p);
  x = (double *)malloc(sizeof(double) * n1);
  xprev = (double *)malloc(sizeof(double) * n1);
  leg = (double *)calloc((size_t)n1 * n1, sizeof(double)); /* leg[i + k*n1] = P_k(x_i) */
  if (!x || !xprev || !leg) { free(x); free(xprev); free(leg); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  for (i = 0; i < n1; i++) x[i] = cos(GLL_PI * i / (double)p);
  res = 1.0;
  while (res > 1.0e-12) {
    for (i = 0; i < n1; i++) xprev[i] = x[i];
    for (i = 0; i < n1; i++) {
      leg[i] = 1.0;
      leg[i + n1] = x[i];
      for (k = 1; k < p; k++) {
        leg[i + (k + 1) * n1] =
            ((2.0 * (k + 1) - 1.0) * x[i] * leg[i + k * n1] - (k + 1.0 - 1.0) * leg[i + (k - 1) * n1]) / (double)(k + 1.0);
      }
    }
    for (i = 0; i < n1; i++) {
      x[i] = xprev[i] - (x[i] * leg[i + p * n1] - leg[i + (p - 1) * n1]) / (n1 * leg[i + p * n1]);
    }
    res = 0.0;
    for (i = 0; i < n1; i++) res += (x[i] - xprev[i]) * (x[i] - xprev[i]);
    res = sqrt(res);
  }
  if (w) {
    for (i = 0; i < n1; i++) {
      const double pn = leg[i + p * n1];
      w[i] = 2.0 / (p * n1 * pn * pn);
    }
  }
  if (xi) {
    for (i = 0; i < n1; i++) xi[i] = x[n1 - 1 - i];
  }
  free(x); free(xprev); free(leg);
  return 0;
}

/* In-place LU of the column-major n x n matrix a (LAPACK dgetrf order). */
static void lu_factor(int n, double *a, int *ipiv)
{
  int col, r, c;
  for (col = 0; col < n; col++) {
    int piv = col;
    double big = fabs(a[col + col * n]);
    for (r = col + 1; r < n; r++) {
      const double v = fabs(a[r + col * n]);
      if (v > big) { big = v; piv = r; }
    }
    ipiv[col] = piv;
    if (a[piv + col * n] != 0.0) {
      if (piv != col) {
        for (c = 0; c < n; c++) {
          const double t = a[col + c * n];
          a[col + c * n] = a[piv + c * n];
          a[piv + c * n] = t;
        }
      }
      {
        const double inv = 1.0 / a[col + col * n];
        for (r = col + 1; r < n; r++) a[r + col * n] *= inv;
      }
    }
    for (c = col + 1; c < n; c++) {
      const double ujc = a[col + c * n];
      if (ujc != 0.0) {
        const double neg = -ujc;
        for (r = col + 1; r < n; r++) a[r + c * n] = a[r + c * n] + a[r + col * n] * neg;
      }
    }
  }
}

/* Solve with the factors (LAPACK dgetrs order: row swaps, unit lower, upper). */
static void lu_solve(int n, const double *lu, const int *ipiv, double *b)
{
  int k, r;
  for (k = 0; k < n; k++) {
    if (ipiv[k] != k) {
      const double t = b[k];
      b[k] = b[ipiv[k]];
      b[ipiv[k]] = t;
    }
  }
  for (k = 0; k < n; k++) {
    if (b[k] != 0.0)
      for (r = k + 1; r < n; r++) b[r] = b[r] - b[k] * lu[r + k * n];
  }
  for (k = n - 1; k >= 0; k--) {
    if (b[k] != 0.0) {
      b[k] = b[k] / lu[k + k * n];
      for (r = 0; r < k; r++) b[r] = b[r] - b[k] * lu[r + k * n];
    }
  }
}

/* coef[i*(p+1) + m]: coefficient of x^m in the Lagrange basis function i. */
static int lagrange_coefficients(int p, double *coef)
{
  const int nb = p + 1;
  double *nodes = (double *)malloc(sizeof(double) * nb);
  double *vdm = (double *)malloc(sizeof(double) * nb * nb);
  double *rhs = (double *)malloc(sizeof(double) * nb);
  int *ipiv = (int *)malloc(sizeof(int) * nb);
  int i, m, ierr;
  if (!nodes || !vdm || !rhs || !ipiv) { ierr = host_fail(SE2W_HOST_ERR_MEM, "out of memory"); goto done; }
  if ((ierr = se2w_gll(p, nodes, NULL))) goto done;
  for (i = 0; i < nb; i++)
    for (m = 0; m < nb; m++) vdm[i + m * nb] = pow(nodes[i], (double)m);
  lu_factor(nb, vdm, ipiv);
  for (i = 0; i < nb; i++) {
    for (m = 0; m < nb; m++) rhs[m] = 0.0;
    rhs[i] = 1.0;
    lu_solve(nb, vdm, ipiv, rhs);
    for (m = 0; m < nb; m++) coef[i * nb + m] = rhs[m];
  }
  ierr = 0;
done:
  free(nodes); free(vdm); free(rhs); free(ipiv);
  return ierr;
}

/* Shared evaluation of sum_m coef[i][m] * mono_m(x) with the 1e-12 cut. */
static int tabulate(int p, int npoints, const double *x, double *out, int derivative)
{
  const int nb = p + 1;
  double *coef, *mono;
  int k, i, m, ierr;
  if (p < 1 || npoints < 0 || (npoints > 0 && (!x || !out))) return host_fail(SE2W_HOST_ERR_ARG, "bad tabulation arguments");
  coef = (double *)malloc(sizeof(double) * nb * nb);
  mono = (double *)malloc(sizeof(double) * nb);
  if (!coef || !mono) { free(coef); free(mono); return host_fail(SE2W_HOST_ERR_MEM, "out of memory"); }
  if ((ierr = lagrange_coefficients(p, coef))) { free(coef); free(mono); return ierr; }
  for (k = 0; k < npoints; k++) {
    for (m = 0; m < nb; m++) {
      if (!derivative) mono[m] = pow((double)x[k], (double)m);
      else mono[m] = (m == 0) ? 0.0 : ((double)m) * pow((double)x[k], (double)(m - 1));
    }
    for (i = 0; i < nb; i++) {
      double acc = 0.0;
      for (m = 0; m < nb; m++) acc += coef[i * nb + m] * mono[m];
      out[k * nb + i] = (fabs(acc) < 1.0e-12) ? 0.0 : acc;
    }
  }
  free(coef); free(mono);
  return 0;
}

int se2w_basis1d(int p, int npoints, const double *x, double *N) { return tabulate(p, npoints, x, N, 0); }

int se2w_basis1d_deriv(int p, int npoints, const double *x, double *dN) { return tabulate(p, npoints, x, dN, 1); }

int se2w_deriv_matrix(int p, double *D)
{
  double nodes[HOST_MAX_NB];
  int ierr;
  if (p < 1 || p + 1 > HOST_MAX_NB) return host_fail(SE2W_HOST_ERR_SUP, "basis degree %d outside 1..%d", p, HOST_MAX_NB - 1);
  if ((ierr = se2w_gll(p, nodes, NULL))) return ierr;
  return tabulate(p, p + 1, nodes, D, 1);
}
