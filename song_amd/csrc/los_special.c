/*
 * los_special.c — FP64 special functions for building the projection tables:
 * spherical Bessel functions j_l(x) for all l <= lmax at one x, Wigner 3j
 * symbols over their first angular momentum, and the clamped cubic spline
 * used for the k3 interpolation of the LOS sources.
 *
 * These stand in for routines of the un-vendored CLASS fork that SONG calls
 * (bessel_j at source/bessel2.c:1706, drc3jj at :1131/:1154 and
 * array_spline_table_columns at source/transfer2.c:2639); they are written
 * from the published methods named at each function, not from CLASS.
 */
#include "los_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/*
 * j_l(x) for l = 0..lmax (out has lmax+1 entries).
 *  x == 0          : j_0 = 1, j_l>0 = 0.
 *  x >  lmax + 1   : upward recurrence j_{l+1} = (2l+1)/x j_l - j_{l-1}, which is
 *                    stable while l < x.
 *  otherwise       : Miller's downward recurrence from an order N well above the
 *                    turning point, normalised against the exact j_0 and j_1
 *                    (least-squares match of both, so a zero of either is harmless).
 */
void los_sph_bessel_array(int lmax, double x, double *out)
{
  if (x == 0.0) {
    out[0] = 1.0;
    for (int l = 1; l <= lmax; ++l) out[l] = 0.0;
    return;
  }
  const double s = sin(x), c = cos(x);
  const double j0 = s / x;
  const double j1 = (s / x - c) / x;
  if (x > (double)lmax + 1.0) {
    out[0] = j0;
    if (lmax >= 1) out[1] = j1;
    for (int l = 1; l < lmax; ++l)
      out[l + 1] = (2.0 * l + 1.0) / x * out[l] - out[l - 1];
    return;
  }
  /* Downward recurrence from an order past both lmax and the turning point x,
     with a margin that covers the Airy transition width ~ x^(1/3). */
  const double top = (double)lmax > x ? (double)lmax : x;
  const int N = (int)top + 40 + (int)(10.0 * cbrt(top + 1.0));
  double fp1 = 0.0, f = 1e-300; /* f_{l}, f_{l-1} as the loop runs */
  for (int l = N; l >= 1; --l) {
    const double fm1 = (2.0 * l + 1.0) / x * f - fp1;
    fp1 = f;
    f = fm1;
    if (fabs(f) > 1e250) {
      f *= 1e-250;
      fp1 *= 1e-250;
      for (int q = l; q <= lmax; ++q) out[q] *= 1e-250;
    }
    if (l - 1 <= lmax) out[l - 1] = f;
  }
  /* f = f_0, fp1 = f_1 (unnormalised, up to 1e250: divide out their size
     before squaring) */
  const double big = fabs(f) > fabs(fp1) ? fabs(f) : fabs(fp1);
  const double g0 = f / big, g1 = fp1 / big;
  const double scale = (j0 * g0 + j1 * g1) / (g0 * g0 + g1 * g1) / big;
  for (int l = 0; l <= lmax; ++l) out[l] *= scale;
}

/*
 * Wigner 3j symbols (j1 j2 j3; m1 m2 m3) for all j1 in [j1min, j1max] with
 * m1 = -m2-m3, via the three-term recursion of Schulten & Gordon,
 * J. Math. Phys. 16, 1961 (1975):
 *   j A(j+1) f(j+1) + B(j) f(j) + (j+1) A(j) f(j-1) = 0,
 *   A(j) = sqrt[(j^2-(j2-j3)^2)((j2+j3+1)^2-j^2)(j^2-m1^2)],
 *   B(j) = -(2j+1)[j2(j2+1) m1 - j3(j3+1) m1 - j(j+1)(m3-m2)].
 * The sequence is built downward from j1max (where A(j1max+1)=0) and upward
 * from j1min (where A(j1min)=0, unless j1min==0), the two halves are matched
 * at their midpoint, normalised by sum (2j1+1) f^2 = 1 and given the
 * Condon-Shortley sign sgn f(j1max) = (-1)^(j2-j3-m1).
 * Returns the count (j1max-j1min+1), or -1 if the symbols all vanish.
 */
int los_wigner3j_j1(int j2, int j3, int m2, int m3, int *j1min_out, int *j1max_out,
                    double *out, int out_cap)
{
  const int m1 = -m2 - m3;
  if (abs(m2) > j2 || abs(m3) > j3) return -1;
  int j1min = abs(j2 - j3);
  if (abs(m1) > j1min) j1min = abs(m1);
  const int j1max = j2 + j3;
  if (j1min > j1max) return -1;
  const int n = j1max - j1min + 1;
  if (n > out_cap) return -1;
  *j1min_out = j1min;
  *j1max_out = j1max;
  if (n == 1) {
    out[0] = ((j2 - j3 - m1) % 2 == 0 ? 1.0 : -1.0) / sqrt(2.0 * j1min + 1.0);
    return 1;
  }
  const double dj2 = j2, dj3 = j3, dm1 = m1, dm2 = m2, dm3 = m3;
#define A_(j) sqrt(((double)(j) * (j) - (dj2 - dj3) * (dj2 - dj3)) * \
                   ((dj2 + dj3 + 1.0) * (dj2 + dj3 + 1.0) - (double)(j) * (j)) * \
                   ((double)(j) * (j) - dm1 * dm1))
#define B_(j) (-(2.0 * (j) + 1.0) * (dj2 * (dj2 + 1.0) * dm1 - dj3 * (dj3 + 1.0) * dm1 - \
                                     (double)(j) * ((j) + 1.0) * (dm3 - dm2)))
  /* Downward pass over the whole range, in the caller's buffer (n <= out_cap). */
  double *down = out;
  down[n - 1] = 1.0;
  {
    int j = j1max;
    double fj = 1.0, fjp1 = 0.0;
    for (int idx = n - 1; idx >= 1; --idx, --j) {
      double a_j = A_(j);
      double a_jp1 = (j + 1 <= j1max) ? A_(j + 1) : 0.0;
      double fjm1 = -((double)j * a_jp1 * fjp1 + B_(j) * fj) / ((j + 1.0) * a_j);
      down[idx - 1] = fjm1;
      fjp1 = fj;
      fj = fjm1;
      double big = fabs(fj) > fabs(fjp1) ? fabs(fj) : fabs(fjp1);
      if (big > 1e100) {
        for (int q = idx - 1; q < n; ++q) down[q] *= 1e-100;
        fj *= 1e-100;
        fjp1 *= 1e-100;
      }
    }
  }
  /* Upward pass, only when the first step is well defined (j1min > 0). */
  if (j1min > 0 && n >= 3) {
    const int mid = n / 2;
    double up_small[64];
    double *up = mid + 1 <= 64 ? up_small : (double *)malloc((size_t)(mid + 1) * sizeof(double));
    if (up == NULL) return -1;
    up[0] = 1.0;
    double fjm1 = 0.0, fj = 1.0;
    int j = j1min;
    for (int idx = 0; idx < mid; ++idx, ++j) {
      double fjp1 = -(B_(j) * fj + (j + 1.0) * A_(j) * fjm1) / ((double)j * A_(j + 1));
      up[idx + 1] = fjp1;
      fjm1 = fj;
      fj = fjp1;
    }
    /* Match at the index with the largest magnitude among the overlap. */
    int best = mid;
    double bmag = fabs(down[mid]);
    for (int q = 1; q <= mid; ++q)
      if (fabs(down[q]) > bmag) { bmag = fabs(down[q]); best = q; }
    if (bmag > 0.0 && up[best] != 0.0) {
      double r = down[best] / up[best];
      for (int q = 0; q < best; ++q) down[q] = up[q] * r;
    }
    if (up != up_small) free(up);
  }
#undef A_
#undef B_
  double norm = 0.0;
  for (int q = 0; q < n; ++q) norm += (2.0 * (j1min + q) + 1.0) * down[q] * down[q];
  double sgn = ((j2 - j3 - m1) % 2 == 0) ? 1.0 : -1.0;
  double scale = sgn / sqrt(norm);
  if (down[n - 1] < 0) scale = -scale;
  for (int q = 0; q < n; ++q) out[q] *= scale;
  return n;
}

/*
 * Second derivatives of cubic splines through (x[i], y[c*n + i]) for ncol
 * columns laid out contiguous in x (the layout of the sources: for each
 * tau, the k3 values are contiguous). End conditions, as the reference asks
 * with _SPLINE_EST_DERIV_ (source/transfer2.c:2645): the first derivative at
 * each end is that of the parabola through the three end points; with
 * natural=1 the second derivative vanishes at both ends. Tridiagonal solve as
 * in Press et al., Numerical Recipes, 2nd ed., sec. 3.3 (routine spline).
 * work needs n doubles.
 */
int los_spline_columns(const double *x, int n, const double *y, int ncol, double *dd,
                       int natural, double *work)
{
  if (n < 3) return _FAILURE_;
  for (int c = 0; c < ncol; ++c) {
    const double *yc = y + (long)c * n;
    double *ddc = dd + (long)c * n;
    los_spline_one(x, n, yc, ddc, natural, work);
  }
  return _SUCCESS_;
}

void los_spline_one(const double *x, int n, const double *y, double *dd, int natural,
                    double *u)
{
  if (natural) {
    dd[0] = 0.0;
    u[0] = 0.0;
  } else {
    double d0 = ((x[2] - x[0]) * (x[2] - x[0]) * (y[1] - y[0]) -
                 (x[1] - x[0]) * (x[1] - x[0]) * (y[2] - y[0])) /
                ((x[2] - x[0]) * (x[1] - x[0]) * (x[2] - x[1]));
    dd[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - d0);
  }
  for (int i = 1; i < n - 1; ++i) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * dd[i - 1] + 2.0;
    dd[i] = (sig - 1.0) / p;
    double t = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * t / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  double qn, un;
  if (natural) {
    qn = 0.0;
    un = 0.0;
  } else {
    const int a = n - 3, b = n - 2, e = n - 1;
    double dn = ((x[e] - x[a]) * (x[e] - x[a]) * (y[e] - y[b]) -
                 (x[e] - x[b]) * (x[e] - x[b]) * (y[e] - y[a])) /
                ((x[e] - x[a]) * (x[e] - x[b]) * (x[b] - x[a]));
    qn = 0.5;
    un = (3.0 / (x[e] - x[b])) * (dn - (y[e] - y[b]) / (x[e] - x[b]));
  }
  dd[n - 1] = (un - qn * u[n - 2]) / (qn * dd[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; --k) dd[k] = dd[k] * dd[k + 1] + u[k];
}
