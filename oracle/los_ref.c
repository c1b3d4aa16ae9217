/*
 * los_ref.c — CPU oracle for SONG's second-order line-of-sight transfer stage.
 *
 * TEST INFRASTRUCTURE. This file is the checker, not the product: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product library (song_amd/) never links, loads or calls it.
 *
 * It restates the reference algorithm in plain C99 + OpenMP, FP64, keeping the
 * loop nest, OpenMP schedule and per-element arithmetic order of
 * source/transfer2.c so that it can also serve as the timed CPU baseline:
 *   transfer2_init            source/transfer2.c:79-621
 *   indices / l list / lm     source/transfer2.c:866-1285
 *   k3 size / list / sizes    source/transfer2.c:1329-1834
 *   compute (T, E, B)         source/transfer2.c:1861-2116
 *   integrate                 source/transfer2.c:2140-2341
 *   time grid                 source/transfer2.c:2366-2585
 *   interpolate in k / time   source/transfer2.c:2612-2763, :2783-2908
 *   store / alloc / free      source/transfer2.c:636-691, :769-838, :2929-3104
 * and the bispectrum module's k3 integral of the transfers (los_ref_k3_integral):
 *   integrate_over_k3 loop    source/bispectra2.c:1182-1410
 *   bessel2_convolution       source/bessel2.c:499-605 (linear, g = NULL)
 * and the power-spectrum module's (k1,k2,k3) sum (los_ref_cl2):
 *   spectra2_cls              source/spectra2.c:241-590 (k3 loop :466-540)
 * The reference cannot be compiled here (its CLASS submodule is empty), so
 * the CLASS helpers it calls are restated from their documented behaviour:
 *   array_spline_table_columns(..., _SPLINE_EST_DERIV_): clamped cubic spline,
 *     end slopes from the parabola through the three end points (Numerical
 *     Recipes 2nd ed. sec. 3.3 solve); pinned by exactness on quadratics and
 *     by scipy's clamped CubicSpline;
 *   size_indexl_indexm / multipole2offset_indexl_indexm /
 *     offset2multipole_indexl_indexm: l-major, m-ascending ragged ordering;
 *     pinned by hand-computed fixtures and the bijection property.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/song_los_types.h"

/* LOS_PLANT = n > 0 builds the oracle with deliberate mistake n, so that the pinning
   tests can show they catch it (1: k3 spline weights a and b swapped; 2: the spline's
   h^2/6 term as h^2/3; 3: the J skip as index_x <= index_xmin_J; 4: the flat k3
   extrapolation fills swapped; 5: linear k3 weights swapped; 6: linear time weights
   swapped; 7: the cubic J weights of ddJ_left and ddJ_right swapped). The oracle proper is
   built with LOS_PLANT = 0. */
#ifndef LOS_PLANT
#define LOS_PLANT 0
#endif

#define R_MIN(a, b) ((a) < (b) ? (a) : (b))
#define R_MAX(a, b) ((a) > (b) ? (a) : (b))

#define r_test(cond, err, ...)                                                  \
  do {                                                                          \
    if (cond) {                                                                 \
      snprintf((err), _ERRORMSGSIZE_, __VA_ARGS__);                             \
      return _FAILURE_;                                                         \
    }                                                                           \
  } while (0)

#define r_alloc(p, size, err)                                                   \
  do {                                                                          \
    (p) = malloc(size);                                                         \
    r_test((p) == NULL && (size) != 0, err, "%s: out of memory", __func__);     \
  } while (0)

/* ---------------------------------------------------------------------------
 * (l,m) ordering helpers (song_tools of the CLASS fork)
 * ------------------------------------------------------------------------- */
static int ref_size_indexl_indexm(const int *lv, int ls, const int *mv, int ms)
{
  int n = 0;
  for (int a = 0; a < ls; ++a)
    for (int b = 0; b < ms; ++b)
      if (mv[b] <= lv[a]) n++;
  return n;
}

static int ref_offset_indexl_indexm(int l, int m, const int *lv, int ls, const int *mv, int ms)
{
  int n = 0;
  for (int a = 0; a < ls; ++a)
    for (int b = 0; b < ms; ++b) {
      if (mv[b] > lv[a]) continue;
      if (lv[a] == l && mv[b] == m) return n;
      n++;
    }
  return -1;
}

static void ref_offset2multipole(int off, const int *lv, int ls, const int *mv, int ms, int *il,
                                 int *im)
{
  int n = 0;
  *il = -1;
  *im = -1;
  for (int a = 0; a < ls; ++a)
    for (int b = 0; b < ms; ++b) {
      if (mv[b] > lv[a]) continue;
      if (n == off) { *il = a; *im = b; return; }
      n++;
    }
}

/* ---------------------------------------------------------------------------
 * Spline with estimated end derivatives, columns contiguous in x.
 * y[c*n + i], dd[c*n + i]; u is scratch of n doubles.
 * ------------------------------------------------------------------------- */
static void ref_spline_column(const double *x, int n, const double *y, double *dd, double *u)
{
  double d0 = ((x[2] - x[0]) * (x[2] - x[0]) * (y[1] - y[0]) -
               (x[1] - x[0]) * (x[1] - x[0]) * (y[2] - y[0])) /
              ((x[2] - x[0]) * (x[1] - x[0]) * (x[2] - x[1]));
  dd[0] = -0.5;
  u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - d0);
  for (int i = 1; i < n - 1; ++i) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double p = sig * dd[i - 1] + 2.0;
    dd[i] = (sig - 1.0) / p;
    double t = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * t / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  const int a = n - 3, b = n - 2, e = n - 1;
  double dn = ((x[e] - x[a]) * (x[e] - x[a]) * (y[e] - y[b]) -
               (x[e] - x[b]) * (x[e] - x[b]) * (y[e] - y[a])) /
              ((x[e] - x[a]) * (x[e] - x[b]) * (x[b] - x[a]));
  double qn = 0.5;
  double un = (3.0 / (x[e] - x[b])) * (dn - (y[e] - y[b]) / (x[e] - x[b]));
  dd[n - 1] = (un - qn * u[n - 2]) / (qn * dd[n - 2] + 1.0);
  for (int k = n - 2; k >= 0; --k) dd[k] = dd[k] * dd[k + 1] + u[k];
}

/* Exported for the oracle's own pinning tests. */
int los_ref_spline_columns(const double *x, int n, const double *y, int ncol, double *dd)
{
  if (n < 3) return _FAILURE_;
  double *u = malloc(n * sizeof(double));
  if (!u) return _FAILURE_;
  for (int c = 0; c < ncol; ++c) ref_spline_column(x, n, y + (long)c * n, dd + (long)c * n, u);
  free(u);
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * k3 step (source/transfer2.c:1428-1452 and :1598-1612)
 * ------------------------------------------------------------------------- */
static int ref_k_step_max(struct precision *ppr, struct precision2 *ppr2,
                          struct bessels2 *pbs2, struct transfers2 *ptr2, double *out)
{
  double k_step_max = 0;
  if (ptr2->k_sampling == class_transfer2_k3_sampling) {
    k_step_max = 2. * M_PI / (ptr2->tau0 - ptr2->tau_rec) * ppr2->q_linstep_song;
    if ((ppr->load_run == _TRUE_) && (ppr2->old_run == _TRUE_))
      k_step_max = 2. * M_PI / ptr2->rs_rec * ppr2->q_linstep_song;
  } else if (ptr2->k_sampling == bessel_k3_sampling) {
    k_step_max = pbs2->xx_step / ptr2->tau0;
  }
  r_test(k_step_max == 0, ptr2->error_message, "stop to avoid infinite loop");
  *out = k_step_max;
  return _SUCCESS_;
}

/* source/transfer2.c:1329-1551 */
static int ref_get_k3_size(struct precision *ppr, struct precision2 *ppr2,
                           struct perturbs2 *ppt2, struct bessels2 *pbs2,
                           struct transfers2 *ptr2, int index_k1, int index_k2)
{
  r_test(index_k2 > index_k1, ptr2->error_message, "we always require k1 >= k2");
  int k_pt_size = ppt2->k3_size[index_k1][index_k2];
  double k_min_pt = ppt2->k3[index_k1][index_k2][0];
  double k_max_pt = ppt2->k3[index_k1][index_k2][k_pt_size - 1];
  double k_min_tr = k_min_pt;
  double k_max_tr = k_max_pt;
  if (ppr->bispectra_k3_extrapolation != no_k3_extrapolation) {
    double physical_k3_range = k_max_pt - k_min_pt;
    double ext_right = 2. * M_PI * ppr->extra_k3_oscillations_right / (ptr2->tau0 - ptr2->tau_rec);
    double ext_left = 2. * M_PI * ppr->extra_k3_oscillations_left / (ptr2->tau0 - ptr2->tau_rec);
    if (physical_k3_range < ext_right) k_max_tr += (ext_right - physical_k3_range);
    k_max_tr = R_MIN(k_max_tr, ppt2->k[ppt2->k_size - 1]);
    if (physical_k3_range < ext_left) k_min_tr -= (ext_left - physical_k3_range);
    k_min_tr = R_MAX(k_min_tr, R_MIN(ppt2->k[0], k_min_pt));
    if (ppr->extra_k3_oscillations_left < 0) k_min_tr = ppt2->k[0];
    if (ppr->extra_k3_oscillations_right < 0) k_max_tr = ppt2->k[ppt2->k_size - 1];
  }
  ptr2->k_min_k1k2[index_k1][index_k2] = k_min_tr;
  ptr2->k_max_k1k2[index_k1][index_k2] = k_max_tr;
  r_test(k_max_tr > pbs2->xx_max / ptr2->tau0, ptr2->error_message,
         "xx_max is not large enough (k_max=%g,tau0=%g,xx_max=%g)", k_max_tr, ptr2->tau0,
         pbs2->xx_max);
  double k_step_max;
  if (ref_k_step_max(ppr, ppr2, pbs2, ptr2, &k_step_max) == _FAILURE_) return _FAILURE_;

  int index_k_pt = 0;
  int index_k_tr = 0;
  double k = k_min_tr;
  index_k_tr++;
  if (ppr->bispectra_k3_extrapolation != no_k3_extrapolation) {
    double k_step = k_step_max;
    r_test(k_step < ppr->smallest_allowed_variation, ptr2->error_message,
           "stopping to avoid segmentation fault, step=%g", k_step);
    while (k < k_min_pt) {
      k += k_step;
      index_k_tr++;
    }
    if (k != k_min_pt) k = k_min_pt;
  }
  index_k_pt++;
  ptr2->k_physical_start_k1k2[index_k1][index_k2] = index_k_tr - 1;
  while ((index_k_pt < k_pt_size) &&
         ((ppt2->k3[index_k1][index_k2][index_k_pt] - k) < k_step_max)) {
    k = ppt2->k3[index_k1][index_k2][index_k_pt];
    index_k_pt++;
    index_k_tr++;
  }
  ptr2->k_physical_size_k1k2[index_k1][index_k2] = index_k_pt;
  while (k < k_max_tr) {
    k += k_step_max;
    index_k_tr++;
    if ((k <= k_max_pt) && (k >= k_min_pt)) ptr2->k_physical_size_k1k2[index_k1][index_k2]++;
  }
  if (k > k_max_tr) ptr2->k_size_k1k2[index_k1][index_k2] = index_k_tr - 1;
  else ptr2->k_size_k1k2[index_k1][index_k2] = index_k_tr;

  r_test(ptr2->k_size_k1k2[index_k1][index_k2] < 2, ptr2->error_message,
         "found less than 2 valid values for the k-grid of modes k1=%f, k2=%f",
         ppt2->k[index_k1], ppt2->k[index_k2]);
  r_test((ptr2->k_physical_size_k1k2[index_k1][index_k2] != ptr2->k_size_k1k2[index_k1][index_k2]) &&
             (ppr->bispectra_k3_extrapolation == no_k3_extrapolation),
         ptr2->error_message, "when no extrapolation is requested, all k's should be physical");
  r_test((ptr2->k_physical_start_k1k2[index_k1][index_k2] != 0) &&
             (ppr->bispectra_k3_extrapolation == no_k3_extrapolation),
         ptr2->error_message, "when no extrapolation is requested, all k's should be physical");
  return _SUCCESS_;
}

/* source/transfer2.c:1566-1737 */
int los_ref_get_k3_list(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                        struct bessels2 *pbs2, struct transfers2 *ptr2, int index_k1,
                        int index_k2, double *k3, int *last_used_index_pt)
{
  int k_pt_size = ppt2->k3_size[index_k1][index_k2];
  double k_min_pt = ppt2->k3[index_k1][index_k2][0];
  int k_tr_size = ptr2->k_size_k1k2[index_k1][index_k2];
  double k_min_tr = ptr2->k_min_k1k2[index_k1][index_k2];
  double k_max_tr = ptr2->k_max_k1k2[index_k1][index_k2];
  double k_step_max;
  if (ref_k_step_max(ppr, ppr2, pbs2, ptr2, &k_step_max) == _FAILURE_) return _FAILURE_;

  int index_k_pt = 0;
  int index_k_tr = 0;
  double k = k_min_tr;
  k3[0] = k_min_tr;
  index_k_tr++;
  if (ppr->bispectra_k3_extrapolation != no_k3_extrapolation) {
    double k_step = k_step_max;
    r_test(k_step < ppr->smallest_allowed_variation, ptr2->error_message,
           "stopping to avoid segmentation fault, step=%g", k_step);
    while (k < k_min_pt) {
      k += k_step;
      k3[index_k_tr] = k;
      index_k_tr++;
    }
    if (k != k_min_pt) {
      k = k_min_pt;
      k3[index_k_tr - 1] = k;
    }
  }
  index_k_pt++;
  while ((index_k_pt < k_pt_size) &&
         ((ppt2->k3[index_k1][index_k2][index_k_pt] - k) < k_step_max)) {
    k = ppt2->k3[index_k1][index_k2][index_k_pt];
    k3[index_k_tr] = k;
    index_k_pt++;
    index_k_tr++;
  }
  *last_used_index_pt = index_k_pt;
  while ((index_k_tr < k_tr_size) && (k < k_max_tr)) {
    k += k_step_max;
    k3[index_k_tr] = k;
    index_k_tr++;
  }
  r_test(k3[k_tr_size - 1] > k_max_tr, ptr2->error_message,
         "bug in k list calculation, k_max_transfers2=%.17f is larger than k_max_tr=%.17f",
         k3[k_tr_size - 1], k_max_tr);
#ifdef DEBUG
  /* DEBUG invariants of the reference (source/transfer2.c:1711-1733) */
  int physical_size = ptr2->k_physical_size_k1k2[index_k1][index_k2];
  int first_physical_index = ptr2->k_physical_start_k1k2[index_k1][index_k2];
  int last_physical_index = first_physical_index + physical_size - 1;
  for (int i = first_physical_index; i <= last_physical_index; ++i)
    r_test((k3[i] <= fabs(ppt2->k[index_k1] - ppt2->k[index_k2])) ||
               (k3[i] >= (ppt2->k[index_k1] + ppt2->k[index_k2])),
           ptr2->error_message, "triangular condition not satisfied (%d,%d)", index_k1, index_k2);
  for (int i = 0; i < k_tr_size - 1; ++i)
    r_test(k3[i] >= k3[i + 1], ptr2->error_message, "k3-grid is not strictly ascending");
#endif
  return _SUCCESS_;
}

/* source/transfer2.c:1752-1834 */
static int ref_get_k3_sizes(struct precision *ppr, struct precision2 *ppr2,
                            struct perturbs2 *ppt2, struct bessels2 *pbs2,
                            struct transfers2 *ptr2)
{
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    for (int i2 = 0; i2 < i1 + 1; ++i2)
      r_test(((ppt2->k3_size[i1][i2] < 2) && (ppr2->sources_k3_interpolation == linear_interpolation)) ||
                 ((ppt2->k3_size[i1][i2] < 4) && (ppr2->sources_k3_interpolation == cubic_interpolation)),
             ptr2->error_message, "index_k1=%d, index_k2=%d: cannot interpolate in k3 with k3_size=%d",
             i1, i2, ppt2->k3_size[i1][i2]);
  int k1_size = ppt2->k_size;
  r_alloc(ptr2->k_size_k1k2, k1_size * sizeof(int *), ptr2->error_message);
  r_alloc(ptr2->k_physical_start_k1k2, k1_size * sizeof(int *), ptr2->error_message);
  r_alloc(ptr2->k_physical_size_k1k2, k1_size * sizeof(int *), ptr2->error_message);
  r_alloc(ptr2->k_min_k1k2, k1_size * sizeof(double *), ptr2->error_message);
  r_alloc(ptr2->k_max_k1k2, k1_size * sizeof(double *), ptr2->error_message);
  for (int i1 = 0; i1 < k1_size; ++i1) {
    int k2_size = i1 + 1;
    r_alloc(ptr2->k_size_k1k2[i1], k2_size * sizeof(int), ptr2->error_message);
    r_alloc(ptr2->k_physical_start_k1k2[i1], k2_size * sizeof(int), ptr2->error_message);
    r_alloc(ptr2->k_physical_size_k1k2[i1], k2_size * sizeof(int), ptr2->error_message);
    r_alloc(ptr2->k_min_k1k2[i1], k2_size * sizeof(double), ptr2->error_message);
    r_alloc(ptr2->k_max_k1k2[i1], k2_size * sizeof(double), ptr2->error_message);
    for (int i2 = 0; i2 <= i1; ++i2)
      if (ref_get_k3_size(ppr, ppr2, ppt2, pbs2, ptr2, i1, i2) == _FAILURE_) return _FAILURE_;
  }
  ptr2->k3_size_max = 0;
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2)
      ptr2->k3_size_max = R_MAX(ptr2->k3_size_max, ptr2->k_size_k1k2[i1][i2]);
  return _SUCCESS_;
}

/* source/transfer2.c:1121-1285 */
static int ref_get_lm_lists(struct precision2 *ppr2, struct perturbs2 *ppt2,
                            struct transfers2 *ptr2)
{
  r_alloc(ptr2->lm_array, ptr2->l_size * sizeof(int *), ptr2->error_message);
  for (int il = 0; il < ptr2->l_size; ++il) {
    int l = ptr2->l[il];
    ptr2->lm_array[il] = calloc(ppr2->index_m_max[l] + 1 > 0 ? ppr2->index_m_max[l] + 1 : 1, sizeof(int));
    r_test(ptr2->lm_array[il] == NULL, ptr2->error_message, "out of memory");
    for (int im = 0; im <= ppr2->index_m_max[l]; ++im)
      ptr2->lm_array[il][im] = ref_offset_indexl_indexm(ptr2->l[il], ptr2->m[im], ptr2->l,
                                                        ptr2->l_size, ptr2->m, ptr2->m_size);
  }
  r_alloc(ptr2->corresponding_index_l, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  r_alloc(ptr2->corresponding_index_m, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  r_alloc(ptr2->index_tt2_monopole, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  r_alloc(ptr2->index_pt2_monopole, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  for (int tt = 0; tt < ptr2->tt2_size; tt++) {
    int il = -1, im = -1;
    const char *lab = NULL;
    if (ppt2->has_source_T && tt >= ptr2->index_tt2_T && tt < ptr2->index_tt2_T + ptr2->n_transfers) {
      ptr2->index_pt2_monopole[tt] = ppt2->index_tp2_T;
      ptr2->index_tt2_monopole[tt] = ptr2->index_tt2_T;
      lab = "T";
    } else if (ppt2->has_source_E && tt >= ptr2->index_tt2_E &&
               tt < ptr2->index_tt2_E + ptr2->n_transfers) {
      ptr2->index_pt2_monopole[tt] = ppt2->index_tp2_E;
      ptr2->index_tt2_monopole[tt] = ptr2->index_tt2_E;
      lab = "E";
    } else if (ppt2->has_source_B && tt >= ptr2->index_tt2_B &&
               tt < ptr2->index_tt2_B + ptr2->n_transfers) {
      ptr2->index_pt2_monopole[tt] = ppt2->index_tp2_B;
      ptr2->index_tt2_monopole[tt] = ptr2->index_tt2_B;
      lab = "B";
    }
    if (lab) {
      int off = tt - ptr2->index_tt2_monopole[tt];
      ref_offset2multipole(off, ptr2->l, ptr2->l_size, ptr2->m, ptr2->m_size, &il, &im);
      if (il >= 0)
        snprintf(ptr2->tt2_labels[tt], _MAX_LENGTH_LABEL_, "%s_%d_%d", lab, ptr2->l[il], ptr2->m[im]);
    }
    r_test((il >= ptr2->l_size) || (im >= ptr2->m_size) || (il < 0) || (im < 0),
           ptr2->error_message, "index_tt=%d: (index_l,index_m)=(%d,%d) out of bounds", tt, il, im);
    ptr2->corresponding_index_l[tt] = il;
    ptr2->corresponding_index_m[tt] = im;
  }
  return _SUCCESS_;
}

/* source/transfer2.c:866-1067 */
static int ref_indices_of_transfers(struct precision *ppr, struct precision2 *ppr2,
                                    struct perturbs2 *ppt2, struct bessels *pbs,
                                    struct bessels2 *pbs2, struct transfers2 *ptr2)
{
  ptr2->l_size = pbs->l_size;
  r_alloc(ptr2->l, ptr2->l_size * sizeof(int), ptr2->error_message);
  for (int il = 0; il < ptr2->l_size; il++) ptr2->l[il] = pbs->l[il];
  ptr2->m_size = ppr2->m_size;
  r_alloc(ptr2->m, ptr2->m_size * sizeof(int), ptr2->error_message);
  for (int im = 0; im < ptr2->m_size; ++im) ptr2->m[im] = ppr2->m[im];

  int index_tt = 0;
  ptr2->n_transfers = ref_size_indexl_indexm(ptr2->l, ptr2->l_size, ppr2->m, ppr2->m_size);
  ptr2->n_nonzero_transfers_E = ptr2->n_transfers;
  if (ppr2->compute_m[0]) ptr2->n_nonzero_transfers_E -= 2;
  if (ppr2->compute_m[1]) ptr2->n_nonzero_transfers_E -= 1;
  ptr2->n_nonzero_transfers_B = ptr2->n_transfers;
  if (ppr2->compute_m[0]) ptr2->n_nonzero_transfers_B -= ptr2->l_size;
  if (ppr2->compute_m[1]) ptr2->n_nonzero_transfers_B -= 1;
  if (ppt2->has_source_T) { ptr2->index_tt2_T = index_tt; index_tt += ptr2->n_transfers; }
  if (ppt2->has_source_E) { ptr2->index_tt2_E = index_tt; index_tt += ptr2->n_transfers; }
  if (ppt2->has_source_B) { ptr2->index_tt2_B = index_tt; index_tt += ptr2->n_transfers; }
  ptr2->tt2_size = index_tt;
  ptr2->tt2_labels = calloc((size_t)ptr2->tt2_size * _MAX_LENGTH_LABEL_, sizeof(char));
  r_test(ptr2->tt2_labels == NULL, ptr2->error_message, "out of memory");
  if (ref_get_lm_lists(ppr2, ppt2, ptr2) == _FAILURE_) return _FAILURE_;
  if (ref_get_k3_sizes(ppr, ppr2, ppt2, pbs2, ptr2) == _FAILURE_) return _FAILURE_;
  ptr2->count_memorised_transfers = 0;
  ptr2->count_allocated_transfers = 0;
  r_alloc(ptr2->transfer, ptr2->tt2_size * sizeof(double ***), ptr2->error_message);
  for (int tt = 0; tt < ptr2->tt2_size; ++tt)
    r_alloc(ptr2->transfer[tt], ppt2->k_size * sizeof(double **), ptr2->error_message);
  ptr2->has_allocated_transfers = calloc(ppt2->k_size, sizeof(short));
  r_test(ptr2->has_allocated_transfers == NULL, ptr2->error_message, "out of memory");
  if (ppr2->store_transfers_to_disk || ppr2->load_transfers_from_disk) {
    r_alloc(ptr2->transfers_files, ptr2->tt2_size * sizeof(FILE *), ptr2->error_message);
    r_alloc(ptr2->transfers_paths, ptr2->tt2_size * sizeof(char *), ptr2->error_message);
    for (int tt = 0; tt < ptr2->tt2_size; ++tt) {
      r_alloc(ptr2->transfers_paths[tt], _FILENAMESIZE_, ptr2->error_message);
      snprintf(ptr2->transfers_paths[tt], _FILENAMESIZE_, "%s/transfers_%03d.dat",
               ptr2->transfers_dir, tt);
      if (ppr2->store_transfers_to_disk) {
        ptr2->transfers_files[tt] = fopen(ptr2->transfers_paths[tt], "wb");
        r_test(ptr2->transfers_files[tt] == NULL, ptr2->error_message, "could not open %s",
               ptr2->transfers_paths[tt]);
      }
    }
  }
  return _SUCCESS_;
}

/* source/transfer2.c:2976-3032 */
static int ref_allocate_k1_level(struct perturbs2 *ppt2, struct transfers2 *ptr2, int i1)
{
  r_test(ptr2->has_allocated_transfers[i1] == _TRUE_, ptr2->error_message,
         "the index_k1=%d level of ptr2->transfers is already allocated", i1);
  for (int tt = 0; tt < ptr2->tt2_size; ++tt) {
    r_alloc(ptr2->transfer[tt][i1], (i1 + 1) * sizeof(double *), ptr2->error_message);
    for (int i2 = 0; i2 <= i1; ++i2) {
      r_alloc(ptr2->transfer[tt][i1][i2], ptr2->k_size_k1k2[i1][i2] * sizeof(double),
              ptr2->error_message);
      ptr2->count_allocated_transfers += ptr2->k_size_k1k2[i1][i2];
    }
  }
  ptr2->has_allocated_transfers[i1] = _TRUE_;
  return _SUCCESS_;
}

static int ref_free_k1_level(struct perturbs2 *ppt2, struct transfers2 *ptr2, int i1)
{
  r_test(ptr2->has_allocated_transfers[i1] == _FALSE_, ptr2->error_message,
         "the index_k1=%d level of ptr2->transfers is already free", i1);
  for (int tt = 0; tt < ptr2->tt2_size; tt++) {
    for (int i2 = 0; i2 <= i1; ++i2) free(ptr2->transfer[tt][i1][i2]);
    free(ptr2->transfer[tt][i1]);
  }
  ptr2->has_allocated_transfers[i1] = _FALSE_;
  return _SUCCESS_;
}

/* perturbations2.c:609-636: the stage frees each k1 level of the sources. */
static int ref_free_sources_k1(struct perturbs2 *ppt2, int i1)
{
  r_test(ppt2->has_allocated_sources[i1] == _FALSE_, ppt2->error_message,
         "the index_k1=%d level of ppt2->sources is already free", i1);
  for (int tp = 0; tp < ppt2->tp2_size; tp++) {
    for (int i2 = 0; i2 <= i1; i2++) free(ppt2->sources[tp][i1][i2]);
    free(ppt2->sources[tp][i1]);
    ppt2->sources[tp][i1] = NULL;
  }
  ppt2->has_allocated_sources[i1] = _FALSE_;
  return _SUCCESS_;
}

/* source/transfer2.c:2612-2763 */
static int ref_interpolate_sources_in_k(struct precision *ppr, struct precision2 *ppr2,
                                        struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                        int i1, int i2, int tp, double *k_grid,
                                        double *spl, double *out)
{
  int k_pt_size = ppt2->k3_size[i1][i2];
  double *k_pt = ppt2->k3[i1][i2];
  int k_tr_size = ptr2->k_size_k1k2[i1][i2];
  double *k_tr = k_grid;
  const double *S = ppt2->sources[tp][i1][i2];
  const int T = ppt2->tau_size;
  if (ppr2->sources_k3_interpolation == cubic_interpolation) {
    r_test(los_ref_spline_columns(k_pt, k_pt_size, S, T, spl) == _FAILURE_, ptr2->error_message,
           "spline failed");
  }
  int physical_size = ptr2->k_physical_size_k1k2[i1][i2];
  int first = ptr2->k_physical_start_k1k2[i1][i2];
  int last = first + physical_size - 1;
  int ik = 0;
  double h = k_pt[ik + 1] - k_pt[ik];
  for (int it = first; it <= last; ++it) {
    while (((ik + 1) < k_pt_size) && (k_pt[ik + 1] < k_tr[it])) {
      ik++;
      h = k_pt[ik + 1] - k_pt[ik];
    }
    r_test(h == 0, ptr2->error_message, "stop to avoid division by zero");
    double b = (k_tr[it] - k_pt[ik]) / h;
    double a = 1 - b;
#if LOS_PLANT == 1 || LOS_PLANT == 5
    if ((LOS_PLANT == 1) == (ppr2->sources_k3_interpolation == cubic_interpolation)) {
      const double w = a;
      a = b;
      b = w;
    }
#endif
    if (ppr2->sources_k3_interpolation == linear_interpolation) {
      for (int t = 0; t < T; t++)
        out[it * T + t] = a * S[t * k_pt_size + ik] + b * S[t * k_pt_size + ik + 1];
    } else {
      for (int t = 0; t < T; t++)
        out[it * T + t] = a * S[t * k_pt_size + ik] + b * S[t * k_pt_size + ik + 1] +
                          ((a * a * a - a) * spl[t * k_pt_size + ik] +
                           (b * b * b - b) * spl[t * k_pt_size + ik + 1]) * h * h /
                              (LOS_PLANT == 2 ? 3.0 : 6.0);
    }
  }
  if (ppr->bispectra_k3_extrapolation != no_k3_extrapolation) {
    for (int it = last + 1; it < k_tr_size; ++it)
      if (ppr->bispectra_k3_extrapolation == flat_k3_extrapolation)
        for (int t = 0; t < T; t++) out[it * T + t] = out[(LOS_PLANT == 4 ? first : last) * T + t];
    for (int it = 0; it < first; ++it)
      if (ppr->bispectra_k3_extrapolation == flat_k3_extrapolation)
        for (int t = 0; t < T; t++) out[it * T + t] = out[(LOS_PLANT == 4 ? last : first) * T + t];
  }
  return _SUCCESS_;
}

/* source/transfer2.c:2366-2585 */
static int ref_get_time_grid(struct precision2 *ppr2, struct perturbs2 *ppt2,
                             struct bessels2 *pbs2, struct transfers2 *ptr2,
                             struct transfer2_workspace *pw)
{
  double k = pw->k;
  int tau_size_pt = ppt2->tau_size;
  double tau_min_pt = ppt2->tau_sampling[0];
  double tau_max_pt = ppt2->tau_sampling[tau_size_pt - 1];
  double tau_step_max = 0;
  int it;
  if (ptr2->tau_sampling == sources_tau_sampling) {
    pw->tau_grid_size = tau_size_pt;
    for (int i = 0; i < pw->tau_grid_size; ++i) pw->tau_grid[i] = ppt2->tau_sampling[i];
    goto fill_delta_tau;
  } else if (ptr2->tau_sampling == custom_tau_sampling) {
    tau_step_max = 2 * M_PI / k * ppr2->tau_linstep_song;
  } else if (ptr2->tau_sampling == bessel_tau_sampling) {
    tau_step_max = pbs2->xx_step / k;
  }
  r_test(tau_step_max == 0, ptr2->error_message, "stopping to prevent segmentation fault");
  {
    int itr = 0, ipt = 0;
    double tau = tau_min_pt;
    itr++;
    ipt++;
    while (ipt < tau_size_pt) {
      while ((ppt2->tau_sampling[ipt] - tau) > tau_step_max) {
        tau += tau_step_max;
        itr++;
      }
      tau = ppt2->tau_sampling[ipt];
      ++itr;
      ++ipt;
    }
    pw->tau_grid_size = itr;
    itr = 0;
    ipt = 0;
    tau = tau_min_pt;
    pw->tau_grid[0] = tau;
    itr++;
    ipt++;
    pw->index_tau_left[0] = 0;
    while (ipt < tau_size_pt) {
      while ((ppt2->tau_sampling[ipt] - tau) > tau_step_max) {
        tau += tau_step_max;
        pw->tau_grid[itr] = tau;
        pw->index_tau_left[itr] = ipt;
        itr++;
      }
      tau = ppt2->tau_sampling[ipt];
      pw->tau_grid[itr] = tau;
      pw->index_tau_left[itr] = ipt;
      itr++;
      ipt++;
    }
    for (itr = 0; itr < pw->tau_grid_size; ++itr)
      if (pw->index_tau_left[itr] == (ppt2->tau_size - 1)) pw->index_tau_left[itr] = ppt2->tau_size - 2;
#ifdef DEBUG /* source/transfer2.c:2541-2559 */
    for (itr = 0; itr < pw->tau_grid_size - 1; ++itr)
      r_test(pw->tau_grid[itr] >= pw->tau_grid[itr + 1], ptr2->error_message,
             "integration grid is not sorted");
    r_test(pw->tau_grid_size < ppt2->tau_size, ptr2->error_message, "integration grid too coarse");
    r_test((pw->tau_grid[0] != tau_min_pt) || (pw->tau_grid[pw->tau_grid_size - 1] != tau_max_pt),
           ptr2->error_message, "limits of the time grid differ from the sources'");
#endif
  }
fill_delta_tau:
  for (it = 0; it < pw->tau_grid_size; ++it) pw->tau0_minus_tau[it] = ptr2->tau0 - pw->tau_grid[it];
  pw->delta_tau[0] = pw->tau_grid[1] - pw->tau_grid[0];
  for (it = 1; it < pw->tau_grid_size - 1; ++it)
    pw->delta_tau[it] = pw->tau_grid[it + 1] - pw->tau_grid[it - 1];
  pw->delta_tau[pw->tau_grid_size - 1] =
      pw->tau_grid[pw->tau_grid_size - 1] - pw->tau_grid[pw->tau_grid_size - 2];
  return _SUCCESS_;
}

/* source/transfer2.c:2783-2908 */
static int ref_interpolate_sources_in_time(struct precision2 *ppr2, struct perturbs2 *ppt2,
                                           struct transfers2 *ptr2, double *in_k,
                                           double *spline, double *out,
                                           struct transfer2_workspace *pw, double *scratch)
{
  int tau_size_pt = ppt2->tau_size;
  double *tau_pt = ppt2->tau_sampling;
  if (ptr2->tau_sampling == sources_tau_sampling) {
    for (int i = 0; i < pw->tau_grid_size; ++i) out[i] = in_k[pw->index_k * tau_size_pt + i];
    return _SUCCESS_;
  }
  const double *col = in_k + (long)pw->index_k * tau_size_pt;
  if (ppr2->sources_time_interpolation == cubic_interpolation) {
    (void)scratch;
    r_test(los_ref_spline_columns(tau_pt, tau_size_pt, col, 1, spline) == _FAILURE_,
           ptr2->error_message, "spline failed");
  }
  for (int itr = 0; itr < pw->tau_grid_size; ++itr) {
    int i = pw->index_tau_left[itr];
    double h = tau_pt[i + 1] - tau_pt[i];
    r_test(h == 0., ptr2->error_message, "stop to avoid division by zero");
    double b = (pw->tau_grid[itr] - tau_pt[i]) / h;
    double a = 1 - b;
#if LOS_PLANT == 6
    {
      const double w = a;
      a = b;
      b = w;
    }
#endif
    if (ppr2->sources_time_interpolation == linear_interpolation)
      out[itr] = a * col[i] + b * col[i + 1];
    else
      out[itr] = a * col[i] + b * col[i + 1] +
                 ((a * a * a - a) * spline[i] + (b * b * b - b) * spline[i + 1]) * h * h / 6.0;
  }
  return _SUCCESS_;
}

/* source/transfer2.c:2140-2341 */
static int ref_integrate(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                         struct bessels2 *pbs2, struct transfers2 *ptr2, int index_l, int index_m,
                         double **src, int index_J, int index_source_monopole,
                         struct transfer2_workspace *pw, double *integral)
{
  int l = ptr2->l[index_l];
  int m = ptr2->m[index_m];
  double k = pw->k;
  *integral = 0;
#ifdef DEBUG /* source/transfer2.c:2168-2177 */
  r_test(k * pw->tau0_minus_tau[0] > pbs2->xx_max, ptr2->error_message, "not enough J's");
#endif
  for (int it = 0; it < pw->tau_grid_size; ++it) {
    double x = k * pw->tau0_minus_tau[it];
    int index_x = (int)(x / pbs2->xx_step);
    double a_J = (pbs2->xx[index_x + 1] - x) / pbs2->xx_step;
    double integrand = 0;
    for (int iL = 0; iL <= pw->L_max; ++iL) {
      int L = pbs2->L[iL];
      if (abs(m) > R_MIN(L, l)) continue;
      int index_x_min = pbs2->index_xmin_J[index_J][iL][index_l][index_m];
      if (LOS_PLANT == 3 ? index_x <= index_x_min : index_x < index_x_min) continue;
      int ixJ = index_x - index_x_min;
#ifdef DEBUG
      /* source/transfer2.c:2239-2246, which tests ixJ >= x_size; J[ixJ + 1] is read
         as well, so the bound here also covers that read */
      int x_size = pbs2->x_size_J[index_J][iL][index_l][index_m];
      r_test((ixJ + 1 >= x_size) || (ixJ < 0), ptr2->error_message,
             "L=%d, l=%d, m=%d, x=%g: wrong indexing of J (%d of %d)", L, l, m, x, ixJ, x_size);
#endif
      double J_Llm;
      double *J = &(pbs2->J_Llm_x[index_J][iL][index_l][index_m][ixJ]);
      double Jl = *J;
      double Jr = *(J + 1);
      if (ppr->bessels_interpolation == linear_interpolation) {
        J_Llm = a_J * Jl + (1 - a_J) * Jr;
      } else {
        double *ddJ = &(pbs2->ddJ_Llm_x[index_J][iL][index_l][index_m][ixJ]);
        double ddl = *ddJ;
        double ddr = *(ddJ + 1);
#if LOS_PLANT == 7
        J_Llm = (a_J * Jl + (1. - a_J) * (Jr - a_J * ((2. - a_J) * ddl + (a_J + 1.) * ddr) *
                                                     pbs2->xx_step * pbs2->xx_step / 6.0));
#else
        J_Llm = (a_J * Jl + (1. - a_J) * (Jr - a_J * ((a_J + 1.) * ddl + (2. - a_J) * ddr) *
                                                     pbs2->xx_step * pbs2->xx_step / 6.0));
#endif
      }
      double source_Lm = src[index_source_monopole + ppt2->lm_array[L][ppr2->index_m[m]]][it];
      integrand += J_Llm * source_Lm;
    }
    *integral += integrand * pw->delta_tau[it];
  }
  *integral *= 0.5;
#ifdef DEBUG /* source/transfer2.c:2332-2337 */
  r_test(isnan(*integral), ptr2->error_message, "found nan in second-order transfer function");
#endif
  return _SUCCESS_;
}

/* source/transfer2.c:1861-2116 */
static int ref_compute(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                       struct bessels2 *pbs2, struct transfers2 *ptr2, int i1, int i2, int ik,
                       int index_l, int index_m, int tt, double **src,
                       struct transfer2_workspace *pw)
{
  int transfer_type = ptr2->index_tt2_monopole[tt];
  if (ppt2->has_source_T && transfer_type == ptr2->index_tt2_T) {
    pw->L_max = ppr2->l_max_los_t;
    if (ref_integrate(ppr, ppr2, ppt2, pbs2, ptr2, index_l, index_m, src, pbs2->index_J_TT,
                      ppt2->index_tp2_T, pw, &(pw->transfer)) == _FAILURE_)
      return _FAILURE_;
  } else if (ppt2->has_source_E && transfer_type == ptr2->index_tt2_E) {
    pw->L_max = ppr2->l_max_los_p;
    double direct = 0, mixing = 0;
    if (ref_integrate(ppr, ppr2, ppt2, pbs2, ptr2, index_l, index_m, src, pbs2->index_J_EE,
                      ppt2->index_tp2_E, pw, &direct) == _FAILURE_)
      return _FAILURE_;
    if (ptr2->m[index_m] != 0)
      if (ref_integrate(ppr, ppr2, ppt2, pbs2, ptr2, index_l, index_m, src, pbs2->index_J_EB,
                        ppt2->index_tp2_B, pw, &mixing) == _FAILURE_)
        return _FAILURE_;
    pw->transfer = direct + mixing;
  } else if (ppt2->has_source_B && transfer_type == ptr2->index_tt2_B) {
    pw->L_max = ppr2->l_max_los_p;
    double direct = 0, mixing = 0;
    if (ref_integrate(ppr, ppr2, ppt2, pbs2, ptr2, index_l, index_m, src, pbs2->index_J_EE,
                      ppt2->index_tp2_B, pw, &direct) == _FAILURE_)
      return _FAILURE_;
    if (ref_integrate(ppr, ppr2, ppt2, pbs2, ptr2, index_l, index_m, src, pbs2->index_J_EB,
                      ppt2->index_tp2_E, pw, &mixing) == _FAILURE_)
      return _FAILURE_;
    pw->transfer = direct - mixing;
  }
  pw->transfer /= 4.;
  pw->transfer /= (2. * ptr2->l[index_l] + 1);
  pw->transfer /= 2.;
  ptr2->transfer[tt][i1][i2][ik] = pw->transfer;
#pragma omp atomic
  ++ptr2->count_memorised_transfers;
  return _SUCCESS_;
}

/* source/transfer2.c:2929-2963 */
static int ref_store_k1(struct perturbs2 *ppt2, struct transfers2 *ptr2, int i1)
{
  for (int tt = 0; tt < ptr2->tt2_size; tt++)
    for (int i2 = 0; i2 <= i1; ++i2)
      fwrite(ptr2->transfer[tt][i1][i2], sizeof(double), ptr2->k_size_k1k2[i1][i2],
             ptr2->transfers_files[tt]);
  return _SUCCESS_;
}

/*
 * Entry point: the oracle's transfer2_init. Same arguments and side effects as
 * the reference (it frees ppt2->sources per k1).
 */
int los_ref_transfer2_init(struct precision *ppr, struct precision2 *ppr2,
                           struct background *pba, struct thermo *pth, struct perturbs *ppt,
                           struct perturbs2 *ppt2, struct bessels *pbs, struct bessels2 *pbs2,
                           struct transfers *ptr, struct transfers2 *ptr2)
{
  (void)ptr;
  if (!ppt2->has_perturbations2 ||
      (!ppt2->has_cmb_spectra && !ppt2->has_cmb_bispectra && !ptr2->stop_at_transfers2)) {
    ptr2->has_cls = _FALSE_;
    return _SUCCESS_;
  }
  ptr2->has_cls = _TRUE_;
  ptr2->tau0 = pba->conformal_age;
  ptr2->tau_rec = pth->tau_rec;
  ptr2->rs_rec = pth->rs_rec;
  if (ref_indices_of_transfers(ppr, ppr2, ppt2, pbs, pbs2, ptr2) == _FAILURE_) return _FAILURE_;
  if (ppr2->load_transfers_from_disk == _TRUE_) return _SUCCESS_;
  r_test(ppr2->load_sources_from_disk || ppr2->store_sources_to_disk, ptr2->error_message,
         "the oracle reads sources from memory only");

  int number_of_threads = 1;
#ifdef _OPENMP
#pragma omp parallel
#pragma omp master
  number_of_threads = omp_get_num_threads();
#endif
  int tau_size_max = ppt2->tau_size;
  if (ptr2->tau_sampling == custom_tau_sampling) {
    double k_max = ptr2->k_max_k1k2[ppt2->k_size - 1][ppt2->k_size - 1];
    double tau_step_min = 2 * M_PI / k_max * ppr2->tau_linstep_song;
    double tau_max = ppt2->tau_sampling[ppt2->tau_size - 1];
    tau_size_max = ppt2->tau_size + ceil(tau_max / tau_step_min) + 1;
  } else if (ptr2->tau_sampling == bessel_tau_sampling) {
    tau_size_max = ppt2->tau_size + pbs2->xx_size;
  }
  struct transfer2_workspace **ppw = calloc(number_of_threads, sizeof *ppw);
  r_test(ppw == NULL, ptr2->error_message, "out of memory");
  for (int th = 0; th < number_of_threads; ++th) {
    struct transfer2_workspace *w = calloc(1, sizeof *w);
    r_test(w == NULL, ptr2->error_message, "out of memory");
    w->k_grid = malloc(ptr2->k3_size_max * sizeof(double));
    w->tau_grid = malloc(tau_size_max * sizeof(double));
    w->tau0_minus_tau = malloc(tau_size_max * sizeof(double));
    w->delta_tau = malloc(tau_size_max * sizeof(double));
    w->index_tau_left = malloc(tau_size_max * sizeof(int));
    w->sources_time_spline = malloc(ppt2->tp2_size * sizeof(double *));
    w->interpolated_sources_in_time = malloc(ppt2->tp2_size * sizeof(double *));
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      w->sources_time_spline[tp] = malloc(R_MAX(tau_size_max, ppt2->tau_size) * sizeof(double));
      w->interpolated_sources_in_time[tp] = malloc(tau_size_max * sizeof(double));
    }
    ppw[th] = w;
  }
  double **spl = malloc(ppt2->tp2_size * sizeof(double *));
  double **isk = malloc(ppt2->tp2_size * sizeof(double *));

  for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
    if (ref_allocate_k1_level(ppt2, ptr2, i1) == _FAILURE_) return _FAILURE_;
    for (int i2 = 0; i2 <= i1; ++i2) {
      int last_used;
      double *kg = malloc(ptr2->k3_size_max * sizeof(double));
      if (los_ref_get_k3_list(ppr, ppr2, ppt2, pbs2, ptr2, i1, i2, kg, &last_used) == _FAILURE_)
        return _FAILURE_;
      for (int th = 0; th < number_of_threads; ++th)
        for (int ik = 0; ik < ptr2->k_size_k1k2[i1][i2]; ++ik) ppw[th]->k_grid[ik] = kg[ik];
      free(kg);
      for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
        spl[tp] = malloc((size_t)ppt2->k3_size[i1][i2] * ppt2->tau_size * sizeof(double));
        isk[tp] = malloc((size_t)ptr2->k_size_k1k2[i1][i2] * ppt2->tau_size * sizeof(double));
        if (ref_interpolate_sources_in_k(ppr, ppr2, ppt2, ptr2, i1, i2, tp, ppw[0]->k_grid,
                                         spl[tp], isk[tp]) == _FAILURE_)
          return _FAILURE_;
      }
      int abort_flag = _FALSE_;
#pragma omp parallel shared(ppw, abort_flag)
      {
        int thread = 0;
#ifdef _OPENMP
        thread = omp_get_thread_num();
#endif
        struct transfer2_workspace *pw = ppw[thread];
        pw->thread = thread;
        pw->index_k1 = i1;
        pw->k1 = ppt2->k[i1];
        pw->index_k2 = i2;
        pw->k2 = ppt2->k[i2];
#pragma omp for schedule(static)
        for (int ik = 0; ik < ptr2->k_size_k1k2[i1][i2]; ++ik) {
          if (abort_flag) continue;
          pw->index_k = ik;
          pw->k = pw->k_grid[ik];
          if (ref_get_time_grid(ppr2, ppt2, pbs2, ptr2, pw) == _FAILURE_) { abort_flag = _TRUE_; continue; }
          for (int tp = 0; tp < ppt2->tp2_size; ++tp)
            if (ref_interpolate_sources_in_time(ppr2, ppt2, ptr2, isk[tp], pw->sources_time_spline[tp],
                                                pw->interpolated_sources_in_time[tp], pw, NULL) == _FAILURE_)
              abort_flag = _TRUE_;
          for (int tt = 0; tt < ptr2->tt2_size && !abort_flag; tt++)
            if (ref_compute(ppr, ppr2, ppt2, pbs2, ptr2, i1, i2, ik, ptr2->corresponding_index_l[tt],
                            ptr2->corresponding_index_m[tt], tt, pw->interpolated_sources_in_time,
                            pw) == _FAILURE_)
              abort_flag = _TRUE_;
        }
      }
      if (abort_flag) return _FAILURE_;
      for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
        free(spl[tp]);
        free(isk[tp]);
      }
    }
    if (ref_free_sources_k1(ppt2, i1) == _FAILURE_) {
      memcpy(ptr2->error_message, ppt2->error_message, _ERRORMSGSIZE_);
      return _FAILURE_;
    }
    if (ppr2->store_transfers_to_disk == _TRUE_) {
      ref_store_k1(ppt2, ptr2, i1);
      if (ref_free_k1_level(ppt2, ptr2, i1) == _FAILURE_) return _FAILURE_;
    }
  }
  free(spl);
  free(isk);
  for (int th = 0; th < number_of_threads; ++th) {
    struct transfer2_workspace *w = ppw[th];
    free(w->k_grid);
    free(w->tau_grid);
    free(w->tau0_minus_tau);
    free(w->delta_tau);
    free(w->index_tau_left);
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      free(w->sources_time_spline[tp]);
      free(w->interpolated_sources_in_time[tp]);
    }
    free(w->sources_time_spline);
    free(w->interpolated_sources_in_time);
    free(w);
  }
  free(ppw);
  if (ppr2->store_transfers_to_disk == _TRUE_)
    for (int tt = 0; tt < ptr2->tt2_size; tt++) fclose(ptr2->transfers_files[tt]);
  if (ppr2->load_transfers_from_disk == _FALSE_)
    r_test(ptr2->count_allocated_transfers != ptr2->count_memorised_transfers, ptr2->error_message,
           "mismatch between allocated (%ld) and used (%ld) space", ptr2->count_allocated_transfers,
           ptr2->count_memorised_transfers);
  if (ptr2->stop_at_transfers2 == _TRUE_) {
    ppt->has_cls = _FALSE_;
    ppt->has_cmb_bispectra = _FALSE_;
    ppt2->has_cmb_spectra = _FALSE_;
    ppt2->has_cmb_bispectra = _FALSE_;
  }
  return _SUCCESS_;
}

/* source/transfer2.c:769-838 */
int los_ref_transfer2_free(struct precision2 *ppr2, struct perturbs2 *ppt2, struct transfers2 *ptr2)
{
  if (ptr2->has_cls != _TRUE_) return _SUCCESS_;
  free(ptr2->tt2_labels);
  if (!ppr2->store_transfers_to_disk && !ppr2->load_transfers_from_disk) {
    for (int i1 = 0; i1 < ppt2->k_size; ++i1)
      if (ptr2->has_allocated_transfers[i1]) ref_free_k1_level(ppt2, ptr2, i1);
    for (int tt = 0; tt < ptr2->tt2_size; ++tt) free(ptr2->transfer[tt]);
  }
  free(ptr2->transfer);
  free(ptr2->has_allocated_transfers);
  for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
    free(ptr2->k_size_k1k2[i1]);
    free(ptr2->k_physical_start_k1k2[i1]);
    free(ptr2->k_physical_size_k1k2[i1]);
    free(ptr2->k_min_k1k2[i1]);
    free(ptr2->k_max_k1k2[i1]);
  }
  free(ptr2->k_size_k1k2);
  free(ptr2->k_physical_start_k1k2);
  free(ptr2->k_physical_size_k1k2);
  free(ptr2->k_min_k1k2);
  free(ptr2->k_max_k1k2);
  free(ptr2->l);
  free(ptr2->m);
  free(ptr2->corresponding_index_l);
  free(ptr2->corresponding_index_m);
  free(ptr2->index_tt2_monopole);
  free(ptr2->index_pt2_monopole);
  for (int il = 0; il < ptr2->l_size; ++il) free(ptr2->lm_array[il]);
  free(ptr2->lm_array);
  if (ppr2->store_transfers_to_disk || ppr2->load_transfers_from_disk) {
    for (int tt = 0; tt < ptr2->tt2_size; ++tt) free(ptr2->transfers_paths[tt]);
    free(ptr2->transfers_files);
    free(ptr2->transfers_paths);
  }
  return _SUCCESS_;
}

/*
 * Bounded-sample timing entry for the CPU baseline: the oracle's bookkeeping
 * (its own transfers2, filled here) and then, for each listed pair, the
 * reference loop body (k3 list, k3 interpolation, OpenMP static loop over k3,
 * compute for every tt). The results of each pair, [tt][k3], are appended to
 * `out` in the order of the list when `out` is not NULL (room for *n_out
 * values is the caller's business: sum over the pairs of tt2_size x k3 count).
 * Returns the number of elements computed in *n_out. Does not free sources.
 */
long los_ref_sizeof_transfers2(void) { return (long)sizeof(struct transfers2); }

/* The oracle's bookkeeping alone (l, m lists, k3 sizes), for fixture tests. */
int los_ref_indices(struct precision *ppr, struct precision2 *ppr2, struct background *pba,
                    struct thermo *pth, struct perturbs2 *ppt2, struct bessels *pbs,
                    struct bessels2 *pbs2, struct transfers2 *ptr2)
{
  ptr2->has_cls = _TRUE_;
  ptr2->tau0 = pba->conformal_age;
  ptr2->tau_rec = pth->tau_rec;
  ptr2->rs_rec = pth->rs_rec;
  return ref_indices_of_transfers(ppr, ppr2, ppt2, pbs, pbs2, ptr2);
}

int los_ref_run_pairs(struct precision *ppr, struct precision2 *ppr2, struct background *pba,
                      struct thermo *pth, struct perturbs2 *ppt2, struct bessels *pbs,
                      struct bessels2 *pbs2, struct transfers2 *ptr2, const int *k1s,
                      const int *k2s, int npairs, long *n_out, double *out)
{
  ptr2->tau0 = pba->conformal_age;
  ptr2->tau_rec = pth->tau_rec;
  ptr2->rs_rec = pth->rs_rec;
  ptr2->has_cls = _TRUE_;
  if (ref_indices_of_transfers(ppr, ppr2, ppt2, pbs, pbs2, ptr2) == _FAILURE_) return _FAILURE_;
  r_test(ptr2->tau_sampling != sources_tau_sampling, ptr2->error_message,
         "sample timing supports sources tau sampling only");
  int nth = 1;
#ifdef _OPENMP
#pragma omp parallel
#pragma omp master
  nth = omp_get_num_threads();
#endif
  const int T = ppt2->tau_size;
  struct transfer2_workspace **ppw = calloc(nth, sizeof *ppw);
  for (int th = 0; th < nth; ++th) {
    struct transfer2_workspace *w = calloc(1, sizeof *w);
    w->k_grid = malloc(ptr2->k3_size_max * sizeof(double));
    w->tau_grid = malloc(T * sizeof(double));
    w->tau0_minus_tau = malloc(T * sizeof(double));
    w->delta_tau = malloc(T * sizeof(double));
    w->index_tau_left = malloc(T * sizeof(int));
    w->sources_time_spline = malloc(ppt2->tp2_size * sizeof(double *));
    w->interpolated_sources_in_time = malloc(ppt2->tp2_size * sizeof(double *));
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      w->sources_time_spline[tp] = malloc(T * sizeof(double));
      w->interpolated_sources_in_time[tp] = malloc(T * sizeof(double));
    }
    ppw[th] = w;
  }
  double **spl = malloc(ppt2->tp2_size * sizeof(double *));
  double **isk = malloc(ppt2->tp2_size * sizeof(double *));
  long count = 0;
  for (int q = 0; q < npairs; ++q) {
    const int i1 = k1s[q], i2 = k2s[q];
    /* scratch output levels for this pair only */
    double ***save = malloc(ptr2->tt2_size * sizeof(double **));
    for (int tt = 0; tt < ptr2->tt2_size; ++tt) {
      save[tt] = ptr2->transfer[tt][i1];
      ptr2->transfer[tt][i1] = calloc(i1 + 1, sizeof(double *));
      ptr2->transfer[tt][i1][i2] = malloc(ptr2->k_size_k1k2[i1][i2] * sizeof(double));
    }
    int last_used;
    double *kg = malloc(ptr2->k3_size_max * sizeof(double));
    if (los_ref_get_k3_list(ppr, ppr2, ppt2, pbs2, ptr2, i1, i2, kg, &last_used) == _FAILURE_)
      return _FAILURE_;
    for (int th = 0; th < nth; ++th)
      memcpy(ppw[th]->k_grid, kg, ptr2->k_size_k1k2[i1][i2] * sizeof(double));
    free(kg);
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      spl[tp] = malloc((size_t)ppt2->k3_size[i1][i2] * T * sizeof(double));
      isk[tp] = malloc((size_t)ptr2->k_size_k1k2[i1][i2] * T * sizeof(double));
      if (ref_interpolate_sources_in_k(ppr, ppr2, ppt2, ptr2, i1, i2, tp, ppw[0]->k_grid, spl[tp],
                                       isk[tp]) == _FAILURE_)
        return _FAILURE_;
    }
    int abort_flag = 0;
#pragma omp parallel shared(ppw, abort_flag)
    {
      int thread = 0;
#ifdef _OPENMP
      thread = omp_get_thread_num();
#endif
      struct transfer2_workspace *pw = ppw[thread];
#pragma omp for schedule(static)
      for (int ik = 0; ik < ptr2->k_size_k1k2[i1][i2]; ++ik) {
        pw->index_k = ik;
        pw->k = pw->k_grid[ik];
        if (ref_get_time_grid(ppr2, ppt2, pbs2, ptr2, pw) == _FAILURE_) { abort_flag = 1; continue; }
        for (int tp = 0; tp < ppt2->tp2_size; ++tp)
          ref_interpolate_sources_in_time(ppr2, ppt2, ptr2, isk[tp], pw->sources_time_spline[tp],
                                          pw->interpolated_sources_in_time[tp], pw, NULL);
        for (int tt = 0; tt < ptr2->tt2_size; tt++)
          if (ref_compute(ppr, ppr2, ppt2, pbs2, ptr2, i1, i2, ik, ptr2->corresponding_index_l[tt],
                          ptr2->corresponding_index_m[tt], tt, pw->interpolated_sources_in_time,
                          pw) == _FAILURE_)
            abort_flag = 1;
      }
    }
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      free(spl[tp]);
      free(isk[tp]);
    }
    for (int tt = 0; tt < ptr2->tt2_size; ++tt) {
      if (out != NULL && !abort_flag) {
        memcpy(out + count, ptr2->transfer[tt][i1][i2],
               (size_t)ptr2->k_size_k1k2[i1][i2] * sizeof(double));
        count += ptr2->k_size_k1k2[i1][i2];
      }
      free(ptr2->transfer[tt][i1][i2]);
      free(ptr2->transfer[tt][i1]);
      ptr2->transfer[tt][i1] = save[tt];
    }
    free(save);
    if (abort_flag) return _FAILURE_;
    if (out == NULL) count += (long)ptr2->tt2_size * ptr2->k_size_k1k2[i1][i2];
  }
  free(spl);
  free(isk);
  for (int th = 0; th < nth; ++th) {
    struct transfer2_workspace *w = ppw[th];
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      free(w->sources_time_spline[tp]);
      free(w->interpolated_sources_in_time[tp]);
    }
    free(w->sources_time_spline);
    free(w->interpolated_sources_in_time);
    free(w->k_grid);
    free(w->tau_grid);
    free(w->tau0_minus_tau);
    free(w->delta_tau);
    free(w->index_tau_left);
    free(w);
  }
  free(ppw);
  *n_out = count;
  return _SUCCESS_;
}

/* The oracle's time grid of one k (source/transfer2.c:2366-2585), for its pinning
   tests: grid, index_tau_left and delta_tau, at most cap points; returns the grid size,
   or -1 (message in ptr2->error_message). */
int los_ref_time_grid(struct precision2 *ppr2, struct perturbs2 *ppt2, struct bessels2 *pbs2,
                      struct transfers2 *ptr2, double k, double *tau, int *left, double *dtau,
                      int cap)
{
  struct transfer2_workspace w;
  memset(&w, 0, sizeof w);
  w.k = k;
  /* room for every point the grid can hold: the sources' points plus one per
     tau_step_max over the whole range */
  double step = 0;
  if (ptr2->tau_sampling == custom_tau_sampling) step = 2 * M_PI / k * ppr2->tau_linstep_song;
  if (ptr2->tau_sampling == bessel_tau_sampling) step = pbs2->xx_step / k;
  long room = 2L * ppt2->tau_size + 2;
  if (step > 0)
    room += (long)ceil((ppt2->tau_sampling[ppt2->tau_size - 1] - ppt2->tau_sampling[0]) / step);
  if (room > 50000000L) return -1;
  const long big = room;
  w.tau_grid = malloc(big * sizeof(double));
  w.tau0_minus_tau = malloc(big * sizeof(double));
  w.delta_tau = malloc(big * sizeof(double));
  w.index_tau_left = malloc(big * sizeof(int));
  int n = -1;
  if (w.tau_grid && w.tau0_minus_tau && w.delta_tau && w.index_tau_left &&
      ref_get_time_grid(ppr2, ppt2, pbs2, ptr2, &w) == _SUCCESS_ && w.tau_grid_size <= cap) {
    n = w.tau_grid_size;
    memcpy(tau, w.tau_grid, n * sizeof(double));
    memcpy(dtau, w.delta_tau, n * sizeof(double));
    if (ptr2->tau_sampling == sources_tau_sampling)
      for (int i = 0; i < n; ++i) left[i] = i;
    else
      memcpy(left, w.index_tau_left, n * sizeof(int));
  }
  free(w.tau_grid);
  free(w.tau0_minus_tau);
  free(w.delta_tau);
  free(w.index_tau_left);
  return n;
}

/* The oracle's (l,m) helpers, exported for their fixture tests. */
int los_ref_size_indexl_indexm(const int *lv, int ls, const int *mv, int ms)
{
  return ref_size_indexl_indexm(lv, ls, mv, ms);
}
int los_ref_offset_indexl_indexm(int l, int m, const int *lv, int ls, const int *mv, int ms)
{
  return ref_offset_indexl_indexm(l, m, lv, ls, mv, ms);
}

/* ---------------------------------------------------------------------------
 * k3 integral of the transfer functions, source/bispectra2.c:1182-1410: for
 * each transfer row q (one l3), pair p and r, bessel2_convolution
 * (source/bessel2.c:499-605) over the pair's k3 grid with its trapezoid
 * measure (:1283-1287), then the M3 coefficient (only for M3 != 0, :1355-1356)
 * and the factor 2 (:1361). Rows with index_L < 0 are the skipped (l3, M3)
 * configurations (:1196-1205) and stay zero. out[(q * n_r + r) * n_pairs + p].
 * ------------------------------------------------------------------------- */
static double ref_bessel2_convolution(const double *kk, const double *dkk, int k_size,
                                      const double *f, const double *j, double x_min,
                                      double xx_step, double r)
{
  double integral = 0;
  for (int ik = 0; ik < k_size; ++ik) {
    double f_in_k = f[ik];
    if (f_in_k == 0.) continue;
    double g_in_k = 1.;
    double k = kk[ik];
    double x = k * r;
    if (x < x_min) continue;
    int index_x = (int)((x - x_min) / xx_step);
    double a = (x_min + xx_step * (index_x + 1) - x) / xx_step;
    double jv = a * j[index_x] + (1. - a) * j[index_x + 1];
    double integrand = k * k * jv * f_in_k * g_in_k;
    integral += integrand * dkk[ik];
  }
  integral *= 0.5;
  return integral;
}

int los_ref_k3_integral(int n_pairs, const long *pair_off, const double *k3, int n_q,
                        const double *transfer, const int *index_L, const long *j_off,
                        const double *j, const double *x_min, double xx_step, int n_r,
                        const double *r, int abs_M3, double M3_coefficient, double *out)
{
  const long n_k3 = n_pairs > 0 ? pair_off[n_pairs] : 0;
  double *dk = malloc((size_t)(n_k3 > 0 ? n_k3 : 1) * sizeof(double));
  if (!dk) return 1;
  for (int p = 0; p < n_pairs; ++p) {
    const long b = pair_off[p];
    const int n = (int)(pair_off[p + 1] - b);
    if (n < 2) { free(dk); return 1; }
    dk[b] = k3[b + 1] - k3[b];
    for (int i = 1; i < n - 1; ++i) dk[b + i] = k3[b + i + 1] - k3[b + i - 1];
    dk[b + n - 1] = k3[b + n - 1] - k3[b + n - 2];
  }
  memset(out, 0, sizeof(double) * (size_t)n_q * n_r * n_pairs);
  for (int q = 0; q < n_q; ++q) {
    const int iL = index_L[q];
    if (iL < 0) continue;
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < n_pairs; ++p) {
      const long b = pair_off[p];
      const int n = (int)(pair_off[p + 1] - b);
      for (int ir = 0; ir < n_r; ++ir) {
        double v = ref_bessel2_convolution(k3 + b, dk + b, n, transfer + (long)q * n_k3 + b,
                                           j + j_off[iL], x_min[iL], xx_step, r[ir]);
        if (abs_M3 != 0) v *= M3_coefficient;
        v *= 2;
        out[((long)q * n_r + ir) * n_pairs + p] = v;
      }
    }
  }
  free(dk);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Second-order C_l, source/spectra2.c:241-590: per l (OpenMP over l, :283),
 * sequential sums over M, k1, k2 <= k1 and the pair's k3 of
 * total_factor * T1 * T2 (:502-516) into cl[l] and cl_m[l][M]. k1/k2 half
 * trapezoid steps :375-380 and :398-404, P(k) = 2 pi^2/k^3 pk :388, :420, the
 * k3 steps with the triangle-edge corrections :486-500. Where the reference
 * reads k3_grid[index_k3 + 1] at the last point (:495, one past the pair's
 * grid, a value left from earlier pairs), the point past the end is taken as
 * beyond k1 + k2. Rows with used[l][M] = 0 are skipped (BB with m = 0).
 * ------------------------------------------------------------------------- */
int los_ref_cl2(int k_size, const double *k, const double *pk, const long *pair_off,
                const double *k3g, int n_l, int m_size, const int *mv, const double *t1,
                const double *t2, const int *used, int rescale, double *cl, double *cl_m)
{
  const double PI = 3.1415926535897932384626433832795e0;
  const int n_pairs = k_size * (k_size + 1) / 2;
  const long n_k3 = pair_off[n_pairs];
#pragma omp parallel for schedule(dynamic)
  for (int il = 0; il < n_l; ++il) {
    double result = 0;
    for (int iM = 0; iM < m_size; ++iM) {
      const int m = mv[iM];
      double result_m = 0;
      if (used[il * m_size + iM]) {
        const double *T1 = t1 + ((long)il * m_size + iM) * n_k3;
        const double *T2 = t2 + ((long)il * m_size + iM) * n_k3;
        int p = 0;
        for (int i1 = 0; i1 < k_size; ++i1) {
          double step_k1;
          if (i1 == 0) step_k1 = (k[1] - k[0]) / 2.;
          else if (i1 == k_size - 1) step_k1 = (k[k_size - 1] - k[k_size - 2]) / 2.;
          else step_k1 = (k[i1 + 1] - k[i1 - 1]) / 2.;
          const double k1 = k[i1];
          const double spectra_k1 = 2 * PI * PI / (k1 * k1 * k1) * pk[i1];
          for (int i2 = 0; i2 <= i1; ++i2, ++p) {
            double step_k2;
            if (i1 == 0) step_k2 = 0.;
            else if (i2 == 0) step_k2 = (k[1] - k[0]) / 2.;
            else if (i2 == i1) step_k2 = (k[i1] - k[i1 - 1]) / 2.;
            else step_k2 = (k[i2 + 1] - k[i2 - 1]) / 2.;
            const double k2 = k[i2];
            const double spectra_k2 = 2 * PI * PI / (k2 * k2 * k2) * pk[i2];
            const long b = pair_off[p];
            const int n = (int)(pair_off[p + 1] - b);
            const double *g = k3g + b;
            int triangular_first = 0, triangular_last = 0;
            for (int i3 = 0; i3 < n; ++i3) {
              const double k3 = g[i3];
              double step_k3;
              if (i3 == 0) step_k3 = (g[1] - g[0]) / 2.;
              else if (i3 == n - 1) step_k3 = (g[n - 1] - g[n - 2]) / 2.;
              else step_k3 = (g[i3 + 1] - g[i3 - 1]) / 2.;
              if (k3 < k1 - k2 || k3 > k1 + k2) step_k3 = 0.;
              if (k3 >= k1 - k2 && triangular_first == 0 && i3 < n - 1) {
                triangular_first = 1;
                step_k3 = (g[i3 + 1] - k3) / 2. + k3 - (k1 - k2);
              }
              const double next = i3 < n - 1 ? g[i3 + 1] : INFINITY;
              if (next > k1 + k2 && triangular_last == 0 && i3 > 0) {
                triangular_last = 1;
                step_k3 = (k3 - g[i3 - 1]) / 2. + (k1 + k2) - k3;
              }
              if (step_k3 == 0) continue;
              const double scale =
                  pow(sqrt(1. - pow((k3 * k3 + k1 * k1 - k2 * k2) / 2. / k3 / k1, 2)), m);
              const double total_factor = 2. * k1 * k2 * k3 * step_k1 * step_k2 * step_k3 *
                                          (rescale && step_k3 != 0. ? pow(scale, 2) : 1.) *
                                          2. / PI / 2. / PI / 2. / PI / 2. / PI * spectra_k1 *
                                          spectra_k2;
              const double integrand = total_factor * T1[b + i3] * T2[b + i3];
              result += integrand;
              result_m += integrand;
            }
          }
        }
      }
      cl_m[il * m_size + iM] = result_m;
    }
    cl[il] = result;
  }
  return 0;
}
