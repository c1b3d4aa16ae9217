/*
 * los_bessel2.c — FP64 builder of the projection functions J_Llm(x) in
 * struct bessels2, the input tables of the LOS stage.
 *
 * Semantics follow SONG's bessel2 module:
 *  - xx = linear grid on [0, xx_max], xx_size = (int)(xx_max/xx_step + 1)
 *    (source/bessel2.c:960-985);
 *  - J types TT (spin 0), EE and EB (spin 2; EB only when some m > 0)
 *    (source/bessel2.c:145-173, :1041-1095);
 *  - J_Llm(x) = (-1)^m (2l+1) sum_l1 i^(l-l1-L) (2l1+1) (l1 L l;0 S -S)
 *               (l1 L l;0 -m m) j_l1(x), with the parity selection of
 *    source/bessel2.c:1460-1555 (even l-l1-L for TT/EE, odd for EB);
 *  - j_l1(x) is taken as zero below the first grid point where |j_l1| reaches
 *    j_l1_cut (source/bessel2.c:1669-1815);
 *  - index_xmin_J by the bisection of source/bessel2.c:1197-1338 on J_Llm_cut,
 *    xx_size-1 when J is negligible everywhere (then the row is one zero),
 *    0 when L == l;
 *  - rows stored from index_xmin_J on; ddJ (cubic case) from a spline with
 *    estimated end derivatives for l == 2 and natural ends otherwise
 *    (source/bessel2.c:395-463).
 * The spherical Bessel functions and 3j symbols come from los_special.c.
 */
#include "los_internal.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef los_jrow jrow;

static int sign_of_power(int n) { return (n % 2 == 0) ? 1 : -1; }

/* J at grid index ix from a row description, with j_l1 values jl (0..l1max). */
static double row_value(const jrow *r, const double *jl, const int *xmin_l1, int ix)
{
  double acc = 0.0;
  for (int t = 0; t < r->nterm; ++t) {
    int l1 = r->l1[t];
    if (ix < xmin_l1[l1]) continue;
    acc += r->coef[t] * jl[l1];
  }
  return acc;
}

/* Shared first phase of both builders: the x grid, the J types, the [J][L][l][m]
   index arrays and the rows with their l1 term lists (3j coefficients). Rows
   whose J vanishes identically get xmin = xx_size - 1, the others xmin = -1. */
int los_bessel2_setup(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, los_jrow **rows_out,
                      int *nrows_out, int *l1max_out, char *errmsg)
{
  /* xx grid */
  pbs2->xx_step = ppr2->bessel_x_step_song;
  pbs2->xx_size = (int)(pbs2->xx_max / pbs2->xx_step + 1);
  los_test(pbs2->xx_size < 2, errmsg, "xx grid has %d points", pbs2->xx_size);
  los_alloc(pbs2->xx, pbs2->xx_size * sizeof(double), errmsg);
  for (int i = 0; i < pbs2->xx_size; ++i)
    pbs2->xx[i] = 0.0 + i * (pbs2->xx_max - 0.0) / (pbs2->xx_size - 1);
  pbs2->j_l1_cut = ppr2->bessel_j_cut_song;
  pbs2->J_Llm_cut = ppr2->bessel_J_cut_song;

  pbs2->m_size = ppr2->m_size;
  los_alloc(pbs2->m, pbs2->m_size * sizeof(int), errmsg);
  for (int i = 0; i < pbs2->m_size; ++i) pbs2->m[i] = ppr2->m[i];
  pbs2->L_max = LOS_MAX(ppr2->l_max_los, ppr2->m_max_song);
  pbs2->L_size = pbs2->L_max + 1;
  los_alloc(pbs2->L, pbs2->L_size * sizeof(int), errmsg);
  for (int i = 0; i < pbs2->L_size; ++i) pbs2->L[i] = i;

  int index_J = 0;
  pbs2->has_J_TT = pbs2->has_J_EE = pbs2->has_J_EB = _FALSE_;
  if (ppt2->has_cmb_temperature) {
    pbs2->has_J_TT = _TRUE_;
    pbs2->index_J_TT = index_J++;
  }
  if (ppt2->has_cmb_polarization_e || ppt2->has_cmb_polarization_b) {
    pbs2->has_J_EE = _TRUE_;
    pbs2->index_J_EE = index_J++;
    if (ppr2->m_max_song > 0) {
      pbs2->has_J_EB = _TRUE_;
      pbs2->index_J_EB = index_J++;
    }
  }
  pbs2->J_size = index_J;

  const int l_size = pbs->l_size;
  const int l1max = pbs->l[l_size - 1] + pbs2->L_max + 2;
  *l1max_out = l1max;

  /* Enumerate rows and their l1 coefficient lists. */
  int nrows = 0;
  for (int J = 0; J < pbs2->J_size; ++J)
    for (int iL = 0; iL < pbs2->L_size; ++iL)
      for (int il = 0; il < l_size; ++il) {
        int mmax = LOS_MIN(ppr2->index_m_max[pbs2->L[iL]], ppr2->index_m_max[pbs->l[il]]);
        nrows += mmax + 1;
      }
  jrow *rows;
  los_calloc(rows, nrows, sizeof(jrow), errmsg);

  /* allocate [J][L][l][m] levels */
  los_alloc(pbs2->index_xmin_J, pbs2->J_size * sizeof(int ***), errmsg);
  los_alloc(pbs2->x_size_J, pbs2->J_size * sizeof(int ***), errmsg);
  los_alloc(pbs2->x_min_J, pbs2->J_size * sizeof(double ***), errmsg);
  los_alloc(pbs2->J_Llm_x, pbs2->J_size * sizeof(double ****), errmsg);
  pbs2->ddJ_Llm_x = NULL;
  if (ppr->bessels_interpolation == cubic_interpolation)
    los_alloc(pbs2->ddJ_Llm_x, pbs2->J_size * sizeof(double ****), errmsg);

  int r = 0;
  for (int J = 0; J < pbs2->J_size; ++J) {
    los_alloc(pbs2->index_xmin_J[J], pbs2->L_size * sizeof(int **), errmsg);
    los_alloc(pbs2->x_size_J[J], pbs2->L_size * sizeof(int **), errmsg);
    los_alloc(pbs2->x_min_J[J], pbs2->L_size * sizeof(double **), errmsg);
    los_alloc(pbs2->J_Llm_x[J], pbs2->L_size * sizeof(double ***), errmsg);
    if (pbs2->ddJ_Llm_x) los_alloc(pbs2->ddJ_Llm_x[J], pbs2->L_size * sizeof(double ***), errmsg);
    enum projection_function_types type =
        (pbs2->has_J_TT && J == pbs2->index_J_TT) ? J_TT
        : (pbs2->has_J_EE && J == pbs2->index_J_EE) ? J_EE : J_EB;
    int S = (type == J_TT) ? 0 : 2;
    for (int iL = 0; iL < pbs2->L_size; ++iL) {
      los_alloc(pbs2->index_xmin_J[J][iL], l_size * sizeof(int *), errmsg);
      los_alloc(pbs2->x_size_J[J][iL], l_size * sizeof(int *), errmsg);
      los_alloc(pbs2->x_min_J[J][iL], l_size * sizeof(double *), errmsg);
      los_alloc(pbs2->J_Llm_x[J][iL], l_size * sizeof(double **), errmsg);
      if (pbs2->ddJ_Llm_x) los_alloc(pbs2->ddJ_Llm_x[J][iL], l_size * sizeof(double **), errmsg);
      for (int il = 0; il < l_size; ++il) {
        int L = pbs2->L[iL];
        int l = pbs->l[il];
        int msz = LOS_MIN(ppr2->index_m_max[L], ppr2->index_m_max[l]) + 1;
        los_alloc(pbs2->index_xmin_J[J][iL][il], LOS_MAX(msz, 1) * sizeof(int), errmsg);
        los_alloc(pbs2->x_size_J[J][iL][il], LOS_MAX(msz, 1) * sizeof(int), errmsg);
        los_alloc(pbs2->x_min_J[J][iL][il], LOS_MAX(msz, 1) * sizeof(double), errmsg);
        los_alloc(pbs2->J_Llm_x[J][iL][il], LOS_MAX(msz, 1) * sizeof(double *), errmsg);
        if (pbs2->ddJ_Llm_x)
          los_alloc(pbs2->ddJ_Llm_x[J][iL][il], LOS_MAX(msz, 1) * sizeof(double *), errmsg);
        for (int im = 0; im < msz; ++im) {
          int m = pbs2->m[im];
          jrow *row = &rows[r++];
          row->J = J; row->iL = iL; row->il = il; row->im = im;
          row->nterm = 0;
          row->xmin = -1;
          int spin_constraint = (abs(S) > l) || (abs(S) > L);
          int eb_constraint = (type == J_EB) && (m == 0);
          if (spin_constraint || eb_constraint) { row->xmin = pbs2->xx_size - 1; continue; }
          double f1[64], f2[64];
          int a1, b1, a2, b2;
          /* (l1 L l; 0 S -S) and (l1 L l; 0 -m m) over l1 */
          int n1 = los_wigner3j_j1(L, l, S, -S, &a1, &b1, f1, 64);
          int n2 = (S != 0 || m != 0) ? los_wigner3j_j1(L, l, -m, m, &a2, &b2, f2, 64) : n1;
          if (S == 0 && m == 0) { a2 = a1; b2 = b1; for (int q = 0; q < n1; ++q) f2[q] = f1[q]; }
          los_test(n1 < 0 || n2 < 0, errmsg, "3j failed for L=%d l=%d m=%d", L, l, m);
          for (int q = 0; q < n1; ++q) {
            int l1 = a1 + q;
            if (f1[q] == 0.0) continue;
            if (l1 < a2 || l1 > b2) continue;
            double s2 = f2[l1 - a2];
            int is_even = ((l - l1 - L) % 2 == 0);
            double pref;
            if (type == J_TT || type == J_EE) {
              if (!is_even) continue;
              pref = sign_of_power((l - l1 - L) / 2);
            } else {
              if (is_even) continue;
              pref = sign_of_power((l - l1 - L - 1) / 2);
            }
            double c = pref * (2 * l1 + 1) * f1[q] * s2;
            if (c == 0.0) continue;
            los_test(row->nterm >= 32, errmsg, "too many l1 terms");
            row->l1[row->nterm] = l1;
            row->coef[row->nterm] = c * sign_of_power(m) * (2 * l + 1);
            row->nterm++;
          }
          if (row->nterm == 0) row->xmin = pbs2->xx_size - 1;
        }
      }
    }
  }
  *rows_out = rows;
  *nrows_out = nrows;
  return _SUCCESS_;
}

/* Row arrays of struct bessels2 once every row's xmin is known: index_xmin_J,
   x_size_J, x_min_J and zeroed J (and ddJ) rows from xmin to the grid end. */
int los_bessel2_alloc_rows(struct bessels2 *pbs2, const los_jrow *rows, int nrows, char *errmsg)
{
  pbs2->count_allocated_Js = 0;
  pbs2->x_size_max_J = 0;
  for (int q = 0; q < nrows; ++q) {
    const jrow *row = &rows[q];
    int xs = pbs2->xx_size - row->xmin;
    pbs2->index_xmin_J[row->J][row->iL][row->il][row->im] = row->xmin;
    pbs2->x_size_J[row->J][row->iL][row->il][row->im] = xs;
    pbs2->x_min_J[row->J][row->iL][row->il][row->im] =
        (row->xmin == pbs2->xx_size - 1) ? pbs2->xx_max : pbs2->xx[row->xmin];
    double *p;
    los_calloc(p, xs, sizeof(double), errmsg);
    pbs2->J_Llm_x[row->J][row->iL][row->il][row->im] = p;
    if (pbs2->ddJ_Llm_x) {
      double *d;
      los_calloc(d, xs, sizeof(double), errmsg);
      pbs2->ddJ_Llm_x[row->J][row->iL][row->il][row->im] = d;
    }
    pbs2->count_allocated_Js += xs;
    if (xs > pbs2->x_size_max_J) pbs2->x_size_max_J = xs;
  }
  return _SUCCESS_;
}

/* Second derivatives of the J rows for cubic interpolation in x: estimated end
   derivatives for l == 2, natural ends otherwise (source/bessel2.c:395-463). */
void los_bessel2_spline_rows(struct bessels *pbs, struct bessels2 *pbs2, const los_jrow *rows,
                             int nrows)
{
  if (!pbs2->ddJ_Llm_x) return;
  double *work = malloc(pbs2->xx_size * sizeof(double));
  for (int q = 0; q < nrows; ++q) {
    const jrow *row = &rows[q];
    int xs = pbs2->x_size_J[row->J][row->iL][row->il][row->im];
    if (xs < 3) continue;
    int natural = (pbs->l[row->il] == 2) ? 0 : 1;
    los_spline_one(pbs2->xx + row->xmin, xs, pbs2->J_Llm_x[row->J][row->iL][row->il][row->im],
                   pbs2->ddJ_Llm_x[row->J][row->iL][row->il][row->im], natural, work);
  }
  free(work);
}

int los_bessel2_build(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, char *errmsg)
{
  jrow *rows;
  int nrows, l1max;
  if (los_bessel2_setup(ppr, ppr2, ppt2, pbs, pbs2, &rows, &nrows, &l1max, errmsg) == _FAILURE_)
    return _FAILURE_;

  /* first grid index where |j_l1| >= j_cut, minus one (the stored range
     starts at the point preceding the threshold) */
  int *xmin_l1;
  los_alloc(xmin_l1, (l1max + 1) * sizeof(int), errmsg);
  {
    double *jl = malloc((l1max + 1) * sizeof(double));
    los_test(jl == NULL, errmsg, "out of memory");
    /* j_l1 grows monotonically for x < l1; bisection over grid indices per l1 */
    for (int l1 = 0; l1 <= l1max; ++l1) {
      if (l1 == 0) { xmin_l1[0] = 0; continue; }
      int hi = 0;
      double xpeak = l1 + 0.5;
      if (xpeak >= pbs2->xx[pbs2->xx_size - 1]) hi = pbs2->xx_size - 1;
      else { hi = (int)(xpeak / pbs2->xx_step); if (hi >= pbs2->xx_size) hi = pbs2->xx_size - 1; }
      los_sph_bessel_array(l1, pbs2->xx[hi], jl);
      if (fabs(jl[l1]) < pbs2->j_l1_cut) { xmin_l1[l1] = pbs2->xx_size - 1; continue; }
      int lo = 0;
      while (hi - lo > 1) {
        int mid = (hi + lo) / 2;
        los_sph_bessel_array(l1, pbs2->xx[mid], jl);
        if (fabs(jl[l1]) >= pbs2->j_l1_cut) hi = mid; else lo = mid;
      }
      xmin_l1[l1] = lo;
    }
    free(jl);
  }


  /* index_xmin_J by bisection on the row values. */
  int abort_flag = 0;
#pragma omp parallel
  {
    double *jl = malloc((l1max + 1) * sizeof(double));
#pragma omp for schedule(dynamic, 4)
    for (int q = 0; q < nrows; ++q) {
      jrow *row = &rows[q];
      if (row->xmin >= 0) continue;
      int L = pbs2->L[row->iL], l = pbs->l[row->il];
      if (L == l) { row->xmin = 0; continue; }
      int l1min = abs(l - L);
      double xup = l1min + 0.5;
      int up;
      if (xup >= pbs2->xx_max) up = pbs2->xx_size - 1;
      else { up = 0; while (pbs2->xx[up] < xup) up++; }
      los_sph_bessel_array(l1max, pbs2->xx[up], jl);
      double Jv = row_value(row, jl, xmin_l1, up);
      if (fabs(Jv) < pbs2->J_Llm_cut) { row->xmin = pbs2->xx_size - 1; continue; }
      int down = 0;
      while (down != up - 1) {
        int mid = (up + down) / 2;
        los_sph_bessel_array(l1max, pbs2->xx[mid], jl);
        Jv = row_value(row, jl, xmin_l1, mid);
        if (fabs(Jv) >= pbs2->J_Llm_cut) up = mid; else down = mid;
      }
      row->xmin = down;
    }
    free(jl);
  }
  if (abort_flag) return _FAILURE_;

  if (los_bessel2_alloc_rows(pbs2, rows, nrows, errmsg) == _FAILURE_) return _FAILURE_;

  /* Fill rows: sweep x, one j_l array per x. */
  const int xs_total = pbs2->xx_size;
  int min_live_xmin = xs_total;
  for (int q = 0; q < nrows; ++q)
    if (rows[q].nterm > 0 && rows[q].xmin < xs_total - 1 && rows[q].xmin < min_live_xmin)
      min_live_xmin = rows[q].xmin;
#pragma omp parallel
  {
    double *jl = malloc((l1max + 1) * sizeof(double));
#pragma omp for schedule(dynamic, 16)
    for (int ix = 0; ix < xs_total; ++ix) {
      if (ix < min_live_xmin) continue;
      los_sph_bessel_array(l1max, pbs2->xx[ix], jl);
      for (int q = 0; q < nrows; ++q) {
        jrow *row = &rows[q];
        if (row->nterm == 0 || row->xmin == xs_total - 1 || ix < row->xmin) continue;
        pbs2->J_Llm_x[row->J][row->iL][row->il][row->im][ix - row->xmin] =
            row_value(row, jl, xmin_l1, ix);
      }
    }
    free(jl);
  }

  los_bessel2_spline_rows(pbs, pbs2, rows, nrows);

  free(rows);
  free(xmin_l1);
  return _SUCCESS_;
}

int los_bessel2_free(struct bessels *pbs, struct bessels2 *pbs2)
{
  if (pbs2->J_Llm_x == NULL) return _SUCCESS_;
  for (int J = 0; J < pbs2->J_size; ++J) {
    for (int iL = 0; iL < pbs2->L_size; ++iL) {
      for (int il = 0; il < pbs->l_size; ++il) {
        /* one row per m of the m-list with m <= min(L,l) */
        int lim = LOS_MIN(pbs2->L[iL], pbs->l[il]);
        for (int im = 0; im < pbs2->m_size && pbs2->m[im] <= lim; ++im) {
          free(pbs2->J_Llm_x[J][iL][il][im]);
          if (pbs2->ddJ_Llm_x) free(pbs2->ddJ_Llm_x[J][iL][il][im]);
        }
        free(pbs2->index_xmin_J[J][iL][il]);
        free(pbs2->x_size_J[J][iL][il]);
        free(pbs2->x_min_J[J][iL][il]);
        free(pbs2->J_Llm_x[J][iL][il]);
        if (pbs2->ddJ_Llm_x) free(pbs2->ddJ_Llm_x[J][iL][il]);
      }
      free(pbs2->index_xmin_J[J][iL]);
      free(pbs2->x_size_J[J][iL]);
      free(pbs2->x_min_J[J][iL]);
      free(pbs2->J_Llm_x[J][iL]);
      if (pbs2->ddJ_Llm_x) free(pbs2->ddJ_Llm_x[J][iL]);
    }
    free(pbs2->index_xmin_J[J]);
    free(pbs2->x_size_J[J]);
    free(pbs2->x_min_J[J]);
    free(pbs2->J_Llm_x[J]);
    if (pbs2->ddJ_Llm_x) free(pbs2->ddJ_Llm_x[J]);
  }
  free(pbs2->index_xmin_J);
  free(pbs2->x_size_J);
  free(pbs2->x_min_J);
  free(pbs2->J_Llm_x);
  if (pbs2->ddJ_Llm_x) free(pbs2->ddJ_Llm_x);
  free(pbs2->xx);
  free(pbs2->L);
  free(pbs2->m);
  pbs2->J_Llm_x = NULL;
  pbs2->ddJ_Llm_x = NULL;
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * The tables built on the GPU (los_bessel2_gpu_fill on the flat los_jbuild
 * inputs): the setup and the row allocation stay here, with the structures.
 * ------------------------------------------------------------------------- */
typedef struct {
  struct bessels2 *pbs2;
} gpu_rows_ctx;

static int gpu_alloc_rows(void *ctx, const los_jrow *rows, int nrows, double **dst, char *errmsg)
{
  struct bessels2 *pbs2 = ((gpu_rows_ctx *)ctx)->pbs2;
  if (los_bessel2_alloc_rows(pbs2, rows, nrows, errmsg) == _FAILURE_) return _FAILURE_;
  for (int q = 0; q < nrows; ++q)
    dst[q] = pbs2->J_Llm_x[rows[q].J][rows[q].iL][rows[q].il][rows[q].im];
  return _SUCCESS_;
}

int los_bessel2_build_gpu(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, int device, char *errmsg)
{
  los_jrow *rows = NULL;
  int nrows = 0, l1max = 0;
  if (los_bessel2_setup(ppr, ppr2, ppt2, pbs, pbs2, &rows, &nrows, &l1max, errmsg) == _FAILURE_)
    return _FAILURE_;
  gpu_rows_ctx ctx = {pbs2};
  los_jbuild d;
  d.device = device;
  d.nx = pbs2->xx_size;
  d.xx = pbs2->xx;
  d.xx_step = pbs2->xx_step;
  d.xx_max = pbs2->xx_max;
  d.j_cut = pbs2->j_l1_cut;
  d.J_cut = pbs2->J_Llm_cut;
  d.L_size = pbs2->L_size;
  d.L = pbs2->L;
  d.l_size = pbs->l_size;
  d.l = pbs->l;
  d.rows = rows;
  d.nrows = nrows;
  d.l1max = l1max;
  d.alloc_rows = gpu_alloc_rows;
  d.ctx = &ctx;
  const int rc = los_bessel2_gpu_fill(&d, errmsg);
  if (rc == _SUCCESS_) los_bessel2_spline_rows(pbs, pbs2, rows, nrows);
  free(rows);
  return rc;
}
