/*
 * los_internal.h — declarations shared by the host-side C sources of the
 * song_los library (not part of the public C-ABI).
 */
#ifndef SONG_LOS_INTERNAL_H
#define SONG_LOS_INTERNAL_H

#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <stddef.h>

#include "../../include/song_los_types.h"
#include "../../include/song_los.h"
#include "../../include/transfer2.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define LOS_MIN(a, b) ((a) < (b) ? (a) : (b))
#define LOS_MAX(a, b) ((a) > (b) ? (a) : (b))

/* Error helpers in the CLASS style: write the message and return _FAILURE_. */
#define los_fail(errbuf, ...)                                                   \
  do {                                                                          \
    snprintf((errbuf), _ERRORMSGSIZE_, __VA_ARGS__);                            \
    return _FAILURE_;                                                           \
  } while (0)

#define los_test(cond, errbuf, ...)                                             \
  do {                                                                          \
    if (cond) {                                                                 \
      char los_tmp__[_ERRORMSGSIZE_];                                           \
      snprintf(los_tmp__, sizeof los_tmp__, __VA_ARGS__);                       \
      snprintf((errbuf), _ERRORMSGSIZE_, "%s: %s", __func__, los_tmp__);        \
      return _FAILURE_;                                                         \
    }                                                                           \
  } while (0)

#define los_alloc(ptr, size, errbuf)                                            \
  do {                                                                          \
    (ptr) = malloc(size);                                                       \
    if ((ptr) == NULL && (size) != 0)                                           \
      los_fail(errbuf, "%s: could not allocate %s (%ld bytes)", __func__, #ptr, \
               (long)(size));                                                   \
  } while (0)

#define los_calloc(ptr, n, size, errbuf)                                        \
  do {                                                                          \
    (ptr) = calloc((n), (size));                                                \
    if ((ptr) == NULL && (n) != 0)                                              \
      los_fail(errbuf, "%s: could not allocate %s", __func__, #ptr);            \
  } while (0)

/* Propagate a failure of a callee, prefixing the caller's name. */
#define los_call(expr, errsrc, errdst)                                          \
  do {                                                                          \
    if ((expr) == _FAILURE_) {                                                  \
      char los_tmp__[_ERRORMSGSIZE_];                                           \
      snprintf(los_tmp__, sizeof los_tmp__, "%s: %s", __func__, (errsrc));     \
      memcpy((errdst), los_tmp__, _ERRORMSGSIZE_);                              \
      return _FAILURE_;                                                         \
    }                                                                           \
  } while (0)

/* ---- engine helpers used by transfer2_init (los_engine.hip) ---- */
double los_engine_batch_budget(los_engine *eng);
double los_engine_batch_bytes(struct precision2 *ppr2, struct perturbs2 *ppt2,
                              struct transfers2 *ptr2, int k1_begin, int k1_end);

/* ---- J-table builders (los_bessel2.c, los_bessel2_gpu.hip) ---- */
/* One J_Llm row: its indices in [J][L][l][m], the l1 terms of
   sum_l1 coef * j_l1(x), and the first stored x index. */
typedef struct {
  int J, iL, il, im;
  int nterm;
  int l1[32];
  double coef[32];
  int xmin;
} los_jrow;
int los_bessel2_setup(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, los_jrow **rows,
                      int *nrows, int *l1max, char *errmsg);
int los_bessel2_alloc_rows(struct bessels2 *pbs2, const los_jrow *rows, int nrows,
                           char *errmsg);
void los_bessel2_spline_rows(struct bessels *pbs, struct bessels2 *pbs2, const los_jrow *rows,
                             int nrows);
int los_bessel2_build_gpu(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, int device,
                          char *errmsg);

/* ---- special functions (los_special.c) ---- */
void los_sph_bessel_array(int lmax, double x, double *out);
int los_wigner3j_j1(int j2, int j3, int m2, int m3, int *j1min, int *j1max,
                    double *out, int out_cap);
int los_spline_columns(const double *x, int n, const double *y, int ncol, double *dd,
                       int natural, double *work);
void los_spline_one(const double *x, int n, const double *y, double *dd, int natural,
                    double *work);

/* ---- (l,m) offsets (los_lm.c), the song_tools helpers of the CLASS fork ---- */
int los_size_l_indexm(int l_max, const int *m_vec, int m_size);
int los_offset_l_indexm(int l, int m, const int *m_vec, int m_size);
int los_size_indexl_indexm(const int *l_vec, int l_size, const int *m_vec, int m_size);
int los_offset_indexl_indexm(int l, int m, const int *l_vec, int l_size, const int *m_vec,
                             int m_size);
int los_offset2multipole_indexl_indexm(int offset, const int *l_vec, int l_size,
                                       const int *m_vec, int m_size, int *index_l,
                                       int *index_m);

#ifdef __cplusplus
}
#endif

#endif
