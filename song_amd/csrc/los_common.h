/*
 * los_common.h — declarations shared by every source of the song_los library
 * (not part of the public C-ABI). It does not define SONG's structures: the
 * HIP sources see them only through the flat descriptor of los_stage.h.
 */
#ifndef SONG_LOS_COMMON_H
#define SONG_LOS_COMMON_H

#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <stddef.h>

#include "../../include/song_los.h"
#include "los_stage.h"

#ifndef _ERRORMSGSIZE_
#define _ERRORMSGSIZE_ 2048
#endif
#ifndef _TRUE_
#define _TRUE_ 1
#define _FALSE_ 0
#endif

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

/* ---- the engine and the server on the flat descriptor (los_engine.hip,
   los_serve.hip); los_transfer2.c fills the descriptor from SONG's structures ---- */
double los_engine_batch_budget(los_engine *eng);
int los_engine_load_stage(los_engine *eng, const los_stage *st, const int *index_k1,
                          const int *index_k2, int npairs, char *errmsg);
int los_engine_store_stage(los_engine *eng, const los_stage *st, char *errmsg);
int los_server_create_stage(int device, int form, double hbm_budget_bytes, int n_slots,
                            const los_stage *st, los_server **out, char *errmsg);

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
/* The GPU part of the J builder (los_bessel2_gpu.hip) on flat inputs: the x
   grid, the l1 cut and J cut, the L and l lists and the rows of the setup.
   It finds each row's first index (rows[q].xmin), asks the caller to allocate
   the rows (alloc_rows: dst[q] receives nx - rows[q].xmin values) and fills them. */
typedef struct {
  int device;
  int nx;
  const double *xx;
  double xx_step, xx_max, j_cut, J_cut;
  int L_size;
  const int *L;
  int l_size;
  const int *l;
  los_jrow *rows;
  int nrows, l1max;
  int (*alloc_rows)(void *ctx, const los_jrow *rows, int nrows, double **dst, char *errmsg);
  void *ctx;
} los_jbuild;
int los_bessel2_gpu_fill(const los_jbuild *d, char *errmsg);

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
