/*
 * los_transfer2.c — SONG's transfer2 interface (include/transfer2.h) on top of
 * the HIP engine of this library.
 *
 * The host side keeps SONG's bookkeeping bit-for-bit (it decides the output
 * layout that downstream modules re-derive): the (l,m) lists, the
 * triangle-limited k3 grids, the trapezoid weights, the ragged
 * ptr2->transfer array, disk storage and the consumption of ppt2->sources.
 * The LOS integral itself (k3 spline, J interpolation, tau reduction) runs in
 * the engine's kernels; no CPU path computes it.
 */
#include "los_internal.h"

#include <math.h>

static int g_default_form = LOS_FORM_FP64;
static int g_default_device = 0;

int los_set_default_form(int form)
{
  if (form < LOS_FORM_FP64 || form > LOS_FORM_FP32C) return _FAILURE_;
  g_default_form = form;
  return _SUCCESS_;
}

int los_set_default_device(int device)
{
  g_default_device = device;
  return _SUCCESS_;
}

/* Offsets of a few struct fields, for language bindings. */
int los_struct_offsets(long *out, int n)
{
  long v[3] = {(long)offsetof(struct transfers2, error_message),
               (long)offsetof(struct transfers2, tt2_size),
               (long)offsetof(struct transfers2, n_transfers)};
  for (int i = 0; i < n && i < 3; ++i) out[i] = v[i];
  return 3;
}

/* Byte offsets of every field of struct transfers2 in declaration order
   (include/transfer2.h:76-238), then its size; returns the number of values. */
int los_transfers2_layout(long *out, int n)
{
#define OFF(f) (long)offsetof(struct transfers2, f)
  const long v[] = {
      OFF(transfer), OFF(index_tt2_T), OFF(index_tt2_E), OFF(index_tt2_B), OFF(tt2_size),
      OFF(tt2_labels), OFF(has_cls), OFF(l_size), OFF(l), OFF(m), OFF(m_size), OFF(n_transfers),
      OFF(n_nonzero_transfers_E), OFF(n_nonzero_transfers_B), OFF(index_tt2_monopole),
      OFF(index_pt2_monopole), OFF(corresponding_index_l), OFF(corresponding_index_m),
      OFF(lm_array), OFF(k_sampling), OFF(tau_sampling), OFF(k_size_k1k2),
      OFF(k_physical_start_k1k2), OFF(k_physical_size_k1k2), OFF(k_min_k1k2), OFF(k_max_k1k2),
      OFF(k3_size_max), OFF(transfers_dir), OFF(transfers_paths), OFF(transfers_files),
      OFF(transfers_status_file), OFF(transfers_status_path), OFF(tau0), OFF(tau_rec),
      OFF(rs_rec), OFF(transfer2_verbose), OFF(error_message), OFF(count_allocated_transfers),
      OFF(count_memorised_transfers), OFF(has_allocated_transfers), OFF(stop_at_transfers2),
      (long)sizeof(struct transfers2)};
#undef OFF
  const int nv = (int)(sizeof v / sizeof v[0]);
  for (int i = 0; i < n && i < nv; ++i) out[i] = v[i];
  return nv;
}

/* ---------------------------------------------------------------------------
 * Producer-side helpers (weak: a SONG link provides its own).
 * source/perturbations2.c:457-508, :521-602, :609-636
 * ------------------------------------------------------------------------- */
__attribute__((weak)) int perturb2_allocate_k1_level(struct perturbs2 *ppt2, int index_k1)
{
  los_test(ppt2->has_allocated_sources[index_k1] == _TRUE_, ppt2->error_message,
           "the index_k1=%d level of ppt2->sources is already allocated", index_k1);
  for (int tp = 0; tp < ppt2->tp2_size; tp++) {
    los_alloc(ppt2->sources[tp][index_k1], (index_k1 + 1) * sizeof(double *), ppt2->error_message);
    for (int i2 = 0; i2 <= index_k1; ++i2) {
      long n = (long)ppt2->tau_size * ppt2->k3_size[index_k1][i2];
      los_calloc(ppt2->sources[tp][index_k1][i2], n, sizeof(double), ppt2->error_message);
      ppt2->count_allocated_sources += n;
    }
  }
  ppt2->has_allocated_sources[index_k1] = _TRUE_;
  return _SUCCESS_;
}

__attribute__((weak)) int perturb2_load_sources_from_disk(struct perturbs2 *ppt2, int index_k1)
{
  los_test(ppt2->sources_paths == NULL, ppt2->error_message, "no sources paths set");
  FILE *fp = fopen(ppt2->sources_paths[index_k1], "rb");
  los_test(fp == NULL, ppt2->error_message, "could not open %s", ppt2->sources_paths[index_k1]);
  for (int tp = 0; tp < ppt2->tp2_size; ++tp)
    for (int i2 = 0; i2 <= index_k1; ++i2) {
      long n = (long)ppt2->tau_size * ppt2->k3_size[index_k1][i2];
      long r = (long)fread(ppt2->sources[tp][index_k1][i2], sizeof(double), n, fp);
      if (r != n) {
        fclose(fp);
        los_fail(ppt2->error_message, "error reading %s: read %ld of %ld",
                 ppt2->sources_paths[index_k1], r, n);
      }
      ppt2->count_memorised_sources += n;
    }
  fclose(fp);
  return _SUCCESS_;
}

/* source/perturbations2.c:15347-15414: one file per k1, for each type and each
   k2 <= k1 the tau_size * k3_size doubles of that (k3, tau) level as they lie in
   memory ([tau][k3], k3 fastest); levels with no k3 write nothing. */
__attribute__((weak)) int perturb2_store_sources_to_disk(struct perturbs2 *ppt2, int index_k1)
{
  los_test(ppt2->sources_paths == NULL, ppt2->error_message, "no sources paths set");
  los_test(ppt2->has_allocated_sources[index_k1] == _FALSE_, ppt2->error_message,
           "the index_k1=%d level of ppt2->sources is not allocated", index_k1);
  FILE *fp = fopen(ppt2->sources_paths[index_k1], "wb");
  los_test(fp == NULL, ppt2->error_message, "could not open %s", ppt2->sources_paths[index_k1]);
  for (int tp = 0; tp < ppt2->tp2_size; ++tp)
    for (int i2 = 0; i2 <= index_k1; ++i2) {
      long n = (long)ppt2->tau_size * ppt2->k3_size[index_k1][i2];
      if (n > 0 && (long)fwrite(ppt2->sources[tp][index_k1][i2], sizeof(double), n, fp) != n) {
        fclose(fp);
        los_fail(ppt2->error_message, "error writing %s", ppt2->sources_paths[index_k1]);
      }
    }
  los_test(fclose(fp) != 0, ppt2->error_message, "error closing %s",
           ppt2->sources_paths[index_k1]);
  return _SUCCESS_;
}

/* source/perturbations2.c:1051-1064: the path of each k1's sources file,
   "<sources_dir>/sources_%03d.dat". */
int los_sources_set_paths(struct perturbs2 *ppt2)
{
  if (ppt2->sources_paths == NULL) {
    los_calloc(ppt2->sources_files, ppt2->k_size, sizeof(FILE *), ppt2->error_message);
    los_calloc(ppt2->sources_paths, ppt2->k_size, sizeof(char *), ppt2->error_message);
    for (int i1 = 0; i1 < ppt2->k_size; ++i1)
      los_alloc(ppt2->sources_paths[i1], _FILENAMESIZE_, ppt2->error_message);
  }
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    snprintf(ppt2->sources_paths[i1], _FILENAMESIZE_, "%s/sources_%03d.dat", ppt2->sources_dir,
             i1);
  return _SUCCESS_;
}

void los_sources_free_paths(struct perturbs2 *ppt2)
{
  if (ppt2->sources_paths)
    for (int i1 = 0; i1 < ppt2->k_size; ++i1) free(ppt2->sources_paths[i1]);
  free(ppt2->sources_paths);
  free(ppt2->sources_files);
  ppt2->sources_paths = NULL;
  ppt2->sources_files = NULL;
}

__attribute__((weak)) int perturb2_free_k1_level(struct perturbs2 *ppt2, int index_k1)
{
  los_test(ppt2->has_allocated_sources[index_k1] == _FALSE_, ppt2->error_message,
           "the index_k1=%d level of ppt2->sources is already free", index_k1);
  for (int tp = 0; tp < ppt2->tp2_size; tp++) {
    for (int i2 = 0; i2 <= index_k1; i2++) free(ppt2->sources[tp][index_k1][i2]);
    free(ppt2->sources[tp][index_k1]);
    ppt2->sources[tp][index_k1] = NULL;
  }
  ppt2->has_allocated_sources[index_k1] = _FALSE_;
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * Multipoles. source/transfer2.c:1081-1103, :1121-1285
 * ------------------------------------------------------------------------- */
int transfer2_get_l_list(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                         struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2)
{
  (void)ppr; (void)ppr2; (void)ppt2; (void)pbs2;
  ptr2->l_size = pbs->l_size;
  los_alloc(ptr2->l, ptr2->l_size * sizeof(int), ptr2->error_message);
  memcpy(ptr2->l, pbs->l, ptr2->l_size * sizeof(int));
  return _SUCCESS_;
}

int transfer2_get_lm_lists(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                           struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2)
{
  (void)ppr; (void)pbs; (void)pbs2;
  los_alloc(ptr2->lm_array, ptr2->l_size * sizeof(int *), ptr2->error_message);
  for (int il = 0; il < ptr2->l_size; ++il) {
    int l = ptr2->l[il];
    int n = ppr2->index_m_max[l] + 1;
    los_calloc(ptr2->lm_array[il], LOS_MAX(n, 1), sizeof(int), ptr2->error_message);
    for (int im = 0; im < n; ++im)
      ptr2->lm_array[il][im] = los_offset_indexl_indexm(ptr2->l[il], ptr2->m[im], ptr2->l,
                                                        ptr2->l_size, ptr2->m, ptr2->m_size);
  }
  los_alloc(ptr2->corresponding_index_l, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  los_alloc(ptr2->corresponding_index_m, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  los_alloc(ptr2->index_tt2_monopole, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  los_alloc(ptr2->index_pt2_monopole, ptr2->tt2_size * sizeof(int), ptr2->error_message);
  /* field blocks in the order T, E, B (source/transfer2.c:916-936) */
  struct { short has; int tt0, tp0; char tag; } fld[3] = {
      {ppt2->has_source_T, ptr2->index_tt2_T, ppt2->index_tp2_T, 'T'},
      {ppt2->has_source_E, ptr2->index_tt2_E, ppt2->index_tp2_E, 'E'},
      {ppt2->has_source_B, ptr2->index_tt2_B, ppt2->index_tp2_B, 'B'}};
  for (int tt = 0; tt < ptr2->tt2_size; tt++) {
    int il = -1, im = -1;
    for (int f = 0; f < 3; ++f) {
      if (!fld[f].has || tt < fld[f].tt0 || tt >= fld[f].tt0 + ptr2->n_transfers) continue;
      ptr2->index_pt2_monopole[tt] = fld[f].tp0;
      ptr2->index_tt2_monopole[tt] = fld[f].tt0;
      los_offset2multipole_indexl_indexm(tt - fld[f].tt0, ptr2->l, ptr2->l_size, ptr2->m,
                                         ptr2->m_size, &il, &im);
      if (il >= 0)
        snprintf(ptr2->tt2_labels[tt], _MAX_LENGTH_LABEL_, "%c_%d_%d", fld[f].tag, ptr2->l[il],
                 ptr2->m[im]);
      break;
    }
    los_test((il >= ptr2->l_size) || (im >= ptr2->m_size) || (il < 0) || (im < 0),
             ptr2->error_message,
             "index_tt=%d: result (index_l,index_m)=(%d,%d) is out of bounds", tt, il, im);
    ptr2->corresponding_index_l[tt] = il;
    ptr2->corresponding_index_m[tt] = im;
  }
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * k3 bookkeeping. source/transfer2.c:1329-1834. The fine grid is the FP64
 * recurrence k += k_step_max, reproduced operation for operation.
 * ------------------------------------------------------------------------- */
static int k3_step(struct precision *ppr, struct precision2 *ppr2, struct bessels2 *pbs2,
                   struct transfers2 *ptr2, double *k_step_max)
{
  double s = 0;
  if (ptr2->k_sampling == class_transfer2_k3_sampling) {
    s = 2. * M_PI / (ptr2->tau0 - ptr2->tau_rec) * ppr2->q_linstep_song;
    if ((ppr->load_run == _TRUE_) && (ppr2->old_run == _TRUE_))
      s = 2. * M_PI / ptr2->rs_rec * ppr2->q_linstep_song;
  } else if (ptr2->k_sampling == bessel_k3_sampling) {
    s = pbs2->xx_step / ptr2->tau0;
  }
  los_test(s == 0, ptr2->error_message, "stop to avoid infinite loop");
  *k_step_max = s;
  return _SUCCESS_;
}

int transfer2_get_k3_size(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2,
                          int index_k1, int index_k2)
{
  (void)pbs;
  los_test(index_k2 > index_k1, ptr2->error_message, "we always require k1 >= k2");
  const double *kpt = ppt2->k3[index_k1][index_k2];
  const int npt = ppt2->k3_size[index_k1][index_k2];
  const double k_min_pt = kpt[0], k_max_pt = kpt[npt - 1];
  double k_min_tr = k_min_pt, k_max_tr = k_max_pt;
  const int extrapolate = ppr->bispectra_k3_extrapolation != no_k3_extrapolation;
  if (extrapolate) {
    double range = k_max_pt - k_min_pt;
    double right = 2. * M_PI * ppr->extra_k3_oscillations_right / (ptr2->tau0 - ptr2->tau_rec);
    double left = 2. * M_PI * ppr->extra_k3_oscillations_left / (ptr2->tau0 - ptr2->tau_rec);
    if (range < right) k_max_tr += (right - range);
    k_max_tr = LOS_MIN(k_max_tr, ppt2->k[ppt2->k_size - 1]);
    if (range < left) k_min_tr -= (left - range);
    k_min_tr = LOS_MAX(k_min_tr, LOS_MIN(ppt2->k[0], k_min_pt));
    if (ppr->extra_k3_oscillations_left < 0) k_min_tr = ppt2->k[0];
    if (ppr->extra_k3_oscillations_right < 0) k_max_tr = ppt2->k[ppt2->k_size - 1];
  }
  ptr2->k_min_k1k2[index_k1][index_k2] = k_min_tr;
  ptr2->k_max_k1k2[index_k1][index_k2] = k_max_tr;
  los_test(k_max_tr > pbs2->xx_max / ptr2->tau0, ptr2->error_message,
           "xx_max is not large enough. The projection functions J(k*(tau0-tau)) do not cover "
           "the needed k-range (k_max=%g,tau0=%g,xx_max=%g)",
           k_max_tr, ptr2->tau0, pbs2->xx_max);
  double step;
  if (k3_step(ppr, ppr2, pbs2, ptr2, &step) == _FAILURE_) return _FAILURE_;

  int ipt = 0, itr = 1;
  double k = k_min_tr;
  if (extrapolate) {
    los_test(step < ppr->smallest_allowed_variation, ptr2->error_message,
             "stopping to avoid segmentation fault, step=%g", step);
    for (; k < k_min_pt; ++itr) k += step;
    if (k != k_min_pt) k = k_min_pt;
  }
  ipt = 1;
  ptr2->k_physical_start_k1k2[index_k1][index_k2] = itr - 1;
  for (; ipt < npt && (kpt[ipt] - k) < step; ++ipt, ++itr) k = kpt[ipt];
  int phys = ipt;
  while (k < k_max_tr) {
    k += step;
    ++itr;
    if ((k <= k_max_pt) && (k >= k_min_pt)) ++phys;
  }
  ptr2->k_physical_size_k1k2[index_k1][index_k2] = phys;
  ptr2->k_size_k1k2[index_k1][index_k2] = (k > k_max_tr) ? itr - 1 : itr;
  los_test(ptr2->k_size_k1k2[index_k1][index_k2] < 2, ptr2->error_message,
           "found less than 2 valid values for the k-grid of modes k1=%f, k2=%f",
           ppt2->k[index_k1], ppt2->k[index_k2]);
  los_test(!extrapolate && (phys != ptr2->k_size_k1k2[index_k1][index_k2] ||
                            ptr2->k_physical_start_k1k2[index_k1][index_k2] != 0),
           ptr2->error_message,
           "when no extrapolation is requested, all k's should be physical");
  return _SUCCESS_;
}

int transfer2_get_k3_list(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2,
                          int index_k1, int index_k2, double *k3, int *last_used_index_pt)
{
  (void)pbs;
  const double *kpt = ppt2->k3[index_k1][index_k2];
  const int npt = ppt2->k3_size[index_k1][index_k2];
  const double k_min_pt = kpt[0];
  const int ntr = ptr2->k_size_k1k2[index_k1][index_k2];
  const double k_min_tr = ptr2->k_min_k1k2[index_k1][index_k2];
  const double k_max_tr = ptr2->k_max_k1k2[index_k1][index_k2];
  double step;
  if (k3_step(ppr, ppr2, pbs2, ptr2, &step) == _FAILURE_) return _FAILURE_;

  int itr = 0;
  double k = k_min_tr;
  k3[itr++] = k_min_tr;
  if (ppr->bispectra_k3_extrapolation != no_k3_extrapolation) {
    los_test(step < ppr->smallest_allowed_variation, ptr2->error_message,
             "stopping to avoid segmentation fault, step=%g", step);
    while (k < k_min_pt) {
      k += step;
      k3[itr++] = k;
    }
    if (k != k_min_pt) {
      k = k_min_pt;
      k3[itr - 1] = k;
    }
  }
  int ipt = 1;
  for (; ipt < npt && (kpt[ipt] - k) < step; ++ipt) {
    k = kpt[ipt];
    k3[itr++] = k;
  }
  *last_used_index_pt = ipt;
  while (itr < ntr && k < k_max_tr) {
    k += step;
    k3[itr++] = k;
  }
  los_test(k3[ntr - 1] > k_max_tr, ptr2->error_message,
           "bug in k list calculation, k_max_transfers2=%.17f is larger than k_max_tr=%.17f",
           k3[ntr - 1], k_max_tr);
  /* DEBUG-build checks of the reference (source/transfer2.c:1711-1733) */
  const int first = ptr2->k_physical_start_k1k2[index_k1][index_k2];
  const int last = first + ptr2->k_physical_size_k1k2[index_k1][index_k2] - 1;
  const double k1 = ppt2->k[index_k1], k2 = ppt2->k[index_k2];
  for (int i = first; i <= last; ++i)
    los_test(k3[i] <= fabs(k1 - k2) || k3[i] >= k1 + k2, ptr2->error_message,
             "triangular condition not satisfied in the physical regime for (%d,%d)", index_k1,
             index_k2);
  for (int i = 0; i < ntr - 1; ++i)
    los_test(k3[i] >= k3[i + 1], ptr2->error_message,
             "k3-grid is not strictly ascending: k[%d]=%g >= k[%d]=%g", i, k3[i], i + 1, k3[i + 1]);
  return _SUCCESS_;
}

int transfer2_get_k3_sizes(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                           struct bessels *pbs, struct bessels2 *pbs2, struct transfers *ptr,
                           struct transfers2 *ptr2)
{
  (void)ptr;
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) {
      int n = ppt2->k3_size[i1][i2];
      los_test((n < 2 && ppr2->sources_k3_interpolation == linear_interpolation) ||
                   (n < 4 && ppr2->sources_k3_interpolation == cubic_interpolation),
               ptr2->error_message,
               "index_k1=%d, index_k2=%d: cannot do interpolation in k3 with just k3_size=%d "
               "values. Increase k3_size.", i1, i2, n);
    }
  const int nk = ppt2->k_size;
  los_alloc(ptr2->k_size_k1k2, nk * sizeof(int *), ptr2->error_message);
  los_alloc(ptr2->k_physical_start_k1k2, nk * sizeof(int *), ptr2->error_message);
  los_alloc(ptr2->k_physical_size_k1k2, nk * sizeof(int *), ptr2->error_message);
  los_alloc(ptr2->k_min_k1k2, nk * sizeof(double *), ptr2->error_message);
  los_alloc(ptr2->k_max_k1k2, nk * sizeof(double *), ptr2->error_message);
  for (int i1 = 0; i1 < nk; ++i1) {
    los_alloc(ptr2->k_size_k1k2[i1], (i1 + 1) * sizeof(int), ptr2->error_message);
    los_alloc(ptr2->k_physical_start_k1k2[i1], (i1 + 1) * sizeof(int), ptr2->error_message);
    los_alloc(ptr2->k_physical_size_k1k2[i1], (i1 + 1) * sizeof(int), ptr2->error_message);
    los_alloc(ptr2->k_min_k1k2[i1], (i1 + 1) * sizeof(double), ptr2->error_message);
    los_alloc(ptr2->k_max_k1k2[i1], (i1 + 1) * sizeof(double), ptr2->error_message);
    for (int i2 = 0; i2 <= i1; ++i2)
      if (transfer2_get_k3_size(ppr, ppr2, ppt2, pbs, pbs2, ptr2, i1, i2) == _FAILURE_)
        return _FAILURE_;
  }
  ptr2->k3_size_max = 0;
  for (int i1 = 0; i1 < nk; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2)
      ptr2->k3_size_max = LOS_MAX(ptr2->k3_size_max, ptr2->k_size_k1k2[i1][i2]);
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * Indices and output levels. source/transfer2.c:866-1067, :636-691, :2976-3104
 * ------------------------------------------------------------------------- */
int transfer2_indices_of_transfers(struct precision *ppr, struct precision2 *ppr2,
                                   struct perturbs2 *ppt2, struct bessels *pbs,
                                   struct bessels2 *pbs2, struct transfers *ptr,
                                   struct transfers2 *ptr2)
{
  los_call(transfer2_get_l_list(ppr, ppr2, ppt2, pbs, pbs2, ptr2), ptr2->error_message,
           ptr2->error_message);
  ptr2->m_size = ppr2->m_size;
  los_alloc(ptr2->m, ptr2->m_size * sizeof(int), ptr2->error_message);
  memcpy(ptr2->m, ppr2->m, ptr2->m_size * sizeof(int));

  int index_tt = 0;
  ptr2->n_transfers = los_size_indexl_indexm(ptr2->l, ptr2->l_size, ppr2->m, ppr2->m_size);
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
  los_calloc(ptr2->tt2_labels, (size_t)ptr2->tt2_size * _MAX_LENGTH_LABEL_, sizeof(char),
             ptr2->error_message);
  if (ptr2->transfer2_verbose > 1)
    printf(" -> will compute tt2_size=%d transfer functions\n", ptr2->tt2_size);

  los_call(transfer2_get_lm_lists(ppr, ppr2, ppt2, pbs, pbs2, ptr2), ptr2->error_message,
           ptr2->error_message);
  los_call(transfer2_get_k3_sizes(ppr, ppr2, ppt2, pbs, pbs2, ptr, ptr2), ptr2->error_message,
           ptr2->error_message);

  ptr2->count_memorised_transfers = 0;
  ptr2->count_allocated_transfers = 0;
  los_alloc(ptr2->transfer, ptr2->tt2_size * sizeof(double ***), ptr2->error_message);
  for (int tt = 0; tt < ptr2->tt2_size; ++tt)
    los_calloc(ptr2->transfer[tt], ppt2->k_size, sizeof(double **), ptr2->error_message);
  los_calloc(ptr2->has_allocated_transfers, ppt2->k_size, sizeof(short), ptr2->error_message);

  if (ppr2->store_transfers_to_disk || ppr2->load_transfers_from_disk) {
    los_alloc(ptr2->transfers_files, ptr2->tt2_size * sizeof(FILE *), ptr2->error_message);
    los_alloc(ptr2->transfers_paths, ptr2->tt2_size * sizeof(char *), ptr2->error_message);
    for (int tt = 0; tt < ptr2->tt2_size; ++tt) {
      los_alloc(ptr2->transfers_paths[tt], _FILENAMESIZE_ * sizeof(char), ptr2->error_message);
      snprintf(ptr2->transfers_paths[tt], _FILENAMESIZE_, "%s/transfers_%03d.dat",
               ptr2->transfers_dir, tt);
      ptr2->transfers_files[tt] = NULL;
      if (ppr2->store_transfers_to_disk) {
        ptr2->transfers_files[tt] = fopen(ptr2->transfers_paths[tt], "wb");
        los_test(ptr2->transfers_files[tt] == NULL, ptr2->error_message,
                 "could not open %s for writing", ptr2->transfers_paths[tt]);
      }
    }
  }
  return _SUCCESS_;
}

int transfer2_allocate_type_level(struct perturbs2 *ppt2, struct transfers2 *ptr2, int index_tt)
{
  los_alloc(ptr2->transfer[index_tt], ppt2->k_size * sizeof(double **), ptr2->error_message);
  for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
    los_alloc(ptr2->transfer[index_tt][i1], (i1 + 1) * sizeof(double *), ptr2->error_message);
    for (int i2 = 0; i2 <= i1; ++i2) {
      los_alloc(ptr2->transfer[index_tt][i1][i2], ptr2->k_size_k1k2[i1][i2] * sizeof(double),
                ptr2->error_message);
      ptr2->count_allocated_transfers += ptr2->k_size_k1k2[i1][i2];
    }
  }
  return _SUCCESS_;
}

int transfer2_free_type_level(struct perturbs2 *ppt2, struct transfers2 *ptr2, int index_tt)
{
  for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
    for (int i2 = 0; i2 <= i1; ++i2) free(ptr2->transfer[index_tt][i1][i2]);
    free(ptr2->transfer[index_tt][i1]);
  }
  free(ptr2->transfer[index_tt]);
  return _SUCCESS_;
}

int transfer2_allocate_k1_level(struct perturbs2 *ppt2, struct transfers2 *ptr2, int index_k1)
{
  los_test(ptr2->has_allocated_transfers[index_k1] == _TRUE_, ptr2->error_message,
           "the index_k1=%d level of ptr2->transfers is already allocated, stop to prevent error",
           index_k1);
  for (int tt = 0; tt < ptr2->tt2_size; ++tt) {
    los_alloc(ptr2->transfer[tt][index_k1], (index_k1 + 1) * sizeof(double *),
              ptr2->error_message);
    for (int i2 = 0; i2 <= index_k1; ++i2) {
      los_alloc(ptr2->transfer[tt][index_k1][i2],
                ptr2->k_size_k1k2[index_k1][i2] * sizeof(double), ptr2->error_message);
      ptr2->count_allocated_transfers += ptr2->k_size_k1k2[index_k1][i2];
    }
  }
  ptr2->has_allocated_transfers[index_k1] = _TRUE_;
  return _SUCCESS_;
}

int transfer2_free_k1_level(struct perturbs2 *ppt2, struct transfers2 *ptr2, int index_k1)
{
  (void)ppt2;
  los_test(ptr2->has_allocated_transfers[index_k1] == _FALSE_, ptr2->error_message,
           "the index_k1=%d level of ptr2->transfers is already free, stop to prevent error",
           index_k1);
  for (int tt = 0; tt < ptr2->tt2_size; tt++) {
    for (int i2 = 0; i2 <= index_k1; ++i2) free(ptr2->transfer[tt][index_k1][i2]);
    free(ptr2->transfer[tt][index_k1]);
    ptr2->transfer[tt][index_k1] = NULL;
  }
  ptr2->has_allocated_transfers[index_k1] = _FALSE_;
  return _SUCCESS_;
}

int transfer2_store_transfers_to_disk(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                      int index_k1)
{
  (void)ppt2;
  for (int tt = 0; tt < ptr2->tt2_size; tt++)
    for (int i2 = 0; i2 <= index_k1; ++i2) {
      size_t n = ptr2->k_size_k1k2[index_k1][i2];
      los_test(fwrite(ptr2->transfer[tt][index_k1][i2], sizeof(double), n,
                      ptr2->transfers_files[tt]) != n,
               ptr2->error_message, "could not write %s", ptr2->transfers_paths[tt]);
    }
  return _SUCCESS_;
}

int transfer2_load_transfers_from_disk(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                       int index_tt)
{
  los_call(transfer2_allocate_type_level(ppt2, ptr2, index_tt), ptr2->error_message,
           ptr2->error_message);
  FILE *fp = fopen(ptr2->transfers_paths[index_tt], "rb");
  los_test(fp == NULL, ptr2->error_message, "could not open %s", ptr2->transfers_paths[index_tt]);
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) {
      int n = ptr2->k_size_k1k2[i1][i2];
      int r = (int)fread(ptr2->transfer[index_tt][i1][i2], sizeof(double), n, fp);
      if (r != n) {
        fclose(fp);
        los_fail(ptr2->error_message,
                 "Could not read in '%s' file, read %d entries but expected %d "
                 "(index_tt=%d,index_k1=%d,index_k2=%d)",
                 ptr2->transfers_paths[index_tt], r, n, index_tt, i1, i2);
      }
      ptr2->count_memorised_transfers += n;
    }
  fclose(fp);
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * Time grid. source/transfer2.c:2366-2585
 * ------------------------------------------------------------------------- */
int transfer2_get_time_grid(struct precision *ppr, struct precision2 *ppr2, struct perturbs *ppt,
                            struct perturbs2 *ppt2, struct bessels *pbs, struct bessels2 *pbs2,
                            struct transfers2 *ptr2, int index_k1, int index_k2, int index_k,
                            struct transfer2_workspace *pw)
{
  (void)ppr; (void)ppt; (void)pbs; (void)index_k1; (void)index_k2; (void)index_k;
  const int npt = ppt2->tau_size;
  const double *tpt = ppt2->tau_sampling;
  if (ptr2->tau_sampling == sources_tau_sampling) {
    pw->tau_grid_size = npt;
    memcpy(pw->tau_grid, tpt, npt * sizeof(double));
  } else {
    double step = 0;
    if (ptr2->tau_sampling == custom_tau_sampling) step = 2 * M_PI / pw->k * ppr2->tau_linstep_song;
    else if (ptr2->tau_sampling == bessel_tau_sampling) step = pbs2->xx_step / pw->k;
    los_test(step == 0, ptr2->error_message, "stopping to prevent segmentation fault");
    int itr = 0;
    double tau = tpt[0];
    pw->tau_grid[itr] = tau;
    pw->index_tau_left[itr] = 0;
    itr++;
    for (int ipt = 1; ipt < npt; ++ipt) {
      while ((tpt[ipt] - tau) > step) {
        tau += step;
        pw->tau_grid[itr] = tau;
        pw->index_tau_left[itr] = ipt;
        itr++;
      }
      tau = tpt[ipt];
      pw->tau_grid[itr] = tau;
      pw->index_tau_left[itr] = ipt;
      itr++;
    }
    pw->tau_grid_size = itr;
    for (int i = 0; i < itr; ++i)
      if (pw->index_tau_left[i] == npt - 1) pw->index_tau_left[i] = npt - 2;
    for (int i = 0; i < itr - 1; ++i)
      los_test(pw->tau_grid[i] >= pw->tau_grid[i + 1], ptr2->error_message,
               "integration grid is not sorted, bug in the code");
  }
  const int n = pw->tau_grid_size;
  for (int i = 0; i < n; ++i) pw->tau0_minus_tau[i] = ptr2->tau0 - pw->tau_grid[i];
  pw->delta_tau[0] = pw->tau_grid[1] - pw->tau_grid[0];
  for (int i = 1; i < n - 1; ++i) pw->delta_tau[i] = pw->tau_grid[i + 1] - pw->tau_grid[i - 1];
  pw->delta_tau[n - 1] = pw->tau_grid[n - 1] - pw->tau_grid[n - 2];
  return _SUCCESS_;
}

/* ---------------------------------------------------------------------------
 * Per-element stages. In this library the source interpolation in k3 and
 * time, the J interpolation and the tau integral run fused in the GPU kernels
 * that transfer2_init launches; these symbols exist so that code linking
 * against SONG's header still links, and they report that clearly.
 * ------------------------------------------------------------------------- */
#define FUSED_MSG "runs fused on the GPU inside transfer2_init(); this library has no per-element CPU path"

int transfer2_interpolate_sources_in_k(struct precision *ppr, struct precision2 *ppr2,
                                       struct perturbs *ppt, struct perturbs2 *ppt2,
                                       struct bessels *pbs, struct bessels2 *pbs2,
                                       struct transfers2 *ptr2, int index_k1, int index_k2,
                                       int index_tp, double *k_grid, double *splines,
                                       double *interpolated_sources)
{
  (void)ppr; (void)ppr2; (void)ppt; (void)ppt2; (void)pbs; (void)pbs2; (void)index_k1;
  (void)index_k2; (void)index_tp; (void)k_grid; (void)splines; (void)interpolated_sources;
  los_fail(ptr2->error_message, "transfer2_interpolate_sources_in_k " FUSED_MSG);
}

int transfer2_interpolate_sources_in_time(struct precision *ppr, struct precision2 *ppr2,
                                          struct perturbs *ppt, struct perturbs2 *ppt2,
                                          struct bessels *pbs, struct bessels2 *pbs2,
                                          struct transfers2 *ptr2, int index_tp2,
                                          double *interpolated_sources_in_k,
                                          double *sources_time_spline,
                                          double *interpolated_sources_in_time,
                                          struct transfer2_workspace *pw)
{
  (void)ppr; (void)ppr2; (void)ppt; (void)ppt2; (void)pbs; (void)pbs2; (void)index_tp2;
  (void)interpolated_sources_in_k; (void)sources_time_spline; (void)interpolated_sources_in_time;
  (void)pw;
  los_fail(ptr2->error_message, "transfer2_interpolate_sources_in_time " FUSED_MSG);
}

int transfer2_compute(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2,
                      int index_k1, int index_k2, int index_k, int index_l, int index_m,
                      int index_tt, double **interpolated_sources_in_time,
                      struct transfer2_workspace *pw)
{
  (void)ppr; (void)ppr2; (void)ppt2; (void)pbs; (void)pbs2; (void)index_k1; (void)index_k2;
  (void)index_k; (void)index_l; (void)index_m; (void)index_tt; (void)interpolated_sources_in_time;
  (void)pw;
  los_fail(ptr2->error_message, "transfer2_compute " FUSED_MSG);
}

int transfer2_integrate(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                        struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2,
                        int index_k1, int index_k2, int index_k, int index_l, int index_m,
                        double **interpolated_sources_in_time, int index_J,
                        int index_source_monopole, struct transfer2_workspace *pw,
                        double *integral)
{
  (void)ppr; (void)ppr2; (void)ppt2; (void)pbs; (void)pbs2; (void)index_k1; (void)index_k2;
  (void)index_k; (void)index_l; (void)index_m; (void)interpolated_sources_in_time;
  (void)index_J; (void)index_source_monopole; (void)pw; (void)integral;
  los_fail(ptr2->error_message, "transfer2_integrate " FUSED_MSG);
}

/* ---------------------------------------------------------------------------
 * The flat descriptor of the GPU side (los_stage.h). This unit is the only
 * one that reads SONG's structures; recompiled against SONG's headers, it
 * adapts the whole library to their layout.
 * ------------------------------------------------------------------------- */
typedef struct {
  struct precision *ppr;
  struct precision2 *ppr2;
  struct perturbs2 *ppt2;
  struct bessels *pbs;
  struct bessels2 *pbs2;
  struct transfers2 *ptr2;
} song_structs;

static int stage_k3_list(const los_stage *st, int i1, int i2, double *out, char *err)
{
  const song_structs *q = (const song_structs *)st->song;
  int last_used = 0;
  if (transfer2_get_k3_list(q->ppr, q->ppr2, q->ppt2, q->pbs, q->pbs2, q->ptr2, i1, i2, out,
                            &last_used) == _FAILURE_) {
    memcpy(err, q->ptr2->error_message, _ERRORMSGSIZE_);
    return _FAILURE_;
  }
  return _SUCCESS_;
}

static void stage_fill(los_stage *st, song_structs *q)
{
  struct precision *ppr = q->ppr;
  struct precision2 *ppr2 = q->ppr2;
  struct perturbs2 *ppt2 = q->ppt2;
  struct bessels2 *pbs2 = q->pbs2;
  struct transfers2 *ptr2 = q->ptr2;
  memset(st, 0, sizeof *st);
  st->tau_sources = ptr2->tau_sampling == sources_tau_sampling;
  st->linear_J = ppr->bessels_interpolation == linear_interpolation;
  st->cubic_k3 = ppr2->sources_k3_interpolation == cubic_interpolation;
  st->tau_size = ppt2->tau_size;
  st->tau = ppt2->tau_sampling;
  st->tau0 = ptr2->tau0;
  st->has_T = ppt2->has_source_T;
  st->has_E = ppt2->has_source_E;
  st->has_B = ppt2->has_source_B;
  st->LmaxT = ppr2->l_max_los_t;
  st->LmaxP = ppr2->l_max_los_p;
  st->m_max = ppr2->m_max_song;
  st->largest_l = ppt2->largest_l;
  st->m_size = ppr2->m_size;
  st->m = ppr2->m;
  st->index_m_max = ppr2->index_m_max;
  st->l_size = ptr2->l_size;
  st->l = ptr2->l;
  st->lm_array = ptr2->lm_array;
  st->src_lm = ppt2->lm_array;
  st->tt2_size = ptr2->tt2_size;
  st->n_transfers = ptr2->n_transfers;
  st->index_tt2_T = ptr2->index_tt2_T;
  st->index_tt2_E = ptr2->index_tt2_E;
  st->index_tt2_B = ptr2->index_tt2_B;
  st->tp2_size = ppt2->tp2_size;
  st->index_tp2_T = ppt2->index_tp2_T;
  st->index_tp2_E = ppt2->index_tp2_E;
  st->index_tp2_B = ppt2->index_tp2_B;
  st->k_size = ppt2->k_size;
  st->sources = ppt2->sources;
  st->has_sources = ppt2->has_allocated_sources;
  st->k3 = ppt2->k3;
  st->k3_size = ppt2->k3_size;
  st->k3_size_max = ptr2->k3_size_max;
  st->k_size_k1k2 = ptr2->k_size_k1k2;
  st->k_phys_start = ptr2->k_physical_start_k1k2;
  st->k_phys_size = ptr2->k_physical_size_k1k2;
  st->k_max_k1k2 = ptr2->k_max_k1k2;
  st->k3_list = stage_k3_list;
  st->song = q;
  st->J_id = pbs2;
  st->J_size = pbs2->J_size;
  st->L_size = pbs2->L_size;
  st->xx_size = pbs2->xx_size;
  st->xx_step = pbs2->xx_step;
  st->xx_max = pbs2->xx_max;
  st->xx = pbs2->xx;
  st->jTT = pbs2->has_J_TT ? pbs2->index_J_TT : -1;
  st->jEE = pbs2->has_J_EE ? pbs2->index_J_EE : -1;
  st->jEB = pbs2->has_J_EB ? pbs2->index_J_EB : -1;
  st->xmin_J = pbs2->index_xmin_J;
  st->x_size_J = pbs2->x_size_J;
  st->J = pbs2->J_Llm_x;
  st->ddJ = st->linear_J ? NULL : pbs2->ddJ_Llm_x;
  st->transfer = ptr2->transfer;
  st->has_transfers = ptr2->has_allocated_transfers;
  st->count_memorised = &ptr2->count_memorised_transfers;
}

/* The engine and server entry points of include/song_los.h that take SONG's
   structures: the descriptor, then the engine's own code. */
int los_engine_load_pairs(los_engine *eng, struct precision *ppr, struct precision2 *ppr2,
                          struct perturbs2 *ppt2, struct bessels *pbs, struct bessels2 *pbs2,
                          struct transfers2 *ptr2, const int *index_k1, const int *index_k2,
                          int npairs, char *errmsg)
{
  song_structs q = {ppr, ppr2, ppt2, pbs, pbs2, ptr2};
  los_stage st;
  stage_fill(&st, &q);
  return los_engine_load_stage(eng, &st, index_k1, index_k2, npairs, errmsg);
}

int los_engine_load(los_engine *eng, struct precision *ppr, struct precision2 *ppr2,
                    struct perturbs2 *ppt2, struct bessels *pbs, struct bessels2 *pbs2,
                    struct transfers2 *ptr2, int k1_begin, int k1_end, char *errmsg)
{
  const int n = k1_end > k1_begin ? (k1_end * (k1_end + 1) - k1_begin * (k1_begin + 1)) / 2 : 0;
  int *k1 = NULL, *k2 = NULL;
  los_alloc(k1, (size_t)(n > 0 ? n : 1) * sizeof(int), errmsg);
  k2 = malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  if (k2 == NULL) {
    free(k1);
    los_fail(errmsg, "los_engine_load: out of memory");
  }
  int c = 0;
  for (int i1 = k1_begin; i1 < k1_end; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) {
      k1[c] = i1;
      k2[c++] = i2;
    }
  const int rc = los_engine_load_pairs(eng, ppr, ppr2, ppt2, pbs, pbs2, ptr2, k1, k2, c, errmsg);
  free(k1);
  free(k2);
  return rc;
}

int los_engine_store(los_engine *eng, struct transfers2 *ptr2, char *errmsg)
{
  los_stage st;
  memset(&st, 0, sizeof st);
  st.transfer = ptr2->transfer;
  st.has_transfers = ptr2->has_allocated_transfers;
  st.count_memorised = &ptr2->count_memorised_transfers;
  return los_engine_store_stage(eng, &st, errmsg);
}

int los_server_create(int device, int form, double hbm_budget_bytes, int n_slots,
                      struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2,
                      los_server **out, char *errmsg)
{
  song_structs q = {ppr, ppr2, ppt2, pbs, pbs2, ptr2};
  los_stage st;
  stage_fill(&st, &q);
  return los_server_create_stage(device, form, hbm_budget_bytes, n_slots, &st, out, errmsg);
}

/* Device bytes of one k1 batch's pairs beyond the resident J tables: the
   sources three times (as SONG lays them out, re-laid out, spline second
   derivatives), and per fine k3 the outputs, the per-(k3, tau) tables of the
   sweep and k-lane kernels (the autotuner may build both) and small lists.
   SONG_LOS_BATCH_BYTES in the environment replaces the budget (bytes), so that
   smaller batches can be asked for. */
static double batch_bytes(const los_stage *st, int k1_begin, int k1_end)
{
  const int NL = (st->LmaxT > st->LmaxP ? st->LmaxT : st->LmaxP) + 1;
  double b = 0;
  for (int i1 = k1_begin; i1 < k1_end; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) {
      const double nf = st->k_size_k1k2[i1][i2];
      b += 3.0 * 8.0 * st->tp2_size * (double)st->tau_size * st->k3_size[i1][i2];
      b += nf * (8.0 * st->tt2_size + 64.0);
      b += nf * st->tau_size * (16.0 + 8.0 * 3 * st->m_size * NL);  /* sweep tables */
      b += nf * st->tau_size * (16.0 + 8.0 * st->tp2_size);         /* k-lane tables */
    }
  return b;
}

static double batch_budget(los_engine *eng)
{
  const char *v = getenv("SONG_LOS_BATCH_BYTES");
  if (v != NULL && atof(v) > 0) return atof(v);
  return los_engine_batch_budget(eng);
}

/* ---------------------------------------------------------------------------
 * transfer2_init / transfer2_free. source/transfer2.c:79-621, :769-838
 * ------------------------------------------------------------------------- */
int transfer2_init(struct precision *ppr, struct precision2 *ppr2, struct background *pba,
                   struct thermo *pth, struct perturbs *ppt, struct perturbs2 *ppt2,
                   struct bessels *pbs, struct bessels2 *pbs2, struct transfers *ptr,
                   struct transfers2 *ptr2)
{
  if (!ppt2->has_perturbations2 ||
      (!ppt2->has_cmb_spectra && !ppt2->has_cmb_bispectra && !ptr2->stop_at_transfers2)) {
    ptr2->has_cls = _FALSE_;
    if (ptr2->transfer2_verbose > 0) printf("Second-order transfer module skipped.\n");
    return _SUCCESS_;
  }
  ptr2->has_cls = _TRUE_;
  if (ptr2->transfer2_verbose > 0) printf("Computing second-order transfer functions\n");
  if ((ppt2->k_size < 4) && (ptr2->stop_at_transfers2 == _FALSE_) && ptr2->transfer2_verbose > 0)
    printf("# WARNING: cannot do cubic interpolation in k1 with less than 4 values.\n");
  ptr2->tau0 = pba->conformal_age;
  ptr2->tau_rec = pth->tau_rec;
  ptr2->rs_rec = pth->rs_rec;

  los_call(transfer2_indices_of_transfers(ppr, ppr2, ppt2, pbs, pbs2, ptr, ptr2),
           ptr2->error_message, ptr2->error_message);
  if (ppr2->load_transfers_from_disk == _TRUE_) {
    if (ptr2->transfer2_verbose > 0)
      printf(" -> leaving transfer2 module; transfer functions will be read from disk\n");
    return _SUCCESS_;
  }

  los_engine *eng = NULL;
  los_call(los_engine_create(g_default_device, &eng, ptr2->error_message), ptr2->error_message,
           ptr2->error_message);
  if (los_engine_set_form(eng, g_default_form, ptr2->error_message) == _FAILURE_) {
    los_engine_destroy(eng);
    return _FAILURE_;
  }

  /* k1 batches sized by device memory */
  song_structs q = {ppr, ppr2, ppt2, pbs, pbs2, ptr2};
  los_stage st;
  stage_fill(&st, &q);
  const int nk = ppt2->k_size;
  int k1_begin = 0;
  int status = _SUCCESS_;
  while (k1_begin < nk && status == _SUCCESS_) {
    int k1_end = k1_begin + 1;
    double budget = batch_budget(eng);
    double bytes = batch_bytes(&st, k1_begin, k1_end);
    while (k1_end < nk && bytes + batch_bytes(&st, k1_end, k1_end + 1) < budget) {
      bytes += batch_bytes(&st, k1_end, k1_end + 1);
      k1_end++;
    }
    for (int i1 = k1_begin; i1 < k1_end && status == _SUCCESS_; ++i1) {
      if (transfer2_allocate_k1_level(ppt2, ptr2, i1) == _FAILURE_) status = _FAILURE_;
      else if (ppr2->load_sources_from_disk || ppr2->store_sources_to_disk) {
        if (perturb2_allocate_k1_level(ppt2, i1) == _FAILURE_ ||
            perturb2_load_sources_from_disk(ppt2, i1) == _FAILURE_) {
          memcpy(ptr2->error_message, ppt2->error_message, _ERRORMSGSIZE_);
          status = _FAILURE_;
        }
      }
    }
    if (status == _SUCCESS_ && ptr2->transfer2_verbose > 1)
      printf("     * computing transfer functions for index_k1 in [%d,%d)\n", k1_begin, k1_end);
    if (status == _SUCCESS_)
      status = los_engine_load(eng, ppr, ppr2, ppt2, pbs, pbs2, ptr2, k1_begin, k1_end,
                               ptr2->error_message);
    float ms = 0;
    if (status == _SUCCESS_) status = los_engine_run(eng, &ms, ptr2->error_message);
    if (status == _SUCCESS_) status = los_engine_store_stage(eng, &st, ptr2->error_message);
    for (int i1 = k1_begin; i1 < k1_end && status == _SUCCESS_; ++i1) {
      if (perturb2_free_k1_level(ppt2, i1) == _FAILURE_) {
        memcpy(ptr2->error_message, ppt2->error_message, _ERRORMSGSIZE_);
        status = _FAILURE_;
        break;
      }
      if (ppr2->store_transfers_to_disk == _TRUE_) {
        if (transfer2_store_transfers_to_disk(ppt2, ptr2, i1) == _FAILURE_ ||
            transfer2_free_k1_level(ppt2, ptr2, i1) == _FAILURE_)
          status = _FAILURE_;
      }
    }
    k1_begin = k1_end;
  }
  los_engine_destroy(eng);
  if (status == _FAILURE_) return _FAILURE_;

  if (ppr2->store_transfers_to_disk == _TRUE_)
    for (int tt = 0; tt < ptr2->tt2_size; tt++) fclose(ptr2->transfers_files[tt]);
  if (ptr2->transfer2_verbose > 1)
    printf(" -> filled ptr2->transfer with %ld values (%g MB)\n", ptr2->count_memorised_transfers,
           ptr2->count_memorised_transfers / 1e6 * 8);
  if (ppr2->load_transfers_from_disk == _FALSE_)
    los_test(ptr2->count_allocated_transfers != ptr2->count_memorised_transfers,
             ptr2->error_message, "there is a mismatch between allocated (%ld) and used (%ld) space!",
             ptr2->count_allocated_transfers, ptr2->count_memorised_transfers);
  if (ptr2->stop_at_transfers2 == _TRUE_) {
    ppt->has_cls = _FALSE_;
    ppt->has_cmb_bispectra = _FALSE_;
    ppt2->has_cmb_spectra = _FALSE_;
    ppt2->has_cmb_bispectra = _FALSE_;
  }
  return _SUCCESS_;
}

int los_transfer2_prepare(struct precision *ppr, struct precision2 *ppr2, struct background *pba,
                          struct thermo *pth, struct perturbs *ppt, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, struct transfers *ptr,
                          struct transfers2 *ptr2)
{
  (void)ppt;
  ptr2->has_cls = _TRUE_;
  ptr2->tau0 = pba->conformal_age;
  ptr2->tau_rec = pth->tau_rec;
  ptr2->rs_rec = pth->rs_rec;
  los_call(transfer2_indices_of_transfers(ppr, ppr2, ppt2, pbs, pbs2, ptr, ptr2),
           ptr2->error_message, ptr2->error_message);
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    los_call(transfer2_allocate_k1_level(ppt2, ptr2, i1), ptr2->error_message,
             ptr2->error_message);
  return _SUCCESS_;
}

int transfer2_free(struct precision2 *ppr2, struct perturbs2 *ppt2, struct transfers2 *ptr2)
{
  if (ptr2->has_cls != _TRUE_) return _SUCCESS_;
  free(ptr2->tt2_labels);
  if (!ppr2->store_transfers_to_disk && !ppr2->load_transfers_from_disk) {
    for (int i1 = 0; i1 < ppt2->k_size; ++i1)
      if (ptr2->has_allocated_transfers[i1] == _TRUE_)
        los_call(transfer2_free_k1_level(ppt2, ptr2, i1), ptr2->error_message,
                 ptr2->error_message);
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
