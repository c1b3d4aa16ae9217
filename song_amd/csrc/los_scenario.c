/*
 * los_scenario.c — seeded synthetic inputs for the LOS stage, laid out in
 * SONG's own structures.
 *
 * What a SONG run would have produced before transfer2_init (main/song.c:32-100):
 *  - precision/precision2 parameters (source/input2.c:1420-1543, :1699-1737,
 *    :1808-1918, :1975-2016 for the derived m-lists and xx_max);
 *  - the (k1,k2) grid and the triangle-limited coarse k3 grids with the rules
 *    of source/perturbations2.c:1965-1984 (linear k3 sampling);
 *  - the source type indices and lm_array of source/perturbations2.c:833-930,
 *    :1149-1161;
 *  - the LOS sources S_Lm(k1,k2,k3,tau) from a seeded model (this replaces the
 *    second-order Boltzmann solver, which is not on the LOS path);
 *  - the J_Llm(x) tables from los_bessel2_build.
 * The RNG is counter-based (splitmix64 + Box-Muller) so that any two
 * scenarios with equal configuration hold bit-identical tables.
 */
#include "los_internal.h"

#include <math.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct los_scenario {
  los_config cfg;
  struct precision pr;
  struct precision2 pr2;
  struct background ba;
  struct thermo th;
  struct perturbs pt;
  struct perturbs2 pt2;
  struct bessels bs;
  struct bessels2 bs2;
  struct transfers tr;
  struct transfers2 tr2;
};

void los_config_default(los_config *c)
{
  memset(c, 0, sizeof *c);
  c->has_T = 1;
  c->has_E = 0;
  c->m_list[0] = 0;
  c->m_size = 1;
  c->l_max_los_t = 4;
  c->l_max_los_p = 4;
  c->k_size = 4;
  c->k_min = 1e-4;
  c->k_max = 0.05;
  c->k_log = 1;
  c->k3_size = 12;
  c->tau_size = 64;
  c->tau_min = 200.0;
  c->tau_max = 14000.0;
  c->tau0 = 14200.0;
  c->tau_rec = 280.0;
  c->rs_rec = 145.0;
  c->l_size = 8;
  c->l_max = 200;
  c->dk_fine = 2e-3;
  c->xx_max = 1000.0;
  c->xx_step = 0.3;
  c->J_cut = 1e-10;
  c->j_cut = 1e-12;
  c->k3_interpolation = cubic_interpolation;
  c->bessels_interpolation = linear_interpolation;
  c->tau_sampling = sources_tau_sampling;
  c->time_interpolation = linear_interpolation;
  c->tau_linstep = 0.1;
  c->k3_sampling = class_transfer2_k3_sampling;
  c->extrapolation = no_k3_extrapolation;
  c->extra_osc_left = 0.0;
  c->extra_osc_right = 0.0;
  c->seed = 20260601ULL;
  c->tau_peak = 280.0;
  c->tau_sigma = 20.0;
  c->ramp_amp = 0.1;
  c->ramp_center = 7000.0;
  c->ramp_width = 2000.0;
  c->verbose = 0;
}

static int cmp_double(const void *a, const void *b)
{
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static unsigned long long splitmix64(unsigned long long x)
{
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

/* N(0,1) for the counter (seed, field, L, m, k1, k2). */
static double gauss_amplitude(unsigned long long seed, int field, int L, int m, int k1, int k2)
{
  unsigned long long key = seed;
  key = splitmix64(key ^ (unsigned long long)(field + 1));
  key = splitmix64(key ^ ((unsigned long long)(L + 1) << 8));
  key = splitmix64(key ^ ((unsigned long long)(m + 1) << 16));
  key = splitmix64(key ^ ((unsigned long long)(k1 + 1) << 24));
  key = splitmix64(key ^ ((unsigned long long)(k2 + 1) << 40));
  unsigned long long a = splitmix64(key), b = splitmix64(key ^ 0xA5A5A5A5ULL);
  double u1 = ((a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

static int build_l_list(const los_config *c, struct bessels *pbs, char *err)
{
  pbs->l_size = c->l_size;
  los_alloc(pbs->l, c->l_size * sizeof(int), err);
  if (c->l_size == 1) { pbs->l[0] = c->l_max; pbs->l_max = c->l_max; return _SUCCESS_; }
  double ratio = log((double)c->l_max / 2.0) / (c->l_size - 1);
  for (int i = 0; i < c->l_size; ++i) {
    int v = (int)floor(2.0 * exp(ratio * i) + 0.5);
    if (i > 0 && v <= pbs->l[i - 1]) v = pbs->l[i - 1] + 1;
    pbs->l[i] = v;
  }
  pbs->l[c->l_size - 1] = c->l_max;
  los_test(pbs->l[c->l_size - 2] >= c->l_max, err,
           "l_size=%d too large for l_max=%d", c->l_size, c->l_max);
  pbs->l_max = c->l_max;
  return _SUCCESS_;
}

int los_scenario_create(const los_config *cfg, los_scenario **out, char *err)
{
  los_scenario *s = calloc(1, sizeof *s);
  los_test(s == NULL, err, "out of memory");
  s->cfg = *cfg;
  const los_config *c = &s->cfg;
  struct precision *ppr = &s->pr;
  struct precision2 *ppr2 = &s->pr2;
  struct perturbs2 *ppt2 = &s->pt2;
  struct bessels *pbs = &s->bs;
  struct bessels2 *pbs2 = &s->bs2;
  struct transfers2 *ptr2 = &s->tr2;

  /* ---------------- precision ---------------- */
  ppr->bessels_interpolation = c->bessels_interpolation;
  ppr->bispectra_k3_extrapolation = c->extrapolation;
  ppr->extra_k3_oscillations_left = c->extra_osc_left;
  ppr->extra_k3_oscillations_right = c->extra_osc_right;
  ppr->smallest_allowed_variation = 1e-14;
  ppr->load_run = _FALSE_;
  ppr->k_max_tau0_over_l_max = 2.0;

  los_test(c->m_size < 1 || c->m_size > _MAX_NUM_AZIMUTHAL_, err, "bad m_size");
  ppr2->m_size = c->m_size;
  for (int i = 0; i < c->m_size; ++i) ppr2->m[i] = c->m_list[i];
  for (int i = 0; i + 1 < c->m_size; ++i)
    los_test(ppr2->m[i] >= ppr2->m[i + 1], err, "m-list not strictly ascending");
  los_test(ppr2->m[0] < 0, err, "negative m");
  ppr2->m_max_song = ppr2->m[ppr2->m_size - 1];
  for (int m = 0; m < _MAX_NUM_AZIMUTHAL_; ++m) {
    ppr2->index_m[m] = -1;
    ppr2->compute_m[m] = _FALSE_;
  }
  for (int m = 0; m <= ppr2->m_max_song; ++m)
    for (int im = 0; im < ppr2->m_size; ++im)
      if (m == ppr2->m[im]) { ppr2->index_m[m] = im; ppr2->compute_m[m] = _TRUE_; }
  ppr2->l_max_los_t = c->has_T ? c->l_max_los_t : 0;
  ppr2->l_max_los_p = c->has_E ? c->l_max_los_p : 0;
  ppr2->l_max_los = LOS_MAX(ppr2->l_max_los_t, ppr2->l_max_los_p);
  ppr2->tau_linstep_song = c->tau_linstep;
  ppr2->k_max_tau0_over_l_max = 2.0;
  ppr2->k3_size = c->k3_size;
  ppr2->q_linstep_song = c->dk_fine * (c->tau0 - c->tau_rec) / (2.0 * M_PI);
  ppr2->sources_time_interpolation = c->time_interpolation;
  ppr2->sources_k3_interpolation = c->k3_interpolation;
  ppr2->bessel_j_cut_song = c->j_cut;
  ppr2->bessel_J_cut_song = c->J_cut;
  ppr2->bessel_x_step_song = c->xx_step;

  if (build_l_list(c, pbs, err) == _FAILURE_) return _FAILURE_;
  /* index_m_max for l up to l_max + max(l_max_los, m_max) (source/input2.c:1890-1913) */
  int lmx = pbs->l_max + LOS_MAX(ppr2->l_max_los, ppr2->m_max_song) + 1;
  ppr2->index_m_max_size = lmx + 1;
  los_alloc(ppr2->index_m_max, (lmx + 1) * sizeof(int), err);
  for (int l = 0; l <= lmx; ++l) {
    if (l < ppr2->m[0]) { ppr2->index_m_max[l] = -1; continue; }
    int im = ppr2->m_size - 1;
    while (ppr2->m[im] > l) im--;
    ppr2->index_m_max[l] = im;
  }

  s->ba.conformal_age = c->tau0;
  s->th.tau_rec = c->tau_rec;
  s->th.rs_rec = c->rs_rec;
  s->pt.has_cls = _TRUE_;
  s->pt.has_cmb_bispectra = _TRUE_;
  s->pt.l_scalar_max = pbs->l_max;

  /* ---------------- perturbs2 ---------------- */
  ppt2->has_perturbations2 = _TRUE_;
  ppt2->has_cmb_spectra = _FALSE_;
  ppt2->has_cmb_bispectra = _TRUE_;
  ppt2->has_cmb_temperature = c->has_T ? _TRUE_ : _FALSE_;
  ppt2->has_cmb_polarization_e = c->has_E ? _TRUE_ : _FALSE_;
  ppt2->has_cmb_polarization_b = _FALSE_;
  ppt2->m_size = ppr2->m_size;
  los_alloc(ppt2->m, ppt2->m_size * sizeof(int), err);
  for (int i = 0; i < ppt2->m_size; ++i) ppt2->m[i] = ppr2->m[i];
  int index_type = 0;
  ppt2->has_source_T = ppt2->has_source_E = ppt2->has_source_B = _FALSE_;
  if (c->has_T) {
    ppt2->has_source_T = _TRUE_;
    ppt2->n_sources_T = los_size_l_indexm(ppr2->l_max_los_t, ppt2->m, ppt2->m_size);
    ppt2->index_tp2_T = index_type;
    index_type += ppt2->n_sources_T;
  }
  if (c->has_E) {
    ppt2->has_source_E = _TRUE_;
    ppt2->n_sources_E = los_size_l_indexm(ppr2->l_max_los_p, ppt2->m, ppt2->m_size);
    ppt2->index_tp2_E = index_type;
    index_type += ppt2->n_sources_E;
    if (ppr2->m_max_song != 0) {
      ppt2->has_source_B = _TRUE_;
      ppt2->n_sources_B = los_size_l_indexm(ppr2->l_max_los_p, ppt2->m, ppt2->m_size);
      ppt2->index_tp2_B = index_type;
      index_type += ppt2->n_sources_B;
    }
  }
  ppt2->tp2_size = index_type;
  los_test(ppt2->tp2_size == 0, err, "no source requested");
  ppt2->largest_l = ppr2->l_max_los;
  los_alloc(ppt2->lm_array, (ppt2->largest_l + 1) * sizeof(int *), err);
  for (int l = 0; l <= ppt2->largest_l; ++l) {
    los_calloc(ppt2->lm_array[l], ppr2->index_m_max[l] + 2, sizeof(int), err);
    for (int im = 0; im <= ppr2->index_m_max[l]; ++im)
      ppt2->lm_array[l][im] = los_offset_l_indexm(l, ppt2->m[im], ppt2->m, ppt2->m_size);
  }

  /* k grid */
  ppt2->k_size = c->k_size;
  los_alloc(ppt2->k, c->k_size * sizeof(double), err);
  for (int i = 0; i < c->k_size; ++i) {
    if (c->k_size == 1) { ppt2->k[i] = c->k_min; break; }
    if (c->k_log == 2) {
      /* log-uniform draws in [k_min, k_max] from the scenario seed (sorted below) */
      unsigned long long u = splitmix64(splitmix64(c->seed ^ 0x6B5F6772696421ULL) + (unsigned)i);
      double t = (double)(u >> 11) * 0x1.0p-53;
      ppt2->k[i] = exp(log(c->k_min) + t * (log(c->k_max) - log(c->k_min)));
    } else if (c->k_log)
      ppt2->k[i] = exp(log(c->k_min) + i * (log(c->k_max) - log(c->k_min)) / (c->k_size - 1));
    else
      ppt2->k[i] = c->k_min + i * (c->k_max - c->k_min) / (c->k_size - 1);
  }
  if (c->k_log == 2) {
    qsort(ppt2->k, c->k_size, sizeof(double), cmp_double);
    for (int i = 1; i < c->k_size; ++i)
      los_test(ppt2->k[i] <= ppt2->k[i - 1], err, "k draws %d and %d coincide", i - 1, i);
  }
  /* coarse k3 grids (source/perturbations2.c:1965-1984, linear sampling) */
  los_alloc(ppt2->k3, c->k_size * sizeof(double **), err);
  los_alloc(ppt2->k3_size, c->k_size * sizeof(int *), err);
  for (int i1 = 0; i1 < c->k_size; ++i1) {
    los_alloc(ppt2->k3[i1], (i1 + 1) * sizeof(double *), err);
    los_alloc(ppt2->k3_size[i1], (i1 + 1) * sizeof(int), err);
    for (int i2 = 0; i2 <= i1; ++i2) {
      double k1 = ppt2->k[i1], k2 = ppt2->k[i2];
      double k3_min = fabs(k1 - k2) + fabs(_MIN_K3_DISTANCE_);
      double k3_max = k1 + k2 - fabs(_MIN_K3_DISTANCE_);
      k3_min = LOS_MAX(k3_min, (k1 + k2) / _MIN_K3_RATIO_);
      k3_min = LOS_MAX(k3_min, ppt2->k[0]);
      k3_max = LOS_MIN(k3_max, ppt2->k[ppt2->k_size - 1]);
      los_test(k3_min >= k3_max, err, "k3_min=%g >= k3_max=%g for k1=%g k2=%g", k3_min, k3_max,
               k1, k2);
      int n = c->k3_size;
      ppt2->k3_size[i1][i2] = n;
      los_alloc(ppt2->k3[i1][i2], n * sizeof(double), err);
      for (int q = 0; q < n; ++q) ppt2->k3[i1][i2][q] = k3_min + q * (k3_max - k3_min) / (n - 1);
    }
  }
  /* time sampling */
  ppt2->tau_size = c->tau_size;
  los_alloc(ppt2->tau_sampling, c->tau_size * sizeof(double), err);
  for (int i = 0; i < c->tau_size; ++i)
    ppt2->tau_sampling[i] =
        exp(log(c->tau_min) + i * (log(c->tau_max) - log(c->tau_min)) / (c->tau_size - 1));

  /* sources */
  los_alloc(ppt2->sources, ppt2->tp2_size * sizeof(double ***), err);
  los_calloc(ppt2->has_allocated_sources, c->k_size, sizeof(short), err);
  for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
    los_alloc(ppt2->sources[tp], c->k_size * sizeof(double **), err);
    for (int i1 = 0; i1 < c->k_size; ++i1) ppt2->sources[tp][i1] = NULL;
  }
  double *gtau;
  los_alloc(gtau, c->tau_size * sizeof(double), err);
  for (int i = 0; i < c->tau_size; ++i) {
    double t = ppt2->tau_sampling[i];
    double gauss = exp(-(t - c->tau_peak) * (t - c->tau_peak) / (2.0 * c->tau_sigma * c->tau_sigma));
    double ramp = 0.5 * (1.0 + tanh((t - c->ramp_center) / c->ramp_width));
    gtau[i] = gauss + c->ramp_amp * ramp;
  }
  /* per type: field and (L,m) */
  int *tp_field, *tp_L, *tp_m;
  los_alloc(tp_field, ppt2->tp2_size * sizeof(int), err);
  los_alloc(tp_L, ppt2->tp2_size * sizeof(int), err);
  los_alloc(tp_m, ppt2->tp2_size * sizeof(int), err);
  for (int f = 0; f < 3; ++f) {
    int has = (f == 0) ? ppt2->has_source_T : (f == 1) ? ppt2->has_source_E : ppt2->has_source_B;
    if (!has) continue;
    int base = (f == 0) ? ppt2->index_tp2_T : (f == 1) ? ppt2->index_tp2_E : ppt2->index_tp2_B;
    int Lmax = (f == 0) ? ppr2->l_max_los_t : ppr2->l_max_los_p;
    for (int L = 0; L <= Lmax; ++L)
      for (int im = 0; im <= ppr2->index_m_max[L]; ++im) {
        int tp = base + ppt2->lm_array[L][im];
        tp_field[tp] = f;
        tp_L[tp] = L;
        tp_m[tp] = ppt2->m[im];
      }
  }
  for (int i1 = 0; i1 < c->k_size; ++i1) {
    for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
      los_alloc(ppt2->sources[tp][i1], (i1 + 1) * sizeof(double *), err);
      for (int i2 = 0; i2 <= i1; ++i2) {
        long n = (long)ppt2->k3_size[i1][i2] * c->tau_size;
        los_alloc(ppt2->sources[tp][i1][i2], n * sizeof(double), err);
      }
    }
    ppt2->has_allocated_sources[i1] = _TRUE_;
    ppt2->count_allocated_sources += 0;
  }
  long npairs = (long)c->k_size * (c->k_size + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
  for (long p = 0; p < npairs * ppt2->tp2_size; ++p) {
    int tp = (int)(p % ppt2->tp2_size);
    long pair = p / ppt2->tp2_size;
    int i1 = (int)floor((sqrt(8.0 * pair + 1.0) - 1.0) / 2.0);
    while ((long)i1 * (i1 + 1) / 2 > pair) i1--;
    while ((long)(i1 + 1) * (i1 + 2) / 2 <= pair) i1++;
    int i2 = (int)(pair - (long)i1 * (i1 + 1) / 2);
    int nk3 = ppt2->k3_size[i1][i2];
    double *dst = ppt2->sources[tp][i1][i2];
    int f = tp_field[tp], L = tp_L[tp], m = tp_m[tp];
    int vanish = (f >= 1 && L < 2) || (f == 2 && m == 0);
    double A = vanish ? 0.0 : gauss_amplitude(c->seed, f, L, m, i1, i2);
    for (int q = 0; q < nk3; ++q) {
      double fk = pow(ppt2->k3[i1][i2][q] / 0.01, -1.5);
      for (int t = 0; t < c->tau_size; ++t) dst[(long)t * nk3 + q] = A * fk * gtau[t];
    }
  }
  free(gtau);
  free(tp_field);
  free(tp_L);
  free(tp_m);

  /* ---------------- bessels ---------------- */
  pbs->x_max = c->xx_max;
  pbs2->xx_max = c->xx_max;
  if (c->j_build) {
    if (los_bessel2_build_gpu(ppr, ppr2, ppt2, pbs, pbs2, 0, err) == _FAILURE_) return _FAILURE_;
  } else if (los_bessel2_build(ppr, ppr2, ppt2, pbs, pbs2, err) == _FAILURE_) {
    return _FAILURE_;
  }

  /* ---------------- transfers2 settings ---------------- */
  ptr2->k_sampling = c->k3_sampling;
  ptr2->tau_sampling = c->tau_sampling;
  ptr2->transfer2_verbose = c->verbose;
  ptr2->stop_at_transfers2 = _FALSE_;
  strcpy(ptr2->transfers_dir, ".");

  *out = s;
  return _SUCCESS_;
}

void los_scenario_structs(los_scenario *s, struct precision **ppr, struct precision2 **ppr2,
                          struct background **pba, struct thermo **pth,
                          struct perturbs **ppt, struct perturbs2 **ppt2,
                          struct bessels **pbs, struct bessels2 **pbs2,
                          struct transfers **ptr, struct transfers2 **ptr2)
{
  if (ppr) *ppr = &s->pr;
  if (ppr2) *ppr2 = &s->pr2;
  if (pba) *pba = &s->ba;
  if (pth) *pth = &s->th;
  if (ppt) *ppt = &s->pt;
  if (ppt2) *ppt2 = &s->pt2;
  if (pbs) *pbs = &s->bs;
  if (pbs2) *pbs2 = &s->bs2;
  if (ptr) *ptr = &s->tr;
  if (ptr2) *ptr2 = &s->tr2;
}

long los_scenario_n_transfers(los_scenario *s)
{
  struct transfers2 *ptr2 = &s->tr2;
  struct perturbs2 *ppt2 = &s->pt2;
  if (!ptr2->has_cls || ptr2->k_size_k1k2 == NULL) return 0;
  long n = 0;
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) n += ptr2->k_size_k1k2[i1][i2];
  return n * ptr2->tt2_size;
}

int los_scenario_gather(los_scenario *s, double *out, long n, char *err)
{
  struct transfers2 *ptr2 = &s->tr2;
  struct perturbs2 *ppt2 = &s->pt2;
  long need = los_scenario_n_transfers(s);
  los_test(n < need, err, "buffer of %ld values, need %ld", n, need);
  long o = 0;
  for (int tt = 0; tt < ptr2->tt2_size; ++tt)
    for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
      /* allocated per k1 (transfer2_init) or per type (load from disk) */
      los_test(ptr2->transfer[tt] == NULL || ptr2->transfer[tt][i1] == NULL, err,
               "transfer level tt=%d k1=%d not allocated", tt, i1);
      for (int i2 = 0; i2 <= i1; ++i2) {
        int nk = ptr2->k_size_k1k2[i1][i2];
        memcpy(out + o, ptr2->transfer[tt][i1][i2], nk * sizeof(double));
        o += nk;
      }
    }
  return _SUCCESS_;
}

int los_scenario_set(los_scenario *s, const char *key, double v, char *err)
{
  if (!strcmp(key, "store_transfers_to_disk")) s->pr2.store_transfers_to_disk = (short)v;
  else if (!strcmp(key, "load_transfers_from_disk")) s->pr2.load_transfers_from_disk = (short)v;
  else if (!strcmp(key, "store_sources_to_disk")) s->pr2.store_sources_to_disk = (short)v;
  else if (!strcmp(key, "load_sources_from_disk")) s->pr2.load_sources_from_disk = (short)v;
  else if (!strcmp(key, "stop_at_transfers2")) s->tr2.stop_at_transfers2 = (short)v;
  else if (!strcmp(key, "has_perturbations2")) s->pt2.has_perturbations2 = (short)v;
  else if (!strcmp(key, "has_cmb_bispectra")) s->pt2.has_cmb_bispectra = (short)v;
  else if (!strcmp(key, "has_cmb_spectra")) s->pt2.has_cmb_spectra = (short)v;
  else if (!strcmp(key, "transfer2_verbose")) s->tr2.transfer2_verbose = (short)v;
  else if (!strcmp(key, "q_linstep_song")) s->pr2.q_linstep_song = v;
  else if (!strcmp(key, "tau_linstep_song")) s->pr2.tau_linstep_song = v;
  else if (!strcmp(key, "k_sampling")) s->tr2.k_sampling = (enum transfer2_k3_sampling)(int)v;
  else if (!strcmp(key, "tau_sampling")) s->tr2.tau_sampling = (enum transfer2_tau_sampling)(int)v;
  else if (!strcmp(key, "sources_k3_interpolation"))
    s->pr2.sources_k3_interpolation = (enum interpolation_methods)(int)v;
  else if (!strcmp(key, "sources_time_interpolation"))
    s->pr2.sources_time_interpolation = (enum interpolation_methods)(int)v;
  else if (!strcmp(key, "bispectra_k3_extrapolation"))
    s->pr.bispectra_k3_extrapolation = (enum bispectra_k3_extrapolation)(int)v;
  else if (!strcmp(key, "extra_k3_oscillations_left")) s->pr.extra_k3_oscillations_left = v;
  else if (!strcmp(key, "extra_k3_oscillations_right")) s->pr.extra_k3_oscillations_right = v;
  else if (!strcmp(key, "xx_max")) s->bs2.xx_max = v;
  else los_fail(err, "los_scenario_set: unknown key '%s'", key);
  return _SUCCESS_;
}

int los_scenario_set_string(los_scenario *s, const char *key, const char *value, char *err)
{
  if (!strcmp(key, "transfers_dir")) {
    snprintf(s->tr2.transfers_dir, _FILENAMESIZE_, "%s", value);
    return _SUCCESS_;
  }
  if (!strcmp(key, "sources_dir")) {
    snprintf(s->pt2.sources_dir, _FILENAMESIZE_, "%s", value);
    los_call(los_sources_set_paths(&s->pt2), s->pt2.error_message, err);
    return _SUCCESS_;
  }
  los_fail(err, "los_scenario_set_string: unknown key '%s'", key);
}

/* Pointers into the tables, for tests that overwrite inputs with analytic ones. */
double *los_scenario_source_ptr(los_scenario *s, int tp, int k1, int k2)
{
  if (!s->pt2.has_allocated_sources[k1]) return NULL;
  return s->pt2.sources[tp][k1][k2];
}

double *los_scenario_J_row(los_scenario *s, int J, int L, int il, int im, int *xmin, int *xsize)
{
  struct bessels2 *b = &s->bs2;
  *xmin = b->index_xmin_J[J][L][il][im];
  *xsize = b->x_size_J[J][L][il][im];
  return b->J_Llm_x[J][L][il][im];
}

int los_scenario_rebuild_J(los_scenario *s, int gpu, int device, double *seconds, char *err)
{
  struct timespec t0, t1;
  los_bessel2_free(&s->bs, &s->bs2);
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int rc = gpu ? los_bessel2_build_gpu(&s->pr, &s->pr2, &s->pt2, &s->bs, &s->bs2, device, err)
               : los_bessel2_build(&s->pr, &s->pr2, &s->pt2, &s->bs, &s->bs2, err);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (seconds) *seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
  return rc;
}

/* Copies of the generated grids (l list, k grid, tau sampling). */
int los_scenario_grids(los_scenario *s, int *l_out, double *k_out, double *tau_out)
{
  if (l_out) memcpy(l_out, s->bs.l, s->bs.l_size * sizeof(int));
  if (k_out) memcpy(k_out, s->pt2.k, s->pt2.k_size * sizeof(double));
  if (tau_out) memcpy(tau_out, s->pt2.tau_sampling, s->pt2.tau_size * sizeof(double));
  return _SUCCESS_;
}

/* Replace the (k1,k2) grid values (same size). */
int los_scenario_override_k(los_scenario *s, const double *k, int n, char *err)
{
  los_test(n != s->pt2.k_size, err, "k grid has %d values, got %d", s->pt2.k_size, n);
  memcpy(s->pt2.k, k, n * sizeof(double));
  return _SUCCESS_;
}

/* Replace the coarse k3 grid of one pair (any size >= 2); the pair's sources
   are reallocated (zero) with the new size. */
int los_scenario_override_k3(los_scenario *s, int i1, int i2, const double *k3, int n, char *err)
{
  struct perturbs2 *ppt2 = &s->pt2;
  los_test(i2 > i1 || i1 >= ppt2->k_size || n < 2, err, "bad pair or size");
  free(ppt2->k3[i1][i2]);
  los_alloc(ppt2->k3[i1][i2], n * sizeof(double), err);
  memcpy(ppt2->k3[i1][i2], k3, n * sizeof(double));
  ppt2->k3_size[i1][i2] = n;
  for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
    free(ppt2->sources[tp][i1][i2]);
    los_calloc(ppt2->sources[tp][i1][i2], (size_t)n * ppt2->tau_size, sizeof(double), err);
  }
  return _SUCCESS_;
}

/* Replace one J row: it is stored from index xmin with n values. */
int los_scenario_override_J(los_scenario *s, int J, int L, int il, int im, int xmin,
                            const double *v, int n, char *err)
{
  struct bessels2 *b = &s->bs2;
  los_test(xmin < 0 || xmin + n != b->xx_size, err, "row must cover [xmin, xx_size)");
  free(b->J_Llm_x[J][L][il][im]);
  los_alloc(b->J_Llm_x[J][L][il][im], n * sizeof(double), err);
  memcpy(b->J_Llm_x[J][L][il][im], v, n * sizeof(double));
  b->index_xmin_J[J][L][il][im] = xmin;
  b->x_size_J[J][L][il][im] = n;
  b->x_min_J[J][L][il][im] = b->xx[xmin];
  return _SUCCESS_;
}

/* Replace the second derivatives of one J row (cubic J interpolation); the row
   keeps its first index and size. */
int los_scenario_override_ddJ(los_scenario *s, int J, int L, int il, int im, const double *v,
                              int n, char *err)
{
  struct bessels2 *b = &s->bs2;
  los_test(b->ddJ_Llm_x == NULL, err, "no second derivatives: J interpolation is linear");
  los_test(n != b->x_size_J[J][L][il][im], err, "the row holds %d values, not %d",
           b->x_size_J[J][L][il][im], n);
  free(b->ddJ_Llm_x[J][L][il][im]);
  los_alloc(b->ddJ_Llm_x[J][L][il][im], n * sizeof(double), err);
  memcpy(b->ddJ_Llm_x[J][L][il][im], v, n * sizeof(double));
  return _SUCCESS_;
}

/* Inverse of los_scenario_gather: write values into ptr2->transfer. */
int los_scenario_scatter(los_scenario *s, const double *in, long n, char *err)
{
  struct transfers2 *ptr2 = &s->tr2;
  struct perturbs2 *ppt2 = &s->pt2;
  long need = los_scenario_n_transfers(s);
  los_test(n < need, err, "buffer of %ld values, need %ld", n, need);
  long o = 0;
  for (int tt = 0; tt < ptr2->tt2_size; ++tt)
    for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
      los_test(!ptr2->has_allocated_transfers[i1], err, "k1 level %d not allocated", i1);
      for (int i2 = 0; i2 <= i1; ++i2) {
        int nk = ptr2->k_size_k1k2[i1][i2];
        memcpy(ptr2->transfer[tt][i1][i2], in + o, nk * sizeof(double));
        o += nk;
      }
    }
  return _SUCCESS_;
}

/* Bookkeeping readout: sizes per pair (k_size, physical start, physical size). */
int los_scenario_k3_sizes(los_scenario *s, int *out, long n, char *err)
{
  struct transfers2 *ptr2 = &s->tr2;
  struct perturbs2 *ppt2 = &s->pt2;
  long need = 3L * ppt2->k_size * (ppt2->k_size + 1) / 2;
  los_test(n < need || ptr2->k_size_k1k2 == NULL, err, "need %ld ints after the indices", need);
  long o = 0;
  for (int i1 = 0; i1 < ppt2->k_size; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) {
      out[o++] = ptr2->k_size_k1k2[i1][i2];
      out[o++] = ptr2->k_physical_start_k1k2[i1][i2];
      out[o++] = ptr2->k_physical_size_k1k2[i1][i2];
    }
  return _SUCCESS_;
}

int los_scenario_store_sources(los_scenario *s, char *err)
{
  struct perturbs2 *ppt2 = &s->pt2;
  los_test(ppt2->sources_paths == NULL, err, "set sources_dir first");
  for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
    if (!ppt2->has_allocated_sources[i1]) continue;
    los_call(perturb2_store_sources_to_disk(ppt2, i1), ppt2->error_message, err);
    los_call(perturb2_free_k1_level(ppt2, i1), ppt2->error_message, err);
  }
  return _SUCCESS_;
}

void los_scenario_destroy(los_scenario *s)
{
  if (!s) return;
  struct perturbs2 *ppt2 = &s->pt2;
  los_sources_free_paths(ppt2);
  for (int tp = 0; tp < ppt2->tp2_size; ++tp) {
    for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
      if (ppt2->has_allocated_sources[i1] && ppt2->sources[tp][i1]) {
        for (int i2 = 0; i2 <= i1; ++i2) free(ppt2->sources[tp][i1][i2]);
        free(ppt2->sources[tp][i1]);
      }
    }
    free(ppt2->sources[tp]);
  }
  free(ppt2->sources);
  free(ppt2->has_allocated_sources);
  for (int i1 = 0; i1 < ppt2->k_size; ++i1) {
    for (int i2 = 0; i2 <= i1; ++i2) free(ppt2->k3[i1][i2]);
    free(ppt2->k3[i1]);
    free(ppt2->k3_size[i1]);
  }
  free(ppt2->k3);
  free(ppt2->k3_size);
  free(ppt2->k);
  free(ppt2->tau_sampling);
  for (int l = 0; l <= ppt2->largest_l; ++l) free(ppt2->lm_array[l]);
  free(ppt2->lm_array);
  free(ppt2->m);
  los_bessel2_free(&s->bs, &s->bs2);
  free(s->bs.l);
  free(s->pr2.index_m_max);
  free(s);
}
