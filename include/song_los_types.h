/*
 * song_los_types.h — the data structures that SONG's second-order line-of-sight
 * (LOS) transfer stage reads and writes, restricted to the fields that stage uses.
 *
 * Field names, types and meanings follow the reference so that code written
 * against SONG's headers compiles against this one:
 *   struct precision2  : include/common2.h:24-214
 *   struct perturbs2   : include/perturbations2.h:397-417, 498-514, 611-644, 847-863, 906-907
 *   struct bessels2    : include/bessel2.h:32-112
 *   struct transfers2  : include/transfer2.h:76-238
 *   struct transfer2_workspace : include/transfer2.h:252-295
 * The first-order CLASS structures (precision, background, thermo, perturbs,
 * bessels, transfers) come from the CLASS submodule, which the reference does
 * not vendor; only the fields read by source/transfer2.c are declared here.
 */
#ifndef SONG_LOS_TYPES_H
#define SONG_LOS_TYPES_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CLASS conventions (common.h of the CLASS fork): functions return _SUCCESS_
   or _FAILURE_ and leave a message in an ErrorMsg buffer. */
#ifndef _SUCCESS_
#define _SUCCESS_ 0
#define _FAILURE_ 1
#endif
#define _TRUE_ 1
#define _FALSE_ 0
#define _ERRORMSGSIZE_ 2048
#define _FILENAMESIZE_ 256
#define _MAX_LENGTH_LABEL_ 64
#define _MAX_NUM_AZIMUTHAL_ 14 /* include/common2.h:18 */
#define _MIN_K3_DISTANCE_ 1e-10 /* include/perturbations2.h:178 */
#define _MIN_K3_RATIO_ 100      /* include/perturbations2.h:187 */
#define _MAX_NUM_TRANSFERS_ 4096 /* include/transfer2.h:59 */

typedef char ErrorMsg[_ERRORMSGSIZE_];

/* Interpolation methods used for the sources (k3, time) and for J(x). */
enum interpolation_methods { linear_interpolation, cubic_interpolation };

/* k3 extrapolation beyond the triangular limits (source/transfer2.c:1365-1400). */
enum bispectra_k3_extrapolation { no_k3_extrapolation, flat_k3_extrapolation };

/* include/transfer2.h:18-21 */
enum transfer2_k3_sampling { bessel_k3_sampling, class_transfer2_k3_sampling };

/* include/transfer2.h:27-32 */
enum transfer2_tau_sampling { sources_tau_sampling, custom_tau_sampling, bessel_tau_sampling };

/* include/bessel2.h:12-16 */
enum projection_function_types { J_TT, J_EE, J_EB };

/* ---------------------------------------------------------------------------
 * First-order CLASS structures: only the fields transfer2.c reads.
 * ------------------------------------------------------------------------- */
struct precision {
  enum interpolation_methods bessels_interpolation;
  enum bispectra_k3_extrapolation bispectra_k3_extrapolation;
  double extra_k3_oscillations_left;
  double extra_k3_oscillations_right;
  double smallest_allowed_variation;
  short load_run;
  double k_max_tau0_over_l_max;
  ErrorMsg error_message;
};

struct background {
  double conformal_age; /* tau0 */
};

struct thermo {
  double tau_rec;
  double rs_rec;
};

struct perturbs {
  short has_cls;
  short has_cmb_bispectra;
  int l_scalar_max;
};

struct bessels {
  int l_size;
  int *l;
  int l_max;
  double x_max;
};

struct transfers {
  short has_cls; /* first-order transfer module; not read by the LOS stage */
};

/* ---------------------------------------------------------------------------
 * struct precision2 (include/common2.h:24-214), fields used by the stage.
 * ------------------------------------------------------------------------- */
struct precision2 {
  int l_max_los_t;
  int l_max_los_p;
  int l_max_los;

  int m[_MAX_NUM_AZIMUTHAL_];
  int m_size;
  int m_max_song;
  int compute_m[_MAX_NUM_AZIMUTHAL_];
  int index_m[_MAX_NUM_AZIMUTHAL_];
  int *index_m_max; /* index_m_max[l]: index in m of the largest m <= l, or -1 */
  int index_m_max_size; /* number of entries in index_m_max (l = 0..size-1) */

  double tau_linstep_song;
  double k_max_tau0_over_l_max;
  int k3_size;
  double q_linstep_song;

  enum interpolation_methods sources_time_interpolation;
  enum interpolation_methods sources_k3_interpolation;

  double bessel_j_cut_song;
  double bessel_J_cut_song;
  double bessel_x_step_song;

  ErrorMsg error_message;
  short store_transfers_to_disk;
  short load_transfers_from_disk;
  short store_sources_to_disk;
  short load_sources_from_disk;
  short old_run;
};

/* ---------------------------------------------------------------------------
 * struct perturbs2, the producer of the LOS sources.
 * ------------------------------------------------------------------------- */
struct perturbs2 {
  short has_perturbations2;
  short has_cmb_spectra;
  short has_cmb_bispectra;
  short has_cmb_temperature;
  short has_cmb_polarization_e;
  short has_cmb_polarization_b;

  short has_source_T;
  short has_source_E;
  short has_source_B;
  int index_tp2_T;
  int index_tp2_E;
  int index_tp2_B;
  int n_sources_T;
  int n_sources_E;
  int n_sources_B;
  int tp2_size;

  /* sources[index_tp2][index_k1][index_k2][index_tau*k3_size + index_k3] */
  double ****sources;
  short *has_allocated_sources;
  long int count_allocated_sources;
  long int count_memorised_sources;

  int largest_l;
  int **lm_array; /* lm_array[l][index_m]: offset of (l,m) in a source hierarchy */

  int m_size;
  int *m;

  double *k;
  int k_size;
  double ***k3;   /* k3[index_k1][index_k2][index_k3] */
  int **k3_size;

  double *tau_sampling;
  int tau_size;

  char sources_dir[_FILENAMESIZE_];
  char **sources_paths;
  FILE **sources_files;

  short perturbations2_verbose;
  ErrorMsg error_message;
};

/* ---------------------------------------------------------------------------
 * struct bessels2 (include/bessel2.h:32-112): the projection functions J_Llm(x).
 * ------------------------------------------------------------------------- */
struct bessels2 {
  short has_J_TT;
  short has_J_EE;
  short has_J_EB;
  long int count_allocated_Js;
  double j_l1_cut;
  double J_Llm_cut;
  int L_max;
  int L_size;
  int *L;
  int m_size;
  int *m;
  double *xx;
  double xx_step;
  int xx_size;
  double xx_max;
  int index_J_TT;
  int index_J_EE;
  int index_J_EB;
  int J_size;
  int ****index_xmin_J; /* [J][L][l][m] first index of xx where J is non-negligible */
  int ****x_size_J;     /* [J][L][l][m] number of stored x values */
  double ****x_min_J;
  int x_size_max_J;
  double *****J_Llm_x;   /* [J][L][l][m][index_x - index_xmin_J] */
  double *****ddJ_Llm_x; /* same indexing, second derivatives for cubic J */
  short bessels2_verbose;
  ErrorMsg error_message;
};

/* ---------------------------------------------------------------------------
 * struct transfers2 (include/transfer2.h:76-238): the output of the stage,
 * every field in SONG's order and type, so that its layout is SONG's (given
 * the CLASS constants above).
 * ------------------------------------------------------------------------- */
struct transfers2 {
  /* transfer[index_tt][index_k1][index_k2][index_k] */
  double ****transfer;
  int index_tt2_T;
  int index_tt2_E;
  int index_tt2_B;
  int tt2_size;
  char (*tt2_labels)[_MAX_LENGTH_LABEL_];
  short has_cls;

  int l_size;
  int *l;
  int *m;
  int m_size;
  int n_transfers;
  int n_nonzero_transfers_E;
  int n_nonzero_transfers_B;

  int *index_tt2_monopole;
  int *index_pt2_monopole;
  int *corresponding_index_l;
  int *corresponding_index_m;
  int **lm_array;

  enum transfer2_k3_sampling k_sampling;
  enum transfer2_tau_sampling tau_sampling;

  int **k_size_k1k2;
  int **k_physical_start_k1k2;
  int **k_physical_size_k1k2;
  double **k_min_k1k2;
  double **k_max_k1k2;
  int k3_size_max;

  char transfers_dir[_FILENAMESIZE_];
  char **transfers_paths;
  FILE **transfers_files;
  FILE *transfers_status_file;                  /* not used by SONG either */
  char transfers_status_path[_FILENAMESIZE_];   /* not used by SONG either */

  double tau0;
  double tau_rec;
  double rs_rec;

  short transfer2_verbose;
  ErrorMsg error_message;
  long int count_allocated_transfers;
  long int count_memorised_transfers;
  short *has_allocated_transfers;
  short stop_at_transfers2;
};

/* include/transfer2.h:252-295 */
struct transfer2_workspace {
  double transfer;
  int index_tt;
  int index_k;
  int index_k1;
  int index_k2;
  double k;
  double k1;
  double k2;
  double cosk1k2;
  int tau_grid_size;
  double *tau_grid;
  double *tau0_minus_tau;
  double *delta_tau;
  double *k_grid;
  double **sources_time_spline;
  double **interpolated_sources_in_time;
  int *index_tau_left;
  int L_max;
  int thread;
};

#ifdef __cplusplus
}
#endif

#endif /* SONG_LOS_TYPES_H */
