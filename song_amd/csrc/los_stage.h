/*
 * los_stage.h — what the GPU side of the LOS stage reads from SONG's
 * structures, as plain C values and pointers.
 *
 * The host translation unit that implements SONG's entry points
 * (los_transfer2.c) is the only one that dereferences struct precision,
 * precision2, perturbs2, bessels, bessels2 and transfers2: it fills this
 * descriptor, and the engine, the server and the kernels read the descriptor
 * only. Recompiling that one unit against SONG's own headers therefore adapts
 * the whole library to SONG's structure layouts. The arrays the pointers
 * reach are SONG's own (same indexing as the fields they come from).
 */
#ifndef SONG_LOS_STAGE_H
#define SONG_LOS_STAGE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct los_stage los_stage;

struct los_stage {
  /* options (ptr2->tau_sampling, ppr->bessels_interpolation,
     ppr2->sources_k3_interpolation) */
  int tau_sources;  /* 1: transfer2_tau_sampling = sources */
  int linear_J;     /* 1: linear interpolation of J in x */
  int cubic_k3;     /* 1: cubic spline of the sources in k3, 0: linear */

  /* time sampling of the sources (ppt2->tau_sampling) and tau0 (ptr2->tau0) */
  int tau_size;
  const double *tau;
  double tau0;

  /* fields (ppt2->has_source_*), L ranges (ppr2->l_max_los_t/p), m list */
  int has_T, has_E, has_B;
  int LmaxT, LmaxP;
  int m_max;       /* ppr2->m_max_song */
  int largest_l;   /* ppt2->largest_l */
  int m_size;
  const int *m;             /* ppr2->m */
  const int *index_m_max;   /* ppr2->index_m_max[l] */
  int l_size;
  const int *l;             /* ptr2->l */
  int **lm_array;           /* ptr2->lm_array[index_l][index_m] */
  int **src_lm;             /* ppt2->lm_array[L][index_m] */
  int tt2_size, n_transfers;
  int index_tt2_T, index_tt2_E, index_tt2_B;

  /* sources: sources[tp][k1][k2][tau * k3_size + k3] */
  int tp2_size;
  int index_tp2_T, index_tp2_E, index_tp2_B;
  int k_size;
  double ****sources;
  const short *has_sources;  /* ppt2->has_allocated_sources[k1] */
  double ***k3;              /* coarse k3 grid per pair */
  int **k3_size;

  /* fine k3 bookkeeping (ptr2) */
  int k3_size_max;
  int **k_size_k1k2;
  int **k_phys_start;
  int **k_phys_size;
  double **k_max_k1k2;
  /* the fine k3 grid of a pair (transfer2_get_k3_list); 0 on success,
     otherwise a message in err */
  int (*k3_list)(const los_stage *st, int i1, int i2, double *out, char *err);
  void *song;  /* what k3_list reads (the caller's structures) */

  /* projection functions (bessels2) */
  const void *J_id;  /* identity of the tables (the J arrays are reused while it stays) */
  int J_size, L_size, xx_size;
  double xx_step, xx_max;
  const double *xx;
  int jTT, jEE, jEB;  /* index_J_* of the types present, -1 otherwise */
  int ****xmin_J;     /* index_xmin_J[J][L][l][m] */
  int ****x_size_J;   /* x_size_J[J][L][l][m] */
  double *****J;      /* J_Llm_x[J][L][l][m][x - xmin] */
  double *****ddJ;    /* ddJ_Llm_x, same indexing (cubic interpolation in x); NULL if linear */

  /* outputs (ptr2) */
  double ****transfer;       /* transfer[tt][k1][k2][k3] */
  short *has_transfers;      /* has_allocated_transfers[k1] */
  long *count_memorised;     /* &count_memorised_transfers */
};

#ifdef __cplusplus
}
#endif

#endif
