/*
 * los_device.h — parameter blocks shared by the host engine and the HIP
 * kernels of the LOS stage. Plain C layout (passed by value to kernels).
 */
#ifndef SONG_LOS_DEVICE_H
#define SONG_LOS_DEVICE_H

#define LOS_MAX_L 8       /* sources L = 0..LOS_MAX_L-1 */
#define LOS_MAX_MI 6      /* entries of the m-list */
#define LOS_NC_MAX 256    /* coarse k3 values per pair handled by the spline kernel */
#define LOS_EXACT_BLOCK_MAX 512 /* (l,m) lanes per workgroup of the exact LOS kernel */
#define LOS_BUF_MAX 0xFFFFF000u /* largest table read through a raw buffer resource */

/* One (k1,k2) pair as laid out on the device. */
typedef struct {
  long src_off;  /* offset of [tp][tau][kc] sources of this pair in src/dd */
  int kc_off;    /* offset of the coarse k3 grid in kc */
  int nc;        /* coarse k3 count */
  int f_off;     /* offset of the fine k3 grid in kf */
  int nf;        /* fine k3 count */
  int k1, k2;    /* indices in ppt2->k */
} los_pair;

/* A job: one block of the output array, [tt][ostride], with n_lm (l,m) lanes
   per field (the first n_lm (l,m) offsets of the projection table). */
typedef struct {
  long obase;          /* offset of the block in out */
  int ostride;         /* columns of the block (fine k3 values of the job) */
  int n_lm;            /* (l,m) lanes */
  int ttT, ttE, ttB;   /* first row of each field inside the block */
} los_job;

/* A tile: npos consecutive entries of the work order, all of one job. */
typedef struct {
  int pos0, npos;
} los_tile;

/* Kernel organisation and tile shape of the fast forms:
   kind 0: "stream" kernel, nk fine k3 per workgroup, ch time steps per LDS
           chunk, persistent = 1 for the chip-sized grid fed by a work queue;
   kind 1: "sweep" kernel, one wave per (tile of nk fine k3, m, 64 l) item,
           persistent = 1 for the chip-sized grid of waves striding over the
           items of their XCD;
   kind 2: "k-lane" kernel, one wave per ((l,m), 64 neighbouring fine k3);
   kind 0 of the plain FP64 form is its one-workgroup-per-k3 kernel. */
typedef struct {
  int kind, nk, ch, persistent;
} los_shape;

/* Everything the LOS kernels read; pointers are device pointers. */
typedef struct {
  /* pairs and fine k3 */
  const los_pair *pairs;
  int n_pairs;
  const double *kc;     /* coarse k3, concatenated per pair */
  const double *kf;     /* fine k3, concatenated per pair (F values) */
  const int *fpair;     /* pair of each fine k3 */
  const int *fsrc;      /* fine index (global) whose k3 feeds the source interpolation */
  const int *fbr;       /* bracket: coarse index i with kc[i] <= k3 <= kc[i+1] */
  const double *src;    /* sources, per pair [tp][tau][kc] */
  const double *dd;     /* spline second derivatives, same layout */
  int F;                /* fine k3 total */
  /* Work order: n_pos positions, forder[s] the fine index at position s. The
     positions of one job are contiguous and sorted by k3, so that workgroups
     that run together on one XCD read neighbouring J rows and share its L2.
     Position s belongs to job pos_job[s] and writes column ocol[s] of it. */
  const int *forder;
  int n_pos;
  const int *pos_job;
  const int *ocol;
  const los_job *jobs;
  int n_lm_max;         /* largest n_lm of the jobs */
  const los_tile *tiles;  /* tiles of the tiled kernels (NK positions or fewer) */
  /* XCD x owns positions [xcd_pos[x], xcd_pos[x+1]) (one fine k3 per
     workgroup) or tiles [xcd_tile[x], xcd_tile[x+1]); the ranges are
     contiguous and balanced by estimated cost, and each XCD takes its range
     from the end down (the costliest, largest k3, first within a job). */
  int xcd_pos[9];
  int xcd_tile[9];
  int *queue;           /* 8 tile counters of the persistent kernels */
  int tp2_size;
  int cubic_k3;         /* 1: cubic spline in k3, 0: linear */

  /* time grid */
  const double *t0mt;   /* tau0 - tau */
  const double *dtau;   /* trapezoid weights (tau[i+1]-tau[i-1], one-sided at ends) */
  int ntau;

  /* projection tables: J[((jt*L_size + L)*X + x)*n_lm + lm] */
  const double *J;
  const float *J32;     /* FP32 copy of J (FP32-compensated form only) */
  const int *xmin;      /* [(jt*L_size + L)*n_lm + lm], huge when the row is absent */
  const int *min_xmin;  /* [lm]: smallest xmin over the rows an element uses */
  const double *xx;     /* x grid, X entries */
  double xx_step;
  int X;
  int L_size;
  int n_lm;             /* (l,m) offsets of the table (its row stride) */
  const int *l_of;      /* l value per lm */
  const int *m_of;      /* m value per lm */
  const int *mi_of;     /* index of m in the m-list per lm */

  /* fields */
  int has_T, has_E, has_B;
  int LmaxT, LmaxP;
  int jTT, jEE, jEB;    /* J type indices (-1 if absent) */
  /* source type of (field, L, m-index), -1 if absent */
  int srcT[LOS_MAX_L][LOS_MAX_MI];
  int srcE[LOS_MAX_L][LOS_MAX_MI];
  int srcB[LOS_MAX_L][LOS_MAX_MI];

  /* Sweep kernels (FMA-FP64 and FP32-compensated forms). One wave handles one
     m and 64 consecutive l: lane slot lp = mi * n_lp + il (n_lp = l count
     rounded up to 64). */
  const void *Jp;        /* J re-laid out [jt][L][x][lp] (FP64 pairs or FP32) */
  const void *Jd;        /* second derivatives of J as FP64 pairs, same layout (cubic J) */
  int cubic_J;           /* 1: cubic interpolation of J in x (plain FP64 sweep kernels) */
  unsigned Jp_bytes;     /* size of the table at Jp (read with 32-bit byte offsets) */
  const int *xminp;      /* [jt][L][lp] first live x index, INT_MAX for empty slots */
  const int *lm_of_lp;   /* (l,m) offset of each slot, -1 for empty slots */
  int n_lp;              /* l slots per m */
  int m_size;            /* entries of the m-list */
  int m_of_mi[LOS_MAX_MI]; /* m value of each m-list entry */
  int n_wave;            /* waves per tile: m_size * n_lp / 64 */
  void *ixa;             /* per (f, t): index_x and a_J ({int, float} or {int, pad, double}) */
  void *sf;              /* per (f, t, mi): [T, E, B][L] sources (x trapezoid weight: fast forms) */
  const int *tau_end_fw; /* per (f, wave): time steps before every row of the wave is dead */

  /* k-lane kernels (one (l,m) per wave, one fine k3 per lane): tables by rank
     r = position in the engine's sorted order, k_npad ranks per row */
  const int *korder;     /* fine index of each rank */
  const int *rank_of_f;  /* rank of each fine index */
  long k_npad;
  void *kixa;            /* [t][r] index_x and a_J */
  void *ksrc;            /* [tp][t][r] interpolated sources (x trapezoid weight: fast forms) */
  const void *Jk;        /* J re-laid out [jt][L][lm][x] (FP64 or FP32) */
  const int *kgrp;       /* lane groups: NLB (l,m) offsets of one m per group, -1 padding */
  int k_ngrp;            /* number of lane groups */

  /* output blocks of the jobs */
  double *out;
} los_params;

#endif
