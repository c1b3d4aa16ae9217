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

/* One (k1,k2) pair as laid out on the device. */
typedef struct {
  long src_off;  /* offset of [tp][tau][kc] sources of this pair in src/dd */
  int kc_off;    /* offset of the coarse k3 grid in kc */
  int nc;        /* coarse k3 count */
  int f_off;     /* offset of the fine k3 grid in kf */
  int nf;        /* fine k3 count */
  int k1, k2;    /* indices in ppt2->k */
} los_pair;

/* Everything the LOS kernels read; pointers are device pointers. */
typedef struct {
  /* pairs and fine k3 */
  const los_pair *pairs;
  const double *kc;     /* coarse k3, concatenated per pair */
  const double *kf;     /* fine k3, concatenated per pair (F values) */
  const int *fpair;     /* pair of each fine k3 */
  const int *fsrc;      /* fine index (global) whose k3 feeds the source interpolation */
  const int *fbr;       /* bracket: coarse index i with kc[i] <= k3 <= kc[i+1] */
  const double *src;    /* sources, per pair [tp][tau][kc] */
  const double *dd;     /* spline second derivatives, same layout */
  int F;                /* fine k3 total */
  /* Work order: the fine k3 sorted by value (all pairs together), dealt out
     so that each XCD of the chip receives a contiguous run of that order:
     blocks that run together on one XCD then read neighbouring J rows and
     share its L2. forder[s] is the fine index at sorted position s. */
  const int *forder;
  int f_per_xcd;        /* sorted positions per XCD (ceil(F/8)) */
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
  int n_lm;
  const int *l_of;      /* l value per lm */
  const int *m_of;      /* m value per lm */
  const int *mi_of;     /* index of m in the m-list per lm */

  /* fields */
  int has_T, has_E, has_B;
  int LmaxT, LmaxP;
  int jTT, jEE, jEB;    /* J type indices (-1 if absent) */
  int ttT, ttE, ttB;    /* first transfer type of each field */
  /* source type of (field, L, m-index), -1 if absent */
  int srcT[LOS_MAX_L][LOS_MAX_MI];
  int srcE[LOS_MAX_L][LOS_MAX_MI];
  int srcB[LOS_MAX_L][LOS_MAX_MI];

  /* output [tt][F] */
  double *out;
} los_params;

#endif
