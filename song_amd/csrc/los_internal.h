/*
 * los_internal.h — declarations shared by the host-side C sources that read
 * SONG's structures (los_transfer2.c, los_bessel2.c, los_scenario.c); not part
 * of the public C-ABI. The HIP sources include los_common.h only.
 */
#ifndef SONG_LOS_INTERNAL_H
#define SONG_LOS_INTERNAL_H

/* Built into SONG (-DLOS_SONG_HEADERS with SONG's include/ and CLASS's on the
   include path), the host units read SONG's own structure definitions;
   otherwise this repository's (same fields, the subset the stage uses). */
#ifdef LOS_SONG_HEADERS
#include "common2.h"
#include "transfer2.h"
#else
#include "../../include/song_los_types.h"
#include "../../include/transfer2.h"
#endif
#include "los_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- J-table builder setup shared by the host and GPU builders (los_bessel2.c) ---- */
int los_bessel2_setup(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, los_jrow **rows,
                      int *nrows, int *l1max, char *errmsg);
int los_bessel2_alloc_rows(struct bessels2 *pbs2, const los_jrow *rows, int nrows,
                           char *errmsg);
void los_bessel2_spline_rows(struct bessels *pbs, struct bessels2 *pbs2, const los_jrow *rows,
                             int nrows);

#ifdef __cplusplus
}
#endif

#endif
