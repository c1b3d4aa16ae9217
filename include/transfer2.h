/*
 * transfer2.h — C-ABI of the MI355X second-order line-of-sight transfer stage.
 *
 * Every function below keeps the name, argument list, return convention
 * (_SUCCESS_/_FAILURE_ with the message in ptr2->error_message) and side
 * effects of the SONG function it replaces; the reference declaration is
 * cited next to each one (all in include/transfer2.h of the reference,
 * definitions in source/transfer2.c). The library behind these entry points
 * computes the LOS integral with HIP kernels on one gfx950 GPU; the host part
 * keeps SONG's bookkeeping (k3 grids, (l,m) lists, output layout, disk
 * storage) bit-for-bit.
 */
#ifndef SONG_LOS_TRANSFER2_H
#define SONG_LOS_TRANSFER2_H

#include "song_los_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/transfer2.h:327-338; source/transfer2.c:79-621 */
int transfer2_init(struct precision *ppr, struct precision2 *ppr2,
                   struct background *pba, struct thermo *pth,
                   struct perturbs *ppt, struct perturbs2 *ppt2,
                   struct bessels *pbs, struct bessels2 *pbs2,
                   struct transfers *ptr, struct transfers2 *ptr2);

/* include/transfer2.h:340-344; source/transfer2.c:769-838 */
int transfer2_free(struct precision2 *ppr2, struct perturbs2 *ppt2,
                   struct transfers2 *ptr2);

/* include/transfer2.h:346-354; source/transfer2.c:866-1067 */
int transfer2_indices_of_transfers(struct precision *ppr, struct precision2 *ppr2,
                                   struct perturbs2 *ppt2, struct bessels *pbs,
                                   struct bessels2 *pbs2, struct transfers *ptr,
                                   struct transfers2 *ptr2);

/* include/transfer2.h:356-363; source/transfer2.c:1081-1103 */
int transfer2_get_l_list(struct precision *ppr, struct precision2 *ppr2,
                         struct perturbs2 *ppt2, struct bessels *pbs,
                         struct bessels2 *pbs2, struct transfers2 *ptr2);

/* include/transfer2.h:366-373; source/transfer2.c:1121-1285 */
int transfer2_get_lm_lists(struct precision *ppr, struct precision2 *ppr2,
                           struct perturbs2 *ppt2, struct bessels *pbs,
                           struct bessels2 *pbs2, struct transfers2 *ptr2);

/* include/transfer2.h:377-385; source/transfer2.c:1752-1834 */
int transfer2_get_k3_sizes(struct precision *ppr, struct precision2 *ppr2,
                           struct perturbs2 *ppt2, struct bessels *pbs,
                           struct bessels2 *pbs2, struct transfers *ptr,
                           struct transfers2 *ptr2);

/* include/transfer2.h:398-412; source/transfer2.c:2612-2763 */
int transfer2_interpolate_sources_in_k(struct precision *ppr, struct precision2 *ppr2,
                                       struct perturbs *ppt, struct perturbs2 *ppt2,
                                       struct bessels *pbs, struct bessels2 *pbs2,
                                       struct transfers2 *ptr2, int index_k1,
                                       int index_k2, int index_tp, double *k_grid,
                                       double *splines, double *interpolated_sources);

/* include/transfer2.h:415-427; source/transfer2.c:2366-2585 */
int transfer2_get_time_grid(struct precision *ppr, struct precision2 *ppr2,
                            struct perturbs *ppt, struct perturbs2 *ppt2,
                            struct bessels *pbs, struct bessels2 *pbs2,
                            struct transfers2 *ptr2, int index_k1, int index_k2,
                            int index_k, struct transfer2_workspace *pw);

/* include/transfer2.h:429-442; source/transfer2.c:2783-2908 */
int transfer2_interpolate_sources_in_time(struct precision *ppr, struct precision2 *ppr2,
                                          struct perturbs *ppt, struct perturbs2 *ppt2,
                                          struct bessels *pbs, struct bessels2 *pbs2,
                                          struct transfers2 *ptr2, int index_tp2,
                                          double *interpolated_sources_in_k,
                                          double *sources_time_spline,
                                          double *interpolated_sources_in_time,
                                          struct transfer2_workspace *pw);

/* include/transfer2.h:444-453; source/transfer2.c:1329-1551 */
int transfer2_get_k3_size(struct precision *ppr, struct precision2 *ppr2,
                          struct perturbs2 *ppt2, struct bessels *pbs,
                          struct bessels2 *pbs2, struct transfers2 *ptr2,
                          int index_k1, int index_k2);

/* include/transfer2.h:456-467; source/transfer2.c:1566-1737 */
int transfer2_get_k3_list(struct precision *ppr, struct precision2 *ppr2,
                          struct perturbs2 *ppt2, struct bessels *pbs,
                          struct bessels2 *pbs2, struct transfers2 *ptr2,
                          int index_k1, int index_k2, double *k3,
                          int *last_used_index_pt);

/* include/transfer2.h:469-484; source/transfer2.c:1861-2116 (CPU, one element) */
int transfer2_compute(struct precision *ppr, struct precision2 *ppr2,
                      struct perturbs2 *ppt2, struct bessels *pbs,
                      struct bessels2 *pbs2, struct transfers2 *ptr2,
                      int index_k1, int index_k2, int index_k, int index_l,
                      int index_m, int index_tt, double **interpolated_sources_in_time,
                      struct transfer2_workspace *pw);

/* include/transfer2.h:486-503; source/transfer2.c:2140-2341 (CPU, one integral) */
int transfer2_integrate(struct precision *ppr, struct precision2 *ppr2,
                        struct perturbs2 *ppt2, struct bessels *pbs,
                        struct bessels2 *pbs2, struct transfers2 *ptr2,
                        int index_k1, int index_k2, int index_k, int index_l,
                        int index_m, double **interpolated_sources_in_time,
                        int index_J, int index_source_monopole,
                        struct transfer2_workspace *pw, double *integral);

/* include/transfer2.h:506-510; source/transfer2.c:2929-2963 */
int transfer2_store_transfers_to_disk(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                      int index_k1);

/* include/transfer2.h:512-516; source/transfer2.c:705-763 */
int transfer2_load_transfers_from_disk(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                       int index_tt);

/* include/transfer2.h:519-523; source/transfer2.c:636-691 */
int transfer2_allocate_type_level(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                  int index_tt);

/* include/transfer2.h:525-529; source/transfer2.c:3038-3060 */
int transfer2_free_type_level(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                              int index_tt);

/* include/transfer2.h:531-535; source/transfer2.c:2976-3032 */
int transfer2_allocate_k1_level(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                                int index_k1);

/* include/transfer2.h:537-541; source/transfer2.c:3069-3104 */
int transfer2_free_k1_level(struct perturbs2 *ppt2, struct transfers2 *ptr2,
                            int index_k1);

/*
 * Producer-side helpers the stage calls on the sources it consumes. SONG
 * defines them in source/perturbations2.c:457-508 (allocate), :521-602
 * (load) and :609-636 (free); this library carries weak definitions with the
 * same behaviour so it also runs standalone, and a SONG link overrides them.
 */
int perturb2_allocate_k1_level(struct perturbs2 *ppt2, int index_k1);
int perturb2_load_sources_from_disk(struct perturbs2 *ppt2, int index_k1);
int perturb2_free_k1_level(struct perturbs2 *ppt2, int index_k1);
/* source/perturbations2.c:15347-15414 (perturb2_store_sources_to_disk): write
   the k1 level of ppt2->sources to ppt2->sources_paths[index_k1], the file
   perturb2_load_sources_from_disk reads back. Weak as the three above. */
int perturb2_store_sources_to_disk(struct perturbs2 *ppt2, int index_k1);

#ifdef __cplusplus
}
#endif

#endif /* SONG_LOS_TRANSFER2_H */
