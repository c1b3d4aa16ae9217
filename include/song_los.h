/*
 * song_los.h — engine-level C-ABI of the MI355X LOS transfer library.
 *
 * transfer2.h exposes SONG's own entry points; this header exposes what sits
 * under them, with plain C types only, for callers that want to keep inputs
 * resident on the GPU (benchmarks, a serving loop, other language bindings):
 *
 *  - los_engine_*   : upload the source and projection tables of a set of
 *                     (k1,k2) pairs once, run the LOS kernels on them as many
 *                     times as wanted, scatter the results into ptr2->transfer.
 *  - los_scenario_* : build seeded synthetic inputs (k grids, source tables,
 *                     J_Llm(x) tables) in SONG's structures, for tests,
 *                     benchmarks and users without a perturbations2 run.
 *  - los_bessel2_build, los_bessel2_build_gpu : FP64 builders of struct
 *                     bessels2 (the J_Llm(x) tables that source/bessel2.c:75-468
 *                     produces in SONG), on the host or on the GPU.
 *
 * All functions return _SUCCESS_ (0) or _FAILURE_ (1) and write a message into
 * the caller's ErrorMsg buffer on failure.
 */
#ifndef SONG_LOS_H
#define SONG_LOS_H

#ifdef __cplusplus
extern "C" {
#endif

/* SONG's and CLASS's structures, by pointer only: this header compiles next to
   SONG's own headers as well as next to song_los_types.h. */
struct precision;
struct precision2;
struct background;
struct thermo;
struct perturbs;
struct perturbs2;
struct bessels;
struct bessels2;
struct transfers;
struct transfers2;

#ifndef _SUCCESS_
#define _SUCCESS_ 0
#define _FAILURE_ 1
#endif

#define LOS_M_LIST_MAX 14 /* entries of los_config.m_list (SONG's _MAX_NUM_AZIMUTHAL_) */

/* ------------------------------------------------------------------------
 * Kernel forms.
 *  LOS_FORM_FP64        : one workgroup per fine k3, per-element operation
 *                         order of source/transfer2.c:2192-2329, no FMA
 *                         contraction: bit-identical to the FP64 CPU path.
 *  LOS_FORM_FP64_FAST   : tiled FP64 kernel with FMA; same math, different
 *                         rounding (1e-10 relative class).
 *  LOS_FORM_FP32C       : source x projection products in FP32 with
 *                         compensated (TwoSum) tau accumulation; x = k(tau0-tau),
 *                         its table index and weight, and the spline
 *                         coefficients stay FP64.
 * ---------------------------------------------------------------------- */
enum los_form { LOS_FORM_FP64 = 0, LOS_FORM_FP64_FAST = 1, LOS_FORM_FP32C = 2 };

typedef struct los_engine los_engine;

/* Create an engine on HIP device `device` (one stream). */
int los_engine_create(int device, los_engine **out, char *errmsg);
void los_engine_destroy(los_engine *eng);

/* Select the kernel form used by later runs (and by transfer2_init). */
int los_engine_set_form(los_engine *eng, int form, char *errmsg);

/*
 * Upload everything the LOS kernels need for the pairs with index_k1 in
 * [k1_begin, k1_end) (all index_k2 <= index_k1): sources ppt2->sources,
 * coarse and fine k3 grids, the time grid and the J tables. Requires
 * transfer2_indices_of_transfers() to have filled ptr2. The sources are
 * read, not freed. J tables are uploaded once per engine and kept resident.
 */
int los_engine_load(los_engine *eng, struct precision *ppr, struct precision2 *ppr2,
                    struct perturbs2 *ppt2, struct bessels *pbs, struct bessels2 *pbs2,
                    struct transfers2 *ptr2, int k1_begin, int k1_end, char *errmsg);

/* The same for an explicit list of (index_k1, index_k2) pairs (0 <= index_k2 <=
   index_k1 < ppt2->k_size), in any order and with any k3 counts: the batched
   form. Results are stored per pair as by los_engine_load. */
int los_engine_load_pairs(los_engine *eng, struct precision *ppr, struct precision2 *ppr2,
                          struct perturbs2 *ppt2, struct bessels *pbs, struct bessels2 *pbs2,
                          struct transfers2 *ptr2, const int *index_k1, const int *index_k2,
                          int n_pairs, char *errmsg);

/* Run the LOS kernels on the loaded pairs; *ms gets the GPU time (HIP events).
   Runs the captured graph when los_engine_capture was called for this form. */
int los_engine_run(los_engine *eng, float *ms, char *errmsg);

/* Pipelined steps (off by default): with on = n >= 2, the next load splits the
   batch into n chunks of whole pairs, and a step of a sweep tile shape runs each
   chunk's source re-layout, k3 spline and per-(k3, tau) tables on a second stream
   while the previous chunk's LOS kernel runs; los_engine_run then reports the step
   time only. on = 0 or 1: one chunk, one pass of each kernel, timed separately.
   On MI355X the two parts share the texture/L2 path of the gathers and the
   chunked order costs the LOS kernel locality, so the default stays off.
   No counterpart in SONG. */
int los_engine_set_pipeline(los_engine *e, int on, char *errmsg);

/* Sweep shapes: with on != 0, the per-(k3, tau) tables are made straight from
   the sources as SONG lays them out, splined in LDS, when one time step of
   every source type fits in 64 KiB of LDS; with on == 0 (the default) the
   sources are first re-laid out and splined in HBM. Same values either way. */
int los_engine_set_fused_prep(los_engine *e, int on, char *errmsg);

/* GPU time of the last eager run's LOS kernel and of everything before it (source
   re-layout, k3 spline, per-(k3,tau) tables of the kernel shape), in ms; -1 after
   a graph or pipelined run (only the total is known). */
int los_engine_last_kernel_ms(los_engine *eng, float *ms_los, float *ms_spline);

/* Kernel organisation and tile shape of a form (enum los_form).
   Plain FP64 form: kind 0, one workgroup per fine k3 and one lane per (l,m);
   kind 2, the k-lane kernel below. Both keep the reference's operations.
   Fast forms: kind 0, the "stream" kernel: nk fine k3 per workgroup, ch time
   steps per LDS chunk, persistent = 1 for the chip-sized grid fed by the device
   work queue; kind 1, the "sweep" kernel: one wave per (nk fine k3, m, 64 l)
   over per-(k3,tau) tables made before it; kind 2, the "k-lane" kernel: one wave
   per ((l,m), 64 neighbouring fine k3) over per-(k3,tau) tables.
   los_engine_autotune times every shape of the current form that fits on the
   loaded batch (reps timed launches each) and keeps the fastest; the choice is
   cached per batch shape in the process. Runs tune on first use unless a shape
   was set after the load. Arguments a kind does not use are ignored. */
int los_engine_autotune(los_engine *eng, int reps, char *errmsg);
int los_engine_set_shape(los_engine *eng, int form, int kind, int nk, int ch, int persistent,
                         char *errmsg);
void los_engine_get_shape(los_engine *eng, int form, int *kind, int *nk, int *ch, int *persistent);

/* Capture the whole run of the loaded batch (source re-layout, k3 spline, LOS
   kernel with its queue reset) in one hipGraph; later los_engine_run calls of
   the same form replay it. A new load drops the graph. */
int los_engine_capture(los_engine *eng, char *errmsg);

/* Copy the results of the loaded pairs into ptr2->transfer[tt][k1][k2][k]
   (levels must be allocated) and update ptr2->count_memorised_transfers. */
int los_engine_store(los_engine *eng, struct transfers2 *ptr2, char *errmsg);

/* Element count (tt x fine k3 summed over loaded pairs). */
long los_engine_n_elements(los_engine *eng);

/* The last run's transfers in device memory, without a copy to the host: out =
   [rows = tt2_size][cols = fine k3 of the loaded pairs, in load order], valid
   until the next load or destroy. Hands the stage's output to a consumer on the
   same device (e.g. los_cl2_sum_noalloc on the T or E rows). No counterpart in
   SONG. Returns _FAILURE_ when nothing is loaded. */
int los_engine_device_output(los_engine *eng, const double **out, long *rows, long *cols);
/* Algorithmic operation count of one run of the loaded batch: per element and
   integral, sum over tau of 5 per live L (J interpolation 3, multiply-add 2)
   plus 2 when any L is live (trapezoid); `live` applies the reference's
   index_xmin_J skip exactly, `dense` counts every L row. */
int los_engine_flops(los_engine *eng, double *live, double *dense);
/* A plain device copy of `bytes` (two launches of los_copy_probe, the second
   timed): the kernel of known traffic that HBM counter readings are scaled by. */
int los_engine_copy_probe(los_engine *eng, double bytes, float *ms, char *errmsg);
/* Algorithmic bytes of one LOS kernel launch of the current form: the coarse
   source values once (the spline's second derivatives are derived data and not
   counted), the live part of the J table once (FP32 for the FP32-compensated
   form), the outputs once. */
double los_engine_alg_bytes(los_engine *eng);
/* The HIP stream the engine launches on (hipStream_t as void*). */
void *los_engine_stream(los_engine *eng);

/* ------------------------------------------------------------------------
 * Serving: a persistent request loop over resident tables.
 *
 * los_server_create loads every (k1,k2) pair of the run (sources, coarse and
 * fine k3 grids, J tables) onto HIP device `device` once and prepares them
 * (k3 spline, per-(k3,tau) tables of the chosen kernel shape). Jobs are lists
 * of pairs with an l_max; they are admitted first come first served while the
 * device bytes of the admitted jobs stay within hbm_budget_bytes, and the jobs
 * admitted together run as one launch on one of n_slots streams, overlapping
 * with the other slots. A job's result is [tt][col]: tt runs over the fields
 * (T, E, B, those present) times the (l,m) offsets with l <= l_max, col over
 * the fine k3 of its pairs in the order given. Results do not depend on what
 * else runs: tiles never mix jobs and the kernel shape is fixed.
 * No counterpart in the reference (SONG runs the stage once per input file).
 * ---------------------------------------------------------------------- */
typedef struct los_server los_server;
int los_server_create(int device, int form, double hbm_budget_bytes, int n_slots,
                      struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, struct transfers2 *ptr2,
                      los_server **out, char *errmsg);
/* Finish the admitted jobs, fail the queued ones, free everything. */
void los_server_destroy(los_server *srv);
/* Queue a job; returns its id (>= 0) or -1 with errmsg set (unknown pair, no
   l <= l_max, or a job larger than the whole budget). Thread-safe. */
long los_server_submit(los_server *srv, const int *index_k1, const int *index_k2, int n_pairs,
                       int l_max, char *errmsg);
/* Pause (hold = 1) or resume (hold = 0) admissions; submitted jobs queue
   meanwhile and are admitted together on resume. */
void los_server_hold(los_server *srv, int hold);
/* Result size of a queued or finished job. */
int los_server_job_shape(los_server *srv, long job, long *n_values, int *n_lm, int *n_cols);
/* Wait up to timeout_ms (< 0: no limit) for a job. 0: done, its n_values
   copied to out and the job released; 1: failed (released) or bad buffer;
   2: timed out. times (if not NULL) gets the steady-clock seconds of submit,
   admission and completion. */
int los_server_wait(los_server *srv, long job, double timeout_ms, double *out, long n,
                    double *times, char *errmsg);
/* out[0..5]: jobs done, launches, most slots busy at once, peak admitted
   bytes, resident bytes, budget. */
int los_server_stats(los_server *srv, double *out, int n);

/* The part of transfer2_init that precedes the integral: copies tau0, tau_rec
   and rs_rec, runs transfer2_indices_of_transfers (l, m, k3 bookkeeping) and
   allocates every k1 level of ptr2->transfer, so that engine runs can be
   stored with los_engine_store. Free with transfer2_free. */
int los_transfer2_prepare(struct precision *ppr, struct precision2 *ppr2, struct background *pba,
                          struct thermo *pth, struct perturbs *ppt, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, struct transfers *ptr,
                          struct transfers2 *ptr2);

/* Paths of the sources files, "<ppt2->sources_dir>/sources_%03d.dat" per k1
   (source/perturbations2.c:1051-1064: the set-up SONG's perturb2_init does when
   sources are stored to or loaded from disk), allocated on the first call;
   los_sources_free_paths releases them. */
int los_sources_set_paths(struct perturbs2 *ppt2);
void los_sources_free_paths(struct perturbs2 *ppt2);

/* Process-wide default form used by transfer2_init (default LOS_FORM_FP64),
   and the device it runs on (default 0). */
int los_set_default_form(int form);
int los_set_default_device(int device);

/* ------------------------------------------------------------------------
 * Synthetic inputs.
 * ---------------------------------------------------------------------- */
typedef struct los_config {
  /* fields */
  int has_T;          /* temperature */
  int has_E;          /* E polarisation (B sources appear when m_max > 0) */
  int m_list[LOS_M_LIST_MAX];
  int m_size;
  int l_max_los_t;    /* sources L = 0..l_max_los_t for T */
  int l_max_los_p;    /* and for E, B */
  /* (k1,k2) grid: k_size values in [k_min,k_max], log (k_log=1) or linear (0) spaced,
     or (k_log=2) drawn log-uniformly from the seed and sorted */
  int k_size;
  double k_min, k_max;
  int k_log;
  /* coarse k3 per pair: k3_size values, linear in the triangle-limited range */
  int k3_size;
  /* time sampling of the sources: tau_size values log-spaced in [tau_min,tau_max] */
  int tau_size;
  double tau_min, tau_max;
  double tau0, tau_rec, rs_rec;
  /* multipoles today: l_size values from 2 to l_max (log then linear) */
  int l_size;
  int l_max;
  /* fine k3 step (sets q_linstep_song so that k_step_max equals it) */
  double dk_fine;
  /* J tables */
  double xx_max, xx_step;
  double J_cut, j_cut;
  /* options of the stage */
  int k3_interpolation;      /* enum interpolation_methods */
  int bessels_interpolation; /* enum interpolation_methods */
  int tau_sampling;          /* enum transfer2_tau_sampling */
  int time_interpolation;    /* enum interpolation_methods */
  double tau_linstep;
  int k3_sampling;           /* enum transfer2_k3_sampling */
  int extrapolation;         /* enum bispectra_k3_extrapolation */
  double extra_osc_left, extra_osc_right;
  /* seeded source model: A ~ N(0,1) per (field,L,m,k1,k2) x (k3/0.01)^-1.5 x
     [exp(-(tau-tau_peak)^2/(2 sigma^2)) + ramp_amp * ramp(tau)] */
  unsigned long long seed;
  double tau_peak, tau_sigma, ramp_amp, ramp_center, ramp_width;
  /* verbosity of the stage */
  int verbose;
  /* J tables: built on the host (0) or on the GPU (1, default device) */
  int j_build;
} los_config;

/* Defaults: a small plain FP64 temperature-only configuration (m = 0, L = 0..4, 4x4 pairs). */
void los_config_default(los_config *cfg);

typedef struct los_scenario los_scenario;

int los_scenario_create(const los_config *cfg, los_scenario **out, char *errmsg);
void los_scenario_destroy(los_scenario *s);
int los_scenario_override_ddJ(los_scenario *s, int J, int L, int il, int im, const double *v,
                              int n, char *errmsg);
void los_scenario_structs(los_scenario *s, struct precision **ppr, struct precision2 **ppr2,
                          struct background **pba, struct thermo **pth,
                          struct perturbs **ppt, struct perturbs2 **ppt2,
                          struct bessels **pbs, struct bessels2 **pbs2,
                          struct transfers **ptr, struct transfers2 **ptr2);

/* Total number of values in ptr2->transfer (after transfer2_init) and a copy of
   them in (tt, k1, k2<=k1, k) order. */
long los_scenario_n_transfers(los_scenario *s);
int los_scenario_gather(los_scenario *s, double *out, long n, char *errmsg);

/* Set a named scalar of the structures (e.g. "store_transfers_to_disk";
   "transfers_dir" and "sources_dir" are set with los_scenario_set_string, the
   latter also setting the sources file paths). */
int los_scenario_set(los_scenario *s, const char *key, double value, char *errmsg);
int los_scenario_set_string(los_scenario *s, const char *key, const char *value,
                            char *errmsg);

/* Input overrides, for tests that need hand-chosen grids or analytic tables:
   the (k1,k2) values (same count), one pair's coarse k3 grid (its sources are
   reset to zero with the new size), one J row stored from index xmin to the end
   of the x grid. Plus the inverse of los_scenario_gather and a readout of the
   k3 bookkeeping (k_size, physical start, physical size per pair). */
int los_scenario_grids(los_scenario *s, int *l_out, double *k_out, double *tau_out);
int los_scenario_override_k(los_scenario *s, const double *k, int n, char *errmsg);
int los_scenario_override_k3(los_scenario *s, int index_k1, int index_k2, const double *k3,
                             int n, char *errmsg);
int los_scenario_override_J(los_scenario *s, int index_J, int index_L, int index_l,
                            int index_m, int xmin, const double *values, int n, char *errmsg);
int los_scenario_scatter(los_scenario *s, const double *in, long n, char *errmsg);
int los_scenario_k3_sizes(los_scenario *s, int *out, long n, char *errmsg);
/* Write every allocated k1 level of the sources to its sources file and free
   it, as SONG's perturbations2 stage does with store_sources_to_disk; needs
   "sources_dir" set. */
int los_scenario_store_sources(los_scenario *s, char *errmsg);
double *los_scenario_source_ptr(los_scenario *s, int index_tp, int index_k1, int index_k2);
double *los_scenario_J_row(los_scenario *s, int index_J, int index_L, int index_l, int index_m,
                           int *xmin, int *xsize);
/* Free the scenario's J tables and build them again on the host (gpu = 0) or on
   HIP device `device` (gpu = 1); *seconds receives the wall time of the build. */
int los_scenario_rebuild_J(los_scenario *s, int gpu, int device, double *seconds,
                           char *errmsg);

/* FP64 builder of the projection-function tables of struct bessels2, with the
   semantics of source/bessel2.c:75-468 (xx grid, index_xmin_J by bisection on
   J_Llm_cut, rows stored from index_xmin_J, ddJ when bessels are cubic). */
int los_bessel2_build(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                      struct bessels *pbs, struct bessels2 *pbs2, char *errmsg);
int los_bessel2_free(struct bessels *pbs, struct bessels2 *pbs2);

/* The same tables built on HIP device `device`: the j_l(x) table of every grid
   point, the index_xmin_J bisections and the row fills run on the GPU
   (source/bessel2.c:1041-1588, :1669-1815 semantics); the rows land in the
   host arrays of struct bessels2 as with los_bessel2_build. No counterpart
   in SONG, whose bessel2 module builds them on the host. */
int los_bessel2_build_gpu(struct precision *ppr, struct precision2 *ppr2, struct perturbs2 *ppt2,
                          struct bessels *pbs, struct bessels2 *pbs2, int device,
                          char *errmsg);

/* Special functions used by the builder (exported for tests). */
void los_sph_bessel_array(int lmax, double x, double *out);
int los_wigner3j_j1(int j2, int j3, int m2, int m3, int *j1min, int *j1max, double *out,
                    int out_cap);

/* ------------------------------------------------------------------------
 * k3 integral of the transfer functions (the consumer of this stage in the
 * bispectrum module): the loop over (l3, k1, k2, r) of
 * bispectra2_intrinsic_integrate_over_k3 (source/bispectra2.c:1182-1410) with
 * bessel2_convolution (source/bessel2.c:499-605, linear interpolation, g = 1)
 * on HIP device `device`, plain FP64 with the reference's operation order:
 *   I[q][r][p] = 2 * (coef * (0.5 * sum_k3 k3*k3*j_L(k3 r)*T_qp(k3) * dk3)).
 * Row q is one transfer row (SONG: tt = index_tt2_k3 + lm_array[l3][M3]) over
 * every pair; index_L[q] is the j table of its L3 (-1: the reference skips the
 * row, which stays zero); j tables are j_l1 from index_xmin_l1 on, with their
 * x_min_l1 (bessel2.h:95-106). Inputs are uploaded once by _create and stay
 * resident; _run launches the kernel `reps` times and gives the HIP-event
 * time per launch; _result copies I back in [q][r][p] order.
 * ---------------------------------------------------------------------- */
typedef struct los_k3int_desc {
  int n_pairs;             /* (k1,k2) pairs */
  const long *pair_off;    /* n_pairs + 1 offsets of each pair's k3 points */
  const double *k3;        /* concatenated k3 grids (transfer2_get_k3_list) */
  int n_q;                 /* transfer rows */
  const double *transfer;  /* [q][pair_off[n_pairs]] */
  const int *index_L;      /* per row: j table of L3, or -1 */
  int n_j;                 /* j tables */
  const long *j_off;       /* n_j + 1 offsets into j */
  const double *j;         /* j_l1[t][x - index_xmin_l1[t]] */
  const double *x_min;     /* x_min_l1[t] */
  double xx_step;          /* x grid step */
  int n_r;                 /* r grid */
  const double *r;
  double coef;             /* M3 coefficient (1 for M3 = 0) */
} los_k3int_desc;

typedef struct los_k3int los_k3int;
int los_k3int_create(const los_k3int_desc *desc, int device, los_k3int **out, char *errmsg);
int los_k3int_run(los_k3int *h, int reps, float *ms, char *errmsg);
int los_k3int_result(los_k3int *h, double *out, long n, char *errmsg);
/* out[0] live terms (T != 0 and k3 r >= x_min), out[1] algorithmic bytes per
   launch (inputs once, the j windows reached once, outputs once). */
int los_k3int_counts(los_k3int *h, double *out, int n);
void los_k3int_destroy(los_k3int *h);

/* ------------------------------------------------------------------------
 * Second-order C_l: the (k1, k2, k3) sum of spectra2_cls
 * (source/spectra2.c:241-590, k3 loop :466-540) on HIP device `device`, FP64,
 * for every l (n_l) and m (m_size) of the run:
 *   cl_m[l][M] = sum_{k1, k2<=k1, k3} 2 k1 k2 k3 dk1 dk2 dk3 [scale^2]
 *                * 2/pi/2/pi/2/pi/2/pi * P(k1) P(k2) * T1 T2
 * with the reference's trapezoid half steps and triangle-edge corrections,
 * P(k) = 2 pi^2 / k^3 pk[k] (pk: primordial_spectrum_at_k, linear) and
 * scale^2 = (1 - cos^2)^m when rescale (ppt2->rescale_cmb_sources). Pairs are
 * k1-major with k2 <= k1 over the whole k grid; transfer1/transfer2 hold the
 * rows [l][M][k3 of all pairs] of the two fields of the spectrum (TT, EE, TE,
 * BB); used[l][M] = 0 skips a row (BB with m = 0, :330-333). C_l itself is the
 * sum over M. Results agree with the sequential sum to rounding.
 * ---------------------------------------------------------------------- */
typedef struct los_cl2_desc {
  int k_size;               /* ppt2->k grid */
  const double *k;
  const double *pk;         /* primordial spectrum at each k */
  const long *pair_off;     /* k_size (k_size+1)/2 + 1 offsets into k3 */
  const double *k3;         /* concatenated k3 grids (transfer2_get_k3_list) */
  int n_l;
  int m_size;
  const int *m;             /* the m values */
  const double *transfer1;  /* [n_l][m_size][n_k3] */
  const double *transfer2;  /* [n_l][m_size][n_k3] */
  const int *used;          /* [n_l][m_size] */
  int rescale;              /* ppt2->rescale_cmb_sources */
} los_cl2_desc;

typedef struct los_cl2 los_cl2;
int los_cl2_create(const los_cl2_desc *desc, int device, los_cl2 **out, char *errmsg);
int los_cl2_run(los_cl2 *h, int reps, float *ms, char *errmsg);
int los_cl2_result(los_cl2 *h, double *cl_m, long n, char *errmsg);
void los_cl2_destroy(los_cl2 *h);

/* The C_l sum of los_cl2_create + los_cl2_run + los_cl2_result in one call
   that allocates nothing. No counterpart in the reference: it is for callers
   that cannot allocate device memory (code under HIP graph capture, serving
   loops). Every pointer of `desc` and cl_m ([n_l][m_size]) are device
   pointers of the current HIP device; n_k3 = pair_off[k_size (k_size+1)/2];
   the work is enqueued on `stream` (a hipStream_t; NULL for the default
   stream) and not waited for. The inputs must satisfy what los_cl2_create
   checks (two or more k3 points per pair, positive k and k3, m >= 0). Limits:
   k_size <= 128, n_k3 <= 2^21, n_l * m_size <= 8192 (larger inputs are
   refused with a message). The per-pair tables and partial sums use a fixed
   10.4 MiB device array per GPU, held until the process ends and shared by
   every call on that GPU: calls that overlap on one GPU race. Same kernels as
   los_cl2_run, so the results are the same bit for bit. */
int los_cl2_sum_noalloc(const los_cl2_desc *desc, long n_k3, double *cl_m, void *stream,
                        char *errmsg);

#ifdef __cplusplus
}
#endif

#endif /* SONG_LOS_H */
