// los_engine_impl.h — internals of the LOS engine shared by the batch engine
// (los_engine.hip) and the job server (los_serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "los_device.h"
#include "los_common.h"

extern "C" hipError_t los_launch_relayout_spline(const los_pair *pairs, int npairs, int nc_max,
                                                 const double *kc, const double *spl,
                                                 const double *raw, double *src, double *dd,
                                                 int ntau, int tp2, int cubic, hipStream_t s);
extern "C" hipError_t los_launch_integrate_exact(const los_params *P, int nthreads, hipStream_t s);
extern "C" hipError_t los_launch_stream(int fp32, const los_params *P, const los_shape *sh,
                                        int n_cu, hipStream_t s);
extern "C" hipError_t los_stream_set_attrs(int tp2_size);
extern "C" int los_shapes(int fp32, int tp2_size, los_shape *out, int cap);
extern "C" hipError_t los_launch_sweep_prep(int fp32, const los_params *P, int f_begin,
                                            int f_end, hipStream_t s);
extern "C" hipError_t los_launch_sweep(int fp32, const los_params *P, int nk, int pers, int pairs,
                                       int n_cu, hipStream_t s);
extern "C" hipError_t los_launch_sweep_prep_exact(const los_params *P, int f_begin, int f_end,
                                                  hipStream_t s);
extern "C" hipError_t los_launch_sweep_exact(const los_params *P, int nk, int pers, int n_cu,
                                             hipStream_t s);
extern "C" hipError_t los_launch_sweep_prep_fused_exact(const los_params *P, const double *raw,
                                                        const double *spl, int nc_max,
                                                        int p_begin, int p_end, hipStream_t s);
extern "C" hipError_t los_launch_sweep_prep_fused(int fp32, const los_params *P, const double *raw,
                                                  const double *spl, int nc_max, int p_begin,
                                                  int p_end, hipStream_t s);
extern "C" int los_sweep_fused_fits(int tp2, int nc_max);
extern "C" hipError_t los_launch_sweep_pairs(int fp32, const void *src, void *dst,
                                             const int *xminp, long rows, int X, int NLP,
                                             hipStream_t s);
extern "C" hipError_t los_launch_klane_prep(int fp32, const los_params *P, hipStream_t s);
extern "C" hipError_t los_launch_klane(int fp32, const los_params *P, int nlb, hipStream_t s);
extern "C" hipError_t los_launch_klane_prep_exact(const los_params *P, hipStream_t s);
extern "C" hipError_t los_launch_klane_exact(const los_params *P, int nlb, hipStream_t s);
extern "C" hipError_t los_launch_round_j32(const double *src, float *dst, long n, hipStream_t s);

struct los_engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int form = LOS_FORM_FP64;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  float ms_spline = 0, ms_los = 0;

  // pipelined steps: the batch in chunks of whole pairs; chunk c's re-layout,
  // spline and sweep tables run on stream2 while chunk c-1's LOS kernel runs
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_chunk;
  int chunks_forced = 0;              // chunk count set by the caller (0: one chunk)
  int n_chunk = 1;                    // chunk count of the loaded batch
  std::vector<int> chunk_pair, chunk_f, chunk_pos;  // [n_chunk + 1] boundaries
  std::vector<int> chunk_tile[5];     // [n_chunk + 1] tile boundaries per tile-size slot

  // projection tables (resident across loads)
  const void *J_owner = nullptr;
  double *dJ = nullptr;
  float *dJ32 = nullptr;
  size_t nJ = 0;
  int *dxmin = nullptr, *dminx = nullptr;
  double *dxx = nullptr;
  int *dl_of = nullptr, *dm_of = nullptr, *dmi_of = nullptr;
  int X = 0, L_size = 0, J_size = 0, n_lm = 0;
  int minx_all = INT_MAX;  // smallest first-live x index over all rows in use
  std::vector<int> hxmin;  // host copy of xmin [jt][L][lm]
  std::vector<int> hm_of;  // m of each (l,m) offset
  std::vector<double> ht0mt;  // tau0 - tau of the loaded time grid
  // the same table laid out for the sweep kernels: [jt][L][x][lp], lane slot
  // lp = mi * n_lp + il (n_lp = l count rounded up to 64)
  double *dJp = nullptr;
  float *dJp32 = nullptr;
  size_t nJp = 0;
  void *dJp2[2] = {nullptr, nullptr};  // (J[x], J[x+1]) pair tables, FP64 and FP32
  // cubic interpolation of J in x (plain FP64 sweep kernels only): second
  // derivatives in the sweep layout, and as (dd[x], dd[x+1]) pairs
  bool cubic_J = false, tables_cubic = false;
  double *dDp = nullptr;
  void *dDp2 = nullptr;
  int *dxminp = nullptr, *dlm_of_lp = nullptr;
  int n_lp = 0, n_wave = 0, m_size = 0;
  std::vector<int> wave_minx;  // smallest first-live x index over the lanes of each wave class
  // and for the k-lane kernels: [jt][L][lm][x]
  double *dJk = nullptr;
  float *dJk32 = nullptr;
  size_t nJk = 0;
  int *dkgrp[3] = {nullptr, nullptr, nullptr};  // lane groups of 1, 2, 4 (l,m) offsets
  int kngrp[3] = {0, 0, 0};
  double J_window_bytes = 0;

  // time grid
  double *dt0mt = nullptr, *ddtau = nullptr;
  int ntau = 0;

  // loaded pairs
  std::vector<los_pair> pairs;
  los_pair *dpairs = nullptr;
  double *dkc = nullptr, *dkf = nullptr, *dspl = nullptr;
  int *dfpair = nullptr, *dfsrc = nullptr, *dfbr = nullptr, *dforder = nullptr;
  int *dposjob = nullptr, *docol = nullptr;
  los_job *djobs = nullptr;
  std::vector<los_tile> htiles[5];     // tiles of 1, 2, 4, 8, 64 positions
  los_tile *dtiles[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double *draw = nullptr, *dsrc = nullptr, *ddd = nullptr, *dout = nullptr;
  long src_total = 0;
  int nc_max = 0;
  int F = 0;
  int tt2_size = 0;
  los_params P;
  bool loaded = false;
  // sweep tables straight from the raw sources (los_sweep_prep_fused) when the
  // batch fits its LDS budget and fused_prep is on (off by default: on c2 it
  // took 18.7 ms against 16.5 ms for re-layout + spline + tables in two passes,
  // its 504-byte records being written from 316,800 one-step workgroups)
  bool fused_prep = false;
  bool pipeline = false;  // pipelined steps (chunks_forced >= 2 and a sweep shape)

  double *staging = nullptr;  // pinned
  size_t staging_bytes = 0;

  int *dtau_end_fw = nullptr;          // per (fine k3, wave class)
  void *dixa = nullptr, *dsf = nullptr;  // per-(k3, tau) tables of the sweep kernels
  int sweep_form = -1;                   // form whose tables they hold (-1: none)
  size_t sweep_bytes = 0;
  int *drank = nullptr;                  // rank (sorted position) of each fine index
  void *dkixa = nullptr, *dksrc = nullptr;  // tables of the k-lane kernels
  int klane_form = -1;                   // form whose tables they hold (-1: none)
  size_t klane_bytes = 0;

  int n_cu = 256;
  int *dqueue = nullptr;               // work-queue counters of the persistent kernels
  std::vector<int> hforder;            // work order (host copy)
  std::vector<double> hkf;             // fine k3 (host copy)
  std::vector<double> pos_cost;        // estimated cost of each sorted position
  std::vector<double> pos_cost_sweep;  // the same for the sweep kernels
  // kernel kind and shape per form (plain FP64, FP64-FMA, FP32C)
  los_shape shape[3] = {{0, 1, 0, 0}, {0, 4, 32, 0}, {0, 4, 32, 0}};
  bool shape_set[3] = {false, false, false};  // chosen by the autotuner or the caller

  // captured run (transpose, spline, LOS) of the loaded batch
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_form = -1;
};

#define HIPCHK(expr, err)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      snprintf((err), _ERRORMSGSIZE_, "%s: %s failed: %s", __func__, #expr,                \
               hipGetErrorString(e_));                                                     \
      return _FAILURE_;                                                                    \
    }                                                                                      \
  } while (0)

template <class T>
static inline void dfree(T *&p)
{
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// Copy host memory to the device through the pinned staging buffer (waits).
int upload(los_engine *e, void *dst, const void *src, size_t bytes, char *err);

template <class T>
static inline int upload_vec(los_engine *e, T *&dptr, const std::vector<T> &v, char *err)
{
  dfree(dptr);
  size_t bytes = std::max<size_t>(1, v.size()) * sizeof(T);
  HIPCHK(hipMalloc((void **)&dptr, bytes), err);
  if (!v.empty()) return upload(e, dptr, v.data(), v.size() * sizeof(T), err);
  return _SUCCESS_;
}


// Engine steps used by the server as well.
void split_xcd(const std::vector<los_tile> &tiles, const std::vector<double> &pos_cost, int out[9]);
int nk_slot(int nk);
std::vector<los_tile> make_tiles(const std::vector<int> &job_end, int nk);
int enqueue_prep(los_engine *e, const los_shape *sh, char *err);
int ensure_j32(los_engine *e, char *err);
int current_shape(los_engine *e, const los_shape **sh, char *err);
bool sweep_fits(const los_engine *e, int form, const los_shape &sh);
int ensure_sweep(los_engine *e, const los_shape *sh, char *err);
int make_sweep_tables(los_engine *e, char *err);
int make_klane_tables(los_engine *e, char *err);
int ensure_klane(los_engine *e, const los_shape *sh, char *err);
int enqueue_los_params(los_engine *e, const los_shape *sh, los_params *P, hipStream_t s,
                       const std::vector<los_tile> &tiles, const std::vector<double> &cost,
                       char *err);
