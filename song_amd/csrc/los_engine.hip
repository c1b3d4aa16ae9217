// los_engine.hip — host side of the LOS engine: device residency of the
// source and projection tables, kernel launches, result scatter.
//
// Data layout in HBM (one engine, one HIP device, one stream):
//   src, dd   per pair [tp][tau][kc] FP64 (the layout of ppt2->sources, so a
//             pair uploads as tp2_size contiguous blocks; dd is its spline)
//   kc, kf    coarse and fine k3 grids, concatenated per pair
//   J         [jt][L][x][lm] FP64: the (l,m) offsets of a J row at one x are
//             contiguous, so the lanes of a wave (one lm each) load 512 B runs
//   xmin      [jt][L][lm] first live x index of each row
//   out       [tt][F] FP64, scattered to ptr2->transfer[tt][k1][k2][k]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "los_device.h"
#include "los_common.h"

#include "los_engine_impl.h"

// Copy host memory to the device through the pinned staging buffer.
int upload(los_engine *e, void *dst, const void *src, size_t bytes, char *err)
{
  const char *s = (const char *)src;
  char *d = (char *)dst;
  while (bytes > 0) {
    size_t n = std::min(bytes, e->staging_bytes);
    memcpy(e->staging, s, n);
    HIPCHK(hipMemcpyAsync(d, e->staging, n, hipMemcpyHostToDevice, e->stream), err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
    s += n;
    d += n;
    bytes -= n;
  }
  return _SUCCESS_;
}

static void drop_graph(los_engine *e)
{
  if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->graph_exec = nullptr;
  e->graph = nullptr;
  e->graph_form = -1;
}

static void free_batch(los_engine *e)
{
  drop_graph(e);
  dfree(e->dtau_end_fw);
  dfree(e->drank);
  if (e->dkixa) (void)hipFree(e->dkixa);
  if (e->dksrc) (void)hipFree(e->dksrc);
  e->dkixa = e->dksrc = nullptr;
  e->klane_form = -1;
  e->klane_bytes = 0;
  if (e->dixa) (void)hipFree(e->dixa);
  if (e->dsf) (void)hipFree(e->dsf);
  e->dixa = e->dsf = nullptr;
  e->sweep_form = -1;
  e->sweep_bytes = 0;
  dfree(e->dpairs);
  dfree(e->dkc);
  dfree(e->dkf);
  dfree(e->dfpair);
  dfree(e->dfsrc);
  dfree(e->dfbr);
  dfree(e->dforder);
  dfree(e->dposjob);
  dfree(e->docol);
  dfree(e->djobs);
  for (int i = 0; i < 5; ++i) {
    dfree(e->dtiles[i]);
    e->htiles[i].clear();
  }
  dfree(e->dspl);
  dfree(e->draw);
  dfree(e->dsrc);
  dfree(e->ddd);
  dfree(e->dout);
  dfree(e->dt0mt);
  dfree(e->ddtau);
  e->loaded = false;
}

static void free_tables(los_engine *e)
{
  dfree(e->dJ);
  dfree(e->dJ32);
  dfree(e->dxmin);
  dfree(e->dminx);
  dfree(e->dxx);
  dfree(e->dl_of);
  dfree(e->dm_of);
  dfree(e->dmi_of);
  dfree(e->dJp);
  dfree(e->dJp32);
  dfree(e->dDp);
  if (e->dDp2) (void)hipFree(e->dDp2);
  e->dDp2 = nullptr;
  e->tables_cubic = false;
  for (int i = 0; i < 2; ++i) {
    if (e->dJp2[i]) (void)hipFree(e->dJp2[i]);
    e->dJp2[i] = nullptr;
  }
  dfree(e->dJk);
  dfree(e->dJk32);
  for (int gs = 0; gs < 3; ++gs) dfree(e->dkgrp[gs]);
  dfree(e->dxminp);
  dfree(e->dlm_of_lp);
  e->J_owner = nullptr;
}

extern "C" int los_engine_create(int device, los_engine **out, char *err)
{
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n), err);
  if (device < 0 || device >= n) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_create: no HIP device %d (%d visible)", device, n);
    return _FAILURE_;
  }
  HIPCHK(hipSetDevice(device), err);
  los_engine *e = new los_engine();
  e->device = device;
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), err);
  HIPCHK(hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking), err);
  for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&e->ev[i]), err);
  HIPCHK(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming), err);
  e->staging_bytes = 64ul << 20;
  HIPCHK(hipHostMalloc((void **)&e->staging, e->staging_bytes, hipHostMallocDefault), err);
  HIPCHK(hipDeviceGetAttribute(&e->n_cu, hipDeviceAttributeMultiprocessorCount, device), err);
  HIPCHK(hipMalloc((void **)&e->dqueue, 64 * sizeof(int)), err);
  *out = e;
  return _SUCCESS_;
}

extern "C" void los_engine_destroy(los_engine *e)
{
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  free_batch(e);
  free_tables(e);
  dfree(e->dqueue);
  if (e->staging) (void)hipHostFree(e->staging);
  for (int i = 0; i < 4; ++i)
    if (e->ev[i]) (void)hipEventDestroy(e->ev[i]);
  if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
  for (hipEvent_t v : e->ev_chunk) (void)hipEventDestroy(v);
  if (e->stream2) {
    (void)hipStreamSynchronize(e->stream2);
    (void)hipStreamDestroy(e->stream2);
  }
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

extern "C" int los_engine_set_form(los_engine *e, int form, char *err)
{
  if (form != LOS_FORM_FP64 && form != LOS_FORM_FP64_FAST && form != LOS_FORM_FP32C) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_set_form: unknown form %d", form);
    return _FAILURE_;
  }
  e->form = form;
  return _SUCCESS_;
}

extern "C" void *los_engine_stream(los_engine *e) { return (void *)e->stream; }

extern "C" double los_engine_batch_budget(los_engine *e)
{
  size_t fr = 0, tot = 0;
  (void)hipSetDevice(e->device);
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 8e9;
  return 0.6 * (double)fr;
}

// Build the device projection table [jt][L][x][lm] for x < X.
static int build_tables(los_engine *e, const los_stage *st, double kmax_all, char *err)
{
  const int J_size = st->J_size, L_size = st->L_size, n_lm = st->n_transfers;
  const double t0mt0 = st->tau0 - st->tau[0];
  const double xmax = kmax_all * t0mt0;
  const int ixmax = (int)(xmax / st->xx_step);
  if (ixmax + 1 > st->xx_size - 1) {
    snprintf(err, _ERRORMSGSIZE_,
             "los_engine_load: not enough J's: k*(tau0-tau_min)=%g exceeds the x grid (xx_max=%g)",
             xmax, st->xx_max);
    return _FAILURE_;
  }
  const int X = std::min(ixmax + 2, st->xx_size);
  const size_t nJ = (size_t)J_size * L_size * X * n_lm;
  std::vector<double> hJ(nJ, 0.0);
  std::vector<double> hD(st->ddJ ? nJ : 0, 0.0);  // second derivatives (cubic J)
  std::vector<int> hxmin((size_t)J_size * L_size * n_lm, INT_MAX);
  double window = 0;
  for (int jt = 0; jt < J_size; ++jt)
    for (int L = 0; L < L_size; ++L)
      for (int il = 0; il < st->l_size; ++il) {
        const int l = st->l[il];
        const int mcount = std::min(st->index_m_max[L], st->index_m_max[l]) + 1;
        for (int im = 0; im < mcount; ++im) {
          const int lm = st->lm_array[il][im];
          const int xm = st->xmin_J[jt][L][il][im];
          const int xs = st->x_size_J[jt][L][il][im];
          hxmin[((size_t)jt * L_size + L) * n_lm + lm] = xm;
          if (xm >= X) continue;
          const double *row = st->J[jt][L][il][im];
          const int xend = std::min(X, xm + xs);
          for (int x = xm; x < xend; ++x)
            hJ[(((size_t)jt * L_size + L) * X + x) * n_lm + lm] = row[x - xm];
          if (st->ddJ) {
            const double *drow = st->ddJ[jt][L][il][im];
            for (int x = xm; x < xend; ++x)
              hD[(((size_t)jt * L_size + L) * X + x) * n_lm + lm] = drow[x - xm];
          }
          window += 8.0 * (xend - xm);
        }
      }
  // smallest first-live index over the rows each (l,m) element uses
  std::vector<int> hminx(n_lm, INT_MAX), hl(n_lm), hm(n_lm), hmi(n_lm);
  for (int il = 0; il < st->l_size; ++il)
    for (int im = 0; im <= st->index_m_max[st->l[il]]; ++im) {
      const int lm = st->lm_array[il][im];
      hl[lm] = st->l[il];
      hm[lm] = st->m[im];
      hmi[lm] = im;
      int mn = INT_MAX;
      auto scan = [&](int jt, int Lmax) {
        if (jt < 0) return;
        for (int L = 0; L <= Lmax && L < L_size; ++L) {
          if (st->m[im] > L) continue;
          mn = std::min(mn, hxmin[((size_t)jt * L_size + L) * n_lm + lm]);
        }
      };
      // rows an element reads: TT for T; EE for E and B; EB for B (any m) and
      // for E when m != 0
      if (st->has_T) scan(st->jTT, st->LmaxT);
      if (st->has_E || st->has_B) {
        scan(st->jEE, st->LmaxP);
        if (st->jEB >= 0 && (st->has_B || st->m[im] != 0)) scan(st->jEB, st->LmaxP);
      }
      hminx[lm] = mn;
    }
  // sweep layout: slot lp = mi * n_lp + il for every (l, m-list entry)
  const int m_size = st->m_size, n_l = st->l_size;
  const int n_lp = (n_l + 63) / 64 * 64, NLP = m_size * n_lp;
  std::vector<int> lm_of_lp(NLP, -1);
  for (int il = 0; il < n_l; ++il)
    for (int im = 0; im < m_size; ++im)
      if (im <= st->index_m_max[st->l[il]]) lm_of_lp[im * n_lp + il] = st->lm_array[il][im];
  const size_t nJp = (size_t)J_size * L_size * X * NLP;
  std::vector<double> hJp(nJp, 0.0), hDp(st->ddJ ? nJp : 0, 0.0);
  std::vector<int> hxminp((size_t)J_size * L_size * NLP, INT_MAX);
  for (size_t r = 0; r < (size_t)J_size * L_size; ++r)
    for (int lp = 0; lp < NLP; ++lp) {
      const int lm = lm_of_lp[lp];
      if (lm < 0) continue;
      hxminp[r * NLP + lp] = hxmin[r * n_lm + lm];
      for (int x = 0; x < X; ++x) hJp[(r * X + x) * NLP + lp] = hJ[(r * X + x) * n_lm + lm];
      if (st->ddJ)
        for (int x = 0; x < X; ++x) hDp[(r * X + x) * NLP + lp] = hD[(r * X + x) * n_lm + lm];
    }
  e->wave_minx.assign(NLP / 64, INT_MAX);
  for (int lp = 0; lp < NLP; ++lp)
    if (lm_of_lp[lp] >= 0) e->wave_minx[lp / 64] = std::min(e->wave_minx[lp / 64], hminx[lm_of_lp[lp]]);

  free_tables(e);
  e->hxmin = hxmin;
  e->hm_of = hm;
  e->n_lp = n_lp;
  e->m_size = m_size;
  e->n_wave = NLP / 64;
  e->nJp = nJp;
  {
    // k-lane layout [jt][L][lm][x]
    std::vector<double> hJk(nJ);
    for (size_t r = 0; r < (size_t)J_size * L_size; ++r)
      for (int x = 0; x < X; ++x)
        for (int lm = 0; lm < n_lm; ++lm)
          hJk[(r * n_lm + lm) * X + x] = hJ[(r * X + x) * n_lm + lm];
    e->nJk = nJ;
    HIPCHK(hipMalloc((void **)&e->dJk, nJ * sizeof(double)), err);
    if (upload(e, e->dJk, hJk.data(), nJ * sizeof(double), err)) return _FAILURE_;
    // lane groups: for each m-list entry, the (l,m) offsets in l order, cut
    // into groups of 1, 2 or 4
    for (int gs = 0; gs < 3; ++gs) {
      const int nlb = 1 << gs;
      std::vector<int> grp;
      for (int im = 0; im < m_size; ++im) {
        std::vector<int> lms;
        for (int il = 0; il < n_l; ++il)
          if (im <= st->index_m_max[st->l[il]]) lms.push_back(st->lm_array[il][im]);
        for (size_t q = 0; q < lms.size(); q += nlb)
          for (int j = 0; j < nlb; ++j) grp.push_back(q + j < lms.size() ? lms[q + j] : -1);
      }
      e->kngrp[gs] = (int)grp.size() / nlb;
      if (upload_vec(e, e->dkgrp[gs], grp, err)) return _FAILURE_;
    }
  }
  HIPCHK(hipMalloc((void **)&e->dJp, nJp * sizeof(double)), err);
  if (upload(e, e->dJp, hJp.data(), nJp * sizeof(double), err)) return _FAILURE_;
  if (st->ddJ) {
    HIPCHK(hipMalloc((void **)&e->dDp, nJp * sizeof(double)), err);
    if (upload(e, e->dDp, hDp.data(), nJp * sizeof(double), err)) return _FAILURE_;
  }
  if (upload_vec(e, e->dxminp, hxminp, err) || upload_vec(e, e->dlm_of_lp, lm_of_lp, err))
    return _FAILURE_;
  e->minx_all = INT_MAX;
  for (int v : hminx) e->minx_all = std::min(e->minx_all, v);
  HIPCHK(hipMalloc((void **)&e->dJ, nJ * sizeof(double)), err);
  if (upload(e, e->dJ, hJ.data(), nJ * sizeof(double), err)) return _FAILURE_;
  if (upload_vec(e, e->dxmin, hxmin, err)) return _FAILURE_;
  if (upload_vec(e, e->dminx, hminx, err)) return _FAILURE_;
  if (upload_vec(e, e->dl_of, hl, err)) return _FAILURE_;
  if (upload_vec(e, e->dm_of, hm, err)) return _FAILURE_;
  if (upload_vec(e, e->dmi_of, hmi, err)) return _FAILURE_;
  std::vector<double> hxx(st->xx, st->xx + X);
  if (upload_vec(e, e->dxx, hxx, err)) return _FAILURE_;
  e->X = X;
  e->L_size = L_size;
  e->J_size = J_size;
  e->n_lm = n_lm;
  e->nJ = nJ;
  e->J_owner = st->J_id;
  e->tables_cubic = st->ddJ != nullptr;
  e->J_window_bytes = window * (st->ddJ ? 2 : 1);
  return _SUCCESS_;
}

// Tiles dealt to the 8 XCDs as contiguous ranges of about equal estimated
// cost (the cost of a tile is that of its last, largest-k3, position).
void split_xcd(const std::vector<los_tile> &tiles, const std::vector<double> &pos_cost,
                      int out[9])
{
  const int nt = (int)tiles.size();
  std::vector<double> pre(nt + 1, 0.0);
  for (int t = 0; t < nt; ++t)
    pre[t + 1] = pre[t] + pos_cost[tiles[t].pos0 + tiles[t].npos - 1];
  out[0] = 0;
  int t = 0;
  for (int x = 1; x < 8; ++x) {
    const double goal = pre[nt] * x / 8.0;
    while (t < nt && pre[t + 1] <= goal) ++t;
    out[x] = t;
  }
  out[8] = nt;
}

int nk_slot(int nk) { return nk == 1 ? 0 : nk == 2 ? 1 : nk == 4 ? 2 : nk == 8 ? 3 : 4; }

// Tiles of at most nk consecutive positions that never straddle two jobs
// (job_end: one past the last position of each job, ascending).
std::vector<los_tile> make_tiles(const std::vector<int> &job_end, int nk)
{
  std::vector<los_tile> t;
  int begin = 0;
  for (int end : job_end) {
    for (int p = begin; p < end; p += nk) t.push_back(los_tile{p, std::min(nk, end - p)});
    begin = end;
  }
  return t;
}

extern "C" int los_engine_load_stage(los_engine *e, const los_stage *st, const int *pk1,
                                     const int *pk2, int npairs, char *err)
{
  HIPCHK(hipSetDevice(e->device), err);
  if (!st->tau_sources) {
    snprintf(err, _ERRORMSGSIZE_,
             "los_engine_load: only transfer2_tau_sampling=sources is supported by this build");
    return _FAILURE_;
  }
  if (!st->linear_J && st->ddJ == nullptr) {
    snprintf(err, _ERRORMSGSIZE_,
             "los_engine_load: cubic bessels_interpolation needs the second derivatives of J "
             "(pbs2->ddJ_Llm_x)");
    return _FAILURE_;
  }
  if (st->tp2_size <= 0 || st->tau_size < 2) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_load: empty source table");
    return _FAILURE_;
  }
  const int LmaxT = st->LmaxT, LmaxP = st->LmaxP;
  if (std::max(LmaxT, LmaxP) >= LOS_MAX_L || st->m_size > LOS_MAX_MI) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_load: l_max_los (%d) or m_size (%d) too large",
             std::max(LmaxT, LmaxP), st->m_size);
    return _FAILURE_;
  }
  for (int q = 0; q < npairs; ++q)
    if (pk1[q] < 0 || pk1[q] >= st->k_size || pk2[q] < 0 || pk2[q] > pk1[q]) {
      snprintf(err, _ERRORMSGSIZE_, "los_engine_load_pairs: pair %d = (%d,%d) is not a (k1,k2) "
               "pair with 0 <= index_k2 <= index_k1 < %d", q, pk1[q], pk2[q], st->k_size);
      return _FAILURE_;
    }
  free_batch(e);
  e->shape_set[0] = e->shape_set[1] = e->shape_set[2] = false;  // shapes are chosen per batch
  if (npairs == 0) {
    // an empty batch: nothing to compute; run, capture, autotune and store do nothing
    e->pairs.clear();
    e->F = 0;
    e->n_chunk = 1;
    e->tt2_size = st->tt2_size;
    e->loaded = true;
    return _SUCCESS_;
  }
  const int ntau = st->tau_size;
  const int tp2 = st->tp2_size;

  // time grid for transfer2_tau_sampling == sources (source/transfer2.c:2405-2412, :2567-2580)
  std::vector<double> t0mt(ntau), dtau(ntau);
  const double *tau = st->tau;
  for (int i = 0; i < ntau; ++i) t0mt[i] = st->tau0 - tau[i];
  dtau[0] = tau[1] - tau[0];
  for (int i = 1; i < ntau - 1; ++i) dtau[i] = tau[i + 1] - tau[i - 1];
  dtau[ntau - 1] = tau[ntau - 1] - tau[ntau - 2];

  // pairs, fine k3 grids and their brackets in the coarse grid
  e->pairs.clear();
  e->nc_max = 0;
  std::vector<double> kc, kf, spl;
  std::vector<int> fpair, fsrc, fbr;
  long src_off = 0;
  std::vector<double> klist(std::max(1, st->k3_size_max));
  for (int q = 0; q < npairs; ++q) {
    {
      const int i1 = pk1[q], i2 = pk2[q];
      if (!st->has_sources[i1]) {
        snprintf(err, _ERRORMSGSIZE_, "los_engine_load: sources of index_k1=%d are not allocated", i1);
        return _FAILURE_;
      }
      los_pair p;
      p.k1 = i1;
      p.k2 = i2;
      p.nc = st->k3_size[i1][i2];
      p.kc_off = (int)kc.size();
      p.src_off = src_off;
      p.f_off = (int)kf.size();
      p.nf = st->k_size_k1k2[i1][i2];
      if (p.nc > LOS_NC_MAX || p.nc < 4) {
        snprintf(err, _ERRORMSGSIZE_, "los_engine_load: coarse k3 count %d outside [4,%d]", p.nc,
                 LOS_NC_MAX);
        return _FAILURE_;
      }
      int last_used = 0;
      (void)last_used;
      if (st->k3_list(st, i1, i2, klist.data(), err) != _SUCCESS_) return _FAILURE_;
      const double *kpt = st->k3[i1][i2];
      const int first = st->k_phys_start[i1][i2];
      const int last = first + st->k_phys_size[i1][i2] - 1;
      // bracket walk of source/transfer2.c:2681-2686
      std::vector<int> br(p.nf, 0);
      int ik = 0;
      for (int it = first; it <= last; ++it) {
        while ((ik + 1) < p.nc && kpt[ik + 1] < klist[it]) ik++;
        if (ik > p.nc - 2) {
          snprintf(err, _ERRORMSGSIZE_, "los_engine_load: k3=%g beyond the coarse grid", klist[it]);
          return _FAILURE_;
        }
        br[it] = ik;
      }
      for (int q = 0; q < p.nc; ++q) kc.push_back(kpt[q]);
      // y-independent part of the tridiagonal sweep of the k3 spline
      {
        const double *x = kpt;
        std::vector<double> sig(p.nc, 0.0), pv(p.nc, 0.0), gam(p.nc, 0.0);
        gam[0] = -0.5;
        for (int i = 1; i < p.nc - 1; ++i) {
          sig[i] = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
          pv[i] = sig[i] * gam[i - 1] + 2.0;
          gam[i] = (sig[i] - 1.0) / pv[i];
        }
        spl.insert(spl.end(), sig.begin(), sig.end());
        spl.insert(spl.end(), pv.begin(), pv.end());
        spl.insert(spl.end(), gam.begin(), gam.end());
      }
      e->nc_max = std::max(e->nc_max, p.nc);
      for (int it = 0; it < p.nf; ++it) {
        const int s = it < first ? first : (it > last ? last : it);
        kf.push_back(klist[it]);
        fpair.push_back((int)e->pairs.size());
        fsrc.push_back(p.f_off + s);
        fbr.push_back(br[s]);
      }
      src_off += (long)tp2 * ntau * p.nc;
      e->pairs.push_back(p);
    }
  }
  e->F = (int)kf.size();
  // chunks of whole pairs with about equal fine k3 counts (pipelined steps)
  {
    const int np = (int)e->pairs.size();
    const int C = std::max(1, std::min(e->chunks_forced, np));
    e->chunk_pair.assign(1, 0);
    e->chunk_f.assign(1, 0);
    for (int c = 1; c < C; ++c) {
      const long goal = (long)e->F * c / C;
      int q = e->chunk_pair.back();
      while (q < np && e->pairs[q].f_off < goal) ++q;
      if (q > e->chunk_pair.back() && q < np) {
        e->chunk_pair.push_back(q);
        e->chunk_f.push_back(e->pairs[q].f_off);
      }
    }
    e->chunk_pair.push_back(np);
    e->chunk_f.push_back(e->F);
    e->n_chunk = (int)e->chunk_pair.size() - 1;
    while ((int)e->ev_chunk.size() < e->n_chunk) {
      hipEvent_t v;
      HIPCHK(hipEventCreateWithFlags(&v, hipEventDisableTiming), err);
      e->ev_chunk.push_back(v);
    }
  }
  // processing order: chunk by chunk, fine k3 sorted by value inside a chunk
  // (ties by index)
  std::vector<int> forder(e->F), chunk_of(e->F, 0);
  for (int c = 0; c < e->n_chunk; ++c)
    for (int f = e->chunk_f[c]; f < e->chunk_f[c + 1]; ++f) chunk_of[f] = c;
  for (int i = 0; i < e->F; ++i) forder[i] = i;
  std::stable_sort(forder.begin(), forder.end(), [&](int a, int b) {
    return chunk_of[a] != chunk_of[b] ? chunk_of[a] < chunk_of[b] : kf[a] < kf[b];
  });
  e->chunk_pos = e->chunk_f;  // chunk c holds positions chunk_f[c] .. chunk_f[c+1]-1
  e->src_total = src_off;
  e->ht0mt = t0mt;
  e->hforder = forder;
  e->hkf = kf;
  e->tt2_size = st->tt2_size;
  e->ntau = ntau;

  // projection tables: built once for the whole k range of the run
  e->cubic_J = !st->linear_J;
  if (e->J_owner != st->J_id || e->tables_cubic != e->cubic_J) {
    double kmax_all = 0;
    for (int i1 = 0; i1 < st->k_size; ++i1)
      for (int i2 = 0; i2 <= i1; ++i2) kmax_all = std::max(kmax_all, st->k_max_k1k2[i1][i2]);
    if (build_tables(e, st, kmax_all, err) == _FAILURE_) return _FAILURE_;
  }

  // estimated cost of each sorted position: the number of time steps before
  // x = k (tau0 - tau) falls below every row's first live index, plus a fixed
  // per-workgroup part
  e->pos_cost.assign(e->F, 0.0);
  for (int sp = 0; sp < e->F; ++sp) {
    const double k = kf[forder[sp]];
    int lo = 0, hi = ntau;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int)(k * t0mt[mid] / st->xx_step) >= e->minx_all) lo = mid + 1; else hi = mid;
    }
    e->pos_cost[sp] = lo + 16.0;
  }
  // sweep kernels: live time steps of each (fine k3, wave class), and the cost
  // of a position as the sum over its waves
  {
    const int W = e->n_wave;
    auto live = [&](double k, int mx) {
      int lo = 0, hi = ntau;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)(k * t0mt[mid] / st->xx_step) >= mx) lo = mid + 1; else hi = mid;
      }
      return lo;
    };
    std::vector<int> tend((size_t)e->F * W);
    for (int f = 0; f < e->F; ++f)
      for (int w = 0; w < W; ++w) tend[(size_t)f * W + w] = live(kf[f], e->wave_minx[w]);
    e->pos_cost_sweep.assign(e->F, 0.0);
    for (int sp = 0; sp < e->F; ++sp) {
      double c = 0;
      for (int w = 0; w < W; ++w) c += tend[(size_t)forder[sp] * W + w] + 4.0;
      e->pos_cost_sweep[sp] = c;
    }
    if (upload_vec(e, e->dtau_end_fw, tend, err)) return _FAILURE_;
  }

  // one job: the whole batch, output [tt][F] at the fine index
  std::vector<int> posjob(e->F, 0);
  std::vector<los_job> jobs(1);
  jobs[0].obase = 0;
  jobs[0].ostride = e->F;
  jobs[0].n_lm = e->n_lm;
  jobs[0].ttT = st->index_tt2_T;
  jobs[0].ttE = st->index_tt2_E;
  jobs[0].ttB = st->index_tt2_B;
  for (int i = 0; i < 5; ++i) {
    // tiles never straddle two chunks; chunk_tile[i] holds each chunk's first tile
    e->htiles[i] = make_tiles(std::vector<int>(e->chunk_pos.begin() + 1, e->chunk_pos.end()),
                              i < 4 ? 1 << i : 64);
    e->chunk_tile[i].assign(e->n_chunk + 1, (int)e->htiles[i].size());
    for (int c = 0, t = 0; c < e->n_chunk; ++c) {
      while (t < (int)e->htiles[i].size() && e->htiles[i][t].pos0 < e->chunk_pos[c]) ++t;
      e->chunk_tile[i][c] = t;
    }
    if (upload_vec(e, e->dtiles[i], e->htiles[i], err)) return _FAILURE_;
  }
  if (upload_vec(e, e->dposjob, posjob, err) || upload_vec(e, e->docol, forder, err) ||
      upload_vec(e, e->djobs, jobs, err))
    return _FAILURE_;

  if (upload_vec(e, e->dt0mt, t0mt, err) || upload_vec(e, e->ddtau, dtau, err) ||
      upload_vec(e, e->dpairs, e->pairs, err) || upload_vec(e, e->dkc, kc, err) ||
      upload_vec(e, e->dkf, kf, err) || upload_vec(e, e->dspl, spl, err) ||
      upload_vec(e, e->dfpair, fpair, err) ||
      upload_vec(e, e->dfsrc, fsrc, err) || upload_vec(e, e->dfbr, fbr, err) ||
      upload_vec(e, e->dforder, forder, err))
    return _FAILURE_;

  // sources as SONG holds them: per pair, tp2 contiguous [tau][kc] blocks; the
  // run re-lays them out to [kc][tau] (dsrc) and splines them (ddd)
  const size_t sbytes = std::max<long>(1, e->src_total) * sizeof(double);
  HIPCHK(hipMalloc((void **)&e->draw, sbytes), err);
  HIPCHK(hipMalloc((void **)&e->dsrc, sbytes), err);
  HIPCHK(hipMalloc((void **)&e->ddd, sbytes), err);
  for (const los_pair &p : e->pairs)
    for (int tp = 0; tp < tp2; ++tp) {
      const size_t n = (size_t)ntau * p.nc;
      if (upload(e, e->draw + p.src_off + (long)tp * n, st->sources[tp][p.k1][p.k2],
                 n * sizeof(double), err))
        return _FAILURE_;
    }
  HIPCHK(hipMalloc((void **)&e->dout, std::max<size_t>(1, (size_t)e->tt2_size * e->F) * sizeof(double)),
         err);

  // kernel parameters
  los_params &P = e->P;
  memset(&P, 0, sizeof P);
  P.pairs = e->dpairs;
  P.n_pairs = (int)e->pairs.size();
  P.kc = e->dkc;
  P.kf = e->dkf;
  P.fpair = e->dfpair;
  P.fsrc = e->dfsrc;
  P.fbr = e->dfbr;
  P.forder = e->dforder;
  P.n_pos = e->F;
  P.pos_job = e->dposjob;
  P.ocol = e->docol;
  P.jobs = e->djobs;
  P.n_lm_max = e->n_lm;
  P.tiles = e->dtiles[0];
  split_xcd(e->htiles[0], e->pos_cost, P.xcd_pos);
  P.Jp = e->dJp;
  P.xminp = e->dxminp;
  P.lm_of_lp = e->dlm_of_lp;
  P.n_lp = e->n_lp;
  P.n_wave = e->n_wave;
  P.m_size = e->m_size;
  for (int im = 0; im < LOS_MAX_MI; ++im) P.m_of_mi[im] = im < st->m_size ? st->m[im] : -1;
  P.tau_end_fw = e->dtau_end_fw;
  {
    std::vector<int> rank(e->F);
    for (int sp = 0; sp < e->F; ++sp) rank[forder[sp]] = sp;
    if (upload_vec(e, e->drank, rank, err)) return _FAILURE_;
  }
  P.korder = e->dforder;
  P.rank_of_f = e->drank;
  P.k_npad = (e->F + 63) / 64 * 64;
  P.kixa = e->dkixa;
  P.ksrc = e->dksrc;
  P.queue = e->dqueue;
  P.src = e->dsrc;
  P.dd = e->ddd;
  P.F = e->F;
  P.tp2_size = tp2;
  P.cubic_k3 = st->cubic_k3;
  P.t0mt = e->dt0mt;
  P.dtau = e->ddtau;
  P.ntau = ntau;
  P.J = e->dJ;
  P.xmin = e->dxmin;
  P.min_xmin = e->dminx;
  P.xx = e->dxx;
  P.xx_step = st->xx_step;
  P.X = e->X;
  P.L_size = e->L_size;
  P.n_lm = e->n_lm;
  P.l_of = e->dl_of;
  P.m_of = e->dm_of;
  P.mi_of = e->dmi_of;
  P.has_T = st->has_T;
  P.has_E = st->has_E;
  P.has_B = st->has_B;
  P.LmaxT = LmaxT;
  P.LmaxP = LmaxP;
  P.jTT = st->jTT;
  P.jEE = st->jEE;
  P.jEB = st->jEB;
  for (int L = 0; L < LOS_MAX_L; ++L)
    for (int im = 0; im < LOS_MAX_MI; ++im) {
      P.srcT[L][im] = P.srcE[L][im] = P.srcB[L][im] = -1;
      if (L > st->largest_l || im >= st->m_size || st->m[im] > L) continue;
      if (im > st->index_m_max[L]) continue;
      const int off = st->src_lm[L][im];
      if (st->has_T && L <= LmaxT) P.srcT[L][im] = st->index_tp2_T + off;
      if (st->has_E && L <= LmaxP) P.srcE[L][im] = st->index_tp2_E + off;
      if (st->has_B && L <= LmaxP) P.srcB[L][im] = st->index_tp2_B + off;
    }
  if ((P.has_T && P.jTT < 0) || ((P.has_E || P.has_B) && P.jEE < 0) ||
      (P.has_B && P.jEB < 0 && st->m_max > 0)) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_load: projection functions missing for a field");
    return _FAILURE_;
  }
  P.out = e->dout;
  HIPCHK(los_stream_set_attrs(tp2), err);
  e->loaded = true;
  return _SUCCESS_;
}

// Re-layout and k3 spline of the loaded sources (enqueued, not waited for).
// Whether a sweep shape takes its tables straight from the raw sources.
static bool use_fused(const los_engine *e, const los_shape *sh)
{
  return sh != nullptr && sh->kind == 1 && e->fused_prep &&
         los_sweep_fused_fits(e->P.tp2_size, e->nc_max);
}

// Source re-layout and k3 spline for the kernels that read them (every shape
// but a sweep shape with fused tables; sh == nullptr: always).
int enqueue_prep(los_engine *e, const los_shape *sh, char *err)
{
  if (use_fused(e, sh)) return _SUCCESS_;
  const int tp2 = e->P.tp2_size;
  const int np = (int)e->pairs.size();
  HIPCHK(los_launch_relayout_spline(e->dpairs, np, e->nc_max, e->dkc, e->dspl, e->draw, e->dsrc,
                                    e->ddd, e->ntau, tp2, e->P.cubic_k3, e->stream),
         err);
  return _SUCCESS_;
}

// Whether the sweep kernels can address the batch's J table (32-bit offsets).
bool sweep_fits(const los_engine *e, int form, const los_shape &sh)
{
  const bool fp32 = form == LOS_FORM_FP32C, pairs = !fp32 || sh.ch == 1;
  const double bytes = (double)e->nJp * (fp32 ? sizeof(float) : sizeof(double)) * (pairs ? 2 : 1);
  return bytes <= (double)LOS_BUF_MAX;
}

// Candidate kernel shapes of a form: the plain FP64 form has its
// one-workgroup-per-k3 kernel (kind 0), the sweep (kind 1) and the k-lane
// kernel (kind 2), all bit-identical to the reference; the fast forms every
// stream, sweep and k-lane shape that fits.
static int form_shapes(const los_engine *e, int form, los_shape *out, int cap)
{
  los_shape all[48];
  int n = 0;
  if (form == LOS_FORM_FP64) {
    all[n++] = los_shape{0, 1, 0, 0};
    for (int nk = 1; nk <= 8; nk *= 2)
      for (int p = 0; p < 2; ++p) all[n++] = los_shape{1, nk, 0, p};
    for (int b = 1; b <= 4; b *= 2) all[n++] = los_shape{2, 64, b, 0};
  } else {
    n = los_shapes(form == LOS_FORM_FP32C, e->loaded ? e->P.tp2_size : 1, all, 48);
  }
  int k = 0;
  for (int i = 0; i < n && k < cap; ++i) {
    // cubic interpolation of J: the plain FP64 sweep kernels only
    if (e->cubic_J && (form != LOS_FORM_FP64 || all[i].kind != 1)) continue;
    if (all[i].kind != 1 || sweep_fits(e, form, all[i])) out[k++] = all[i];
  }
  return k;
}

// Shapes the autotuner times: all but the persistent sweep, which measured
// slower than the dispatched sweep on every workload tried (the hardware's
// dynamic assignment of finished waves balances better than a fixed stride);
// it stays available through los_engine_set_shape.
static bool tuned_shape(const los_shape &s) { return !(s.kind == 1 && s.persistent); }

// The LOS kernel of the current form with shape sh, on work list P (tiles and
// cost: host copies of its tiles and per-position costs). A sweep or k-lane
// shape's per-(k3, tau) tables must be made already.
int enqueue_los_params(los_engine *e, const los_shape *sh, los_params *P, hipStream_t s,
                       const std::vector<los_tile> &tiles, const std::vector<double> &cost,
                       char *err)
{
  const int fp32 = e->form == LOS_FORM_FP32C;
  if (sh->kind == 2) {
    const int gs = sh->ch == 4 ? 2 : sh->ch == 2 ? 1 : 0;
    P->Jk = e->form == LOS_FORM_FP32C ? (const void *)e->dJk32 : (const void *)e->dJk;
    P->kgrp = e->dkgrp[gs];
    P->k_ngrp = e->kngrp[gs];
    split_xcd(tiles, cost, P->xcd_tile);
    if (e->form == LOS_FORM_FP64)
      HIPCHK(los_launch_klane_exact(P, sh->ch, s), err);
    else
      HIPCHK(los_launch_klane(fp32, P, sh->ch, s), err);
    return _SUCCESS_;
  }
  if (e->form == LOS_FORM_FP64 && sh->kind == 1) {
    P->Jp = e->dJp2[0];
    P->Jd = e->dDp2;
    P->cubic_J = e->cubic_J ? 1 : 0;
    P->Jp_bytes = (unsigned)(e->nJp * 2 * sizeof(double));
    split_xcd(tiles, cost, P->xcd_tile);
    HIPCHK(los_launch_sweep_exact(P, sh->nk, sh->persistent, e->n_cu, s), err);
    return _SUCCESS_;
  }
  if (e->form == LOS_FORM_FP64) {
    split_xcd(tiles, cost, P->xcd_pos);
    HIPCHK(los_launch_integrate_exact(P, 0, s), err);
    return _SUCCESS_;
  }
  split_xcd(tiles, cost, P->xcd_tile);
  if (sh->kind == 1) {
    const bool pairs = !fp32 || sh->ch == 1;
    P->Jp = !fp32 ? e->dJp2[0] : pairs ? e->dJp2[1] : (const void *)e->dJp32;
    P->Jp_bytes = (unsigned)(e->nJp * (fp32 ? sizeof(float) : sizeof(double)) * (pairs ? 2 : 1));
    HIPCHK(los_launch_sweep(fp32, P, sh->nk, sh->persistent, pairs, e->n_cu, s), err);
  } else {
    HIPCHK(los_launch_stream(fp32, P, sh, e->n_cu, s), err);
  }
  return _SUCCESS_;
}

// Per-(k3, tau) tables of a sweep or k-lane shape for the loaded batch
// (enqueued on the engine stream).
static int enqueue_tables(los_engine *e, const los_shape *sh, char *err, int f_begin = 0,
                          int f_end = -1, hipStream_t s = nullptr, int p_begin = 0,
                          int p_end = -1)
{
  const int fp32 = e->form == LOS_FORM_FP32C;
  if (f_end < 0) f_end = e->F;
  if (p_end < 0) p_end = (int)e->pairs.size();
  if (!s) s = e->stream;
  if (sh->kind == 1) {
    if (e->sweep_form != e->form) {
      snprintf(err, _ERRORMSGSIZE_, "enqueue_tables: sweep tables not allocated for this form");
      return _FAILURE_;
    }
    if (use_fused(e, sh)) {
      if (e->form == LOS_FORM_FP64)
        HIPCHK(los_launch_sweep_prep_fused_exact(&e->P, e->draw, e->dspl, e->nc_max, p_begin,
                                                 p_end, s),
               err);
      else
        HIPCHK(los_launch_sweep_prep_fused(fp32, &e->P, e->draw, e->dspl, e->nc_max, p_begin,
                                           p_end, s),
               err);
    } else if (e->form == LOS_FORM_FP64)
      HIPCHK(los_launch_sweep_prep_exact(&e->P, f_begin, f_end, s), err);
    else
      HIPCHK(los_launch_sweep_prep(fp32, &e->P, f_begin, f_end, s), err);
  } else if (sh->kind == 2) {
    if (e->klane_form != e->form) {
      snprintf(err, _ERRORMSGSIZE_, "enqueue_tables: k-lane tables not allocated for this form");
      return _FAILURE_;
    }
    if (e->form == LOS_FORM_FP64)
      HIPCHK(los_launch_klane_prep_exact(&e->P, e->stream), err);
    else
      HIPCHK(los_launch_klane_prep(fp32, &e->P, e->stream), err);
  }
  return _SUCCESS_;
}

// The LOS stage of the loaded batch with shape sh: its tables (if any), then
// the kernel. `mid`, if given, is recorded between the two so that the kernel
// can be timed alone.
static int enqueue_los(los_engine *e, const los_shape *sh, char *err, hipEvent_t mid = nullptr)
{
  if (enqueue_tables(e, sh, err)) return _FAILURE_;
  if (mid) HIPCHK(hipEventRecord(mid, e->stream), err);
  los_params P = e->P;
  const int slot = nk_slot(sh->kind == 2 ? 64 : sh->nk);
  P.tiles = e->dtiles[slot];
  const std::vector<double> &cost =
      sh->kind == 1 ? e->pos_cost_sweep : e->pos_cost;
  return enqueue_los_params(e, sh, &P, e->stream, e->htiles[slot], cost, err);
}

// Whether a step of shape sh runs pipelined (chunk c+1's preparation on
// stream2 while chunk c's LOS kernel runs): sweep shapes of batches with more
// than one chunk.
static bool pipelined(const los_engine *e, const los_shape *sh)
{
  return sh->kind == 1 && e->n_chunk > 1;
}

// One pipelined step: per chunk, the source re-layout and spline of its pairs
// and its sweep tables on stream2, then (after an event) its LOS kernel on the
// engine stream. stream2 first waits for everything already on the engine
// stream, so a step never overwrites tables that the previous step still reads.
static int enqueue_pipelined(los_engine *e, const los_shape *sh, char *err)
{
  const int tp2 = e->P.tp2_size;
  const int slot = nk_slot(sh->nk);
  HIPCHK(hipEventRecord(e->ev_fork, e->stream), err);
  HIPCHK(hipStreamWaitEvent(e->stream2, e->ev_fork, 0), err);
  for (int c = 0; c < e->n_chunk; ++c) {
    const int p0 = e->chunk_pair[c], p1 = e->chunk_pair[c + 1];
    if (!use_fused(e, sh))
      HIPCHK(los_launch_relayout_spline(e->dpairs + p0, p1 - p0, e->nc_max, e->dkc, e->dspl,
                                        e->draw, e->dsrc, e->ddd, e->ntau, tp2, e->P.cubic_k3,
                                        e->stream2),
             err);
    if (enqueue_tables(e, sh, err, e->chunk_f[c], e->chunk_f[c + 1], e->stream2, p0, p1))
      return _FAILURE_;
    HIPCHK(hipEventRecord(e->ev_chunk[c], e->stream2), err);
    HIPCHK(hipStreamWaitEvent(e->stream, e->ev_chunk[c], 0), err);
    los_params P = e->P;
    const int t0 = e->chunk_tile[slot][c], t1 = e->chunk_tile[slot][c + 1];
    P.tiles = e->dtiles[slot] + t0;
    const std::vector<los_tile> tiles(e->htiles[slot].begin() + t0, e->htiles[slot].begin() + t1);
    if (enqueue_los_params(e, sh, &P, e->stream, tiles, e->pos_cost_sweep, err)) return _FAILURE_;
  }
  return _SUCCESS_;
}

// Per-(k3, tau) tables of the current shape for every loaded fine k3 (waits).
int make_sweep_tables(los_engine *e, char *err)
{
  const los_shape *sh = &e->shape[e->form];
  if (enqueue_tables(e, sh, err)) return _FAILURE_;
  HIPCHK(hipStreamSynchronize(e->stream), err);
  return _SUCCESS_;
}

// Device tables a sweep shape needs (allocated outside any capture).
int ensure_sweep(los_engine *e, const los_shape *sh, char *err)
{
  if (sh == nullptr || sh->kind != 1) return _SUCCESS_;
  const int fp32 = e->form == LOS_FORM_FP32C;
  if (!sweep_fits(e, e->form, *sh)) {
    snprintf(err, _ERRORMSGSIZE_,
             "sweep shape: the J table of this batch (%zu entries) exceeds the 32-bit offsets "
             "of the sweep kernels; use another kernel organisation",
             e->nJp);
    return _FAILURE_;
  }
  if (fp32 && e->dJp32 == nullptr) {
    HIPCHK(hipMalloc((void **)&e->dJp32, std::max<size_t>(1, e->nJp) * sizeof(float)), err);
    HIPCHK(los_launch_round_j32(e->dJp, e->dJp32, (long)e->nJp, e->stream), err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
  }
  if (!fp32 && e->dJp2[0] == nullptr) {
    // FP64: (J[x], J[x+1]) pairs, zero below each row's index_xmin_J, one load per row
    const int NLP = e->m_size * e->n_lp;
    HIPCHK(hipMalloc(&e->dJp2[0], std::max<size_t>(1, e->nJp) * 2 * sizeof(double)), err);
    HIPCHK(los_launch_sweep_pairs(0, e->dJp, e->dJp2[0], e->dxminp, (long)e->J_size * e->L_size,
                                  e->X, NLP, e->stream),
           err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
  }
  if (!fp32 && e->cubic_J && e->dDp2 == nullptr) {
    // second derivatives as (dd[x], dd[x+1]) pairs with the same zeroing
    const int NLP = e->m_size * e->n_lp;
    HIPCHK(hipMalloc(&e->dDp2, std::max<size_t>(1, e->nJp) * 2 * sizeof(double)), err);
    HIPCHK(los_launch_sweep_pairs(0, e->dDp, e->dDp2, e->dxminp, (long)e->J_size * e->L_size,
                                  e->X, NLP, e->stream),
           err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
  }
  if (fp32 && sh->ch == 1 && e->dJp2[1] == nullptr) {
    // FP32 pairs of the rounded table, same zeroing
    const int NLP = e->m_size * e->n_lp;
    HIPCHK(hipMalloc(&e->dJp2[1], std::max<size_t>(1, e->nJp) * 2 * sizeof(float)), err);
    HIPCHK(los_launch_sweep_pairs(1, e->dJp32, e->dJp2[1], e->dxminp, (long)e->J_size * e->L_size,
                                  e->X, NLP, e->stream),
           err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
  }
  if (e->sweep_form == e->form) return _SUCCESS_;
  if (e->dixa) (void)hipFree(e->dixa);
  if (e->dsf) (void)hipFree(e->dsf);
  e->dixa = e->dsf = nullptr;
  e->sweep_form = -1;
  const int NL = std::max(e->P.LmaxT, e->P.LmaxP) + 1;
  const size_t n = (size_t)e->F * e->ntau;
  const size_t rb = fp32 ? 4 : 8;
  const size_t ixa_b = std::max<size_t>(1, n) * (fp32 ? 8 : 16);
  const size_t sf_b = std::max<size_t>(1, n * e->m_size * NL) * 3 * rb;
  HIPCHK(hipMalloc(&e->dixa, ixa_b), err);
  HIPCHK(hipMalloc(&e->dsf, sf_b), err);
  e->P.ixa = e->dixa;
  e->P.sf = e->dsf;
  e->sweep_form = e->form;
  e->sweep_bytes = ixa_b + sf_b;
  return _SUCCESS_;
}

// Device tables a k-lane shape needs (allocated outside any capture).
int ensure_klane(los_engine *e, const los_shape *sh, char *err)
{
  if (sh == nullptr || sh->kind != 2) return _SUCCESS_;
  const int fp32 = e->form == LOS_FORM_FP32C;
  if (fp32 && e->dJk32 == nullptr) {
    HIPCHK(hipMalloc((void **)&e->dJk32, std::max<size_t>(1, e->nJk) * sizeof(float)), err);
    HIPCHK(los_launch_round_j32(e->dJk, e->dJk32, (long)e->nJk, e->stream), err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
  }
  if (e->klane_form == e->form) return _SUCCESS_;
  if (e->dkixa) (void)hipFree(e->dkixa);
  if (e->dksrc) (void)hipFree(e->dksrc);
  e->dkixa = e->dksrc = nullptr;
  e->klane_form = -1;
  const size_t n = (size_t)e->P.k_npad * e->ntau;
  const size_t ixa_b = std::max<size_t>(1, n) * (fp32 ? 8 : 16);
  const size_t src_b = std::max<size_t>(1, n * e->P.tp2_size) * (fp32 ? 4 : 8);
  HIPCHK(hipMalloc(&e->dkixa, ixa_b), err);
  HIPCHK(hipMalloc(&e->dksrc, src_b), err);
  e->P.kixa = e->dkixa;
  e->P.ksrc = e->dksrc;
  e->klane_form = e->form;
  e->klane_bytes = ixa_b + src_b;
  return _SUCCESS_;
}

static int ensure_tables(los_engine *e, const los_shape *sh, char *err)
{
  return ensure_sweep(e, sh, err) || ensure_klane(e, sh, err) ? _FAILURE_ : _SUCCESS_;
}

int ensure_j32(los_engine *e, char *err)
{
  if (e->form == LOS_FORM_FP32C && e->dJ32 == nullptr) {
    // FP32 copy of the resident J table, made once per table upload
    HIPCHK(hipMalloc((void **)&e->dJ32, std::max<size_t>(1, e->nJ) * sizeof(float)), err);
    HIPCHK(los_launch_round_j32(e->dJ, e->dJ32, (long)e->nJ, e->stream), err);
    HIPCHK(hipStreamSynchronize(e->stream), err);
  }
  e->P.J32 = e->dJ32;
  return _SUCCESS_;
}

// Autotuner cache: best shape per (form, (l,m) count, source types, sources L,
// time steps, fine-k3 count rounded to a power of two).
typedef std::tuple<int, int, int, int, int, int, int> tune_key;
static std::map<tune_key, los_shape> g_tuned;
static std::mutex g_tuned_mu;

static tune_key key_of(const los_engine *e)
{
  int fb = 0;
  while ((1 << fb) < e->F) ++fb;
  return tune_key(e->form, e->n_lm, e->P.tp2_size, std::max(e->P.LmaxT, e->P.LmaxP), e->ntau, fb,
                  e->cubic_J ? 1 : 0);
}

extern "C" int los_engine_autotune(los_engine *e, int reps, char *err)
{
  if (!e->loaded) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_autotune: nothing loaded");
    return _FAILURE_;
  }
  if (e->F == 0) return _SUCCESS_;
  HIPCHK(hipSetDevice(e->device), err);
  const int form = e->form;
  const tune_key key = key_of(e);
  {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end() && reps >= 0) {
      e->shape[form] = it->second;
      e->shape_set[form] = true;
      return _SUCCESS_;
    }
  }
  if (reps < 1) reps = 1;
  if (ensure_j32(e, err)) return _FAILURE_;
  los_shape all[48], cand[48];
  const int na = form_shapes(e, form, all, 48);
  int nc = 0;
  for (int i = 0; i < na; ++i)
    if (tuned_shape(all[i])) cand[nc++] = all[i];
  if (nc == 0) {
    if (e->cubic_J)
      snprintf(err, _ERRORMSGSIZE_, "los_engine_autotune: cubic bessels_interpolation is "
               "implemented in the plain FP64 form (LOS_FORM_FP64) only");
    else
      snprintf(err, _ERRORMSGSIZE_,
               "los_engine_autotune: no tile shape fits %d source types in LDS", e->P.tp2_size);
    return _FAILURE_;
  }
  float best = 1e30f;
  int ibest = 0;
  for (int i = 0; i < nc; ++i) {
    // each shape timed on whole steps: its own preparation and tables, then its kernel
    if (ensure_tables(e, &cand[i], err)) return _FAILURE_;
    if (enqueue_prep(e, &cand[i], err) || enqueue_los(e, &cand[i], err))  // warm-up
      return _FAILURE_;
    HIPCHK(hipEventRecord(e->ev[1], e->stream), err);
    for (int r = 0; r < reps; ++r)
      if (enqueue_prep(e, &cand[i], err) || enqueue_los(e, &cand[i], err)) return _FAILURE_;
    HIPCHK(hipEventRecord(e->ev[2], e->stream), err);
    HIPCHK(hipEventSynchronize(e->ev[2]), err);
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[1], e->ev[2]), err);
    if (ms < best) {
      best = ms;
      ibest = i;
    }
  }
  e->shape[form] = cand[ibest];
  e->shape_set[form] = true;
  drop_graph(e);
  std::lock_guard<std::mutex> g(g_tuned_mu);
  g_tuned[key] = cand[ibest];
  return _SUCCESS_;
}

extern "C" int los_engine_set_shape(los_engine *e, int form, int kind, int nk, int ch,
                                    int persistent, char *err)
{
  if (form < LOS_FORM_FP64 || form > LOS_FORM_FP32C) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_set_shape: unknown form %d", form);
    return _FAILURE_;
  }
  los_shape cand[48];
  const int nc = form_shapes(e, form, cand, 48);
  for (int i = 0; i < nc; ++i)
    if (cand[i].kind == kind &&
        (kind == 2   ? cand[i].ch == ch
         : kind == 1 ? (cand[i].nk == nk && cand[i].persistent == (persistent != 0) &&
                        cand[i].ch == ch)
                     : (form == LOS_FORM_FP64 ||
                        (cand[i].nk == nk && cand[i].ch == ch &&
                         cand[i].persistent == (persistent != 0))))) {
      e->shape[form] = cand[i];
      e->shape_set[form] = true;
      drop_graph(e);
      return _SUCCESS_;
    }
  snprintf(err, _ERRORMSGSIZE_, "los_engine_set_shape: form %d has no shape kind=%d nk=%d ch=%d "
           "persistent=%d for this batch", form, kind, nk, ch, persistent);
  return _FAILURE_;
}

extern "C" void los_engine_get_shape(los_engine *e, int form, int *kind, int *nk, int *ch,
                                     int *persistent)
{
  if (form < LOS_FORM_FP64 || form > LOS_FORM_FP32C) form = LOS_FORM_FP64;
  const los_shape &s = e->shape[form];
  if (kind) *kind = s.kind;
  if (nk) *nk = s.nk;
  if (ch) *ch = s.ch;
  if (persistent) *persistent = s.persistent;
}

// Shape of the current form, tuned on first use, with its tables allocated.
int current_shape(los_engine *e, const los_shape **sh, char *err)
{
  if (!e->shape_set[e->form])
    if (los_engine_autotune(e, 2, err)) return _FAILURE_;
  *sh = &e->shape[e->form];
  if (e->cubic_J && (e->form != LOS_FORM_FP64 || (*sh)->kind != 1)) {
    snprintf(err, _ERRORMSGSIZE_, "cubic bessels_interpolation is implemented in the plain FP64 "
             "form's sweep kernels only (form %d, kernel kind %d)", e->form, (*sh)->kind);
    return _FAILURE_;
  }
  return ensure_tables(e, *sh, err);
}

extern "C" int los_engine_capture(los_engine *e, char *err)
{
  if (!e->loaded) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_capture: nothing loaded");
    return _FAILURE_;
  }
  if (e->F == 0) return _SUCCESS_;  // empty batch: runs do nothing
  HIPCHK(hipSetDevice(e->device), err);
  const los_shape *sh;
  if (ensure_j32(e, err) || current_shape(e, &sh, err)) return _FAILURE_;
  drop_graph(e);
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal), err);
  int rc;
  if (pipelined(e, sh) && e->pipeline) {
    rc = enqueue_pipelined(e, sh, err);
  } else {
    rc = enqueue_prep(e, sh, err);
    if (rc == _SUCCESS_) rc = enqueue_los(e, sh, err);
  }
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(e->stream, &g);
  if (rc != _SUCCESS_) {
    if (g) (void)hipGraphDestroy(g);
    return _FAILURE_;
  }
  HIPCHK(ec, err);
  e->graph = g;
  HIPCHK(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0), err);
  e->graph_form = e->form;
  return _SUCCESS_;
}

extern "C" int los_engine_run(los_engine *e, float *ms, char *err)
{
  if (!e->loaded) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_run: nothing loaded");
    return _FAILURE_;
  }
  if (e->F == 0) {
    e->ms_spline = e->ms_los = 0;
    if (ms) *ms = 0;
    return _SUCCESS_;
  }
  HIPCHK(hipSetDevice(e->device), err);
  const los_shape *sh;
  if (ensure_j32(e, err) || current_shape(e, &sh, err)) return _FAILURE_;
  if (e->graph_exec && e->graph_form == e->form) {
    // captured run: one graph launch; only the total time is known
    HIPCHK(hipEventRecord(e->ev[0], e->stream), err);
    HIPCHK(hipGraphLaunch(e->graph_exec, e->stream), err);
    HIPCHK(hipEventRecord(e->ev[2], e->stream), err);
    HIPCHK(hipEventSynchronize(e->ev[2]), err);
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e->ev[0], e->ev[2]), err);
    e->ms_spline = -1;
    e->ms_los = -1;
    if (ms) *ms = t;
    return _SUCCESS_;
  }
  if (pipelined(e, sh) && e->pipeline) {
    // pipelined step: only the total time is known
    HIPCHK(hipEventRecord(e->ev[0], e->stream), err);
    if (enqueue_pipelined(e, sh, err)) return _FAILURE_;
    HIPCHK(hipEventRecord(e->ev[2], e->stream), err);
    HIPCHK(hipEventSynchronize(e->ev[2]), err);
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e->ev[0], e->ev[2]), err);
    e->ms_spline = -1;
    e->ms_los = -1;
    if (ms) *ms = t;
    return _SUCCESS_;
  }
  // ev0 .. ev1: source re-layout, k3 spline and the shape's per-(k3, tau)
  // tables; ev1 .. ev2: the LOS kernel alone
  HIPCHK(hipEventRecord(e->ev[0], e->stream), err);
  if (enqueue_prep(e, sh, err)) return _FAILURE_;
  if (enqueue_los(e, sh, err, e->ev[1])) return _FAILURE_;
  HIPCHK(hipEventRecord(e->ev[2], e->stream), err);
  HIPCHK(hipEventSynchronize(e->ev[2]), err);
  HIPCHK(hipEventElapsedTime(&e->ms_spline, e->ev[0], e->ev[1]), err);
  HIPCHK(hipEventElapsedTime(&e->ms_los, e->ev[1], e->ev[2]), err);
  if (ms) *ms = e->ms_spline + e->ms_los;
  return _SUCCESS_;
}

extern "C" int los_engine_set_fused_prep(los_engine *e, int on, char *err)
{
  (void)err;
  if ((on != 0) != e->fused_prep) drop_graph(e);
  e->fused_prep = on != 0;
  return _SUCCESS_;
}

extern "C" int los_engine_set_pipeline(los_engine *e, int on, char *err)
{
  if (on < 0) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_set_pipeline: %d chunks", on);
    return _FAILURE_;
  }
  if ((on >= 2) != e->pipeline) drop_graph(e);
  e->pipeline = on >= 2;
  e->chunks_forced = on >= 2 ? on : 0;
  return _SUCCESS_;
}

extern "C" int los_engine_last_kernel_ms(los_engine *e, float *ms_los, float *ms_spline)
{
  if (ms_los) *ms_los = e->ms_los;
  if (ms_spline) *ms_spline = e->ms_spline;
  return _SUCCESS_;
}

// Results into SONG's transfer arrays, with the reference's NaN check
// (source/transfer2.c:2332-2337, on in its default DEBUG build): a NaN
// transfer fails the store with the reference's message.
extern "C" int los_engine_store_stage(los_engine *e, const los_stage *st, char *err)
{
  if (!e->loaded) {
    snprintf(err, _ERRORMSGSIZE_, "los_engine_store: nothing loaded");
    return _FAILURE_;
  }
  if (e->F == 0) return _SUCCESS_;
  HIPCHK(hipSetDevice(e->device), err);
  const size_t n = (size_t)e->tt2_size * e->F;
  std::vector<double> h(n);
  HIPCHK(hipMemcpyAsync(h.data(), e->dout, n * sizeof(double), hipMemcpyDeviceToHost, e->stream), err);
  HIPCHK(hipStreamSynchronize(e->stream), err);
  for (int tt = 0; tt < e->tt2_size; ++tt)
    for (const los_pair &p : e->pairs) {
      if (!st->has_transfers[p.k1]) {
        snprintf(err, _ERRORMSGSIZE_, "los_engine_store: k1 level %d not allocated", p.k1);
        return _FAILURE_;
      }
      const double *v = h.data() + (size_t)tt * e->F + p.f_off;
      for (int k = 0; k < p.nf; ++k)
        if (std::isnan(v[k])) {
          snprintf(err, _ERRORMSGSIZE_,
                   "found nan in second-order transfer function (index_tt=%d, index_k1=%d, "
                   "index_k2=%d, index_k3=%d)", tt, p.k1, p.k2, k);
          return _FAILURE_;
        }
      memcpy(st->transfer[tt][p.k1][p.k2], v, p.nf * sizeof(double));
      *st->count_memorised += p.nf;
    }
  return _SUCCESS_;
}

extern "C" long los_engine_n_elements(los_engine *e) { return (long)e->tt2_size * e->F; }

// The last run's transfers where they are, in device memory: [tt][F], column f
// = the loaded pairs' fine k3 in load order (each pair's grid in turn).
extern "C" int los_engine_device_output(los_engine *e, const double **out, long *rows,
                                        long *cols)
{
  if (out) *out = e->loaded ? e->dout : nullptr;
  if (rows) *rows = e->loaded ? e->tt2_size : 0;
  if (cols) *cols = e->loaded ? e->F : 0;
  return e->loaded && e->dout ? _SUCCESS_ : _FAILURE_;
}

extern "C" hipError_t los_launch_copy_probe(const double *src, double *dst, long n, hipStream_t s);

// One launch of a plain copy kernel of `bytes` on the engine's stream (device
// to device), timed with HIP events: the reference of known traffic for HBM
// counter readings of the LOS kernels.
extern "C" int los_engine_copy_probe(los_engine *e, double bytes, float *ms, char *err)
{
  HIPCHK(hipSetDevice(e->device), err);
  const long n = (long)(bytes / 8);
  double *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc((void **)&a, std::max<long>(1, n) * 8), err);
  HIPCHK(hipMalloc((void **)&b, std::max<long>(1, n) * 8), err);
  HIPCHK(los_launch_copy_probe(b, a, n, e->stream), err);  // first touch
  HIPCHK(hipEventRecord(e->ev[0], e->stream), err);
  HIPCHK(los_launch_copy_probe(a, b, n, e->stream), err);
  HIPCHK(hipEventRecord(e->ev[1], e->stream), err);
  HIPCHK(hipEventSynchronize(e->ev[1]), err);
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e->ev[0], e->ev[1]), err);
  if (ms) *ms = t;
  (void)hipFree(a);
  (void)hipFree(b);
  return _SUCCESS_;
}

// Algorithmic FP64/FP32 operation count of one run. For every output element
// and each integral it sums (source/transfer2.c:2140-2341), over the time
// steps, 5 per live L (3 for the J interpolation, 2 for the multiply-add with
// the source) plus 2 when any L is live (the trapezoid multiply-add); `live`
// applies the reference's skip rule exactly, `dense` counts every L row.
extern "C" int los_engine_flops(los_engine *e, double *live, double *dense)
{
  if (!e->loaded) return _FAILURE_;
  const los_params &P = e->P;
  const int ntau = e->ntau, n_lm = e->n_lm, L_size = e->L_size;
  const double xs = P.xx_step;
  // the integrals of each (l,m): lists of J row indices (jt * L_size + L)
  std::vector<std::vector<std::vector<int>>> calls(n_lm);
  for (int lm = 0; lm < n_lm; ++lm) {
    const int m = e->hm_of[lm];
    auto rows = [&](int jt, int Lmax) {
      std::vector<int> r;
      for (int L = m; L <= Lmax && L < L_size; ++L) r.push_back(jt * L_size + L);
      return r;
    };
    if (P.has_T && P.jTT >= 0) calls[lm].push_back(rows(P.jTT, P.LmaxT));
    if (P.has_E && P.jEE >= 0) {
      calls[lm].push_back(rows(P.jEE, P.LmaxP));
      if (m != 0 && P.jEB >= 0 && P.has_B) calls[lm].push_back(rows(P.jEB, P.LmaxP));
    }
    if (P.has_B && P.jEE >= 0) {
      calls[lm].push_back(rows(P.jEE, P.LmaxP));
      if (P.jEB >= 0) calls[lm].push_back(rows(P.jEB, P.LmaxP));
    }
  }
  double d = 0;
  for (int lm = 0; lm < n_lm; ++lm)
    for (const auto &c : calls[lm]) d += (double)ntau * (5.0 * c.size() + 2.0);
  d *= e->F;
  const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<double> part(nth, 0.0);
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nth; ++w)
    th.emplace_back([&, w]() {
      std::vector<int> ix(ntau);
      double acc = 0;
      for (int f = (int)w; f < e->F; f += (int)nth) {
        const double k = e->hkf[f];
        for (int t = 0; t < ntau; ++t) ix[t] = (int)(k * e->ht0mt[t] / xs);
        auto steps = [&](int xm) {  // leading time steps with index_x >= xm
          int lo = 0, hi = ntau;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ix[mid] >= xm) lo = mid + 1; else hi = mid;
          }
          return lo;
        };
        for (int lm = 0; lm < n_lm; ++lm)
          for (const auto &c : calls[lm]) {
            int most = 0;
            for (int row : c) {
              const int n = steps(e->hxmin[(size_t)row * n_lm + lm]);
              acc += 5.0 * n;
              most = std::max(most, n);
            }
            acc += 2.0 * most;
          }
      }
      part[w] = acc;
    });
  for (auto &x : th) x.join();
  double l = 0;
  for (double v : part) l += v;
  if (live) *live = l;
  if (dense) *dense = d;
  return _SUCCESS_;
}

extern "C" double los_engine_alg_bytes(los_engine *e)
{
  // compulsory bytes: each coarse source value once (the spline second
  // derivatives are derived from them and not counted), the live part of the J
  // table once at the precision the form reads it in, each output once
  const double src = 8.0 * e->src_total;
  const double jb = e->J_window_bytes * (e->form == LOS_FORM_FP32C ? 0.5 : 1.0);
  return src + jb + 8.0 * (double)e->tt2_size * e->F;
}
