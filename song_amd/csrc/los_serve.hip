// los_serve.hip — the LOS stage as a service: a persistent request loop over
// resident tables.
//
// At creation every (k1,k2) pair of the run is loaded into one engine: the
// sources are re-laid out and splined in k3 once, the projection tables and,
// for the sweep kernels, the per-(k3, tau) tables are made once, and all of it
// stays in HBM. A job asks for the transfers of a list of pairs up to some
// l_max. Jobs wait in a FIFO queue; a scheduler thread admits them while the
// device bytes of the admitted jobs (outputs and work lists) stay within a
// fixed HBM budget, and launches every job admitted together as one kernel
// launch on one of several streams ("slots"). Jobs that arrive while others
// run are admitted as soon as a slot and budget free up, so their work
// overlaps on the device (continuous batching). Tiles never mix jobs and the
// kernel shape is fixed for the server's lifetime, so each job's result is
// bit-identical whether it runs alone or in a mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include "los_engine_impl.h"

namespace {

double now_s()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct job_rec {
  long id = 0;
  int n_lm = 0, n_cols = 0, fields = 0;
  std::vector<int> pos;  // fine indices, sorted by k3
  std::vector<int> col;  // output column of each position
  size_t n_values = 0;   // fields * n_lm * n_cols
  size_t bytes = 0;      // device footprint while admitted
  double t_submit = 0, t_admit = 0, t_done = 0;
  int state = 0;         // 0 queued, 1 running, 2 done, 3 failed
  std::string err;
  std::vector<double> result;
};

struct slot_rec {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  double *dout = nullptr;
  size_t out_cap = 0;
  char *ddesc = nullptr;  // forder | pos_job | ocol | jobs | tiles
  size_t desc_cap = 0;
  int *dqueue = nullptr;
  char *hdesc = nullptr;  // pinned staging of the descriptors
  size_t hdesc_cap = 0;
  double *hout = nullptr;  // pinned results
  size_t hout_cap = 0;
  std::vector<long> jobs;
  std::vector<size_t> obase;
  bool busy = false;
};

}  // namespace

struct los_server {
  los_engine *e = nullptr;
  los_shape shape{};
  int has_T = 0, has_E = 0, has_B = 0;
  std::map<std::pair<int, int>, int> pair_of;  // (index_k1, index_k2) -> loaded pair
  std::vector<int> lm_l;                       // l of each (l,m) offset
  std::vector<double> fcost;                   // estimated cost per fine index
  double budget = 0, in_flight = 0, peak_in_flight = 0;
  size_t resident = 0;
  int max_batch = 64;
  std::vector<slot_rec> slots;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<long> queue;
  std::map<long, job_rec> jobs;
  long next_id = 0;
  bool stop = false;
  bool hold = false;  // admissions paused (jobs keep queueing)
  std::thread th;
  long n_batches = 0, n_done = 0;
  int max_busy = 0;
  std::string fatal;
};

#define SRVCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      snprintf(err, _ERRORMSGSIZE_, "%s: %s failed: %s", __func__, #expr,                 \
               hipGetErrorString(e_));                                                    \
      return _FAILURE_;                                                                   \
    }                                                                                     \
  } while (0)

template <class T>
static int grow_device(T *&p, size_t &cap, size_t need, char *err)
{
  if (need <= cap) return _SUCCESS_;
  if (p) (void)hipFree((void *)p);
  p = nullptr;
  cap = 0;
  SRVCHK(hipMalloc((void **)&p, need));
  cap = need;
  return _SUCCESS_;
}

template <class T>
static int grow_pinned(T *&p, size_t &cap, size_t need, char *err)
{
  if (need <= cap) return _SUCCESS_;
  if (p) (void)hipHostFree((void *)p);
  p = nullptr;
  cap = 0;
  SRVCHK(hipHostMalloc((void **)&p, need, hipHostMallocDefault));
  cap = need;
  return _SUCCESS_;
}

static size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Launch the admitted jobs `ids` on slot s (mu held).
static int launch_batch(los_server *S, slot_rec &sl, const std::vector<long> &ids, char *err)
{
  los_engine *e = S->e;
  std::vector<int> forder, posjob, ocol, job_end;
  std::vector<los_job> jd;
  std::vector<double> cost;
  sl.obase.clear();
  size_t obase = 0;
  int n_lm_max = 0;
  for (size_t q = 0; q < ids.size(); ++q) {
    const job_rec &j = S->jobs[ids[q]];
    los_job d;
    d.obase = (long)obase;
    d.ostride = j.n_cols;
    d.n_lm = j.n_lm;
    int row = 0;
    d.ttT = d.ttE = d.ttB = 0;
    if (S->has_T) { d.ttT = row; row += j.n_lm; }
    if (S->has_E) { d.ttE = row; row += j.n_lm; }
    if (S->has_B) { d.ttB = row; row += j.n_lm; }
    jd.push_back(d);
    sl.obase.push_back(obase);
    obase += j.n_values;
    n_lm_max = std::max(n_lm_max, j.n_lm);
    for (size_t i = 0; i < j.pos.size(); ++i) {
      forder.push_back(j.pos[i]);
      posjob.push_back((int)q);
      ocol.push_back(j.col[i]);
      cost.push_back(S->fcost[j.pos[i]]);
    }
    job_end.push_back((int)forder.size());
  }
  const int nk = S->shape.kind == 2 ? 64 : S->shape.nk;
  std::vector<los_tile> tiles = make_tiles(job_end, nk);

  // descriptors, packed into one upload
  const size_t b0 = align256(forder.size() * sizeof(int));
  const size_t b1 = b0 + align256(posjob.size() * sizeof(int));
  const size_t b2 = b1 + align256(ocol.size() * sizeof(int));
  const size_t b3 = b2 + align256(jd.size() * sizeof(los_job));
  const size_t b4 = b3 + align256(tiles.size() * sizeof(los_tile));
  if (grow_device(sl.ddesc, sl.desc_cap, b4, err) || grow_pinned(sl.hdesc, sl.hdesc_cap, b4, err) ||
      grow_device(sl.dout, sl.out_cap, std::max<size_t>(1, obase) * sizeof(double), err) ||
      grow_pinned(sl.hout, sl.hout_cap, std::max<size_t>(1, obase) * sizeof(double), err))
    return _FAILURE_;
  memcpy(sl.hdesc, forder.data(), forder.size() * sizeof(int));
  memcpy(sl.hdesc + b0, posjob.data(), posjob.size() * sizeof(int));
  memcpy(sl.hdesc + b1, ocol.data(), ocol.size() * sizeof(int));
  memcpy(sl.hdesc + b2, jd.data(), jd.size() * sizeof(los_job));
  memcpy(sl.hdesc + b3, tiles.data(), tiles.size() * sizeof(los_tile));
  SRVCHK(hipMemcpyAsync(sl.ddesc, sl.hdesc, b4, hipMemcpyHostToDevice, sl.stream));

  los_params P = e->P;
  P.forder = (const int *)sl.ddesc;
  P.n_pos = (int)forder.size();
  P.pos_job = (const int *)(sl.ddesc + b0);
  P.ocol = (const int *)(sl.ddesc + b1);
  P.jobs = (const los_job *)(sl.ddesc + b2);
  P.tiles = (const los_tile *)(sl.ddesc + b3);
  P.n_lm_max = n_lm_max;
  P.out = sl.dout;
  P.queue = sl.dqueue;
  if (enqueue_los_params(e, &S->shape, &P, sl.stream, tiles, cost, err)) return _FAILURE_;
  SRVCHK(hipMemcpyAsync(sl.hout, sl.dout, obase * sizeof(double), hipMemcpyDeviceToHost, sl.stream));
  SRVCHK(hipEventRecord(sl.done, sl.stream));
  sl.jobs = ids;
  sl.busy = true;
  S->n_batches++;
  return _SUCCESS_;
}

static void fail_jobs(los_server *S, const std::vector<long> &ids, const char *msg)
{
  for (long id : ids) {
    job_rec &j = S->jobs[id];
    j.state = 3;
    j.err = msg;
    j.t_done = now_s();
    S->in_flight -= (double)j.bytes;
  }
}

static void scheduler(los_server *S)
{
  char err[_ERRORMSGSIZE_];
  (void)hipSetDevice(S->e->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    bool progressed = false;
    int busy = 0;
    // completions
    for (slot_rec &sl : S->slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) {
        ++busy;
        continue;
      }
      const double t = now_s();
      for (size_t i = 0; i < sl.jobs.size(); ++i) {
        job_rec &j = S->jobs[sl.jobs[i]];
        if (q == hipSuccess) {
          j.result.assign(sl.hout + sl.obase[i], sl.hout + sl.obase[i] + j.n_values);
          j.state = 2;
        } else {
          j.state = 3;
          j.err = std::string("device error: ") + hipGetErrorString(q);
        }
        j.t_done = t;
        S->in_flight -= (double)j.bytes;
        S->n_done++;
      }
      sl.jobs.clear();
      sl.busy = false;
      progressed = true;
    }
    // admissions: FIFO, as many queued jobs per free slot as the budget allows
    for (slot_rec &sl : S->slots) {
      if (sl.busy || S->queue.empty() || S->hold) continue;
      std::vector<long> ids;
      while (!S->queue.empty() && (int)ids.size() < S->max_batch) {
        job_rec &j = S->jobs[S->queue.front()];
        if (S->in_flight + (double)j.bytes > S->budget) break;
        S->in_flight += (double)j.bytes;
        j.state = 1;
        j.t_admit = now_s();
        ids.push_back(j.id);
        S->queue.pop_front();
      }
      if (ids.empty()) break;
      S->peak_in_flight = std::max(S->peak_in_flight, S->in_flight);
      if (launch_batch(S, sl, ids, err)) {
        fail_jobs(S, ids, err);
        sl.busy = false;
      } else {
        ++busy;
      }
      progressed = true;
    }
    S->max_busy = std::max(S->max_busy, busy);
    if (progressed) S->cv.notify_all();
    if (S->stop && (S->queue.empty() || S->hold) && busy == 0) break;
    if (busy > 0)
      S->cv.wait_for(lk, std::chrono::microseconds(50));
    else if (!progressed)
      S->cv.wait(lk);
  }
}

// every (k1,k2) pair of the run
static int load_all(los_engine *e, const los_stage *st, char *err)
{
  std::vector<int> k1, k2;
  for (int i1 = 0; i1 < st->k_size; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2) {
      k1.push_back(i1);
      k2.push_back(i2);
    }
  return los_engine_load_stage(e, st, k1.data(), k2.data(), (int)k1.size(), err);
}

extern "C" int los_server_create_stage(int device, int form, double hbm_budget_bytes,
                                       int n_slots, const los_stage *st, los_server **out,
                                       char *err)
{
  if (n_slots < 1 || n_slots > 8) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_create: %d slots (1 to 8)", n_slots);
    return _FAILURE_;
  }
  if (!(hbm_budget_bytes > 0)) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_create: the HBM budget must be positive");
    return _FAILURE_;
  }
  los_server *S = new los_server();
  S->budget = hbm_budget_bytes;
  if (los_engine_create(device, &S->e, err) || los_engine_set_form(S->e, form, err) ||
      load_all(S->e, st, err)) {
    if (S->e) los_engine_destroy(S->e);
    delete S;
    return _FAILURE_;
  }
  los_engine *e = S->e;
  // resident preparation: source re-layout and k3 spline, kernel shape, and the
  // per-(k3, tau) tables of a sweep shape
  const los_shape *sh = nullptr;
  int rc = enqueue_prep(e, nullptr, err);
  if (rc == _SUCCESS_ && hipStreamSynchronize(e->stream) != hipSuccess) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_create: source preparation failed");
    rc = _FAILURE_;
  }
  if (rc == _SUCCESS_) rc = ensure_j32(e, err);
  if (rc == _SUCCESS_) rc = current_shape(e, &sh, err);
  if (rc == _SUCCESS_) rc = make_sweep_tables(e, err);  // tables of a sweep or k-lane shape
  if (rc != _SUCCESS_) {
    los_engine_destroy(e);
    delete S;
    return _FAILURE_;
  }
  if (sh) S->shape = *sh;
  // what the service no longer reads: the sources as SONG lays them out and the
  // engine's own output array
  dfree(e->draw);
  dfree(e->dout);
  e->P.out = nullptr;
  S->has_T = e->P.has_T;
  S->has_E = e->P.has_E;
  S->has_B = e->P.has_B;
  for (size_t q = 0; q < e->pairs.size(); ++q)
    S->pair_of[std::make_pair(e->pairs[q].k1, e->pairs[q].k2)] = (int)q;
  S->lm_l.assign(e->n_lm, 0);
  for (int il = 0; il < st->l_size; ++il)
    for (int im = 0; im <= st->index_m_max[st->l[il]]; ++im)
      S->lm_l[st->lm_array[il][im]] = st->l[il];
  const std::vector<double> &pc =
      S->shape.kind == 1 ? e->pos_cost_sweep : e->pos_cost;
  S->fcost.assign(e->F, 0.0);
  for (int sp = 0; sp < e->F; ++sp) S->fcost[e->hforder[sp]] = pc[sp];
  S->resident = (size_t)(16.0 * e->src_total) + e->nJ * 8 + e->nJp * 8 + e->nJk * 8 +
                (e->dJp32 ? e->nJp * 4 : 0) + (e->dJ32 ? e->nJ * 4 : 0) +
                (e->dJk32 ? e->nJk * 4 : 0) + (e->dJp2[0] ? e->nJp * 16 : 0) +
                (e->dJp2[1] ? e->nJp * 8 : 0) + e->sweep_bytes + e->klane_bytes;

  S->slots.resize(n_slots);
  for (slot_rec &sl : S->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
        hipMalloc((void **)&sl.dqueue, 64 * sizeof(int)) != hipSuccess) {
      snprintf(err, _ERRORMSGSIZE_, "los_server_create: stream set-up failed");
      los_server_destroy(S);
      return _FAILURE_;
    }
  }
  S->th = std::thread(scheduler, S);
  *out = S;
  return _SUCCESS_;
}

extern "C" void los_server_destroy(los_server *S)
{
  if (!S) return;
  if (S->th.joinable()) {
    {
      std::lock_guard<std::mutex> g(S->mu);
      // queued jobs are dropped; admitted ones finish
      for (long id : S->queue) {
        job_rec &j = S->jobs[id];
        j.state = 3;
        j.err = "server stopped";
      }
      S->queue.clear();
      S->stop = true;
    }
    S->cv.notify_all();
    S->th.join();
  }
  for (slot_rec &sl : S->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.dout) (void)hipFree(sl.dout);
    if (sl.ddesc) (void)hipFree(sl.ddesc);
    if (sl.dqueue) (void)hipFree(sl.dqueue);
    if (sl.hdesc) (void)hipHostFree(sl.hdesc);
    if (sl.hout) (void)hipHostFree(sl.hout);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  if (S->e) los_engine_destroy(S->e);
  delete S;
}

// Pause (hold = 1) or resume (0) admissions; queued jobs wait meanwhile.
extern "C" void los_server_hold(los_server *S, int hold)
{
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->hold = hold != 0;
  }
  S->cv.notify_all();
}

extern "C" long los_server_submit(los_server *S, const int *index_k1, const int *index_k2,
                                  int n_pairs, int l_max, char *err)
{
  if (n_pairs < 1) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_submit: a job needs at least one pair");
    return -1;
  }
  los_engine *e = S->e;
  job_rec j;
  j.fields = S->has_T + S->has_E + S->has_B;
  for (int lm = 0; lm < (int)S->lm_l.size(); ++lm)
    if (S->lm_l[lm] <= l_max) j.n_lm = lm + 1;
  if (j.n_lm == 0) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_submit: no multipole l <= l_max = %d", l_max);
    return -1;
  }
  for (int q = 0; q < n_pairs; ++q) {
    auto it = S->pair_of.find(std::make_pair(index_k1[q], index_k2[q]));
    if (it == S->pair_of.end()) {
      snprintf(err, _ERRORMSGSIZE_, "los_server_submit: pair (%d,%d) is not served",
               index_k1[q], index_k2[q]);
      return -1;
    }
    const los_pair &p = e->pairs[it->second];
    for (int k = 0; k < p.nf; ++k) {
      j.pos.push_back(p.f_off + k);
      j.col.push_back(j.n_cols + k);
    }
    j.n_cols += p.nf;
  }
  // work order of the job: its fine k3 by value
  std::vector<int> order(j.pos.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return e->hkf[j.pos[a]] < e->hkf[j.pos[b]]; });
  std::vector<int> pos(order.size()), col(order.size());
  for (size_t i = 0; i < order.size(); ++i) {
    pos[i] = j.pos[order[i]];
    col[i] = j.col[order[i]];
  }
  j.pos.swap(pos);
  j.col.swap(col);
  j.n_values = (size_t)j.fields * j.n_lm * j.n_cols;
  j.bytes = j.n_values * sizeof(double) + j.pos.size() * 3 * sizeof(int) + 4096;
  if ((double)j.bytes > S->budget) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_submit: the job needs %zu device bytes, more than "
             "the whole budget (%.0f)", j.bytes, S->budget);
    return -1;
  }
  std::lock_guard<std::mutex> g(S->mu);
  if (S->stop) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_submit: server stopped");
    return -1;
  }
  j.id = S->next_id++;
  j.t_submit = now_s();
  const long id = j.id;
  S->jobs[id] = std::move(j);
  S->queue.push_back(id);
  S->cv.notify_all();
  return id;
}

extern "C" int los_server_job_shape(los_server *S, long id, long *n_values, int *n_lm,
                                    int *n_cols)
{
  std::lock_guard<std::mutex> g(S->mu);
  auto it = S->jobs.find(id);
  if (it == S->jobs.end()) return _FAILURE_;
  if (n_values) *n_values = (long)it->second.n_values;
  if (n_lm) *n_lm = it->second.n_lm;
  if (n_cols) *n_cols = it->second.n_cols;
  return _SUCCESS_;
}

// 0: done (result copied, job released), 1: failed, 2: timed out (job kept).
extern "C" int los_server_wait(los_server *S, long id, double timeout_ms, double *out, long n,
                               double *times, char *err)
{
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->jobs.find(id);
  if (it == S->jobs.end()) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_wait: unknown job %ld", id);
    return 1;
  }
  auto ready = [&] { return it->second.state >= 2; };
  if (timeout_ms < 0)
    S->cv.wait(lk, ready);
  else if (!S->cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready))
    return 2;
  job_rec &j = it->second;
  if (times) {
    times[0] = j.t_submit;
    times[1] = j.t_admit;
    times[2] = j.t_done;
  }
  int rc = 0;
  if (j.state == 3) {
    snprintf(err, _ERRORMSGSIZE_, "job %ld failed: %s", id, j.err.c_str());
    rc = 1;
  } else if (n < (long)j.n_values) {
    snprintf(err, _ERRORMSGSIZE_, "los_server_wait: buffer of %ld values, job has %zu", n,
             j.n_values);
    return 1;  // job kept
  } else {
    memcpy(out, j.result.data(), j.n_values * sizeof(double));
  }
  S->jobs.erase(it);
  return rc;
}

// [jobs done, batches launched, most slots busy at once, peak admitted bytes,
//  resident bytes, budget]
extern "C" int los_server_stats(los_server *S, double *out, int n)
{
  std::lock_guard<std::mutex> g(S->mu);
  const double v[6] = {(double)S->n_done, (double)S->n_batches, (double)S->max_busy,
                       S->peak_in_flight, (double)S->resident, S->budget};
  for (int i = 0; i < n && i < 6; ++i) out[i] = v[i];
  return _SUCCESS_;
}
