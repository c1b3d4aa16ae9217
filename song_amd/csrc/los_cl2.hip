// los_cl2.hip — the second-order angular power spectrum integral over the
// transfer functions on the GPU (gfx950), FP64: the (k1, k2, k3) sum of SONG's
// spectra2_cls (source/spectra2.c:241-590, the k3 loop :466-540), for every l
// and m of the run:
//
//   C_l^m = sum_{k1} sum_{k2<=k1} sum_{k3} 2 k1 k2 k3 dk1 dk2 dk3 [scale^2]
//           * 2/pi/2/pi/2/pi/2/pi * P1 P2 * T1_lm(k1,k2,k3) T2_lm(k1,k2,k3)
//
// with dk1, dk2 the half trapezoid steps of the k grid (dk2 = 0 for k1 = 0,
// the k1 = k2 end of the k2 range taking the left half step only), dk3 the
// half trapezoid steps of the pair's k3 grid with the triangle-edge
// corrections (:486-500), P = 2 pi^2 / k^3 times the primordial spectrum, and
// scale^2 = (1 - cos^2)^m, cos = (k3^2 + k1^2 - k2^2) / (2 k3 k1), when the
// sources were rescaled (ppt2->rescale_cmb_sources).
//
// Organisation: the per-pair quantities (k1, k2, their half steps and P
// values, computed on the host with the reference's expressions) and each
// pair's two triangle-edge points (the first point with k3 >= k1 - k2 before
// the last, and the first point after the first whose successor lies beyond
// k1 + k2, the point past the end counting as beyond) are resolved once, so
// every k3 term is independent: los_cl2_partial runs CL2_CHUNKS workgroups of
// 256 threads per (l, m) row over the flat k3 index and reduces in LDS in a
// fixed order; los_cl2_finish adds a row's chunk sums in order. Results are
// deterministic; the sum order differs from the reference's sequential one,
// so they agree to rounding (relative 1e-12 of the sum of |terms|).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "los_common.h"

#define CL2_THREADS 256

#define CCHK(call)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      snprintf(errmsg, _ERRORMSGSIZE_, "%s: %s failed: %s", __func__, #call,                 \
               hipGetErrorString(e_));                                                       \
      return _FAILURE_;                                                                      \
    }                                                                                        \
  } while (0)

#define CL2_CHUNKS 32

struct los_cl2 {
  int device;
  int k_size, n_pairs, n_l, m_size;
  long n_k3;
  int rescale;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  double *k3, *t1, *t2, *out, *partial, *pq;  // pq: per pair k1 k2 s1 s2 sp1 sp2
  long *pair_off;
  int *pair_of, *edge, *m, *used;             // edge: per pair first, last
};

// One (k1, k2, k3) term without the transfers; the reference's factor order
// (:502-511). sqrt of a negative 1 - cos^2 (rounding at a triangle edge) is
// NaN there, and pow(NaN, 0) = 1, as in the reference.
__device__ __forceinline__ double cl2_factor(double k1, double k2, double k3, double s1,
                                             double s2, double s3, double sp1, double sp2,
                                             int m, int rescale)
{
  double sc2 = 1.;
  if (rescale && m != 0) {
    const double c = (k3 * k3 + k1 * k1 - k2 * k2) / 2. / k3 / k1;
    const double w = 1. - c * c;
    if (w < 0) {
      sc2 = __builtin_nan("");
    } else {
      sc2 = w;
      for (int i = 1; i < m; ++i) sc2 *= w;
    }
  }
  return 2. * k1 * k2 * k3 * s1 * s2 * s3 * sc2 * 2. / M_PI / 2. / M_PI / 2. / M_PI / 2. /
         M_PI * sp1 * sp2;
}

__global__ void __launch_bounds__(CL2_THREADS)
los_cl2_partial(const long *__restrict__ pair_off, const int *__restrict__ pair_of,
                const int *__restrict__ edge, const double *__restrict__ pq,
                const double *__restrict__ k3g, long n_k3, const double *__restrict__ t1,
                const double *__restrict__ t2, const int *__restrict__ mv, int m_size,
                const int *__restrict__ used, int rescale, double *__restrict__ partial)
{
  __shared__ double part[CL2_THREADS];
  const int chunk = blockIdx.x % CL2_CHUNKS;
  const long row = blockIdx.x / CL2_CHUNKS;
  const int m = mv[row % m_size];
  double acc = 0.;
  if (used[row]) {
    const double *T1 = t1 + row * n_k3;
    const double *T2 = t2 + row * n_k3;
    for (long i = (long)chunk * CL2_THREADS + threadIdx.x; i < n_k3;
         i += (long)CL2_CHUNKS * CL2_THREADS) {
      const int p = pair_of[i];
      const long b = pair_off[p];
      const int n = (int)(pair_off[p + 1] - b);
      const int j = (int)(i - b);
      const double *g = k3g + b;
      const double *q = pq + 6 * (long)p;
      const double k1 = q[0], k2 = q[1];
      const double k3 = g[j];
      double s3;
      // half trapezoid step and triangle-edge corrections (:486-500)
      if (j == edge[2 * p + 1]) s3 = (k3 - g[j - 1]) / 2. + (k1 + k2) - k3;
      else if (j == edge[2 * p]) s3 = (g[j + 1] - k3) / 2. + k3 - (k1 - k2);
      else if (k3 < k1 - k2 || k3 > k1 + k2) s3 = 0.;
      else if (j == 0) s3 = (g[1] - g[0]) / 2.;
      else if (j == n - 1) s3 = (g[n - 1] - g[n - 2]) / 2.;
      else s3 = (g[j + 1] - g[j - 1]) / 2.;
      if (s3 == 0) continue;
      const double f = cl2_factor(k1, k2, k3, q[2], q[3], s3, q[4], q[5], m, rescale);
      acc += f * T1[i] * T2[i];
    }
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = CL2_THREADS / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[row * CL2_CHUNKS + chunk] = part[0];
}

__global__ void __launch_bounds__(64)
los_cl2_finish(const double *__restrict__ partial, long rows, double *__restrict__ out)
{
  const long row = (long)blockIdx.x * 64 + threadIdx.x;
  if (row >= rows) return;
  double v = 0.;
  for (int c = 0; c < CL2_CHUNKS; ++c) v += partial[row * CL2_CHUNKS + c];
  out[row] = v;
}

template <class T>
static hipError_t cl2_upload(T **dst, const T *src, long n)
{
  hipError_t e = hipMalloc((void **)dst, (size_t)(n > 0 ? n : 1) * sizeof(T));
  if (e != hipSuccess || n <= 0) return e;
  return hipMemcpy(*dst, src, (size_t)n * sizeof(T), hipMemcpyHostToDevice);
}

extern "C" {

void los_cl2_destroy(los_cl2 *h)
{
  if (!h) return;
  hipSetDevice(h->device);
  void *bufs[] = {h->k3, h->t1, h->t2, h->out, h->partial, h->pq,
                  h->pair_off, h->pair_of, h->edge, h->m, h->used};
  for (void *b : bufs)
    if (b) hipFree(b);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->stream) hipStreamDestroy(h->stream);
  free(h);
}

int los_cl2_create(const los_cl2_desc *d, int device, los_cl2 **out, char *errmsg)
{
  *out = NULL;
  los_test(d->k_size < 2, errmsg, "the k grid needs two points");
  los_test(d->n_l < 0 || d->m_size < 1, errmsg, "bad l or m counts");
  for (int i = 0; i < d->m_size; ++i)
    los_test(d->m[i] < 0, errmsg, "m[%d] = %d: azimuthal numbers are non-negative", i, d->m[i]);
  const int n_pairs = d->k_size * (d->k_size + 1) / 2;
  los_test(d->pair_off[0] != 0, errmsg, "pair_off[0] must be 0");
  // source/spectra2.c:463-465
  for (int p = 0; p < n_pairs; ++p)
    los_test(d->pair_off[p + 1] - d->pair_off[p] < 2, errmsg,
             "pair %d: integration grid has less than two elements, cannot use trapezoidal "
             "integration", p);
  for (int i = 0; i < d->k_size; ++i)
    los_test(!(d->k[i] > 0), errmsg, "k[%d] = %g must be positive", i, d->k[i]);
  const long n_k3 = d->pair_off[n_pairs];
  for (long i = 0; i < n_k3; ++i)
    los_test(!(d->k3[i] > 0), errmsg, "k3[%ld] = %g must be positive", i, d->k3[i]);
  // per pair: k1, k2, half steps and P values (:375-380, :388, :398-404, :420),
  // and the two triangle-edge points of its k3 grid (:490-500)
  const double PI = M_PI;
  const double *k = d->k;
  const int ks = d->k_size;
  std::vector<double> pq(6 * (size_t)n_pairs);
  std::vector<int> edge(2 * (size_t)n_pairs), pair_of(n_k3 > 0 ? n_k3 : 1);
  for (int i1 = 0, p = 0; i1 < ks; ++i1)
    for (int i2 = 0; i2 <= i1; ++i2, ++p) {
      double s1, s2;
      if (i1 == 0) s1 = (k[1] - k[0]) / 2.;
      else if (i1 == ks - 1) s1 = (k[ks - 1] - k[ks - 2]) / 2.;
      else s1 = (k[i1 + 1] - k[i1 - 1]) / 2.;
      if (i1 == 0) s2 = 0.;
      else if (i2 == 0) s2 = (k[1] - k[0]) / 2.;
      else if (i2 == i1) s2 = (k[i1] - k[i1 - 1]) / 2.;
      else s2 = (k[i2 + 1] - k[i2 - 1]) / 2.;
      const double k1 = k[i1], k2 = k[i2];
      double *q = &pq[6 * (size_t)p];
      q[0] = k1;
      q[1] = k2;
      q[2] = s1;
      q[3] = s2;
      q[4] = 2 * PI * PI / (k1 * k1 * k1) * d->pk[i1];
      q[5] = 2 * PI * PI / (k2 * k2 * k2) * d->pk[i2];
      const long b = d->pair_off[p];
      const int n = (int)(d->pair_off[p + 1] - b);
      const double *g = d->k3 + b;
      int first = -1, last = -1;
      for (int j = 0; j < n; ++j) {
        if (first < 0 && g[j] >= k1 - k2 && j < n - 1) first = j;
        const bool next_out = (j == n - 1) || (g[j + 1] > k1 + k2);
        if (last < 0 && next_out && j > 0) last = j;
        pair_of[b + j] = p;
      }
      edge[2 * (size_t)p] = first;
      edge[2 * (size_t)p + 1] = last;
    }
  const long rows = (long)d->n_l * d->m_size;
  los_cl2 *h = (los_cl2 *)calloc(1, sizeof *h);
  los_test(h == NULL, errmsg, "out of host memory");
  h->device = device;
  h->k_size = d->k_size;
  h->n_pairs = n_pairs;
  h->n_l = d->n_l;
  h->m_size = d->m_size;
  h->n_k3 = n_k3;
  h->rescale = d->rescale;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreate(&h->stream);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = cl2_upload(&h->pq, pq.data(), 6L * n_pairs);
  if (e == hipSuccess) e = cl2_upload(&h->edge, edge.data(), 2L * n_pairs);
  if (e == hipSuccess) e = cl2_upload(&h->pair_of, pair_of.data(), n_k3);
  if (e == hipSuccess) e = cl2_upload(&h->pair_off, d->pair_off, (long)n_pairs + 1);
  if (e == hipSuccess) e = cl2_upload(&h->k3, d->k3, n_k3);
  if (e == hipSuccess) e = cl2_upload(&h->t1, d->transfer1, rows * n_k3);
  if (e == hipSuccess) e = cl2_upload(&h->t2, d->transfer2, rows * n_k3);
  if (e == hipSuccess) e = cl2_upload(&h->m, d->m, (long)d->m_size);
  if (e == hipSuccess) e = cl2_upload(&h->used, d->used, rows);
  if (e == hipSuccess) e = hipMalloc((void **)&h->out, (size_t)(rows > 0 ? rows : 1) * 8);
  if (e == hipSuccess)
    e = hipMalloc((void **)&h->partial, (size_t)(rows > 0 ? rows : 1) * CL2_CHUNKS * 8);
  if (e != hipSuccess) {
    snprintf(errmsg, _ERRORMSGSIZE_, "los_cl2_create: %s", hipGetErrorString(e));
    los_cl2_destroy(h);
    return _FAILURE_;
  }
  *out = h;
  return _SUCCESS_;
}

int los_cl2_run(los_cl2 *h, int reps, float *ms, char *errmsg)
{
  CCHK(hipSetDevice(h->device));
  if (ms) *ms = 0;
  const long rows = (long)h->n_l * h->m_size;
  if (rows == 0) return _SUCCESS_;
  if (reps < 1) reps = 1;
  CCHK(hipEventRecord(h->ev0, h->stream));
  los_test(rows * CL2_CHUNKS > 0x7fffffffL, errmsg, "los_cl2_run: too many rows");
  for (int i = 0; i < reps; ++i) {
    hipLaunchKernelGGL(los_cl2_partial, dim3((unsigned)(rows * CL2_CHUNKS)), dim3(CL2_THREADS),
                       0, h->stream, h->pair_off, h->pair_of, h->edge, h->pq, h->k3, h->n_k3,
                       h->t1, h->t2, h->m, h->m_size, h->used, h->rescale, h->partial);
    hipLaunchKernelGGL(los_cl2_finish, dim3((unsigned)((rows + 63) / 64)), dim3(64), 0,
                       h->stream, h->partial, rows, h->out);
  }
  CCHK(hipGetLastError());
  CCHK(hipEventRecord(h->ev1, h->stream));
  CCHK(hipEventSynchronize(h->ev1));
  float t = 0;
  CCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
  if (ms) *ms = t / reps;
  return _SUCCESS_;
}

int los_cl2_result(los_cl2 *h, double *cl_m, long n, char *errmsg)
{
  const long rows = (long)h->n_l * h->m_size;
  los_test(n != rows, errmsg, "los_cl2_result: %ld values asked, %ld held", n, rows);
  if (rows == 0) return _SUCCESS_;
  CCHK(hipSetDevice(h->device));
  CCHK(hipMemcpy(cl_m, h->out, (size_t)rows * sizeof(double), hipMemcpyDeviceToHost));
  return _SUCCESS_;
}


// ---------------------------------------------------------------------------
// The same sum for callers that cannot allocate device memory: every input and
// the output are the caller's device arrays; the per-pair table, the pair of
// each k3 point and the partial sums live in `scratch`, a file-scope device
// array (one per GPU, held until the process ends). Each kernel writes every
// scratch value it later reads, so nothing needs clearing between calls;
// calls that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
#define CL2_NA_KSIZE_MAX 128                                   // k grid points
#define CL2_NA_PAIRS_MAX (CL2_NA_KSIZE_MAX * (CL2_NA_KSIZE_MAX + 1) / 2)
#define CL2_NA_K3_MAX (1L << 21)                               // k3 points of all pairs
#define CL2_NA_ROWS_MAX 8192L                                  // (l, m) rows
#define CL2_NA_OFF_EDGE (((6L * CL2_NA_PAIRS_MAX * 8) + 255) / 256 * 256)
#define CL2_NA_OFF_PAIR (CL2_NA_OFF_EDGE + ((2L * CL2_NA_PAIRS_MAX * 4) + 255) / 256 * 256)
#define CL2_NA_OFF_PART (CL2_NA_OFF_PAIR + ((CL2_NA_K3_MAX * 4) + 255) / 256 * 256)
#define CL2_NA_BYTES (CL2_NA_OFF_PART + CL2_NA_ROWS_MAX * CL2_CHUNKS * 8)

static __device__ __align__(256) unsigned char scratch[CL2_NA_BYTES];

// Per pair (one thread each): k1, k2, half steps and P values, the two
// triangle-edge points and the pair of each of its k3 points, with the host
// expressions of los_cl2_create (this unit is compiled without contraction).
__global__ void __launch_bounds__(64)
los_cl2_pairs(int k_size, const double *__restrict__ k, const double *__restrict__ pk,
              const long *__restrict__ pair_off, const double *__restrict__ k3g,
              double *__restrict__ pq, int *__restrict__ edge, int *__restrict__ pair_of)
{
  const int p = blockIdx.x * 64 + threadIdx.x;
  if (p >= k_size * (k_size + 1) / 2) return;
  int i1 = (int)((sqrt(8.0 * p + 1.0) - 1.0) / 2.0);
  while (i1 * (i1 + 1) / 2 > p) --i1;
  while ((i1 + 1) * (i1 + 2) / 2 <= p) ++i1;
  const int i2 = p - i1 * (i1 + 1) / 2;
  const int ks = k_size;
  double s1, s2;
  if (i1 == 0) s1 = (k[1] - k[0]) / 2.;
  else if (i1 == ks - 1) s1 = (k[ks - 1] - k[ks - 2]) / 2.;
  else s1 = (k[i1 + 1] - k[i1 - 1]) / 2.;
  if (i1 == 0) s2 = 0.;
  else if (i2 == 0) s2 = (k[1] - k[0]) / 2.;
  else if (i2 == i1) s2 = (k[i1] - k[i1 - 1]) / 2.;
  else s2 = (k[i2 + 1] - k[i2 - 1]) / 2.;
  const double k1 = k[i1], k2 = k[i2];
  const double PI = M_PI;
  double *q = pq + 6 * (long)p;
  q[0] = k1;
  q[1] = k2;
  q[2] = s1;
  q[3] = s2;
  q[4] = 2 * PI * PI / (k1 * k1 * k1) * pk[i1];
  q[5] = 2 * PI * PI / (k2 * k2 * k2) * pk[i2];
  const long b = pair_off[p];
  const int n = (int)(pair_off[p + 1] - b);
  const double *g = k3g + b;
  int first = -1, last = -1;
  for (int j = 0; j < n; ++j) {
    if (first < 0 && g[j] >= k1 - k2 && j < n - 1) first = j;
    const bool next_out = (j == n - 1) || (g[j + 1] > k1 + k2);
    if (last < 0 && next_out && j > 0) last = j;
    pair_of[b + j] = p;
  }
  edge[2 * p] = first;
  edge[2 * p + 1] = last;
}

namespace {
std::mutex na_mutex;
void *na_base[64];  // device address of `scratch` on each device, once looked up
}  // namespace

extern "C" int los_cl2_sum_noalloc(const los_cl2_desc *d, long n_k3, double *cl_m, void *stream,
                                   char *errmsg)
{
  los_test(d->k_size < 2 || d->k_size > CL2_NA_KSIZE_MAX, errmsg,
           "k_size = %d: 2 to %d points", d->k_size, CL2_NA_KSIZE_MAX);
  los_test(n_k3 < 0 || n_k3 > CL2_NA_K3_MAX, errmsg, "n_k3 = %ld: at most %ld k3 points", n_k3,
           (long)CL2_NA_K3_MAX);
  const long rows = (long)d->n_l * d->m_size;
  los_test(d->n_l < 0 || d->m_size < 1 || rows > CL2_NA_ROWS_MAX, errmsg,
           "%d l x %d m: at most %ld (l, m) rows", d->n_l, d->m_size, (long)CL2_NA_ROWS_MAX);
  int dev = 0;
  CCHK(hipGetDevice(&dev));
  los_test(dev < 0 || dev >= 64, errmsg, "device %d", dev);
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(na_mutex);
    if (na_base[dev] == nullptr) {
      void *ptr = nullptr;
      CCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      na_base[dev] = ptr;
    }
    base = na_base[dev];
  }
  if (rows == 0) return _SUCCESS_;
  unsigned char *b = (unsigned char *)base;
  double *pq = (double *)b;
  int *edge = (int *)(b + CL2_NA_OFF_EDGE);
  int *pair_of = (int *)(b + CL2_NA_OFF_PAIR);
  double *partial = (double *)(b + CL2_NA_OFF_PART);
  hipStream_t s = (hipStream_t)stream;
  const int n_pairs = d->k_size * (d->k_size + 1) / 2;
  hipLaunchKernelGGL(los_cl2_pairs, dim3((n_pairs + 63) / 64), dim3(64), 0, s, d->k_size, d->k,
                     d->pk, d->pair_off, d->k3, pq, edge, pair_of);
  hipLaunchKernelGGL(los_cl2_partial, dim3((unsigned)(rows * CL2_CHUNKS)), dim3(CL2_THREADS), 0,
                     s, d->pair_off, pair_of, edge, pq, d->k3, n_k3, d->transfer1, d->transfer2,
                     d->m, d->m_size, d->used, d->rescale, partial);
  hipLaunchKernelGGL(los_cl2_finish, dim3((unsigned)((rows + 63) / 64)), dim3(64), 0, s, partial,
                     rows, cl_m);
  CCHK(hipGetLastError());
  return _SUCCESS_;
}

}  // extern "C"
