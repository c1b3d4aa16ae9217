// los_k3int.hip — the k3 integral of the second-order transfer functions on
// the GPU (gfx950), FP64: the inner part of SONG's bispectrum stage,
// bispectra2_intrinsic_integrate_over_k3 (source/bispectra2.c:1182-1410), whose
// work is bessel2_convolution (source/bessel2.c:499-605) with linear Bessel
// interpolation, for every (l3, r, k1, k2):
//
//   I[q][r][p] = 2 * (coef * (0.5 * sum_k3 ((k3 * k3 * j_L(k3 r)) * T_qp(k3)) * dk3))
//
// with j_L(x) = a j[ix] + (1 - a) j[ix+1], ix = (int)((x - x_min)/step),
// a = (x_min + step (ix+1) - x)/step, terms with T = 0 or x < x_min skipped,
// dk3 the trapezoid measure of the pair's k3 grid (:1283-1287), coef the M3
// coefficient (the reference multiplies by it only for M3 != 0; with coef = 1
// the product is exact, so applying it always is the same result).
//
// Organisation: one wave per (row q, pair p, 64 consecutive r), one lane per r.
// The pair's k3, dk3 and T(k3) are wave-uniform (scalar loads); each lane
// gathers its two j nodes. Each lane sums its k3 terms in the reference's order,
// and the unit is compiled without contraction, so the plain FP64 result is the
// reference's bit for bit.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <vector>

#include "los_common.h"

#define KCHK(call)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      snprintf(errmsg, _ERRORMSGSIZE_, "%s: %s failed: %s", __func__, #call,                 \
               hipGetErrorString(e_));                                                       \
      return _FAILURE_;                                                                      \
    }                                                                                        \
  } while (0)

struct los_k3int {
  int device;
  int n_pairs, n_q, n_r, n_j;
  long n_k3;
  double xx_step, coef;
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  // device buffers
  long *pair_off;
  double *k3, *dk3, *transfer, *j, *x_min, *r, *out;
  long *j_off;
  int *index_L;
  long live_terms;  // terms with T != 0 and x >= x_min (host count)
  double alg_bytes;
};

__global__ void __launch_bounds__(64)
los_k3int_kernel(const long *__restrict__ pair_off, const double *__restrict__ k3,
                 const double *__restrict__ dk3, const double *__restrict__ transfer, long n_k3,
                 const int *__restrict__ index_L, const long *__restrict__ j_off,
                 const double *__restrict__ jtab, const double *__restrict__ x_min, double step,
                 const double *__restrict__ rr, int n_r, int n_pairs, int n_rc, double coef,
                 double *__restrict__ out)
{
  const int rc = blockIdx.x % n_rc;
  const int p = (blockIdx.x / n_rc) % n_pairs;
  const int q = blockIdx.x / (n_rc * n_pairs);
  const int ir = rc * 64 + threadIdx.x;
  if (ir >= n_r) return;
  const int iL = index_L[q];
  double integral = 0.0;
  if (iL >= 0) {
    const double r = rr[ir];
    const double xmin = x_min[iL];
    const double *J = jtab + j_off[iL];
    const double *T = transfer + (long)q * n_k3;
    const long b = pair_off[p], e = pair_off[p + 1];
    for (long i = b; i < e; ++i) {
      const double f = T[i];
      if (f == 0.) continue;
      const double k = k3[i];
      const double x = k * r;
      if (x < xmin) continue;
      const int ix = (int)((x - xmin) / step);
      const double a = (xmin + step * (ix + 1) - x) / step;
      const double jv = a * J[ix] + (1. - a) * J[ix + 1];
      const double integrand = k * k * jv * f;
      integral += integrand * dk3[i];
    }
    integral *= 0.5;
    integral *= coef;
    integral *= 2;
  }
  out[((long)q * n_r + ir) * n_pairs + p] = integral;
}

template <class T>
static hipError_t upload(T **dst, const T *src, long n)
{
  hipError_t e = hipMalloc((void **)dst, (size_t)(n > 0 ? n : 1) * sizeof(T));
  if (e != hipSuccess || n <= 0) return e;
  return hipMemcpy(*dst, src, (size_t)n * sizeof(T), hipMemcpyHostToDevice);
}

extern "C" {

void los_k3int_destroy(los_k3int *h)
{
  if (!h) return;
  hipSetDevice(h->device);
  void *bufs[] = {h->pair_off, h->k3, h->dk3, h->transfer, h->j, h->x_min,
                  h->r,        h->out, h->j_off, h->index_L};
  for (void *b : bufs)
    if (b) hipFree(b);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  if (h->stream) hipStreamDestroy(h->stream);
  free(h);
}

int los_k3int_create(const los_k3int_desc *d, int device, los_k3int **out, char *errmsg)
{
  *out = NULL;
  los_test(d->n_pairs < 0 || d->n_q < 0 || d->n_r < 0 || d->n_j < 0, errmsg,
           "negative sizes");
  los_test(!(d->xx_step > 0), errmsg, "xx_step must be positive");
  const long n_k3 = d->n_pairs > 0 ? d->pair_off[d->n_pairs] : 0;
  los_test(d->n_pairs > 0 && d->pair_off[0] != 0, errmsg, "pair_off[0] must be 0");
  // source/bispectra2.c:1278-1281: trapezoidal integration needs two k3 points
  for (int p = 0; p < d->n_pairs; ++p)
    los_test(d->pair_off[p + 1] - d->pair_off[p] < 2, errmsg,
             "pair %d: integration grid has less than two elements, cannot use trapezoidal "
             "integration", p);
  double kmax = 0, rmax = 0;
  for (long i = 0; i < n_k3; ++i) {
    los_test(!(d->k3[i] >= 0), errmsg, "k3[%ld] = %g is negative", i, d->k3[i]);
    if (d->k3[i] > kmax) kmax = d->k3[i];
  }
  for (int i = 0; i < d->n_r; ++i) {
    los_test(!(d->r[i] >= 0), errmsg, "r[%d] = %g is negative", i, d->r[i]);
    if (d->r[i] > rmax) rmax = d->r[i];
  }
  // every j table a row uses must hold the nodes of the largest argument
  // (source/bessel2.c:530-533 checks k_max r <= x_max under DEBUG)
  for (int q = 0; q < d->n_q; ++q) {
    const int iL = d->index_L[q];
    if (iL < 0) continue;
    los_test(iL >= d->n_j, errmsg, "row %d: index_L %d outside the %d j tables", q, iL, d->n_j);
    const long size = d->j_off[iL + 1] - d->j_off[iL];
    const double x = kmax * rmax, xmin = d->x_min[iL];
    if (x < xmin) continue;
    const double ixd = (x - xmin) / d->xx_step;
    los_test(ixd + 1 >= (double)size, errmsg,
             "row %d: k3_max r_max = %g is beyond the j table of index %d (x_min %g, %ld points)",
             q, x, iL, xmin, size);
  }
  los_k3int *h = (los_k3int *)calloc(1, sizeof *h);
  los_test(h == NULL, errmsg, "out of host memory");
  h->device = device;
  h->n_pairs = d->n_pairs;
  h->n_q = d->n_q;
  h->n_r = d->n_r;
  h->n_j = d->n_j;
  h->n_k3 = n_k3;
  h->xx_step = d->xx_step;
  h->coef = d->coef;
  // trapezoid measure of each pair's grid (source/bispectra2.c:1283-1287)
  std::vector<double> dk(n_k3 > 0 ? n_k3 : 1);
  for (int p = 0; p < d->n_pairs; ++p) {
    const long b = d->pair_off[p], e = d->pair_off[p + 1];
    dk[b] = d->k3[b + 1] - d->k3[b];
    for (long i = b + 1; i < e - 1; ++i) dk[i] = d->k3[i + 1] - d->k3[i - 1];
    dk[e - 1] = d->k3[e - 1] - d->k3[e - 2];
  }
  // algorithmic counts: live terms, bytes (inputs once, used j windows once, outputs)
  // live terms: T != 0 and k3 r >= x_min; fl(k3 r) is monotone in r, so the
  // count over r is a binary search on the sorted r grid with the same test
  long live = 0;
  std::vector<double> rs(d->r, d->r + d->n_r);
  std::sort(rs.begin(), rs.end());
  std::vector<char> used(d->n_j > 0 ? d->n_j : 1, 0);
  for (int q = 0; q < d->n_q; ++q) {
    const int iL = d->index_L[q];
    if (iL < 0) continue;
    used[iL] = 1;
    const double xmin = d->x_min[iL];
    const double *T = d->transfer + (long)q * n_k3;
    for (long i = 0; i < n_k3; ++i) {
      if (T[i] == 0.) continue;
      const double k = d->k3[i];
      int lo = 0, hi = d->n_r;  // first sorted r with k r >= xmin
      while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (k * rs[mid] >= xmin) hi = mid; else lo = mid + 1;
      }
      live += d->n_r - lo;
    }
  }
  double jbytes = 0;
  for (int t = 0; t < d->n_j; ++t)
    if (used[t]) {
      const long size = d->j_off[t + 1] - d->j_off[t];
      const double ixd = (kmax * rmax - d->x_min[t]) / d->xx_step + 2;
      jbytes += 8.0 * (ixd < 0 ? 0 : (ixd < size ? ixd : size));
    }
  h->live_terms = live;
  h->alg_bytes = 8.0 * n_k3 * (2 + d->n_q) + jbytes + 8.0 * d->n_q * d->n_r * d->n_pairs;
  const long n_jv = d->n_j > 0 ? d->j_off[d->n_j] : 0;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreate(&h->stream);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = upload(&h->pair_off, d->pair_off, (long)d->n_pairs + 1);
  if (e == hipSuccess) e = upload(&h->k3, d->k3, n_k3);
  if (e == hipSuccess) e = upload(&h->dk3, dk.data(), n_k3);
  if (e == hipSuccess) e = upload(&h->transfer, d->transfer, n_k3 * d->n_q);
  if (e == hipSuccess) e = upload(&h->index_L, d->index_L, (long)d->n_q);
  if (e == hipSuccess) e = upload(&h->j_off, d->j_off, (long)d->n_j + 1);
  if (e == hipSuccess) e = upload(&h->j, d->j, n_jv);
  if (e == hipSuccess) e = upload(&h->x_min, d->x_min, (long)d->n_j);
  if (e == hipSuccess) e = upload(&h->r, d->r, (long)d->n_r);
  if (e == hipSuccess)
    e = hipMalloc((void **)&h->out, (size_t)d->n_q * d->n_r * d->n_pairs * sizeof(double) + 8);
  if (e != hipSuccess) {
    snprintf(errmsg, _ERRORMSGSIZE_, "los_k3int_create: %s", hipGetErrorString(e));
    los_k3int_destroy(h);
    return _FAILURE_;
  }
  *out = h;
  return _SUCCESS_;
}

int los_k3int_run(los_k3int *h, int reps, float *ms, char *errmsg)
{
  KCHK(hipSetDevice(h->device));
  if (ms) *ms = 0;
  const long n_out = (long)h->n_q * h->n_r * h->n_pairs;
  if (n_out == 0) return _SUCCESS_;
  const int n_rc = (h->n_r + 63) / 64;
  const long blocks = (long)n_rc * h->n_pairs * h->n_q;
  los_test(blocks > 0x7fffffffL, errmsg, "los_k3int_run: %ld workgroups is too many", blocks);
  if (reps < 1) reps = 1;
  KCHK(hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < reps; ++i)
    hipLaunchKernelGGL(los_k3int_kernel, dim3((unsigned)blocks), dim3(64), 0, h->stream,
                       h->pair_off, h->k3, h->dk3, h->transfer, h->n_k3, h->index_L, h->j_off,
                       h->j, h->x_min, h->xx_step, h->r, h->n_r, h->n_pairs, n_rc, h->coef,
                       h->out);
  KCHK(hipGetLastError());
  KCHK(hipEventRecord(h->ev1, h->stream));
  KCHK(hipEventSynchronize(h->ev1));
  float t = 0;
  KCHK(hipEventElapsedTime(&t, h->ev0, h->ev1));
  if (ms) *ms = t / reps;
  return _SUCCESS_;
}

int los_k3int_result(los_k3int *h, double *out, long n, char *errmsg)
{
  const long n_out = (long)h->n_q * h->n_r * h->n_pairs;
  los_test(n != n_out, errmsg, "los_k3int_result: %ld values asked, %ld held", n, n_out);
  if (n_out == 0) return _SUCCESS_;
  KCHK(hipSetDevice(h->device));
  KCHK(hipMemcpy(out, h->out, (size_t)n_out * sizeof(double), hipMemcpyDeviceToHost));
  return _SUCCESS_;
}

int los_k3int_counts(los_k3int *h, double *out, int n)
{
  const double v[2] = {(double)h->live_terms, h->alg_bytes};
  for (int i = 0; i < n && i < 2; ++i) out[i] = v[i];
  return 2;
}

}  // extern "C"
