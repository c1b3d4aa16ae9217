// los_bessel2_gpu.hip — the J_Llm(x) tables of struct bessels2 built on the
// GPU (gfx950), FP64.
//
// Same product as los_bessel2_build (los_bessel2.c), with the semantics of
// SONG's bessel2 module (source/bessel2.c:75-468, :1041-1588, :1669-1815):
// the host keeps the small setup (x grid, J types, the rows and their l1 term
// lists from 3j symbols, los_bessel2_setup); the GPU does the heavy part:
//   los_jl_table   j_l(x) for every l <= l1max at every grid point, one thread
//                  per x, with the recurrences of los_sph_bessel_array
//                  (upward above the turning point, Miller's downward
//                  recurrence normalised on j_0 and j_1 below it);
//   los_jl_xmin    per l1, the grid index below which j_l1 is cut, by the
//                  bisection of the host builder (source/bessel2.c:1669-1815);
//   los_jrow_xmin  per row, index_xmin_J by the bisection on |J| >= J_cut
//                  (source/bessel2.c:1197-1338);
//   los_jrow_fill  every row from its index_xmin_J to the grid end.
// The rows are then copied into the host arrays of struct bessels2 (cubic
// bessel interpolation adds their second derivatives on the host).
//
// Compiled without floating-point contraction, so every operation is the host
// builder's; the results differ from the host tables only through the device
// sin, cos and cbrt (a few units in the last place of j_0 and j_1).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <thread>
#include <vector>

#include "los_common.h"

#define BCHK(call)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      snprintf(errmsg, _ERRORMSGSIZE_, "los_bessel2_build_gpu: %s failed: %s", #call,        \
               hipGetErrorString(e_));                                                       \
      goto fail;                                                                             \
    }                                                                                        \
  } while (0)

// j_l(xx[ix]) for l = 0..lmax into jl[l * nx + ix]: the host's
// los_sph_bessel_array(lmax, x, .) with the output array in global memory.
__global__ void __launch_bounds__(64)
los_jl_table(int lmax, const double *__restrict__ xx, int nx, double *__restrict__ jl)
{
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= nx) return;
  const double x = xx[ix];
  double *out = jl + ix;
  const long ld = nx;
  if (x == 0.0) {
    out[0] = 1.0;
    for (int l = 1; l <= lmax; ++l) out[l * ld] = 0.0;
    return;
  }
  const double s = sin(x), c = cos(x);
  const double j0 = s / x;
  const double j1 = (s / x - c) / x;
  if (x > (double)lmax + 1.0) {
    out[0] = j0;
    if (lmax >= 1) out[ld] = j1;
    double jm = j0, jc = j1;
    for (int l = 1; l < lmax; ++l) {
      const double jn = (2.0 * l + 1.0) / x * jc - jm;
      out[(l + 1) * ld] = jn;
      jm = jc;
      jc = jn;
    }
    return;
  }
  const double top = (double)lmax > x ? (double)lmax : x;
  const int N = (int)top + 40 + (int)(10.0 * cbrt(top + 1.0));
  double fp1 = 0.0, f = 1e-300;
  // Each rescale multiplies the stored values by 1e-250, as the host does. A
  // stored value is at most 1e250 in size, so after three rescales it is an
  // exact (signed) zero and stays one: the loop stops below the position of the
  // third-latest earlier rescale (r3), which leaves every value bit-identical.
  int r1 = INT_MAX, r2 = INT_MAX, r3 = INT_MAX;
  for (int l = N; l >= 1; --l) {
    const double fm1 = (2.0 * l + 1.0) / x * f - fp1;
    fp1 = f;
    f = fm1;
    if (fabs(f) > 1e250) {
      f *= 1e-250;
      fp1 *= 1e-250;
      const int qend = r3 - 1 < lmax ? r3 - 1 : lmax;
      for (int q = l; q <= qend; ++q) out[q * ld] *= 1e-250;
      r3 = r2;
      r2 = r1;
      r1 = l;
    }
    if (l - 1 <= lmax) out[(l - 1) * ld] = f;
  }
  const double big = fabs(f) > fabs(fp1) ? fabs(f) : fabs(fp1);
  const double g0 = f / big, g1 = fp1 / big;
  const double scale = (j0 * g0 + j1 * g1) / (g0 * g0 + g1 * g1) / big;
  // final normalisation, eight values in flight per lane
  for (int l0 = 0; l0 <= lmax; l0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (l0 + u <= lmax) v[u] = out[(l0 + u) * ld];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (l0 + u <= lmax) out[(l0 + u) * ld] = v[u] * scale;
  }
}

// Per l1: the stored range of j_l1 starts one grid point before |j_l1| first
// reaches j_cut (bisection below the turning point); xx_size - 1 when j_l1
// stays below the cut on the whole grid.
__global__ void __launch_bounds__(64)
los_jl_xmin(int l1max, const double *__restrict__ xx, int nx, double xx_step, double j_cut,
            const double *__restrict__ jl, int *__restrict__ xmin_l1)
{
  const int l1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (l1 > l1max) return;
  if (l1 == 0) {
    xmin_l1[0] = 0;
    return;
  }
  const double *j = jl + (long)l1 * nx;
  int hi;
  const double xpeak = l1 + 0.5;
  if (xpeak >= xx[nx - 1]) {
    hi = nx - 1;
  } else {
    hi = (int)(xpeak / xx_step);
    if (hi >= nx) hi = nx - 1;
  }
  if (fabs(j[hi]) < j_cut) {
    xmin_l1[l1] = nx - 1;
    return;
  }
  int lo = 0;
  while (hi - lo > 1) {
    const int mid = (hi + lo) / 2;
    if (fabs(j[mid]) >= j_cut) hi = mid;
    else lo = mid;
  }
  xmin_l1[l1] = lo;
}

// sum over the row's terms of coef * j_l1(xx[ix]), terms below their cut skipped
__device__ __forceinline__ double los_jrow_value(const los_jrow &r, const double *__restrict__ jl,
                                                 long nx, const int *__restrict__ xmin_l1, int ix)
{
  double acc = 0.0;
  for (int t = 0; t < r.nterm; ++t) {
    const int l1 = r.l1[t];
    if (ix < xmin_l1[l1]) continue;
    acc += r.coef[t] * jl[l1 * nx + ix];
  }
  return acc;
}

// Per row still open (xmin < 0): index_xmin_J. 0 when L == l; otherwise the
// bisection between grid 0 and the first point at or past |l - L| + 1/2 on
// |J| >= J_cut; xx_size - 1 when J is below the cut there.
__global__ void __launch_bounds__(64)
los_jrow_xmin(los_jrow *__restrict__ rows, int nrows, const int *__restrict__ Lv,
              const int *__restrict__ lv, const double *__restrict__ xx, int nx, double xx_max,
              double J_cut, const double *__restrict__ jl, const int *__restrict__ xmin_l1)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrows) return;
  los_jrow &r = rows[q];
  if (r.xmin >= 0) return;
  const int L = Lv[r.iL], l = lv[r.il];
  if (L == l) {
    r.xmin = 0;
    return;
  }
  const int l1min = abs(l - L);
  const double xup = l1min + 0.5;
  // the first grid point at or past xup (the grid is increasing: start from the
  // estimate xup / step and step to it)
  int up;
  if (xup >= xx_max) {
    up = nx - 1;
  } else {
    up = (int)(xup / (xx[1] - xx[0]));
    if (up > nx - 1) up = nx - 1;
    while (up > 0 && xx[up - 1] >= xup) up--;
    while (xx[up] < xup) up++;
  }
  double Jv = los_jrow_value(r, jl, nx, xmin_l1, up);
  if (fabs(Jv) < J_cut) {
    r.xmin = nx - 1;
    return;
  }
  int down = 0;
  while (down != up - 1) {
    const int mid = (up + down) / 2;
    Jv = los_jrow_value(r, jl, nx, xmin_l1, mid);
    if (fabs(Jv) >= J_cut) up = mid;
    else down = mid;
  }
  r.xmin = down;
}

// Row order[y] (y = blockIdx.y) from its xmin on into out[off[q] + ix - xmin]
// (rows with no live term keep the zeros the buffer was cleared with). The
// order is l-major, so rows running together read the same j_l1 table rows.
__global__ void __launch_bounds__(256)
los_jrow_fill(const los_jrow *__restrict__ rows, const int *__restrict__ order,
              const long *__restrict__ off, int nx, const double *__restrict__ jl,
              const int *__restrict__ xmin_l1, double *__restrict__ out)
{
  const int q = order[blockIdx.y];
  const los_jrow &r = rows[q];
  if (r.nterm == 0 || r.xmin >= nx - 1) return;
  double *o = out + off[q];
  for (int ix = r.xmin + blockIdx.x * blockDim.x + threadIdx.x; ix < nx;
       ix += gridDim.x * blockDim.x)
    o[ix - r.xmin] = los_jrow_value(r, jl, nx, xmin_l1, ix);
}

extern "C" int los_bessel2_gpu_fill(const los_jbuild *d, char *errmsg)
{
  los_jrow *rows = d->rows;
  const int nrows = d->nrows, l1max = d->l1max, device = d->device;
  const int nx = d->nx;
  std::vector<double *> dst(nrows > 0 ? nrows : 1, nullptr);
  double *dxx = nullptr, *djl = nullptr, *dout = nullptr;
  int *dxmin = nullptr, *dL = nullptr, *dl = nullptr, *dorder = nullptr;
  los_jrow *drows = nullptr;
  long *doff = nullptr;
  std::vector<long> off(nrows + 1, 0);
  std::vector<int> order(nrows);
  double *stg[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = _FAILURE_;
  hipStream_t s = nullptr;

  BCHK(hipSetDevice(device));
  BCHK(hipStreamCreate(&s));
  BCHK(hipMalloc((void **)&dxx, (size_t)nx * sizeof(double)));
  BCHK(hipMalloc((void **)&djl, (size_t)(l1max + 1) * nx * sizeof(double)));
  BCHK(hipMalloc((void **)&dxmin, (size_t)(l1max + 1) * sizeof(int)));
  BCHK(hipMalloc((void **)&drows, (size_t)(nrows > 0 ? nrows : 1) * sizeof(los_jrow)));
  BCHK(hipMalloc((void **)&dL, (size_t)d->L_size * sizeof(int)));
  BCHK(hipMalloc((void **)&dl, (size_t)d->l_size * sizeof(int)));
  BCHK(hipMemcpyAsync(dxx, d->xx, (size_t)nx * sizeof(double), hipMemcpyHostToDevice, s));
  BCHK(hipMemcpyAsync(drows, rows, (size_t)nrows * sizeof(los_jrow), hipMemcpyHostToDevice, s));
  BCHK(hipMemcpyAsync(dL, d->L, (size_t)d->L_size * sizeof(int), hipMemcpyHostToDevice, s));
  BCHK(hipMemcpyAsync(dl, d->l, (size_t)d->l_size * sizeof(int), hipMemcpyHostToDevice, s));

  hipLaunchKernelGGL(los_jl_table, dim3((nx + 63) / 64), dim3(64), 0, s, l1max, dxx, nx, djl);
  BCHK(hipGetLastError());
  hipLaunchKernelGGL(los_jl_xmin, dim3((l1max + 64) / 64), dim3(64), 0, s, l1max, dxx, nx,
                     d->xx_step, d->j_cut, djl, dxmin);
  BCHK(hipGetLastError());
  if (nrows > 0) {
    hipLaunchKernelGGL(los_jrow_xmin, dim3((nrows + 63) / 64), dim3(64), 0, s, drows, nrows, dL,
                       dl, dxx, nx, d->xx_max, d->J_cut, djl, dxmin);
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(rows, drows, (size_t)nrows * sizeof(los_jrow), hipMemcpyDeviceToHost, s));
  }
  BCHK(hipStreamSynchronize(s));

  if (d->alloc_rows(d->ctx, rows, nrows, dst.data(), errmsg) == _FAILURE_) goto fail_quiet;
  for (int q = 0; q < nrows; ++q) off[q + 1] = off[q] + (nx - rows[q].xmin);
  if (nrows > 0 && off[nrows] > 0) {
    BCHK(hipMalloc((void **)&dout, (size_t)off[nrows] * sizeof(double)));
    BCHK(hipMalloc((void **)&doff, (size_t)(nrows + 1) * sizeof(long)));
    BCHK(hipMemcpyAsync(doff, off.data(), (size_t)(nrows + 1) * sizeof(long),
                        hipMemcpyHostToDevice, s));
    BCHK(hipMemsetAsync(dout, 0, (size_t)off[nrows] * sizeof(double), s));
    // rows l-major (stable in the setup order within one l)
    {
      int n = 0;
      for (int il = 0; il < d->l_size; ++il)
        for (int q = 0; q < nrows; ++q)
          if (rows[q].il == il) order[n++] = q;
    }
    BCHK(hipMalloc((void **)&dorder, (size_t)nrows * sizeof(int)));
    BCHK(hipMemcpyAsync(dorder, order.data(), (size_t)nrows * sizeof(int), hipMemcpyHostToDevice,
                        s));
    // rows in the grid's y dimension, at most 65535 per launch
    for (int r0 = 0; r0 < nrows; r0 += 65535) {
      const int nr = nrows - r0 < 65535 ? nrows - r0 : 65535;
      hipLaunchKernelGGL(los_jrow_fill, dim3(64, nr), dim3(256), 0, s, drows, dorder + r0, doff,
                         nx, djl, dxmin, dout);
      BCHK(hipGetLastError());
    }
    // rows back into the host arrays: chunks through two pinned staging
    // buffers, each scattered into the rows by host threads while the next
    // chunk is in flight
    const long total = off[nrows];
    const long chunk = std::min<long>(total, 4L << 20);  // values per chunk (32 MiB)
    for (int b = 0; b < 2; ++b) {
      BCHK(hipHostMalloc((void **)&stg[b], (size_t)chunk * sizeof(double), 0));
      BCHK(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
    }
    const long nchunk = (total + chunk - 1) / chunk;
    auto fetch = [&](long c) {
      const long c0 = c * chunk, n = std::min(chunk, total - c0);
      hipError_t e = hipMemcpyAsync(stg[c & 1], dout + c0, (size_t)n * sizeof(double),
                                    hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipEventRecord(ev[c & 1], s);
      return e;
    };
    BCHK(fetch(0));
    const int nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (long c = 0; c < nchunk; ++c) {
      if (c + 1 < nchunk) BCHK(fetch(c + 1));
      BCHK(hipEventSynchronize(ev[c & 1]));
      const long c0 = c * chunk, c1 = std::min(total, c0 + chunk);
      const double *src = stg[c & 1];
      std::vector<std::thread> th;
      for (int k = 0; k < nthr; ++k)
        th.emplace_back([&, k]() {
          const long a = c0 + (c1 - c0) * k / nthr, z = c0 + (c1 - c0) * (k + 1) / nthr;
          // first row that ends after a
          long q = std::upper_bound(off.begin(), off.end(), a) - off.begin() - 1;
          for (long i = a; i < z && q < nrows;) {
            const long e = std::min(z, off[q + 1]);
            memcpy(dst[q] + (i - off[q]), src + (i - c0), (size_t)(e - i) * sizeof(double));
            i = e;
            ++q;
          }
        });
      for (auto &x : th) x.join();
    }
  }
  rc = _SUCCESS_;
  goto done;
fail:
fail_quiet:
  rc = _FAILURE_;
done:
  if (s) (void)hipStreamSynchronize(s);
  (void)hipFree(dxx);
  (void)hipFree(djl);
  (void)hipFree(dout);
  (void)hipFree(dxmin);
  (void)hipFree(dL);
  (void)hipFree(dl);
  (void)hipFree(drows);
  (void)hipFree(doff);
  (void)hipFree(dorder);
  for (int b = 0; b < 2; ++b) {
    if (stg[b]) (void)hipHostFree(stg[b]);
    if (ev[b]) (void)hipEventDestroy(ev[b]);
  }
  if (s) (void)hipStreamDestroy(s);
  return rc;
}
