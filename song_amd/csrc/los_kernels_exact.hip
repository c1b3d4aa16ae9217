// los_kernels_exact.hip — plain FP64 kernels of the LOS stage for gfx950.
//
// This translation unit is compiled with -ffp-contract=off and keeps, for
// every output element, the floating-point operation sequence of the
// reference CPU code, so the results are bit-identical to an FP64 CPU run:
//   k3 spline coefficients  source/transfer2.c:2639 (clamped spline, NR solve)
//   spline evaluation        source/transfer2.c:2700-2706
//   x, index_x, a_J          source/transfer2.c:2195-2202
//   J lerp, integrand        source/transfer2.c:2259-2305
//   trapezoid, 1/2           source/transfer2.c:2317, :2329
//   E = EE+EB, B = EE-EB     source/transfer2.c:1986, :2059
//   /4 /(2l+1) /2            source/transfer2.c:2092-2098
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "los_device.h"

extern "C" {

// ---------------------------------------------------------------------------
// Spline second derivatives in k3, one thread per (pair, tp, tau) column.
// Column layout [tau][kc] inside each (pair, tp) block: column c of pair p
// starts at src_off + (tp*ntau + tau)*nc and is contiguous in kc.
// End slopes from the parabola through the three end points.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
los_spline_k3_exact(const los_pair *pairs, long ncols, const double *kc_all, const double *src,
                    double *dd, int ntau, int tp2_size)
{
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  // every pair owns tp2_size*ntau columns; column r of a pair is (tp*ntau + tau)
  const long per_pair = (long)tp2_size * ntau;
  const int p = (int)(c / per_pair);
  const long r = c - (long)p * per_pair;
  const int n = pairs[p].nc;
  const double *x = kc_all + pairs[p].kc_off;
  const long off = pairs[p].src_off + r * n;
  const double *y = src + off;
  double *d = dd + off;
  double u[LOS_NC_MAX];

  double d0 = ((x[2] - x[0]) * (x[2] - x[0]) * (y[1] - y[0]) -
               (x[1] - x[0]) * (x[1] - x[0]) * (y[2] - y[0])) /
              ((x[2] - x[0]) * (x[1] - x[0]) * (x[2] - x[1]));
  double ddprev = -0.5;
  d[0] = ddprev;
  u[0] = (3.0 / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - d0);
  for (int i = 1; i < n - 1; ++i) {
    double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    double pp = sig * ddprev + 2.0;
    ddprev = (sig - 1.0) / pp;
    d[i] = ddprev;
    double t = (y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * t / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / pp;
  }
  const int a = n - 3, b = n - 2, e = n - 1;
  double dn = ((x[e] - x[a]) * (x[e] - x[a]) * (y[e] - y[b]) -
               (x[e] - x[b]) * (x[e] - x[b]) * (y[e] - y[a])) /
              ((x[e] - x[a]) * (x[e] - x[b]) * (x[b] - x[a]));
  double qn = 0.5;
  double un = (3.0 / (x[e] - x[b])) * (dn - (y[e] - y[b]) / (x[e] - x[b]));
  double next = (un - qn * u[n - 2]) / (qn * ddprev + 1.0);
  d[n - 1] = next;
  for (int k = n - 2; k >= 0; --k) {
    next = d[k] * next + u[k];
    d[k] = next;
  }
}

// ---------------------------------------------------------------------------
// LOS integral, one workgroup per fine k3, one lane per (l,m) offset.
// ---------------------------------------------------------------------------
#define CH 64

__device__ __forceinline__ int index_x_of(double k, double t0mt, double xx_step)
{
  double x = k * t0mt;
  return (int)(x / xx_step);
}

__global__ void __launch_bounds__(1024)
los_integrate_exact(los_params P)
{
  extern __shared__ double smem[];
  double *sS = smem;                                  // [tp2][CH]
  double *sA = sS + (long)P.tp2_size * CH;            // a_J[CH]
  int *sIX = (int *)(sA + CH);                        // index_x[CH]
  __shared__ int s_tau_end;

  const int f = blockIdx.x;
  const int pidx = P.fpair[f];
  const los_pair pr = P.pairs[pidx];
  const double k = P.kf[f];
  const double ks = P.kf[P.fsrc[f]];
  const int br = P.fbr[f];
  const double *kc = P.kc + pr.kc_off;
  const int nc = pr.nc;
  const double h = kc[br + 1] - kc[br];
  const double bw = (ks - kc[br]) / h;
  const double aw = 1 - bw;
  const double ca = aw * aw * aw - aw;
  const double cb = bw * bw * bw - bw;

  const int lm = threadIdx.x;
  const bool active = lm < P.n_lm;

  // Live tau range of this lane: index_x falls with tau, so the terms with
  // index_x >= min_xmin form a prefix of the time grid.
  int tau_end = 0;
  if (active) {
    const int mx = P.min_xmin[lm];
    int lo = 0, hi = P.ntau; // first tau with index_x < mx is in [lo, hi]
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (index_x_of(k, P.t0mt[mid], P.xx_step) >= mx) lo = mid + 1; else hi = mid;
    }
    tau_end = lo;
  }
  if (threadIdx.x == 0) s_tau_end = 0;
  __syncthreads();
  if (active && tau_end > 0) atomicMax(&s_tau_end, tau_end);
  __syncthreads();
  const int tau_end_block = s_tau_end;

  int l = 0, m = 0, mi = 0;
  if (active) {
    l = P.l_of[lm];
    m = P.m_of[lm];
    mi = P.mi_of[lm];
  }

  double accT = 0, accEd = 0, accEm = 0, accBd = 0, accBm = 0;

  for (int c0 = 0; c0 < tau_end_block; c0 += CH) {
    __syncthreads();
    const int nch = min(CH, P.ntau - c0);
    // sources at this k3 for every type, spline-evaluated in the reference's order
    for (int idx = threadIdx.x; idx < P.tp2_size * CH; idx += blockDim.x) {
      const int tp = idx / CH, c = idx - tp * CH;
      if (c >= nch) continue;
      const long base = pr.src_off + ((long)tp * P.ntau + (c0 + c)) * nc;
      const double *S = P.src + base;
      double v;
      if (P.cubic_k3) {
        const double *D = P.dd + base;
        v = aw * S[br] + bw * S[br + 1] + (ca * D[br] + cb * D[br + 1]) * h * h / 6.0;
      } else {
        v = aw * S[br] + bw * S[br + 1];
      }
      sS[idx] = v;
    }
    for (int c = threadIdx.x; c < nch; c += blockDim.x) {
      const double x = k * P.t0mt[c0 + c];
      const int ix = (int)(x / P.xx_step);
      sIX[c] = ix;
      sA[c] = (P.xx[ix + 1] - x) / P.xx_step;
    }
    __syncthreads();
    if (!active) continue;
    const int cend = min(nch, tau_end - c0);
    for (int c = 0; c < cend; ++c) {
      const int ix = sIX[c];
      const double aJ = sA[c];
      const double dt = P.dtau[c0 + c];
      if (P.has_T) {
        double integrand = 0;
        for (int L = 0; L <= P.LmaxT; ++L) {
          if (m > L) continue;
          const long row = (long)P.jTT * P.L_size + L;
          if (ix < P.xmin[row * P.n_lm + lm]) continue;
          const double *J = P.J + (row * P.X + ix) * P.n_lm + lm;
          const double Jv = aJ * J[0] + (1 - aJ) * J[P.n_lm];
          integrand += Jv * sS[P.srcT[L][mi] * CH + c];
        }
        accT += integrand * dt;
      }
      if (P.has_E || P.has_B) {
        double iEd = 0, iEm = 0, iBd = 0, iBm = 0;
        for (int L = 0; L <= P.LmaxP; ++L) {
          if (m > L) continue;
          const long rEE = (long)P.jEE * P.L_size + L;
          if (ix >= P.xmin[rEE * P.n_lm + lm]) {
            const double *J = P.J + (rEE * P.X + ix) * P.n_lm + lm;
            const double Jv = aJ * J[0] + (1 - aJ) * J[P.n_lm];
            if (P.has_E) iEd += Jv * sS[P.srcE[L][mi] * CH + c];
            if (P.has_B) iBd += Jv * sS[P.srcB[L][mi] * CH + c];
          }
          if (P.jEB >= 0 && m != 0) {
            const long rEB = (long)P.jEB * P.L_size + L;
            if (ix >= P.xmin[rEB * P.n_lm + lm]) {
              const double *J = P.J + (rEB * P.X + ix) * P.n_lm + lm;
              const double Jv = aJ * J[0] + (1 - aJ) * J[P.n_lm];
              if (P.has_E && P.has_B) iEm += Jv * sS[P.srcB[L][mi] * CH + c];
              if (P.has_B) iBm += Jv * sS[P.srcE[L][mi] * CH + c];
            }
          }
        }
        accEd += iEd * dt;
        accEm += iEm * dt;
        accBd += iBd * dt;
        accBm += iBm * dt;
      }
    }
  }
  if (!active) return;
  const double den = 2. * l + 1;
  if (P.has_T) {
    double v = accT;
    v *= 0.5;
    v /= 4.;
    v /= den;
    v /= 2.;
    P.out[(long)(P.ttT + lm) * P.F + f] = v;
  }
  if (P.has_E) {
    double d = accEd;
    d *= 0.5;
    double mx = 0;
    if (m != 0) {
      mx = accEm;
      mx *= 0.5;
    }
    double v = d + mx;
    v /= 4.;
    v /= den;
    v /= 2.;
    P.out[(long)(P.ttE + lm) * P.F + f] = v;
  }
  if (P.has_B) {
    double d = accBd;
    d *= 0.5;
    double mx = accBm;
    mx *= 0.5;
    double v = d - mx;
    v /= 4.;
    v /= den;
    v /= 2.;
    P.out[(long)(P.ttB + lm) * P.F + f] = v;
  }
}

// ---------------------------------------------------------------------------
// Launchers (host).
// ---------------------------------------------------------------------------
hipError_t los_launch_spline_exact(const los_pair *pairs, int npairs, const double *kc,
                                   const double *src, double *dd, int ntau, int tp2,
                                   hipStream_t s)
{
  const long ncols = (long)npairs * tp2 * ntau;
  if (ncols == 0) return hipSuccess;
  const int block = 256;
  const long grid = (ncols + block - 1) / block;
  hipLaunchKernelGGL(los_spline_k3_exact, dim3((unsigned)grid), dim3(block), 0, s, pairs, ncols,
                     kc, src, dd, ntau, tp2);
  return hipGetLastError();
}

hipError_t los_launch_integrate_exact(const los_params *P, int nthreads, hipStream_t s)
{
  if (P->F == 0) return hipSuccess;
  const size_t shmem = (size_t)P->tp2_size * CH * sizeof(double) + CH * (sizeof(double) + sizeof(int));
  hipLaunchKernelGGL(los_integrate_exact, dim3(P->F), dim3(nthreads), shmem, s, *P);
  return hipGetLastError();
}

}  // extern "C"
