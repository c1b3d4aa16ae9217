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

#include <climits>

#include "los_device.h"
#include "los_klane.h"
#include "los_sweep.h"

extern "C" {

// ---------------------------------------------------------------------------
// Source re-layout and k3 spline in one pass. A workgroup (one wave) takes
// LOS_RS_TW consecutive tau of one (pair, tp) block: it stages their
// [tau][kc] rows (k3 fastest, as SONG stores them, so the read is one
// contiguous run) in LDS, writes them out as [kc][tau] (tau-contiguous for
// the table kernels), then each lane solves the spline of its tau column in
// place in LDS and the second derivatives are written out as [kc][tau].
// The y-independent recurrence values sig_i, p_i and gamma_i = (sig_i-1)/p_i
// of the tridiagonal sweep are computed on the host with the same operations
// and passed per pair (spl[3*kc_off + {0,1,2}*nc + i]); the column's u_i
// overwrite its y_i once y_i is consumed, and the back substitution
// overwrites u_i. LDS row stride `ls` is odd (in doubles), so lanes that walk
// their own columns hit distinct banks.
// ---------------------------------------------------------------------------
#define LOS_RS_TW 64

__global__ void __launch_bounds__(LOS_RS_TW)
los_relayout_spline(const los_pair *pairs, const double *kc_all, const double *spl,
                    const double *raw, double *src, double *dd, int ntau, int tp2, int nblk,
                    int cubic, int ls)
{
  HIP_DYNAMIC_SHARED(double, tile)  // [LOS_RS_TW][ls]
  const int mat = blockIdx.x / nblk;  // pair * tp2 + tp
  const int p = mat / tp2, tp = mat - p * tp2;
  const int t0 = (blockIdx.x - mat * nblk) * LOS_RS_TW;
  const int n = pairs[p].nc;
  const int nt = min(LOS_RS_TW, ntau - t0);
  const long base = pairs[p].src_off + (long)tp * ntau * n;
  const int lane = threadIdx.x;

  // rows t0 .. t0+nt-1 of [tau][kc] are contiguous: nt * n doubles, read in
  // batches of LOS_RS_B loads per lane so that they are all in flight at once
  {
    constexpr int LOS_RS_B = 32;
    const double *r = raw + base + (long)t0 * n;
    const int total = nt * n;
    for (int i0 = 0; i0 < total; i0 += LOS_RS_TW * LOS_RS_B) {
      double v[LOS_RS_B];
#pragma unroll
      for (int u = 0; u < LOS_RS_B; ++u) {
        const int i = i0 + u * LOS_RS_TW + lane;
        v[u] = i < total ? r[i] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < LOS_RS_B; ++u) {
        const int i = i0 + u * LOS_RS_TW + lane;
        if (i < total) {
          const int t = i / n;
          tile[t * ls + (i - t * n)] = v[u];
        }
      }
    }
  }
  __syncthreads();
  const bool live = lane < nt;
  double *so = src + base + t0 + lane;
  if (live)
    for (int kc = 0; kc < n; ++kc) so[(long)kc * ntau] = tile[lane * ls + kc];
  if (!cubic) return;

  if (live) {
    const double *x = kc_all + pairs[p].kc_off;
    const double *sig = spl + 3L * pairs[p].kc_off;
    const double *pv = sig + n;
    const double *gam = pv + n;
    double *y = tile + lane * ls;  // y_i, then u_i, then the second derivatives
    const int a = n - 3, b = n - 2, e = n - 1;
    const double ya = y[a], yb = y[b], ye = y[e];
    const double y0 = y[0], y1 = y[1], y2 = y[2];
    double d0 = ((x[2] - x[0]) * (x[2] - x[0]) * (y1 - y0) - (x[1] - x[0]) * (x[1] - x[0]) * (y2 - y0)) /
                ((x[2] - x[0]) * (x[1] - x[0]) * (x[2] - x[1]));
    double u = (3.0 / (x[1] - x[0])) * ((y1 - y0) / (x[1] - x[0]) - d0);
    y[0] = u;
    double ym = y0, yc = y1;
    for (int i = 1; i < n - 1; ++i) {
      const double yn = y[i + 1];
      double t = (yn - yc) / (x[i + 1] - x[i]) - (yc - ym) / (x[i] - x[i - 1]);
      u = (6.0 * t / (x[i + 1] - x[i - 1]) - sig[i] * u) / pv[i];
      y[i] = u;
      ym = yc;
      yc = yn;
    }
    double dn = ((x[e] - x[a]) * (x[e] - x[a]) * (ye - yb) - (x[e] - x[b]) * (x[e] - x[b]) * (ye - ya)) /
                ((x[e] - x[a]) * (x[e] - x[b]) * (x[b] - x[a]));
    double qn = 0.5;
    double un = (3.0 / (x[e] - x[b])) * (dn - (ye - yb) / (x[e] - x[b]));
    double next = (un - qn * u) / (qn * gam[n - 2] + 1.0);
    y[e] = next;
    for (int k = n - 2; k >= 0; --k) {
      next = gam[k] * next + y[k];
      y[k] = next;
    }
    double *dout = dd + base + t0 + lane;
    for (int kc = 0; kc < n; ++kc) dout[(long)kc * ntau] = y[kc];
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

}  // extern "C"

// NL = number of source multipoles L handled (max(l_max_los_t, l_max_los_p)+1).
// All 2*3*NL J loads of a time step are issued before any of them is used, so
// a wave keeps many independent loads in flight; the reference's skips are
// then applied as selects on the summation, which leaves the per-element
// operation order untouched. Rows that the reference skips still have valid
// addresses (the table is zero-filled below each row's first live index).
template <int NL>
__global__ void __launch_bounds__(LOS_EXACT_BLOCK_MAX)
los_integrate_exact(los_params P)
{
  HIP_DYNAMIC_SHARED(double, smem)
  double *sS = smem;                                  // [tp2][CH]
  double *sA = sS + (long)P.tp2_size * CH;            // a_J[CH]
  int *sIX = (int *)(sA + CH);                        // index_x[CH]
  __shared__ int s_tau_end;

  // workgroups are dispatched round-robin over the 8 XCDs: XCD (b % 8) gets
  // its contiguous range of sorted positions, costliest (largest k3) first
  const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3;
  if (q >= P.xcd_pos[xcd + 1] - P.xcd_pos[xcd]) return;
  const int spos = P.xcd_pos[xcd + 1] - 1 - q;
  const los_job job = P.jobs[P.pos_job[spos]];
  if ((int)(blockIdx.y * blockDim.x) >= job.n_lm) return;  // no lane of this job here
  const int f = P.forder[spos];
  const long ocell = job.obase + P.ocol[spos];
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

  const int lm = blockIdx.y * blockDim.x + threadIdx.x;
  const bool active = lm < job.n_lm;

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
  // Per-lane row constants, hoisted out of the tau loop: first live x index of
  // each (J type, L) row (INT_MAX for rows the reference skips: m > L, EB at
  // m = 0), the row's base offset in J, and the LDS row of its source.
  int xmT[NL], xmE[NL], xmM[NL];
  int sT[NL], sE[NL], sB[NL];
  int rowT[NL], rowE[NL], rowM[NL];  // (jt*L_size + L) * X
#pragma unroll
  for (int L = 0; L < NL; ++L) {
    xmT[L] = xmE[L] = xmM[L] = INT_MAX;
    sT[L] = sE[L] = sB[L] = 0;
    rowT[L] = rowE[L] = rowM[L] = 0;
    if (!active || m > L || L >= P.L_size) continue;
    if (P.has_T && L <= P.LmaxT) {
      const long row = (long)P.jTT * P.L_size + L;
      xmT[L] = P.xmin[row * P.n_lm + lm];
      rowT[L] = (int)row * P.X;
      sT[L] = P.srcT[L][mi] * CH;
    }
    if ((P.has_E || P.has_B) && L <= P.LmaxP) {
      const long rE = (long)P.jEE * P.L_size + L;
      xmE[L] = P.xmin[rE * P.n_lm + lm];
      rowE[L] = (int)rE * P.X;
      if (P.has_E) sE[L] = P.srcE[L][mi] * CH;
      if (P.has_B) sB[L] = P.srcB[L][mi] * CH;
      if (P.jEB >= 0 && P.has_B) {
        const long rM = (long)P.jEB * P.L_size + L;
        xmM[L] = P.xmin[rM * P.n_lm + lm];
        rowM[L] = (int)rM * P.X;
      }
    }
  }

  const double *Jl = P.J + lm;
  double accT = 0, accEd = 0, accEm = 0, accBd = 0, accBm = 0;

  for (int c0 = 0; c0 < tau_end_block; c0 += CH) {
    __syncthreads();
    const int nch = min(CH, P.ntau - c0);
    // sources at this k3 for every type, spline-evaluated in the reference's order
    for (int idx = threadIdx.x; idx < P.tp2_size * CH; idx += blockDim.x) {
      const int tp = idx / CH, c = idx - tp * CH;
      if (c >= nch) continue;
      // tau-contiguous layout [tp][kc][tau]
      const long base = pr.src_off + ((long)tp * nc + br) * P.ntau + (c0 + c);
      const double *S = P.src + base;
      double v;
      if (P.cubic_k3) {
        const double *D = P.dd + base;
        v = aw * S[0] + bw * S[P.ntau] + (ca * D[0] + cb * D[P.ntau]) * h * h / 6.0;
      } else {
        v = aw * S[0] + bw * S[P.ntau];
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
      // loads of live rows only (exec-masked: dead lanes fetch nothing)
      double JT0[NL], JT1[NL], JE0[NL], JE1[NL], JM0[NL], JM1[NL];
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        JT0[L] = JT1[L] = JE0[L] = JE1[L] = JM0[L] = JM1[L] = 0.0;
        if (ix >= xmT[L]) {
          const double *a = Jl + (long)(rowT[L] + ix) * P.n_lm;
          JT0[L] = a[0];
          JT1[L] = a[P.n_lm];
        }
        if (ix >= xmE[L]) {
          const double *b = Jl + (long)(rowE[L] + ix) * P.n_lm;
          JE0[L] = b[0];
          JE1[L] = b[P.n_lm];
        }
        if (ix >= xmM[L]) {
          const double *d = Jl + (long)(rowM[L] + ix) * P.n_lm;
          JM0[L] = d[0];
          JM1[L] = d[P.n_lm];
        }
      }
      if (P.has_T) {
        double integrand = 0;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (ix >= xmT[L]) {  // xmT = INT_MAX for rows the reference skips
            const double Jv = aJ * JT0[L] + (1 - aJ) * JT1[L];
            integrand += Jv * sS[sT[L] + c];
          }
        }
        accT += integrand * dt;
      }
      if (P.has_E || P.has_B) {
        double iEd = 0, iEm = 0, iBd = 0, iBm = 0;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (ix >= xmE[L]) {
            const double Jv = aJ * JE0[L] + (1 - aJ) * JE1[L];
            if (P.has_E) iEd += Jv * sS[sE[L] + c];
            if (P.has_B) iBd += Jv * sS[sB[L] + c];
          }
          if (ix >= xmM[L]) {
            const double Jv = aJ * JM0[L] + (1 - aJ) * JM1[L];
            // E adds the B->E mixing only for m != 0 (source/transfer2.c:1961);
            // B always subtracts the E->B mixing (source/transfer2.c:2034-2059)
            if (m != 0) iEm += Jv * sS[sB[L] + c];
            iBm += Jv * sS[sE[L] + c];
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
    P.out[ocell + (long)(job.ttT + lm) * job.ostride] = v;
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
    P.out[ocell + (long)(job.ttE + lm) * job.ostride] = v;
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
    P.out[ocell + (long)(job.ttB + lm) * job.ostride] = v;
  }
}

// ---------------------------------------------------------------------------
// Launchers (host).
// ---------------------------------------------------------------------------
template <int NL>
static hipError_t launch_exact_nl(const los_params *P, dim3 grid, int threads, size_t shmem,
                                  hipStream_t s)
{
  if (shmem > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)los_integrate_exact<NL>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(los_integrate_exact<NL>, grid, dim3(threads), shmem, s, *P);
  return hipGetLastError();
}

extern "C" {

// Re-layout [tau][kc] -> [kc][tau] of every (pair, tp) block and, if cubic,
// the k3 spline second derivatives dd in the same layout.
hipError_t los_launch_relayout_spline(const los_pair *pairs, int npairs, int nc_max,
                                      const double *kc, const double *spl, const double *raw,
                                      double *src, double *dd, int ntau, int tp2, int cubic,
                                      hipStream_t s)
{
  if (npairs == 0 || ntau == 0) return hipSuccess;
  const int ls = nc_max | 1;
  const size_t lds = (size_t)LOS_RS_TW * ls * sizeof(double);
  if (lds > 65536) {
    const hipError_t r = hipFuncSetAttribute((const void *)los_relayout_spline,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (r != hipSuccess) return r;
  }
  const int nblk = (ntau + LOS_RS_TW - 1) / LOS_RS_TW;
  const long grid = (long)npairs * tp2 * nblk;
  hipLaunchKernelGGL(los_relayout_spline, dim3((unsigned)grid), dim3(LOS_RS_TW), lds, s, pairs, kc,
                     spl, raw, src, dd, ntau, tp2, nblk, cubic, ls);
  return hipGetLastError();
}

// sweep organisation of the plain FP64 form (same per-element operations)
hipError_t los_launch_sweep_prep_exact(const los_params *P, int f_begin, int f_end, hipStream_t s)
{
  return launch_sweep_prep<double, false>(P, f_begin, f_end, s);
}

hipError_t los_launch_sweep_prep_fused_exact(const los_params *P, const double *raw,
                                             const double *spl, int nc_max, int p_begin,
                                             int p_end, hipStream_t s)
{
  return launch_sweep_prep_fused<double, false>(P, raw, spl, nc_max, p_begin, p_end, s);
}

int los_sweep_fused_fits(int tp2, int nc_max) { return sweep_fused_fits(tp2, nc_max) ? 1 : 0; }

hipError_t los_launch_sweep_exact(const los_params *P, int nk, int pers, int n_cu, hipStream_t s)
{
  switch (nk) {
    case 1: return launch_sweep_exact<1>(P, pers != 0, n_cu, s);
    case 2: return launch_sweep_exact<2>(P, pers != 0, n_cu, s);
    case 4: return launch_sweep_exact<4>(P, pers != 0, n_cu, s);
    case 8: return launch_sweep_exact<8>(P, pers != 0, n_cu, s);
    default: return hipErrorInvalidValue;
  }
}

// k-lane organisation of the plain FP64 form (same per-element operations).
hipError_t los_launch_klane_prep_exact(const los_params *P, hipStream_t s)
{
  return launch_klane_prep<double, false>(P, s);
}

hipError_t los_launch_klane_exact(const los_params *P, int nlb, hipStream_t s)
{
  if (P->n_pos == 0 || P->n_lm_max == 0) return hipSuccess;
  const dim3 grid = klane_grid(P);
  if (grid.x == 0) return hipSuccess;
  const int NL = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  switch (NL * 8 + nlb) {
#define LOS_KLX1(n, b) \
  case n * 8 + b: hipLaunchKernelGGL((los_klane_exact<n, b>), grid, dim3(64), 0, s, *P); break;
#define LOS_KLX(n) LOS_KLX1(n, 1) LOS_KLX1(n, 2) LOS_KLX1(n, 4)
    LOS_KLX(1) LOS_KLX(2) LOS_KLX(3) LOS_KLX(4) LOS_KLX(5) LOS_KLX(6) LOS_KLX(7) LOS_KLX(8)
#undef LOS_KLX
#undef LOS_KLX1
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t los_launch_integrate_exact(const los_params *P, int nthreads, hipStream_t s)
{
  (void)nthreads;
  if (P->n_pos == 0 || P->n_lm_max == 0) return hipSuccess;
  // (l,m) lanes split over gridDim.y blocks of at most LOS_EXACT_BLOCK_MAX lanes
  const int nblk = (P->n_lm_max + LOS_EXACT_BLOCK_MAX - 1) / LOS_EXACT_BLOCK_MAX;
  const int per = (P->n_lm_max + nblk - 1) / nblk;
  const int threads = ((per + 63) / 64) * 64;
  const size_t shmem = (size_t)P->tp2_size * CH * sizeof(double) + CH * (sizeof(double) + sizeof(int));
  const int NLr = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  int most = 0;
  for (int x = 0; x < 8; ++x) most = most > P->xcd_pos[x + 1] - P->xcd_pos[x] ? most : P->xcd_pos[x + 1] - P->xcd_pos[x];
  const dim3 grid(8 * most, nblk);
  switch (NLr) {
    case 1: return launch_exact_nl<1>(P, grid, threads, shmem, s);
    case 2: return launch_exact_nl<2>(P, grid, threads, shmem, s);
    case 3: return launch_exact_nl<3>(P, grid, threads, shmem, s);
    case 4: return launch_exact_nl<4>(P, grid, threads, shmem, s);
    case 5: return launch_exact_nl<5>(P, grid, threads, shmem, s);
    case 6: return launch_exact_nl<6>(P, grid, threads, shmem, s);
    case 7: return launch_exact_nl<7>(P, grid, threads, shmem, s);
    case 8: return launch_exact_nl<8>(P, grid, threads, shmem, s);
    default: return hipErrorInvalidValue;
  }
}

}  // extern "C"
