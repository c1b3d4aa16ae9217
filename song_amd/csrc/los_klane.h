// los_klane.h — LOS integral with one (l,m) per wave and one fine k3 per lane.
//
// A wave owns one (l,m) offset and a tile of up to 64 consecutive positions of
// the work order (neighbouring k3, one job). Every lane computes its own
// output element(s) from start to end, in the reference's loop order: time
// steps outside, source multipoles L inside (source/transfer2.c:2184-2329).
// Because the wave's lanes share the (l,m), they share the first live index
// index_xmin_J of every J row: a row that is dead for the whole wave at a time
// step is skipped with a uniform branch, and the wave stops at the last time
// step where any of its rows is live for any of its k3. Only the live work of
// the reference's skip rule (source/transfer2.c:2217, :2232) is done, apart
// from lanes whose k3 leaves the live range a few steps before their
// neighbours.
//
// Tables, indexed by the position's rank r in the engine's sorted order:
//   ixa[t][r]           index_x and a_J of (k3, tau) (source/transfer2.c:2195-2202)
//   sk[(tp*ntau + t)][r] sources interpolated in k3 (source/transfer2.c:2700-2706),
//                        times the trapezoid weight for the fast forms
//   Jk[jt][L][lm][x]    J rows, x contiguous, zero outside [xmin, X)
// Each lane gathers J[ix] and J[ix+1] of its row; neighbouring k3 land within a
// few cache lines at any time step.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "los_device.h"

template <class R>
struct KlaneIXA;
template <>
struct KlaneIXA<float> {
  int ix;
  float a;
};
template <>
struct KlaneIXA<double> {
  int ix, pad;
  double a;
};

// ---------------------------------------------------------------------------
// Tables of the k-lane kernels for every loaded fine k3 (rank r = sorted
// position). A workgroup of 256 threads makes a tile of 64 ranks x 64 time
// steps: it reads the sources along tau (their layout) and writes the tables
// along r through an LDS transpose. With WITH_DT the sources are multiplied
// by the trapezoid weight (fast forms); without, they are the reference's
// interpolated sources with its operations (plain FP64 form).
// ---------------------------------------------------------------------------
template <class R, bool WITH_DT>
__global__ void __launch_bounds__(256)
los_klane_prep(los_params P)
{
  __shared__ double tile[64][65];
  __shared__ long r_off[64];
  __shared__ int r_nc[64], r_br[64];
  __shared__ double r_k[64], r_aw[64], r_bw[64], r_ca[64], r_cb[64], r_h[64];
  const int r0 = blockIdx.x * 64, t0 = blockIdx.y * 64;
  const int tid = threadIdx.x;
  const long npr = P.k_npad;
  if (tid < 64) {
    const int r = min(r0 + tid, P.F - 1);
    const int f = P.korder[r];
    const los_pair pr = P.pairs[P.fpair[f]];
    const int br = P.fbr[f];
    const double *kc = P.kc + pr.kc_off;
    const double ks = P.kf[P.fsrc[f]];
    const double h = kc[br + 1] - kc[br];
    const double bw = (ks - kc[br]) / h;
    const double aw = 1 - bw;
    r_off[tid] = pr.src_off;
    r_nc[tid] = pr.nc;
    r_br[tid] = br;
    r_k[tid] = P.kf[f];
    r_h[tid] = h;
    r_aw[tid] = aw;
    r_bw[tid] = bw;
    r_ca[tid] = aw * aw * aw - aw;
    r_cb[tid] = bw * bw * bw - bw;
  }
  __syncthreads();
  const int lane = tid & 63, grp = tid >> 6;  // 4 groups of 64
  // index_x and a_J: lanes along r
  for (int i = grp; i < 64; i += 4) {
    const int t = t0 + i, r = r0 + lane;
    if (t >= P.ntau || r >= P.F) continue;
    const double x = r_k[lane] * P.t0mt[t];
    const int ix = (int)(x / P.xx_step);
    KlaneIXA<R> o;
    o.ix = ix;
    o.a = (R)((P.xx[ix + 1] - x) / P.xx_step);
    ((KlaneIXA<R> *)P.kixa)[(long)t * npr + r] = o;
  }
  R *out = (R *)P.ksrc;
  for (int tp = 0; tp < P.tp2_size; ++tp) {
    // read along tau: lane = time step, rows = ranks
    for (int i = grp; i < 64; i += 4) {
      const int t = t0 + lane;
      double v = 0;
      if (t < P.ntau && r0 + i < P.F) {
        const long base = r_off[i] + ((long)tp * r_nc[i] + r_br[i]) * P.ntau + t;
        const double *S = P.src + base;
        if (P.cubic_k3) {
          const double *D = P.dd + base;
          v = r_aw[i] * S[0] + r_bw[i] * S[P.ntau] +
              (r_ca[i] * D[0] + r_cb[i] * D[P.ntau]) * r_h[i] * r_h[i] / 6.0;
        } else {
          v = r_aw[i] * S[0] + r_bw[i] * S[P.ntau];
        }
        if (WITH_DT) v = v * P.dtau[t];
      }
      tile[i][lane] = v;
    }
    __syncthreads();
    // write along r: lane = rank, rows = time steps
    for (int i = grp; i < 64; i += 4) {
      const int t = t0 + i, r = r0 + lane;
      if (t < P.ntau && r < npr) out[((long)tp * P.ntau + t) * npr + r] = (R)tile[lane][i];
    }
    __syncthreads();
  }
}

// Largest value of v over the wave.
__device__ __forceinline__ int klane_wave_max(int v)
{
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Number of leading time steps with index_x >= mx for the lane's k3 (index_x
// falls with t).
template <class R>
__device__ __forceinline__ int klane_live_steps(const los_params &P, int r, int mx)
{
  const KlaneIXA<R> *ixa = (const KlaneIXA<R> *)P.kixa;
  int lo = 0, hi = P.ntau;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ixa[(long)mid * P.k_npad + r].ix >= mx) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Common prologue of the k-lane kernels. Workgroup b runs on XCD b % 8;
// q = b / 8 walks that XCD's tiles, costliest first, with the lane group
// fastest, so that the waves of one tile run together and share its sources in
// the XCD's L2. A lane group is NLB (l,m) offsets with the same m (consecutive
// l): every lane handles one k3 for all of them, so each source value it
// loads serves NLB elements.
// ---------------------------------------------------------------------------
template <class R, int NL, int NLB>
struct KlaneWave {
  int r, sp, m, tend;
  bool valid;
  int lm[NLB];
  bool act[NLB];
  int xmT[NLB][NL], xmE[NLB][NL], xmM[NLB][NL];
  int tpT[NL], tpE[NL], tpB[NL];
  los_tile tl;
  los_job job;

  // false: nothing to do for this workgroup (uniform)
  __device__ __forceinline__ bool init(const los_params &P, bool exact_mix_rule)
  {
    const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3;
    const int tq = q / P.k_ngrp, g = q - tq * P.k_ngrp;
    if (tq >= P.xcd_tile[xcd + 1] - P.xcd_tile[xcd]) return false;
    tl = P.tiles[P.xcd_tile[xcd + 1] - 1 - tq];
    job = P.jobs[P.pos_job[tl.pos0]];
    int first = -1;
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
      lm[j] = P.kgrp[g * NLB + j];
      act[j] = lm[j] >= 0 && lm[j] < job.n_lm;
      if (act[j] && first < 0) first = lm[j];
    }
    if (first < 0) return false;
    m = P.m_of[first];
    const int mi = P.mi_of[first];
    const int lane = threadIdx.x;
    valid = lane < tl.npos;
    sp = tl.pos0 + (valid ? lane : tl.npos - 1);
    r = P.rank_of_f[P.forder[sp]];
    int xall = INT_MAX;
#pragma unroll
    for (int L = 0; L < NL; ++L) {
      tpT[L] = tpE[L] = tpB[L] = 0;
      const bool rowok = m <= L && L < P.L_size;
      if (rowok && P.has_T && L <= P.LmaxT) tpT[L] = P.srcT[L][mi];
      if (rowok && (P.has_E || P.has_B) && L <= P.LmaxP) {
        if (P.has_E) tpE[L] = P.srcE[L][mi];
        if (P.has_B) tpB[L] = P.srcB[L][mi];
      }
#pragma unroll
      for (int j = 0; j < NLB; ++j) {
        xmT[j][L] = xmE[j][L] = xmM[j][L] = INT_MAX;
        if (!act[j] || !rowok) continue;
        if (P.has_T && L <= P.LmaxT && P.jTT >= 0)
          xmT[j][L] = P.xmin[((long)P.jTT * P.L_size + L) * P.n_lm + lm[j]];
        if ((P.has_E || P.has_B) && L <= P.LmaxP) {
          if (P.jEE >= 0) xmE[j][L] = P.xmin[((long)P.jEE * P.L_size + L) * P.n_lm + lm[j]];
          if (P.jEB >= 0 && (!exact_mix_rule || P.has_B))
            xmM[j][L] = P.xmin[((long)P.jEB * P.L_size + L) * P.n_lm + lm[j]];
        }
        xall = min(xall, min(xmT[j][L], min(xmE[j][L], xmM[j][L])));
      }
    }
    tend = klane_wave_max(valid ? klane_live_steps<R>(P, r, xall) : 0);
    return true;
  }
};

// ---------------------------------------------------------------------------
// Fast forms: products in R, partial sums over FOLD time steps in Pol::part_t,
// folded into Pol::acc_t (TwoSum for the FP32-compensated form). Every row's
// values are loaded at every time step, live or not (rows the reference skips
// are zero-filled in the table, their source type is any valid one), so that
// a lane has its loads of a step in flight together; the skip rule is applied
// as a select.
// ---------------------------------------------------------------------------
template <class Pol, int NL, int NLB>
__global__ void __launch_bounds__(64)
los_klane(los_params P)
{
  typedef typename Pol::real R;
  typedef typename Pol::part_t Q;
  typedef typename Pol::acc_t A;
  constexpr int FOLD = 32;
  KlaneWave<R, NL, NLB> W;
  if (!W.init(P, false)) return;

  const long np = P.k_npad;
  const R *Jk = (const R *)P.Jk;
  const long rowstep = (long)P.n_lm * P.X;
  const bool doT = P.has_T && P.jTT >= 0, doP = (P.has_E || P.has_B) && P.jEE >= 0;
  const int jEB = P.jEB >= 0 ? P.jEB : P.jEE;
  const KlaneIXA<R> *ixa = (const KlaneIXA<R> *)P.kixa + W.r;
  const R *sk = (const R *)P.ksrc + W.r;
  const long tstride = (long)P.ntau * np;
  const bool emix = W.m != 0;

  A aT[NLB], aE[NLB], aB[NLB];
#pragma unroll
  for (int j = 0; j < NLB; ++j) aT[j] = aE[j] = aB[j] = A{};
  for (int t0 = 0; t0 < W.tend; t0 += FOLD) {
    const int t1 = min(W.tend, t0 + FOLD);
    Q pT[NLB], pE[NLB], pB[NLB];
#pragma unroll
    for (int j = 0; j < NLB; ++j) pT[j] = pE[j] = pB[j] = 0;
    for (int t = t0; t < t1; ++t) {
      const KlaneIXA<R> v = ixa[(long)t * np];
      const int ix = v.ix;
      const R a = v.a;
      const long so = (long)t * np;
      R ts[NL], se[NL], sb[NL];
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        if (doT) ts[L] = sk[(long)W.tpT[L] * tstride + so];
        if (doP) {
          se[L] = sk[(long)W.tpE[L] * tstride + so];
          sb[L] = sk[(long)W.tpB[L] * tstride + so];
        }
      }
#pragma unroll
      for (int j = 0; j < NLB; ++j) {
        if (!W.act[j]) continue;
        const long lmo = (long)W.lm[j] * P.X + ix;
        R tj0[NL], tj1[NL], ej0[NL], ej1[NL], qj0[NL], qj1[NL];
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (doT) {
            const R *row = Jk + ((long)P.jTT * P.L_size + L) * rowstep + lmo;
            tj0[L] = row[0];
            tj1[L] = row[1];
          }
          if (doP) {
            const R *re = Jk + ((long)P.jEE * P.L_size + L) * rowstep + lmo;
            const R *rq = Jk + ((long)jEB * P.L_size + L) * rowstep + lmo;
            ej0[L] = re[0];
            ej1[L] = re[1];
            qj0[L] = rq[0];
            qj1[L] = rq[1];
          }
        }
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (doT) {
            const R Jv = tj1[L] + a * (tj0[L] - tj1[L]);
            pT[j] += ix >= W.xmT[j][L] ? Jv * ts[L] : (R)0;
          }
          if (doP) {
            R Je = ej1[L] + a * (ej0[L] - ej1[L]);
            R Jq = qj1[L] + a * (qj0[L] - qj1[L]);
            Je = ix >= W.xmE[j][L] ? Je : (R)0;
            Jq = ix >= W.xmM[j][L] ? Jq : (R)0;
            pE[j] += Je * se[L] + (emix ? Jq * sb[L] : (R)0);
            pB[j] += Je * sb[L] - Jq * se[L];
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
      Pol::fold(aT[j], pT[j]);
      Pol::fold(aE[j], pE[j]);
      Pol::fold(aB[j], pB[j]);
    }
  }
  if (!W.valid) return;
  const long o = W.job.obase + P.ocol[W.sp];
#pragma unroll
  for (int j = 0; j < NLB; ++j) {
    if (!W.act[j]) continue;
    const int lm = W.lm[j];
    const double den = 2. * P.l_of[lm] + 1;
    if (P.has_T) P.out[o + (long)(W.job.ttT + lm) * W.job.ostride] = Pol::value(aT[j]) * 0.5 / 4. / den / 2.;
    if (P.has_E) P.out[o + (long)(W.job.ttE + lm) * W.job.ostride] = Pol::value(aE[j]) * 0.5 / 4. / den / 2.;
    if (P.has_B) P.out[o + (long)(W.job.ttB + lm) * W.job.ostride] = Pol::value(aB[j]) * 0.5 / 4. / den / 2.;
  }
}

// grid: 8 x (most tiles of one XCD) x (lane groups)
static inline dim3 klane_grid(const los_params *P)
{
  long most = 0;
  for (int x = 0; x < 8; ++x) {
    const long n = P->xcd_tile[x + 1] - P->xcd_tile[x];
    most = n > most ? n : most;
  }
  return dim3((unsigned)(8 * most * P->k_ngrp), 1);
}

template <class R, bool WITH_DT>
static hipError_t launch_klane_prep(const los_params *P, hipStream_t s)
{
  if (P->F == 0 || P->ntau == 0) return hipSuccess;
  const dim3 grid((unsigned)(P->k_npad / 64), (unsigned)((P->ntau + 63) / 64));
  hipLaunchKernelGGL((los_klane_prep<R, WITH_DT>), grid, dim3(256), 0, s, *P);
  return hipGetLastError();
}



// ---------------------------------------------------------------------------
// Plain FP64 form: the reference's operation order per element (compile this
// instantiation without floating-point contraction): per time step the
// integrand of each integral is the sum over live L, in L order, of
// (a_J J_i + (1 - a_J) J_{i+1}) S; the integral adds integrand x dtau; E adds
// the B->E mixing integral only for m != 0, B always subtracts the E->B one;
// then x 1/2, / 4, / (2l+1), / 2 (source/transfer2.c:2259-2329, :1861-2098).
// Loads are issued for every row, live or not, as in the fast kernel; the
// reference's skips are conditions on the additions.
// ---------------------------------------------------------------------------
template <int NL, int NLB>
__global__ void __launch_bounds__(64)
los_klane_exact(los_params P)
{
  KlaneWave<double, NL, NLB> W;
  if (!W.init(P, true)) return;

  const long np = P.k_npad;
  const double *Jk = (const double *)P.Jk;
  const long rowstep = (long)P.n_lm * P.X;
  const bool doT = P.has_T && P.jTT >= 0, doP = (P.has_E || P.has_B) && P.jEE >= 0;
  const int jEB = P.jEB >= 0 ? P.jEB : P.jEE;
  const KlaneIXA<double> *ixa = (const KlaneIXA<double> *)P.kixa + W.r;
  const double *sk = (const double *)P.ksrc + W.r;
  const long tstride = (long)P.ntau * np;
  const int m = W.m;

  double accT[NLB], accEd[NLB], accEm[NLB], accBd[NLB], accBm[NLB];
#pragma unroll
  for (int j = 0; j < NLB; ++j) accT[j] = accEd[j] = accEm[j] = accBd[j] = accBm[j] = 0;
  for (int t = 0; t < W.tend; ++t) {
    const KlaneIXA<double> v = ixa[(long)t * np];
    const int ix = v.ix;
    const double aJ = v.a;
    const double dt = P.dtau[t];
    const long so = (long)t * np;
    double ts[NL], se[NL], sb[NL];
#pragma unroll
    for (int L = 0; L < NL; ++L) {
      if (doT) ts[L] = sk[(long)W.tpT[L] * tstride + so];
      if (doP) {
        se[L] = sk[(long)W.tpE[L] * tstride + so];
        sb[L] = sk[(long)W.tpB[L] * tstride + so];
      }
    }
#pragma unroll
    for (int j = 0; j < NLB; ++j) {
      if (!W.act[j]) continue;
      const long lmo = (long)W.lm[j] * P.X + ix;
      double tj0[NL], tj1[NL], ej0[NL], ej1[NL], qj0[NL], qj1[NL];
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        if (doT) {
          const double *row = Jk + ((long)P.jTT * P.L_size + L) * rowstep + lmo;
          tj0[L] = row[0];
          tj1[L] = row[1];
        }
        if (doP) {
          const double *re = Jk + ((long)P.jEE * P.L_size + L) * rowstep + lmo;
          const double *rq = Jk + ((long)jEB * P.L_size + L) * rowstep + lmo;
          ej0[L] = re[0];
          ej1[L] = re[1];
          qj0[L] = rq[0];
          qj1[L] = rq[1];
        }
      }
      if (doT) {
        double integrand = 0;
#pragma unroll
        for (int L = 0; L < NL; ++L)
          if (ix >= W.xmT[j][L]) integrand += (aJ * tj0[L] + (1 - aJ) * tj1[L]) * ts[L];
        accT[j] += integrand * dt;
      }
      if (doP) {
        double iEd = 0, iEm = 0, iBd = 0, iBm = 0;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (ix >= W.xmE[j][L]) {
            const double Jv = aJ * ej0[L] + (1 - aJ) * ej1[L];
            if (P.has_E) iEd += Jv * se[L];
            if (P.has_B) iBd += Jv * sb[L];
          }
          if (ix >= W.xmM[j][L]) {
            const double Jv = aJ * qj0[L] + (1 - aJ) * qj1[L];
            if (m != 0) iEm += Jv * sb[L];
            iBm += Jv * se[L];
          }
        }
        accEd[j] += iEd * dt;
        accEm[j] += iEm * dt;
        accBd[j] += iBd * dt;
        accBm[j] += iBm * dt;
      }
    }
  }
  if (!W.valid) return;
  const long ocell = W.job.obase + P.ocol[W.sp];
#pragma unroll
  for (int j = 0; j < NLB; ++j) {
    if (!W.act[j]) continue;
    const int lm = W.lm[j];
    const double den = 2. * P.l_of[lm] + 1;
    if (P.has_T) {
      double v = accT[j];
      v *= 0.5;
      v /= 4.;
      v /= den;
      v /= 2.;
      P.out[ocell + (long)(W.job.ttT + lm) * W.job.ostride] = v;
    }
    if (P.has_E) {
      double d = accEd[j];
      d *= 0.5;
      double mx = 0;
      if (m != 0) {
        mx = accEm[j];
        mx *= 0.5;
      }
      double v = d + mx;
      v /= 4.;
      v /= den;
      v /= 2.;
      P.out[ocell + (long)(W.job.ttE + lm) * W.job.ostride] = v;
    }
    if (P.has_B) {
      double d = accBd[j];
      d *= 0.5;
      double mx = accBm[j];
      mx *= 0.5;
      double v = d - mx;
      v /= 4.;
      v /= den;
      v /= 2.;
      P.out[ocell + (long)(W.job.ttB + lm) * W.job.ostride] = v;
    }
  }
}
