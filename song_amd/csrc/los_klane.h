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
// position), one thread per (r, t), r fastest. With WITH_DT the sources are
// multiplied by the trapezoid weight (fast forms); without, they are the
// reference's interpolated sources (plain FP64 form, same operations).
// ---------------------------------------------------------------------------
template <class R, bool WITH_DT>
__global__ void __launch_bounds__(256)
los_klane_prep(los_params P)
{
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long npr = P.k_npad;
  if (id >= npr * P.ntau) return;
  const int t = (int)(id / npr);
  const int r = (int)(id - (long)t * npr);
  if (r >= P.F) return;
  const int f = P.korder[r];
  const double k = P.kf[f];
  const double x = k * P.t0mt[t];
  const int ix = (int)(x / P.xx_step);
  KlaneIXA<R> o;
  o.ix = ix;
  o.a = (R)((P.xx[ix + 1] - x) / P.xx_step);
  ((KlaneIXA<R> *)P.kixa)[id] = o;

  const los_pair pr = P.pairs[P.fpair[f]];
  const int br = P.fbr[f];
  const double *kc = P.kc + pr.kc_off;
  const double ks = P.kf[P.fsrc[f]];
  const double h = kc[br + 1] - kc[br];
  const double bw = (ks - kc[br]) / h;
  const double aw = 1 - bw;
  const double ca = aw * aw * aw - aw;
  const double cb = bw * bw * bw - bw;
  const double dt = P.dtau[t];
  R *out = (R *)P.ksrc;
  for (int tp = 0; tp < P.tp2_size; ++tp) {
    const long base = pr.src_off + ((long)tp * pr.nc + br) * P.ntau + t;
    const double *S = P.src + base;
    double v;
    if (P.cubic_k3) {
      const double *D = P.dd + base;
      v = aw * S[0] + bw * S[P.ntau] + (ca * D[0] + cb * D[P.ntau]) * h * h / 6.0;
    } else {
      v = aw * S[0] + bw * S[P.ntau];
    }
    out[((long)tp * P.ntau + t) * npr + r] = (R)(WITH_DT ? v * dt : v);
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
// Fast forms: products in R, partial sums over FOLD time steps in Pol::part_t,
// folded into Pol::acc_t (TwoSum for the FP32-compensated form).
// blockIdx.x: XCD-aware tile index; blockIdx.y: (l,m) offset.
// ---------------------------------------------------------------------------
template <class Pol, int NL>
__global__ void __launch_bounds__(64)
los_klane(los_params P)
{
  typedef typename Pol::real R;
  typedef typename Pol::part_t Q;
  typedef typename Pol::acc_t A;
  constexpr int FOLD = 32;

  const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3;
  if (q >= P.xcd_tile[xcd + 1] - P.xcd_tile[xcd]) return;
  const los_tile tl = P.tiles[P.xcd_tile[xcd + 1] - 1 - q];
  const los_job job = P.jobs[P.pos_job[tl.pos0]];
  const int lm = blockIdx.y;
  if (lm >= job.n_lm) return;
  const int lane = threadIdx.x;
  const bool valid = lane < tl.npos;
  const int sp = tl.pos0 + (valid ? lane : tl.npos - 1);
  const int r = P.rank_of_f[P.forder[sp]];
  const int m = P.m_of[lm], mi = P.mi_of[lm];
  const long np = P.k_npad;

  // rows of this (l,m): first live index and source type per L (uniform)
  int xmT[NL], xmE[NL], xmM[NL], tpT[NL], tpE[NL], tpB[NL];
  int xall = INT_MAX;
#pragma unroll
  for (int L = 0; L < NL; ++L) {
    xmT[L] = xmE[L] = xmM[L] = INT_MAX;
    tpT[L] = tpE[L] = tpB[L] = 0;
    if (m > L || L >= P.L_size) continue;
    if (P.has_T && L <= P.LmaxT && P.jTT >= 0) {
      xmT[L] = P.xmin[((long)P.jTT * P.L_size + L) * P.n_lm + lm];
      tpT[L] = P.srcT[L][mi];
    }
    if ((P.has_E || P.has_B) && L <= P.LmaxP) {
      if (P.jEE >= 0) xmE[L] = P.xmin[((long)P.jEE * P.L_size + L) * P.n_lm + lm];
      if (P.jEB >= 0) xmM[L] = P.xmin[((long)P.jEB * P.L_size + L) * P.n_lm + lm];
      if (P.has_E) tpE[L] = P.srcE[L][mi];
      if (P.has_B) tpB[L] = P.srcB[L][mi];
    }
    xall = min(xall, min(xmT[L], min(xmE[L], xmM[L])));
  }
  const int tend = klane_wave_max(valid ? klane_live_steps<R>(P, r, xall) : 0);

  const R *Jk = (const R *)P.Jk;
  const long rowstep = (long)P.n_lm * P.X;
  const R *rT = Jk + ((long)P.jTT * P.L_size) * rowstep + (long)lm * P.X;
  const R *rE = Jk + ((long)P.jEE * P.L_size) * rowstep + (long)lm * P.X;
  const R *rM = Jk + ((long)(P.jEB >= 0 ? P.jEB : P.jEE) * P.L_size) * rowstep + (long)lm * P.X;
  const KlaneIXA<R> *ixa = (const KlaneIXA<R> *)P.kixa + r;
  const R *sk = (const R *)P.ksrc + r;
  const long tstride = (long)P.ntau * np;
  const bool doT = P.has_T, doP = P.has_E || P.has_B, mix = P.jEB >= 0;
  const bool emix = m != 0;

  A aT = A{}, aE = A{}, aB = A{};
  for (int t0 = 0; t0 < tend; t0 += FOLD) {
    const int t1 = min(tend, t0 + FOLD);
    Q pT = 0, pE = 0, pB = 0;
    for (int t = t0; t < t1; ++t) {
      const KlaneIXA<R> v = ixa[(long)t * np];
      const int ix = v.ix;
      const R a = v.a;
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        if (m > L) continue;
        if (doT && L <= P.LmaxT) {
          const bool live = ix >= xmT[L];
          if (__ballot(live)) {
            const R *row = rT + (long)L * rowstep;
            const R j0 = row[ix], j1 = row[ix + 1];
            const R s = sk[(long)tpT[L] * tstride + (long)t * np];
            const R Jv = j1 + a * (j0 - j1);
            pT += live ? Jv * s : (R)0;
          }
        }
        if (doP && L <= P.LmaxP) {
          const bool le = ix >= xmE[L], lq = mix && ix >= xmM[L];
          if (__ballot(le || lq)) {
            const R *re = rE + (long)L * rowstep;
            const R *rq = rM + (long)L * rowstep;
            const R e0 = re[ix], e1 = re[ix + 1];
            const R q0 = rq[ix], q1 = rq[ix + 1];
            const R sE = tpE[L] >= 0 && P.has_E ? sk[(long)tpE[L] * tstride + (long)t * np] : (R)0;
            const R sB = P.has_B ? sk[(long)tpB[L] * tstride + (long)t * np] : (R)0;
            R Je = e1 + a * (e0 - e1);
            R Jq = q1 + a * (q0 - q1);
            Je = le ? Je : (R)0;
            Jq = lq ? Jq : (R)0;
            pE += Je * sE + (emix ? Jq * sB : (R)0);
            pB += Je * sB - Jq * sE;
          }
        }
      }
    }
    Pol::fold(aT, pT);
    Pol::fold(aE, pE);
    Pol::fold(aB, pB);
  }
  if (!valid) return;
  const double den = 2. * P.l_of[lm] + 1;
  const long o = job.obase + P.ocol[sp];
  if (P.has_T) P.out[o + (long)(job.ttT + lm) * job.ostride] = Pol::value(aT) * 0.5 / 4. / den / 2.;
  if (P.has_E) P.out[o + (long)(job.ttE + lm) * job.ostride] = Pol::value(aE) * 0.5 / 4. / den / 2.;
  if (P.has_B) P.out[o + (long)(job.ttB + lm) * job.ostride] = Pol::value(aB) * 0.5 / 4. / den / 2.;
}

template <class R, bool WITH_DT>
static hipError_t launch_klane_prep(const los_params *P, hipStream_t s)
{
  const long n = P->k_npad * P->ntau;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((los_klane_prep<R, WITH_DT>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     s, *P);
  return hipGetLastError();
}

// grid: (8 * most tiles of one XCD, (l,m) offsets of the largest job)
static inline dim3 klane_grid(const los_params *P)
{
  int most = 0;
  for (int x = 0; x < 8; ++x) {
    const int n = P->xcd_tile[x + 1] - P->xcd_tile[x];
    most = n > most ? n : most;
  }
  return dim3(8 * most, P->n_lm_max);
}

// ---------------------------------------------------------------------------
// Plain FP64 form: the reference's operation order per element (compile this
// instantiation without floating-point contraction): per time step the
// integrand of each integral is the sum over live L, in L order, of
// (a_J J_i + (1 - a_J) J_{i+1}) S; the integral adds integrand x dtau; E adds
// the B->E mixing integral only for m != 0, B always subtracts the E->B one;
// then x 1/2, / 4, / (2l+1), / 2 (source/transfer2.c:2259-2329, :1861-2098).
// ---------------------------------------------------------------------------
template <int NL>
__global__ void __launch_bounds__(64)
los_klane_exact(los_params P)
{
  const int xcd = blockIdx.x & 7, q = blockIdx.x >> 3;
  if (q >= P.xcd_tile[xcd + 1] - P.xcd_tile[xcd]) return;
  const los_tile tl = P.tiles[P.xcd_tile[xcd + 1] - 1 - q];
  const los_job job = P.jobs[P.pos_job[tl.pos0]];
  const int lm = blockIdx.y;
  if (lm >= job.n_lm) return;
  const int lane = threadIdx.x;
  const bool valid = lane < tl.npos;
  const int sp = tl.pos0 + (valid ? lane : tl.npos - 1);
  const int r = P.rank_of_f[P.forder[sp]];
  const int m = P.m_of[lm], mi = P.mi_of[lm];
  const long np = P.k_npad;

  int xmT[NL], xmE[NL], xmM[NL], tpT[NL], tpE[NL], tpB[NL];
  int xall = INT_MAX;
#pragma unroll
  for (int L = 0; L < NL; ++L) {
    xmT[L] = xmE[L] = xmM[L] = INT_MAX;
    tpT[L] = tpE[L] = tpB[L] = 0;
    if (m > L || L >= P.L_size) continue;
    if (P.has_T && L <= P.LmaxT) {
      xmT[L] = P.xmin[((long)P.jTT * P.L_size + L) * P.n_lm + lm];
      tpT[L] = P.srcT[L][mi];
    }
    if ((P.has_E || P.has_B) && L <= P.LmaxP) {
      xmE[L] = P.xmin[((long)P.jEE * P.L_size + L) * P.n_lm + lm];
      if (P.has_E) tpE[L] = P.srcE[L][mi];
      if (P.has_B) tpB[L] = P.srcB[L][mi];
      if (P.jEB >= 0 && P.has_B) xmM[L] = P.xmin[((long)P.jEB * P.L_size + L) * P.n_lm + lm];
    }
    xall = min(xall, min(xmT[L], min(xmE[L], xmM[L])));
  }
  const int tend = klane_wave_max(valid ? klane_live_steps<double>(P, r, xall) : 0);

  const double *Jk = (const double *)P.Jk;
  const long rowstep = (long)P.n_lm * P.X;
  const double *rT = Jk + ((long)P.jTT * P.L_size) * rowstep + (long)lm * P.X;
  const double *rE = Jk + ((long)P.jEE * P.L_size) * rowstep + (long)lm * P.X;
  const double *rM = Jk + ((long)(P.jEB >= 0 ? P.jEB : P.jEE) * P.L_size) * rowstep + (long)lm * P.X;
  const KlaneIXA<double> *ixa = (const KlaneIXA<double> *)P.kixa + r;
  const double *sk = (const double *)P.ksrc + r;
  const long tstride = (long)P.ntau * np;

  double accT = 0, accEd = 0, accEm = 0, accBd = 0, accBm = 0;
  for (int t = 0; t < tend; ++t) {
    const KlaneIXA<double> v = ixa[(long)t * np];
    const int ix = v.ix;
    const double aJ = v.a;
    const double dt = P.dtau[t];
    if (P.has_T) {
      double integrand = 0;
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        const bool live = ix >= xmT[L];  // INT_MAX for rows the reference skips
        if (__ballot(live)) {
          const double *row = rT + (long)L * rowstep;
          const double j0 = row[ix], j1 = row[ix + 1];
          const double s = sk[(long)tpT[L] * tstride + (long)t * np];
          const double Jv = aJ * j0 + (1 - aJ) * j1;
          if (live) integrand += Jv * s;
        }
      }
      accT += integrand * dt;
    }
    if (P.has_E || P.has_B) {
      double iEd = 0, iEm = 0, iBd = 0, iBm = 0;
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        const bool le = ix >= xmE[L], lq = ix >= xmM[L];
        if (__ballot(le || lq)) {
          const double sE = P.has_E ? sk[(long)tpE[L] * tstride + (long)t * np] : 0.0;
          const double sB = P.has_B ? sk[(long)tpB[L] * tstride + (long)t * np] : 0.0;
          const double *re = rE + (long)L * rowstep;
          const double e0 = re[ix], e1 = re[ix + 1];
          if (le) {
            const double Jv = aJ * e0 + (1 - aJ) * e1;
            if (P.has_E) iEd += Jv * sE;
            if (P.has_B) iBd += Jv * sB;
          }
          const double *rq = rM + (long)L * rowstep;
          const double q0 = rq[ix], q1 = rq[ix + 1];
          if (lq) {
            const double Jv = aJ * q0 + (1 - aJ) * q1;
            if (m != 0) iEm += Jv * sB;
            iBm += Jv * sE;
          }
        }
      }
      accEd += iEd * dt;
      accEm += iEm * dt;
      accBd += iBd * dt;
      accBm += iBm * dt;
    }
  }
  if (!valid) return;
  const double den = 2. * P.l_of[lm] + 1;
  const long ocell = job.obase + P.ocol[sp];
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
