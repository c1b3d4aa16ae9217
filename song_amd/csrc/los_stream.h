// los_stream.h — LOS integral for the fast forms (FP64 with FMA, FP32 with
// compensated accumulation) on gfx950.
//
// A workgroup owns NK fine k3 values that are neighbours in the k-sorted order
// of all pairs, and all (l,m) offsets (one per lane). Loop order per tau chunk:
// (J type, L) row outside, tau inside. At one tau the NK arguments
// x_j = k_j (tau0 - tau) lie within one or two J rows, so each lane loads the
// rows lo, lo+1 (and lo+2 when the group straddles a row) once for all NK
// elements, and keeps them in registers while lo stays the same from one tau
// to the next (at early times x moves by less than a row per step).
//
// The reference skips a (tau, L) term when index_x < index_xmin_J of the row
// (source/transfer2.c:2232); the kernels apply the same test per element, as a
// select on the interpolated J, so that every skipped term is exactly zero
// (near the first live index J is of the order of bessel_J_cut_song, which is
// not negligible for pairs whose whole transfer is of that order). Sources are
// staged in LDS already multiplied by the trapezoid weight.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "los_device.h"

#define LOS_STREAM_BLOCK 512  // max lanes per workgroup

struct StreamFP64 {
  typedef double real;
  typedef double part_t;
  struct acc_t {
    double s;
  };
  static __device__ __forceinline__ void fold(acc_t &a, double p) { a.s += p; }
  static __device__ __forceinline__ double value(const acc_t &a) { return a.s; }
};

struct StreamFP32C {
  typedef float real;
  typedef float part_t;
  struct acc_t {
    float s, c;  // TwoSum (Knuth) running sum and compensation
  };
  static __device__ __forceinline__ void fold(acc_t &a, float t)
  {
    const float s = a.s + t;
    const float bp = s - a.s;
    const float e = (a.s - (s - bp)) + (t - bp);
    a.s = s;
    a.c += e;
  }
  static __device__ __forceinline__ double value(const acc_t &a) { return (double)a.s + (double)a.c; }
};

template <class R>
__device__ __forceinline__ const R *stream_jtab(const los_params &P);
template <>
__device__ __forceinline__ const double *stream_jtab<double>(const los_params &P) { return P.J; }
template <>
__device__ __forceinline__ const float *stream_jtab<float>(const los_params &P) { return P.J32; }

// One tile: the (at most NK) fine k3 at positions tiles[tile].pos0, ... of the
// work order, all of one job, and the (l,m) lanes yb*blockDim.x + threadIdx.x.
// Every thread of the workgroup calls it with the same arguments.
template <class Pol, int NK, int CH>
__device__ __forceinline__ void stream_tile(const los_params &P, int tile, int yb, double *smem_d)
{
  typedef typename Pol::real R;
  typedef typename Pol::part_t Q;
  typedef typename Pol::acc_t A;

  // per-element data (element j = sorted position tile*NK + j)
  __shared__ long e_off[NK], e_out[NK];
  __shared__ int e_nc[NK], e_br[NK];
  __shared__ double e_k[NK], e_aw[NK], e_bw[NK], e_ca[NK], e_cb[NK], e_h[NK];
  __shared__ int s_nv, s_tau_end, s_minx;
  // per chunk: index_x and a_J per (tau, element), first/last index_x per tau
  __shared__ int sIX[CH][NK];
  __shared__ double sAJ[CH][NK];
  __shared__ int sLO[CH], sHI[CH];
  R *sS = (R *)smem_d;  // [NK][tp2 + 1][CH], sources x trapezoid weight, then a zero row

  const int tid = threadIdx.x;
  const los_tile tl = P.tiles[tile];
  const los_job job = P.jobs[P.pos_job[tl.pos0]];
  if (yb * (int)blockDim.x >= job.n_lm) return;  // no lane of this job here (uniform)
  if (tid == 0) {
    s_nv = tl.npos;
    s_tau_end = 0;
    s_minx = INT_MAX;
  }
  if (tid < NK) {
    const int sp = tl.pos0 + min(tid, tl.npos - 1);  // missing elements repeat the last one
    const int f = P.forder[sp];
    const los_pair pr = P.pairs[P.fpair[f]];
    const int br = P.fbr[f];
    const double *kc = P.kc + pr.kc_off;
    const double ks = P.kf[P.fsrc[f]];
    const double h = kc[br + 1] - kc[br];
    const double bw = (ks - kc[br]) / h;
    const double aw = 1 - bw;
    e_out[tid] = job.obase + P.ocol[sp];
    e_k[tid] = P.kf[f];
    e_off[tid] = pr.src_off;
    e_nc[tid] = pr.nc;
    e_br[tid] = br;
    e_h[tid] = h;
    e_aw[tid] = aw;
    e_bw[tid] = bw;
    e_ca[tid] = aw * aw * aw - aw;
    e_cb[tid] = bw * bw * bw - bw;
  }
  const int lm = yb * blockDim.x + tid;
  const bool active = lm < job.n_lm;
  __syncthreads();
  if (active) atomicMin(&s_minx, P.min_xmin[lm]);
  __syncthreads();
  const int nv = s_nv;
  // live prefix of the block: the last element (largest k) lives longest
  if (tid == 0) {
    const double k = e_k[nv - 1];
    const int mx = s_minx;
    int lo = 0, hi = P.ntau;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int)(k * P.t0mt[mid] / P.xx_step) >= mx) lo = mid + 1; else hi = mid;
    }
    s_tau_end = lo;
  }
  __syncthreads();
  const int tau_end = s_tau_end;

  int m = 0, mi = 0, l = 0;
  if (active) {
    l = P.l_of[lm];
    m = P.m_of[lm];
    mi = P.mi_of[lm];
  }
  const int zero_row = P.tp2_size * CH;  // one extra zero row after each element slab
  const int NL = max(P.LmaxT, P.LmaxP) + 1;
  const long slab = (long)P.tp2_size * CH + CH;  // per element, incl. the zero row
  const R *Jb = stream_jtab<R>(P) + lm;
  const long nlm = P.n_lm;

  A aT[NK], aE[NK], aB[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) aT[j] = aE[j] = aB[j] = A{};

  for (int c0 = 0; c0 < tau_end; c0 += CH) {
    const int nch = min(CH, tau_end - c0);
    __syncthreads();
    for (int idx = tid; idx < NK * P.tp2_size * CH; idx += blockDim.x) {
      const int j = idx / (P.tp2_size * CH);
      const int rem = idx - j * P.tp2_size * CH;
      const int tp = rem / CH, c = rem - tp * CH;
      R v = 0;
      if (c < nch && j < nv) {
        const long base = e_off[j] + ((long)tp * e_nc[j] + e_br[j]) * P.ntau + (c0 + c);
        const double *S = P.src + base;
        double s;
        if (P.cubic_k3) {
          const double *D = P.dd + base;
          const double h = e_h[j];
          s = e_aw[j] * S[0] + e_bw[j] * S[P.ntau] +
              (e_ca[j] * D[0] + e_cb[j] * D[P.ntau]) * h * h / 6.0;
        } else {
          s = e_aw[j] * S[0] + e_bw[j] * S[P.ntau];
        }
        v = (R)(s * P.dtau[c0 + c]);
      }
      sS[j * slab + rem] = v;
    }
    for (int idx = tid; idx < NK * CH; idx += blockDim.x) {
      const int j = idx / CH, c = idx - j * CH;
      sS[j * slab + P.tp2_size * CH + c] = 0;  // zero row
      int ix = 0;
      double a = 0;
      if (c < nch) {
        const int jj = min(j, nv - 1);
        const double x = e_k[jj] * P.t0mt[c0 + c];
        ix = (int)(x / P.xx_step);
        a = (P.xx[ix + 1] - x) / P.xx_step;
      }
      sIX[c][j] = ix;
      sAJ[c][j] = a;
    }
    __syncthreads();
    for (int c = tid; c < nch; c += blockDim.x) {
      sLO[c] = sIX[c][0];
      sHI[c] = sIX[c][NK - 1];
    }
    __syncthreads();
    if (!active) continue;

#pragma unroll 1
    for (int L = 0; L < NL; ++L) {
      // LDS offsets of the lane's sources (the zero row for absent rows, whose
      // J is 0 anyway); E takes the B->E mixing only for m != 0
      if (m > L || L >= P.L_size) continue;
      int oT = zero_row, oE = zero_row, oB = zero_row, oBm = zero_row;
      if (P.has_T && L <= P.LmaxT) oT = P.srcT[L][mi] * CH;
      if (L <= P.LmaxP) {
        if (P.has_E) oE = P.srcE[L][mi] * CH;
        if (P.has_B) oB = P.srcB[L][mi] * CH;
        if (P.has_B && m != 0) oBm = oB;
      }
      // first live x index of each row this lane reads
      const int xmT = P.jTT >= 0 ? P.xmin[((long)P.jTT * P.L_size + L) * nlm + lm] : INT_MAX;
      const int xmE = P.jEE >= 0 ? P.xmin[((long)P.jEE * P.L_size + L) * nlm + lm] : INT_MAX;
      const int xmM = P.jEB >= 0 ? P.xmin[((long)P.jEB * P.L_size + L) * nlm + lm] : INT_MAX;
      // ---- temperature rows ----
      if (P.has_T && L <= P.LmaxT) {
        const R *Jr = Jb + (long)(P.jTT * P.L_size + L) * P.X * nlm;
        Q pT[NK];
#pragma unroll
        for (int j = 0; j < NK; ++j) pT[j] = 0;
        int cur = -1;
        R v0 = 0, v1 = 0, v2 = 0;
        bool have2 = false;
        for (int c = 0; c < nch; ++c) {
          const int lo = sLO[c], hi = sHI[c];
          if (lo != cur) {
            v0 = Jr[(long)lo * nlm];
            v1 = Jr[(long)(lo + 1) * nlm];
            cur = lo;
            have2 = false;
          }
          if (hi < xmT) continue;  // every element skips this row at this tau
          if (hi == lo) {
#pragma unroll
            for (int j = 0; j < NK; ++j) {
              const R Jv = v1 + (R)sAJ[c][j] * (v0 - v1);
              pT[j] += Jv * sS[j * slab + oT + c];
            }
          } else if (hi == lo + 1) {
            if (!have2) {
              v2 = Jr[(long)(lo + 2) * nlm];
              have2 = true;
            }
#pragma unroll
            for (int j = 0; j < NK; ++j) {
              const bool d = sIX[c][j] != lo;
              const R a = d ? v1 : v0, b = d ? v2 : v1;
              R Jv = b + (R)sAJ[c][j] * (a - b);
              Jv = (d ? lo + 1 : lo) >= xmT ? Jv : (R)0;
              pT[j] += Jv * sS[j * slab + oT + c];
            }
          } else {
#pragma unroll
            for (int j = 0; j < NK; ++j) {
              const int ix = sIX[c][j];
              const R a = Jr[(long)ix * nlm], b = Jr[(long)(ix + 1) * nlm];
              R Jv = b + (R)sAJ[c][j] * (a - b);
              Jv = ix >= xmT ? Jv : (R)0;
              pT[j] += Jv * sS[j * slab + oT + c];
            }
          }
        }
#pragma unroll
        for (int j = 0; j < NK; ++j) Pol::fold(aT[j], pT[j]);
      }
      // ---- polarisation rows: EE feeds E (S_E) and B (S_B); EB feeds E (S_B, m != 0)
      //      and B (-S_E) ----
      if ((P.has_E || P.has_B) && L <= P.LmaxP) {
        const R *Je = Jb + (long)(P.jEE * P.L_size + L) * P.X * nlm;
        const bool mix = P.jEB >= 0;
        const R *Jm = Jb + (long)((mix ? P.jEB : P.jEE) * P.L_size + L) * P.X * nlm;
        Q pE[NK], pB[NK];
#pragma unroll
        for (int j = 0; j < NK; ++j) pE[j] = pB[j] = 0;
        int cur = -1;
        R e0 = 0, e1 = 0, e2 = 0, q0 = 0, q1 = 0, q2 = 0;
        bool have2 = false;
        for (int c = 0; c < nch; ++c) {
          const int lo = sLO[c], hi = sHI[c];
          if (hi < xmE && hi < xmM) continue;
          if (lo != cur) {
            e0 = Je[(long)lo * nlm];
            e1 = Je[(long)(lo + 1) * nlm];
            if (mix) {
              q0 = Jm[(long)lo * nlm];
              q1 = Jm[(long)(lo + 1) * nlm];
            }
            cur = lo;
            have2 = false;
          }
          if (hi == lo + 1 && !have2) {
            e2 = Je[(long)(lo + 2) * nlm];
            if (mix) q2 = Jm[(long)(lo + 2) * nlm];
            have2 = true;
          }
#pragma unroll
          for (int j = 0; j < NK; ++j) {
            R Jee, Jeb = 0;
            const R a = (R)sAJ[c][j];
            int ix;
            if (hi == lo) {
              ix = lo;
              Jee = e1 + a * (e0 - e1);
              if (mix) Jeb = q1 + a * (q0 - q1);
            } else if (hi == lo + 1) {
              const bool d = sIX[c][j] != lo;
              ix = d ? lo + 1 : lo;
              const R x0 = d ? e1 : e0, x1 = d ? e2 : e1;
              Jee = x1 + a * (x0 - x1);
              if (mix) {
                const R y0 = d ? q1 : q0, y1 = d ? q2 : q1;
                Jeb = y1 + a * (y0 - y1);
              }
            } else {
              ix = sIX[c][j];
              const R x0 = Je[(long)ix * nlm], x1 = Je[(long)(ix + 1) * nlm];
              Jee = x1 + a * (x0 - x1);
              if (mix) {
                const R y0 = Jm[(long)ix * nlm], y1 = Jm[(long)(ix + 1) * nlm];
                Jeb = y1 + a * (y0 - y1);
              }
            }
            Jee = ix >= xmE ? Jee : (R)0;
            Jeb = ix >= xmM ? Jeb : (R)0;
            const R sE_ = sS[j * slab + oE + c];
            const R sB_ = sS[j * slab + oB + c];
            const R sBm = sS[j * slab + oBm + c];
            pE[j] += Jee * sE_ + Jeb * sBm;
            pB[j] += Jee * sB_ - Jeb * sE_;
          }
        }
#pragma unroll
        for (int j = 0; j < NK; ++j) {
          Pol::fold(aE[j], pE[j]);
          Pol::fold(aB[j], pB[j]);
        }
      }
    }
  }
  if (!active) return;
  const double den = 2. * l + 1;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    if (j >= nv) continue;
    const long o = e_out[j];
    if (P.has_T) P.out[o + (long)(job.ttT + lm) * job.ostride] = Pol::value(aT[j]) * 0.5 / 4. / den / 2.;
    if (P.has_E) P.out[o + (long)(job.ttE + lm) * job.ostride] = Pol::value(aE[j]) * 0.5 / 4. / den / 2.;
    if (P.has_B) P.out[o + (long)(job.ttB + lm) * job.ostride] = Pol::value(aB[j]) * 0.5 / 4. / den / 2.;
  }
}

// Workgroup b runs on XCD b % 8 and takes the (b / 8)-th tile of that XCD's
// range, costliest first.
template <class Pol, int NK, int CH>
__global__ void __launch_bounds__(LOS_STREAM_BLOCK)
los_integrate_stream(los_params P)
{
  HIP_DYNAMIC_SHARED(double, smem_d)
  const int x = blockIdx.x & 7, q = blockIdx.x >> 3;
  if (q >= P.xcd_tile[x + 1] - P.xcd_tile[x]) return;
  stream_tile<Pol, NK, CH>(P, P.xcd_tile[x + 1] - 1 - q, blockIdx.y, smem_d);
}

// Persistent form: a grid sized to the chip; the workgroups of XCD x take
// (tile, lane-block) items of that XCD's range from the device counter
// queue[x] until it runs past the range (every workgroup then exits). The
// counters are cleared by los_clear_queue before each launch.
template <class Pol, int NK, int CH>
__global__ void __launch_bounds__(LOS_STREAM_BLOCK)
los_integrate_stream_q(los_params P)
{
  HIP_DYNAMIC_SHARED(double, smem_d)
  __shared__ int s_item;
  const int x = blockIdx.x & 7;
  const int nt = P.xcd_tile[x + 1] - P.xcd_tile[x];
  const int nblk = (P.n_lm_max + blockDim.x - 1) / blockDim.x;
  const int total = nt * nblk;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&P.queue[x], 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= total) break;
    const int t = it / nblk;
    stream_tile<Pol, NK, CH>(P, P.xcd_tile[x + 1] - 1 - t, it - t * nblk, smem_d);
  }
}

__global__ void los_clear_queue(int *q)
{
  if (threadIdx.x < 8) q[threadIdx.x] = 0;
}


// Lanes per workgroup: the (l,m) count split into the fewest blocks of at most
// LOS_STREAM_BLOCK lanes, rounded up to whole waves.
static inline int stream_threads(int n_lm, int *nblk)
{
  *nblk = (n_lm + LOS_STREAM_BLOCK - 1) / LOS_STREAM_BLOCK;
  const int per = (n_lm + *nblk - 1) / *nblk;
  return ((per + 63) / 64) * 64;
}

template <class Pol, int NK, int CH>
static size_t stream_shmem(const los_params *P)
{
  return (size_t)NK * ((size_t)P->tp2_size * CH + CH) * sizeof(typename Pol::real);
}

// Raise the dynamic LDS limit of both kernels of a shape to what a batch with
// tp2_size source types needs (a host call, made at load time: outside any
// stream capture). Shapes whose need exceeds the workgroup LDS are skipped:
// los_shapes never offers them.
template <class Pol, int NK, int CH>
static hipError_t stream_attr_one(int tp2_size)
{
  los_params P;
  P.tp2_size = tp2_size;
  const size_t need = stream_shmem<Pol, NK, CH>(&P);
  if (need <= 64 * 1024 || need + 8 * 1024 > 160 * 1024) return hipSuccess;
  hipError_t e = hipFuncSetAttribute((const void *)los_integrate_stream<Pol, NK, CH>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void *)los_integrate_stream_q<Pol, NK, CH>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)need);
}

// P->xcd_tile must hold the split for tiles of NK positions; P->queue the 8
// counters (persistent form); n_cu the compute units of the device.
template <class Pol, int NK, int CH>
static hipError_t launch_stream(const los_params *P, int persistent, int n_cu, hipStream_t s)
{
  if (P->n_pos == 0 || P->n_lm_max == 0) return hipSuccess;
  int nblk;
  const int threads = stream_threads(P->n_lm_max, &nblk);
  const size_t shmem = stream_shmem<Pol, NK, CH>(P);
  if (shmem + 8 * 1024 > 160 * 1024) return hipErrorInvalidValue;
  if (!persistent) {
    int most = 0;
    for (int x = 0; x < 8; ++x) {
      const int n = P->xcd_tile[x + 1] - P->xcd_tile[x];
      most = n > most ? n : most;
    }
    if (most == 0) return hipSuccess;
    hipLaunchKernelGGL((los_integrate_stream<Pol, NK, CH>), dim3(8 * most, nblk), dim3(threads),
                       shmem, s, *P);
    return hipGetLastError();
  }
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void *)los_integrate_stream_q<Pol, NK, CH>, threads, shmem);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  hipLaunchKernelGGL(los_clear_queue, dim3(1), dim3(64), 0, s, P->queue);
  hipLaunchKernelGGL((los_integrate_stream_q<Pol, NK, CH>), dim3(8 * ((n_cu * per_cu + 7) / 8)),
                     dim3(threads), shmem, s, *P);
  return hipGetLastError();
}
