// los_sweep.h — LOS integral for the fast forms (FP64 with FMA, FP32 with
// compensated accumulation) on gfx950, "sweep" organisation.
//
// A wave (one workgroup) owns a tile of at most NK fine k3 of one job, one m
// and 64 consecutive l (lane slot lp = mi * n_lp + il). Everything that depends
// on (k3, tau) alone is wave-uniform and read with scalar loads from tables
// made by los_sweep_prep: index_x and a_J (source/transfer2.c:2195-2202) and
// the sources interpolated in k3 (source/transfer2.c:2700-2706) times the
// trapezoid weight (source/transfer2.c:2317), for the wave's m and every L.
// Each lane then gathers its J row values, J[jt][L][x][lp] with the slots of
// one m contiguous so that a wave reads 64 consecutive values per x, and the
// NK elements of the tile share a J load whenever their index_x coincide
// (neighbouring k3 in the sorted work order). The reference's skip of rows
// below index_xmin_J (source/transfer2.c:2232) is a select on the
// interpolated J. No LDS and no barriers: waves are independent.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "los_device.h"

template <class R>
struct SweepIXA;
template <>
struct SweepIXA<float> {
  int ix;
  float a;
};
template <>
struct SweepIXA<double> {
  int ix, pad;
  double a;
};


// Both interpolation nodes of a J row at one x slot, J[x] and J[x+1], side by
// side so that a lane reads them with one load.
template <class R>
struct alignas(2 * sizeof(R)) SweepPair {
  R j0, j1;
};

// The J table is read through a raw buffer resource: a load adds the lane's
// byte offset (lane slot) to a wave-uniform 32-bit byte offset (row and x), so
// the address costs one scalar add and the row offsets fit in single SGPRs.
// The table must be smaller than LOS_BUF_MAX bytes (the host checks).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sweep_rsrc(const void *base, unsigned bytes)
{
  // word 3 of a gfx9-family raw buffer descriptor: 32-bit data, no swizzle
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                           0x00020000);
}

// J[x] and J[x+1] of a row at one lane slot, `off` = byte offset of (row, x).
// FP64: one 16-byte load from the pair table; FP32: two loads from the plain
// table (`xstep` bytes apart), which share cache lines with the neighbouring
// time steps. The pair is returned whole, so a conditional load needs no
// register moves (and no wait) where the branches join.
template <class R>
__device__ __forceinline__ SweepPair<R> sweep_row(__amdgpu_buffer_rsrc_t J, unsigned lane,
                                                  unsigned off, unsigned xstep);
template <>
__device__ __forceinline__ SweepPair<double> sweep_row<double>(__amdgpu_buffer_rsrc_t J,
                                                               unsigned lane, unsigned off,
                                                               unsigned)
{
  return __builtin_bit_cast(SweepPair<double>,
                            __builtin_amdgcn_raw_buffer_load_b128(J, lane, off, 0));
}
// FP32 pair table: one 8-byte load.
__device__ __forceinline__ SweepPair<float> sweep_pair_f32(__amdgpu_buffer_rsrc_t J, unsigned lane,
                                                          unsigned off)
{
  return __builtin_bit_cast(SweepPair<float>,
                            __builtin_amdgcn_raw_buffer_load_b64(J, lane, off, 0));
}
template <>
__device__ __forceinline__ SweepPair<float> sweep_row<float>(__amdgpu_buffer_rsrc_t J,
                                                             unsigned lane, unsigned off,
                                                             unsigned xstep)
{
  SweepPair<float> p;
  p.j0 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(J, lane, off, 0));
  p.j1 = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(J, lane, off + xstep, 0));
  return p;
}

template <class R, bool PAIRS>
__device__ __forceinline__ SweepPair<R> sweep_load(__amdgpu_buffer_rsrc_t J, unsigned lane,
                                                   unsigned off, unsigned xstep)
{
  if constexpr (sizeof(R) == 4 && PAIRS)
    return sweep_pair_f32(J, lane, off);
  else
    return sweep_row<R>(J, lane, off, xstep);
}

// Value of lane i of v for the whole wave (v_readlane; every lane must call it).
__device__ __forceinline__ double sweep_lane(double v, int i)
{
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, i);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), i);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int sweep_lane(int v, int i) { return __builtin_amdgcn_readlane(v, i); }
__device__ __forceinline__ float sweep_lane(float v, int i)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), i));
}

// Pair table [row][x][lp] from the sweep table [row][x][lp]: (J[x], J[x+1])
// where x >= index_xmin_J of (row, lp), (0, 0) below it (x + 1 = X reads 0).
// A time step whose index_x lies below the row's index_xmin_J, which the
// reference skips (source/transfer2.c:2232), then adds exactly zero: the
// kernels reading this table need no per-row test.
template <class R>
__global__ void __launch_bounds__(256)
los_sweep_pairs(const R *src, SweepPair<R> *dst, const int *xminp, long rows, int X, int NLP)
{
  const long n = rows * X * NLP;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long lp = i % NLP;
    const long rx = i / NLP;
    const int x = (int)(rx % X);
    const long r = rx / X;
    SweepPair<R> p;
    if (x >= xminp[r * NLP + lp]) {
      p.j0 = src[i];
      p.j1 = x + 1 < X ? src[i + NLP] : (R)0;
    } else {
      p.j0 = p.j1 = (R)0;
    }
    dst[i] = p;
  }
}

// ---------------------------------------------------------------------------
// Per (fine k3, time step) tables of the sweep kernels: index_x and a_J
// (source/transfer2.c:2195-2202) and, per (m, field), the sources of every L
// interpolated in k3 (source/transfer2.c:2700-2706), times the trapezoid
// weight (source/transfer2.c:2317) in the fast forms:
//   ixa[f][t] = {index_x, a_J},   sf[f][t][mi][field][L] (field T, E, B).
// A workgroup takes LOS_SP_FC consecutive fine k3 (fine indices are grouped by
// pair) and LOS_SP_TB consecutive time steps; its threads are (mi, field, t),
// t fastest. They walk the fine k3 in order, each keeping the two coarse
// columns of the current bracket (sources and spline second derivatives,
// every L) in registers, so that moving to the next bracket loads only the new
// right-hand column. Per fine k3 the values are staged in LDS in the table's
// order and the workgroup writes its LOS_SP_TB records as one contiguous block.
// ---------------------------------------------------------------------------
#define LOS_SP_TB 32
#define LOS_SP_FC 16
#define LOS_SP_MAXT (LOS_SP_TB * 3 * LOS_MAX_MI)

template <class R, bool WITH_DT, int NL>
__global__ void __launch_bounds__(LOS_SP_MAXT)
los_sweep_prep(los_params P, int ntb, int f_begin, int f_end)
{
  HIP_DYNAMIC_SHARED(double, stage_raw)
  R *stage = (R *)stage_raw;  // [tt][mi][field][L]
  const int nmc = P.m_size * 3;
  const int fc = blockIdx.x / ntb;
  const int t0 = (blockIdx.x - fc * ntb) * LOS_SP_TB;
  const int f0 = f_begin + fc * LOS_SP_FC, f1 = min(f_end, f0 + LOS_SP_FC);
  const int tt = threadIdx.x % LOS_SP_TB, mc = threadIdx.x / LOS_SP_TB;
  const long nt = P.ntau;
  const int nlive = min(LOS_SP_TB, (int)(nt - t0));
  const bool live = tt < nlive;
  const int t = t0 + (live ? tt : 0);
  const int mi = mc / 3, c = mc - mi * 3;

  // source type of each L for this (m, field); absent ones read type 0 and give 0
  int tp[NL];
#pragma unroll
  for (int L = 0; L < NL; ++L) {
    int v = -1;
    if (c == 0 && P.has_T && L <= P.LmaxT) v = P.srcT[L][mi];
    if (c == 1 && P.has_E && L <= P.LmaxP) v = P.srcE[L][mi];
    if (c == 2 && P.has_B && L <= P.LmaxP) v = P.srcB[L][mi];
    tp[L] = v;
  }
  const bool cubic = P.cubic_k3 != 0;
  const double dt = P.dtau[t];
  const double t0mt = P.t0mt[t];
  const int nval = nlive * nmc * NL;
  R *st = stage + (tt * nmc + mc) * NL;

  // the bracket's two columns stay in registers: they are only ever indexed by
  // the unrolled L (an array whose address is passed on would live in scratch)
  double s0[NL], s1[NL], d0[NL], d1[NL];
  int cur = -2, curp = -1;
  long tstep = 0;
  const double *scol = nullptr, *dcol = nullptr, *kc = nullptr;
  // the next fine k3's pair, bracket and interpolation point, read one step ahead
  int np = P.fpair[f0], nbr = P.fbr[f0];
  double nks = P.kf[P.fsrc[f0]];
  for (int f = f0; f < f1; ++f) {
    const int pi = np, br = nbr;
    const double ks = nks;
    if (f + 1 < f1) {
      np = P.fpair[f + 1];
      nbr = P.fbr[f + 1];
      nks = P.kf[P.fsrc[f + 1]];
    }
    if (pi != curp) {
      const los_pair pr = P.pairs[pi];
      tstep = (long)pr.nc * nt;
      scol = P.src + pr.src_off + t;
      dcol = P.dd + pr.src_off + t;
      kc = P.kc + pr.kc_off;
      curp = pi;
      cur = -2;
    }
    if (br != cur) {
      if (br == cur + 1) {
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          s0[L] = s1[L];
          d0[L] = d1[L];
        }
      } else {
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          const long o = (long)max(tp[L], 0) * tstep + (long)br * nt;
          s0[L] = scol[o];
          d0[L] = cubic ? dcol[o] : 0.0;
        }
      }
#pragma unroll
      for (int L = 0; L < NL; ++L) {
        const long o = (long)max(tp[L], 0) * tstep + (long)(br + 1) * nt;
        s1[L] = scol[o];
        d1[L] = cubic ? dcol[o] : 0.0;
      }
      cur = br;
    }
    const double h = kc[br + 1] - kc[br];
    const double bw = (ks - kc[br]) / h;
    const double aw = 1 - bw;
    const double ca = aw * aw * aw - aw, cb = bw * bw * bw - bw;
#pragma unroll
    for (int L = 0; L < NL; ++L) {
      double v = cubic ? aw * s0[L] + bw * s1[L] + (ca * d0[L] + cb * d1[L]) * h * h / 6.0
                       : aw * s0[L] + bw * s1[L];
      if (WITH_DT) v = v * dt;
      st[L] = (R)(tp[L] >= 0 ? v : 0.0);
    }
    if (mc == 0 && live) {
      const double k = P.kf[f];
      const double x = k * t0mt;
      const int ix = (int)(x / P.xx_step);
      SweepIXA<R> w;
      w.ix = ix;
      w.a = (R)((P.xx[ix + 1] - x) / P.xx_step);
      ((SweepIXA<R> *)P.ixa)[(long)f * nt + t] = w;
    }
    __syncthreads();
    R *dst = (R *)P.sf + ((long)f * nt + t0) * nmc * NL;
    for (int i = threadIdx.x; i < nval; i += blockDim.x) dst[i] = stage[i];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// The same tables straight from the sources as SONG lays them out, with the k3
// spline in the same kernel (no re-laid sources or second derivatives in HBM).
// A workgroup takes one pair and TB consecutive time steps: it stages the
// pair's [tp][tau][kc] rows of those steps in LDS (each (tp) block of TB rows
// is one contiguous run), solves the spline of every (tp, tau) column in LDS
// with the operations of los_relayout_spline (so the second derivatives are
// the same values), and writes the records ixa[f][t] and sf[f][t][mi][field][L]
// of every fine k3 of the pair with the expressions of los_sweep_prep.
// LDS: values and second derivatives, [tp * TB + t][ls] each, ls odd.
// ---------------------------------------------------------------------------
#define LOS_SPF_THREADS 256

template <class R, bool WITH_DT, int NL>
__global__ void __launch_bounds__(LOS_SPF_THREADS)
los_sweep_prep_fused(los_params P, const double *__restrict__ raw, const double *__restrict__ spl,
                     int p_begin, int nchunk, int TB, int ls)
{
  HIP_DYNAMIC_SHARED(double, lds)
  const int p = p_begin + (int)(blockIdx.x / nchunk);
  const int t0 = (int)(blockIdx.x % nchunk) * TB;
  const los_pair pr = P.pairs[p];
  const int n = pr.nc, tp2 = P.tp2_size;
  const long ntau = P.ntau;
  const int nt = min(TB, (int)(ntau - t0));
  double *yv = lds;                         // [tp * TB + t][ls] source values
  double *dv = lds + (long)tp2 * TB * ls;   // the same, second derivatives
  const bool cubic = P.cubic_k3 != 0;

  // each (tp) block of nt rows [tau][kc] is one contiguous run of nt * n values
  const int run = nt * n;
  for (int q = threadIdx.x; q < tp2 * run; q += LOS_SPF_THREADS) {
    const int tp = q / run, r = q - tp * run;
    const int t = r / n, kc = r - t * n;
    const double v = raw[pr.src_off + (long)tp * ntau * n + (long)t0 * n + r];
    yv[(tp * TB + t) * ls + kc] = v;
    dv[(tp * TB + t) * ls + kc] = v;
  }
  __syncthreads();
  if (cubic) {
    const double *x = P.kc + pr.kc_off;
    const double *sig = spl + 3L * pr.kc_off;
    const double *pv = sig + n;
    const double *gam = pv + n;
    for (int col = threadIdx.x; col < tp2 * nt; col += LOS_SPF_THREADS) {
      const int tp = col / nt, t = col - tp * nt;
      double *y = dv + (tp * TB + t) * ls;  // y_i, then u_i, then the second derivatives
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
        double tt = (yn - yc) / (x[i + 1] - x[i]) - (yc - ym) / (x[i] - x[i - 1]);
        u = (6.0 * tt / (x[i + 1] - x[i - 1]) - sig[i] * u) / pv[i];
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
    }
    __syncthreads();
  }

  // records of every fine k3 of the pair: [f][t][mi][field][L], t in the chunk
  const int nmc = P.m_size * 3;
  const int rec = nmc * NL;
  const int per_f = nt * rec;
  const double *kcg = P.kc + pr.kc_off;
  R *sf = (R *)P.sf;
  for (int q = threadIdx.x; q < pr.nf * per_f; q += LOS_SPF_THREADS) {
    const int fl = q / per_f, r = q - fl * per_f;
    const int t = r / rec, w = r - t * rec;
    const int mc = w / NL, L = w - mc * NL;
    const int mi = mc / 3, c = mc - mi * 3;
    int tp = -1;
    if (c == 0 && P.has_T && L <= P.LmaxT) tp = P.srcT[L][mi];
    if (c == 1 && P.has_E && L <= P.LmaxP) tp = P.srcE[L][mi];
    if (c == 2 && P.has_B && L <= P.LmaxP) tp = P.srcB[L][mi];
    const int f = pr.f_off + fl;
    const int br = P.fbr[f];
    const double ks = P.kf[P.fsrc[f]];
    const double h = kcg[br + 1] - kcg[br];
    const double bw = (ks - kcg[br]) / h;
    const double aw = 1 - bw;
    const double ca = aw * aw * aw - aw, cb = bw * bw * bw - bw;
    const int row = (max(tp, 0) * TB + t) * ls;
    const double s0 = yv[row + br], s1 = yv[row + br + 1];
    double v;
    if (cubic) {
      const double d0 = dv[row + br], d1 = dv[row + br + 1];
      v = aw * s0 + bw * s1 + (ca * d0 + cb * d1) * h * h / 6.0;
    } else {
      v = aw * s0 + bw * s1;
    }
    if (WITH_DT) v = v * P.dtau[t0 + t];
    sf[((long)f * ntau + t0) * rec + r] = (R)(tp >= 0 ? v : 0.0);
  }
  for (int q = threadIdx.x; q < pr.nf * nt; q += LOS_SPF_THREADS) {
    const int fl = q / nt, t = q - fl * nt;
    const int f = pr.f_off + fl;
    const double k = P.kf[f];
    const double x = k * P.t0mt[t0 + t];
    const int ix = (int)(x / P.xx_step);
    SweepIXA<R> wv;
    wv.ix = ix;
    wv.a = (R)((P.xx[ix + 1] - x) / P.xx_step);
    ((SweepIXA<R> *)P.ixa)[(long)f * ntau + t0 + t] = wv;
  }
}

// Steps per workgroup of los_sweep_prep_fused: as many as fit in 64 KiB of LDS
// (values and second derivatives of every source type), 1 to 8.
static inline int sweep_fused_steps(int tp2, int ls)
{
  const long per = 2L * tp2 * ls * (long)sizeof(double);
  long tb = per > 0 ? (64L * 1024) / per : 8;
  return (int)(tb < 1 ? 1 : tb > 8 ? 8 : tb);
}

// Whether los_sweep_prep_fused can take a batch: one step of every source type
// (values and second derivatives) must fit in 64 KiB of LDS.
static inline bool sweep_fused_fits(int tp2, int nc_max)
{
  return 2L * tp2 * (nc_max | 1) * (long)sizeof(double) <= 64L * 1024;
}

// Tables of the pairs p_begin .. p_end - 1 from the raw sources.
template <class R, bool WITH_DT>
static hipError_t launch_sweep_prep_fused(const los_params *P, const double *raw, const double *spl,
                                          int nc_max, int p_begin, int p_end, hipStream_t s)
{
  if (P->n_pairs == 0 || p_end <= p_begin || P->ntau == 0) return hipSuccess;
  if (!sweep_fused_fits(P->tp2_size, nc_max)) return hipErrorInvalidValue;
  const int NL = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  const int ls = nc_max | 1;
  const int TB = sweep_fused_steps(P->tp2_size, ls);
  const size_t lds = 2 * (size_t)P->tp2_size * TB * ls * sizeof(double);
  const int nchunk = (P->ntau + TB - 1) / TB;
  const dim3 grid((unsigned)((long)(p_end - p_begin) * nchunk)), block(LOS_SPF_THREADS);
  switch (NL) {
#define LOS_SPF(k)                                                                             \
  case k:                                                                                     \
    hipLaunchKernelGGL((los_sweep_prep_fused<R, WITH_DT, k>), grid, block, lds, s, *P, raw, spl, \
                       p_begin, nchunk, TB, ls);                                              \
    break;
    LOS_SPF(1) LOS_SPF(2) LOS_SPF(3) LOS_SPF(4) LOS_SPF(5) LOS_SPF(6) LOS_SPF(7) LOS_SPF(8)
#undef LOS_SPF
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sweep kernel. Block = one wave; blockIdx.y = wave class (mi, l chunk);
// blockIdx.x = XCD-aware tile index (tiles of XCD b % 8, costliest first).
// ---------------------------------------------------------------------------
// PAIRS: the J table is a pair table zeroed below each row's index_xmin_J
// (always for FP64; an option for FP32), so no per-row test is needed;
// otherwise (FP32 plain table) the interpolated J is selected to zero there.
template <class Pol, int NK, int NL, bool PAIRS>
__device__ __forceinline__ void sweep_item(const los_params &P, int xcd, int q, int w)
{
  typedef typename Pol::real R;
  typedef typename Pol::part_t Q;
  typedef typename Pol::acc_t A;
  constexpr int FOLD = 32;  // time steps per partial sum

  if (q >= P.xcd_tile[xcd + 1] - P.xcd_tile[xcd]) return;
  const int tile = P.xcd_tile[xcd + 1] - 1 - q;
  const int nchunk = P.n_lp >> 6;
  const int mi = w / nchunk;
  const int m = P.m_of_mi[mi];
  const int lp = mi * P.n_lp + (w - mi * nchunk) * 64 + threadIdx.x;
  const los_tile tl = P.tiles[tile];
  const los_job job = P.jobs[P.pos_job[tl.pos0]];
  const int lm = P.lm_of_lp[lp];
  const bool active = lm >= 0 && lm < job.n_lm;
  if (__ballot(active) == 0) return;

  const int NLP = P.m_size * P.n_lp;
  int f[NK];
  int tend = 0;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    f[j] = P.forder[tl.pos0 + min(j, tl.npos - 1)];
    tend = max(tend, P.tau_end_fw[(long)f[j] * P.n_wave + w]);
  }

  // first live x index of each row the lane reads (INT_MAX: never live)
  int xmT[NL], xmE[NL], xmM[NL];
#pragma unroll
  for (int L = 0; L < NL; ++L) {
    xmT[L] = xmE[L] = xmM[L] = INT_MAX;
    if (PAIRS || !active || m > L || L >= P.L_size) continue;
    if (P.has_T && L <= P.LmaxT && P.jTT >= 0)
      xmT[L] = P.xminp[((long)P.jTT * P.L_size + L) * NLP + lp];
    if ((P.has_E || P.has_B) && L <= P.LmaxP) {
      if (P.jEE >= 0) xmE[L] = P.xminp[((long)P.jEE * P.L_size + L) * NLP + lp];
      if (P.jEB >= 0) xmM[L] = P.xminp[((long)P.jEB * P.L_size + L) * NLP + lp];
    }
  }
  // pair table (FP64, or FP32 with PAIRS) or FP32 plain table, all [row][x][lp]
  const __amdgpu_buffer_rsrc_t Jr = sweep_rsrc(P.Jp, P.Jp_bytes);
  constexpr unsigned esz = PAIRS ? 2u * sizeof(R) : (unsigned)sizeof(R);  // bytes per x slot
  const unsigned lane = (unsigned)lp * esz;
  const unsigned xstep = (unsigned)NLP * esz;
  const unsigned rowb = (unsigned)P.X * xstep;
  const SweepIXA<R> *__restrict__ ixa = (const SweepIXA<R> *)P.ixa;
  const R *__restrict__ sf = (const R *)P.sf;
  const bool doT = P.has_T, doP = P.has_E || P.has_B, mix = P.jEB >= 0;
  const unsigned bT = doT ? (unsigned)(P.jTT * P.L_size) * rowb : 0u;
  const unsigned bE = doP ? (unsigned)(P.jEE * P.L_size) * rowb : 0u;
  const unsigned bQ = mix ? (unsigned)(P.jEB * P.L_size) * rowb : 0u;
  const R emix = m != 0 ? (R)1 : (R)0;  // E takes the B->E mixing only for m != 0

  A aT[NK], aE[NK], aB[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) aT[j] = aE[j] = aB[j] = A{};
  // J rows at the current index_x; rows this wave never reads stay zero
  SweepPair<R> tv[NL], ev[NL], qv[NL];
#pragma unroll
  for (int L = 0; L < NL; ++L) tv[L] = ev[L] = qv[L] = SweepPair<R>{(R)0, (R)0};
  // Per-step values shared by all lanes come through the vector path and are
  // broadcast with readlane where used: lane j holds element j's (index_x, a_J),
  // read one step ahead; lane i of src[j] holds value i of element j's
  // [field T, E, B][L] sources of the current step, loaded with the J rows.
  const int li = threadIdx.x;
  int fl = f[0];  // this lane's element (lanes below NK)
#pragma unroll
  for (int j = 1; j < NK; ++j) fl = li == j ? f[j] : fl;
  SweepIXA<R> nxt{};
  if (li < NK && tend > 0) nxt = ixa[(long)fl * P.ntau];

  for (int t0 = 0; t0 < tend; t0 += FOLD) {
    const int t1 = min(tend, t0 + FOLD);
    Q pT[NK], pE[NK], pB[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) pT[j] = pE[j] = pB[j] = 0;
    for (int t = t0; t < t1; ++t) {
      int ix[NK];
      R a[NK];
#pragma unroll
      for (int j = 0; j < NK; ++j) {
        ix[j] = sweep_lane(nxt.ix, j);
        a[j] = sweep_lane(nxt.a, j);
      }
      if (li < NK && t + 1 < tend) nxt = ixa[(long)fl * P.ntau + t + 1];
      R src[NK];
#pragma unroll
      for (int j = 0; j < NK; ++j)
        src[j] = li < 3 * NL ? sf[(((long)f[j] * P.ntau + t) * P.m_size + mi) * 3 * NL + li] : (R)0;
      // the NK elements in turn; the J rows are loaded again only when an
      // element's index_x differs from the previous one's (neighbouring k3
      // usually share it), so one set of row registers serves the tile
#pragma unroll
      for (int j = 0; j < NK; ++j) {
        if (j == 0 || ix[j] != ix[j - 1]) {
          const unsigned ox = (unsigned)ix[j] * xstep;
#pragma unroll
          for (int L = 0; L < NL; ++L) {
            if (m > L) continue;
            if (doT && L <= P.LmaxT) tv[L] = sweep_load<R, PAIRS>(Jr, lane, bT + L * rowb + ox, xstep);
            if (doP && L <= P.LmaxP) {
              ev[L] = sweep_load<R, PAIRS>(Jr, lane, bE + L * rowb + ox, xstep);
              if (mix) qv[L] = sweep_load<R, PAIRS>(Jr, lane, bQ + L * rowb + ox, xstep);
            }
          }
        }
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (m > L) continue;
          const R sT = sweep_lane(src[j], L), sE = sweep_lane(src[j], NL + L),
                  sB = sweep_lane(src[j], 2 * NL + L);
          if (doT && L <= P.LmaxT) {
            R Jv = tv[L].j1 + a[j] * (tv[L].j0 - tv[L].j1);
            if (!PAIRS) Jv = ix[j] >= xmT[L] ? Jv : (R)0;
            pT[j] += Jv * sT;
          }
          if (doP && L <= P.LmaxP) {
            R Je = ev[L].j1 + a[j] * (ev[L].j0 - ev[L].j1);
            R Jm = qv[L].j1 + a[j] * (qv[L].j0 - qv[L].j1);
            if (!PAIRS) {
              Je = ix[j] >= xmE[L] ? Je : (R)0;
              Jm = ix[j] >= xmM[L] ? Jm : (R)0;
            }
            pE[j] += Je * sE + Jm * (emix * sB);
            pB[j] += Je * sB - Jm * sE;
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      Pol::fold(aT[j], pT[j]);
      Pol::fold(aE[j], pE[j]);
      Pol::fold(aB[j], pB[j]);
    }
  }
  if (!active) return;
  const double den = 2. * P.l_of[lm] + 1;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    if (j >= tl.npos) continue;
    const long o = job.obase + P.ocol[tl.pos0 + j];
    if (P.has_T) P.out[o + (long)(job.ttT + lm) * job.ostride] = Pol::value(aT[j]) * 0.5 / 4. / den / 2.;
    if (P.has_E) P.out[o + (long)(job.ttE + lm) * job.ostride] = Pol::value(aE[j]) * 0.5 / 4. / den / 2.;
    if (P.has_B) P.out[o + (long)(job.ttB + lm) * job.ostride] = Pol::value(aB[j]) * 0.5 / 4. / den / 2.;
  }
}

// Item order shared by both sweep kernels. Dispatch form (PERS false): one
// workgroup per item, grid (8 x tiles of the largest XCD range, waves per
// tile). Persistent form (PERS true): the grid is the number of waves the GPU
// holds at once; wave p of an XCD takes that XCD's items p, p + stride, ...
// (item i = wave class i / ntile, tile position i % ntile), so the waves
// resident together work on neighbouring k3 of one class and advance through
// tau together, and the J rows they read stay in the XCD's L2.
#define LOS_SWEEP_ITEMS(CALL)                                   \
  const int xcd = blockIdx.x & 7;                               \
  if (!PERS) {                                                  \
    CALL(P, xcd, (int)(blockIdx.x >> 3), (int)blockIdx.y);      \
    return;                                                     \
  }                                                             \
  const int ntile = P.xcd_tile[xcd + 1] - P.xcd_tile[xcd];      \
  const int stride = (int)(gridDim.x >> 3);                     \
  const int nitems = ntile * P.n_wave;                          \
  for (int i = (int)(blockIdx.x >> 3); i < nitems; i += stride) { \
    const int w = i / ntile;                                    \
    CALL(P, xcd, i - w * ntile, w);                             \
  }

// Sweep kernel, one wave per workgroup (see sweep_item).
template <class Pol, int NK, int NL, bool PERS, bool PAIRS>
__global__ void __launch_bounds__(64)
los_sweep(los_params P)
{
#define LOS_CALL_FAST(...) sweep_item<Pol, NK, NL, PAIRS>(__VA_ARGS__)
  LOS_SWEEP_ITEMS(LOS_CALL_FAST)
#undef LOS_CALL_FAST
}

// Tables of the fine k3 f_begin .. f_end - 1.
template <class R, bool WITH_DT>
static hipError_t launch_sweep_prep(const los_params *P, int f_begin, int f_end, hipStream_t s)
{
  if (P->n_pairs == 0 || f_end <= f_begin || P->ntau == 0) return hipSuccess;
  const int NL = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  const int ntb = (P->ntau + LOS_SP_TB - 1) / LOS_SP_TB;
  const long nfc = ((long)(f_end - f_begin) + LOS_SP_FC - 1) / LOS_SP_FC;
  const dim3 grid((unsigned)(nfc * ntb)), block(LOS_SP_TB * 3 * P->m_size);
  const size_t lds = (size_t)LOS_SP_TB * 3 * P->m_size * NL * sizeof(R);
  switch (NL) {
#define LOS_SP(k) \
  case k:                                                                                  \
    hipLaunchKernelGGL((los_sweep_prep<R, WITH_DT, k>), grid, block, lds, s, *P, ntb, f_begin,  \
                       f_end);                                                                \
    break;
    LOS_SP(1) LOS_SP(2) LOS_SP(3) LOS_SP(4) LOS_SP(5) LOS_SP(6) LOS_SP(7) LOS_SP(8)
#undef LOS_SP
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Grid of a sweep launch: dispatch form (8 x largest XCD tile range, waves
// per tile) or persistent form (8 x the waves each XCD holds at once of kernel
// k, from the occupancy of its 64-thread workgroups on n_cu CUs).
static hipError_t sweep_grid(const los_params *P, bool pers, const void *k, int n_cu, dim3 *grid)
{
  int most = 0;
  for (int x = 0; x < 8; ++x) {
    const int n = P->xcd_tile[x + 1] - P->xcd_tile[x];
    most = n > most ? n : most;
  }
  if (most == 0 || P->n_wave == 0) {
    *grid = dim3(0, 0);
    return hipSuccess;
  }
  if (!pers) {
    *grid = dim3(8 * most, P->n_wave);
    return hipSuccess;
  }
  int per_cu = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  const long waves = ((long)n_cu * per_cu + 7) / 8;                    // per XCD
  const long items = (long)most * P->n_wave;
  *grid = dim3((unsigned)(8 * (waves < items ? waves : items)), 1);
  return hipSuccess;
}

// P->tiles / P->xcd_tile must describe tiles of at most NK positions. FP64
// always reads the pair table; FP32 reads the pair table when `pairs`.
template <class Pol, int NK>
static hipError_t launch_sweep(const los_params *P, bool pers, bool pairs, int n_cu, hipStream_t s)
{
  if (P->n_pos == 0 || P->n_wave == 0) return hipSuccess;
  constexpr bool f64 = sizeof(typename Pol::real) == 8;
  if (f64) pairs = true;
  const int NL = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  switch (NL * 4 + (pers ? 2 : 0) + (pairs ? 1 : 0)) {
#define LOS_SW1(n, b, q)                                                                \
  case n * 4 + (b ? 2 : 0) + (q ? 1 : 0): {                                             \
    if (f64 && !q) return hipErrorInvalidValue;                                         \
    constexpr bool qq = f64 ? true : q;                                                 \
    dim3 grid;                                                                          \
    const hipError_t e =                                                                \
        sweep_grid(P, b, (const void *)los_sweep<Pol, NK, n, b, qq>, n_cu, &grid);      \
    if (e != hipSuccess) return e;                                                      \
    if (grid.x == 0) return hipSuccess;                                                 \
    hipLaunchKernelGGL((los_sweep<Pol, NK, n, b, qq>), grid, dim3(64), 0, s, *P);       \
    break;                                                                              \
  }
#define LOS_SW(n) LOS_SW1(n, false, false) LOS_SW1(n, false, true) LOS_SW1(n, true, false) \
                  LOS_SW1(n, true, true)
    LOS_SW(1) LOS_SW(2) LOS_SW(3) LOS_SW(4) LOS_SW(5) LOS_SW(6) LOS_SW(7) LOS_SW(8)
#undef LOS_SW
#undef LOS_SW1
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plain FP64 form in the sweep organisation (instantiate without
// floating-point contraction): the tables hold the reference's interpolated
// sources (no trapezoid weight), and each element follows the reference's
// operations: per time step the integrand of each integral is the sum over
// live L, in L order, of (a_J J_i + (1 - a_J) J_{i+1}) S; the integral adds
// integrand x dtau; E adds the B->E mixing integral only for m != 0, B always
// subtracts the E->B one; then x 1/2, / 4, / (2l+1), / 2
// (source/transfer2.c:2259-2329, :1861-2098). CUBIC: J from the cubic form of
// source/transfer2.c:2262-2268, a_J Jl + (1 - a_J)(Jr - a_J((a_J + 1) ddl +
// (2 - a_J) ddr) step^2 / 6), with (dd[x], dd[x+1]) pairs zeroed as the J pairs.
// ---------------------------------------------------------------------------
template <int NK, int NL, bool CUBIC>
__device__ __forceinline__ void sweep_exact_item(const los_params &P, int xcd, int q, int w)
{
  if (q >= P.xcd_tile[xcd + 1] - P.xcd_tile[xcd]) return;
  const int tile = P.xcd_tile[xcd + 1] - 1 - q;
  const int nchunk = P.n_lp >> 6;
  const int mi = w / nchunk;
  const int m = P.m_of_mi[mi];
  const int lp = mi * P.n_lp + (w - mi * nchunk) * 64 + threadIdx.x;
  const los_tile tl = P.tiles[tile];
  const los_job job = P.jobs[P.pos_job[tl.pos0]];
  const int lm = P.lm_of_lp[lp];
  const bool active = lm >= 0 && lm < job.n_lm;
  if (__ballot(active) == 0) return;

  const int NLP = P.m_size * P.n_lp;
  int f[NK];
  int tend = 0;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    f[j] = P.forder[tl.pos0 + min(j, tl.npos - 1)];
    tend = max(tend, P.tau_end_fw[(long)f[j] * P.n_wave + w]);
  }
  // The pair table is zero below each row's index_xmin_J (los_sweep_pairs), so
  // a time step the reference skips for a row (source/transfer2.c:2232) adds
  // (a_J 0 + (1 - a_J) 0) S = +-0 here, which leaves the sum's bits unchanged
  // (a sum that starts at +0 never becomes -0); rows the reference never
  // reads (m > L, L beyond the field's L_max, the B->E mixing without B
  // sources) are skipped by wave-uniform tests.
  const bool mixB = P.jEB >= 0 && P.has_B;
  const __amdgpu_buffer_rsrc_t Jr = sweep_rsrc(P.Jp, P.Jp_bytes);
  const __amdgpu_buffer_rsrc_t Dr = sweep_rsrc(CUBIC ? P.Jd : P.Jp, P.Jp_bytes);
  const double step = P.xx_step;
  const unsigned lane = (unsigned)lp * 16u;
  const unsigned xstep = (unsigned)NLP * 16u;
  const unsigned rowb = (unsigned)P.X * xstep;
  const SweepIXA<double> *__restrict__ ixa = (const SweepIXA<double> *)P.ixa;
  const double *__restrict__ sf = (const double *)P.sf;
  const bool doT = P.has_T && P.jTT >= 0, doP = (P.has_E || P.has_B) && P.jEE >= 0;
  const int jEB = P.jEB >= 0 ? P.jEB : P.jEE;
  const unsigned bT = doT ? (unsigned)(P.jTT * P.L_size) * rowb : 0u;
  const unsigned bE = doP ? (unsigned)(P.jEE * P.L_size) * rowb : 0u;
  const unsigned bQ = doP ? (unsigned)(jEB * P.L_size) * rowb : 0u;

  double accT[NK], accEd[NK], accEm[NK], accBd[NK], accBm[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) accT[j] = accEd[j] = accEm[j] = accBd[j] = accBm[j] = 0;
  // J rows at the current index_x (rows this wave never reads stay zero), and
  // their second derivatives for cubic J
  SweepPair<double> tv[NL], ev[NL], qv[NL];
  SweepPair<double> td[CUBIC ? NL : 1], ed[CUBIC ? NL : 1], qd[CUBIC ? NL : 1];
#pragma unroll
  for (int L = 0; L < NL; ++L) tv[L] = ev[L] = qv[L] = SweepPair<double>{0., 0.};
#pragma unroll
  for (int L = 0; L < (CUBIC ? NL : 1); ++L) td[L] = ed[L] = qd[L] = SweepPair<double>{0., 0.};
  // J at a_J from a row's pair (and second-derivative pair), the reference's operations
  auto jval = [&](double aJ, const SweepPair<double> &v, const SweepPair<double> &d) {
    if (CUBIC)
      return aJ * v.j0 + (1. - aJ) * (v.j1 - aJ * ((aJ + 1.) * d.j0 + (2. - aJ) * d.j1) * step *
                                                 step / 6.0);
    return aJ * v.j0 + (1 - aJ) * v.j1;
  };
  // The per-step values that all lanes share come through the vector path and
  // are broadcast with readlane where they are used: lane j holds element j's
  // (index_x, a_J), read one step ahead, and lane i of src[j] holds value i of
  // element j's interpolated sources [field T, E, B][L] of the current step,
  // whose loads go out together with the step's J rows (scalar loads would
  // wait one after the other).
  const int li = threadIdx.x;
  int fl = f[0];  // this lane's element (lanes below NK)
#pragma unroll
  for (int j = 1; j < NK; ++j) fl = li == j ? f[j] : fl;
  SweepIXA<double> nxt{0, 0, 0.};
  if (li < NK && tend > 0) nxt = ixa[(long)fl * P.ntau];
  for (int t = 0; t < tend; ++t) {
    const double dt = P.dtau[t];
    int ix[NK];
    double a[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      ix[j] = sweep_lane(nxt.ix, j);
      a[j] = sweep_lane(nxt.a, j);
    }
    if (li < NK && t + 1 < tend) nxt = ixa[(long)fl * P.ntau + t + 1];
    double src[NK];
#pragma unroll
    for (int j = 0; j < NK; ++j)
      src[j] = li < 3 * NL ? sf[(((long)f[j] * P.ntau + t) * P.m_size + mi) * 3 * NL + li] : 0.;
#pragma unroll
    for (int j = 0; j < NK; ++j) {
      if (j == 0 || ix[j] != ix[j - 1]) {
        const unsigned ox = (unsigned)ix[j] * xstep;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (m > L) continue;
          if (doT && L <= P.LmaxT) {
            tv[L] = sweep_row<double>(Jr, lane, bT + L * rowb + ox, xstep);
            if (CUBIC) td[CUBIC ? L : 0] = sweep_row<double>(Dr, lane, bT + L * rowb + ox, xstep);
          }
          if (doP && L <= P.LmaxP) {
            ev[L] = sweep_row<double>(Jr, lane, bE + L * rowb + ox, xstep);
            qv[L] = sweep_row<double>(Jr, lane, bQ + L * rowb + ox, xstep);
            if (CUBIC) {
              ed[CUBIC ? L : 0] = sweep_row<double>(Dr, lane, bE + L * rowb + ox, xstep);
              qd[CUBIC ? L : 0] = sweep_row<double>(Dr, lane, bQ + L * rowb + ox, xstep);
            }
          }
        }
      }
      const double aJ = a[j];
      if (doT) {
        double integrand = 0;
#pragma unroll
        for (int L = 0; L < NL; ++L)
          if (m <= L && L <= P.LmaxT)
            integrand += jval(aJ, tv[L], td[CUBIC ? L : 0]) * sweep_lane(src[j], L);
        accT[j] += integrand * dt;
      }
      if (doP) {
        double iEd = 0, iEm = 0, iBd = 0, iBm = 0;
#pragma unroll
        for (int L = 0; L < NL; ++L) {
          if (m > L || L > P.LmaxP) continue;
          const double sE = sweep_lane(src[j], NL + L), sB = sweep_lane(src[j], 2 * NL + L);
          {
            const double Jv = jval(aJ, ev[L], ed[CUBIC ? L : 0]);
            if (P.has_E) iEd += Jv * sE;
            if (P.has_B) iBd += Jv * sB;
          }
          if (mixB) {
            const double Jv = jval(aJ, qv[L], qd[CUBIC ? L : 0]);
            if (m != 0) iEm += Jv * sB;
            iBm += Jv * sE;
          }
        }
        accEd[j] += iEd * dt;
        accEm[j] += iEm * dt;
        accBd[j] += iBd * dt;
        accBm[j] += iBm * dt;
      }
    }
  }
  if (!active) return;
  const double den = 2. * P.l_of[lm] + 1;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    if (j >= tl.npos) continue;
    const long o = job.obase + P.ocol[tl.pos0 + j];
    if (P.has_T) {
      double v = accT[j];
      v *= 0.5;
      v /= 4.;
      v /= den;
      v /= 2.;
      P.out[o + (long)(job.ttT + lm) * job.ostride] = v;
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
      P.out[o + (long)(job.ttE + lm) * job.ostride] = v;
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
      P.out[o + (long)(job.ttB + lm) * job.ostride] = v;
    }
  }
}

template <int NK, int NL, bool PERS, bool CUBIC = false>
__global__ void __launch_bounds__(64)
los_sweep_exact(los_params P)
{
#define LOS_CALL_EXACT(...) sweep_exact_item<NK, NL, CUBIC>(__VA_ARGS__)
  LOS_SWEEP_ITEMS(LOS_CALL_EXACT)
#undef LOS_CALL_EXACT
}

template <int NK>
static hipError_t launch_sweep_exact(const los_params *P, bool pers, int n_cu, hipStream_t s)
{
  if (P->n_pos == 0 || P->n_wave == 0) return hipSuccess;
  const int NL = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  switch (NL * 4 + (pers ? 2 : 0) + (P->cubic_J ? 1 : 0)) {
#define LOS_SWX1(n, b, c)                                                             \
  case n * 4 + (b ? 2 : 0) + (c ? 1 : 0): {                                          \
    dim3 grid;                                                                        \
    const hipError_t e =                                                              \
        sweep_grid(P, b, (const void *)los_sweep_exact<NK, n, b, c>, n_cu, &grid);    \
    if (e != hipSuccess) return e;                                                    \
    if (grid.x == 0) return hipSuccess;                                               \
    hipLaunchKernelGGL((los_sweep_exact<NK, n, b, c>), grid, dim3(64), 0, s, *P);     \
    break;                                                                            \
  }
#define LOS_SWX(n) LOS_SWX1(n, false, false) LOS_SWX1(n, true, false) \
                   LOS_SWX1(n, false, true) LOS_SWX1(n, true, true)
    LOS_SWX(1) LOS_SWX(2) LOS_SWX(3) LOS_SWX(4) LOS_SWX(5) LOS_SWX(6) LOS_SWX(7) LOS_SWX(8)
#undef LOS_SWX
#undef LOS_SWX1
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
