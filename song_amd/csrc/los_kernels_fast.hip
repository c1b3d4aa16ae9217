// los_kernels_fast.hip — the LOS integral in its FP64-with-FMA and
// FP32-compensated forms (los_stream.h) for every tile shape the engine's
// autotuner chooses from, plus the FP32 copy of the J table.
#include <hip/hip_runtime.h>

#include "los_stream.h"
#include "los_sweep.h"
#include "los_klane.h"

__global__ void __launch_bounds__(256) los_round_j32(const double *src, float *dst, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// tile shapes: NK fine k3 per workgroup x CH tau values per LDS chunk
#define LOS_SHAPES(X) X(2, 16) X(2, 32) X(4, 16) X(4, 32) X(8, 16) X(8, 32)

template <class Pol>
static hipError_t launch_shape(const los_params *P, const los_shape *sh, int n_cu,
                               hipStream_t s)
{
#define LOS_CASE(NK_, CH_) \
  if (sh->nk == NK_ && sh->ch == CH_) return launch_stream<Pol, NK_, CH_>(P, sh->persistent, n_cu, s);
  LOS_SHAPES(LOS_CASE)
#undef LOS_CASE
  return hipErrorInvalidValue;
}

template <class Pol>
static hipError_t set_attrs(int tp2_size)
{
  hipError_t e = hipSuccess;
#define LOS_ATTR(NK_, CH_) \
  if (e == hipSuccess) e = stream_attr_one<Pol, NK_, CH_>(tp2_size);
  LOS_SHAPES(LOS_ATTR)
#undef LOS_ATTR
  return e;
}

// Plain copy of n doubles (the reference kernel of known traffic that HBM
// counters are calibrated against).
__global__ void __launch_bounds__(256) los_copy_probe(const double *src, double *dst, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

extern "C" {

hipError_t los_launch_copy_probe(const double *src, double *dst, long n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(los_copy_probe, dim3(8192), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t los_launch_round_j32(const double *src, float *dst, long n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(los_round_j32, dim3(2048), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

// fp32: 0 for the FP64-with-FMA form, 1 for the FP32-compensated form
hipError_t los_launch_stream(int fp32, const los_params *P, const los_shape *sh, int n_cu,
                             hipStream_t s)
{
  return fp32 ? launch_shape<StreamFP32C>(P, sh, n_cu, s) : launch_shape<StreamFP64>(P, sh, n_cu, s);
}

hipError_t los_stream_set_attrs(int tp2_size)
{
  hipError_t e = set_attrs<StreamFP64>(tp2_size);
  return e == hipSuccess ? set_attrs<StreamFP32C>(tp2_size) : e;
}

// Sweep organisation: per-(k3, tau) tables, then the sweep kernel with tiles of
// nk positions (1, 2 or 4).
hipError_t los_launch_sweep_prep(int fp32, const los_params *P, int f_begin, int f_end,
                                 hipStream_t s)
{
  return fp32 ? launch_sweep_prep<float, true>(P, f_begin, f_end, s)
              : launch_sweep_prep<double, true>(P, f_begin, f_end, s);
}

// Pair table of the sweep kernels from the sweep J table.
hipError_t los_launch_sweep_pairs(int fp32, const void *src, void *dst, const int *xminp,
                                  long rows, int X, int NLP, hipStream_t s)
{
  if (rows * X * NLP == 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL(los_sweep_pairs<float>, dim3(4096), dim3(256), 0, s, (const float *)src,
                       (SweepPair<float> *)dst, xminp, rows, X, NLP);
  else
    hipLaunchKernelGGL(los_sweep_pairs<double>, dim3(4096), dim3(256), 0, s, (const double *)src,
                       (SweepPair<double> *)dst, xminp, rows, X, NLP);
  return hipGetLastError();
}

hipError_t los_launch_sweep_prep_fused(int fp32, const los_params *P, const double *raw,
                                       const double *spl, int nc_max, int p_begin, int p_end,
                                       hipStream_t s)
{
  return fp32 ? launch_sweep_prep_fused<float, true>(P, raw, spl, nc_max, p_begin, p_end, s)
              : launch_sweep_prep_fused<double, true>(P, raw, spl, nc_max, p_begin, p_end, s);
}

hipError_t los_launch_sweep(int fp32, const los_params *P, int nk, int pers, int pairs, int n_cu,
                            hipStream_t s)
{
  const bool p = pers != 0, q = pairs != 0;
  switch (nk) {
    case 1:
      return fp32 ? launch_sweep<StreamFP32C, 1>(P, p, q, n_cu, s)
                  : launch_sweep<StreamFP64, 1>(P, p, q, n_cu, s);
    case 2:
      return fp32 ? launch_sweep<StreamFP32C, 2>(P, p, q, n_cu, s)
                  : launch_sweep<StreamFP64, 2>(P, p, q, n_cu, s);
    case 4:
      return fp32 ? launch_sweep<StreamFP32C, 4>(P, p, q, n_cu, s)
                  : launch_sweep<StreamFP64, 4>(P, p, q, n_cu, s);
    case 8:
      return fp32 ? launch_sweep<StreamFP32C, 8>(P, p, q, n_cu, s)
                  : launch_sweep<StreamFP64, 8>(P, p, q, n_cu, s);
    default: return hipErrorInvalidValue;
  }
}

// k-lane organisation (fast forms): tables then kernel, tiles of 64 positions.
hipError_t los_launch_klane_prep(int fp32, const los_params *P, hipStream_t s)
{
  return fp32 ? launch_klane_prep<float, true>(P, s) : launch_klane_prep<double, true>(P, s);
}

hipError_t los_launch_klane(int fp32, const los_params *P, int nlb, hipStream_t s)
{
  if (P->n_pos == 0 || P->n_lm_max == 0) return hipSuccess;
  const dim3 grid = klane_grid(P);
  if (grid.x == 0) return hipSuccess;
  const int NL = (P->LmaxT > P->LmaxP ? P->LmaxT : P->LmaxP) + 1;
  const int key = NL * 8 + nlb;
  switch (key) {
#define LOS_KL1(n, b)                                                                    \
  case n * 8 + b:                                                                        \
    if (fp32)                                                                            \
      hipLaunchKernelGGL((los_klane<StreamFP32C, n, b>), grid, dim3(64), 0, s, *P);      \
    else                                                                                 \
      hipLaunchKernelGGL((los_klane<StreamFP64, n, b>), grid, dim3(64), 0, s, *P);       \
    break;
#define LOS_KL(n) LOS_KL1(n, 1) LOS_KL1(n, 2) LOS_KL1(n, 4)
    LOS_KL(1) LOS_KL(2) LOS_KL(3) LOS_KL(4) LOS_KL(5) LOS_KL(6) LOS_KL(7) LOS_KL(8)
#undef LOS_KL
#undef LOS_KL1
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Candidate shapes: the stream shapes (both dispatch orders) that fit the LDS of
// one workgroup, then the sweep shapes.
int los_shapes(int fp32, int tp2_size, los_shape *out, int cap)
{
  int n = 0;
  const size_t real = fp32 ? sizeof(float) : sizeof(double);
#define LOS_LIST(NK_, CH_)                                                        \
  if ((size_t)NK_ * ((size_t)tp2_size * CH_ + CH_) * real + 8 * 1024 <= 160 * 1024) \
    for (int p = 0; p < 2 && n < cap; ++p) out[n++] = los_shape{0, NK_, CH_, p};
  LOS_SHAPES(LOS_LIST)
#undef LOS_LIST
  // sweep: ch = 1 reads the FP32 pair table (FP64 always reads its pair table)
  for (int nk = 1; nk <= 8; nk *= 2)
    for (int c = 0; c <= fp32; ++c)
      for (int p = 0; p < 2 && n < cap; ++p) out[n++] = los_shape{1, nk, c, p};
  for (int b = 1; b <= 4 && n < cap; b *= 2) out[n++] = los_shape{2, 64, b, 0};
  return n;
}

}  // extern "C"
