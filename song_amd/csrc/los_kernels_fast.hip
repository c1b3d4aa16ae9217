// los_kernels_fast.hip — the LOS integral in its FP64-with-FMA and
// FP32-compensated forms (los_stream.h), plus the FP32 copy of the J table.
#include <hip/hip_runtime.h>

#include "los_stream.h"

#ifndef LOS_NK_FP64
#define LOS_NK_FP64 8
#endif
#ifndef LOS_NK_FP32
#define LOS_NK_FP32 8
#endif

__global__ void __launch_bounds__(256) los_round_j32(const double *src, float *dst, long n)
{
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

extern "C" {

hipError_t los_launch_round_j32(const double *src, float *dst, long n, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(los_round_j32, dim3(2048), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t los_launch_stream_fp64(const los_params *P, hipStream_t s)
{
  return launch_stream<StreamFP64, LOS_NK_FP64>(P, s);
}

hipError_t los_launch_stream_fp32c(const los_params *P, hipStream_t s)
{
  return launch_stream<StreamFP32C, LOS_NK_FP32>(P, s);
}

}  // extern "C"
