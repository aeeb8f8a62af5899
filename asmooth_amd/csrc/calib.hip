// A plain streaming copy of known size.  Hardware byte counters
// (FETCH_SIZE / WRITE_SIZE) read for it calibrate the counters read for the
// smoothing kernels: its algorithmic traffic is exactly `bytes` in and out.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/asmooth.h"

namespace {

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

extern "C" int asm_calibration_copy(void* dst, const void* src, int64_t bytes)
{
    if (!dst || !src || bytes <= 0 || bytes % 16 != 0) return ASM_ERR_ARG;
    const int64_t n = bytes / 16;
    copy_kernel<<<2048, 256, 0, nullptr>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
    if (hipGetLastError() != hipSuccess) return ASM_ERR_HIP;
    return hipDeviceSynchronize() == hipSuccess ? ASM_OK : ASM_ERR_HIP;
}

extern "C" int asm_device_alloc(void** p, int64_t bytes)
{
    if (!p || bytes <= 0) return ASM_ERR_ARG;
    return hipMalloc(p, static_cast<size_t>(bytes)) == hipSuccess ? ASM_OK : ASM_ERR_HIP;
}

extern "C" int asm_device_free(void* p)
{
    return hipFree(p) == hipSuccess ? ASM_OK : ASM_ERR_HIP;
}

extern "C" int asm_device_to_host(void* host, const void* dev, int64_t bytes)
{
    if (!host || !dev || bytes < 0) return ASM_ERR_ARG;
    return hipMemcpy(host, dev, static_cast<size_t>(bytes), hipMemcpyDeviceToHost) == hipSuccess ? ASM_OK
                                                                                                  : ASM_ERR_HIP;
}

extern "C" int asm_host_to_device(void* dev, const void* host, int64_t bytes)
{
    if (!host || !dev || bytes < 0) return ASM_ERR_ARG;
    return hipMemcpy(dev, host, static_cast<size_t>(bytes), hipMemcpyHostToDevice) == hipSuccess ? ASM_OK
                                                                                                  : ASM_ERR_HIP;
}
