// TEST INFRASTRUCTURE: the renderer's math functions evaluated on the GPU
// from flat input records (rdeval.h), for comparison with the host build.
#include <hip/hip_runtime.h>

#include "../../psychopath_amd/csrc/render_device.h"
#include "rdeval.h"

__global__ void rdgpu_kernel(int which, const float* in, float* out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o[RDEVAL_OUT_STRIDE];
    rdeval::eval(which, in + i * RDEVAL_IN_STRIDE, o);
    for (int k = 0; k < RDEVAL_OUT_STRIDE; k++) out[i * RDEVAL_OUT_STRIDE + k] = o[k];
}

extern "C" int rdgpu_eval(int which, const float* in_host, float* out_host, int64_t n) {
    if (n <= 0 || n > (1 << 24) || which < 0 || which > 8) return -1;
    float *din = nullptr, *dout = nullptr;
    const size_t ib = (size_t)n * RDEVAL_IN_STRIDE * sizeof(float), ob = (size_t)n * RDEVAL_OUT_STRIDE * sizeof(float);
    if (hipMalloc(&din, ib) != hipSuccess) return -2;
    if (hipMalloc(&dout, ob) != hipSuccess) {
        (void)hipFree(din);
        return -2;
    }
    int rc = 0;
    if (hipMemcpy(din, in_host, ib, hipMemcpyHostToDevice) != hipSuccess) rc = -3;
    if (!rc) {
        hipLaunchKernelGGL(rdgpu_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, which, din, dout, n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -4;
    }
    if (!rc && hipMemcpy(out_host, dout, ob, hipMemcpyDeviceToHost) != hipSuccess) rc = -5;
    (void)hipFree(din);
    (void)hipFree(dout);
    return rc;
}
