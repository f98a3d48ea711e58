// Measurement hook: a plain device copy of a known size, whose hardware
// counter readings (FETCH_SIZE, WRITE_SIZE under rocprofv3 --pmc) scale the
// same counters of the engine's kernels to bytes.
#include <string>
#include "engine.h"

namespace mdc {

// Each thread copies 16-byte words with a grid stride: every byte of src
// is read once and every byte of dst written once.
__global__ void __launch_bounds__(256) k_copy_calibration(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          size_t words) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < words; k += stride) dst[k] = src[k];
}


// reps timed launches after one untimed launch; buffers freed on every path
void copy_calibration(int device, size_t bytes, int reps, double* us_per_copy) {
    if (bytes == 0 || bytes % 16 != 0 || reps < 1)
        throw Error(E_INVAL, "copy calibration needs a positive multiple of 16 bytes and reps >= 1");
    HIPCHECK(hipSetDevice(device));
    struct Bufs {
        void *a = nullptr, *b = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Bufs() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
        }
    } B;
    if (hipMalloc(&B.a, bytes) != hipSuccess || hipMalloc(&B.b, bytes) != hipSuccess)
        throw Error(E_NOMEM, "out of device memory for the copy calibration");
    HIPCHECK(hipMemset(B.a, 0, bytes));
    HIPCHECK(hipEventCreate(&B.e0));
    HIPCHECK(hipEventCreate(&B.e1));
    const size_t words = bytes / 16;
    const unsigned grid = 4096;
    hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, 0, (const float4*)B.a, (float4*)B.b, words);
    HIPCHECK(hipEventRecord(B.e0, 0));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, 0, (const float4*)B.a, (float4*)B.b, words);
    HIPCHECK(hipEventRecord(B.e1, 0));
    HIPCHECK(hipEventSynchronize(B.e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, B.e0, B.e1));
    if (us_per_copy) *us_per_copy = 1e3 * (double)ms / reps;
}

}  // namespace mdc
