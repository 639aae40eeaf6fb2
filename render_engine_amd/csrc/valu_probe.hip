// VALU issue-rate probe for the rooflines of the traversal kernels: independent FP32 multiply /
// add / max chains with no memory traffic in the loop (20 VALU wave-instructions per iteration,
// no contraction), 8192 blocks of 256 threads; prints the rate from HIP events. Run it under
// rocprofv3 --pmc SQ_INSTS_VALU to read the counter's own count for the same dispatches.
#include <hip/hip_runtime.h>
#include <cstdio>

__global__ void __launch_bounds__(256) k_valu_probe(float* out, int iters, float s) {
    float a0 = threadIdx.x * 1e-3f, a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f, a4 = a0 + 4.f, a5 = a0 + 5.f, a6 = a0 + 6.f,
          a7 = a0 + 7.f;
    for (int i = 0; i < iters; ++i) {
        a0 = a0 * s + 0.5f; a1 = a1 * s + 0.5f; a2 = a2 * s + 0.5f; a3 = a3 * s + 0.5f;
        a4 = a4 * s + 0.5f; a5 = a5 * s + 0.5f; a6 = a6 * s + 0.5f; a7 = a7 * s + 0.5f;
        a0 = fmaxf(a0, a4); a1 = fmaxf(a1, a5); a2 = fminf(a2, a6); a3 = fminf(a3, a7);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

int main() {
    const int blocks = 256 * 32, threads = 256, iters = 20000;
    float* d = nullptr;
    if (hipMalloc((void**)&d, (size_t)blocks * threads * sizeof(float)) != hipSuccess) { std::printf("hipMalloc failed\n"); return 1; }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    for (int r = 0; r < 3; ++r) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(threads), 0, 0, d, iters, 0.999f);
        (void)hipEventRecord(e1);
        if (hipEventSynchronize(e1) != hipSuccess) { std::printf("launch failed\n"); return 1; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double winst = (double)blocks * (threads / 64) * iters * 20.0;
        std::printf("run %d: %.3f ms, %.1f G VALU wave-instructions/s\n", r, ms, winst / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(d);
    return 0;
}
