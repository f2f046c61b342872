// copy_probe.hip -- plain device copy of a known size, for calibrating the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters against HBM bytes and for
// measuring the attainable copy bandwidth on the box.
//
//   copy_probe <GB> [repeats]
//
// Reads and writes exactly GB gigabytes per launch (float4 grid-stride copy)
// and prints the bandwidth of each launch measured with HIP events.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

__global__ void __launch_bounds__(256) copy_kernel(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

int main(int argc, char **argv)
{
    double gb = argc > 1 ? atof(argv[1]) : 4.0;
    int reps = argc > 2 ? atoi(argv[2]) : 3;
    long n = (long)(gb * 1e9 / sizeof(float4));
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
        fprintf(stderr, "copy_probe: allocation failed\n");
        return 1;
    }
    (void)hipMemset(a, 0, n * sizeof(float4));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(copy_kernel, dim3(256 * 8 * 16), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        printf("copy_probe %ld bytes read + %ld written in %.3f ms: %.1f GB/s (read+write)\n",
               n * (long)sizeof(float4), n * (long)sizeof(float4), ms,
               2.0 * n * sizeof(float4) / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
