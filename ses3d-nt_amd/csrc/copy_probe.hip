// copy_probe.hip -- plain device streams of a known size, for calibrating the
// rocprofv3 FETCH_SIZE / WRITE_SIZE counters against HBM bytes and for
// measuring the attainable HBM bandwidth of the device.
//
//   copy_probe <GB> [repeats] [mode]
//
// mode (default "copy"):
//   copy     float4 grid-stride copy: reads and writes GB gigabytes
//   copy_nt  the same with non-temporal loads and stores
//   copy_1d  one float4 per lane, one pass (no grid-stride loop)
//   read     float4 grid-stride sum of GB gigabytes (one float per workgroup
//            written)
//   write    float4 grid-stride fill of GB gigabytes
// Each launch's bandwidth is measured with HIP events and printed.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>

__global__ void __launch_bounds__(256) copy_kernel(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

typedef float floatv4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) copy_nt_kernel(const float4 *__restrict__ src,
                                                      float4 *__restrict__ dst, long n)
{
    const floatv4 *s = (const floatv4 *)src;
    floatv4 *d = (floatv4 *)dst;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(s + t), d + t);
}

__global__ void __launch_bounds__(256) copy_1d_kernel(const float4 *__restrict__ src,
                                                      float4 *__restrict__ dst, long n)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

__global__ void __launch_bounds__(256) read_kernel(const float4 *__restrict__ src, float *out, long n)
{
    float s = 0.f;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const float4 v = src[t];
        s += (v.x + v.y) + (v.z + v.w);
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) write_kernel(float4 *__restrict__ dst, long n, float v)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = make_float4(v, v, v, v);
}

int main(int argc, char **argv)
{
    double gb = argc > 1 ? atof(argv[1]) : 4.0;
    int reps = argc > 2 ? atoi(argv[2]) : 3;
    const char *mode = argc > 3 ? argv[3] : "copy";
    long n = (long)(gb * 1e9 / sizeof(float4));
    const unsigned grid = 256 * 8 * 16;
    float4 *a = nullptr, *b = nullptr;
    float *part = nullptr;
    if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess || hipMalloc(&b, n * sizeof(float4)) != hipSuccess ||
        hipMalloc(&part, grid * sizeof(float)) != hipSuccess) {
        fprintf(stderr, "copy_probe: allocation failed\n");
        return 1;
    }
    (void)hipMemset(a, 0, n * sizeof(float4));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int r = 0; r < reps; ++r) {
        double bytes = 2.0 * n * sizeof(float4);
        (void)hipEventRecord(e0, 0);
        if (!strcmp(mode, "copy_nt")) {
            hipLaunchKernelGGL(copy_nt_kernel, dim3(grid), dim3(256), 0, 0, a, b, n);
        } else if (!strcmp(mode, "copy_1d")) {
            hipLaunchKernelGGL(copy_1d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, b, n);
        } else if (!strcmp(mode, "read")) {
            hipLaunchKernelGGL(read_kernel, dim3(grid), dim3(256), 0, 0, a, part, n);
            bytes = 1.0 * n * sizeof(float4);
        } else if (!strcmp(mode, "write")) {
            hipLaunchKernelGGL(write_kernel, dim3(grid), dim3(256), 0, 0, b, n, 1.0f);
            bytes = 1.0 * n * sizeof(float4);
        } else {
            hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, 0, a, b, n);
        }
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        printf("copy_probe %s: %.0f bytes moved in %.3f ms: %.1f GB/s\n", mode, bytes, ms,
               bytes / (ms * 1e-3) / 1e9);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(part);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
