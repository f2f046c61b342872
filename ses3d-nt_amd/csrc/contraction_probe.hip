// contraction_probe.hip -- the GLL derivative contraction on its own, on the
// VALU and on the matrix cores, for deciding where the element kernel's
// contraction belongs.
//
//   contraction_probe layout
//       checks the operand and result lane layout of the FP32 matrix product
//       v_mfma_f32_16x16x4f32 that the mfma32 mode relies on, against a host
//       product of the same operands (prints OK or the first mismatch)
//   contraction_probe <mode> [columns] [repeats] [workgroups]
//       out(i, col) = sum_q D(i, q) V(q, col), i, q = 0..4 (GLL order 4), for
//       `columns` columns staged in LDS per workgroup (default 1200: one
//       element's 25 lines x 3 velocity components x 16 sources), repeated
//       `repeats` times, in `workgroups` workgroups of 256 lanes; prints the
//       kernel time (HIP events) and the useful rate 2*5*5 flop per column.
//   modes:
//     valu64  FP64 coefficients, FP64 sums (the element kernel's form)
//     valu32  FP32 coefficients, FP32 sums
//     mfma32  FP32 matrix cores, 16 columns per pair of 16x16x4 products
//             (rows 5..15 and k 5..7 padded with zeros), the coefficients
//             split into FP32 high and low parts (two pairs of products) so
//             that no coefficient is rounded to FP32
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define CHK(x)                                                                     \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
            exit(1);                                                               \
        }                                                                          \
    } while (0)

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int P1 = 5;          // GLL order 4
constexpr int MAXCOL = 2048;   // columns per workgroup held in LDS

struct Coef {
    double d[P1 * P1];   // D(i, q) at d[i * P1 + q]
};

// --------------------------------------------------------------------------
// lane layout check: one wave, A(16x4), B(4x16) given per lane, D(16x16)
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(64) layout_kernel(const float *A, const float *B, float *D)
{
    const int l = threadIdx.x;
    // operands as documented for v_mfma_f32_16x16x4f32: lane l holds
    // A(l % 16, l / 16) and B(l / 16, l % 16); result element
    // D(4 * (l / 16) + r, l % 16) in component r
    const float a = A[(l % 16) * 4 + l / 16];
    const float b = B[(l / 16) * 16 + l % 16];
    floatx4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    for (int r = 0; r < 4; ++r) D[(4 * (l / 16) + r) * 16 + l % 16] = c[r];
}

static int layout_check()
{
    float hA[64], hB[64], hD[256], ref[256];
    for (int q = 0; q < 64; ++q) {
        hA[q] = (float)((q * 7) % 13 - 6);
        hB[q] = (float)((q * 5) % 11 - 5);
    }
    for (int i = 0; i < 16; ++i)
        for (int j = 0; j < 16; ++j) {
            float s = 0.f;
            for (int k = 0; k < 4; ++k) s += hA[i * 4 + k] * hB[k * 16 + j];
            ref[i * 16 + j] = s;
        }
    float *A, *B, *D;
    CHK(hipMalloc(&A, sizeof hA));
    CHK(hipMalloc(&B, sizeof hB));
    CHK(hipMalloc(&D, sizeof hD));
    CHK(hipMemcpy(A, hA, sizeof hA, hipMemcpyHostToDevice));
    CHK(hipMemcpy(B, hB, sizeof hB, hipMemcpyHostToDevice));
    CHK(hipMemset(D, 0, sizeof hD));
    hipLaunchKernelGGL(layout_kernel, dim3(1), dim3(64), 0, 0, A, B, D);
    CHK(hipGetLastError());
    CHK(hipMemcpy(hD, D, sizeof hD, hipMemcpyDeviceToHost));
    for (int q = 0; q < 256; ++q)
        if (hD[q] != ref[q]) {
            printf("layout MISMATCH at D(%d,%d): %g, host %g\n", q / 16, q % 16, hD[q], ref[q]);
            return 1;
        }
    printf("layout OK: v_mfma_f32_16x16x4f32 lane layout as assumed\n");
    CHK(hipFree(A));
    CHK(hipFree(B));
    CHK(hipFree(D));
    return 0;
}

// --------------------------------------------------------------------------
// contraction kernels: V(q, col) in LDS, repeated over shifted column sets
// --------------------------------------------------------------------------
__device__ inline void stage(float (*V)[MAXCOL], const float *src, int ncol)
{
    for (int t = threadIdx.x; t < P1 * ncol; t += blockDim.x) V[t / ncol][t % ncol] = src[t];
    __syncthreads();
}

__global__ void __launch_bounds__(256) valu64_kernel(const float *src, Coef D, int ncol, int reps, double *out)
{
    __shared__ float V[P1][MAXCOL];
    stage(V, src, ncol);
    double sum = 0.0;
    for (int r = 0; r < reps; ++r)
        for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
            const int cc = (col + r) % ncol;
            double v[P1];
#pragma unroll
            for (int q = 0; q < P1; ++q) v[q] = (double)V[q][cc];
#pragma unroll
            for (int i = 0; i < P1; ++i) {
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < P1; ++q) acc += D.d[i * P1 + q] * v[q];
                sum += acc;
            }
        }
    out[blockIdx.x * blockDim.x + threadIdx.x] = sum;
}

__global__ void __launch_bounds__(256) valu32_kernel(const float *src, Coef D, int ncol, int reps, double *out)
{
    __shared__ float V[P1][MAXCOL];
    stage(V, src, ncol);
    float d[P1 * P1];
#pragma unroll
    for (int q = 0; q < P1 * P1; ++q) d[q] = (float)D.d[q];
    float sum = 0.f;
    for (int r = 0; r < reps; ++r)
        for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
            const int cc = (col + r) % ncol;
            float v[P1];
#pragma unroll
            for (int q = 0; q < P1; ++q) v[q] = V[q][cc];
#pragma unroll
            for (int i = 0; i < P1; ++i) {
                float acc = 0.f;
#pragma unroll
                for (int q = 0; q < P1; ++q) acc += d[i * P1 + q] * v[q];
                sum += acc;
            }
        }
    out[blockIdx.x * blockDim.x + threadIdx.x] = (double)sum;
}

__global__ void __launch_bounds__(256) mfma32_kernel(const float *src, Coef D, int ncol, int reps, double *out)
{
    __shared__ float V[P1][MAXCOL];
    stage(V, src, ncol);
    const int l = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    const int row = l % 16, kq = l / 16;
    // A operands: D(row, k) for k = kq (first product) and k = 4 + kq
    // (second; only k = 4 exists), high and low FP32 parts, zero outside 5x5
    auto dpart = [&](int i, int q, bool lo) -> float {
        if (i >= P1 || q >= P1) return 0.f;
        const double x = D.d[i * P1 + q];
        const float hi = (float)x;
        return lo ? (float)(x - (double)hi) : hi;
    };
    const float a0h = dpart(row, kq, false), a1h = dpart(row, 4 + kq, false);
    const float a0l = dpart(row, kq, true), a1l = dpart(row, 4 + kq, true);
    const int nblk = ncol / 16;
    floatx4 sum = {0.f, 0.f, 0.f, 0.f};
    for (int r = 0; r < reps; ++r)
        for (int b = wave; b < nblk; b += nwave) {
            const int cc = (b * 16 + l % 16 + r) % ncol;
            const float b0 = V[kq][cc];
            const float b1 = kq == 0 ? V[4][cc] : 0.f;
            floatx4 c = {0.f, 0.f, 0.f, 0.f};
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0h, b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1h, b1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a0l, b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(a1l, b1, c, 0, 0, 0);
            sum += c;
        }
    out[blockIdx.x * blockDim.x + threadIdx.x] = (double)(sum[0] + sum[1] + sum[2] + sum[3]);
}

int main(int argc, char **argv)
{
    const char *mode = argc > 1 ? argv[1] : "valu64";
    if (!strcmp(mode, "layout")) return layout_check();
    const int ncol = argc > 2 ? atoi(argv[2]) : 1200;
    const int reps = argc > 3 ? atoi(argv[3]) : 200;
    const int nwg = argc > 4 ? atoi(argv[4]) : 2048;
    if (ncol < 16 || ncol > MAXCOL || ncol % 16) {
        fprintf(stderr, "columns must be a multiple of 16 in 16..%d\n", MAXCOL);
        return 2;
    }
    if (reps < 1 || nwg < 1) {
        fprintf(stderr, "repeats and workgroups must be positive\n");
        return 2;
    }
    // GLL order 4 derivative matrix entries (any fixed 5x5 FP64 matrix serves
    // for timing)
    Coef D;
    for (int i = 0; i < P1; ++i)
        for (int q = 0; q < P1; ++q) D.d[i * P1 + q] = std::cos(0.37 * i + 1.3 * q) / 3.0;
    std::vector<float> h((size_t)P1 * ncol);
    for (size_t q = 0; q < h.size(); ++q) h[q] = (float)std::sin(0.01 * (double)q);
    float *src;
    double *out;
    CHK(hipMalloc(&src, sizeof(float) * h.size()));
    CHK(hipMalloc(&out, sizeof(double) * nwg * 256));
    CHK(hipMemcpy(src, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int trial = 0; trial < 4; ++trial) {
        CHK(hipEventRecord(e0, 0));
        if (!strcmp(mode, "valu64"))
            hipLaunchKernelGGL(valu64_kernel, dim3(nwg), dim3(256), 0, 0, src, D, ncol, reps, out);
        else if (!strcmp(mode, "valu32"))
            hipLaunchKernelGGL(valu32_kernel, dim3(nwg), dim3(256), 0, 0, src, D, ncol, reps, out);
        else if (!strcmp(mode, "mfma32"))
            hipLaunchKernelGGL(mfma32_kernel, dim3(nwg), dim3(256), 0, 0, src, D, ncol, reps, out);
        else {
            fprintf(stderr, "unknown mode %s\n", mode);
            return 2;
        }
        CHK(hipGetLastError());
        CHK(hipEventRecord(e1, 0));
        CHK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        if (trial > 0 && ms < best) best = ms;   // the first launch warms up
    }
    const double cols = (double)ncol * reps * nwg;
    printf("{\"mode\": \"%s\", \"columns_per_workgroup\": %d, \"repeats\": %d, \"workgroups\": %d, "
           "\"ms\": %.4f, \"columns_per_s\": %.4e, \"useful_tflops\": %.3f}\n",
           mode, ncol, reps, nwg, best, cols / (best * 1e-3), 50.0 * cols / (best * 1e-3) / 1e12);
    CHK(hipFree(src));
    CHK(hipFree(out));
    return 0;
}
