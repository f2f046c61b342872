// s3d_elem.hip -- element-kernel launchers for one GLL order, S3D_ELEM_P.
// Compiled once per order (2..7) so that the orders' instantiations of
// elem_kernel / elem_kernel_batch build in parallel. A launch description
// (s3d_launch.h) selects the instantiation: element tile, memory words per
// mechanism, stress compensation, kernel variant, BF16 statics. The choices
// are template parameters so that no run-time branch inside the kernel
// leaves register joins that wait for every load in flight.
#include "s3d_elem.h"
#include "s3d_launch.h"

#include <type_traits>

#ifndef S3D_ELEM_P
#error "compile with -DS3D_ELEM_P=<GLL order 2..7>"
#endif

namespace s3d {
namespace {

template <int P, int E>
unsigned tiles(long nel) { return (unsigned)((nel + Shape<P, E>::EPB - 1) / Shape<P, E>::EPB); }

// stress compensation (FP32 only)
template <int P, int E, typename R, int VM, bool S16, int LL>
void launch_ks(int ks, const ElemArgs<R> &ea, hipStream_t st, unsigned ge)
{
    constexpr int NT = Shape<P, E>::NT;
    if (ks == 1)
        hipLaunchKernelGGL((elem_kernel<P, E, R, VM, S16, 1, LL>), dim3(ge), dim3(NT), 0, st, ea);
    else if (ks == 2)
        hipLaunchKernelGGL((elem_kernel<P, E, R, VM, S16, 2, LL>), dim3(ge), dim3(NT), 0, st, ea);
    else
        hipLaunchKernelGGL((elem_kernel<P, E, R, VM, S16, 0, LL>), dim3(ge), dim3(NT), 0, st, ea);
}

// kernel variant and stress compensation: FP32 only; FP64 and the
// BF16-static kernels use variant 0 without compensation words
template <int P, int E, typename R, int VM, bool S16>
void launch_ll(int ks, int ll, const ElemArgs<R> &ea, hipStream_t st, unsigned ge)
{
    if constexpr (std::is_same<R, float>::value) {
        if constexpr (!S16) {
            if (ll == 1) return launch_ks<P, E, R, VM, S16, 1>(ks, ea, st, ge);
            if (ll == 2) return launch_ks<P, E, R, VM, S16, 2>(ks, ea, st, ge);
        }
        launch_ks<P, E, R, VM, S16, 0>(ks, ea, st, ge);
    } else {
        hipLaunchKernelGGL((elem_kernel<P, E, R, VM, S16, 0, 0>), dim3(ge), dim3(Shape<P, E>::NT), 0, st, ea);
    }
}

template <int P, int E, typename R>
void launch_vm(const ElemLaunch &l, const ElemArgs<R> &ea, hipStream_t st)
{
    const unsigned ge = tiles<P, E>(l.nel);
    if constexpr (std::is_same<R, float>::value && E == 0) {
        if (l.bf16_static) {
            if (l.vm == 6) return launch_ll<P, 0, float, 6, true>(l.ks, 0, ea, st, ge);
            if (l.vm == 5) return launch_ll<P, 0, float, 5, true>(l.ks, 0, ea, st, ge);
            return launch_ll<P, 0, float, 0, true>(l.ks, 0, ea, st, ge);
        }
    }
    if (l.vm == 6) return launch_ll<P, E, R, 6, false>(l.ks, l.ll, ea, st, ge);
    if (l.vm == 5) return launch_ll<P, E, R, 5, false>(l.ks, l.ll, ea, st, ge);
    launch_ll<P, E, R, 0, false>(l.ks, l.ll, ea, st, ge);
}

template <int P, int E, typename R, int VM, int KS>
void batch_ll(const ElemBatchLaunch &l, hipStream_t st)
{
    constexpr int NT = Shape<P, E>::NT;
    const ElemArgs<R> *args = (const ElemArgs<R> *)l.args;
    if constexpr (std::is_same<R, float>::value) {
        if (l.ll == 1) {
            hipLaunchKernelGGL((elem_kernel_batch<P, E, R, VM, KS, 1>), dim3(l.nwg), dim3(NT), 0, st, args, l.own);
            return;
        }
        if (l.ll == 2) {
            hipLaunchKernelGGL((elem_kernel_batch<P, E, R, VM, KS, 2>), dim3(l.nwg), dim3(NT), 0, st, args, l.own);
            return;
        }
    }
    hipLaunchKernelGGL((elem_kernel_batch<P, E, R, VM, KS, 0>), dim3(l.nwg), dim3(NT), 0, st, args, l.own);
}

template <int P, int E, typename R, int VM>
void batch_ks(const ElemBatchLaunch &l, hipStream_t st)
{
    if constexpr (std::is_same<R, float>::value) {
        if (l.ks == 1) return batch_ll<P, E, R, VM, 1>(l, st);
        if (l.ks == 2) return batch_ll<P, E, R, VM, 2>(l, st);
    }
    batch_ll<P, E, R, VM, 0>(l, st);
}

template <int P, int E, typename R>
void batch_vm(const ElemBatchLaunch &l, hipStream_t st)
{
    if (l.vm == 6) return batch_ks<P, E, R, 6>(l, st);
    if (l.vm == 5) return batch_ks<P, E, R, 5>(l, st);
    batch_ks<P, E, R, 0>(l, st);
}

}  // namespace

// Element tiles: 1, 2 and 4 elements per workgroup where they fit 1024
// lanes, besides the default. The BF16-static kernel exists for the
// default tile only.
template <int P, typename R>
hipError_t launch_elem_kernel(const ElemLaunch &l, const ElemArgs<R> &ea, hipStream_t st)
{
    const int epb = (std::is_same<R, float>::value && l.bf16_static) ? 0 : l.epb;
    bool done = false;
    if constexpr (extra_tile<P, 1>())
        if (epb == 1) { launch_vm<P, 1, R>(l, ea, st); done = true; }
    if constexpr (extra_tile<P, 2>())
        if (epb == 2) { launch_vm<P, 2, R>(l, ea, st); done = true; }
    if constexpr (extra_tile<P, 4>())
        if (epb == 4) { launch_vm<P, 4, R>(l, ea, st); done = true; }
    if (!done) launch_vm<P, 0, R>(l, ea, st);
    return hipGetLastError();
}

// Batches of FP64 members use the default tile; FP32 groups take their
// largest member's.
template <int P, typename R>
hipError_t launch_elem_batch(const ElemBatchLaunch &l, hipStream_t st)
{
    bool done = false;
    if constexpr (std::is_same<R, float>::value) {
        if constexpr (extra_tile<P, 1>())
            if (l.epb == 1) { batch_vm<P, 1, R>(l, st); done = true; }
        if constexpr (extra_tile<P, 2>())
            if (l.epb == 2) { batch_vm<P, 2, R>(l, st); done = true; }
        if constexpr (extra_tile<P, 4>())
            if (l.epb == 4) { batch_vm<P, 4, R>(l, st); done = true; }
    }
    if (!done) batch_vm<P, 0, R>(l, st);
    return hipGetLastError();
}

template hipError_t launch_elem_kernel<S3D_ELEM_P, float>(const ElemLaunch &, const ElemArgs<float> &, hipStream_t);
template hipError_t launch_elem_kernel<S3D_ELEM_P, double>(const ElemLaunch &, const ElemArgs<double> &, hipStream_t);
template hipError_t launch_elem_batch<S3D_ELEM_P, float>(const ElemBatchLaunch &, hipStream_t);
template hipError_t launch_elem_batch<S3D_ELEM_P, double>(const ElemBatchLaunch &, hipStream_t);

}  // namespace s3d
