// s3d_launch.h -- element-kernel launchers. They are defined in s3d_elem.hip,
// which is compiled once per GLL order (2..7) so that the orders' kernel
// instantiations build in parallel, and called by the library.
#pragma once

#include "s3d_elem.h"

namespace s3d {

// A tile of e elements per workgroup fits the 1024 lanes of a workgroup.
template <int P>
constexpr bool tile_ok(int e) { return e * (P + 1) * (P + 1) * (P + 1) <= 1024; }

// A tile of E elements per workgroup other than the default (which is the
// E = 0 instantiation).
template <int P, int E>
constexpr bool extra_tile() { return tile_ok<P>(E) && default_epb<P>() != E; }

// What selects the element-kernel instantiation of one launch.
struct ElemLaunch {
    long nel;          // elements
    int vm;            // memory words per relaxation mechanism: 0 (elastic), 6 or 5
    int ks;            // stress compensation: 0 off, 1 FP32 words, 2 BF16 words
    int ll;            // kernel variant of the FP32 kernels: 0 planar LDS staging,
                       // 1 node-interleaved staging, 2 planar with early state stores
    int bf16_static;   // FP32 moduli and tau stored in BF16 (default tile, variant 0)
    int epb;           // elements per workgroup, 0 the default tile
};

// One batched element launch: a group's argument blocks for the current
// phase (device memory) and its workgroup -> (member, member's workgroup) map.
struct ElemBatchLaunch {
    int vm, ks, ll, epb;   // as in ElemLaunch, shared by the group's members
    unsigned nwg;          // workgroups of the launch
    const void *args;      // ElemArgs<R>[members]
    const int2 *own;       // [nwg]
};

// Both return the launch status (hipGetLastError after the launch).
template <int P, typename R>
hipError_t launch_elem_kernel(const ElemLaunch &l, const ElemArgs<R> &ea, hipStream_t st);

template <int P, typename R>
hipError_t launch_elem_batch(const ElemBatchLaunch &l, hipStream_t st);

}  // namespace s3d
