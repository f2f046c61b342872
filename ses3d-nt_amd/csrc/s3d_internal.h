// s3d_internal.h -- device data layout and kernel argument blocks.
//
// Layout in HBM (one simulation):
//   element-local fields, e = ex + nx*(ey + ny*ez), n = i + P1*(j + P1*k):
//     block fields, element-major f[e][c][n]: the state block (stress, its
//       compensation and the memory variables in one block, see
//       state_components) and the moduli lam/mu/tau; the element force in
//       two element-major parts, the k < p nodes and the k = p layers
//       (force_index).
//       One element's components are one
//       contiguous run, so the workgroup of a tile of consecutive elements
//       reads and writes one contiguous region per block field instead of one
//       per component (fewer concurrent DRAM streams: measured 8 % faster
//       than f[c][e][n]);
//     density and anisotropy coefficients, component-major f[c][e][n].
//     Either way the (p+1)^3 nodes of one element are contiguous and a
//     workgroup's loads/stores of an element are fully coalesced.
//   unique-node fields (velocity, its compensation, 1/M):
//       g[b][o], b = bx + (nx+1)*(by + (ny+1)*bz), o = li + p*(lj + p*lk)
//     element block b "owns" the p^3 nodes with local indices 0..p-1; the
//     padding blocks bx = nx (by = ny, bz = nz) hold the last node plane.
//     Global node (gx, gy, gz) = (bx*p + li, by*p + lj, bz*p + lk).
//   1-D tables: per (element, GLL index) along each axis, FP64.
//
// Precision split of the FP32 step: the state (velocity, stress, memory
// variables), the material and the element force are stored and updated in
// FP32, but every operator coefficient that is the same for many nodes (the
// GLL derivative matrix and weights, the 2/dx Jacobian factors, the
// spherical tables) is held in FP64 and applied in FP64 registers. An FP32
// rounding of such a shared coefficient is a systematic error (a wave-speed or
// time-axis scaling that grows into a phase error with simulated time);
// rounding data-dependent results is not. The time step is folded into the
// per-node coefficients (dt*lambda, dt*mu, dt*A/B/C, dt/M) in FP64 before
// their one FP32 rounding, for the same reason.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace s3d {

constexpr int MAXP1 = 8;       // p <= 7
constexpr int NRDISS = 3;      // relaxation mechanisms

// Components of the element-major state block: stress (6), its compensation
// (6 FP32 words for kahan_sigma 1, 3 packed BF16 pairs for 2) and the memory
// variables of a visco-elastic model, in that order. vm: memory words per
// relaxation mechanism, 0 (elastic), 6 (xx xy xz yy yz zz, the reference's
// form) or 5 (the deviatoric part, xx xy xz yy yz; zz = -(xx + yy)).
constexpr int state_components(int kahan_sigma, int vm)
{
    return 6 + (kahan_sigma == 1 ? 6 : kahan_sigma == 2 ? 3 : 0) + NRDISS * vm;
}

// strain-rate / stress component order used everywhere: xx xy xz yy yz zz
enum { XX = 0, XY, XZ, YY, YZ, ZZ };

// Precision of the operator coefficients and of the sums they enter.
template <typename R> struct CoefType { using T = double; };
template <typename R> using coef_t = typename CoefType<R>::T;

template <typename C>
struct ThTab {                 // per (ex, i): colatitude dependent factors
    C sin_t, inv_sin, cot, pad;
};
template <typename C>
struct RTab {                  // per (ez, k): radius dependent factors
    C r, inv_r, r2, pad;
};

using bf16 = __bf16;
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// two floats -> one dword of BF16 (round to nearest even), a in the low half
__device__ inline unsigned pack_bf16(float a, float b)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector((floatx2){a, b}, bf16x2));
}
__device__ inline float bf16_lo(unsigned w) { return __uint_as_float(w << 16); }
__device__ inline float bf16_hi(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

// Device code sees every pointer of the kernel-argument structs below as a
// pointer to global memory. Batched launches read these structs from device
// memory, and a pointer loaded from there would otherwise be generic: every
// access through it would become a flat instruction (address-space checks,
// LDS and vector-memory counters both), ~10 % slower kernels. The host sees
// plain pointers of the same size.
#ifdef __HIP_DEVICE_COMPILE__
template <typename T> using gptr = T __attribute__((address_space(1))) *;
#else
template <typename T> using gptr = T *;
#endif

// Static per-node fields stored in BF16 (all null: stored in the step
// precision).
struct Static16 {
    gptr<const bf16> lam, mu, tau, ca, cb, cc;
};

template <typename R>
struct ElemArgs {
    int nx, ny, nz, nel;
    int remap;                 // XCD-aware block order
    gptr<const int> it_dev;    // iterations completed so far (device counter);
                               // this launch computes iteration *it_dev + 1
    int nt;
    int kahan_sigma;           // 0 off, 1 FP32 compensation, 2 BF16 compensation
    gptr<const R> vx, vy, vz;  // unique-node velocity
    // state block read by the step (parts of one element-major block, see
    // em_index): stress, its compensation, memory variables
    gptr<const R> sig;               // 6 components
    gptr<const R> sig_lo;            // 6 compensation words (kahan_sigma 1) or null
    gptr<const unsigned> sig_lo2;    // 3 words (kahan_sigma 2): BF16 pairs (xx,xy)
                                     // (xz,yy) (yz,zz), first in the low half
    gptr<const R> mem;               // 18 memory variables (visco) or null
    // state block the step writes: the same block (one state buffer) or the
    // other one of a pair (two state buffers, swapped every step)
    gptr<R> sig_w, sig_lo_w, mem_w;
    gptr<unsigned> sig_lo2_w;
    // [nel][NP] moduli times dt: elastic lam, mu; visco-elastic
    // lam - 2/3 tau mu and mu (1 + tau) (relaxed-to-unrelaxed form)
    gptr<const R> lam, mu, tau;
    gptr<const R> ca, cb, cc;  // anisotropy dt*A, dt*B, dt*C or null
    Static16 s16;              // the same fields in BF16 when lam is null
    gptr<R> force;             // element force, k < p nodes then k = p layers (force_index)
    gptr<double> dgrad;        // [nel][6][NP] displacement gradient (xx xy xz yy yz zz) or null
    double dt;                 // time step (displacement gradient increment)
    gptr<const ThTab<coef_t<R>>> tth;   // [nx*P1]
    gptr<const RTab<coef_t<R>>> tr;     // [nz*P1]
    gptr<const double> px, py, pz;  // 1-D taper profiles [n*P1]
    double alpha;              // taper slope
    int taper_on;
    gptr<const int> src_slot;  // [nel], -1 = none, or null
    gptr<const int> src_chan;  // [nslot][9] channel row in stf, -1 = none
    gptr<const double> src_coef;  // [nslot][9][NP]
    gptr<const double> stf;    // [nchan][nt]
    coef_t<R> d1, d2, d3;      // (2/dx)J, (2/dy)J, -(2/dz)J
    R rate[NRDISS];            // dt/tau_p
    coef_t<R> w[MAXP1];        // GLL weights
    coef_t<R> Dx[MAXP1 * MAXP1];  // Dx[i*P1+q] = 2 l_q'(xi_i) / dx
    coef_t<R> Dy[MAXP1 * MAXP1];
    coef_t<R> Dz[MAXP1 * MAXP1];
    coef_t<R> dlw[MAXP1 * MAXP1]; // dlw[i*P1+q] = l_i'(xi_q) w_q
};

template <typename R>
struct NodeArgs {
    int nx, ny, nz, nel;
    long nblk;                 // (nx+1)(ny+1)(nz+1)
    gptr<R> vx, vy, vz;
    gptr<R> lx, ly, lz;        // compensation terms in FP32 (kahan_v 1) or null
    gptr<bf16> hx, hy, hz;     // compensation terms in BF16 (kahan_v 2) or null
    gptr<const R> force;       // element force (force_index)
    gptr<const R> dt_inv_mm;   // unique, dt / M
    gptr<const bf16> dt_inv_mm16;  // the same in BF16 when dt_inv_mm is null
    gptr<const double> px, py, pz;
    double alpha;
    int taper_on;
};

// Compensation term of a compensated velocity component at unique node u,
// from FP32 words (l), BF16 words (h) or none (both null): the component's
// value is v - lo.
template <typename R>
__device__ __forceinline__ double vel_lo(gptr<const R> l, gptr<const bf16> h, long u)
{
    return l ? (double)l[u] : h ? (double)(float)h[u] : 0.0;
}

// taper = exp(-0.1 * prof^2) after corner normalisation
// (reference src/geometric_paras.f90:303-316). 0.1 is the reference's
// single-precision literal promoted to double.
__host__ __device__ inline double taper_value(double px, double py, double pz,
                                              double alpha)
{
    double prof = px + py + pz;
    if (prof == 0.0) return 1.0;
    double den = 1.0 + prof / alpha;
    double pz2 = 2.0 * pz / den;
    double py2 = 2.0 * py / den;
    double px2 = 2.0 * px / den;
    double p2 = px2 + py2 + pz2;
    return exp(-((double)0.1f * (p2 * p2)));
}

// 1 - taper, from the same profile: the kernels apply the taper as
// f - f * damp. A factor such as 0.9999 rounded to FP32 keeps only a few
// digits of its distance from 1, and the taper acts every step, so the
// damping is carried to full relative precision instead.
__host__ __device__ inline double taper_damp(double px, double py, double pz, double alpha)
{
    double prof = px + py + pz;
    if (prof == 0.0) return 0.0;
    double den = 1.0 + prof / alpha;
    double pz2 = 2.0 * pz / den;
    double py2 = 2.0 * py / den;
    double px2 = 2.0 * px / den;
    double p2 = px2 + py2 + pz2;
    return -expm1(-((double)0.1f * (p2 * p2)));
}

}  // namespace s3d
