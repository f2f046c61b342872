// s3d_kernels.hip -- gfx950 kernels of the forward time step.
//
// One time step (reference src/ses3d-nt.F90:240-270 around
// src/evolution.f90:87-444) is two kernels plus a tiny receiver kernel:
//
//   elem_kernel  one workgroup per EPB elements, one lane per element-local
//                node. Gathers the element's velocity into LDS, forms the
//                strain rate (K1-K4), the stress rate and the stress / memory
//                variable update (K5-K7, SLS), adds the moment-tensor glut,
//                stages the effective stress in LDS and contracts the weak
//                divergence (K8). Writes the element force, the tapered
//                stress and the memory variables; all HBM traffic coalesced.
//   node_kernel  one lane per unique node. Sums the (up to 8) element copies
//                of the force in the reference's theta -> phi -> r order
//                (K10, src/communicate_fields.F90:48-79; deterministic, no
//                atomics), multiplies by 1/M (K9), integrates the velocity
//                (K12, optional compensated summation) and tapers it (K13).
//   record_kernel one wave per receiver: sum vm * field over the element.
//
// HBM bytes per element-local node and step (FP32, Kahan on v, visco):
//   stress 24 r + 24 w, memory variables 72 r + 72 w (60 + 60 in the
//   five-word deviatoric form), lambda/mu/tau 12 r,
//   force 12 w + 12 r, velocity ~(3 r gather) + unique-node v, lo, 1/M
//   (~24*u r + 24*u w with u = unique/element-local ~ 0.52).
#include "s3d_elem.h"

namespace s3d {

// vel_lo for plain pointers (the set-up and field kernels)
template <typename R>
__device__ __forceinline__ double lo_at(const R *l, const bf16 *h, long u)
{
    return l ? (double)l[u] : h ? (double)(float)h[u] : 0.0;
}

// ---------------------------------------------------------------------------
// node kernel: assembly, 1/M, velocity integration, taper
// ---------------------------------------------------------------------------
// KV: compensated velocity integration, 0 off, 1 FP32 compensation words,
// 2 BF16 compensation words (the term carries the rounding error of the
// sum, below half an FP32 unit of it, so 8 bits of it keep the sum within
// a small fraction of that unit).
template <int P, typename R, int KV>
__device__ __forceinline__ void node_body(const NodeArgs<R> &a, long tid)
{
    constexpr int P1 = P + 1, PB = P * P * P;
    constexpr int NL = P1 * P1, NM = NL * P;   // force layout, see force_index
    if (tid >= a.nblk * PB) return;
    // 32-bit index arithmetic (s3d_create keeps unique nodes below 2^32 and
    // elements below 2^31): 64-bit divisions by the run-time block counts
    // took about 40 % of the kernel's VALU instructions
    const unsigned t = (unsigned)tid;
    const unsigned b = t / PB;
    const unsigned o = t - b * PB;
    const unsigned li = o % P, lj = (o / P) % P, lk = o / (P * P);
    const unsigned nx = a.nx, ny = a.ny, nz = a.nz;
    const unsigned nbx = nx + 1, nby = ny + 1;
    const unsigned bq = b / nbx;
    const unsigned bx = b - bq * nbx, bz = bq / nby, by = bq - bz * nby;
    if ((bx == nx && li) || (by == ny && lj) || (bz == nz && lk)) return;

    // Copies: element (bx-dx, by-dy, bz-dz) at local (li+dx*P, lj+dy*P,
    // lk+dz*P), from clamped (always valid) addresses: dz = 0 in the k < p
    // part of the force, dz = 1 (lk = 0, local k = p) in the k = p layers
    // (force_index). Only present copies are loaded (lane-predicated loads:
    // an interior node, the common case, issues one load instead of eight);
    // absent ones are +0. Adding +0 is exact: the sum equals the reference's
    // theta -> phi -> r sweep.
    const bool px0 = bx < nx, px1 = li == 0 && bx > 0;
    const bool py0 = by < ny, py1 = lj == 0 && by > 0;
    const bool pz0 = bz < nz, pz1 = lk == 0 && bz > 0;
    const unsigned ex0 = bx < nx ? bx : nx - 1, ex1 = bx > 0 ? bx - 1 : 0;
    const unsigned ey0 = by < ny ? by : ny - 1, ey1 = by > 0 ? by - 1 : 0;
    const unsigned ez0 = bz < nz ? bz : nz - 1, ez1 = bz > 0 ? bz - 1 : 0;
    const long top = (long)a.nel * 3 * NM;   // start of the k = p layers
    long addr[8];
    bool pres[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int dx = m & 1, dy = (m >> 1) & 1, dz = m >> 2;
        const unsigned e = (dx ? ex1 : ex0) + nx * ((dy ? ey1 : ey0) + ny * (dz ? ez1 : ez0));
        const unsigned i2 = li + dx * P < P ? li + dx * P : P;
        const unsigned j2 = lj + dy * P < P ? lj + dy * P : P;
        addr[m] = dz ? top + (long)e * (3 * NL) + (i2 + P1 * j2) : (long)e * (3 * NM) + (i2 + P1 * (j2 + P1 * lk));
        pres[m] = (dx ? px1 : px0) && (dy ? py1 : py0) && (dz ? pz1 : pz0);
    }
    R acc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R v[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = R(0);
        v[0] = a.force[addr[0] + c * NM];   // clamped: in bounds also for a padding block
#pragma unroll
        for (int m = 1; m < 8; ++m)
            if (pres[m]) v[m] = a.force[addr[m] + c * (m < 4 ? NM : NL)];
        v[0] = pres[0] ? v[0] : R(0);
        acc[c] = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    }
    const long u = tid;
    R im;
    if (a.dt_inv_mm)
        im = a.dt_inv_mm[u];
    else
        im = (R)(float)a.dt_inv_mm16[u];
    R dmp = R(0);   // 1 - taper
    if (a.taper_on) {
        const unsigned ex = bx < nx ? bx : nx - 1, i = bx < nx ? li : P;
        const unsigned ey = by < ny ? by : ny - 1, j = by < ny ? lj : P;
        const unsigned ez = bz < nz ? bz : nz - 1, k = bz < nz ? lk : P;
        dmp = (R)taper_damp(a.px[ex * P1 + i], a.py[ey * P1 + j], a.pz[ez * P1 + k], a.alpha);
    }
    R *v[3] = {a.vx, a.vy, a.vz};
    if constexpr (KV == 1) {
        R *l[3] = {a.lx, a.ly, a.lz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R s = v[c][u], cc = l[c][u];
            kahan_add(s, cc, acc[c] * im);
            v[c][u] = s - s * dmp;
            l[c][u] = cc - cc * dmp;
        }
    } else if constexpr (KV == 2) {
        bf16 *h[3] = {a.hx, a.hy, a.hz};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            R s = v[c][u], cc = (R)(float)h[c][u];
            kahan_add(s, cc, acc[c] * im);
            v[c][u] = s - s * dmp;
            h[c][u] = (bf16)(float)(cc - cc * dmp);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const R s = v[c][u] + acc[c] * im;
            v[c][u] = s - s * dmp;
        }
    }
}

template <int P, typename R, int KV>
__global__ void __launch_bounds__(256)
node_kernel(const NodeArgs<R> a)
{
    node_body<P, R, KV>(a, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <int P, typename R, int KV>
__global__ void __launch_bounds__(256)
node_kernel_batch(const NodeArgs<R> *__restrict__ args, const int2 *__restrict__ own)
{
    const int2 o = own[blockIdx.x];
    node_body<P, R, KV>(args[o.x], (long)o.y * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// receivers: one wave per receiver, all its channels
// ---------------------------------------------------------------------------
// The curl formulas (src/config/channel_mod.f90:620-774 for receivers,
// src/volume_snapshot.f90:238-245 for snapshots) from the parts of a step's
// start velocity and the velocity vx, vy at its end; comp 0, 1, 2 =
// rot_vx, rot_vy, rot_vz.
__device__ inline double curl_formula(int comp, const double part[6], double vxn, double vyn, double r,
                                      double rs, double ct, double st)
{
    if (comp == 0) return -vyn / r + part[0] / rs + part[3];
    if (comp == 1) return -part[4] + (vxn - part[1]) / r;
    return (part[2] + vyn * ct - part[5] / st) / r;
}

template <typename R>
struct RecArgs {
    int nx, ny, nel, nt;
    int ns;     // components of the state block
    gptr<const R> vx, vy, vz, lx, ly, lz, sig;
    gptr<const bf16> hx, hy, hz;  // BF16 velocity compensation (lx.. null) or null
    gptr<const int> rec_rot;      // curl slot per receiver, -1 none (null: no curl receivers)
    gptr<const double> rot_buf;   // per curl slot [10][NP], see curl_strain_kernel
    gptr<const int> rec_e, rec_nch, rec_off, rec_ch;  // channels of receiver r: rec_ch[rec_off[r] + q]
    gptr<const double> rec_vm;
    gptr<double> traces;          // [sum of channels][nt], receiver r's at rec_off[r]
    gptr<const int> it_dev;
    // sources, for the echo channels: slot per element (null: no sources),
    // channel row per slot and component, coefficients, source-time table
    gptr<const int> src_slot, src_chan;
    gptr<const double> src_coef, stf;
};

template <int P, typename R>
__device__ __forceinline__ void record_body(const RecArgs<R> &a, int r)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    // records iteration it = *it_dev + 1 (the counter advances after this kernel)
    const int it = *a.it_dev + 1;
    if (it < 1 || it > a.nt) return;
    const int lane = threadIdx.x;
    const int e = a.rec_e[r];
    const int ex = e % a.nx, ey = (e / a.nx) % a.ny, ez = e / (a.nx * a.ny);
    // source-echo channels: add(it) of this step wrote, per component c with
    // a source in this element, src_c = -stf(it) * coef_c (an assignment: the
    // last source of the element wins, as in add_m** / add_f*,
    // src/config/channel_mod.f90:859-1095); record_m** / record_f* sum vm *
    // src_c (:871-1085)
    const int slot = a.src_slot ? a.src_slot[e] : -1;
    double sd[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const int ch = slot >= 0 ? a.src_chan[slot * 9 + c] : -1;
        sd[c] = ch >= 0 ? a.stf[(long)ch * a.nt + (it - 1)] : 0.0;
    }
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};   // vx, vy, vz, pressure, rot_vx, rot_vy, rot_vz
    double accs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // mxx .. mzz, fx, fy, fz
    const int rs = a.rec_rot ? a.rec_rot[r] : -1;
    for (int n = lane; n < NP; n += 64) {
        const int i = n % P1, j = (n / P1) % P1, k = n / (P1 * P1);
        const long u = unique_index<P>(ex, ey, ez, i, j, k, a.nx, a.ny);
        const double w = a.rec_vm[(long)r * NP + n];
        // compensated velocity: value = hi - lo
        const double vxn = (double)a.vx[u] - vel_lo(a.lx, a.hx, u);
        const double vyn = (double)a.vy[u] - vel_lo(a.ly, a.hy, u);
        acc[0] += w * vxn;
        acc[1] += w * vyn;
        acc[2] += w * ((double)a.vz[u] - vel_lo(a.lz, a.hz, u));
        const double tr = (double)a.sig[em_index(e, a.ns, XX, NP, n)] + (double)a.sig[em_index(e, a.ns, YY, NP, n)] +
                          (double)a.sig[em_index(e, a.ns, ZZ, NP, n)];
        acc[3] += -((double)(1.0f / 3.0f) * tr) * w;
        if (rs >= 0) {
            // curl channels (src/config/channel_mod.f90:620-774): strain-rate
            // parts of the step's start velocity, this velocity at its end
            const double *b = a.rot_buf + (long)rs * 10 * NP + n;
            double part[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) part[c] = b[c * NP];
            const double rr = b[6 * NP], rst = b[7 * NP], ct = b[8 * NP], st = b[9 * NP];
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[4 + c] += w * curl_formula(c, part, vxn, vyn, rr, rst, ct, st);
        }
        if (slot >= 0) {
#pragma unroll
            for (int c = 0; c < 9; ++c)
                if (sd[c] != 0.0)
                    accs[c] += w * -(sd[c] * a.src_coef[((long)slot * 9 + c) * NP + n]);
        }
    }
#pragma unroll
    for (int c = 0; c < 7; ++c)
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_down(acc[c], off, 64);
#pragma unroll
    for (int c = 0; c < 9; ++c)
        for (int off = 32; off > 0; off >>= 1) accs[c] += __shfl_down(accs[c], off, 64);
    if (lane == 0) {
        const int nch = a.rec_nch[r];
        const int off = a.rec_off[r];
        for (int q = 0; q < nch; ++q) {
            const int ch = a.rec_ch[off + q];
            double val = 0.0;
            switch (ch) {
            case 0: val = acc[0]; break;          // vx
            case 1: val = acc[1]; break;          // vy
            case 2: val = acc[2]; break;          // vz
            case 3: val = -acc[0]; break;         // n = -vx
            case 4: val = acc[1]; break;          // e = vy
            case 5: val = acc[2]; break;          // z = vz
            case 6: val = acc[3]; break;          // pressure
            case 7: val = acc[4]; break;          // rot_vx
            case 8: val = acc[5]; break;          // rot_vy
            case 9: val = acc[6]; break;          // rot_vz
            default:                              // mxx .. mzz, fx, fy, fz
                if (ch >= 10 && ch < 19) val = accs[ch - 10];
                break;
            }
            a.traces[((long)off + q) * a.nt + (it - 1)] = val;
        }
    }
}

// Strain-rate parts of the curl receiver channels (reference record_rot_v*,
// src/config/channel_mod.f90:620-774, which read the step's strain-rate
// arrays): for each curl receiver's element, from the velocity at the start
// of the step (hi - lo, FP64), the raw derivatives (src/evolution.f90:126-159),
// their spherical completion (:165-173) and the symmetrisation (:206-208) as
// the step forms them; stored per node: EYZ, EZY, EZX, EXZ, EXY, EYX (EYZ,
// EXZ, EXY symmetrised; EZY, EZX, EYX completed only), r, r sin(theta),
// cot(theta), sin(theta). Launched before the node kernel updates v.
template <typename R>
struct CurlArgs {
    int nx, ny, nrot;
    gptr<const R> vx, vy, vz, lx, ly, lz;
    gptr<const bf16> hx, hy, hz;   // BF16 velocity compensation (lx.. null) or null
    gptr<const ThTab<double>> tth;
    gptr<const RTab<double>> tr;
    gptr<const int> rot_e;       // element of each curl slot
    gptr<double> buf;            // [nrot][10][NP]
    double Dx[MAXP1 * MAXP1], Dy[MAXP1 * MAXP1], Dz[MAXP1 * MAXP1];   // D[i*P1 + q] = 2 l'_q(xi_i) / d
};

template <typename R>
__device__ __forceinline__ double curl_vel(gptr<const R> v, gptr<const R> l, gptr<const bf16> h, long u)
{
    return (double)v[u] - vel_lo(l, h, u);
}

// The strain-rate parts the curl formulas use, at node (i, j, k) of element
// (ex, ey, ez), from the current velocity: eyz, exz, exy as the step's
// symmetrisation leaves them and ezy, ezx, eyx as its spherical completion
// leaves them (src/evolution.f90:165-208), in that order in part[0..5].
template <int P, typename R>
__device__ inline void curl_parts(const CurlArgs<R> &a, int ex, int ey, int ez, int i, int j, int k,
                                  double part[6])
{
    constexpr int P1 = P + 1;
    double exy = 0.0, exz = 0.0, eyx = 0.0, eyz = 0.0, ezx = 0.0, ezy = 0.0;
    for (int q = 0; q < P1; ++q) {
        const double d1 = a.Dx[i * P1 + q];
        const long u = unique_index<P>(ex, ey, ez, q, j, k, a.nx, a.ny);
        exy = exy + curl_vel(a.vy, a.ly, a.hy, u) * d1;
        exz = exz + curl_vel(a.vz, a.lz, a.hz, u) * d1;
    }
    for (int q = 0; q < P1; ++q) {
        const double d2 = a.Dy[j * P1 + q];
        const long u = unique_index<P>(ex, ey, ez, i, q, k, a.nx, a.ny);
        eyx = eyx + curl_vel(a.vx, a.lx, a.hx, u) * d2;
        eyz = eyz + curl_vel(a.vz, a.lz, a.hz, u) * d2;
    }
    for (int q = 0; q < P1; ++q) {
        const double d3 = a.Dz[k * P1 + q];
        const long u = unique_index<P>(ex, ey, ez, i, j, q, a.nx, a.ny);
        ezx = ezx + curl_vel(a.vx, a.lx, a.hx, u) * d3;
        ezy = ezy + curl_vel(a.vy, a.ly, a.hy, u) * d3;
    }
    const long u = unique_index<P>(ex, ey, ez, i, j, k, a.nx, a.ny);
    const double vxn = curl_vel(a.vx, a.lx, a.hx, u), vyn = curl_vel(a.vy, a.ly, a.hy, u);
    const double r = a.tr[ez * P1 + k].r, st = a.tth[ex * P1 + i].sin_t, ct = a.tth[ex * P1 + i].cot;
    const double rs = r * st;
    ezx = -ezx;
    exz = (exz - vxn) / r;
    ezy = -ezy;
    eyz = -vyn / r + eyz / rs;
    exy = exy / r;
    eyx = -(vyn * ct) / r + eyx / rs;
    part[0] = (eyz + ezy) / 2.0;
    part[1] = (exz + ezx) / 2.0;
    part[2] = (exy + eyx) / 2.0;
    part[3] = ezy;
    part[4] = ezx;
    part[5] = eyx;
}

template <int P, typename R>
__global__ void __launch_bounds__(64) curl_strain_kernel(const CurlArgs<R> a)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const int s = blockIdx.x;
    if (s >= a.nrot) return;
    const int e = a.rot_e[s];
    const int ex = e % a.nx, ey = (e / a.nx) % a.ny, ez = e / (a.nx * a.ny);
    for (int n = threadIdx.x; n < NP; n += 64) {
        const int i = n % P1, j = (n / P1) % P1, k = n / (P1 * P1);
        double part[6];
        curl_parts<P, R>(a, ex, ey, ez, i, j, k, part);
        const double r = a.tr[ez * P1 + k].r, st = a.tth[ex * P1 + i].sin_t, ct = a.tth[ex * P1 + i].cot;
        double *b = a.buf + (long)s * 10 * NP + n;
#pragma unroll
        for (int c = 0; c < 6; ++c) b[c * NP] = part[c];
        b[6 * NP] = r;
        b[7 * NP] = r * st;
        b[8 * NP] = ct;
        b[9 * NP] = st;
    }
}

// Volume curl fields, first half: the parts of every element-local node from
// the current velocity (the start velocity of the next step), element-major
// parts[e][6][NP]. One workgroup per element.
template <int P, typename R>
__global__ void __launch_bounds__(64) curl_parts_kernel(const CurlArgs<R> a, long nel, double *parts)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const long e = blockIdx.x;
    if (e >= nel) return;
    const int ex = (int)(e % a.nx), ey = (int)((e / a.nx) % a.ny), ez = (int)(e / ((long)a.nx * a.ny));
    for (int n = threadIdx.x; n < NP; n += 64) {
        const int i = n % P1, j = (n / P1) % P1, k = n / (P1 * P1);
        double part[6];
        curl_parts<P, R>(a, ex, ey, ez, i, j, k, part);
#pragma unroll
        for (int c = 0; c < 6; ++c) parts[(e * 6 + c) * NP + n] = part[c];
    }
}

// Volume curl fields, second half: rot_v{x,y,z} (comp 0, 1, 2) of every
// element-local node from the stored parts and the current velocity, in the
// reference's 6-D layout out[ex + nx (ey + ny (ez + nz (i + P1 (j + P1 k))))].
template <int P, typename R>
__global__ void __launch_bounds__(256) curl_field_kernel(const CurlArgs<R> a, int nz, long nn,
                                                         const double *parts, int comp, double *out)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nn) return;
    const long e = t / NP;
    const int n = (int)(t - e * NP);
    const int ex = (int)(e % a.nx), ey = (int)((e / a.nx) % a.ny), ez = (int)(e / ((long)a.nx * a.ny));
    const int i = n % P1, j = (n / P1) % P1, k = n / (P1 * P1);
    double part[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) part[c] = parts[(e * 6 + c) * NP + n];
    const long u = unique_index<P>(ex, ey, ez, i, j, k, a.nx, a.ny);
    const double vxn = curl_vel(a.vx, a.lx, a.hx, u), vyn = curl_vel(a.vy, a.ly, a.hy, u);
    const double r = a.tr[ez * P1 + k].r, st = a.tth[ex * P1 + i].sin_t, ct = a.tth[ex * P1 + i].cot;
    const long o = ex + (long)a.nx * (ey + (long)a.ny * (ez + (long)nz * (i + (long)P1 * (j + (long)P1 * k))));
    out[o] = curl_formula(comp, part, vxn, vyn, r, r * st, ct, st);
}

template <int P, typename R>
__global__ void __launch_bounds__(64) record_kernel(const RecArgs<R> a)
{
    record_body<P, R>(a, blockIdx.x);
}

template <int P, typename R>
__global__ void __launch_bounds__(64) record_kernel_batch(const RecArgs<R> *__restrict__ args, const int2 *__restrict__ own)
{
    const int2 o = own[blockIdx.x];
    record_body<P, R>(args[o.x], o.y);
}

// End of one iteration: advance the device step counter. A single lane
// stores through the vector path.
__global__ void __launch_bounds__(64) advance_kernel(int *it_dev)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) it_dev[0] = it_dev[0] + 1;
}

// End of one batched iteration: advance every simulation's counter.
__global__ void __launch_bounds__(64) advance_many_kernel(int *const *its, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) its[s][0] = its[s][0] + 1;
}

// ---------------------------------------------------------------------------
// setup and I/O kernels
// ---------------------------------------------------------------------------

// Conversions between the reference layout [n][e] (FP64) and one component
// of an element-major block field of nc components (dst/src point at the
// component's first entry; element e's entries start at e * nc * NP).
template <typename R>
__global__ void to_device_layout(const double *src, R *dst, long nel, int NP, int nc)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long e = t / NP;
    int n = (int)(t - e * NP);
    dst[e * nc * NP + n] = (R)src[(long)n * nel + e];
}

template <typename R>
__global__ void from_device_layout(const R *src, double *dst, long nel, int NP, int nc)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long n = t / nel;
    long e = t - n * nel;
    dst[t] = (double)src[e * nc * NP + n];
}

template <typename R>
__global__ void clear_component(R *p, long nel, int NP, int nc)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long e = t / NP;
    int n = (int)(t - e * NP);
    p[e * nc * NP + n] = R(0);
}

// one half of a packed BF16 stress compensation component of a block of nc
// -> reference layout
__global__ void lo_pair_from_device(const unsigned *src, int half, double *dst, long nel, int NP, int nc)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long n = t / nel;
    long e = t - n * nel;
    const unsigned w = src[e * nc * NP + n];
    dst[t] = (double)(half ? bf16_hi(w) : bf16_lo(w));
}

// clear one half of one packed BF16 stress compensation component of a block
// of nc
__global__ void lo_pair_clear(unsigned *p, int half, long nel, int NP, int nc)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long e = t / NP;
    int n = (int)(t - e * NP);
    p[e * nc * NP + n] &= half ? 0x0000ffffu : 0xffff0000u;
}

// Adjoint density sensitivity (reference ses3d_grad,
// src/SOURCE/ses3d_grad.F90:212-213), all arrays in the reference layout
// [n][e]: g += inc_dt * (fx*vx + fy*vy + fz*vz) / jac, the velocity (hi - lo)
// gathered from the unique nodes.
template <int P, typename R>
__global__ void grad_rho_kernel(double *g, const double *fx, const double *fy, const double *fz,
                                const R *vx, const R *vy, const R *vz, const R *lx, const R *ly,
                                const R *lz, const bf16 *hx, const bf16 *hy, const bf16 *hz,
                                int nx, int ny, int nz, double inc_dt, double jac)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const long nel = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long n = t / nel;
    long e = t - n * nel;
    int i = (int)(n % P1), j = (int)((n / P1) % P1), k = (int)(n / (P1 * P1));
    int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((long)nx * ny));
    long u = unique_index<P>(ex, ey, ez, i, j, k, nx, ny);
    const double ax = (double)vx[u] - lo_at(lx, hx, u);
    const double ay = (double)vy[u] - lo_at(ly, hy, u);
    const double az = (double)vz[u] - lo_at(lz, hz, u);
    g[t] = g[t] + inc_dt * ((fx[t] * ax + fy[t] * ay) + fz[t] * az) / jac;
}

// unique-node field -> reference layout (every copy)
template <int P, typename R>
__global__ void expand_unique(const R *src, const R *lo, const bf16 *lo16, double *dst, int nx, int ny, int nz)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const long nel = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long n = t / nel;
    long e = t - n * nel;
    int i = (int)(n % P1), j = (int)((n / P1) % P1), k = (int)(n / (P1 * P1));
    int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((long)nx * ny));
    long u = unique_index<P>(ex, ey, ez, i, j, k, nx, ny);
    dst[t] = (double)src[u] - lo_at(lo, lo16, u);
}

// reference layout -> unique-node field, taking the owner element's copy
template <int P, typename R>
__global__ void gather_unique(const double *src, R *dst, R *lo, bf16 *lo16, int nx, int ny, int nz, long nblk)
{
    constexpr int P1 = P + 1, PB = P * P * P;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long b = t / PB;
    int o = (int)(t - b * PB);
    if (b >= nblk) return;
    int li = o % P, lj = (o / P) % P, lk = o / (P * P);
    int nbx = nx + 1, nby = ny + 1;
    int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long)nbx * nby));
    if ((bx == nx && li) || (by == ny && lj) || (bz == nz && lk)) return;
    int ex = bx < nx ? bx : nx - 1, i = bx < nx ? li : P;
    int ey = by < ny ? by : ny - 1, j = by < ny ? lj : P;
    int ez = bz < nz ? bz : nz - 1, k = bz < nz ? lk : P;
    long nel = (long)nx * ny * nz;
    long e = ex + (long)nx * (ey + (long)ny * ez);
    int n = i + P1 * (j + P1 * k);
    double v = src[(long)n * nel + e];
    dst[t] = (R)v;
    if (lo) lo[t] = (R)(v - (double)(R)v) * R(-1);
    if (lo16) lo16[t] = (bf16)(float)(-(v - (double)(R)v));
}

// FP64 -> storage precision (BF16 through the nearest float)
template <typename S>
__device__ inline S narrow(double x)
{
    if constexpr (sizeof(S) == 2)
        return (S)(float)x;
    else
        return (S)x;
}

// Material from the FP64 staging arrays in reference layout: tau
// (= 6.0814 Q^-1.015, src/model_paras.F90:407), rho, dt*A, dt*B, dt*C and the
// moduli times dt, elastic: lambda, mu; visco-elastic (dtau non-null):
// kappa - 2 mu_tau / 3 = lambda - 2/3 tau mu and mu_tau = mu (1 + tau)
// (src/evolution.f90:216-284). All formed in FP64 and rounded once: a
// constant tau makes 1 + tau the same number in every node of a region, and
// an FP32 rounding of it would be a systematic modulus error.
template <typename R, typename S>
__global__ void material_kernel(const double *rhoinv, const double *lam, const double *mu,
                                const double *q, const double *A, const double *B, const double *C,
                                S *dlam, S *dmu, S *dtau, R *drho, S *dA, S *dB, S *dC,
                                long nel, int NP, int ncm, double dt)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nel * NP) return;
    long e = t / NP;
    int n = (int)(t - e * NP);
    long s = (long)n * nel + e;
    long tm = em_index(e, ncm, 0, NP, n);   // lam/mu/tau: element-major block of ncm
    const double l = lam ? lam[s] : 0.0, m = mu ? mu[s] : 0.0;
    if (dtau) {
        double qq = q ? q[s] : 1.79769313486231570815e+308;
        const double tau = (double)6.0814f * pow(qq, (double)-1.015f);
        dtau[tm] = narrow<S>(tau);
        dlam[tm] = narrow<S>(dt * (l - (2.0 / 3.0) * tau * m));
        dmu[tm] = narrow<S>(dt * (m * (1.0 + tau)));
    } else {
        dlam[tm] = narrow<S>(dt * l);
        dmu[tm] = narrow<S>(dt * m);
    }
    double ri = rhoinv ? rhoinv[s] : 0.0;
    drho[t] = (R)(ri > 0.0 ? 1.0 / ri : 0.0);
    if (dA) {
        dA[t] = A ? narrow<S>(dt * A[s]) : narrow<S>(0.0);
        dB[t] = B ? narrow<S>(dt * B[s]) : narrow<S>(0.0);
        dC[t] = C ? narrow<S>(dt * C[s]) : narrow<S>(0.0);
    }
}

// Mass matrix -J r^2 sin(theta) rho W per copy (src/model_paras.F90:411),
// assembled with the same theta -> phi -> r summation tree as the force.
// Stores M (FP64) and dt/M (storage precision S).
template <int P, typename S>
__global__ void mass_kernel(const double *rhoinv, const double *sin_t, const double *rad,
                            const double *w, double jac, double dt, double *mm64, S *dt_inv_mm,
                            int nx, int ny, int nz, long nblk)
{
    constexpr int P1 = P + 1, PB = P * P * P;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long b = t / PB;
    int o = (int)(t - b * PB);
    if (b >= nblk) return;
    int li = o % P, lj = (o / P) % P, lk = o / (P * P);
    int nbx = nx + 1, nby = ny + 1;
    int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long)nbx * nby));
    if ((bx == nx && li) || (by == ny && lj) || (bz == nz && lk)) {
        dt_inv_mm[t] = narrow<S>(0.0);
        mm64[t] = 0.0;
        return;
    }
    const long nel = (long)nx * ny * nz;
    bool px[2] = {bx < nx, li == 0 && bx > 0};
    bool py[2] = {by < ny, lj == 0 && by > 0};
    bool pz[2] = {bz < nz, lk == 0 && bz > 0};
    double zsum = 0;
    bool zhas = false;
    for (int dz = 0; dz < 2; ++dz) {
        if (!pz[dz]) continue;
        double ysum = 0;
        bool yhas = false;
        for (int dy = 0; dy < 2; ++dy) {
            if (!py[dy]) continue;
            double xsum = 0;
            bool xhas = false;
            for (int dx = 0; dx < 2; ++dx) {
                if (!px[dx]) continue;
                int ex = bx - dx, ey = by - dy, ez = bz - dz;
                int i = li + dx * P, j = lj + dy * P, k = lk + dz * P;
                long e = ex + (long)nx * (ey + (long)ny * ez);
                int n = i + P1 * (j + P1 * k);
                double ri = rhoinv ? rhoinv[(long)n * nel + e] : 0.0;
                double rho = ri > 0.0 ? 1.0 / ri : 0.0;
                double r = rad[ez * P1 + k];
                double ww = w[i] * w[j] * w[k];
                double val = -(jac * (r * r) * sin_t[ex * P1 + i] * rho * ww);
                xsum = xhas ? xsum + val : val;
                xhas = true;
            }
            ysum = yhas ? ysum + xsum : xsum;
            yhas = true;
        }
        zsum = zhas ? zsum + ysum : ysum;
        zhas = true;
    }
    mm64[t] = zsum;
    dt_inv_mm[t] = narrow<S>(zsum != 0.0 ? dt / zsum : 0.0);
}

// Courant number partial maxima (src/model_paras.F90:529-561), FP64.
template <int P>
__global__ void cfl_kernel(const double *rhoinv, const double *lam, const double *mu,
                           const double *theta, const double *phi, const double *rad,
                           int nx, int ny, int nz, double *partial)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const long nel = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double m[3] = {0, 0, 0};
    if (t < nel * NP) {
        long n = t / nel, e = t - n * nel;
        int i = (int)(n % P1), j = (int)((n / P1) % P1), k = (int)(n / (P1 * P1));
        int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((long)nx * ny));
        double ri = rhoinv ? rhoinv[t] : 0.0;
        double l = lam ? lam[t] : 0.0, u = mu ? mu[t] : 0.0;
        double cp = sqrt((l + 2.0 * u) * ri), cs = sqrt(u * ri);
        double vmax = cs < cp ? cp : cs;
        double r = rad[ez * P1 + k];
        if (i >= 1) m[0] = vmax / (r * (theta[ex * P1 + i] - theta[ex * P1 + i - 1]));
        if (j >= 1) m[1] = vmax / (r * sin(theta[ex * P1 + i]) * (phi[ey * P1 + j] - phi[ey * P1 + j - 1]));
        if (k >= 1) m[2] = vmax / (rad[ez * P1 + k - 1] - r);
    }
    __shared__ double sm[3][256];
    for (int c = 0; c < 3; ++c) sm[c][threadIdx.x] = m[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c)
                sm[c][threadIdx.x] = fmax(sm[c][threadIdx.x], sm[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) partial[blockIdx.x * 3 + c] = sm[c][0];
}

// v range and kinetic energy partials over unique / element-local nodes
template <int P, typename R>
__global__ void vrange_kernel(const R *vx, const R *vy, const R *vz, int nx, int ny, int nz,
                              long nblk, double *partial)
{
    constexpr int PB = P * P * P;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long b = t / PB;
    int o = (int)(t - b * PB);
    double mn = 1e308, mx = -1e308;
    if (b < nblk) {
        int li = o % P, lj = (o / P) % P, lk = o / (P * P);
        int nbx = nx + 1, nby = ny + 1;
        int bx = (int)(b % nbx), by = (int)((b / nbx) % nby), bz = (int)(b / ((long)nbx * nby));
        if (!((bx == nx && li) || (by == ny && lj) || (bz == nz && lk))) {
            double a0 = vx[t], a1 = vy[t], a2 = vz[t];
            mn = fmin(a0, fmin(a1, a2));
            mx = fmax(a0, fmax(a1, a2));
        }
    }
    __shared__ double s0[256], s1[256];
    s0[threadIdx.x] = mn;
    s1[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s0[threadIdx.x] = fmin(s0[threadIdx.x], s0[threadIdx.x + s]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s0[0];
        partial[2 * blockIdx.x + 1] = s1[0];
    }
}

// min and max over n partial (min, max) pairs, one workgroup of 256 lanes;
// the result pair goes to out[0], out[1]
__global__ void __launch_bounds__(256) range_reduce_kernel(const double *partial, int n, double *out)
{
    double mn = 1e308, mx = -1e308;
    for (int q = threadIdx.x; q < n; q += 256) {
        mn = fmin(mn, partial[2 * q]);
        mx = fmax(mx, partial[2 * q + 1]);
    }
    __shared__ double s0[256], s1[256];
    s0[threadIdx.x] = mn;
    s1[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s0[threadIdx.x] = fmin(s0[threadIdx.x], s0[threadIdx.x + s]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

template <int P, typename R>
__global__ void kinetic_kernel(const R *vx, const R *vy, const R *vz, const R *rho,
                               const ThTab<coef_t<R>> *tth, const RTab<coef_t<R>> *tr, const double *w,
                               int nx, int ny, int nz, double *partial)
{
    constexpr int P1 = P + 1, NP = P1 * P1 * P1;
    const long nel = (long)nx * ny * nz;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (t < nel * NP) {
        long e = t / NP;
        int n = (int)(t - e * NP);
        int i = n % P1, j = (n / P1) % P1, k = n / (P1 * P1);
        int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((long)nx * ny));
        long u = unique_index<P>(ex, ey, ez, i, j, k, nx, ny);
        double a0 = vx[u], a1 = vy[u], a2 = vz[u];
        double r = (double)tr[ez * P1 + k].r;
        acc = (a0 * a0 + a1 * a1 + a2 * a2) * (r * r * (double)tth[ex * P1 + i].sin_t) *
              (double)rho[t] * (w[i] * w[j] * w[k]);
    }
    __shared__ double s0[256];
    s0[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) s0[threadIdx.x] += s0[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s0[0];
}

}  // namespace s3d
