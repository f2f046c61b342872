// s3d_elem.h -- the element kernel (strain rate, stress and memory-variable
// update, weak divergence; see s3d_kernels.hip for the step it belongs to)
// and the helpers it shares with the other kernels. Templates and inline
// functions only: the element-kernel launchers instantiate it once per GLL
// order in their own translation units (s3d_elem.hip), which build in
// parallel.
#pragma once

#include "s3d_internal.h"

namespace s3d {

// Element tile: EPB elements per workgroup, one lane per element-local node.
// EPB = 0 selects the default, as many elements as fit in 256 lanes.
template <int P>
constexpr int default_epb() { return (256 / ((P + 1) * (P + 1) * (P + 1))) > 0 ? (256 / ((P + 1) * (P + 1) * (P + 1))) : 1; }

template <int P, int E = 0>
struct Shape {
    static constexpr int P1 = P + 1;
    static constexpr int NP = P1 * P1 * P1;
    static constexpr int PB = P * P * P;
    static constexpr int EPB = E > 0 ? E : default_epb<P>();
    static constexpr int NT = ((EPB * NP + 63) / 64) * 64;
};

__device__ inline int remap_block(int b, int G, int remap)
{
    // Dispatch sends consecutive workgroups to consecutive XCDs (8 on MI355X).
    // Remapping gives each XCD a contiguous range of elements, so element
    // neighbours (shared velocity / force nodes) meet in the same L2.
    if (!remap || (G & 7)) return b;
    return (b & 7) * (G >> 3) + (b >> 3);
}

template <int P>
__device__ inline long unique_index(int ex, int ey, int ez, int i, int j, int k,
                                    int nx, int ny)
{
    int bx = ex + (i == P), by = ey + (j == P), bz = ez + (k == P);
    int li = (i == P) ? 0 : i, lj = (j == P) ? 0 : j, lk = (k == P) ? 0 : k;
    long b = (long)bx + (long)(nx + 1) * ((long)by + (long)(ny + 1) * bz);
    return b * (P * P * P) + li + P * (lj + P * lk);
}

template <typename R>
__device__ inline void kahan_add(R &s, R &c, R inc)
{
    R y = inc - c;
    R t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---------------------------------------------------------------------------
// element kernel
// ---------------------------------------------------------------------------
// Direct HBM -> LDS load of one dword per lane (global_load_lds_dword):
// lane l writes LDS byte address lds + 4 l. Written as inline assembly so the
// compiler's wait accounting never sees it: callers issue these loads before
// any compiler-issued load (vector memory operations complete in issue
// order, so the compiler's own waits stay exact) and read the LDS words
// only after wait_vmem_zero. The statement writes M0, so M0 is in its clobber
// list: the compiler keeps nothing there across it. LLVM reserves M0 and warns
// about the clobber; in the gfx950 code of this library the only
// instructions that touch M0 are these statements' own.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void lds_dma_dword(const void *g, unsigned lds)
{
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off"
                 :: "v"(g), "s"(__builtin_amdgcn_readfirstlane(lds)) : "memory", "m0");
}
#pragma clang diagnostic pop

// s_waitcnt vmcnt(0), returning a zero that the reads of LDS words written by
// lds_dma_dword must add to their index: the data dependency orders those
// reads after the wait without a memory clobber (which would make the
// compiler reload every argument read through a pointer).
__device__ __forceinline__ int wait_vmem_zero()
{
    int z;
    asm volatile("s_waitcnt vmcnt(0)\n\tv_mov_b32 %0, 0" : "=v"(z));
    return z;
}

typedef float lds4 __attribute__((ext_vector_type(4)));

// LDS layout of the tile's staged velocity and effective stress (the
// autotuner's choice; every layout gives the same arithmetic):
//   LL = 0  planar: lds[c][node], one 4-byte read per component;
//   LL = 1  node-interleaved (FP32 only): 48 bytes per node, the velocity in
//           the first 16-byte vector and the stress as the three triples the
//           divergence contracts, (xx xy xz | xy yy yz | xz yz zz), one
//           16-byte read per contraction term instead of three 4-byte reads.
template <typename R, int LL, int N>
struct Stage {
    static constexpr bool INTERLEAVED = LL == 1 && sizeof(R) == 4;
    R planar[INTERLEAVED ? 1 : 6][INTERLEAVED ? 1 : N];
    lds4 inter[INTERLEAVED ? N : 1][INTERLEAVED ? 3 : 1];

    __device__ void put_v(int n, R v0, R v1, R v2)
    {
        if constexpr (INTERLEAVED) {
            inter[n][0] = (lds4){v0, v1, v2, 0.f};
        } else {
            planar[0][n] = v0;
            planar[1][n] = v1;
            planar[2][n] = v2;
        }
    }
    // velocity components at node n
    __device__ void get_v(int n, R *v) const
    {
        if constexpr (INTERLEAVED) {
            const lds4 w = inter[n][0];
            v[0] = w.x; v[1] = w.y; v[2] = w.z;
        } else {
            v[0] = planar[0][n]; v[1] = planar[1][n]; v[2] = planar[2][n];
        }
    }
    // effective stress s[xx xy xz yy yz zz]
    __device__ void put_s(int n, const R *sv)
    {
        if constexpr (INTERLEAVED) {
            inter[n][0] = (lds4){sv[XX], sv[XY], sv[XZ], 0.f};
            inter[n][1] = (lds4){sv[XY], sv[YY], sv[YZ], 0.f};
            inter[n][2] = (lds4){sv[XZ], sv[YZ], sv[ZZ], 0.f};
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) planar[c][n] = sv[c];
        }
    }
    // the stress components that contraction d (0 theta, 1 phi, 2 r) reads at
    // node n: (xx xy xz), (xy yy yz), (xz yz zz)
    __device__ void get_s(int d, int n, R *w3) const
    {
        if constexpr (INTERLEAVED) {
            const lds4 w = inter[n][d];
            w3[0] = w.x; w3[1] = w.y; w3[2] = w.z;
        } else {
            constexpr int comp[3][3] = {{XX, XY, XZ}, {XY, YY, YZ}, {XZ, YZ, ZZ}};
            w3[0] = planar[comp[d][0]][n];
            w3[1] = planar[comp[d][1]][n];
            w3[2] = planar[comp[d][2]][n];
        }
    }
};

// Per-lane state of one element tile (EPB elements, one lane per
// element-local node). VM: memory words per mechanism (state_components).
template <typename R, int VM>
struct TileState {
    R v[3];             // velocity gathered from the unique nodes
    R sg[6];            // stress
    R lam, mu, tau;     // dt-scaled moduli, tau
    R M[VM ? NRDISS * VM : 1];
    R lo[6];            // stress compensation, kahan_sigma 1
    unsigned lo2[3];    // stress compensation, kahan_sigma 2 (packed BF16 pairs)
};

// Element-major block fields (the state block, lam/mu/tau, the displacement
// gradient; the element force has its own two-part layout, force_index):
// component c of element e occupies the NP entries at (e * nc + c) * NP.
// The state block's parts are addressed through pointers at their first
// components (sig, sig_lo, mem) with nc = state_components(KS, VM).
__device__ __forceinline__ long em_index(long e, int nc, int c, int NP, int node)
{
    return (e * nc + c) * NP + node;
}

// Element force layout: first the nodes with k < p of every element,
// [e][c][(p+1)^2 p], then the k = p layers, [e][c][(p+1)^2]. The node kernel
// reads each part once: an element's k = p layer is the z-1 copy of the
// owned block above it, gathered a plane of blocks after the rest of the
// element, and in the plain element-major layout it shared cache lines with
// what had already been read.
template <int P>
__device__ __forceinline__ long force_index(long e, int c, int node, long nel)
{
    constexpr int NL = (P + 1) * (P + 1), NM = NL * P;
    return node < NM ? (e * 3 + c) * NM + node : nel * 3 * NM + (e * 3 + c) * NL + (node - NM);
}

// Issue the loads of tile `tile` into `s` (no waits). Lanes without an
// element (the tail of the last tile) read element 0's addresses, so every
// load is unconditional and in bounds; such lanes store nothing. KS: stress
// compensation 0 off, 1 FP32 words, 2 BF16 words in pairs (a template
// parameter: a run-time choice leaves register joins that wait for every
// load in flight).
template <int P, int E, typename R, int VM, bool S16, bool LOADM, int KS>
__device__ __forceinline__ void load_tile(const ElemArgs<R> &a, int tile, TileState<R, VM> &s)
{
    using S = Shape<P, E>;
    constexpr int P1 = S::P1, NP = S::NP, EPB = S::EPB;
    constexpr int NS = state_components(KS, VM);
    constexpr bool VISCO = VM != 0;
    const int t = threadIdx.x;
    const int slot = t / NP;
    const int node = t - slot * NP;
    const int e = tile * EPB + slot;
    const bool valid = slot < EPB && e < a.nel;
    const int ee = valid ? e : 0;
    const int ex = ee % a.nx, ey = (ee / a.nx) % a.ny, ez = ee / (a.nx * a.ny);
    const int i = node % P1, j = (node / P1) % P1, k = node / (P1 * P1);
    const long u = unique_index<P>(ex, ey, ez, i, j, k, a.nx, a.ny);
    s.v[0] = a.vx[u];
    s.v[1] = a.vy[u];
    s.v[2] = a.vz[u];
#pragma unroll
    for (int c = 0; c < 6; ++c) s.sg[c] = a.sig[em_index(ee, NS, c, NP, node)];
    s.tau = R(0);
    // lam, mu (, tau): one element-major block, the pointers at its components
    const long mb = em_index(ee, VISCO ? 3 : 2, 0, NP, node);
    if constexpr (S16) {
        s.lam = (R)(float)a.s16.lam[mb];
        s.mu = (R)(float)a.s16.mu[mb];
        if constexpr (VISCO) s.tau = (R)(float)a.s16.tau[mb];
    } else {
        s.lam = a.lam[mb];
        s.mu = a.mu[mb];
        if constexpr (VISCO) s.tau = a.tau[mb];
    }
    if constexpr (VISCO && LOADM) {
#pragma unroll
        for (int c = 0; c < NRDISS * VM; ++c) s.M[c] = a.mem[em_index(ee, NS, c, NP, node)];
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) s.lo[c] = R(0);
#pragma unroll
    for (int q = 0; q < 3; ++q) s.lo2[q] = 0u;
    if constexpr (KS == 2) {
#pragma unroll
        for (int q = 0; q < 3; ++q) s.lo2[q] = a.sig_lo2[em_index(ee, NS, q, NP, node)];
    } else if constexpr (KS == 1) {
#pragma unroll
        for (int c = 0; c < 6; ++c) s.lo[c] = a.sig_lo[em_index(ee, NS, c, NP, node)];
    }
}

// Geometry rows per element (FP64): sin, 1/sin, cot of theta_q; r, 1/r, r^2
// of r_q; the three taper profiles.
enum { G_SIN = 0, G_ISIN, G_COT, G_R, G_IR, G_R2, G_PX, G_PY, G_PZ, G_N };

// Element kernel body: the workgroup advances tile `blk` (EPB elements, one
// lane per element-local node). S16: the static per-node fields are stored
// in BF16 (ElemArgs::s16). KS: stress compensation (see load_tile). LL: the
// kernel variant, 0 planar LDS staging, 1 node-interleaved staging (see
// Stage), 2 planar staging with the new stress, compensation and memory
// variables stored before the weak divergence instead of after it (their
// registers are then free during the divergence, which raises the occupancy
// of some orders and compensation modes); the arithmetic is the same. VM:
// memory words per mechanism, 0 elastic, 6 the six components, 5 their
// deviatoric part (see the stress rate below).
template <int P, int E, typename R, int VM, bool S16, int KS, int LL>
__device__ __forceinline__ void elem_body(const ElemArgs<R> &a, int blk)
{
    using S = Shape<P, E>;
    using C = coef_t<R>;
    constexpr int P1 = S::P1, NP = S::NP, EPB = S::EPB, NE = EPB;
    constexpr int NS = state_components(KS, VM);
    constexpr bool VISCO = VM != 0;
    constexpr int NM = NRDISS * VM;   // memory words per node
    constexpr bool EARLY = LL == 2;
    __shared__ Stage<R, LL == 1 ? 1 : 0, EPB * NP> lds;
    __shared__ C cf[4][P1 * P1];     // Dx, Dy, Dz, dl*w
    __shared__ C cw[P1];             // GLL weights
    __shared__ C gtab[NE][G_N * P1]; // geometry rows of the tile's elements
    __shared__ int sst[NE];          // moment-tensor source slot per element

    const int t = threadIdx.x;
    const int slot = t / NP;
    const int node = t - slot * NP;
    const bool inside = slot < EPB;
    const int i = node % P1, j = (node / P1) % P1, k = node / (P1 * P1);
    const int L = (inside ? slot : 0) * NP;
    const int tile0 = blk;

    // ---- memory variables of the FP32 kernels: streamed straight into LDS
    // (lane t's words at mbuf[c][t]), so they hold no registers until the
    // stress rate; issued before every other load ----
    constexpr bool MDMA = VISCO && sizeof(R) == 4;
    __shared__ R mbuf[MDMA ? NM : 1][MDMA ? S::NT : 1];
    if constexpr (MDMA) {
        // operands in registers first: each asm statement clobbers memory,
        // and argument blocks read through a pointer would be reloaded
        const int e = tile0 * EPB + slot;
        const R *mem = a.mem + em_index((inside && e < a.nel) ? e : 0, NS, 0, NP, node);
        const unsigned base = (unsigned)(size_t)&mbuf[0][(t >> 6) << 6];
#pragma unroll
        for (int c = 0; c < NM; ++c)
            lds_dma_dword(mem + c * NP, base + (unsigned)(c * S::NT * sizeof(R)));
    }

    // ---- tables of the tile: every lane loads from clamped (in-bounds)
    // indices, the lanes owning an entry store it after the state loads are
    // issued, so table and state loads share one round trip ----
    static_assert(P1 * P1 <= S::NT, "coefficient fill needs (p+1)^2 lanes");
    static_assert(NE * P1 <= S::NT, "one geometry entry per lane");
    const int tq = t < P1 * P1 ? t : 0;
    const C tD0 = a.Dx[tq], tD1 = a.Dy[tq], tD2 = a.Dz[tq], tD3 = a.dlw[tq];
    const C tW = a.w[t < P1 ? t : 0];
    const int xg = t < NE * P1 ? t : 0;
    const int ges = xg / P1, gq = xg - ges * P1;
    int ge = tile0 * EPB + ges;
    ge = ge < a.nel ? ge : a.nel - 1;
    const int gex = ge % a.nx, gey = (ge / a.nx) % a.ny, gez = ge / (a.nx * a.ny);
    const ThTab<C> th = a.tth[gex * P1 + gq];
    const RTab<C> rr = a.tr[gez * P1 + gq];
    // the profiles exist (zero) also without a taper
    const double gpx = a.px[gex * P1 + gq], gpy = a.py[gey * P1 + gq], gpz = a.pz[gez * P1 + gq];
    const int se = tile0 * EPB + (t < NE ? t : 0);
    int sslot = -1;
    if (a.src_slot) sslot = a.src_slot[se < a.nel ? se : a.nel - 1];

    // ---- state of the tile ----
    TileState<R, VM> cur;
    load_tile<P, E, R, VM, S16, !MDMA, KS>(a, tile0, cur);

    if (t < P1 * P1) {
        cf[0][t] = tD0;
        cf[1][t] = tD1;
        cf[2][t] = tD2;
        cf[3][t] = tD3;
    }
    if (t < P1) cw[t] = tW;
    if (t < NE * P1) {
        C *g = gtab[ges];
        g[G_SIN * P1 + gq] = th.sin_t;
        g[G_ISIN * P1 + gq] = th.inv_sin;
        g[G_COT * P1 + gq] = th.cot;
        g[G_R * P1 + gq] = rr.r;
        g[G_IR * P1 + gq] = rr.inv_r;
        g[G_R2 * P1 + gq] = rr.r2;
        g[G_PX * P1 + gq] = a.taper_on ? (C)gpx : C(0);
        g[G_PY * P1 + gq] = a.taper_on ? (C)gpy : C(0);
        g[G_PZ * P1 + gq] = a.taper_on ? (C)gpz : C(0);
    }
    if (t < NE) sst[t] = se < a.nel ? sslot : -1;
    // iteration it = *it_dev + 1 uses the STF sample it-1 (zero for it = 1)
    const int it = *a.it_dev + 1;

    {
        const int tile = tile0;
        const int es = inside ? slot : 0;
        const int e = tile * EPB + slot;
        const bool valid = inside && e < a.nel;
        const long eb = (long)(valid ? e : 0) * NP + node;    // element-local address
        const long ev = valid ? e : 0;
        const C *gs = gtab[es];

        if (inside) lds.put_v(L + node, cur.v[0], cur.v[1], cur.v[2]);
        // tables and velocity in LDS
        __syncthreads();

        // ---- K1: tensor-product derivatives along theta, phi, r ----
        C dxv[3] = {0, 0, 0}, dyv[3] = {0, 0, 0}, dzv[3] = {0, 0, 0};
        {
            const C *Dxi = cf[0] + i * P1;
            const C *Dyj = cf[1] + j * P1;
            const C *Dzk = cf[2] + k * P1;
#pragma unroll
            for (int q = 0; q < P1; ++q) {
                const int nxq = L + q + P1 * (j + P1 * k);
                const int nyq = L + i + P1 * (q + P1 * k);
                const int nzq = L + i + P1 * (j + P1 * q);
                const C cx = Dxi[q], cy = Dyj[q], cz = Dzk[q];
                R vx[3], vy[3], vz[3];
                lds.get_v(nxq, vx);
                lds.get_v(nyq, vy);
                lds.get_v(nzq, vz);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    dxv[c] += (C)vx[c] * cx;
                    dyv[c] += (C)vy[c] * cy;
                    dzv[c] += (C)vz[c] * cz;
                }
            }
        }

        // ---- K2: spherical completion, K4: symmetrisation ----
        R exx, eyy, ezz, exy, exz, eyz;
        {
            const C ux = cur.v[0], uy = cur.v[1], uz = cur.v[2];
            const C inv_r = gs[G_IR * P1 + k];
            const C cot = gs[G_COT * P1 + i];
            const C inv_rs = inv_r * gs[G_ISIN * P1 + i];
            const C ezx = -dzv[0];
            const C exz_c = (dxv[2] - ux) * inv_r;
            const C ezy = -dzv[1];
            const C eyz_c = -uy * inv_r + dyv[2] * inv_rs;
            const C exy_c = dxv[1] * inv_r;
            const C eyx = -uy * cot * inv_r + dyv[0] * inv_rs;
            const C gxx = (uz + dxv[0]) * inv_r;
            const C gyy = dyv[1] * inv_rs + (uz + ux * cot) * inv_r;
            const C gzz = -dzv[2];
            const C gxy = (exy_c + eyx) * 0.5;
            const C gxz = (exz_c + ezx) * 0.5;
            const C gyz = (eyz_c + ezy) * 0.5;
            exx = (R)gxx;
            eyy = (R)gyy;
            ezz = (R)gzz;
            exy = (R)gxy;
            exz = (R)gxz;
            eyz = (R)gyz;
            // K3: the displacement gradient integrates the velocity gradient
            // (src/evolution.f90:179-187); kept as its symmetric part, the
            // only form the reference outputs (src/volume_snapshot.f90:220-236),
            // in FP64, and only when requested
            if (a.dgrad && valid) {
                double *g = a.dgrad + em_index(ev, 6, 0, NP, node);
                g[0 * NP] += a.dt * (double)gxx;
                g[1 * NP] += a.dt * (double)gxy;
                g[2 * NP] += a.dt * (double)gxz;
                g[3 * NP] += a.dt * (double)gyy;
                g[4 * NP] += a.dt * (double)gyz;
                g[5 * NP] += a.dt * (double)gzz;
            }
        }

        // ---- K5: stress rate times dt (lam, mu, A, B, C carry the dt) ----
        R rt[6];
        const R tmu = R(2) * cur.mu;
        if constexpr (!VISCO) {
            const R Lr = cur.lam * (exx + eyy + ezz);
            rt[XX] = Lr + tmu * exx;
            rt[YY] = Lr + tmu * eyy;
            rt[ZZ] = Lr + tmu * ezz;
            rt[XY] = tmu * exy;
            rt[XZ] = tmu * exz;
            rt[YZ] = tmu * eyz;
        } else {
            // SLS in the loop form of the commented-out src/evolution.f90:216-284
            if constexpr (MDMA) {
                // this lane's words were written by its own wave's loads
                const int z = wait_vmem_zero();
#pragma unroll
                for (int c = 0; c < NM; ++c) cur.M[c] = mbuf[c][t + z];
            }
            // lam holds kappa - 2 mu_tau / 3 and mu holds mu_tau = mu (1 + tau),
            // both formed in FP64; 2 mu itself only scales the memory variables
            const R tmut = tmu;                                    // 2 mu_tau
            const R tm = tmut / (R(1) + cur.tau);                  // 2 mu
            if constexpr (VM == 6) {
                R Md[6];
#pragma unroll
                for (int c = 0; c < 6; ++c) Md[c] = cur.M[c] + cur.M[6 + c] + cur.M[12 + c];
                const R Msum = Md[XX] + Md[YY] + Md[ZZ];
                const R Lr = cur.lam * (exx + eyy + ezz) - tm * Msum * R(1.0 / 3.0);
                rt[XX] = Lr + tmut * exx + tm * Md[XX];
                rt[YY] = Lr + tmut * eyy + tm * Md[YY];
                rt[ZZ] = Lr + tmut * ezz + tm * Md[ZZ];
                rt[XY] = tmut * exy + tm * Md[XY];
                rt[XZ] = tmut * exz + tm * Md[XZ];
                rt[YZ] = tmut * eyz + tm * Md[YZ];
            } else {
                // The memory variables enter the stress rate only as
                // Md - tr(Md)/3 I, and every component relaxes at the same
                // rate with the same forcing (t3 e), so their deviatoric
                // part evolves on its own, driven by the deviatoric strain
                // rate: five words per mechanism hold it (xx xy xz yy yz,
                // zz = -(xx + yy)), and the trace is never formed.
                R Dd[5];
#pragma unroll
                for (int c = 0; c < 5; ++c) Dd[c] = cur.M[c] + cur.M[5 + c] + cur.M[10 + c];
                const R Dzz = -(Dd[XX] + Dd[YY]);
                const R Lr = cur.lam * (exx + eyy + ezz);
                rt[XX] = Lr + tmut * exx + tm * Dd[XX];
                rt[YY] = Lr + tmut * eyy + tm * Dd[YY];
                rt[ZZ] = Lr + tmut * ezz + tm * Dzz;
                rt[XY] = tmut * exy + tm * Dd[XY];
                rt[XZ] = tmut * exz + tm * Dd[XZ];
                rt[YZ] = tmut * eyz + tm * Dd[YZ];
            }
        }
        if (S16 ? a.s16.ca != nullptr : a.ca != nullptr) {   // anisotropy (src/evolution.f90:290-296)
            R pa, pb, pc;
            if constexpr (S16) {
                pa = (R)(float)a.s16.ca[eb];
                pb = (R)(float)a.s16.cb[eb];
                pc = (R)(float)a.s16.cc[eb];
            } else {
                pa = a.ca[eb];
                pb = a.cb[eb];
                pc = a.cc[eb];
            }
            rt[XX] += pc * ezz + pa * (eyy + exx);
            rt[YY] += pc * ezz + pa * (eyy + exx);
            rt[ZZ] += pc * (eyy + exx);
            rt[XZ] += R(2) * pb * exz;
            rt[YZ] += R(2) * pb * eyz;
        }

        // ---- K6: integrate the stress ----
        R sg[6], lo[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            sg[c] = cur.sg[c];
            lo[c] = cur.lo[c];
        }
        if constexpr (KS != 0) {
            if constexpr (KS == 2) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    lo[2 * q] = (R)bf16_lo(cur.lo2[q]);
                    lo[2 * q + 1] = (R)bf16_hi(cur.lo2[q]);
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) kahan_add(sg[c], lo[c], rt[c]);
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) sg[c] = sg[c] + rt[c];
        }

        // ---- K7: moment-tensor glut and single force of the previous add() ----
        R src[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int sslot = sst[es];
        if (valid && sslot >= 0 && it > 1 && it <= a.nt) {
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                const int ch = a.src_chan[sslot * 9 + c];
                if (ch >= 0) {
                    const double d = a.stf[(long)ch * a.nt + (it - 2)];
                    src[c] = (R)(-(d * a.src_coef[((long)sslot * 9 + c) * NP + node]));
                }
            }
        }

        __syncthreads();   // LDS velocity no longer needed
        if (inside) {
            R se[6];   // effective stress, with the glut
#pragma unroll
            for (int c = 0; c < 6; ++c) se[c] = sg[c] + src[c];
            lds.put_s(L + node, se);
        }

        // ---- memory variables (loop form of src/evolution.f90:415-442) ----
        R M[VISCO ? NM : 1];
        if constexpr (VISCO) {
            // M (1 - dt/tau_p) - dt tau / (3 tau_p) e  ==  M - dt/tau_p (M + tau/3 e):
            // the decrement form keeps the FP32 relaxation rate exact to one
            // rounding of dt/tau_p (1 - dt/tau_p would lose its low bits)
            const R t3 = cur.tau * R(1.0 / 3.0);
            R ev[6] = {exx, exy, exz, eyy, eyz, ezz};
            if constexpr (VM == 5) {   // deviatoric strain rate
                const R tr3 = (exx + eyy + ezz) * R(1.0 / 3.0);
                ev[XX] = exx - tr3;
                ev[YY] = eyy - tr3;
            }
#pragma unroll
            for (int m = 0; m < NRDISS; ++m) {
#pragma unroll
                for (int c = 0; c < VM; ++c)
                    M[VM * m + c] = cur.M[VM * m + c] - a.rate[m] * (cur.M[VM * m + c] + t3 * ev[c]);
            }
        }

        // ---- taper factor of this node (K13) ----
        R dmp = R(0);   // 1 - taper
        {
            const C tx = gs[G_PX * P1 + i], ty = gs[G_PY * P1 + j], tz = gs[G_PZ * P1 + k];
            if (tx + ty + tz != C(0)) dmp = (R)taper_damp(tx, ty, tz, a.alpha);
        }

        // ---- store tapered stress (K13) and memory variables ----
        auto store_state = [&]() {
            if constexpr (KS == 1) {
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    a.sig_w[em_index(ev, NS, c, NP, node)] = sg[c] - sg[c] * dmp;
                    a.sig_lo_w[em_index(ev, NS, c, NP, node)] = lo[c] - lo[c] * dmp;
                }
            } else if constexpr (KS == 2) {
#pragma unroll
                for (int c = 0; c < 6; ++c) a.sig_w[em_index(ev, NS, c, NP, node)] = sg[c] - sg[c] * dmp;
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    a.sig_lo2_w[em_index(ev, NS, q, NP, node)] = pack_bf16((float)(lo[2 * q] - lo[2 * q] * dmp),
                                                       (float)(lo[2 * q + 1] - lo[2 * q + 1] * dmp));
            } else {
#pragma unroll
                for (int c = 0; c < 6; ++c) a.sig_w[em_index(ev, NS, c, NP, node)] = sg[c] - sg[c] * dmp;
            }
            if constexpr (VISCO) {
#pragma unroll
                for (int c = 0; c < NM; ++c) a.mem_w[em_index(ev, NS, c, NP, node)] = M[c];
            }
        };
        if constexpr (EARLY) {
            if (valid) store_state();
        }

        __syncthreads();

        // ---- K8: weak divergence ----
        C ax[P1], ay[P1], az[P1];
        {
            const C *dli = cf[3] + i * P1;
            const C *dlj = cf[3] + j * P1;
            const C *dlk = cf[3] + k * P1;
#pragma unroll
            for (int q = 0; q < P1; ++q) {
                ax[q] = dli[q] * gs[G_SIN * P1 + q];
                ay[q] = dlj[q];
                az[q] = dlk[q] * gs[G_R2 * P1 + q];
            }
        }
        C Sx[3] = {0, 0, 0}, Sy[3] = {0, 0, 0}, Sz[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < P1; ++q) {
            const int nxq = L + q + P1 * (j + P1 * k);
            const int nyq = L + i + P1 * (q + P1 * k);
            const int nzq = L + i + P1 * (j + P1 * q);
            // x contraction: sxx, sxy, sxz feed sx, sy, sz; y: sxy, syy,
            // syz; z: sxz, syz, szz
            R wx[3], wy[3], wz[3];
            lds.get_s(0, nxq, wx);
            lds.get_s(1, nyq, wy);
            lds.get_s(2, nzq, wz);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                Sx[c] += ax[q] * (C)wx[c];
                Sy[c] += ay[q] * (C)wy[c];
                Sz[c] += az[q] * (C)wz[c];
            }
        }
        const C rk = gs[G_R * P1 + k];
        const C Cx = cw[j] * cw[k] * rk * a.d1;
        const C Cy = cw[i] * cw[k] * rk * a.d2;
        const C Cz = cw[i] * cw[j] * gs[G_SIN * P1 + i] * a.d3;

        if (valid) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                a.force[force_index<P>(ev, c, node, a.nel)] = (R)((C)src[6 + c] + Cz * Sz[c] + Cy * Sy[c] + Cx * Sx[c]);

            if constexpr (!EARLY) store_state();
        }
    }
}

// Plane-pair tile order: launch position t runs the tile at position q of
// plane 2m + s, s fastest, so the tile above follows the tile below and
// finds the velocity blocks they share (its own, the lower tile's z+1 face)
// still in L2; in plane order they came back from memory a plane of tiles
// later. A permutation of the tiles (identity where a tile would straddle
// two planes, and for the last plane of an odd count): the arithmetic is
// unchanged. Fewer velocity re-reads from memory; the time gain at
// 192x192x96 is within run-to-run noise (at most about 0.5 %).
template <int EPB>
__device__ inline int plane_pair_tile(int t, int nx, int ny, int nz)
{
    const int pe = nx * ny;
    if (pe % EPB) return t;
    const int TP = pe / EPB;   // tiles per plane
    if (t >= (nz & ~1) * TP) return t;
    const int m = t / (2 * TP), r = t - m * 2 * TP;
    return (2 * m + (r & 1)) * TP + (r >> 1);
}

template <int P, int E, typename R, int VM, bool S16, int KS, int LL>
__global__ void __launch_bounds__((Shape<P, E>::NT))
elem_kernel(const ElemArgs<R> a)
{
    const int t = remap_block(blockIdx.x, gridDim.x, a.remap);
    elem_body<P, E, R, VM, S16, KS, LL>(a, plane_pair_tile<Shape<P, E>::EPB>(t, a.nx, a.ny, a.nz));
}

template <int P, int E, typename R, int VM, int KS, int LL>
__global__ void __launch_bounds__((Shape<P, E>::NT))
elem_kernel_batch(const ElemArgs<R> *__restrict__ args, const int2 *__restrict__ own)
{
    // (simulation, its workgroup index), in XCD-aware order over the whole
    // batch: each XCD works through a contiguous range of simulations
    const int2 o = own[remap_block(blockIdx.x, gridDim.x, 1)];
    if (o.x < 0) return;   // padding workgroup
    const ElemArgs<R> &a = args[o.x];
    elem_body<P, E, R, VM, false, KS, LL>(a, plane_pair_tile<Shape<P, E>::EPB>(o.y, a.nx, a.ny, a.nz));
}

}  // namespace s3d
