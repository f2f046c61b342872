// s3d_api.hip -- C ABI (include/ses3d_hip.h) over the gfx950 kernels.
//
// Host-side setup runs in FP64 with the reference's formulas (cited per
// function); the step precision (FP64 or FP32) only applies to the state and
// to the rounded static tables the kernels read.
#include "../../include/ses3d_hip.h"
#include "s3d_internal.h"
#include "s3d_kernels.hip"
#include "s3d_launch.h"

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace s3d;

// Scratch memory of s3d_vrange_noalloc: one (min, max) pair per 256 unique
// nodes, then the result pair. Each device holds its own copy until the
// process ends; host code reaches it only through the per-device address
// that hipGetSymbolAddress returns (g_scratch), never through this name.
constexpr long S3D_NOALLOC_PAIRS = 1L << 20;   // 16 MiB of FP64 pairs
static __device__ __align__(256) double scratch[2 * (S3D_NOALLOC_PAIRS + 1)];

namespace {

thread_local std::string g_err;

// Per-device address of `scratch`, looked up once per device by s3d_create
// with that device current.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch[MAX_DEVICES] = {};

int fail(const std::string &m)
{
    g_err = m;
    return 1;
}

#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess)                                                      \
            return fail(std::string(#x) + ": " + hipGetErrorString(e_));          \
    } while (0)

constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double EARTH_RADIUS = 6371000.0;   // src/parameters.F90:64
constexpr double TAU_P[NRDISS] = {1000.0, 70.0, 5.25};  // src/geometric_paras.f90:42
constexpr double HUGE64 = 1.79769313486231570815e+308;

// -------------------------------------------------------------------------
// GLL basis (src/gll_mod.f90:61-275): points are -1, 1 and the roots of
// P_N' found by Newton iteration in long double; weights 2/(N(N+1)P_N^2).
// -------------------------------------------------------------------------
void legendre(int n, long double x, long double *pn, long double *dpn)
{
    long double p0 = 1.0L, p1 = x;
    for (int k = 2; k <= n; ++k) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *dpn = (long double)n * (p0 - x * p1) / (1.0L - x * x);
}

int gll_tables(int N, double *knots, double *weights, double *dl)
{
    if (N < 1 || N > 7) return 1;
    long double xs[MAXP1];
    xs[0] = -1.0L;
    xs[N] = 1.0L;
    for (int i = 1; i < N; ++i) {
        long double x = -cosl((long double)PI * i / N);
        for (int it = 0; it < 100; ++it) {
            long double pn, dpn;
            legendre(N, x, &pn, &dpn);
            long double d2 = (2.0L * x * dpn - (long double)N * (N + 1) * pn) / (1.0L - x * x);
            long double step = dpn / d2;
            x -= step;
            if (fabsl(step) < 1e-30L) break;
        }
        xs[i] = x;
    }
    for (int i = 0; i <= N / 2; ++i) {
        long double m = 0.5L * (xs[N - i] - xs[i]);
        xs[i] = -m;
        xs[N - i] = m;
    }
    if (N % 2 == 0) xs[N / 2] = 0.0L;
    for (int i = 0; i <= N; ++i) {
        knots[i] = (double)xs[i];
        long double pn, dpn;
        legendre(N, xs[i], &pn, &dpn);
        weights[i] = (double)(2.0L / ((long double)N * (N + 1) * pn * pn));
    }
    if (N == 1) weights[0] = weights[1] = 0.5;   // reference placeholder values
    if (dl) {
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j) {
                double y;
                if (i == j) {
                    y = 0.0;
                    for (int k = 0; k <= N; ++k)
                        if (k != i) y += 1.0 / (knots[i] - knots[k]);
                } else {
                    double num = 1.0, den = 1.0;
                    for (int k = 0; k <= N; ++k) {
                        if (k == i) continue;
                        if (k != j) num *= knots[j] - knots[k];
                        den *= knots[i] - knots[k];
                    }
                    y = num / den;
                }
                dl[i + (N + 1) * j] = y;
            }
    }
    return 0;
}

// coordinates_1d (src/coordinates_mod.f90:114-141); out[e*P1 + j]
void coords_1d(const double *knots, int p, int elms, double mn, double mx, double *out)
{
    const int P1 = p + 1;
    double delta = (mx - mn) / elms;
    for (int i = 0; i < elms; ++i)
        for (int j = 0; j < P1; ++j) out[i * P1 + j] = mn + ((double)i + 0.5 * (1.0 + knots[j])) * delta;
    out[(elms - 1) * P1 + p] = mx;
}

double deg2rad(double d) { return d * PI / 180.0; }   // coordinate_utilities_mod.f90:74-84

void lagrange_1d(const double *knots, int P1, double x, double *l)
{
    // src/interpolation_mod.f90:83-171
    double w[MAXP1];
    for (int j = 0; j < P1; ++j) {
        double a = 1.0, b = 1.0;
        for (int k = 0; k < j; ++k) a *= knots[j] - knots[k];
        for (int k = j + 1; k < P1; ++k) b *= knots[j] - knots[k];
        w[j] = 1.0 / (a * b);
    }
    for (int j = 0; j < P1; ++j)
        if (x == knots[j]) {
            for (int q = 0; q < P1; ++q) l[q] = q == j ? 1.0 : 0.0;
            return;
        }
    double prod = 1.0;
    for (int k = 0; k < P1; ++k) prod *= x - knots[k];
    for (int j = 0; j < P1; ++j) l[j] = prod / (x - knots[j]) * w[j];
}

}  // namespace

struct Slot {
    int e;
    int chan[9];
    std::vector<double> coef;   // 9 * NP
};

struct s3d_ctx {
    s3d_grid g;
    s3d_options o;
    int p, P1, NP, PB;
    int ns = 6;    // components of the state block
    long nel, nn, nblk, nuniq;
    bool fp32 = false, visco = false, aniso = false, inited = false;
    int kv = 0;                // velocity compensation words: 0 none, 1 FP32, 2 BF16
    int dev = 0;
    hipStream_t stream = nullptr;
    double *stage[7] = {};
    bool q_finite = false, abc_nonzero = false;
    // geometry (host, FP64)
    double knots[MAXP1], weights[MAXP1], dl[MAXP1 * MAXP1];
    double x_glb_min, x_glb_max, y_glb_min, y_glb_max, z_glb_min, z_glb_max;
    double dx, dy, dz, jac, x_loc_min, x_loc_max, y_loc_min, y_loc_max, z_loc_min, z_loc_max;
    std::vector<double> x1, y1, z1, px1, py1, pz1;
    double courant = 0.0;
    // device state
    void *v[3] = {}, *lo[3] = {}, *dt_inv_mm = nullptr, *sig = nullptr, *sig_lo = nullptr,
         *mem = nullptr, *lam = nullptr, *mu = nullptr, *tau = nullptr, *rho = nullptr,
         *ca = nullptr, *cb = nullptr, *cc = nullptr, *force = nullptr, *tth = nullptr, *mat = nullptr,
         *grad_rho = nullptr, *grad_stage = nullptr,   // FP64 density; 3 snapshot components
         *dgrad = nullptr,   // FP64 displacement gradient [nel][6][NP], see s3d_enable_displacement_gradient
         *curl_parts = nullptr,   // FP64 [nel][6][NP], see s3d_capture_curl_strain
         *tr = nullptr, *wdev = nullptr;
    double *mm64 = nullptr, *dpx = nullptr, *dpy = nullptr, *dpz = nullptr;
    double bytes = 0;
    // sources
    std::vector<int> slot_of;
    std::vector<Slot> slots;
    std::vector<double> stf;   // nchan * nt
    int nchan = 0;
    bool src_dirty = false;
    int *d_slot = nullptr, *d_chan = nullptr;
    double *d_coef = nullptr, *d_stf = nullptr;
    // receivers: receiver r records rec_nch[r] channels rec_ch[rec_off[r] + q],
    // traced at rows rec_off[r] + q of the trace buffer [channels][nt]
    std::vector<int> rec_e, rec_nch, rec_off, rec_ch;
    std::vector<double> rec_vm;
    bool rec_dirty = false;
    int *d_rec_e = nullptr, *d_rec_nch = nullptr, *d_rec_off = nullptr, *d_rec_ch = nullptr;
    double *d_rec_vm = nullptr, *d_traces = nullptr;
    // curl receiver channels: slot per receiver (-1 none), element per slot,
    // strain-rate parts [nrot][10][NP]
    int nrot = 0;
    int *d_rec_rot = nullptr, *d_rot_e = nullptr;
    double *d_rot_buf = nullptr;
    int nrec_dev = 0;
    int it = 0;
    int *d_it = nullptr;       // device copy of it, advanced by advance_kernel
    // captured chunk of steps (hipGraph), one per starting state block
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    bool use_graph = true;
    // state blocks: with two (nbuf 2) the element kernel reads blk[cur] and
    // writes blk[1 - cur], and cur flips every step; with one it updates
    // blk[0] in place. sig, sig_lo and mem view blk[cur].
    void *blk[2] = {nullptr, nullptr};
    int nbuf = 2;
    int cur = 0;
    int epb = 0;               // element tile (0 = default), see s3d_autotune
    int lds_layout = 0;        // LDS staging layout of the FP32 element kernel, see Stage
    bool early_stores = false; // FP32 element kernel stores the state before the divergence
    int nm = 6;                // memory words per relaxation mechanism (6, or 5: deviatoric)
    int curl_it = -1;          // steps completed at the last s3d_capture_curl_strain
    bool fields_set = false;   // s3d_set_field was called (state not at rest)
    // s3d_create_shared: the simulation whose static fields (moduli, tau, rho,
    // A/B/C, dt/M, FP64 mass, 1-D tables, taper profiles) this one uses, and
    // on the owner the number of simulations using its fields
    s3d_ctx *model = nullptr;
    int sharers = 0;
    int batches = 0;           // batches this simulation is a member of
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev[3];   // start/stop pairs per kernel class
    double ms[3] = {0, 0, 0};
    int cnt[3] = {0, 0, 0};
};

namespace {

int dalloc(s3d_ctx *c, void **p, size_t bytes, bool zero = true)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail("hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    }
    c->bytes += (double)bytes;
    if (zero) {
        e = hipMemsetAsync(*p, 0, bytes ? bytes : 16, c->stream);
        if (e != hipSuccess) return fail(std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    }
    return 0;
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

size_t rsize(const s3d_ctx *c);

// Per-simulation state: velocity (+ compensation), stress (+ compensation),
// memory variables, element force and the device step counter, zeroed.
// The parts of state block k: stress, compensation words, memory variables.
struct StateParts {
    char *sig, *lo, *mem;
};

StateParts state_parts(const s3d_ctx *c, int k)
{
    const size_t rs = c->fp32 ? sizeof(float) : sizeof(double);
    const int ncl = c->ns - 6 - (c->visco ? NRDISS * c->nm : 0);
    char *b = (char *)c->blk[k];
    return {b, ncl ? b + rs * 6 * c->NP : nullptr, c->visco ? b + rs * (6 + ncl) * c->NP : nullptr};
}

void set_state_block(s3d_ctx *c, int k)
{
    c->cur = k;
    const StateParts s = state_parts(c, k);
    c->sig = s.sig;
    c->sig_lo = s.lo;
    c->mem = s.mem;
}

// the block the next step writes
int next_block(const s3d_ctx *c) { return c->nbuf == 2 ? 1 - c->cur : c->cur; }

// after a step's element kernel: its output block becomes the state
void advance_state(s3d_ctx *c) { set_state_block(c, next_block(c)); }

int alloc_state(s3d_ctx *c)
{
    const size_t rs = rsize(c);
    for (int q = 0; q < 3; ++q) {
        if (dalloc(c, &c->v[q], rs * c->nuniq)) return 1;
        if (c->kv && dalloc(c, &c->lo[q], (c->kv == 2 ? sizeof(bf16) : rs) * c->nuniq)) return 1;
    }
    // stress, its compensation and the memory variables: one element-major
    // block of ns words (FP32: 4-byte words, the BF16 compensation in pairs;
    // FP64 has no compensation); sig_lo and mem point at their first
    // components and are not separate allocations
    c->ns = state_components(c->o.kahan_sigma, c->visco ? c->nm : 0);
    for (int k = 0; k < c->nbuf; ++k)
        if (dalloc(c, &c->blk[k], rs * c->ns * c->nn)) return 1;
    set_state_block(c, 0);
    if (dalloc(c, &c->force, rs * 3 * c->nn, false)) return 1;
    if (dalloc(c, (void **)&c->d_it, sizeof(int))) return 1;
    return 0;
}

s3d_ctx *model_of(s3d_ctx *c) { return c->model ? c->model : c; }

// 1-D colatitude and radius tables of the element kernel, computed in FP64
// and stored in the kernel's coefficient precision C.
template <typename C>
int upload_tables(s3d_ctx *c, const std::vector<double> &sin1)
{
    const int P1 = c->p + 1;
    std::vector<ThTab<C>> tt(c->g.nx * P1);
    std::vector<RTab<C>> tr(c->g.nz * P1);
    for (size_t q = 0; q < tt.size(); ++q) {
        double s = sin1[q];
        tt[q] = {(C)s, (C)(1.0 / s), (C)(std::cos(c->x1[q]) / s), (C)0};
    }
    for (size_t q = 0; q < tr.size(); ++q) {
        double r = c->z1[q];
        tr[q] = {(C)r, (C)(1.0 / r), (C)(r * r), (C)0};
    }
    if (dalloc(c, &c->tth, sizeof(tt[0]) * tt.size())) return 1;
    if (dalloc(c, &c->tr, sizeof(tr[0]) * tr.size())) return 1;
    HIPCHK(hipMemcpy(c->tth, tt.data(), sizeof(tt[0]) * tt.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->tr, tr.data(), sizeof(tr[0]) * tr.size(), hipMemcpyHostToDevice));
    return 0;
}

size_t rsize(const s3d_ctx *c) { return c->fp32 ? sizeof(float) : sizeof(double); }

// the element kernel's variant (elem_body's LL) for the layout and store order
// the velocity compensation arrays as FP32 or BF16 words (null when the
// simulation keeps the other kind or none)
template <typename R>
R *lo32(const s3d_ctx *c, int q) { return c->kv == 1 ? (R *)c->lo[q] : nullptr; }
bf16 *lo16(const s3d_ctx *c, int q) { return c->kv == 2 ? (bf16 *)c->lo[q] : nullptr; }

int kernel_variant(const s3d_ctx *c) { return c->early_stores ? 2 : c->lds_layout; }

inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---- template dispatch helpers ----
// Element kernel arguments for a step that reads state block rd and writes
// block wr (by default the current block and the next step's output block).
template <int P, typename R>
ElemArgs<R> elem_args(s3d_ctx *c, int rd = -1, int wr = -1)
{
    if (rd < 0) rd = c->cur;
    if (wr < 0) wr = next_block(c);
    ElemArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    const int P1 = P + 1;
    a.nx = c->g.nx; a.ny = c->g.ny; a.nz = c->g.nz; a.nel = (int)c->nel;
    a.remap = c->o.xcd_remap;
    a.it_dev = (decltype(a.it_dev))(c->d_it);
    a.nt = c->g.nt;
    a.kahan_sigma = c->o.kahan_sigma;
    a.vx = (decltype(a.vx))((const R *)c->v[0]); a.vy = (decltype(a.vy))((const R *)c->v[1]); a.vz = (decltype(a.vz))((const R *)c->v[2]);
    const StateParts in = state_parts(c, rd), out = state_parts(c, wr);
    a.sig = (decltype(a.sig))((const R *)in.sig);
    a.sig_w = (decltype(a.sig_w))((R *)out.sig);
    if (c->o.kahan_sigma == 2) {
        a.sig_lo2 = (decltype(a.sig_lo2))((const unsigned *)in.lo);
        a.sig_lo2_w = (decltype(a.sig_lo2_w))((unsigned *)out.lo);
    } else {
        a.sig_lo = (decltype(a.sig_lo))((const R *)in.lo);
        a.sig_lo_w = (decltype(a.sig_lo_w))((R *)out.lo);
    }
    a.mem = (decltype(a.mem))((const R *)in.mem);
    a.mem_w = (decltype(a.mem_w))((R *)out.mem);
    if (c->o.bf16_static) {
        a.s16.lam = (decltype(a.s16.lam))((const bf16 *)c->lam); a.s16.mu = (decltype(a.s16.mu))((const bf16 *)c->mu); a.s16.tau = (decltype(a.s16.tau))((const bf16 *)c->tau);
        a.s16.ca = (decltype(a.s16.ca))((const bf16 *)c->ca); a.s16.cb = (decltype(a.s16.cb))((const bf16 *)c->cb); a.s16.cc = (decltype(a.s16.cc))((const bf16 *)c->cc);
    } else {
        a.lam = (decltype(a.lam))((const R *)c->lam); a.mu = (decltype(a.mu))((const R *)c->mu); a.tau = (decltype(a.tau))((const R *)c->tau);
        a.ca = (decltype(a.ca))((const R *)c->ca); a.cb = (decltype(a.cb))((const R *)c->cb); a.cc = (decltype(a.cc))((const R *)c->cc);
    }
    a.force = (decltype(a.force))((R *)c->force);
    a.dgrad = (decltype(a.dgrad))(c->dgrad);
    a.dt = c->g.dt;
    using C = coef_t<R>;
    a.tth = (decltype(a.tth))((const ThTab<C> *)c->tth);
    a.tr = (decltype(a.tr))((const RTab<C> *)c->tr);
    a.px = (decltype(a.px))(c->dpx); a.py = (decltype(a.py))(c->dpy); a.pz = (decltype(a.pz))(c->dpz);
    a.alpha = c->g.taper_slope;
    a.taper_on = c->g.taper_width > 0;
    a.src_slot = (decltype(a.src_slot))(c->d_slot);
    a.src_chan = (decltype(a.src_chan))(c->d_chan);
    a.src_coef = (decltype(a.src_coef))(c->d_coef);
    a.stf = (decltype(a.stf))(c->d_stf);
    a.d1 = (C)((2.0 / c->dx) * c->jac);
    a.d2 = (C)((2.0 / c->dy) * c->jac);
    a.d3 = (C)(-((2.0 / c->dz) * c->jac));
    for (int m = 0; m < NRDISS; ++m) a.rate[m] = (R)(c->g.dt / TAU_P[m]);
    for (int q = 0; q < P1; ++q) a.w[q] = (C)c->weights[q];
    for (int i = 0; i < P1; ++i)
        for (int q = 0; q < P1; ++q) {
            // dl(q,i) = l_q'(xi_i) is stored at c->dl[q + P1*i]
            a.Dx[i * P1 + q] = (C)(2.0 * c->dl[q + P1 * i] / c->dx);
            a.Dy[i * P1 + q] = (C)(2.0 * c->dl[q + P1 * i] / c->dy);
            a.Dz[i * P1 + q] = (C)(2.0 * c->dl[q + P1 * i] / c->dz);
            a.dlw[i * P1 + q] = (C)(c->dl[i + P1 * q] * c->weights[q]);
        }
    return a;
}

template <int P, typename R>
NodeArgs<R> node_args(s3d_ctx *c)
{
    NodeArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    a.nx = c->g.nx; a.ny = c->g.ny; a.nz = c->g.nz; a.nel = (int)c->nel;
    a.nblk = c->nblk;
    a.vx = (decltype(a.vx))((R *)c->v[0]); a.vy = (decltype(a.vy))((R *)c->v[1]); a.vz = (decltype(a.vz))((R *)c->v[2]);
    a.lx = (decltype(a.lx))(lo32<R>(c, 0)); a.ly = (decltype(a.ly))(lo32<R>(c, 1)); a.lz = (decltype(a.lz))(lo32<R>(c, 2));
    a.hx = (decltype(a.hx))(lo16(c, 0)); a.hy = (decltype(a.hy))(lo16(c, 1)); a.hz = (decltype(a.hz))(lo16(c, 2));
    a.force = (decltype(a.force))((const R *)c->force);
    if (c->o.bf16_static)
        a.dt_inv_mm16 = (decltype(a.dt_inv_mm16))((const bf16 *)c->dt_inv_mm);
    else
        a.dt_inv_mm = (decltype(a.dt_inv_mm))((const R *)c->dt_inv_mm);
    a.px = (decltype(a.px))(c->dpx); a.py = (decltype(a.py))(c->dpy); a.pz = (decltype(a.pz))(c->dpz);
    a.alpha = c->g.taper_slope;
    a.taper_on = c->g.taper_width > 0;
    return a;
}

template <typename R>
RecArgs<R> rec_args(s3d_ctx *c)
{
    RecArgs<R> a;
    a.nx = c->g.nx; a.ny = c->g.ny; a.nel = (int)c->nel; a.nt = c->g.nt;
    a.vx = (decltype(a.vx))((const R *)c->v[0]); a.vy = (decltype(a.vy))((const R *)c->v[1]); a.vz = (decltype(a.vz))((const R *)c->v[2]);
    a.lx = (decltype(a.lx))(lo32<R>(c, 0)); a.ly = (decltype(a.ly))(lo32<R>(c, 1)); a.lz = (decltype(a.lz))(lo32<R>(c, 2));
    a.hx = (decltype(a.hx))(lo16(c, 0)); a.hy = (decltype(a.hy))(lo16(c, 1)); a.hz = (decltype(a.hz))(lo16(c, 2));
    a.sig = (decltype(a.sig))((const R *)c->sig);
    a.ns = c->ns;
    a.rec_e = (decltype(a.rec_e))(c->d_rec_e); a.rec_nch = (decltype(a.rec_nch))(c->d_rec_nch); a.rec_ch = (decltype(a.rec_ch))(c->d_rec_ch);
    a.rec_off = (decltype(a.rec_off))(c->d_rec_off);
    a.src_slot = (decltype(a.src_slot))(c->d_slot);
    a.src_chan = (decltype(a.src_chan))(c->d_chan);
    a.src_coef = (decltype(a.src_coef))(c->d_coef);
    a.stf = (decltype(a.stf))(c->d_stf);
    a.rec_vm = (decltype(a.rec_vm))(c->d_rec_vm);
    a.rec_rot = (decltype(a.rec_rot))(c->nrot > 0 ? c->d_rec_rot : nullptr);
    a.rot_buf = (decltype(a.rot_buf))(c->d_rot_buf);
    a.traces = (decltype(a.traces))(c->d_traces);
    a.it_dev = (decltype(a.it_dev))(c->d_it);
    return a;
}

void ev_begin(s3d_ctx *c, int k)
{
    if (!c->timing) return;
    hipEvent_t e;
    if (hipEventCreate(&e) == hipSuccess) {
        (void)hipEventRecord(e, c->stream);
        c->ev[k].push_back(e);
    }
}
void ev_end(s3d_ctx *c, int k) { ev_begin(c, k); }

// One element-kernel launch; the instantiation (tile, memory words, stress
// compensation, kernel variant, BF16 statics) is chosen in s3d_elem.hip.
template <int P, typename R>
int launch_elem(s3d_ctx *c, const ElemArgs<R> &ea, hipStream_t st, int epb)
{
    ElemLaunch l;
    l.nel = c->nel;
    l.vm = c->visco ? c->nm : 0;
    l.ks = c->o.kahan_sigma;
    l.ll = kernel_variant(c);
    l.bf16_static = c->o.bf16_static;
    l.epb = epb;
    const hipError_t e = launch_elem_kernel<P, R>(l, ea, st);
    if (e != hipSuccess) return fail(std::string("element kernel launch: ") + hipGetErrorString(e));
    return 0;
}

// Element tiles offered to the autotuner: kernels exist for 1, 2 and 4
// elements per workgroup (where that fits 1024 lanes) besides the default;
// BF16-static uses the default tile only.
template <int P>
std::vector<int> tile_candidates(bool bf16_static)
{
    std::vector<int> v = {0};
    if (bf16_static) return v;
    for (int e : {1, 2, 4})
        if (e * (P + 1) * (P + 1) * (P + 1) <= 1024 && e != default_epb<P>()) v.push_back(e);
    return v;
}

void drop_graph(s3d_ctx *c);

// Time every tile variant x block order on this mesh (state at rest, so
// the trial launches write zeros over zeros) and keep the fastest.
template <int P, typename R>
int autotune_p(s3d_ctx *c, int *tile_out, double *ms_out)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    int best_tile = c->epb, best_remap = c->o.xcd_remap, best_ll = kernel_variant(c);
    const int nll = (c->fp32 && !c->o.bf16_static) ? 3 : 1;
    for (int tile : tile_candidates<P>(c->o.bf16_static))
        for (int ll = 0; ll < nll; ++ll)
            for (int remap = 0; remap < 2; ++remap) {
                c->o.xcd_remap = remap;
                c->lds_layout = ll == 1 ? 1 : 0;
                c->early_stores = ll == 2;
                ElemArgs<R> ea = elem_args<P, R>(c);
                if (int rc = launch_elem<P, R>(c, ea, c->stream, tile)) return rc;   // warm-up
                HIPCHK(hipEventRecord(e0, c->stream));
                for (int r = 0; r < 3; ++r)
                    if (int rc = launch_elem<P, R>(c, ea, c->stream, tile)) return rc;
                HIPCHK(hipEventRecord(e1, c->stream));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (ms / 3.0 < best) {
                    best = ms / 3.0;
                    best_tile = tile;
                    best_remap = remap;
                    best_ll = ll;
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->epb = best_tile;
    c->o.xcd_remap = best_remap;
    c->lds_layout = best_ll == 1 ? 1 : 0;
    c->early_stores = best_ll == 2;
    drop_graph(c);
    if (tile_out) *tile_out = best_tile == 0 ? default_epb<P>() : best_tile;
    if (ms_out) *ms_out = best;
    return 0;
}

template <int P, typename R>
CurlArgs<R> curl_args(s3d_ctx *c)
{
    CurlArgs<R> a;
    const int P1 = P + 1;
    a.nx = c->g.nx; a.ny = c->g.ny; a.nrot = c->nrot;
    a.vx = (decltype(a.vx))((const R *)c->v[0]); a.vy = (decltype(a.vy))((const R *)c->v[1]); a.vz = (decltype(a.vz))((const R *)c->v[2]);
    a.lx = (decltype(a.lx))(lo32<R>(c, 0)); a.ly = (decltype(a.ly))(lo32<R>(c, 1)); a.lz = (decltype(a.lz))(lo32<R>(c, 2));
    a.hx = (decltype(a.hx))(lo16(c, 0)); a.hy = (decltype(a.hy))(lo16(c, 1)); a.hz = (decltype(a.hz))(lo16(c, 2));
    a.tth = (decltype(a.tth))((const ThTab<double> *)c->tth);
    a.tr = (decltype(a.tr))((const RTab<double> *)c->tr);
    a.rot_e = (decltype(a.rot_e))(c->d_rot_e);
    a.buf = (decltype(a.buf))(c->d_rot_buf);
    for (int i = 0; i < P1; ++i)
        for (int q = 0; q < P1; ++q) {
            a.Dx[i * P1 + q] = 2.0 * c->dl[q + P1 * i] / c->dx;
            a.Dy[i * P1 + q] = 2.0 * c->dl[q + P1 * i] / c->dy;
            a.Dz[i * P1 + q] = 2.0 * c->dl[q + P1 * i] / c->dz;
        }
    return a;
}

template <int P, typename R>
int launch_step(s3d_ctx *c)
{
    if (c->nrot > 0) {
        // before the node kernel: the curl channels use the start velocity
        hipLaunchKernelGGL((curl_strain_kernel<P, R>), dim3(c->nrot), dim3(64), 0, c->stream,
                           curl_args<P, R>(c));
        HIPCHK(hipGetLastError());
    }
    ElemArgs<R> ea = elem_args<P, R>(c);
    ev_begin(c, 0);
    if (int rc = launch_elem<P, R>(c, ea, c->stream, c->epb)) return rc;
    ev_end(c, 0);
    advance_state(c);   // the receivers read the stress this step wrote
    NodeArgs<R> na = node_args<P, R>(c);
    const unsigned gn = nblocks(c->nuniq, 256);
    ev_begin(c, 1);
    if (c->kv == 1)
        hipLaunchKernelGGL((node_kernel<P, R, 1>), dim3(gn), dim3(256), 0, c->stream, na);
    else if (c->kv == 2)
        hipLaunchKernelGGL((node_kernel<P, R, 2>), dim3(gn), dim3(256), 0, c->stream, na);
    else
        hipLaunchKernelGGL((node_kernel<P, R, 0>), dim3(gn), dim3(256), 0, c->stream, na);
    ev_end(c, 1);
    HIPCHK(hipGetLastError());
    if (c->nrec_dev > 0) {
        ev_begin(c, 2);
        hipLaunchKernelGGL((record_kernel<P, R>), dim3(c->nrec_dev), dim3(64), 0, c->stream,
                           rec_args<R>(c));
        ev_end(c, 2);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, c->stream, c->d_it);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename R>
int dispatch_step(s3d_ctx *c)
{
    switch (c->p) {
    case 2: return launch_step<2, R>(c);
    case 3: return launch_step<3, R>(c);
    case 4: return launch_step<4, R>(c);
    case 5: return launch_step<5, R>(c);
    case 6: return launch_step<6, R>(c);
    case 7: return launch_step<7, R>(c);
    default: return fail("unsupported lpd");
    }
}

int launch_one(s3d_ctx *c) { return c->fp32 ? dispatch_step<float>(c) : dispatch_step<double>(c); }

void drop_graph_slot(s3d_ctx *c, int k)
{
    if (c->graph_exec[k]) (void)hipGraphExecDestroy(c->graph_exec[k]);
    if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
    c->graph_exec[k] = nullptr;
    c->graph[k] = nullptr;
    c->graph_steps[k] = 0;
}

void drop_graph(s3d_ctx *c)
{
    drop_graph_slot(c, 0);
    drop_graph_slot(c, 1);
}

// Capture `chunk` (even) consecutive iterations into one graph, for the
// current state block: an even number of steps ends on the block it started
// from, so the graph replays whenever that block is current. The kernels
// read the iteration number from the device counter, so the same graph
// replays any chunk of the run.
int capture_graph(s3d_ctx *c, int chunk)
{
    const int k = c->cur;
    drop_graph_slot(c, k);
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int s = 0; s < chunk && rc == 0; ++s) rc = launch_one(c);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        set_state_block(c, k);
        return rc;
    }
    if (e != hipSuccess) {
        set_state_block(c, k);
        return fail(std::string("graph capture: ") + hipGetErrorString(e));
    }
    c->graph[k] = g;
    e = hipGraphInstantiate(&c->graph_exec[k], g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        drop_graph_slot(c, k);
        return fail(std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    c->graph_steps[k] = chunk;
    return 0;
}

// Generic P-dispatch for setup / IO kernels.
template <template <int, typename> class F, typename R, typename... A>
int with_p(int p, A &&...args)
{
    switch (p) {
    case 2: return F<2, R>::run(args...);
    case 3: return F<3, R>::run(args...);
    case 4: return F<4, R>::run(args...);
    case 5: return F<5, R>::run(args...);
    case 6: return F<6, R>::run(args...);
    case 7: return F<7, R>::run(args...);
    default: return fail("unsupported lpd");
    }
}

template <int P, typename R>
struct MassRun {
    static int run(s3d_ctx *c, const double *sin_t, const double *rad, const double *w)
    {
        if (c->o.bf16_static)
            hipLaunchKernelGGL((mass_kernel<P, bf16>), dim3(nblocks(c->nuniq, 256)), dim3(256), 0, c->stream,
                               c->stage[S3D_RHOINV], sin_t, rad, w, c->jac, c->g.dt, c->mm64,
                               (bf16 *)c->dt_inv_mm, c->g.nx, c->g.ny, c->g.nz, c->nblk);
        else
            hipLaunchKernelGGL((mass_kernel<P, R>), dim3(nblocks(c->nuniq, 256)), dim3(256), 0, c->stream,
                               c->stage[S3D_RHOINV], sin_t, rad, w, c->jac, c->g.dt, c->mm64,
                               (R *)c->dt_inv_mm, c->g.nx, c->g.ny, c->g.nz, c->nblk);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct CflRun {
    static int run(s3d_ctx *c, const double *th, const double *ph, const double *rad, double *out)
    {
        unsigned nb = nblocks(c->nn, 256);
        double *part = nullptr;
        HIPCHK(hipMalloc(&part, sizeof(double) * 3 * nb));
        hipLaunchKernelGGL((cfl_kernel<P>), dim3(nb), dim3(256), 0, c->stream, c->stage[S3D_RHOINV],
                           c->stage[S3D_LAMBDA], c->stage[S3D_MU], th, ph, rad, c->g.nx, c->g.ny,
                           c->g.nz, part);
        std::vector<double> h(3 * (size_t)nb);
        hipError_t e = hipMemcpyAsync(h.data(), part, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(part);
        if (e != hipSuccess) return fail(std::string("cfl: ") + hipGetErrorString(e));
        double m[3] = {0, 0, 0};
        for (unsigned b = 0; b < nb; ++b)
            for (int q = 0; q < 3; ++q) m[q] = std::max(m[q], h[3 * (size_t)b + q]);
        double dt = c->g.dt;
        *out = std::max(m[0] * dt, std::max(m[1] * dt, m[2] * dt));
        return 0;
    }
};

template <int P, typename R>
struct ExpandRun {
    static int run(s3d_ctx *c, const void *src, const void *lo, double *dst)
    {
        hipLaunchKernelGGL((expand_unique<P, R>), dim3(nblocks(c->nn, 256)), dim3(256), 0, c->stream,
                           (const R *)src, (const R *)(c->kv == 1 ? lo : nullptr),
                           (const bf16 *)(c->kv == 2 ? lo : nullptr),
                           dst, c->g.nx, c->g.ny, c->g.nz);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct CurlStrainRun {
    static int run(s3d_ctx *c, hipStream_t st)
    {
        hipLaunchKernelGGL((curl_strain_kernel<P, R>), dim3(c->nrot), dim3(64), 0, st, curl_args<P, R>(c));
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct CurlPartsRun {
    static int run(s3d_ctx *c)
    {
        hipLaunchKernelGGL((curl_parts_kernel<P, R>), dim3((unsigned)c->nel), dim3(64), 0, c->stream,
                           curl_args<P, R>(c), c->nel, (double *)c->curl_parts);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct CurlFieldRun {
    static int run(s3d_ctx *c, int comp, double *out)
    {
        hipLaunchKernelGGL((curl_field_kernel<P, R>), dim3(nblocks(c->nn, 256)), dim3(256), 0, c->stream,
                           curl_args<P, R>(c), c->g.nz, c->nn, (const double *)c->curl_parts, comp, out);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct GradRun {
    static int run(s3d_ctx *c, const double *f, double inc_dt)
    {
        const long nn = c->nn;
        hipLaunchKernelGGL((grad_rho_kernel<P, R>), dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                           (double *)c->grad_rho, f, f + nn, f + 2 * nn, (const R *)c->v[0],
                           (const R *)c->v[1], (const R *)c->v[2], (const R *)lo32<R>(c, 0),
                           (const R *)lo32<R>(c, 1), (const R *)lo32<R>(c, 2), (const bf16 *)lo16(c, 0),
                           (const bf16 *)lo16(c, 1), (const bf16 *)lo16(c, 2), c->g.nx, c->g.ny, c->g.nz,
                           inc_dt, c->jac);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct GatherRun {
    static int run(s3d_ctx *c, const double *src, void *dst, void *lo)
    {
        hipLaunchKernelGGL((gather_unique<P, R>), dim3(nblocks(c->nuniq, 256)), dim3(256), 0, c->stream,
                           src, (R *)dst, (R *)(c->kv == 1 ? lo : nullptr), (bf16 *)(c->kv == 2 ? lo : nullptr),
                           c->g.nx, c->g.ny, c->g.nz, c->nblk);
        HIPCHK(hipGetLastError());
        return 0;
    }
};

template <int P, typename R>
struct VrangeRun {
    static int run(s3d_ctx *c, double *out)
    {
        unsigned nb = nblocks(c->nuniq, 256);
        double *part = nullptr;
        HIPCHK(hipMalloc(&part, sizeof(double) * 2 * nb));
        hipLaunchKernelGGL((vrange_kernel<P, R>), dim3(nb), dim3(256), 0, c->stream, (const R *)c->v[0],
                           (const R *)c->v[1], (const R *)c->v[2], c->g.nx, c->g.ny, c->g.nz, c->nblk, part);
        std::vector<double> h(2 * (size_t)nb);
        hipError_t e = hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(part);
        if (e != hipSuccess) return fail(std::string("vrange: ") + hipGetErrorString(e));
        double mn = 1e308, mx = -1e308;
        for (unsigned b = 0; b < nb; ++b) {
            mn = std::min(mn, h[2 * (size_t)b]);
            mx = std::max(mx, h[2 * (size_t)b + 1]);
        }
        out[0] = mn;
        out[1] = mx;
        return 0;
    }
};

// s3d_vrange_noalloc: partials and the result pair in the device's scratch
template <int P, typename R>
struct VrangeNoallocRun {
    static int run(s3d_ctx *c, double *part, unsigned nb, double *out)
    {
        hipLaunchKernelGGL((vrange_kernel<P, R>), dim3(nb), dim3(256), 0, c->stream, (const R *)c->v[0],
                           (const R *)c->v[1], (const R *)c->v[2], c->g.nx, c->g.ny, c->g.nz, c->nblk, part);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(range_reduce_kernel, dim3(1), dim3(256), 0, c->stream, (const double *)part,
                           (int)nb, part + 2 * (size_t)nb);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, part + 2 * (size_t)nb, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    }
};

template <int P, typename R>
struct KineticRun {
    static int run(s3d_ctx *c, double *out)
    {
        unsigned nb = nblocks(c->nn, 256);
        double *part = nullptr;
        HIPCHK(hipMalloc(&part, sizeof(double) * nb));
        hipLaunchKernelGGL((kinetic_kernel<P, R>), dim3(nb), dim3(256), 0, c->stream, (const R *)c->v[0],
                           (const R *)c->v[1], (const R *)c->v[2], (const R *)c->rho,
                           (const ThTab<coef_t<R>> *)c->tth, (const RTab<coef_t<R>> *)c->tr,
                           (const double *)c->wdev,
                           c->g.nx, c->g.ny, c->g.nz, part);
        std::vector<double> h(nb);
        hipError_t e = hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(part);
        if (e != hipSuccess) return fail(std::string("kinetic: ") + hipGetErrorString(e));
        double s = 0.0;
        for (unsigned b = 0; b < nb; ++b) s += h[b];
        *out = 0.5 * c->jac * s;
        return 0;
    }
};

void drop_graph(s3d_ctx *c);

int upload_sources(s3d_ctx *c)
{
    if (!c->src_dirty) return 0;
    drop_graph(c);
    dfree(c->d_slot); dfree(c->d_chan); dfree(c->d_coef); dfree(c->d_stf);
    c->d_slot = nullptr; c->d_chan = nullptr; c->d_coef = nullptr; c->d_stf = nullptr;
    if (!c->slots.empty()) {
        std::vector<int> chan(9 * c->slots.size());
        std::vector<double> coef(9 * (size_t)c->NP * c->slots.size());
        for (size_t s = 0; s < c->slots.size(); ++s) {
            for (int q = 0; q < 9; ++q) chan[9 * s + q] = c->slots[s].chan[q];
            std::memcpy(&coef[9 * (size_t)c->NP * s], c->slots[s].coef.data(), sizeof(double) * 9 * c->NP);
        }
        HIPCHK(hipMalloc(&c->d_slot, sizeof(int) * c->nel));
        HIPCHK(hipMalloc(&c->d_chan, sizeof(int) * chan.size()));
        HIPCHK(hipMalloc(&c->d_coef, sizeof(double) * coef.size()));
        HIPCHK(hipMalloc(&c->d_stf, sizeof(double) * std::max<size_t>(1, c->stf.size())));
        HIPCHK(hipMemcpy(c->d_slot, c->slot_of.data(), sizeof(int) * c->nel, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_chan, chan.data(), sizeof(int) * chan.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice));
        if (!c->stf.empty())
            HIPCHK(hipMemcpy(c->d_stf, c->stf.data(), sizeof(double) * c->stf.size(), hipMemcpyHostToDevice));
    }
    c->src_dirty = false;
    return 0;
}

int upload_receivers(s3d_ctx *c)
{
    if (!c->rec_dirty) return 0;
    drop_graph(c);
    const int n = (int)c->rec_e.size();
    const int nchan = (int)c->rec_ch.size();   // channels of all receivers
    const size_t tr = (size_t)nchan * c->g.nt;
    double *traces = nullptr;
    HIPCHK(hipMalloc(&traces, sizeof(double) * std::max<size_t>(1, tr)));
    HIPCHK(hipMemsetAsync(traces, 0, sizeof(double) * std::max<size_t>(1, tr), c->stream));
    // receivers are only appended: the rows recorded so far come first
    if (c->d_traces && c->nrec_dev > 0) {
        const size_t old_rows = (size_t)c->rec_off[c->nrec_dev - 1] + c->rec_nch[c->nrec_dev - 1];
        HIPCHK(hipMemcpyAsync(traces, c->d_traces, sizeof(double) * old_rows * c->g.nt,
                              hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    dfree(c->d_traces); dfree(c->d_rec_e); dfree(c->d_rec_nch); dfree(c->d_rec_off); dfree(c->d_rec_ch);
    dfree(c->d_rec_vm);
    c->d_traces = traces;
    HIPCHK(hipMalloc(&c->d_rec_e, sizeof(int) * n));
    HIPCHK(hipMalloc(&c->d_rec_nch, sizeof(int) * n));
    HIPCHK(hipMalloc(&c->d_rec_off, sizeof(int) * n));
    HIPCHK(hipMalloc(&c->d_rec_ch, sizeof(int) * std::max(1, nchan)));
    HIPCHK(hipMalloc(&c->d_rec_vm, sizeof(double) * c->NP * (size_t)n));
    HIPCHK(hipMemcpy(c->d_rec_e, c->rec_e.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rec_nch, c->rec_nch.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rec_off, c->rec_off.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rec_ch, c->rec_ch.data(), sizeof(int) * nchan, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_rec_vm, c->rec_vm.data(), sizeof(double) * c->NP * (size_t)n, hipMemcpyHostToDevice));
    // curl receiver channels
    dfree(c->d_rec_rot); dfree(c->d_rot_e); dfree(c->d_rot_buf);
    c->d_rec_rot = c->d_rot_e = nullptr;
    c->d_rot_buf = nullptr;
    std::vector<int> slot(n, -1), rot_e;
    for (int r = 0; r < n; ++r)
        for (int q = 0; q < c->rec_nch[r]; ++q) {
            const int ch = c->rec_ch[c->rec_off[r] + q];
            if (ch >= S3D_CH_ROT_VX && ch <= S3D_CH_ROT_VZ && slot[r] < 0) {
                slot[r] = (int)rot_e.size();
                rot_e.push_back(c->rec_e[r]);
            }
        }
    c->nrot = (int)rot_e.size();
    if (c->nrot > 0) {
        HIPCHK(hipMalloc(&c->d_rec_rot, sizeof(int) * n));
        HIPCHK(hipMalloc(&c->d_rot_e, sizeof(int) * c->nrot));
        HIPCHK(hipMalloc(&c->d_rot_buf, sizeof(double) * 10 * c->NP * (size_t)c->nrot));
        HIPCHK(hipMemcpy(c->d_rec_rot, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_rot_e, rot_e.data(), sizeof(int) * c->nrot, hipMemcpyHostToDevice));
    }
    c->nrec_dev = n;
    c->rec_dirty = false;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batches: simulations of one device advanced together. Simulations with the
// same (order, precision, memory-variable form, compensation) form a group;
// each group costs
// one element, one node and one receiver launch per step, with per-simulation
// arguments in device arrays and workgroups partitioned by prefix sums.
// ---------------------------------------------------------------------------
struct BatchGroup {
    int p = 0;
    int epb = 0;
    int ll = 0;                 // element kernel variant (elem_body's LL)
    int ks = 0;                 // stress compensation mode
    int vm = 0;                 // memory words per mechanism (0 elastic, 6, 5)
    int kv = 0;                 // velocity compensation words (0, 1 FP32, 2 BF16)
    bool fp32 = false, visco = false;
    std::vector<s3d_ctx *> sims;
    // element and receiver arguments in two sets of sims.size() entries: set
    // `phase` serves the batch's even (0) or odd (1) steps since it was
    // built, when members with two state buffers read and write the other
    // block of their pair
    void *d_eargs = nullptr, *d_nargs = nullptr, *d_rargs = nullptr;
    int phase = 0;
    // per workgroup of each batched launch: (simulation, its workgroup index)
    int2 *d_eown = nullptr, *d_nown = nullptr, *d_rown = nullptr;
    int ewg = 0, nwg = 0, rwg = 0;
};

struct s3d_batch {
    std::vector<s3d_ctx *> sims;
    std::vector<BatchGroup> groups;
    int dev = 0;
    hipStream_t stream = nullptr;
    int **d_its = nullptr;
    bool built = false;
    // captured chunk of steps, one per phase (see BatchGroup)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    int phase = 0;
    hipEvent_t ev = nullptr;
};

namespace {

template <int P, int E, typename R>
int batch_build_group_e(BatchGroup &G)
{
    using S = Shape<P, E>;
    const int n = (int)G.sims.size();
    std::vector<ElemArgs<R>> ea(2 * n);
    std::vector<NodeArgs<R>> na(n);
    std::vector<RecArgs<R>> ra(2 * n);
    std::vector<int> es(n + 1, 0), ns(n + 1, 0), rs(n + 1, 0);
    for (int q = 0; q < n; ++q) {
        s3d_ctx *c = G.sims[q];
        // even steps read the current block and write the next one; odd
        // steps read that one and write the block after it
        const int r0 = c->cur, w0 = next_block(c);
        const int r1 = w0, w1 = c->nbuf == 2 ? r0 : w0;
        ea[q] = elem_args<P, R>(c, r0, w0);
        ea[n + q] = elem_args<P, R>(c, r1, w1);
        ea[q].remap = ea[n + q].remap = 0;
        na[q] = node_args<P, R>(c);
        for (int h = 0; h < 2; ++h) {
            RecArgs<R> &r = ra[h * n + q];
            r = rec_args<R>(c);
            // the receivers read the stress their step wrote
            r.sig = (decltype(r.sig))((const R *)state_parts(c, h ? w1 : w0).sig);
        }
        es[q + 1] = es[q] + (int)nblocks(c->nel, S::EPB);
        ns[q + 1] = ns[q] + (int)nblocks(c->nuniq, 256);
        rs[q + 1] = rs[q] + c->nrec_dev;
    }
    G.ewg = es[n];
    G.nwg = ns[n];
    G.rwg = rs[n];
    // Simulations that share one model (s3d_create_shared) run
    // source-innermost: the workgroups of one element tile (or node range)
    // for every member follow each other, so the shared moduli, tau and dt/M
    // come from HBM once per tile and from cache for the other members.
    bool interleave = n > 1;
    for (int q = 1; q < n && interleave; ++q) interleave = model_of(G.sims[q]) == model_of(G.sims[0]);
    // the element launch is padded to a multiple of 8 workgroups (idle ones
    // own simulation -1) so its XCD-aware block order applies
    auto owners = [n](const std::vector<int> &st, int pad8, bool inter) {
        const int total = pad8 ? (st[n] + 7) / 8 * 8 : st[n];
        std::vector<int2> o(total, make_int2(-1, 0));
        if (inter) {
            const int per = st[1] - st[0];   // members share the mesh and tile
            for (int w = 0; w < per; ++w)
                for (int q = 0; q < n; ++q) o[(size_t)w * n + q] = make_int2(q, w);
        } else {
            for (int q = 0; q < n; ++q)
                for (int w = st[q]; w < st[q + 1]; ++w) o[w] = make_int2(q, w - st[q]);
        }
        return o;
    };
    // receivers differ between members: their launch stays member-major
    const std::vector<int2> eo = owners(es, 1, interleave), no = owners(ns, 0, interleave),
                            ro = owners(rs, 0, false);
    G.ewg = (int)eo.size();
    HIPCHK(hipMalloc(&G.d_eargs, sizeof(ElemArgs<R>) * 2 * n));
    HIPCHK(hipMalloc(&G.d_nargs, sizeof(NodeArgs<R>) * n));
    HIPCHK(hipMalloc(&G.d_rargs, sizeof(RecArgs<R>) * 2 * n));
    HIPCHK(hipMalloc(&G.d_eown, sizeof(int2) * (eo.size() + 1)));
    HIPCHK(hipMalloc(&G.d_nown, sizeof(int2) * (no.size() + 1)));
    HIPCHK(hipMalloc(&G.d_rown, sizeof(int2) * (ro.size() + 1)));
    HIPCHK(hipMemcpy(G.d_eargs, ea.data(), sizeof(ElemArgs<R>) * 2 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(G.d_nargs, na.data(), sizeof(NodeArgs<R>) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(G.d_rargs, ra.data(), sizeof(RecArgs<R>) * 2 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(G.d_eown, eo.data(), sizeof(int2) * eo.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(G.d_nown, no.data(), sizeof(int2) * no.size(), hipMemcpyHostToDevice));
    if (!ro.empty()) HIPCHK(hipMemcpy(G.d_rown, ro.data(), sizeof(int2) * ro.size(), hipMemcpyHostToDevice));
    return 0;
}

// The group's workgroup map for its tile: FP32 groups take their largest
// member's tile, FP64 groups the default (as launch_elem_batch does).
template <int P, typename R>
int batch_build_group(BatchGroup &G)
{
    if constexpr (std::is_same<R, float>::value) {
        if constexpr (extra_tile<P, 1>()) if (G.epb == 1) return batch_build_group_e<P, 1, R>(G);
        if constexpr (extra_tile<P, 2>()) if (G.epb == 2) return batch_build_group_e<P, 2, R>(G);
        if constexpr (extra_tile<P, 4>()) if (G.epb == 4) return batch_build_group_e<P, 4, R>(G);
    }
    return batch_build_group_e<P, 0, R>(G);
}

template <int P, typename R>
int batch_launch_group(const BatchGroup &G, hipStream_t st)
{
    ElemBatchLaunch l;
    l.vm = G.vm;
    l.ks = G.ks;
    l.ll = G.ll;
    l.epb = G.epb;
    l.nwg = (unsigned)G.ewg;
    l.args = (const ElemArgs<R> *)G.d_eargs + (size_t)G.phase * G.sims.size();
    l.own = (const int2 *)G.d_eown;
    const hipError_t e = launch_elem_batch<P, R>(l, st);
    if (e != hipSuccess) return fail(std::string("batched element kernel launch: ") + hipGetErrorString(e));
    return 0;
}

template <int P, typename R>
int batch_launch_node(const BatchGroup &G, hipStream_t st)
{
    if (G.kv == 1)
        hipLaunchKernelGGL((node_kernel_batch<P, R, 1>), dim3(G.nwg), dim3(256), 0, st,
                           (const NodeArgs<R> *)G.d_nargs, (const int2 *)G.d_nown);
    else if (G.kv == 2)
        hipLaunchKernelGGL((node_kernel_batch<P, R, 2>), dim3(G.nwg), dim3(256), 0, st,
                           (const NodeArgs<R> *)G.d_nargs, (const int2 *)G.d_nown);
    else
        hipLaunchKernelGGL((node_kernel_batch<P, R, 0>), dim3(G.nwg), dim3(256), 0, st,
                           (const NodeArgs<R> *)G.d_nargs, (const int2 *)G.d_nown);
    HIPCHK(hipGetLastError());
    if (G.rwg > 0) {
        hipLaunchKernelGGL((record_kernel_batch<P, R>), dim3(G.rwg), dim3(64), 0, st,
                           (const RecArgs<R> *)G.d_rargs + (size_t)G.phase * G.sims.size(),
                           (const int2 *)G.d_rown);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

int batch_group_call(int what, BatchGroup &G, hipStream_t st)
{
#define S3D_BATCH_CASE(PP)                                                                  \
    case PP:                                                                                \
        if (what == 0) return G.fp32 ? batch_build_group<PP, float>(G) : batch_build_group<PP, double>(G); \
        if (what == 1) return G.fp32 ? batch_launch_group<PP, float>(G, st) : batch_launch_group<PP, double>(G, st); \
        return G.fp32 ? batch_launch_node<PP, float>(G, st) : batch_launch_node<PP, double>(G, st);
    switch (G.p) {
        S3D_BATCH_CASE(2)
        S3D_BATCH_CASE(3)
        S3D_BATCH_CASE(4)
        S3D_BATCH_CASE(5)
        S3D_BATCH_CASE(6)
        S3D_BATCH_CASE(7)
    default: return fail("unsupported lpd");
    }
#undef S3D_BATCH_CASE
}

void batch_free_groups(s3d_batch *b)
{
    for (auto &G : b->groups) {
        void *bufs[] = {G.d_eargs, G.d_nargs, G.d_rargs, G.d_eown, G.d_nown, G.d_rown};
        for (void *p : bufs) dfree(p);
    }
    b->groups.clear();
    for (int k = 0; k < 2; ++k) {
        if (b->graph_exec[k]) (void)hipGraphExecDestroy(b->graph_exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->graph_exec[k] = nullptr;
        b->graph[k] = nullptr;
        b->graph_steps[k] = 0;
    }
    b->phase = 0;
    b->built = false;
}

// one batched iteration: every group's element kernel, then every group's
// node and receiver kernels, then every counter advances
int batch_launch_step(s3d_batch *b)
{
    // members with curl receivers: the strain parts of the start velocity,
    // before the element kernels change it
    for (s3d_ctx *c : b->sims)
        if (c->nrot > 0) {
            const int rc = c->fp32 ? with_p<CurlStrainRun, float>(c->p, c, b->stream)
                                   : with_p<CurlStrainRun, double>(c->p, c, b->stream);
            if (rc) return rc;
        }
    for (auto &G : b->groups) {
        G.phase = b->phase;
        if (int rc = batch_group_call(1, G, b->stream)) return rc;
    }
    for (auto &G : b->groups)
        if (int rc = batch_group_call(2, G, b->stream)) return rc;
    const int n = (int)b->sims.size();
    hipLaunchKernelGGL(advance_many_kernel, dim3((n + 63) / 64), dim3(64), 0, b->stream, b->d_its, n);
    HIPCHK(hipGetLastError());
    for (s3d_ctx *c : b->sims) advance_state(c);
    b->phase ^= 1;
    return 0;
}

int batch_build(s3d_batch *b)
{
    for (s3d_ctx *c : b->sims) {
        if (c->src_dirty || c->rec_dirty) b->built = false;
    }
    if (b->built) return 0;
    batch_free_groups(b);
    for (s3d_ctx *c : b->sims) {
        if (upload_sources(c) || upload_receivers(c)) return 1;
        drop_graph(c);
    }
    // one launch per kind of step (order, precision, memory variables,
    // compensation);
    // the element tile and LDS layout give bit-identical results, so a
    // group takes those of its largest member rather than splitting by them
    for (s3d_ctx *c : b->sims) {
        bool placed = false;
        for (auto &G : b->groups)
            if (G.p == c->p && G.fp32 == c->fp32 && G.visco == c->visco &&
                G.vm == (c->visco ? c->nm : 0) &&
                G.kv == c->kv && G.ks == c->o.kahan_sigma) {
                G.sims.push_back(c);
                placed = true;
                break;
            }
        if (!placed) {
            BatchGroup G;
            G.p = c->p;
            G.fp32 = c->fp32;
            G.visco = c->visco;
            G.vm = c->visco ? c->nm : 0;
            G.kv = c->kv;
            G.ks = c->o.kahan_sigma;
            G.sims.push_back(c);
            b->groups.push_back(G);
        }
    }
    for (auto &G : b->groups) {
        const s3d_ctx *big = G.sims[0];
        for (const s3d_ctx *c : G.sims)
            if (c->nel > big->nel) big = c;
        G.epb = big->epb;
        G.ll = big->fp32 ? kernel_variant(big) : 0;
    }
    for (auto &G : b->groups)
        if (int rc = batch_group_call(0, G, b->stream)) return rc;
    b->built = true;
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *s3d_last_error(void) { return g_err.c_str(); }

int s3d_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return 0;
}

int s3d_gll(int lpd, double *knots, double *weights, double *dl)
{
    if (gll_tables(lpd, knots, weights, dl)) return fail("lpd must be 1..7");
    return 0;
}

int s3d_wavelet(int wavelet, double onset, double width, double dt, int nt, double *out)
{
    // src/misc/source_signature_mod.f90:50-147, t = (i-1) dt
    for (int i = 0; i < nt; ++i) {
        double t = (double)i * dt, s = 0.0;
        switch (wavelet) {
        case S3D_RICKER: {
            double q = PI / width, a2 = -(q * q), u = t - onset;
            s = 2.0 * (0.5 + a2 * (u * u)) * std::exp(a2 * (u * u));
        } break;
        case S3D_GAUSS: {
            double u = (t - onset) / width;
            s = std::exp(-(u * u));
        } break;
        case S3D_SIN3:
            if (onset < t && t < onset + width) {
                double sv = std::sin(PI * (t - onset) / width);
                s = sv * sv * sv;
            }
            break;
        case S3D_BRUNE:
            if (t >= onset) s = (t - onset) / (width * width) * std::exp(-(t - onset) / width);
            break;
        case S3D_DELTA: s = (t == onset) ? 1.0 : 0.0; break;
        default: return fail("unknown wavelet");
        }
        out[i] = s;
    }
    return 0;
}

int s3d_create(const s3d_grid *grid, const s3d_options *opt, s3d_ctx **out)
{
    if (!grid || !out) return fail("null argument");
    if (grid->lpd < 2 || grid->lpd > 7) return fail("lpd must be 2..7 on the GPU path");
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return fail("element counts must be >= 1");
    if (grid->nt < 1) return fail("nt must be >= 1");
    if (grid->taper_width > 0 && grid->taper_slope == 0.0) return fail("taper_slope must be nonzero");
    {
        // the kernels index elements in 31 bits and unique nodes in 32
        const long pb = (long)grid->lpd * grid->lpd * grid->lpd;
        const long nel = (long)grid->nx * grid->ny * grid->nz;
        const long nuniq = (long)(grid->nx + 1) * (grid->ny + 1) * (grid->nz + 1) * pb;
        if (nel > 0x7fffffffL || nuniq > 0xffffffffL)
            return fail("mesh too large: more than 2^31 - 1 elements or 2^32 - 1 unique nodes");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
    s3d_ctx *c = new s3d_ctx();
    c->g = *grid;
    s3d_options d{};
    d.precision = S3D_FP64;
    c->o = opt ? *opt : d;
    c->fp32 = c->o.precision == S3D_FP32;
    if (c->o.kahan_sigma < 0 || c->o.kahan_sigma > 2) { delete c; return fail("kahan_sigma must be 0, 1 or 2"); }
    if (c->o.kahan_v < 0 || c->o.kahan_v > 2) { delete c; return fail("kahan_v must be 0, 1 or 2"); }
    if (!c->fp32) { c->o.kahan_v = 0; c->o.kahan_sigma = 0; c->o.bf16_static = 0; }
    c->kv = c->o.kahan_v;
    c->o.bf16_static = c->o.bf16_static != 0;
    if (c->o.state_buffers < 0 || c->o.state_buffers > 2) {
        delete c;
        return fail("state_buffers must be 0 (default), 1 or 2");
    }
    c->nbuf = c->o.state_buffers == 1 ? 1 : 2;
    if (c->o.memory_vars != 0 && c->o.memory_vars != 5 && c->o.memory_vars != 6) {
        delete c;
        return fail("memory_vars must be 0 (default, 6), 6 or 5");
    }
    c->nm = c->o.memory_vars == 5 ? 5 : 6;
    c->dev = c->o.device;
    if (c->dev < 0 || c->dev >= ndev) { delete c; return fail("bad device ordinal"); }
    hipError_t e = hipSetDevice(c->dev);
    if (e == hipSuccess) e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) { delete c; return fail(std::string("stream: ") + hipGetErrorString(e)); }
    if (c->dev < MAX_DEVICES) {
        // the device's copy of the s3d_vrange_noalloc scratch array
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (!g_scratch[c->dev]) {
            void *ptr = nullptr;
            e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                (void)hipStreamDestroy(c->stream);
                delete c;
                return fail(std::string("scratch address: ") + hipGetErrorString(e));
            }
            g_scratch[c->dev] = (double *)ptr;
        }
    }
    c->p = grid->lpd;
    c->P1 = c->p + 1;
    c->NP = c->P1 * c->P1 * c->P1;
    c->PB = c->p * c->p * c->p;
    c->nel = (long)grid->nx * grid->ny * grid->nz;
    c->nn = c->nel * c->NP;
    c->nblk = (long)(grid->nx + 1) * (grid->ny + 1) * (grid->nz + 1);
    c->nuniq = c->nblk * c->PB;
    gll_tables(c->p, c->knots, c->weights, c->dl);
    *out = c;
    return 0;
}

int s3d_destroy(s3d_ctx *c)
{
    if (!c) return 0;
    if (c->sharers > 0)
        return fail(std::to_string(c->sharers) + " simulation(s) created with s3d_create_shared still use "
                    "this model; destroy them first");
    if (c->batches > 0)
        return fail("the simulation is a member of " + std::to_string(c->batches) +
                    " batch(es); destroy them first");
    (void)hipSetDevice(c->dev);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->model) {
        // the static fields belong to the model's simulation
        void **shared[] = {&c->dt_inv_mm, (void **)&c->mm64, &c->lam, &c->mu, &c->tau, &c->rho, &c->ca,
                           &c->cb, &c->cc, &c->tth, &c->tr, &c->wdev, (void **)&c->dpx,
                           (void **)&c->dpy, (void **)&c->dpz};
        for (void **q : shared) *q = nullptr;
        c->model->sharers--;
        c->model = nullptr;
    }
    for (int q = 0; q < 7; ++q) dfree(c->stage[q]);
    c->sig = c->sig_lo = c->mem = nullptr;   // views of the state blocks
    void *bufs[] = {c->v[0], c->v[1], c->v[2], c->lo[0], c->lo[1], c->lo[2], c->dt_inv_mm, c->blk[0], c->blk[1],
                    c->mat, c->rho, c->ca, c->cb, c->cc, c->grad_rho, c->grad_stage, c->dgrad, c->curl_parts,
                    c->force, c->tth, c->tr, c->wdev, c->mm64, c->dpx, c->dpy, c->dpz,
                    c->d_slot, c->d_chan, c->d_coef, c->d_stf, c->d_rec_e, c->d_rec_nch,
                    c->d_rec_off, c->d_rec_ch, c->d_rec_vm, c->d_traces, c->d_rec_rot, c->d_rot_e, c->d_rot_buf};
    for (void *b : bufs) dfree(b);
    for (int k = 0; k < 3; ++k)
        for (hipEvent_t e : c->ev[k]) (void)hipEventDestroy(e);
    drop_graph(c);
    dfree(c->d_it);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int s3d_create_shared(s3d_ctx *base, s3d_ctx **out)
{
    if (!base || !out) return fail("null argument");
    if (!base->inited) return fail("the model's simulation must be initialised (s3d_init) first");
    s3d_ctx *m = model_of(base);
    HIPCHK(hipSetDevice(m->dev));
    s3d_ctx *c = new s3d_ctx();
    // mesh, options, host geometry and element tile of the model's simulation
    c->g = m->g;
    c->o = m->o;
    c->nbuf = m->nbuf;
    c->nm = m->nm;
    c->p = m->p; c->P1 = m->P1; c->NP = m->NP; c->PB = m->PB;
    c->nel = m->nel; c->nn = m->nn; c->nblk = m->nblk; c->nuniq = m->nuniq;
    c->fp32 = m->fp32; c->visco = m->visco; c->aniso = m->aniso; c->kv = m->kv;
    c->q_finite = m->q_finite; c->abc_nonzero = m->abc_nonzero;
    c->dev = m->dev;
    std::copy(m->knots, m->knots + MAXP1, c->knots);
    std::copy(m->weights, m->weights + MAXP1, c->weights);
    std::copy(m->dl, m->dl + MAXP1 * MAXP1, c->dl);
    c->x_glb_min = m->x_glb_min; c->x_glb_max = m->x_glb_max;
    c->y_glb_min = m->y_glb_min; c->y_glb_max = m->y_glb_max;
    c->z_glb_min = m->z_glb_min; c->z_glb_max = m->z_glb_max;
    c->dx = m->dx; c->dy = m->dy; c->dz = m->dz; c->jac = m->jac;
    c->x_loc_min = m->x_loc_min; c->x_loc_max = m->x_loc_max;
    c->y_loc_min = m->y_loc_min; c->y_loc_max = m->y_loc_max;
    c->z_loc_min = m->z_loc_min; c->z_loc_max = m->z_loc_max;
    c->x1 = m->x1; c->y1 = m->y1; c->z1 = m->z1;
    c->px1 = m->px1; c->py1 = m->py1; c->pz1 = m->pz1;
    c->courant = m->courant;
    c->epb = m->epb;
    c->lds_layout = m->lds_layout;
    c->early_stores = m->early_stores;
    // static fields: the model's, not copied
    c->dt_inv_mm = m->dt_inv_mm; c->mm64 = m->mm64; c->lam = m->lam; c->mu = m->mu; c->tau = m->tau;
    c->rho = m->rho; c->ca = m->ca; c->cb = m->cb; c->cc = m->cc; c->tth = m->tth; c->tr = m->tr;
    c->wdev = m->wdev; c->dpx = m->dpx; c->dpy = m->dpy; c->dpz = m->dpz;
    c->model = m;
    m->sharers++;
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
        s3d_destroy(c);
        return fail(std::string("stream: ") + hipGetErrorString(e));
    }
    if (alloc_state(c)) {
        const std::string msg = s3d_last_error();
        s3d_destroy(c);
        return fail(msg);
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        s3d_destroy(c);
        return fail(std::string("create_shared: ") + hipGetErrorString(e));
    }
    c->slot_of.assign(c->nel, -1);
    c->inited = true;
    c->it = 0;
    *out = c;
    return 0;
}

int s3d_set_model_slab(s3d_ctx *c, int param, int ez0, int nez, const double *data)
{
    if (!c || !data) return fail("null argument");
    if (param < 0 || param > 6) return fail("unknown model parameter");
    if (c->inited) return fail("model must be set before s3d_init");
    if (ez0 < 0 || nez < 1 || ez0 + nez > c->g.nz) return fail("slab outside the mesh");
    HIPCHK(hipSetDevice(c->dev));
    if (!c->stage[param]) {
        if (dalloc(c, (void **)&c->stage[param], sizeof(double) * c->nn, false)) return 1;
        c->bytes -= sizeof(double) * (double)c->nn;   // staging is transient
        if (param == S3D_Q) {
            // default HUGE everywhere (src/model_paras.F90:76)
            std::vector<double> h(c->nel, HUGE64);
            for (int n = 0; n < c->NP; ++n)
                HIPCHK(hipMemcpy(c->stage[param] + (size_t)n * c->nel, h.data(), sizeof(double) * c->nel,
                                 hipMemcpyHostToDevice));
        } else {
            HIPCHK(hipMemset(c->stage[param], 0, sizeof(double) * c->nn));
        }
    }
    const size_t layer = (size_t)c->g.nx * c->g.ny;
    const size_t width = layer * nez * sizeof(double);
    HIPCHK(hipMemcpy2D(c->stage[param] + layer * ez0, sizeof(double) * c->nel, data, width, width,
                       c->NP, hipMemcpyHostToDevice));
    const size_t n = layer * nez * c->NP;
    if (param == S3D_Q) {
        for (size_t q = 0; q < n; ++q)
            if (data[q] != HUGE64) { c->q_finite = true; break; }
    } else if (param == S3D_A || param == S3D_B || param == S3D_C) {
        for (size_t q = 0; q < n; ++q)
            if (data[q] != 0.0) { c->abc_nonzero = true; break; }
    }
    return 0;
}

int s3d_set_model(s3d_ctx *c, int param, const double *data)
{
    if (!c) return fail("null argument");
    return s3d_set_model_slab(c, param, 0, c->g.nz, data);
}

int s3d_init(s3d_ctx *c)
{
    if (!c) return fail("null argument");
    if (c->inited) return fail("s3d_init called twice");
    if (!c->stage[S3D_RHOINV] || !c->stage[S3D_LAMBDA] || !c->stage[S3D_MU])
        return fail("rhoinv, lambda and mu must be set before s3d_init");
    HIPCHK(hipSetDevice(c->dev));
    const s3d_grid &g = c->g;
    const int P1 = c->P1;
    // src/config/model.f90:89-239 with one rank
    c->x_glb_min = deg2rad(90.0 - g.lat_max);
    c->x_glb_max = deg2rad(90.0 - g.lat_min);
    c->y_glb_min = deg2rad(g.lon_min);
    c->y_glb_max = deg2rad(g.lon_max);
    c->z_glb_min = g.rad_min;
    c->z_glb_max = g.rad_max;
    double dxg = c->x_glb_max - c->x_glb_min, dyg = c->y_glb_max - c->y_glb_min,
           dzg = c->z_glb_max - c->z_glb_min;
    c->dx = dxg / g.nx;
    c->dy = dyg / g.ny;
    c->dz = dzg / g.nz;
    c->x_loc_min = c->x_glb_min;
    c->x_loc_max = c->x_glb_min + dxg;
    c->y_loc_min = c->y_glb_min;
    c->y_loc_max = c->y_glb_min + dyg;
    c->z_loc_min = g.rad_min;
    c->z_loc_max = g.rad_min + dzg;
    c->jac = c->dx * c->dy * c->dz / 8.0;
    c->x1.assign((size_t)g.nx * P1, 0.0);
    c->y1.assign((size_t)g.ny * P1, 0.0);
    c->z1.assign((size_t)g.nz * P1, 0.0);
    coords_1d(c->knots, c->p, g.nx, c->x_loc_min, c->x_loc_max, c->x1.data());
    coords_1d(c->knots, c->p, g.ny, c->y_loc_min, c->y_loc_max, c->y1.data());
    coords_1d(c->knots, c->p, g.nz, c->z_loc_max, c->z_loc_min, c->z1.data());
    // damping profiles (src/geometric_paras.f90:190-279)
    const int pml = g.taper_width;
    const double alpha = g.taper_slope;
    auto prof = [&](const std::vector<double> &x1, int n, double d, double lo, double hi, bool do_lo,
                    std::vector<double> &out) {
        out.assign((size_t)n * P1, 0.0);
        if (pml <= 0) return;
        for (int e = 0; e < n; ++e)
            for (int m = 0; m < P1; ++m) {
                double t;
                if (do_lo && e < pml) {
                    t = ((lo + pml * d) - x1[e * P1 + m]) / (pml * d);
                    out[e * P1 + m] = alpha * (t * t);
                }
                if (e >= n - pml) {
                    t = ((hi - pml * d) - x1[e * P1 + m]) / (pml * d);
                    out[e * P1 + m] = alpha * (t * t);
                }
            }
    };
    prof(c->x1, g.nx, c->dx, c->x_loc_min, c->x_loc_max, true, c->px1);
    prof(c->y1, g.ny, c->dy, c->y_loc_min, c->y_loc_max, true, c->py1);
    // r: only the bottom strip; the surface is free (src/geometric_paras.f90:256-279)
    c->pz1.assign((size_t)g.nz * P1, 0.0);
    if (pml > 0)
        for (int e = std::max(0, g.nz - pml); e < g.nz; ++e)
            for (int m = 0; m < P1; ++m) {
                double t = ((c->z_loc_min + pml * c->dz) - c->z1[e * P1 + m]) / (pml * c->dz);
                c->pz1[e * P1 + m] = alpha * (t * t);
            }

    // ---- 1-D tables ----
    const size_t rs = rsize(c);
    std::vector<double> sin1(g.nx * P1), th1 = c->x1, ph1 = c->y1, r1 = c->z1;
    for (size_t q = 0; q < sin1.size(); ++q) sin1[q] = std::sin(c->x1[q]);
    if (c->fp32 ? upload_tables<coef_t<float>>(c, sin1) : upload_tables<double>(c, sin1)) return 1;
    {
        if (dalloc(c, &c->wdev, sizeof(double) * P1)) return 1;
        HIPCHK(hipMemcpy(c->wdev, c->weights, sizeof(double) * P1, hipMemcpyHostToDevice));
    }
    auto up = [&](double **dst, const std::vector<double> &h) -> int {
        if (dalloc(c, (void **)dst, sizeof(double) * h.size(), false)) return 1;
        HIPCHK(hipMemcpy(*dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        return 0;
    };
    if (up(&c->dpx, c->px1) || up(&c->dpy, c->py1) || up(&c->dpz, c->pz1)) return 1;
    double *d_sin = nullptr, *d_r = nullptr, *d_w = nullptr, *d_th = nullptr, *d_ph = nullptr;
    HIPCHK(hipMalloc(&d_sin, sizeof(double) * sin1.size()));
    HIPCHK(hipMalloc(&d_r, sizeof(double) * r1.size()));
    HIPCHK(hipMalloc(&d_w, sizeof(double) * P1));
    HIPCHK(hipMalloc(&d_th, sizeof(double) * th1.size()));
    HIPCHK(hipMalloc(&d_ph, sizeof(double) * ph1.size()));
    HIPCHK(hipMemcpy(d_sin, sin1.data(), sizeof(double) * sin1.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_r, r1.data(), sizeof(double) * r1.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_w, c->weights, sizeof(double) * P1, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_th, th1.data(), sizeof(double) * th1.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_ph, ph1.data(), sizeof(double) * ph1.size(), hipMemcpyHostToDevice));

    // ---- state and static fields ----
    c->visco = c->q_finite;
    c->aniso = c->abc_nonzero;
    if (alloc_state(c)) return 1;
    const size_t ss = c->o.bf16_static ? 2 : rs;   // static fields
    if (dalloc(c, &c->dt_inv_mm, ss * c->nuniq)) return 1;
    if (dalloc(c, (void **)&c->mm64, sizeof(double) * c->nuniq)) return 1;
    // lam, mu (and tau) as one element-major block; lam/mu/tau point at the
    // components' first entries
    const int ncm = c->visco ? 3 : 2;
    if (dalloc(c, &c->mat, ss * ncm * c->nn, false)) return 1;
    c->lam = c->mat;
    c->mu = (char *)c->mat + ss * c->NP;
    c->tau = c->visco ? (char *)c->mat + 2 * ss * c->NP : nullptr;
    if (dalloc(c, &c->rho, rs * c->nn, false)) return 1;
    if (c->aniso) {
        if (dalloc(c, &c->ca, ss * c->nn, false) || dalloc(c, &c->cb, ss * c->nn, false) ||
            dalloc(c, &c->cc, ss * c->nn, false))
            return 1;
    }
    const unsigned gm = nblocks(c->nn, 256);
    if (c->fp32 && c->o.bf16_static)
        hipLaunchKernelGGL((material_kernel<float, bf16>), dim3(gm), dim3(256), 0, c->stream,
                           c->stage[S3D_RHOINV], c->stage[S3D_LAMBDA], c->stage[S3D_MU], c->stage[S3D_Q],
                           c->stage[S3D_A], c->stage[S3D_B], c->stage[S3D_C], (bf16 *)c->lam,
                           (bf16 *)c->mu, (bf16 *)c->tau, (float *)c->rho, (bf16 *)c->ca, (bf16 *)c->cb,
                           (bf16 *)c->cc, c->nel, c->NP, ncm, c->g.dt);
    else if (c->fp32)
        hipLaunchKernelGGL((material_kernel<float, float>), dim3(gm), dim3(256), 0, c->stream,
                           c->stage[S3D_RHOINV], c->stage[S3D_LAMBDA], c->stage[S3D_MU], c->stage[S3D_Q],
                           c->stage[S3D_A], c->stage[S3D_B], c->stage[S3D_C], (float *)c->lam,
                           (float *)c->mu, (float *)c->tau, (float *)c->rho, (float *)c->ca, (float *)c->cb,
                           (float *)c->cc, c->nel, c->NP, ncm, c->g.dt);
    else
        hipLaunchKernelGGL((material_kernel<double, double>), dim3(gm), dim3(256), 0, c->stream, c->stage[S3D_RHOINV],
                           c->stage[S3D_LAMBDA], c->stage[S3D_MU], c->stage[S3D_Q], c->stage[S3D_A],
                           c->stage[S3D_B], c->stage[S3D_C], (double *)c->lam, (double *)c->mu,
                           (double *)c->tau, (double *)c->rho, (double *)c->ca, (double *)c->cb,
                           (double *)c->cc, c->nel, c->NP, ncm, c->g.dt);
    HIPCHK(hipGetLastError());
    int rc = c->fp32 ? with_p<MassRun, float>(c->p, c, (const double *)d_sin, (const double *)d_r, (const double *)d_w)
                     : with_p<MassRun, double>(c->p, c, (const double *)d_sin, (const double *)d_r, (const double *)d_w);
    if (rc) return rc;
    rc = with_p<CflRun, double>(c->p, c, (const double *)d_th, (const double *)d_ph, (const double *)d_r,
                                &c->courant);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(d_sin); (void)hipFree(d_r); (void)hipFree(d_w); (void)hipFree(d_th); (void)hipFree(d_ph);
    for (int q = 0; q < 7; ++q) { dfree(c->stage[q]); c->stage[q] = nullptr; }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->slot_of.assign(c->nel, -1);
    c->inited = true;
    c->it = 0;
    return 0;
}

int s3d_autotune(s3d_ctx *c, int *tile, double *ms)
{
    if (!c) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    if (c->it != 0 || c->fields_set)
        return fail("s3d_autotune needs the state at rest: call it after s3d_init, before s3d_set_field / s3d_step");
    HIPCHK(hipSetDevice(c->dev));
    int rc;
    switch (c->p) {
#define S3D_TUNE(PP) case PP: rc = c->fp32 ? autotune_p<PP, float>(c, tile, ms) : autotune_p<PP, double>(c, tile, ms); break;
        S3D_TUNE(2) S3D_TUNE(3) S3D_TUNE(4) S3D_TUNE(5) S3D_TUNE(6) S3D_TUNE(7)
#undef S3D_TUNE
    default: return fail("unsupported lpd");
    }
    return rc;
}

int s3d_set_tile(s3d_ctx *c, int tile)
{
    if (!c) return fail("null argument");
    if (tile != 0 && tile != 1 && tile != 2 && tile != 4) return fail("tile must be 0, 1, 2 or 4");
    if (tile != 0 && c->o.bf16_static) return fail("BF16-static simulations use the default tile");
    const int P1 = c->p + 1;
    if (tile * P1 * P1 * P1 > 1024) return fail("tile does not fit 1024 lanes at this order");
    const int dflt = std::max(1, 256 / (P1 * P1 * P1));   // default_epb
    c->epb = tile == dflt ? 0 : tile;
    drop_graph(c);
    return 0;
}

int s3d_set_lds_layout(s3d_ctx *c, int layout)
{
    if (!c) return fail("null argument");
    if (layout != 0 && layout != 1) return fail("LDS layout must be 0 (planar) or 1 (node-interleaved)");
    if (layout != 0 && (!c->fp32 || c->o.bf16_static))
        return fail("FP64 and BF16-static simulations use the planar LDS layout");
    if (layout == 1 && c->early_stores)
        return fail("the node-interleaved LDS layout stores the state after the divergence");
    c->lds_layout = layout;
    drop_graph(c);
    return 0;
}

int s3d_set_store_order(s3d_ctx *c, int early)
{
    if (!c) return fail("null argument");
    if (early != 0 && early != 1) return fail("store order must be 0 (after the divergence) or 1 (before)");
    if (early && (!c->fp32 || c->o.bf16_static))
        return fail("FP64 and BF16-static simulations store the state after the divergence");
    if (early && c->lds_layout == 1)
        return fail("the node-interleaved LDS layout stores the state after the divergence");
    c->early_stores = early != 0;
    drop_graph(c);
    return 0;
}

int s3d_get_store_order(s3d_ctx *c, int *early)
{
    if (!c || !early) return fail("null argument");
    *early = c->early_stores ? 1 : 0;
    return 0;
}

int s3d_get_tuning(s3d_ctx *c, int *tile, int *xcd_remap, int *lds_layout)
{
    if (!c || !tile || !xcd_remap || !lds_layout) return fail("null argument");
    int d = 1;
    switch (c->p) {
    case 2: d = default_epb<2>(); break;
    case 3: d = default_epb<3>(); break;
    case 4: d = default_epb<4>(); break;
    case 5: d = default_epb<5>(); break;
    case 6: d = default_epb<6>(); break;
    case 7: d = default_epb<7>(); break;
    default: break;
    }
    *tile = c->epb == 0 ? d : c->epb;
    *xcd_remap = c->o.xcd_remap;
    *lds_layout = c->lds_layout;
    return 0;
}

int s3d_cfl(s3d_ctx *c, double *courant)
{
    if (!c || !courant) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    *courant = c->courant;
    return 0;
}

int s3d_locate(s3d_ctx *c, double lat, double lon, double depth, int *exyz, double *vm)
{
    // construct_point_cls (src/config/point.f90:150-254)
    if (!c || !exyz || !vm) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    double x = deg2rad(90.0 - lat), y = deg2rad(lon), z = EARTH_RADIUS - depth;
    if (!(c->x_glb_min < x && x < c->x_glb_max && c->y_glb_min < y && y < c->y_glb_max &&
          c->z_glb_min < z && z <= c->z_glb_max))
        return 1;
    int ex = (c->g.nx - 1) - (int)std::floor((c->x_loc_max - x) / c->dx);
    int ey = (c->g.ny - 1) - (int)std::floor((c->y_loc_max - y) / c->dy);
    int ez = (int)std::floor((c->z_loc_max - z) / c->dz);
    if (ex < 0 || ex >= c->g.nx || ey < 0 || ey >= c->g.ny || ez < 0 || ez >= c->g.nz) return 1;
    double sx = 2.0 * ((x - c->x_loc_min) / c->dx - ex) - 1.0;
    double sy = 2.0 * ((y - c->y_loc_min) / c->dy - ey) - 1.0;
    double sz = 2.0 * ((c->z_loc_max - z) / c->dz - ez) - 1.0;
    double lx[MAXP1], ly[MAXP1], lz[MAXP1];
    lagrange_1d(c->knots, c->P1, sx, lx);
    lagrange_1d(c->knots, c->P1, sy, ly);
    lagrange_1d(c->knots, c->P1, sz, lz);
    const int P1 = c->P1;
    for (int k = 0; k < P1; ++k)
        for (int j = 0; j < P1; ++j)
            for (int i = 0; i < P1; ++i) vm[i + P1 * (j + P1 * k)] = lx[i] * ly[j] * lz[k];
    exyz[0] = ex;
    exyz[1] = ey;
    exyz[2] = ez;
    return 0;
}

// A batch caches its members' device pointers (sources, receivers, traces)
// in its launch arguments and graph; adding a source or a receiver to a
// member reallocates those buffers, so members are closed to both, and to
// stepping on their own, until the batch is destroyed.
static int refuse_batch_member(const s3d_ctx *c, const char *what)
{
    return fail(std::string(what) + ": the simulation is a member of " + std::to_string(c->batches) +
                " batch(es); destroy the batch first");
}

static int add_source(s3d_ctx *c, const int *exyz, const double *vm, const double *stf, int ncomp, int first)
{
    if (!c || !exyz || !vm || !stf) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    if (c->batches > 0) return refuse_batch_member(c, "add_source");
    int ex = exyz[0], ey = exyz[1], ez = exyz[2];
    if (ex < 0 || ex >= c->g.nx || ey < 0 || ey >= c->g.ny || ez < 0 || ez >= c->g.nz)
        return fail("source element outside the mesh");
    const long e = ex + (long)c->g.nx * (ey + (long)c->g.ny * ez);
    int sl = c->slot_of[e];
    if (sl < 0) {
        Slot s;
        s.e = (int)e;
        for (int q = 0; q < 9; ++q) s.chan[q] = -1;
        s.coef.assign(9 * (size_t)c->NP, 0.0);
        c->slots.push_back(s);
        sl = (int)c->slots.size() - 1;
        c->slot_of[e] = sl;
    }
    Slot &S = c->slots[sl];
    const int P1 = c->P1, nt = c->g.nt;
    for (int q = 0; q < ncomp; ++q) {
        const double *row = stf + (size_t)q * nt;
        bool nz = false;
        for (int i = 0; i < nt; ++i)
            if (row[i] != 0.0) { nz = true; break; }
        if (!nz) continue;   // dropped channel (src/config/source.f90:320-323)
        const int comp = first + q;
        c->stf.insert(c->stf.end(), row, row + nt);
        S.chan[comp] = c->nchan++;
        for (int k = 0; k < P1; ++k)
            for (int j = 0; j < P1; ++j)
                for (int i = 0; i < P1; ++i) {
                    const int n = i + P1 * (j + P1 * k);
                    double cf;
                    if (comp < 6) {
                        // vm / (jac W r^2 sin(theta)) (src/config/channel_mod.f90:866)
                        double r = c->z1[ez * P1 + k], st = std::sin(c->x1[ex * P1 + i]);
                        double rr = r * (r * st);
                        double W = c->weights[i] * c->weights[j] * c->weights[k];
                        cf = vm[n] / (c->jac * W * rr);
                    } else {
                        cf = vm[n];   // src/config/channel_mod.f90:1028
                    }
                    S.coef[(size_t)comp * c->NP + n] = cf;
                }
    }
    c->src_dirty = true;
    return 0;
}

int s3d_add_source_mt(s3d_ctx *c, const int *exyz, const double *vm, const double *stf6)
{
    return add_source(c, exyz, vm, stf6, 6, 0);
}

int s3d_add_source_sf(s3d_ctx *c, const int *exyz, const double *vm, const double *stf3)
{
    return add_source(c, exyz, vm, stf3, 3, 6);
}

int s3d_add_receiver(s3d_ctx *c, const int *exyz, const double *vm, int nch, const int *channels, int *id)
{
    if (!c || !exyz || !vm || !channels) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    if (nch < 1 || nch > S3D_MAX_CHANNELS)
        return fail("a receiver records 1.." + std::to_string(S3D_MAX_CHANNELS) + " channels");
    int ex = exyz[0], ey = exyz[1], ez = exyz[2];
    if (ex < 0 || ex >= c->g.nx || ey < 0 || ey >= c->g.ny || ez < 0 || ez >= c->g.nz)
        return fail("receiver element outside the mesh");
    for (int q = 0; q < nch; ++q)
        if (channels[q] < S3D_CH_VX || channels[q] > S3D_CH_FZ)
            return fail("unsupported receiver channel " + std::to_string(channels[q]));
    if (c->batches > 0) return refuse_batch_member(c, "add_receiver");
    c->rec_e.push_back((int)(ex + (long)c->g.nx * (ey + (long)c->g.ny * ez)));
    c->rec_nch.push_back(nch);
    c->rec_off.push_back((int)c->rec_ch.size());
    c->rec_ch.insert(c->rec_ch.end(), channels, channels + nch);
    c->rec_vm.insert(c->rec_vm.end(), vm, vm + c->NP);
    c->rec_dirty = true;
    if (id) *id = (int)c->rec_e.size() - 1;
    return 0;
}

int s3d_step(s3d_ctx *c, int nsteps)
{
    if (!c) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    if (nsteps < 0) return fail("negative step count");
    if (c->batches > 0) return refuse_batch_member(c, "step");
    if (c->it + nsteps > c->g.nt) return fail("stepping past nt");
    HIPCHK(hipSetDevice(c->dev));
    if (upload_sources(c) || upload_receivers(c)) return 1;
    // Launch-bound meshes replay a captured chunk of GRAPH_CHUNK iterations;
    // the rest (and every step while kernel timing is on) launches directly.
    constexpr int GRAPH_CHUNK = 16;
    int left = nsteps;
    static_assert(GRAPH_CHUNK % 2 == 0, "a chunk must end on the state block it starts from");
    if (c->use_graph && !c->timing && left >= GRAPH_CHUNK) {
        if (c->graph_steps[c->cur] != GRAPH_CHUNK) {
            int rc = capture_graph(c, GRAPH_CHUNK);
            if (rc) return rc;
        }
        while (left >= GRAPH_CHUNK) {
            HIPCHK(hipGraphLaunch(c->graph_exec[c->cur], c->stream));
            c->it += GRAPH_CHUNK;
            left -= GRAPH_CHUNK;
        }
    }
    for (; left > 0; --left) {
        int rc = launch_one(c);
        if (rc) return rc;
        c->it += 1;
    }
    return 0;
}

int s3d_current_step(s3d_ctx *c, int *it)
{
    if (!c || !it) return fail("null argument");
    *it = c->it;
    return 0;
}

int s3d_get_trace(s3d_ctx *c, int rec, int ch, double *out)
{
    if (!c || !out) return fail("null argument");
    if (rec < 0 || rec >= (int)c->rec_e.size()) return fail("no such receiver");
    if (ch < 0 || ch >= c->rec_nch[rec]) return fail("no such channel");
    HIPCHK(hipSetDevice(c->dev));
    if (upload_receivers(c)) return 1;
    HIPCHK(hipMemcpyAsync(out, c->d_traces + ((size_t)c->rec_off[rec] + ch) * c->g.nt, sizeof(double) * c->g.nt,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

// Component of the memory-variable part of the state block that holds field
// S3D_MXX + 6 * mechanism + comp (comp < 5 in the five-word form).
static int mem_word(const s3d_ctx *c, int field)
{
    const int f = field - S3D_MXX;
    return c->nm == 5 ? 5 * (f / 6) + f % 6 : f;
}

int s3d_get_field(s3d_ctx *c, int field, double *out)
{
    if (!c || !out) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    HIPCHK(hipSetDevice(c->dev));
    const long nn = c->nn;
    if (field >= S3D_GRAD_RHO && field <= S3D_GRAD_CSV) {
        // grad_cs/cp/csh/csv: the reference builds them from forward strain
        // rates it never loads; they read as zero
        if (field == S3D_GRAD_RHO && c->grad_rho) {
            HIPCHK(hipMemcpyAsync(out, c->grad_rho, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } else {
            std::memset(out, 0, sizeof(double) * nn);
        }
        return 0;
    }
    if (field == S3D_TAPER) {
        const int P1 = c->P1;
        for (int k = 0; k < P1; ++k)
            for (int j = 0; j < P1; ++j)
                for (int i = 0; i < P1; ++i)
                    for (int ez = 0; ez < c->g.nz; ++ez)
                        for (int ey = 0; ey < c->g.ny; ++ey)
                            for (int ex = 0; ex < c->g.nx; ++ex) {
                                long n = ex + (long)c->g.nx * (ey + (long)c->g.ny * (ez + (long)c->g.nz *
                                                                                              (i + (long)P1 * (j + (long)P1 * k))));
                                out[n] = c->g.taper_width > 0
                                             ? taper_value(c->px1[ex * P1 + i], c->py1[ey * P1 + j],
                                                           c->pz1[ez * P1 + k], c->g.taper_slope)
                                             : 1.0;
                            }
        return 0;
    }
    double *buf = nullptr;
    HIPCHK(hipMalloc(&buf, sizeof(double) * nn));
    int rc = 0;
    if (field >= S3D_VX && field <= S3D_VZ) {
        rc = c->fp32 ? with_p<ExpandRun, float>(c->p, c, (const void *)c->v[field], (const void *)c->lo[field], buf)
                     : with_p<ExpandRun, double>(c->p, c, (const void *)c->v[field], (const void *)nullptr, buf);
    } else if (field == S3D_MM) {
        rc = with_p<ExpandRun, double>(c->p, c, (const void *)c->mm64, (const void *)nullptr, buf);
    } else if (field >= S3D_ROT_VX && field <= S3D_ROT_VZ) {
        if (!c->curl_parts || c->curl_it != c->it - 1) {
            (void)hipFree(buf);
            return fail("volume curl fields need s3d_capture_curl_strain just before the last step");
        }
        const int comp = field - S3D_ROT_VX;
        rc = c->fp32 ? with_p<CurlFieldRun, float>(c->p, c, comp, buf)
                     : with_p<CurlFieldRun, double>(c->p, c, comp, buf);
    } else if (field >= S3D_EXX && field <= S3D_EZZ) {
        if (!c->dgrad) {
            (void)hipFree(buf);
            return fail("the displacement gradient is not integrated: call s3d_enable_displacement_gradient "
                        "before stepping");
        }
        hipLaunchKernelGGL(from_device_layout<double>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                           (const double *)c->dgrad + (long)c->NP * (field - S3D_EXX), buf, c->nel, c->NP, 6);
    } else if (field >= S3D_MXX && field < S3D_MXX + 18 && c->mem && c->nm == 5 &&
               (field - S3D_MXX) % 6 == ZZ) {
        // five-word form: zz is not stored, it is -(xx + yy)
        (void)hipFree(buf);
        std::vector<double> xx(nn), yy(nn);
        const int m0 = S3D_MXX + 6 * ((field - S3D_MXX) / 6);
        if (int rc = s3d_get_field(c, m0 + XX, xx.data())) return rc;
        if (int rc = s3d_get_field(c, m0 + YY, yy.data())) return rc;
        for (long q = 0; q < nn; ++q) out[q] = -(xx[q] + yy[q]);
        return 0;
    } else if ((field >= S3D_SXX && field <= S3D_SZZ) || (field >= S3D_MXX && field < S3D_MXX + 18)) {
        const void *src = nullptr;
        const void *lo = nullptr;
        const int nc = c->ns;   // components of the state block
        if (field <= S3D_SZZ) {
            src = (const char *)c->sig + rsize(c) * c->NP * (field - S3D_SXX);
            if (c->sig_lo && c->o.kahan_sigma == 1) lo = (const char *)c->sig_lo + rsize(c) * c->NP * (field - S3D_SXX);
            if (c->sig_lo && c->o.kahan_sigma == 2) lo = (const char *)c->sig_lo + 4 * c->NP * ((field - S3D_SXX) / 2);
        } else if (c->mem) {
            src = (const char *)c->mem + rsize(c) * c->NP * mem_word(c, field);
        }
        if (!src) {
            HIPCHK(hipMemsetAsync(buf, 0, sizeof(double) * nn, c->stream));
        } else if (c->fp32) {
            hipLaunchKernelGGL(from_device_layout<float>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                               (const float *)src, buf, c->nel, c->NP, nc);
            if (lo) {
                // value = hi - lo
                double *tmp = nullptr;
                HIPCHK(hipMalloc(&tmp, sizeof(double) * nn));
                if (c->o.kahan_sigma == 2)
                    hipLaunchKernelGGL(lo_pair_from_device, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                                       (const unsigned *)lo, (field - S3D_SXX) % 2, tmp, c->nel, c->NP, nc);
                else
                    hipLaunchKernelGGL(from_device_layout<float>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                                       (const float *)lo, tmp, c->nel, c->NP, nc);
                std::vector<double> a(nn), b(nn);
                HIPCHK(hipMemcpyAsync(a.data(), buf, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipMemcpyAsync(b.data(), tmp, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
                (void)hipFree(tmp);
                for (long q = 0; q < nn; ++q) out[q] = a[q] - b[q];
                (void)hipFree(buf);
                return 0;
            }
        } else {
            hipLaunchKernelGGL(from_device_layout<double>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                               (const double *)src, buf, c->nel, c->NP, nc);
        }
    } else {
        (void)hipFree(buf);
        return fail("unknown field id " + std::to_string(field));
    }
    if (rc) { (void)hipFree(buf); return rc; }
    HIPCHK(hipMemcpyAsync(out, buf, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(buf);
    return 0;
}

int s3d_set_field(s3d_ctx *c, int field, const double *in)
{
    if (!c || !in) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    HIPCHK(hipSetDevice(c->dev));
    c->fields_set = true;
    const long nn = c->nn;
    double *buf = nullptr;
    HIPCHK(hipMalloc(&buf, sizeof(double) * nn));
    HIPCHK(hipMemcpyAsync(buf, in, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    int rc = 0;
    if (field >= S3D_VX && field <= S3D_VZ) {
        rc = c->fp32 ? with_p<GatherRun, float>(c->p, c, (const double *)buf, c->v[field], c->lo[field])
                     : with_p<GatherRun, double>(c->p, c, (const double *)buf, c->v[field], (void *)nullptr);
    } else if ((field >= S3D_SXX && field <= S3D_SZZ) || (field >= S3D_MXX && field < S3D_MXX + 18)) {
        void *dst = nullptr;
        const int nc = c->ns;   // components of the state block
        if (field <= S3D_SZZ) {
            dst = (char *)c->sig + rsize(c) * c->NP * (field - S3D_SXX);
            if (c->sig_lo && c->o.kahan_sigma == 1)
                hipLaunchKernelGGL(clear_component<float>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                                   (float *)c->sig_lo + c->NP * (field - S3D_SXX), c->nel, c->NP, nc);
            if (c->sig_lo && c->o.kahan_sigma == 2)
                hipLaunchKernelGGL(lo_pair_clear, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                                   (unsigned *)c->sig_lo + c->NP * ((field - S3D_SXX) / 2), (field - S3D_SXX) % 2,
                                   c->nel, c->NP, nc);
        } else if (c->mem && c->nm == 5 && (field - S3D_MXX) % 6 == ZZ) {
            rc = fail("with five memory words per mechanism the zz memory variables are -(xx + yy) "
                      "and cannot be set");
        } else if (c->mem) {
            dst = (char *)c->mem + rsize(c) * c->NP * mem_word(c, field);
        } else {
            rc = fail("memory variables exist only for a visco-elastic model");
        }
        if (dst) {
            if (c->fp32)
                hipLaunchKernelGGL(to_device_layout<float>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                                   (const double *)buf, (float *)dst, c->nel, c->NP, nc);
            else
                hipLaunchKernelGGL(to_device_layout<double>, dim3(nblocks(nn, 256)), dim3(256), 0, c->stream,
                                   (const double *)buf, (double *)dst, c->nel, c->NP, nc);
        }
    } else {
        rc = fail("field " + std::to_string(field) + " cannot be set");
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (rc) return rc;
    if (e != hipSuccess) return fail(std::string("set_field: ") + hipGetErrorString(e));
    return 0;
}

int s3d_grad_add(s3d_ctx *c, const double *vx_fw, const double *vy_fw, const double *vz_fw, int increment)
{
    if (!c || !vx_fw || !vy_fw || !vz_fw) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    if (increment == 0) return fail("grad_add: increment must be nonzero");
    HIPCHK(hipSetDevice(c->dev));
    const long nn = c->nn;
    if (!c->grad_rho && dalloc(c, &c->grad_rho, sizeof(double) * nn)) return 1;
    // the snapshot staging buffer is kept for the next saved step
    if (!c->grad_stage && dalloc(c, &c->grad_stage, sizeof(double) * 3 * nn, false)) return 1;
    double *f = (double *)c->grad_stage;
    const double *src[3] = {vx_fw, vy_fw, vz_fw};
    int rc = 0;
    for (int q = 0; q < 3 && !rc; ++q)
        if (hipMemcpyAsync(f + q * nn, src[q], sizeof(double) * nn, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            rc = fail("grad_add: upload failed");
    // increment*dt as the reference forms it, before the product and the
    // division by J
    const double inc_dt = (double)increment * c->g.dt;
    if (!rc)
        rc = c->fp32 ? with_p<GradRun, float>(c->p, c, (const double *)f, inc_dt)
                     : with_p<GradRun, double>(c->p, c, (const double *)f, inc_dt);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc) return rc;
    if (e != hipSuccess) return fail(std::string("grad_add: ") + hipGetErrorString(e));
    return 0;
}

int s3d_enable_displacement_gradient(s3d_ctx *c)
{
    if (!c) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    if (c->batches > 0) return refuse_batch_member(c, "enable_displacement_gradient");
    if (c->dgrad) return 0;
    HIPCHK(hipSetDevice(c->dev));
    if (dalloc(c, &c->dgrad, sizeof(double) * 6 * c->nn)) return 1;
    drop_graph(c);   // the captured launches carry the element kernel's arguments
    return 0;
}

int s3d_capture_curl_strain(s3d_ctx *c)
{
    if (!c) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    HIPCHK(hipSetDevice(c->dev));
    if (!c->curl_parts && dalloc(c, &c->curl_parts, sizeof(double) * 6 * c->nn, false)) return 1;
    const int rc = c->fp32 ? with_p<CurlPartsRun, float>(c->p, c) : with_p<CurlPartsRun, double>(c->p, c);
    if (rc) return rc;
    c->curl_it = c->it;
    return 0;
}

int s3d_vrange(s3d_ctx *c, double *out2)
{
    if (!c || !out2) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    HIPCHK(hipSetDevice(c->dev));
    return c->fp32 ? with_p<VrangeRun, float>(c->p, c, out2) : with_p<VrangeRun, double>(c->p, c, out2);
}

int s3d_vrange_noalloc(s3d_ctx *c, double *out2)
{
    if (!c || !out2) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    const long nb = (c->nuniq + 255) / 256;
    if (nb > S3D_NOALLOC_PAIRS)
        return fail("vrange_noalloc: " + std::to_string(c->nuniq) + " unique nodes exceed the " +
                    std::to_string(256 * S3D_NOALLOC_PAIRS) + " of the fixed scratch buffer");
    double *part = c->dev < MAX_DEVICES ? g_scratch[c->dev] : nullptr;
    if (!part) return fail("vrange_noalloc: no scratch buffer on this device");
    HIPCHK(hipSetDevice(c->dev));
    return c->fp32 ? with_p<VrangeNoallocRun, float>(c->p, c, part, (unsigned)nb, out2)
                   : with_p<VrangeNoallocRun, double>(c->p, c, part, (unsigned)nb, out2);
}

int s3d_kinetic_energy(s3d_ctx *c, double *out)
{
    if (!c || !out) return fail("null argument");
    if (!c->inited) return fail("s3d_init first");
    HIPCHK(hipSetDevice(c->dev));
    return c->fp32 ? with_p<KineticRun, float>(c->p, c, out) : with_p<KineticRun, double>(c->p, c, out);
}

int s3d_synchronize(s3d_ctx *c)
{
    if (!c) return fail("null argument");
    HIPCHK(hipSetDevice(c->dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int s3d_timing_enable(s3d_ctx *c, int enable)
{
    if (!c) return fail("null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) {
        for (hipEvent_t e : c->ev[k]) (void)hipEventDestroy(e);
        c->ev[k].clear();
        c->ms[k] = 0;
        c->cnt[k] = 0;
    }
    c->timing = enable != 0;
    return 0;
}

int s3d_kernel_times(s3d_ctx *c, double *ms3, int *count3)
{
    if (!c || !ms3 || !count3) return fail("null argument");
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) {
        auto &v = c->ev[k];
        for (size_t q = 0; q + 1 < v.size(); q += 2) {
            float m = 0.f;
            HIPCHK(hipEventElapsedTime(&m, v[q], v[q + 1]));
            c->ms[k] += m;
            c->cnt[k] += 1;
        }
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
        v.clear();
        ms3[k] = c->ms[k];
        count3[k] = c->cnt[k];
    }
    return 0;
}

int s3d_batch_create(int n, s3d_ctx *const *sims, s3d_batch **out)
{
    if (!sims || !out || n < 1) return fail("a batch needs at least one simulation");
    for (int q = 0; q < n; ++q) {
        if (!sims[q] || !sims[q]->inited) return fail("batch members must be initialised (s3d_init)");
        if (sims[q]->dev != sims[0]->dev) return fail("batch members must share one device");
        if (sims[q]->o.bf16_static) return fail("BF16 static fields are not supported in a batch");
        for (int r = 0; r < q; ++r)
            if (sims[r] == sims[q]) return fail("a simulation appears twice in the batch");
    }
    HIPCHK(hipSetDevice(sims[0]->dev));
    s3d_batch *b = new s3d_batch();
    b->sims.assign(sims, sims + n);
    b->dev = sims[0]->dev;
    hipError_t e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&b->d_its, sizeof(int *) * n);
    if (e == hipSuccess) {
        std::vector<int *> its(n);
        for (int q = 0; q < n; ++q) its[q] = sims[q]->d_it;
        e = hipMemcpy(b->d_its, its.data(), sizeof(int *) * n, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return fail(std::string("batch: ") + hipGetErrorString(e));
    }
    for (int q = 0; q < n; ++q) sims[q]->batches++;
    *out = b;
    return 0;
}

int s3d_batch_step(s3d_batch *b, int nsteps)
{
    if (!b) return fail("null argument");
    if (nsteps < 0) return fail("negative step count");
    for (s3d_ctx *c : b->sims)
        if (c->it + nsteps > c->g.nt) return fail("stepping past nt in a batch member");
    HIPCHK(hipSetDevice(b->dev));
    if (int rc = batch_build(b)) return rc;
    // order the batch stream after each member's own stream
    for (s3d_ctx *c : b->sims) {
        HIPCHK(hipEventRecord(b->ev, c->stream));
        HIPCHK(hipStreamWaitEvent(b->stream, b->ev, 0));
    }
    constexpr int GRAPH_CHUNK = 16;
    int left = nsteps;
    // an even chunk ends in the phase (and on the state blocks) it started
    // from, so the graph captured for a phase replays whenever it is current
    static_assert(GRAPH_CHUNK % 2 == 0, "a chunk must end in the phase it starts from");
    if (left >= GRAPH_CHUNK) {
        const int k = b->phase;
        if (b->graph_steps[k] != GRAPH_CHUNK) {
            std::vector<int> cur0;
            for (s3d_ctx *c : b->sims) cur0.push_back(c->cur);
            auto restore = [&]() {
                for (size_t q = 0; q < b->sims.size(); ++q) set_state_block(b->sims[q], cur0[q]);
                b->phase = k;
            };
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            int rc = 0;
            for (int q = 0; q < GRAPH_CHUNK && rc == 0; ++q) rc = batch_launch_step(b);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                restore();
                return rc;
            }
            if (e != hipSuccess) {
                restore();
                return fail(std::string("batch capture: ") + hipGetErrorString(e));
            }
            b->graph[k] = g;
            e = hipGraphInstantiate(&b->graph_exec[k], g, nullptr, nullptr, 0);
            if (e != hipSuccess) return fail(std::string("batch instantiate: ") + hipGetErrorString(e));
            b->graph_steps[k] = GRAPH_CHUNK;
        }
        while (left >= GRAPH_CHUNK) {
            HIPCHK(hipGraphLaunch(b->graph_exec[k], b->stream));
            for (s3d_ctx *c : b->sims) c->it += GRAPH_CHUNK;
            left -= GRAPH_CHUNK;
        }
    }
    for (; left > 0; --left) {
        if (int rc = batch_launch_step(b)) return rc;
        for (s3d_ctx *c : b->sims) c->it += 1;
    }
    // and each member's stream after the batch
    HIPCHK(hipEventRecord(b->ev, b->stream));
    for (s3d_ctx *c : b->sims) HIPCHK(hipStreamWaitEvent(c->stream, b->ev, 0));
    return 0;
}

int s3d_batch_synchronize(s3d_batch *b)
{
    if (!b) return fail("null argument");
    HIPCHK(hipSetDevice(b->dev));
    HIPCHK(hipStreamSynchronize(b->stream));
    return 0;
}

int s3d_batch_destroy(s3d_batch *b)
{
    if (!b) return 0;
    (void)hipSetDevice(b->dev);
    (void)hipStreamSynchronize(b->stream);
    batch_free_groups(b);
    dfree(b->d_its);
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (s3d_ctx *c : b->sims) c->batches--;
    delete b;
    return 0;
}

int s3d_device_bytes(s3d_ctx *c, double *bytes)
{
    if (!c || !bytes) return fail("null argument");
    *bytes = c->bytes;
    return 0;
}

}  // extern "C"
