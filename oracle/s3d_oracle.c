/*
 * s3d_oracle.c -- FP64 CPU restatement of the Ses3d-NT forward time step.
 *
 * TEST INFRASTRUCTURE ONLY: what the HIP path is compared with, never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every routine restates one piece of the reference (cited per function) in
 * IEEE double precision, keeping the reference's 6-D element-local layout,
 * its operation order inside each expression and its single-precision
 * literals (0.1 in the taper, 6.0814 and -1.015 in tau, 1/3 in pressure).
 * Compiled with -ffp-contract=off so no FMA is introduced behind our back.
 *
 * Pins (see tests/test_oracle_*.py):
 *   GLL tables        -> reference gll_mod built in oracle/_ref, closed forms
 *   derivative matrix -> reference get_dlgll in oracle/_ref, polynomial exactness
 *   coordinates       -> reference coordinates_1d in oracle/_ref, seamlessness
 *   interpolation     -> reference interp_mod in oracle/_ref, polynomial exactness
 *   wavelets          -> reference source_signature_mod in oracle/_ref
 *   strain operator   -> rigid rotations / translations give zero strain rate,
 *                        pure dilatation gives the identity (closed forms)
 *   weak divergence   -> partition of unity, discrete integration by parts
 *                        against an independently integrated volume integral
 *   mass matrix       -> sum equals -rho * volume of the spherical section
 *   assembly          -> brute-force sum over copies keyed by node coordinates
 *   taper / CFL       -> closed-form profile values, the printed Courant number
 *   time stepping     -> zero source gives zero field, linearity in the moment
 *                        tensor, P arrival time r/vp in a homogeneous medium
 *   visco (SLS)       -> tau = 0 reduces to the elastic step, the memory
 *                        variable recurrence against its closed form
 */
#include "s3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXP 8
#define NRDISS 3

/* tau_p of the three relaxation mechanisms (src/geometric_paras.f90:41-42). */
static const double TAU_P[NRDISS] = {1000.0, 70.0, 5.25};

/* Single-precision literals that the FP64 reference build promotes. */
#define F32_0_1 ((double)0.1f)        /* src/geometric_paras.f90:316 */
#define F32_6_0814 ((double)6.0814f)  /* src/model_paras.F90:407      */
#define F32_M1_015 ((double)-1.015f)  /* src/model_paras.F90:407      */
#define F32_THIRD ((double)(1.0f / 3.0f)) /* src/config/channel_mod.f90:795 */

static const double PI = 3.141592653589793238462643383279502884;
static const double EARTH_RADIUS = 6371000.0; /* src/parameters.F90:64 */

typedef struct {
    int ex, ey, ez;
    double *vm;      /* (p+1)^3 */
    int kind;        /* 0 = moment tensor, 1 = single force */
    int ncomp;
    int comp_id[6];  /* component index: 0..5 tensor, 6..8 force */
    double *data;    /* ncomp * nt */
} ora_source;

typedef struct {
    int ex, ey, ez;
    double *vm;
    int nch;
    int ch[32];
    double *data;    /* nch * nt */
} ora_receiver;

struct ora_ctx {
    ora_config cfg;
    int p, P1, nel;
    long nn;          /* number of element-local nodes */
    long s_i, s_j, s_k;
    double knots[MAXP], weights[MAXP];
    double dl[MAXP * MAXP];  /* dl[i + P1*j] = l_i'(xi_j) */
    /* derived scalars (src/config/model.f90:89-239) */
    double x_glb_min, x_glb_max, y_glb_min, y_glb_max, z_glb_min, z_glb_max;
    double dx, dy, dz, jac;
    double x_loc_min, x_loc_max, y_loc_min, y_loc_max, z_loc_min, z_loc_max;
    double *x1, *y1, *z1;    /* 1-D coordinates (nx*P1 etc.), elem fastest */
    double *px1, *py1, *pz1; /* 1-D taper profiles */
    /* 6-D static fields */
    double *theta, *phi, *r, *sin_theta, *cot_theta, *r_sin_theta,
        *rr_sin_theta, *w3, *taper;
    double *rhoinv, *lambda, *mu, *a, *b, *c, *qq;
    double *two_mu, *kappa, *rho, *cp, *cs, *tau, *mu_tau, *mm;
    int anisotropy, visco;
    /* state */
    double *v[3], *s[3], *sig[6], *e[9], *mem; /* mem: 6*NRDISS fields */
    double *du[9];   /* displacement gradient (K3), ORA_DXUX order, or NULL */
    double *src[9];  /* xx xy xz yy yz zz x y z */
    double *work[6];
    int it;           /* iterations completed */
    int nsrc, nrec;
    ora_source *srcs;
    ora_receiver *recs;
};

/* strain-rate slots in ctx->e */
enum { EXX = 0, EYY, EZZ, EXY, EXZ, EYZ, EYX, EZX, EZY };

/* ------------------------------------------------------------------------ */
/* GLL basis: src/gll_mod.f90:61-275                                          */
/* ------------------------------------------------------------------------ */

/* Legendre P_n and its derivative at x (three-term recurrence). */
static void legendre(int n, long double x, long double *pn, long double *dpn)
{
    long double p0 = 1.0L, p1 = x;
    if (n == 0) { *pn = 1.0L; *dpn = 0.0L; return; }
    for (int k = 2; k <= n; ++k) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1; p1 = p2;
    }
    *pn = p1;
    /* (1-x^2) P_n' = n (P_{n-1} - x P_n) */
    *dpn = (long double)n * (p0 - x * p1) / (1.0L - x * x);
}

/*
 * GLL points are -1, 1 and the roots of P_N'. They are found here by Newton
 * iteration on P_N' in long double, starting from the Chebyshev-Gauss-Lobatto
 * points, and rounded once to double. Weights are 2 / (N (N+1) P_N(x)^2).
 * For N = 1 the reference returns the placeholder weights 0.5, 0.5
 * (src/gll_mod.f90:66-68); this restatement keeps that behaviour.
 */
int ora_gll(int lpd, double *knots, double *weights, double *dl)
{
    if (lpd < 1 || lpd > 7) return 1;
    const int N = lpd;
    long double xs[MAXP];
    xs[0] = -1.0L; xs[N] = 1.0L;
    for (int i = 1; i < N; ++i) {
        long double x = -cosl(PI * (long double)i / (long double)N);
        for (int it = 0; it < 100; ++it) {
            /* f = P_N'(x); f' = P_N''(x) from Legendre's ODE:
             * (1-x^2) P'' = 2x P' - N(N+1) P */
            long double pn, dpn;
            legendre(N, x, &pn, &dpn);
            long double d2 = (2.0L * x * dpn - (long double)N * (N + 1) * pn) /
                             (1.0L - x * x);
            long double dx = dpn / d2;
            x -= dx;
            if (fabsl(dx) < 1e-30L) break;
        }
        xs[i] = x;
    }
    /* enforce exact symmetry and the exact zero of odd-count rules */
    for (int i = 0; i <= N / 2; ++i) {
        long double m = 0.5L * (xs[N - i] - xs[i]);
        xs[i] = -m; xs[N - i] = m;
    }
    if (N % 2 == 0) xs[N / 2] = 0.0L;
    for (int i = 0; i <= N; ++i) {
        knots[i] = (double)xs[i];
        long double pn, dpn;
        legendre(N, xs[i], &pn, &dpn);
        weights[i] = (double)(2.0L / ((long double)N * (N + 1) * pn * pn));
    }
    if (N == 1) { weights[0] = 0.5; weights[1] = 0.5; }
    /* dl(i,j) = l_i'(x_j) (src/gll_mod.f90:251-275) */
    if (dl) {
        for (int i = 0; i <= N; ++i) {
            for (int j = 0; j <= N; ++j) {
                double y;
                if (i == j) {
                    y = 0.0;
                    for (int k = 0; k <= N; ++k)
                        if (k != i) y = y + 1.0 / (knots[i] - knots[k]);
                } else {
                    double num = 1.0, den = 1.0;
                    for (int k = 0; k <= N; ++k) {
                        if (k == i) continue;
                        if (k != j) num = num * (knots[j] - knots[k]);
                        den = den * (knots[i] - knots[k]);
                    }
                    y = num / den;
                }
                dl[i + (N + 1) * j] = y;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */

int ora_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

static double *zalloc(long n)
{
    double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    return p;
}

static double *falloc(long n, double v)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (p) for (long i = 0; i < n; ++i) p[i] = v;
    return p;
}

#define IDX(c, ex, ey, ez, i, j, k)                                          \
    ((long)(ex) + (long)(c)->cfg.nx * ((long)(ey) + (long)(c)->cfg.ny *        \
     ((long)(ez) + (long)(c)->cfg.nz * ((long)(i) + (long)(c)->P1 *            \
     ((long)(j) + (long)(c)->P1 * (long)(k))))))

/* deg2rad, lat2colat: src/misc/coordinate_utilities_mod.f90:52-84 */
static double deg2rad(double d) { return d * PI / 180.0; }
static double lat2colat(double lat) { return 90.0 - lat; }

/*
 * coordinates_1d (src/coordinates_mod.f90:114-141): element i spans
 * min + (i + (1+xi)/2) * delta with delta = (max-min)/elms, and the last node
 * is forced to max.
 */
static void coordinates_1d(const double *knots, int lpd, int elms, double mn,
                           double mx, double *out /* elms*(lpd+1), elm fastest */)
{
    double kn[MAXP];
    for (int j = 0; j <= lpd; ++j) kn[j] = 0.5 * (1.0 + knots[j]);
    double delta = (mx - mn) / elms;
    for (int i = 0; i < elms; ++i)
        for (int j = 0; j <= lpd; ++j)
            out[i + elms * j] = mn + ((double)i + kn[j]) * delta;
    out[(elms - 1) + elms * lpd] = mx;
}

ora_ctx *ora_create(const ora_config *cfg)
{
    if (!cfg || cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return NULL;
    ora_ctx *c = (ora_ctx *)calloc(1, sizeof(ora_ctx));
    if (!c) return NULL;
    c->cfg = *cfg;
    c->p = cfg->lpd;
    c->P1 = cfg->lpd + 1;
    if (ora_gll(c->p, c->knots, c->weights, c->dl)) { free(c); return NULL; }
    c->nel = cfg->nx * cfg->ny * cfg->nz;
    c->nn = (long)c->nel * c->P1 * c->P1 * c->P1;
    c->s_i = c->nel;
    c->s_j = (long)c->nel * c->P1;
    c->s_k = (long)c->nel * c->P1 * c->P1;
    long n = c->nn;
    double **z6[] = {&c->theta, &c->phi, &c->r, &c->sin_theta, &c->cot_theta,
                     &c->r_sin_theta, &c->rr_sin_theta, &c->w3, &c->taper,
                     &c->rhoinv, &c->lambda, &c->mu, &c->a, &c->b, &c->c,
                     &c->two_mu, &c->kappa, &c->rho, &c->cp, &c->cs, &c->tau,
                     &c->mu_tau, &c->mm};
    for (size_t q = 0; q < sizeof(z6) / sizeof(z6[0]); ++q) *z6[q] = zalloc(n);
    c->qq = falloc(n, 1.79769313486231570815e+308); /* HUGE(0.0_real64) */
    for (int q = 0; q < 3; ++q) { c->v[q] = zalloc(n); c->s[q] = zalloc(n); }
    for (int q = 0; q < 6; ++q) { c->sig[q] = zalloc(n); c->work[q] = zalloc(n); }
    for (int q = 0; q < 9; ++q) { c->e[q] = zalloc(n); c->src[q] = zalloc(n); }
    c->mem = zalloc(n * 6 * NRDISS);
    c->x1 = zalloc((long)cfg->nx * c->P1);
    c->y1 = zalloc((long)cfg->ny * c->P1);
    c->z1 = zalloc((long)cfg->nz * c->P1);
    c->px1 = zalloc((long)cfg->nx * c->P1);
    c->py1 = zalloc((long)cfg->ny * c->P1);
    c->pz1 = zalloc((long)cfg->nz * c->P1);
    return c;
}

void ora_destroy(ora_ctx *c)
{
    if (!c) return;
    double *f[] = {c->theta, c->phi, c->r, c->sin_theta, c->cot_theta,
                   c->r_sin_theta, c->rr_sin_theta, c->w3, c->taper, c->rhoinv,
                   c->lambda, c->mu, c->a, c->b, c->c, c->qq, c->two_mu,
                   c->kappa, c->rho, c->cp, c->cs, c->tau, c->mu_tau, c->mm,
                   c->mem, c->x1, c->y1, c->z1, c->px1, c->py1, c->pz1};
    for (size_t q = 0; q < sizeof(f) / sizeof(f[0]); ++q) free(f[q]);
    for (int q = 0; q < 3; ++q) { free(c->v[q]); free(c->s[q]); }
    for (int q = 0; q < 6; ++q) { free(c->sig[q]); free(c->work[q]); }
    for (int q = 0; q < 9; ++q) { free(c->e[q]); free(c->src[q]); free(c->du[q]); }
    for (int q = 0; q < c->nsrc; ++q) { free(c->srcs[q].vm); free(c->srcs[q].data); }
    for (int q = 0; q < c->nrec; ++q) { free(c->recs[q].vm); free(c->recs[q].data); }
    free(c->srcs); free(c->recs);
    free(c);
}

int ora_set_model(ora_ctx *c, int param, const double *f)
{
    double *dst = NULL;
    switch (param) {
    case ORA_RHOINV: dst = c->rhoinv; break;
    case ORA_LAMBDA: dst = c->lambda; break;
    case ORA_MU: dst = c->mu; break;
    case ORA_A: dst = c->a; break;
    case ORA_B: dst = c->b; break;
    case ORA_C: dst = c->c; break;
    case ORA_Q: dst = c->qq; break;
    default: return 1;
    }
    memcpy(dst, f, sizeof(double) * (size_t)c->nn);
    return 0;
}

/*
 * communicate_local_boundaries (src/communicate_fields.F90:48-79): sum shared
 * nodes along theta, then phi, then r; afterwards every copy holds the sum.
 */
static void assemble(ora_ctx *c, double *f)
{
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz, p = c->p, P1 = c->P1;
    for (int ex = 1; ex < nx; ++ex)
#pragma omp parallel for collapse(2)
        for (int k = 0; k < P1; ++k)
            for (int j = 0; j < P1; ++j)
                for (int ez = 0; ez < nz; ++ez)
                    for (int ey = 0; ey < ny; ++ey) {
                        long a0 = IDX(c, ex, ey, ez, 0, j, k);
                        long b0 = IDX(c, ex - 1, ey, ez, p, j, k);
                        f[a0] = f[a0] + f[b0];
                        f[b0] = f[a0];
                    }
    for (int ey = 1; ey < ny; ++ey)
#pragma omp parallel for collapse(2)
        for (int k = 0; k < P1; ++k)
            for (int i = 0; i < P1; ++i)
                for (int ez = 0; ez < nz; ++ez)
                    for (int ex = 0; ex < nx; ++ex) {
                        long a0 = IDX(c, ex, ey, ez, i, 0, k);
                        long b0 = IDX(c, ex, ey - 1, ez, i, p, k);
                        f[a0] = f[a0] + f[b0];
                        f[b0] = f[a0];
                    }
    for (int ez = 1; ez < nz; ++ez)
#pragma omp parallel for collapse(2)
        for (int j = 0; j < P1; ++j)
            for (int i = 0; i < P1; ++i)
                for (int ey = 0; ey < ny; ++ey)
                    for (int ex = 0; ex < nx; ++ex) {
                        long a0 = IDX(c, ex, ey, ez, i, j, 0);
                        long b0 = IDX(c, ex, ey, ez - 1, i, j, p);
                        f[a0] = f[a0] + f[b0];
                        f[b0] = f[a0];
                    }
}

/*
 * One-axis damping profile (src/geometric_paras.f90:190-279): inside the
 * first/last `pml` elements, alpha * (delta / (pml*d))^2 where delta is the
 * distance (in coordinate units) to the inner edge of the strip.
 */
static void profile_1d(const double *x1, int n, int P1, int pml, double alpha,
                       double d, double lo_edge, int do_lo, double hi_edge,
                       int do_hi, double *out)
{
    for (int q = 0; q < n * P1; ++q) out[q] = 0.0;
    if (pml <= 0) return;
    if (do_lo)
        for (int i = 0; i < pml && i < n; ++i)
            for (int m = 0; m < P1; ++m) {
                double delta = (lo_edge + pml * d) - x1[i + n * m];
                double t = delta / (pml * d);
                out[i + n * m] = alpha * (t * t);
            }
    if (do_hi)
        for (int i = n - pml; i < n; ++i) {
            if (i < 0) continue;
            for (int m = 0; m < P1; ++m) {
                double delta = (hi_edge - pml * d) - x1[i + n * m];
                double t = delta / (pml * d);
                out[i + n * m] = alpha * (t * t);
            }
        }
}

/*
 * Setup: init_geometric_parameters (src/geometric_paras.f90:108-325),
 * init_auxiliary_model_parameters (src/model_paras.F90:389-416),
 * init_temp (src/evolution.f90:48-84).
 */
int ora_init(ora_ctx *c)
{
    const ora_config *g = &c->cfg;
    const int nx = g->nx, ny = g->ny, nz = g->nz, P1 = c->P1;
    /* src/config/model.f90:89-239 with one MPI rank (coords 0, px=py=pz=1) */
    c->x_glb_min = deg2rad(lat2colat(g->lat_max));
    c->x_glb_max = deg2rad(lat2colat(g->lat_min));
    c->y_glb_min = deg2rad(g->lon_min);
    c->y_glb_max = deg2rad(g->lon_max);
    c->z_glb_min = g->rad_min;
    c->z_glb_max = g->rad_max;
    double dx_glb = c->x_glb_max - c->x_glb_min;
    double dy_glb = c->y_glb_max - c->y_glb_min;
    double dz_glb = c->z_glb_max - c->z_glb_min;
    c->dx = dx_glb / nx;
    c->dy = dy_glb / ny;
    c->dz = dz_glb / nz;
    double dx_loc = dx_glb / 1, dy_loc = dy_glb / 1, dz_loc = dz_glb / 1;
    c->x_loc_min = c->x_glb_min + 0 * dx_loc;
    c->x_loc_max = c->x_glb_min + 1 * dx_loc;
    c->y_loc_min = c->y_glb_min + 0 * dy_loc;
    c->y_loc_max = c->y_glb_min + 1 * dy_loc;
    c->z_loc_min = g->rad_min + 0 * dz_loc;
    c->z_loc_max = g->rad_min + 1 * dz_loc;
    c->jac = c->dx * c->dy * c->dz / 8.0;

    coordinates_1d(c->knots, c->p, nx, c->x_loc_min, c->x_loc_max, c->x1);
    coordinates_1d(c->knots, c->p, ny, c->y_loc_min, c->y_loc_max, c->y1);
    /* r runs from the surface downwards: min and max are swapped */
    coordinates_1d(c->knots, c->p, nz, c->z_loc_max, c->z_loc_min, c->z1);

    const int pml = g->taper_width;
    const double alpha = g->taper_slope;
    /* one rank: every side is a global boundary; the surface (low z index)
     * is free, only the bottom is tapered (src/geometric_paras.f90:256-279) */
    profile_1d(c->x1, nx, P1, pml, alpha, c->dx, c->x_loc_min, 1, c->x_loc_max, 1, c->px1);
    profile_1d(c->y1, ny, P1, pml, alpha, c->dy, c->y_loc_min, 1, c->y_loc_max, 1, c->py1);
    {
        for (int q = 0; q < nz * P1; ++q) c->pz1[q] = 0.0;
        if (pml > 0)
            for (int k = nz - pml; k < nz; ++k) {
                if (k < 0) continue;
                for (int m = 0; m < P1; ++m) {
                    double delta = (c->z_loc_min + pml * c->dz) - c->z1[k + nz * m];
                    double t = delta / (pml * c->dz);
                    c->pz1[k + nz * m] = alpha * (t * t);
                }
            }
    }

    int aniso = 0, visco = 0;
#pragma omp parallel for collapse(3) reduction(|| : aniso, visco)
    for (int k = 0; k < P1; ++k)
        for (int j = 0; j < P1; ++j)
            for (int i = 0; i < P1; ++i)
                for (int ez = 0; ez < nz; ++ez)
                    for (int ey = 0; ey < ny; ++ey)
                        for (int ex = 0; ex < nx; ++ex) {
                            long n = IDX(c, ex, ey, ez, i, j, k);
                            double th = c->x1[ex + nx * i];
                            double ph = c->y1[ey + ny * j];
                            double rr = c->z1[ez + nz * k];
                            c->theta[n] = th; c->phi[n] = ph; c->r[n] = rr;
                            double st = sin(th);
                            c->sin_theta[n] = st;
                            c->cot_theta[n] = cos(th) / st;
                            c->r_sin_theta[n] = rr * st;
                            c->rr_sin_theta[n] = rr * c->r_sin_theta[n];
                            c->w3[n] = c->weights[i] * c->weights[j] * c->weights[k];
                            /* corner normalisation, src/geometric_paras.f90:303-316 */
                            double px = c->px1[ex + nx * i];
                            double py = c->py1[ey + ny * j];
                            double pz = c->pz1[ez + nz * k];
                            double prof = px + py + pz;
                            double pz2 = 2.0 * pz / (1.0 + prof / alpha);
                            double py2 = 2.0 * py / (1.0 + prof / alpha);
                            double px2 = 2.0 * px / (1.0 + prof / alpha);
                            double prof2 = px2 + py2 + pz2;
                            c->taper[n] = exp(-(F32_0_1 * (prof2 * prof2)));
                            if (pml <= 0) c->taper[n] = 1.0;
                            /* auxiliary model parameters */
                            double mu = c->mu[n], lam = c->lambda[n], ri = c->rhoinv[n];
                            c->two_mu[n] = 2.0 * mu;
                            c->kappa[n] = lam + 2.0 * mu / 3.0;
                            c->rho[n] = ri > 0.0 ? 1.0 / ri : 0.0;
                            c->cp[n] = sqrt((lam + 2.0 * mu) * ri);
                            c->cs[n] = sqrt(mu * ri);
                            c->tau[n] = F32_6_0814 * pow(c->qq[n], F32_M1_015);
                            c->mu_tau[n] = (1.0 + c->tau[n]) * mu;
                            c->mm[n] = -(c->jac * (rr * rr) * st * c->rho[n] * c->w3[n]);
                            if (c->a[n] != 0.0 || c->b[n] != 0.0 || c->c[n] != 0.0) aniso = 1;
                            if (c->qq[n] != 1.79769313486231570815e+308) visco = 1;
                        }
    c->anisotropy = aniso;
    c->visco = visco;
    assemble(c, c->mm);
    c->it = 0;
    return 0;
}

/* cfl_loc, src/model_paras.F90:529-561 */
double ora_cfl(ora_ctx *c)
{
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz, p = c->p;
    double m1 = 0.0, m2 = 0.0, m3 = 0.0;
    for (int k = 0; k <= p; ++k)
        for (int j = 0; j <= p; ++j)
            for (int i = 0; i <= p; ++i)
                for (int ez = 0; ez < nz; ++ez)
                    for (int ey = 0; ey < ny; ++ey)
                        for (int ex = 0; ex < nx; ++ex) {
                            long n = IDX(c, ex, ey, ez, i, j, k);
                            double vmax = c->cs[n] < c->cp[n] ? c->cp[n] : c->cs[n];
                            if (i >= 1) {
                                long n0 = IDX(c, ex, ey, ez, i - 1, j, k);
                                double arc = c->r[n] * (c->theta[n] - c->theta[n0]);
                                double q = vmax / arc; if (q > m1) m1 = q;
                            }
                            if (j >= 1) {
                                long n0 = IDX(c, ex, ey, ez, i, j - 1, k);
                                double arc = c->r_sin_theta[n] * (c->phi[n] - c->phi[n0]);
                                double q = vmax / arc; if (q > m2) m2 = q;
                            }
                            if (k >= 1) {
                                long n0 = IDX(c, ex, ey, ez, i, j, k - 1);
                                double arc = c->r[n0] - c->r[n];
                                double q = vmax / arc; if (q > m3) m3 = q;
                            }
                        }
    double dt = c->cfg.dt;
    double c1 = m1 * dt, c2 = m2 * dt, c3 = m3 * dt;
    double mx = c1;
    if (c2 > mx) mx = c2;
    if (c3 > mx) mx = c3;
    return mx;
}

/* ------------------------------------------------------------------------ */
/* Point location: src/config/point.f90:150-254,553-657,826-889             */
/* Lagrange weights: src/interpolation_mod.f90:83-205                        */
/* ------------------------------------------------------------------------ */

static void lagrange_1d(const double *knots, int P1, double x, double *l)
{
    double w[MAXP];
    for (int j = 0; j < P1; ++j) {
        double a = 1.0, b = 1.0;
        for (int k = 0; k < j; ++k) a = a * (knots[j] - knots[k]);
        for (int k = j + 1; k < P1; ++k) b = b * (knots[j] - knots[k]);
        w[j] = 1.0 / (a * b);
    }
    int hit = -1;
    for (int j = 0; j < P1; ++j) if (x == knots[j]) hit = j;
    if (hit >= 0) {
        for (int j = 0; j < P1; ++j) l[j] = (j == hit) ? 1.0 : 0.0;
        return;
    }
    double prod = 1.0;
    for (int k = 0; k < P1; ++k) prod = prod * (x - knots[k]);
    for (int j = 0; j < P1; ++j) l[j] = prod / (x - knots[j]) * w[j];
}

int ora_lagrange(int lpd, double x, double *l)
{
    double k[MAXP], w[MAXP];
    if (ora_gll(lpd, k, w, NULL)) return 1;
    lagrange_1d(k, lpd + 1, x, l);
    return 0;
}

int ora_locate(ora_ctx *c, double lat, double lon, double depth, int *exyz,
               double *vm)
{
    double x = deg2rad(lat2colat(lat));
    double y = deg2rad(lon);
    double z = EARTH_RADIUS - depth;
    if (!(c->x_glb_min < x && x < c->x_glb_max && c->y_glb_min < y &&
          y < c->y_glb_max && c->z_glb_min < z && z <= c->z_glb_max))
        return 1;
    const int nxl = c->cfg.nx - 1, nyl = c->cfg.ny - 1;
    int ex = nxl - (int)floor((c->x_loc_max - x) / c->dx);
    int ey = nyl - (int)floor((c->y_loc_max - y) / c->dy);
    int ez = (int)floor((c->z_loc_max - z) / c->dz);
    if (ex < 0 || ex >= c->cfg.nx || ey < 0 || ey >= c->cfg.ny || ez < 0 ||
        ez >= c->cfg.nz)
        return 1;
    double sx = 2.0 * ((x - c->x_loc_min) / c->dx - ex) - 1.0;
    double sy = 2.0 * ((y - c->y_loc_min) / c->dy - ey) - 1.0;
    double sz = 2.0 * ((c->z_loc_max - z) / c->dz - ez) - 1.0;
    double lx[MAXP], ly[MAXP], lz[MAXP];
    lagrange_1d(c->knots, c->P1, sx, lx);
    lagrange_1d(c->knots, c->P1, sy, ly);
    lagrange_1d(c->knots, c->P1, sz, lz);
    const int P1 = c->P1;
    for (int k = 0; k < P1; ++k)
        for (int j = 0; j < P1; ++j)
            for (int i = 0; i < P1; ++i)
                vm[i + P1 * (j + P1 * k)] = lx[i] * ly[j] * lz[k];
    exyz[0] = ex; exyz[1] = ey; exyz[2] = ez;
    return 0;
}

/* Wavelets: src/misc/source_signature_mod.f90:50-147 */
int ora_stf(int wavelet, double onset, double width, double dt, int nt, double *out)
{
    for (int i = 0; i < nt; ++i) {
        double t = (double)i * dt;
        double s = 0.0;
        switch (wavelet) {
        case 0: { /* RICKER */
            double q = PI / width;
            double a2 = -(q * q);
            double u = t - onset;
            s = 2.0 * (0.5 + a2 * (u * u)) * exp(a2 * (u * u));
        } break;
        case 1: { /* GAUSS */
            double u = (t - onset) / width;
            s = exp(-(u * u));
        } break;
        case 2: /* SIN**3 */
            if (onset < t && t < onset + width) {
                double sv = sin(PI * (t - onset) / width);
                s = sv * sv * sv;
            }
            break;
        case 3: /* BRUNE */
            if (t >= onset) s = (t - onset) / (width * width) * exp(-(t - onset) / width);
            break;
        case 4: /* DELTA */
            s = (t == onset) ? 1.0 : 0.0;
            break;
        default: return 1;
        }
        out[i] = s;
    }
    return 0;
}

static int push_source(ora_ctx *c, const int *exyz, const double *vm,
                       const double *stf, int ncomp_in, int kind)
{
    const int P3 = c->P1 * c->P1 * c->P1, nt = c->cfg.nt;
    ora_source s;
    memset(&s, 0, sizeof(s));
    s.ex = exyz[0]; s.ey = exyz[1]; s.ez = exyz[2];
    s.kind = kind;
    s.vm = (double *)malloc(sizeof(double) * P3);
    memcpy(s.vm, vm, sizeof(double) * P3);
    s.data = (double *)malloc(sizeof(double) * (size_t)ncomp_in * nt);
    for (int q = 0; q < ncomp_in; ++q) {
        int nonzero = 0;
        for (int i = 0; i < nt; ++i) if (stf[(long)q * nt + i] != 0.0) nonzero = 1;
        if (!nonzero) continue; /* dropped channel */
        memcpy(s.data + (long)s.ncomp * nt, stf + (long)q * nt, sizeof(double) * nt);
        s.comp_id[s.ncomp++] = kind == 0 ? q : 6 + q;
    }
    c->srcs = (ora_source *)realloc(c->srcs, sizeof(ora_source) * (c->nsrc + 1));
    c->srcs[c->nsrc++] = s;
    return 0;
}

int ora_add_source_mt(ora_ctx *c, const int *exyz, const double *vm, const double *stf6)
{
    return push_source(c, exyz, vm, stf6, 6, 0);
}

int ora_add_source_sf(ora_ctx *c, const int *exyz, const double *vm, const double *stf3)
{
    return push_source(c, exyz, vm, stf3, 3, 1);
}

int ora_add_receiver(ora_ctx *c, const int *exyz, const double *vm, int nch,
                     const int *channels)
{
    if (nch < 1 || nch > 32) return -1;
    const int P3 = c->P1 * c->P1 * c->P1, nt = c->cfg.nt;
    ora_receiver r;
    memset(&r, 0, sizeof(r));
    r.ex = exyz[0]; r.ey = exyz[1]; r.ez = exyz[2];
    r.vm = (double *)malloc(sizeof(double) * P3);
    memcpy(r.vm, vm, sizeof(double) * P3);
    r.nch = nch;
    for (int q = 0; q < nch; ++q) r.ch[q] = channels[q];
    r.data = zalloc((long)nch * nt);
    c->recs = (ora_receiver *)realloc(c->recs, sizeof(ora_receiver) * (c->nrec + 1));
    c->recs[c->nrec] = r;
    return c->nrec++;
}

/* ------------------------------------------------------------------------ */
/* ses3d_evolution: src/evolution.f90:87-444                                 */
/* ------------------------------------------------------------------------ */
int ora_evolution(ora_ctx *c)
{
    const int nx = c->cfg.nx, ny = c->cfg.ny, nz = c->cfg.nz, p = c->p, P1 = c->P1;
    const long si = c->s_i, sj = c->s_j, sk = c->s_k;
    const double dt = c->cfg.dt;
    const double dxe = c->dx, dye = c->dy, dze = c->dz, jac = c->jac;
    double *vx = c->v[0], *vy = c->v[1], *vz = c->v[2];
    double **e = c->e;
    double *sx = c->s[0], *sy = c->s[1], *sz = c->s[2];
    const double *dl = c->dl;
    const int nel = c->nel;

    /* K1: raw derivatives along theta, phi, r (src/evolution.f90:126-159) */
#pragma omp parallel for collapse(3)
    for (int k = 0; k < P1; ++k)
        for (int j = 0; j < P1; ++j)
            for (int i = 0; i < P1; ++i) {
                long b = (long)i * si + (long)j * sj + (long)k * sk;
                for (int el = 0; el < nel; ++el) {
                    long n = b + el;
                    double exx = 0.0, exy = 0.0, exz = 0.0, eyx = 0.0, eyy = 0.0,
                           eyz = 0.0, ezx = 0.0, ezy = 0.0, ezz = 0.0;
                    for (int q = 0; q < P1; ++q) {
                        double d1 = 2.0 * dl[q + P1 * i] / dxe;
                        long nq = (long)el + (long)q * si + (long)j * sj + (long)k * sk;
                        exx = exx + vx[nq] * d1;
                        exy = exy + vy[nq] * d1;
                        exz = exz + vz[nq] * d1;
                    }
                    for (int q = 0; q < P1; ++q) {
                        double d2 = 2.0 * dl[q + P1 * j] / dye;
                        long nq = (long)el + (long)i * si + (long)q * sj + (long)k * sk;
                        eyy = eyy + vy[nq] * d2;
                        eyx = eyx + vx[nq] * d2;
                        eyz = eyz + vz[nq] * d2;
                    }
                    for (int q = 0; q < P1; ++q) {
                        double d3 = 2.0 * dl[q + P1 * k] / dze;
                        long nq = (long)el + (long)i * si + (long)j * sj + (long)q * sk;
                        ezz = ezz + vz[nq] * d3;
                        ezy = ezy + vy[nq] * d3;
                        ezx = ezx + vx[nq] * d3;
                    }
                    /* K2: spherical completion (src/evolution.f90:165-173) */
                    double r = c->r[n], rs = c->r_sin_theta[n], ct = c->cot_theta[n];
                    double vxn = vx[n], vyn = vy[n], vzn = vz[n];
                    ezz = -ezz;
                    eyy = eyy / rs + (vzn + vxn * ct) / r;
                    exx = (vzn + exx) / r;
                    ezx = -ezx;
                    exz = (exz - vxn) / r;
                    ezy = -ezy;
                    eyz = -vyn / r + eyz / rs;
                    exy = exy / r;
                    eyx = -(vyn * ct) / r + eyx / rs;
                    /* K3: integrate the velocity gradient to the displacement
                     * gradient (src/evolution.f90:179-187), when requested */
                    if (c->du[0]) {
                        c->du[8][n] = c->du[8][n] + dt * ezz;   /* dzuz */
                        c->du[4][n] = c->du[4][n] + dt * eyy;   /* dyuy */
                        c->du[0][n] = c->du[0][n] + dt * exx;   /* dxux */
                        c->du[6][n] = c->du[6][n] + dt * ezx;   /* dzux */
                        c->du[2][n] = c->du[2][n] + dt * exz;   /* dxuz */
                        c->du[7][n] = c->du[7][n] + dt * ezy;   /* dzuy */
                        c->du[5][n] = c->du[5][n] + dt * eyz;   /* dyuz */
                        c->du[1][n] = c->du[1][n] + dt * exy;   /* dxuy */
                        c->du[3][n] = c->du[3][n] + dt * eyx;   /* dyux */
                    }
                    /* K4: symmetrise (src/evolution.f90:206-208) */
                    exy = (exy + eyx) / 2.0;
                    exz = (exz + ezx) / 2.0;
                    eyz = (eyz + ezy) / 2.0;
                    e[EXX][n] = exx; e[EYY][n] = eyy; e[EZZ][n] = ezz;
                    e[EXY][n] = exy; e[EXZ][n] = exz; e[EYZ][n] = eyz;
                    e[EYX][n] = eyx; e[EZX][n] = ezx; e[EZY][n] = ezy;
                }
            }

    /* K5-K7: stress rates, integration, moment-tensor glut
     * (src/evolution.f90:216-325; visco: the loop form of :216-284) */
    double **w = c->work; /* effective stresses xx xy xz yy yz zz */
    const int visco = c->visco;
#pragma omp parallel for
    for (long n = 0; n < c->nn; ++n) {
        double exx = e[EXX][n], eyy = e[EYY][n], ezz = e[EZZ][n];
        double exy = e[EXY][n], exz = e[EXZ][n], eyz = e[EYZ][n];
        double A = c->a[n], B = c->b[n], C = c->c[n], tmu = c->two_mu[n];
        double rxx, ryy, rzz, rxy, rxz, ryz;
        if (!visco) {
            double L = c->lambda[n] * (exx + eyy + ezz);
            rxx = L + tmu * exx + C * ezz + A * (eyy + exx);
            ryy = L + tmu * eyy + C * ezz + A * (eyy + exx);
            rzz = L + tmu * ezz + C * (eyy + exx);
            rxy = tmu * exy;
            rxz = tmu * exz + 2.0 * B * exz;
            ryz = tmu * eyz + 2.0 * B * eyz;
        } else {
            const double *M = c->mem;
            const long nn = c->nn;
            double Mdx = 0.0, Mdy = 0.0, Mdz = 0.0, Mxy = 0.0, Mxz = 0.0, Myz = 0.0;
            for (int m = 0; m < NRDISS; ++m) {
                Mdx = Mdx + M[(long)(6 * m + 0) * nn + n];
                Mxy = Mxy + M[(long)(6 * m + 1) * nn + n];
                Mxz = Mxz + M[(long)(6 * m + 2) * nn + n];
                Mdy = Mdy + M[(long)(6 * m + 3) * nn + n];
                Myz = Myz + M[(long)(6 * m + 4) * nn + n];
                Mdz = Mdz + M[(long)(6 * m + 5) * nn + n];
            }
            double Msum = Mdx + Mdy + Mdz;
            double mut = c->mu_tau[n];
            double L = (c->kappa[n] - 2.0 * mut / 3.0) * (exx + eyy + ezz) - tmu * Msum / 3.0;
            rxx = L + 2.0 * mut * exx + tmu * Mdx + C * ezz + A * (eyy + exx);
            ryy = L + 2.0 * mut * eyy + tmu * Mdy + C * ezz + A * (eyy + exx);
            rzz = L + 2.0 * mut * ezz + tmu * Mdz + C * (eyy + exx);
            rxy = 2.0 * mut * exy + tmu * Mxy;
            rxz = 2.0 * mut * exz + tmu * Mxz + 2.0 * B * exz;
            ryz = 2.0 * mut * eyz + tmu * Myz + 2.0 * B * eyz;
        }
        double *sg[6];
        for (int q = 0; q < 6; ++q) sg[q] = c->sig[q];
        sg[0][n] = sg[0][n] + dt * rxx;
        sg[1][n] = sg[1][n] + dt * rxy;
        sg[2][n] = sg[2][n] + dt * rxz;
        sg[3][n] = sg[3][n] + dt * ryy;
        sg[4][n] = sg[4][n] + dt * ryz;
        sg[5][n] = sg[5][n] + dt * rzz;
        for (int q = 0; q < 6; ++q) w[q][n] = sg[q][n] + c->src[q][n];
        sx[n] = c->src[6][n];
        sy[n] = c->src[7][n];
        sz[n] = c->src[8][n];
    }

    /* K8: weak divergence (src/evolution.f90:332-367) */
    const double d1 = (2.0 / dxe) * jac;
    const double d2 = (2.0 / dye) * jac;
    const double d3 = -((2.0 / dze) * jac);
    const double *wxx = w[0], *wxy = w[1], *wxz = w[2], *wyy = w[3], *wyz = w[4], *wzz = w[5];
    const double *W = c->w3, *RS = c->r_sin_theta, *R = c->r, *RRS = c->rr_sin_theta;
#pragma omp parallel for collapse(3)
    for (int i = 0; i < P1; ++i)
        for (int j = 0; j < P1; ++j)
            for (int k = 0; k < P1; ++k) {
                long b = (long)i * si + (long)j * sj + (long)k * sk;
                for (int q = 0; q < P1; ++q) {
                    long bx = (long)q * si + (long)j * sj + (long)k * sk;
                    long by = (long)i * si + (long)q * sj + (long)k * sk;
                    long bz = (long)i * si + (long)j * sj + (long)q * sk;
                    double dli = dl[i + P1 * q], dlj = dl[j + P1 * q], dlk = dl[k + P1 * q];
                    for (int el = 0; el < nel; ++el) {
                        double dmx = W[bx + el] * dli * RS[bx + el] * d1;
                        double dmy = W[by + el] * dlj * R[by + el] * d2;
                        double dmz = W[bz + el] * dlk * RRS[bz + el] * d3;
                        long n = b + el;
                        sx[n] = sx[n] + wxz[bz + el] * dmz + wxy[by + el] * dmy + wxx[bx + el] * dmx;
                        sy[n] = sy[n] + wyz[bz + el] * dmz + wyy[by + el] * dmy + wxy[bx + el] * dmx;
                        sz[n] = sz[n] + wzz[bz + el] * dmz + wyz[by + el] * dmy + wxz[bx + el] * dmx;
                    }
                }
            }

    /* K9: divide by the mass matrix (src/evolution.f90:372-374) */
#pragma omp parallel for
    for (long n = 0; n < c->nn; ++n) {
        sx[n] = sx[n] / c->mm[n];
        sy[n] = sy[n] / c->mm[n];
        sz[n] = sz[n] / c->mm[n];
    }
    /* K10: assembly (src/evolution.f90:379-381) */
    assemble(c, sx);
    assemble(c, sy);
    assemble(c, sz);

    /* K12, K13: velocity update and taper (src/evolution.f90:393-409),
     * then the memory variables (loop form of src/evolution.f90:415-442) */
#pragma omp parallel for
    for (long n = 0; n < c->nn; ++n) {
        double t = c->taper[n];
        vx[n] = vx[n] + dt * sx[n];
        vy[n] = vy[n] + dt * sy[n];
        vz[n] = vz[n] + dt * sz[n];
        vx[n] = vx[n] * t;
        vy[n] = vy[n] * t;
        vz[n] = vz[n] * t;
        for (int q = 0; q < 6; ++q) c->sig[q][n] = c->sig[q][n] * t;
        if (visco) {
            double *M = c->mem;
            const long nn = c->nn;
            const double ev[6] = {e[EXX][n], e[EXY][n], e[EXZ][n],
                                  e[EYY][n], e[EYZ][n], e[EZZ][n]};
            for (int m = 0; m < NRDISS; ++m) {
                double Lm = dt * c->tau[n] / (NRDISS * TAU_P[m]);
                for (int q = 0; q < 6; ++q) {
                    long id = (long)(6 * m + q) * nn + n;
                    M[id] = M[id] - dt * M[id] / TAU_P[m] - Lm * ev[q];
                }
            }
        }
    }
    (void)nx; (void)ny; (void)nz; (void)p;
    return 0;
}

/* source_cls%add (src/config/source.f90:359-370 -> channel_mod add_m* and add_f*) */
static void sources_add(ora_ctx *c, int it /* 1-based */)
{
    const int P1 = c->P1;
    for (int s = 0; s < c->nsrc; ++s) {
        ora_source *S = &c->srcs[s];
        for (int q = 0; q < S->ncomp; ++q) {
            int comp = S->comp_id[q];
            double d = S->data[(long)q * c->cfg.nt + (it - 1)];
            double *f = c->src[comp];
            for (int k = 0; k < P1; ++k)
                for (int j = 0; j < P1; ++j)
                    for (int i = 0; i < P1; ++i) {
                        long n = IDX(c, S->ex, S->ey, S->ez, i, j, k);
                        double vmv = S->vm[i + P1 * (j + P1 * k)];
                        if (comp < 6)
                            f[n] = -(d * vmv / (c->jac * c->w3[n] * c->rr_sin_theta[n]));
                        else
                            f[n] = -(d * vmv);
                    }
        }
    }
}

/* receiver_cls%record (src/config/receiver.f90:136-142 -> record_*) */
static void receivers_record(ora_ctx *c, int it)
{
    const int P1 = c->P1;
    for (int r = 0; r < c->nrec; ++r) {
        ora_receiver *R = &c->recs[r];
        for (int q = 0; q < R->nch; ++q) {
            double acc = 0.0;
            for (int k = 0; k < P1; ++k)
                for (int j = 0; j < P1; ++j)
                    for (int i = 0; i < P1; ++i) {
                        long n = IDX(c, R->ex, R->ey, R->ez, i, j, k);
                        double vmv = R->vm[i + P1 * (j + P1 * k)];
                        double f;
                        switch (R->ch[q]) {
                        case ORA_CH_VX: f = vmv * c->v[0][n]; break;
                        case ORA_CH_N: f = -1.0 * vmv * c->v[0][n]; break;
                        case ORA_CH_VY: case ORA_CH_E: f = vmv * c->v[1][n]; break;
                        case ORA_CH_VZ: case ORA_CH_Z: f = vmv * c->v[2][n]; break;
                        case ORA_CH_PRESSURE:
                            f = -(F32_THIRD * (c->sig[0][n] + c->sig[3][n] + c->sig[5][n])) * vmv;
                            break;
                        case ORA_CH_ROT_VX: {
                            double el = -c->v[1][n] / c->r[n] + c->e[EYZ][n] / c->r_sin_theta[n] + c->e[EZY][n];
                            f = vmv * el;
                        } break;
                        case ORA_CH_ROT_VY: {
                            double el = -c->e[EZX][n] + (c->v[0][n] - c->e[EXZ][n]) / c->r[n];
                            f = vmv * el;
                        } break;
                        case ORA_CH_ROT_VZ: {
                            double el = (c->e[EXY][n] + c->v[1][n] * c->cot_theta[n] -
                                         c->e[EYX][n] / c->sin_theta[n]) / c->r[n];
                            f = vmv * el;
                        } break;
                        default:
                            /* mxx .. mzz, fx, fy, fz: the glut or force that
                               sources_add(it) just assigned to this element */
                            if (R->ch[q] >= ORA_CH_MXX && R->ch[q] <= ORA_CH_FZ)
                                f = vmv * c->src[R->ch[q] - ORA_CH_MXX][n];
                            else
                                f = 0.0;
                        }
                        acc = acc + f;
                    }
            R->data[(long)q * c->cfg.nt + (it - 1)] = acc;
        }
    }
}

int ora_step(ora_ctx *c, int n)
{
    for (int s = 0; s < n; ++s) {
        if (c->it >= c->cfg.nt) return 1;
        int it = c->it + 1;
        ora_evolution(c);
        sources_add(c, it);
        receivers_record(c, it);
        c->it = it;
    }
    return 0;
}

int ora_current_it(ora_ctx *c) { return c->it; }
int ora_visco_enabled(ora_ctx *c) { return c->visco; }

static double *field_ptr(ora_ctx *c, int f)
{
    switch (f) {
    case ORA_VX: return c->v[0];
    case ORA_VY: return c->v[1];
    case ORA_VZ: return c->v[2];
    case ORA_SXX: case ORA_SXY: case ORA_SXZ: case ORA_SYY: case ORA_SYZ: case ORA_SZZ:
        return c->sig[f - ORA_SXX];
    case ORA_EXX: return c->e[EXX];
    case ORA_EYY: return c->e[EYY];
    case ORA_EZZ: return c->e[EZZ];
    case ORA_EXY: return c->e[EXY];
    case ORA_EXZ: return c->e[EXZ];
    case ORA_EYZ: return c->e[EYZ];
    case ORA_EYX: return c->e[EYX];
    case ORA_EZX: return c->e[EZX];
    case ORA_EZY: return c->e[EZY];
    case ORA_MM: return c->mm;
    case ORA_TAPER: return c->taper;
    case ORA_THETA: return c->theta;
    case ORA_PHI: return c->phi;
    case ORA_R: return c->r;
    case ORA_SX: return c->s[0];
    case ORA_SY: return c->s[1];
    case ORA_SZ: return c->s[2];
    default:
        if (f >= ORA_MXX && f < ORA_MXX + 6 * NRDISS) return c->mem + (long)(f - ORA_MXX) * c->nn;
        if (f >= ORA_DXUX && f < ORA_DXUX + 9) return c->du[f - ORA_DXUX];
        return NULL;
    }
}

int ora_enable_displacement_gradient(ora_ctx *c)
{
    for (int q = 0; q < 9; ++q)
        if (!c->du[q]) {
            c->du[q] = (double *)calloc((size_t)c->nn, sizeof(double));
            if (!c->du[q]) return 1;
        }
    return 0;
}

int ora_get_field(ora_ctx *c, int f, double *out)
{
    double *p = field_ptr(c, f);
    if (!p) return 1;
    memcpy(out, p, sizeof(double) * (size_t)c->nn);
    return 0;
}

int ora_set_field(ora_ctx *c, int f, const double *in)
{
    double *p = field_ptr(c, f);
    if (!p) return 1;
    memcpy(p, in, sizeof(double) * (size_t)c->nn);
    return 0;
}

int ora_get_trace(ora_ctx *c, int rec, int ch, double *out)
{
    if (rec < 0 || rec >= c->nrec || ch < 0 || ch >= c->recs[rec].nch) return 1;
    memcpy(out, c->recs[rec].data + (long)ch * c->cfg.nt, sizeof(double) * c->cfg.nt);
    return 0;
}

int ora_vrange(ora_ctx *c, double *out2)
{
    double mn = c->v[0][0], mx = c->v[0][0];
    for (int q = 0; q < 3; ++q)
        for (long n = 0; n < c->nn; ++n) {
            double x = c->v[q][n];
            if (x < mn) mn = x;
            if (x > mx) mx = x;
        }
    out2[0] = mn; out2[1] = mx;
    return 0;
}

double ora_kinetic_energy(ora_ctx *c)
{
    double s = 0.0;
    for (long n = 0; n < c->nn; ++n) {
        double v2 = c->v[0][n] * c->v[0][n] + c->v[1][n] * c->v[1][n] + c->v[2][n] * c->v[2][n];
        s = s + v2 * c->rr_sin_theta[n] * c->rho[n] * c->w3[n];
    }
    return 0.5 * c->jac * s;
}

int ora_scalars(ora_ctx *c, double *o)
{
    o[0] = c->dx; o[1] = c->dy; o[2] = c->dz; o[3] = c->jac;
    o[4] = c->x_loc_min; o[5] = c->x_loc_max; o[6] = c->y_loc_min;
    o[7] = c->y_loc_max; o[8] = c->z_loc_min; o[9] = c->z_loc_max;
    return 0;
}

int ora_coords_1d(ora_ctx *c, double *th, double *ph, double *r)
{
    memcpy(th, c->x1, sizeof(double) * (size_t)c->cfg.nx * c->P1);
    memcpy(ph, c->y1, sizeof(double) * (size_t)c->cfg.ny * c->P1);
    memcpy(r, c->z1, sizeof(double) * (size_t)c->cfg.nz * c->P1);
    return 0;
}

int ora_profiles_1d(ora_ctx *c, double *px, double *py, double *pz)
{
    memcpy(px, c->px1, sizeof(double) * (size_t)c->cfg.nx * c->P1);
    memcpy(py, c->py1, sizeof(double) * (size_t)c->cfg.ny * c->P1);
    memcpy(pz, c->pz1, sizeof(double) * (size_t)c->cfg.nz * c->P1);
    return 0;
}
