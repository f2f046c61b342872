/*
 * ses3d_hip.h -- C ABI of the MI355X (gfx950) time-stepping library that
 * replaces Ses3d-NT's forward hot path.
 *
 * The reference exposes this path as a Fortran module API, not a plugin
 * (reference src/evolution.f90:41, src/elastic_vars.F90:89-94,
 * src/config/channel_mod.f90:433-446). Each entry point below names the
 * reference procedure it stands in for. Everything is plain C: opaque
 * handles, plain pointers and sizes, int status codes (0 = success, nonzero
 * = error; s3d_last_error() gives the message). The library owns device
 * memory; the caller owns host buffers. Calls for one context come from one
 * host thread and are ordered on the context's HIP stream.
 *
 * 6-D arrays exchanged with the host are float64 in the reference layout:
 * column-major f(0:nx-1, 0:ny-1, 0:nz-1, 0:p, 0:p, 0:p), element index ex
 * fastest, GLL index k slowest (reference src/elastic_vars.F90:112). Here
 * nx, ny, nz are element COUNTS (the reference's nx_loc is count-1,
 * src/config/grid.F90:233).
 */
#ifndef SES3D_HIP_H
#define SES3D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh and time axis: the &grid, &model and &time NAMELIST groups
 * (reference src/parser/parser.f90:148-260). */
typedef struct s3d_grid {
    int nx, ny, nz;        /* elements along theta, phi, r            */
    int lpd;               /* Lagrange polynomial degree, 2..7        */
    double lat_min, lat_max, lon_min, lon_max;  /* degrees            */
    double rad_min, rad_max;                    /* metres             */
    int taper_width;       /* absorbing strip, elements               */
    double taper_slope;
    double dt;             /* seconds                                 */
    int nt;                /* number of time steps                    */
} s3d_grid;

/* Arithmetic of the time step. These have no counterpart in the reference,
 * whose precision is the compile-time real_kind (src/parameters.F90:49). */
enum { S3D_FP64 = 0, S3D_FP32 = 1 };
typedef struct s3d_options {
    int precision;         /* S3D_FP64 or S3D_FP32                        */
    int kahan_v;           /* FP32: compensated velocity integration,
                              0 off, 1 FP32 compensation words, 2 BF16
                              compensation words                          */
    int kahan_sigma;       /* FP32: compensated stress integration,
                              0 off, 1 FP32 compensation words,
                              2 BF16 compensation words                   */
    int device;            /* HIP device ordinal                          */
    int xcd_remap;         /* 1: give each XCD a contiguous element range */
    int bf16_static;       /* FP32: store the moduli, tau and dt/M in BF16 */
    int state_buffers;     /* stress, compensation and memory variables:
                              2 (and 0, the default): two blocks, each step
                              reads one and writes the other, about 10 %
                              faster element kernel for the state's size
                              again in memory; 1: one block updated in place */
    int memory_vars;       /* visco-elastic memory variables per relaxation
                              mechanism: 6 (and 0, the default) the six
                              components xx xy xz yy yz zz as the reference
                              holds them (src/evolution.f90:95); 5 their
                              deviatoric part xx xy xz yy yz, zz = -(xx + yy):
                              only that part enters the stress, so the
                              stress and velocity follow the same equations
                              with 24 B per node less state traffic in FP32;
                              the diagonal S3D_MXX fields are then the
                              deviatoric parts and zz is read only         */
} s3d_options;

/* Model parameter ids (reference src/model_paras.F90:39-46). */
enum { S3D_RHOINV = 0, S3D_LAMBDA, S3D_MU, S3D_A, S3D_B, S3D_C, S3D_Q };

/* Field ids for s3d_get_field / s3d_set_field (reference
 * src/elastic_vars.F90:89-94; channel field_* views,
 * src/config/channel_mod.f90:505-1189). */
enum {
    S3D_VX = 0, S3D_VY, S3D_VZ,
    S3D_SXX, S3D_SXY, S3D_SXZ, S3D_SYY, S3D_SYZ, S3D_SZZ,  /* *_pml stresses */
    S3D_MM = 18, S3D_TAPER = 19,
    /* adjoint sensitivity densities (reference grad_mod,
       src/SOURCE/ses3d_grad.F90:33-37), see s3d_grad_add */
    S3D_GRAD_RHO = 20, S3D_GRAD_CS, S3D_GRAD_CP, S3D_GRAD_CSH, S3D_GRAD_CSV,
    /* displacement gradient, symmetric part (reference snapshot attributes
       exx .. ezz, src/volume_snapshot.f90:220-236), read only, after
       s3d_enable_displacement_gradient */
    S3D_EXX = 26, S3D_EXY, S3D_EXZ, S3D_EYY, S3D_EYZ, S3D_EZZ,
    S3D_MXX = 32,  /* memory variable: S3D_MXX + 6*mechanism + comp,
                      comp order xx, xy, xz, yy, yz, zz; mechanism 0..2 */
    /* volume curl fields (reference snapshot attributes rot_vx, rot_vy,
       rot_vz, src/volume_snapshot.f90:238-245), read only, after a step
       preceded by s3d_capture_curl_strain */
    S3D_ROT_VX = 50, S3D_ROT_VY, S3D_ROT_VZ
};

/* Receiver channels (reference src/config/channel_mod.f90:270-336). The curl
 * channels rot_vx/vy/vz (:620-774) combine the strain rates the step forms
 * from its start velocity with the velocity at its end, as the reference's
 * records do. The source-echo channels
 * mxx .. mzz and fx, fy, fz (record_m**, record_f*, :871-1085) sample the
 * moment-tensor glut and the single force that the step's add(it) wrote into
 * the receiver's element, i.e. sum(vm * src_**) with the source-time sample
 * of this step; they read 0 in an element without such a source component. */
enum {
    S3D_CH_VX = 0, S3D_CH_VY, S3D_CH_VZ, S3D_CH_N, S3D_CH_E, S3D_CH_Z,
    S3D_CH_PRESSURE, S3D_CH_ROT_VX, S3D_CH_ROT_VY, S3D_CH_ROT_VZ,
    S3D_CH_MXX, S3D_CH_MXY, S3D_CH_MXZ, S3D_CH_MYY, S3D_CH_MYZ, S3D_CH_MZZ,
    S3D_CH_FX, S3D_CH_FY, S3D_CH_FZ
};
/* Channels one receiver may record (the reference's max_attributes is 30,
 * src/examples/dummy_parameters.f90:39). */
#define S3D_MAX_CHANNELS 32

/* Wavelets (reference src/misc/source_signature_mod.f90:50-147). */
enum { S3D_RICKER = 0, S3D_GAUSS, S3D_SIN3, S3D_BRUNE, S3D_DELTA };

typedef struct s3d_ctx s3d_ctx;

/* Last error message of the calling thread. */
const char *s3d_last_error(void);

/* Number of HIP devices visible (0 when none). */
int s3d_device_count(int *count);

/* Allocate the device state of one simulation.
 * Replaces allocate_geometric_parameters, allocate_model_parameters and
 * allocate_elastic_variables (reference src/ses3d-nt.F90:128-146). Meshes
 * with more than 2^31 - 1 elements or 2^32 - 1 unique nodes are refused
 * (the kernels index them in 32 bits; such a mesh would not fit on one GPU
 * anyway). */
int s3d_create(const s3d_grid *grid, const s3d_options *opt, s3d_ctx **out);

/* Release everything the context owns. Fails while simulations created with
 * s3d_create_shared still use this context's model. */
int s3d_destroy(s3d_ctx *ctx);

/* A further simulation on the mesh and model of an initialised one: it uses
 * `model`'s static fields (moduli, tau, A/B/C, dt/M, mass matrix, geometry
 * tables, taper) and options and element tile, and owns only its state
 * (velocity, stress, memory variables, compensation, force), its sources,
 * receivers and step counter; it is ready to step (no s3d_init). The model's
 * simulation must outlive it. No counterpart in the reference, which
 * superposes several sources in one wavefield; this is the batch axis of
 * many sources sharing one Earth model (batched source-innermost by
 * s3d_batch_create). */
int s3d_create_shared(s3d_ctx *model, s3d_ctx **out);

/* Model input: element layers ez0 .. ez0+nez-1 of one parameter, float64 in
 * the reference layout with nz replaced by nez. Parameters never set keep the
 * reference's defaults (0; Q = HUGE, i.e. elastic), src/model_paras.F90:70-76.
 * Replaces read_model_parameters / mpi_read_parallel
 * (reference src/model_paras.F90:97-137). */
int s3d_set_model_slab(s3d_ctx *ctx, int param, int ez0, int nez,
                       const double *data);
int s3d_set_model(s3d_ctx *ctx, int param, const double *data);

/* Geometry, taper, auxiliary model parameters, assembled mass matrix and the
 * derivative matrix, all in FP64, then rounded once to the step precision.
 * Replaces init_geometric_parameters (src/geometric_paras.f90:108),
 * init_auxiliary_model_parameters (src/model_paras.F90:389) and init_temp
 * (src/evolution.f90:48). Must follow the model upload. */
int s3d_init(s3d_ctx *ctx);

/* Courant number of the model for grid->dt (reference cfl_loc,
 * src/model_paras.F90:529-561). Requires s3d_init. */
int s3d_cfl(s3d_ctx *ctx, double *courant);

/* Autotuner: times the element-kernel tile variants (elements per
 * workgroup; FP32 and FP64), the XCD-aware block order, and for FP32 the
 * LDS staging layout and the store order (s3d_set_store_order) on this
 * mesh and keeps the fastest; *tile gets the elements per workgroup chosen
 * and *ms the element kernel time (s3d_get_tuning gives the rest). Needs
 * the state at rest (after s3d_init, before s3d_set_field and s3d_step):
 * the trial launches then leave every field unchanged. No counterpart in
 * the reference. */
int s3d_autotune(s3d_ctx *ctx, int *tile, double *ms);

/* Element tile of the step without timing: 0 default, or 1, 2 or 4
 * elements per workgroup (where that fits 1024 lanes); FP32 and FP64, not
 * BF16-static simulations (default tile only). Every tile gives
 * bit-identical results. No counterpart in the reference. */
int s3d_set_tile(s3d_ctx *ctx, int tile);

/* LDS layout of the FP32 element kernel without timing: 0 planar (one array
 * per staged component), 1 node-interleaved (the components one contraction
 * term reads in one 16-byte vector). Every layout gives bit-identical
 * results; FP64 and BF16-static simulations use 0. s3d_autotune searches
 * tile x block order x layout. No counterpart in the reference. */
int s3d_set_lds_layout(s3d_ctx *ctx, int layout);

/* Where the FP32 element kernel stores the new stress, its compensation and
 * the memory variables: 0 after the weak divergence (default), 1 before it
 * (their registers are then free during the divergence, which raises the
 * occupancy of some GLL orders: order 6 with FP32 compensation of the
 * velocity only runs about 6 % faster). Bit-identical results; planar LDS
 * layout only; FP64 and BF16-static simulations use 0. s3d_autotune
 * searches it with the layout. No counterpart in the reference. */
int s3d_set_store_order(s3d_ctx *ctx, int early);
int s3d_get_store_order(s3d_ctx *ctx, int *early);

/* The element tile (elements per workgroup), block order (1: XCD-aware) and
 * LDS layout in use, e.g. as chosen by s3d_autotune. */
int s3d_get_tuning(s3d_ctx *ctx, int *tile, int *xcd_remap, int *lds_layout);

/* GLL points, weights and derivative matrix dl[i + (p+1)*j] = l_i'(xi_j)
 * (reference get_knots / get_weights / get_dlgll, src/gll_mod.f90:178-275). */
int s3d_gll(int lpd, double *knots, double *weights, double *dl);

/* Point location: element (0-based ex, ey, ez) and the (p+1)^3 Lagrange
 * weights vm (column-major i, j, k) of a point given by lat, lon (degrees) and
 * depth (metres). Returns 0 when the point is inside the domain, 1 when not.
 * Replaces construct_point_cls (src/config/point.f90:150-254). */
int s3d_locate(s3d_ctx *ctx, double lat, double lon, double depth,
               int *exyz, double *vm);

/* Source-time function sampled at t = (i-1)*dt, i = 1..nt (reference
 * make_source_signature_, src/config/source.f90:564-652). */
int s3d_wavelet(int wavelet, double onset, double width, double dt, int nt,
                double *out);

/* Moment-tensor point source: stf6[c*nt + it-1] is channel c (xx, xy, xz,
 * yy, yz, zz) at step it. All-zero channels are dropped, and a later source
 * overwrites an earlier one's component in the same element, like the
 * reference's assignment in add_m** (src/config/source.f90:274-327,
 * src/config/channel_mod.f90:859-999). Sources are refused on a member of a
 * batch (destroy the batch first). */
int s3d_add_source_mt(s3d_ctx *ctx, const int *exyz, const double *vm,
                      const double *stf6);
/* Single-force point source: stf3[c*nt + it-1] for x, y, z
 * (src/config/channel_mod.f90:1023-1095). */
int s3d_add_source_sf(s3d_ctx *ctx, const int *exyz, const double *vm,
                      const double *stf3);

/* Receiver with nch (1 .. S3D_MAX_CHANNELS) channel codes (S3D_CH_*);
 * returns its index in *id. Refused on a member of a batch (destroy the
 * batch first).
 * Replaces parse_receiver_file's channel setup (src/config/receiver.f90:435-492). */
int s3d_add_receiver(s3d_ctx *ctx, const int *exyz, const double *vm,
                     int nch, const int *channels, int *id);

/* Advance nsteps iterations of the forward loop body
 * (reference src/ses3d-nt.F90:240-270): ses3d_evolution, every source's
 * add(it) and every receiver's record(it). Refused on a member of a batch,
 * which s3d_batch_step advances. */
int s3d_step(s3d_ctx *ctx, int nsteps);

/* Iterations completed so far. */
int s3d_current_step(s3d_ctx *ctx, int *it);

/* nt samples of one receiver channel (reference channel data(:),
 * src/config/channel_mod.f90:420-426). Samples not yet recorded are 0. */
int s3d_get_trace(s3d_ctx *ctx, int receiver, int channel, double *out);

/* Full-field copy to / from the host in the reference layout
 * (reference field_* views and volume_snapshot get_field,
 * src/volume_snapshot.f90:192-298). Setting a velocity component takes the
 * value of each unique node from the highest-index element that holds it. */
int s3d_get_field(s3d_ctx *ctx, int field, double *out6d);
int s3d_set_field(s3d_ctx *ctx, int field, const double *in6d);

/* One step of the adjoint workflow's sensitivity accumulation (reference
 * ses3d_grad, src/SOURCE/ses3d_grad.F90:212-213, called from the adjoint loop
 * at src/ses3d-nt.F90:318): per element-local node
 *   grad_rho += increment*dt * (vx_fw*vx + vy_fw*vy + vz_fw*vz) / J,
 * J = dx dy dz / 8, with the forward velocity snapshot vx_fw, vy_fw, vz_fw
 * given in the global 6-D layout (FP64) and vx, vy, vz the current
 * (adjoint) velocity on the device. grad_rho is FP64 on the device, zero
 * until the first call, read with s3d_get_field(S3D_GRAD_RHO); the first
 * call also allocates a staging buffer of three FP64 fields, both kept until
 * s3d_destroy (4 x 8 B per element-local node). The caller
 * picks the steps and snapshots (the reference's saving_vector, :159-174).
 * increment is the reference's signed timestep_increment (nonzero; the
 * reference's default -1 integrates with -dt).
 * The reference accumulates grad_cp, grad_cs, grad_csh and grad_csv from
 * forward strain rates that it never loads (:184-191); S3D_GRAD_CS,
 * S3D_GRAD_CP, S3D_GRAD_CSH and S3D_GRAD_CSV read as zero. */
int s3d_grad_add(s3d_ctx *ctx, const double *vx_fw6d, const double *vy_fw6d, const double *vz_fw6d,
                 int increment);

/* Integrate the displacement gradient from the next step on (reference K3,
 * src/evolution.f90:179-187, which the reference runs every step for its
 * exx .. ezz snapshots, src/volume_snapshot.f90:220-236): each step adds
 * dt times the symmetric part of the velocity gradient that the element
 * kernel forms, in FP64 (6 x 8 B per element-local node, allocated and
 * zeroed by the first call, kept until s3d_destroy; 96 B of extra traffic
 * per node and step). Read with s3d_get_field(S3D_EXX .. S3D_EZZ). Call it
 * before the first step to match the reference. Refused for a batch
 * member. */
int s3d_enable_displacement_gradient(s3d_ctx *ctx);

/* Store the strain-rate parts that the volume curl fields combine with the
 * velocity, formed from the current velocity, i.e. the start velocity of the
 * next step (the reference's snapshot of rot_v* after step it mixes the
 * strain rates ses3d_evolution formed at the start of step it with the
 * velocity at its end, src/volume_snapshot.f90:238-245). After exactly one
 * more step, s3d_get_field(S3D_ROT_VX .. S3D_ROT_VZ) returns the fields.
 * The first call allocates 6 x 8 B per element-local node, kept until
 * s3d_destroy. No counterpart call in the reference, whose arrays hold the
 * strain rates anyway. */
int s3d_capture_curl_strain(s3d_ctx *ctx);

/* min and max of vx, vy, vz (reference v_ptp_loc, src/elastic_vars.F90:179-185). */
int s3d_vrange(s3d_ctx *ctx, double *out2);

/* The same min and max as s3d_vrange, computed without allocating anything:
 * no hipMalloc, hipFree or allocating library call. No counterpart in the
 * reference; it is for callers that cannot allocate device memory, such as
 * a time loop that reports the velocity range every step (as the reference's
 * loop does, src/ses3d-nt.F90:262-266) under a harness that forbids
 * allocation. The partial results live in one fixed scratch array per device
 * (16 MiB, held until the process ends), so the mesh may have at most
 * 268 435 456 unique nodes (256 per partial pair); larger meshes get an
 * error. Calls on one device share that array: calls that overlap on one
 * device (contexts on different host threads or streams) race. */
int s3d_vrange_noalloc(s3d_ctx *ctx, double *out2);

/* Kinetic energy (reference kinetic_energy, src/elastic_vars.F90:219-246). */
int s3d_kinetic_energy(s3d_ctx *ctx, double *out);

/* Block until all work queued on the context's stream is done. */
int s3d_synchronize(s3d_ctx *ctx);

/* Kernel timing with HIP events on the context's stream: while enabled,
 * every launch of the element and node kernels is bracketed by events.
 * s3d_kernel_times returns the summed milliseconds and launch counts of
 * (0) the element kernel, (1) the node kernel, (2) the receiver kernel since
 * the last reset. No counterpart in the reference (which prints no timing). */
int s3d_timing_enable(s3d_ctx *ctx, int enable);
int s3d_kernel_times(s3d_ctx *ctx, double *ms3, int *count3);

/* Device bytes held by the context. */
int s3d_device_bytes(s3d_ctx *ctx, double *bytes);

/* Batches: several initialised simulations on one device advance together.
 * Every step runs, for all members with the same GLL order / precision /
 * visco / compensation, ONE element kernel, ONE node kernel and ONE receiver
 * kernel (workgroups partitioned across members), and chunks of steps replay
 * one captured hipGraph. Members keep their own state, sources, receivers
 * and step counters, and may differ in mesh size and GLL order. No
 * counterpart in the reference, which runs one event per process group;
 * this is the batch axis of the workloads with many independent regional
 * simulations. A member may not be stepped on its own while its batch steps.
 * When every member of a group shares one model (s3d_create_shared), the
 * group's workgroups run source-innermost: all members' workgroups of one
 * element tile follow each other, and the shared static fields are read
 * from HBM once per tile. */
typedef struct s3d_batch s3d_batch;
int s3d_batch_create(int n, s3d_ctx *const *sims, s3d_batch **out);
int s3d_batch_step(s3d_batch *batch, int nsteps);
int s3d_batch_synchronize(s3d_batch *batch);
/* Members cannot be destroyed while a batch holds them: destroy the batch
 * first. */
int s3d_batch_destroy(s3d_batch *batch);

#ifdef __cplusplus
}
#endif
#endif
