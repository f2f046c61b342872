!> ISO_C_BINDING interfaces of the C ABI in include/ses3d_hip.h.
!!
!! Every procedure returns an INTEGER(c_int) status, 0 on success; the
!! message of a failure comes from s3d_error_message(). check() turns a
!! nonzero status into an abort, like the reference's abort
!! (src/misc/error_mod.F90:79-102).
MODULE ses3d_hip_c
    USE, INTRINSIC :: ISO_C_BINDING
    USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : ERROR_UNIT
    IMPLICIT NONE

    PRIVATE

    TYPE, BIND(C), PUBLIC :: s3d_grid
        INTEGER(c_int) :: nx, ny, nz, lpd
        REAL(c_double) :: lat_min, lat_max, lon_min, lon_max
        REAL(c_double) :: rad_min, rad_max
        INTEGER(c_int) :: taper_width
        REAL(c_double) :: taper_slope
        REAL(c_double) :: dt
        INTEGER(c_int) :: nt
    END TYPE

    TYPE, BIND(C), PUBLIC :: s3d_options
        INTEGER(c_int) :: precision = 1
        INTEGER(c_int) :: kahan_v = 1
        INTEGER(c_int) :: kahan_sigma = 0
        INTEGER(c_int) :: device = 0
        INTEGER(c_int) :: xcd_remap = 1
        INTEGER(c_int) :: bf16_static = 0
        INTEGER(c_int) :: state_buffers = 0
        INTEGER(c_int) :: memory_vars = 0
    END TYPE

    INTEGER(c_int), PARAMETER, PUBLIC :: S3D_FP64 = 0, S3D_FP32 = 1
    INTEGER(c_int), PARAMETER, PUBLIC :: S3D_RHOINV = 0, S3D_LAMBDA = 1, &
        S3D_MU = 2, S3D_A = 3, S3D_B = 4, S3D_C = 5, S3D_Q = 6
    INTEGER(c_int), PARAMETER, PUBLIC :: S3D_VX = 0, S3D_VY = 1, S3D_VZ = 2, &
        S3D_SXX = 3, S3D_SXY = 4, S3D_SXZ = 5, S3D_SYY = 6, S3D_SYZ = 7, S3D_SZZ = 8, &
        S3D_MM = 18, S3D_TAPER = 19, S3D_GRAD_RHO = 20, S3D_GRAD_CS = 21, &
        S3D_GRAD_CP = 22, S3D_GRAD_CSH = 23, S3D_GRAD_CSV = 24, &
        S3D_EXX = 26, S3D_EXY = 27, S3D_EXZ = 28, S3D_EYY = 29, S3D_EYZ = 30, S3D_EZZ = 31, &
        S3D_ROT_VX = 50, S3D_ROT_VY = 51, S3D_ROT_VZ = 52
    INTEGER(c_int), PARAMETER, PUBLIC :: S3D_CH_VX = 0, S3D_CH_VY = 1, &
        S3D_CH_VZ = 2, S3D_CH_N = 3, S3D_CH_E = 4, S3D_CH_Z = 5, &
        S3D_CH_PRESSURE = 6, S3D_CH_ROT_VX = 7, S3D_CH_ROT_VY = 8, S3D_CH_ROT_VZ = 9, &
        S3D_CH_MXX = 10, S3D_CH_MXY = 11, S3D_CH_MXZ = 12, S3D_CH_MYY = 13, &
        S3D_CH_MYZ = 14, S3D_CH_MZZ = 15, S3D_CH_FX = 16, S3D_CH_FY = 17, S3D_CH_FZ = 18
    INTEGER(c_int), PARAMETER, PUBLIC :: S3D_MAX_CHANNELS = 32
    INTEGER(c_int), PARAMETER, PUBLIC :: S3D_RICKER = 0, S3D_GAUSS = 1, &
        S3D_SIN3 = 2, S3D_BRUNE = 3, S3D_DELTA = 4

    PUBLIC :: s3d_last_error, s3d_device_count, s3d_create, s3d_destroy, &
              s3d_set_model_slab, s3d_set_model, s3d_init, s3d_cfl, s3d_gll, &
              s3d_locate, s3d_wavelet, s3d_add_source_mt, s3d_add_source_sf, &
              s3d_add_receiver, s3d_step, s3d_current_step, s3d_get_trace, &
              s3d_get_field, s3d_set_field, s3d_vrange, s3d_kinetic_energy, &
              s3d_synchronize, s3d_autotune, s3d_set_tile, s3d_create_shared, &
              s3d_batch_create, s3d_batch_step, s3d_batch_synchronize, s3d_batch_destroy, &
              s3d_grad_add, s3d_vrange_noalloc, s3d_set_lds_layout, s3d_get_tuning, &
              s3d_enable_displacement_gradient, s3d_capture_curl_strain, &
              s3d_set_store_order, s3d_get_store_order, &
              s3d_timing_enable, s3d_kernel_times, s3d_device_bytes
    PUBLIC :: check, s3d_error_message

    INTERFACE
        FUNCTION s3d_last_error() BIND(C, NAME='s3d_last_error')
            IMPORT :: c_ptr
            TYPE(c_ptr) :: s3d_last_error
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_device_count(n) BIND(C, NAME='s3d_device_count')
            IMPORT :: c_int
            INTEGER(c_int), INTENT(OUT) :: n
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_create(grid, opt, ctx) BIND(C, NAME='s3d_create')
            IMPORT :: c_int, c_ptr, s3d_grid, s3d_options
            TYPE(s3d_grid), INTENT(IN) :: grid
            TYPE(s3d_options), INTENT(IN) :: opt
            TYPE(c_ptr), INTENT(OUT) :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_destroy(ctx) BIND(C, NAME='s3d_destroy')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_set_model_slab(ctx, param, ez0, nez, data) &
                BIND(C, NAME='s3d_set_model_slab')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: param, ez0, nez
            REAL(c_double), INTENT(IN) :: data(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_set_model(ctx, param, data) BIND(C, NAME='s3d_set_model')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: param
            REAL(c_double), INTENT(IN) :: data(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_init(ctx) BIND(C, NAME='s3d_init')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_cfl(ctx, courant) BIND(C, NAME='s3d_cfl')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: courant
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_gll(lpd, knots, weights, dl) BIND(C, NAME='s3d_gll')
            IMPORT :: c_int, c_double
            INTEGER(c_int), VALUE :: lpd
            REAL(c_double), INTENT(OUT) :: knots(*), weights(*), dl(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_locate(ctx, lat, lon, depth, exyz, vm) BIND(C, NAME='s3d_locate')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), VALUE :: lat, lon, depth
            INTEGER(c_int), INTENT(OUT) :: exyz(3)
            REAL(c_double), INTENT(OUT) :: vm(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_wavelet(wavelet, onset, width, dt, nt, out) &
                BIND(C, NAME='s3d_wavelet')
            IMPORT :: c_int, c_double
            INTEGER(c_int), VALUE :: wavelet, nt
            REAL(c_double), VALUE :: onset, width, dt
            REAL(c_double), INTENT(OUT) :: out(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_add_source_mt(ctx, exyz, vm, stf6) BIND(C, NAME='s3d_add_source_mt')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(IN) :: exyz(3)
            REAL(c_double), INTENT(IN) :: vm(*), stf6(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_add_source_sf(ctx, exyz, vm, stf3) BIND(C, NAME='s3d_add_source_sf')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(IN) :: exyz(3)
            REAL(c_double), INTENT(IN) :: vm(*), stf3(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_add_receiver(ctx, exyz, vm, nch, channels, id) &
                BIND(C, NAME='s3d_add_receiver')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(IN) :: exyz(3)
            REAL(c_double), INTENT(IN) :: vm(*)
            INTEGER(c_int), VALUE :: nch
            INTEGER(c_int), INTENT(IN) :: channels(*)
            INTEGER(c_int), INTENT(OUT) :: id
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_step(ctx, nsteps) BIND(C, NAME='s3d_step')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: nsteps
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_current_step(ctx, it) BIND(C, NAME='s3d_current_step')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(OUT) :: it
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_get_trace(ctx, rec, ch, out) BIND(C, NAME='s3d_get_trace')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: rec, ch
            REAL(c_double), INTENT(OUT) :: out(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_get_field(ctx, field, out) BIND(C, NAME='s3d_get_field')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: field
            REAL(c_double), INTENT(OUT) :: out(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_set_field(ctx, field, in) BIND(C, NAME='s3d_set_field')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: field
            REAL(c_double), INTENT(IN) :: in(*)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_grad_add(ctx, vx_fw, vy_fw, vz_fw, increment) &
                BIND(C, NAME='s3d_grad_add')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(IN) :: vx_fw(*), vy_fw(*), vz_fw(*)
            INTEGER(c_int), VALUE :: increment
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_vrange(ctx, out2) BIND(C, NAME='s3d_vrange')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: out2(2)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_kinetic_energy(ctx, out) BIND(C, NAME='s3d_kinetic_energy')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: out
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_synchronize(ctx) BIND(C, NAME='s3d_synchronize')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_autotune(ctx, tile, ms) BIND(C, NAME='s3d_autotune')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(OUT) :: tile
            REAL(c_double), INTENT(OUT) :: ms
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_set_tile(ctx, tile) BIND(C, NAME='s3d_set_tile')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: tile
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_set_lds_layout(ctx, layout) BIND(C, NAME='s3d_set_lds_layout')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: layout
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_get_tuning(ctx, tile, xcd_remap, lds_layout) &
                BIND(C, NAME='s3d_get_tuning')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(OUT) :: tile, xcd_remap, lds_layout
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_enable_displacement_gradient(ctx) &
                BIND(C, NAME='s3d_enable_displacement_gradient')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_capture_curl_strain(ctx) BIND(C, NAME='s3d_capture_curl_strain')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_set_store_order(ctx, early) BIND(C, NAME='s3d_set_store_order')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: early
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_get_store_order(ctx, early) BIND(C, NAME='s3d_get_store_order')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), INTENT(OUT) :: early
        END FUNCTION
        ! the v range without any allocation (no counterpart in the reference)
        INTEGER(c_int) FUNCTION s3d_vrange_noalloc(ctx, out2) BIND(C, NAME='s3d_vrange_noalloc')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: out2(2)
        END FUNCTION
        ! kernel timing with HIP events and device memory held
        INTEGER(c_int) FUNCTION s3d_timing_enable(ctx, enable) BIND(C, NAME='s3d_timing_enable')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: ctx
            INTEGER(c_int), VALUE :: enable
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_kernel_times(ctx, ms3, count3) BIND(C, NAME='s3d_kernel_times')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: ms3(3)
            INTEGER(c_int), INTENT(OUT) :: count3(3)
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_device_bytes(ctx, bytes) BIND(C, NAME='s3d_device_bytes')
            IMPORT :: c_int, c_ptr, c_double
            TYPE(c_ptr), VALUE :: ctx
            REAL(c_double), INTENT(OUT) :: bytes
        END FUNCTION
        ! further simulations on one model, and batches (no counterpart in
        ! the reference; see include/ses3d_hip.h)
        INTEGER(c_int) FUNCTION s3d_create_shared(model, out) BIND(C, NAME='s3d_create_shared')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: model
            TYPE(c_ptr), INTENT(OUT) :: out
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_batch_create(n, sims, out) BIND(C, NAME='s3d_batch_create')
            IMPORT :: c_int, c_ptr
            INTEGER(c_int), VALUE :: n
            TYPE(c_ptr), INTENT(IN) :: sims(*)
            TYPE(c_ptr), INTENT(OUT) :: out
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_batch_step(batch, nsteps) BIND(C, NAME='s3d_batch_step')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: batch
            INTEGER(c_int), VALUE :: nsteps
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_batch_synchronize(batch) BIND(C, NAME='s3d_batch_synchronize')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: batch
        END FUNCTION
        INTEGER(c_int) FUNCTION s3d_batch_destroy(batch) BIND(C, NAME='s3d_batch_destroy')
            IMPORT :: c_int, c_ptr
            TYPE(c_ptr), VALUE :: batch
        END FUNCTION
        FUNCTION c_strlen(s) BIND(C, NAME='strlen')
            IMPORT :: c_ptr, c_size_t
            TYPE(c_ptr), VALUE :: s
            INTEGER(c_size_t) :: c_strlen
        END FUNCTION
    END INTERFACE

CONTAINS

    !> Message of the last failed call.
    FUNCTION s3d_error_message() RESULT(msg)
        CHARACTER(LEN=:), ALLOCATABLE :: msg
        TYPE(c_ptr) :: p
        CHARACTER(KIND=c_char), POINTER :: chars(:)
        INTEGER :: n, i
        p = s3d_last_error()
        n = INT(c_strlen(p))
        CALL C_F_POINTER(p, chars, [n])
        ALLOCATE(CHARACTER(LEN=n) :: msg)
        DO i = 1, n
            msg(i:i) = chars(i)
        END DO
    END FUNCTION

    !> Abort with the library's message when a call failed.
    SUBROUTINE check(status, what)
        INTEGER(c_int), INTENT(IN) :: status
        CHARACTER(LEN=*), INTENT(IN) :: what
        IF (status /= 0) THEN
            WRITE(ERROR_UNIT, '("ERROR: ", A, ": ", A)') what, s3d_error_message()
            ERROR STOP 1
        END IF
    END SUBROUTINE

END MODULE ses3d_hip_c
