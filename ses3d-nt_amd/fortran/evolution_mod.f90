!> Time evolution with the reference's interface (src/evolution.f90:41):
!! init_temp() prepares the step, ses3d_evolution() advances it once.
!! The device context lives here; model upload, geometry, mass matrix and the
!! derivative matrix are set up by init_temp (reference init_temp,
!! init_geometric_parameters and init_auxiliary_model_parameters).
MODULE evolution_mod
    USE, INTRINSIC :: ISO_C_BINDING
    USE ses3d_hip_c
    USE configuration_mod, ONLY : rk, configuration, abort
    IMPLICIT NONE
    PRIVATE

    TYPE(c_ptr), PUBLIC :: ctx = c_null_ptr

    PUBLIC :: init_temp, ses3d_evolution, v_ptp_loc, finalize_evolution

CONTAINS

    !> Model parameters from raw files: the reference's global 6-D layout,
    !! native endianness, float64 or float32 (src/mpi_parallel_io.F90:61-341).
    SUBROUTINE read_model_file(file, n, out)
        CHARACTER(LEN=*), INTENT(IN) :: file
        INTEGER(8), INTENT(IN) :: n
        REAL(c_double), INTENT(OUT) :: out(n)
        INTEGER :: lun, ios
        INTEGER(8) :: sz
        REAL(4), ALLOCATABLE :: f4(:)
        OPEN(NEWUNIT=lun, FILE=file, ACCESS='stream', FORM='unformatted', &
             STATUS='old', ACTION='read', IOSTAT=ios)
        IF (ios /= 0) CALL abort('Model-file "' // TRIM(file) // '" does not exist!')
        INQUIRE(UNIT=lun, SIZE=sz)
        IF (sz == 8 * n) THEN
            READ(lun) out
        ELSE IF (sz == 4 * n) THEN
            ALLOCATE(f4(n))
            READ(lun) f4
            out = REAL(f4, c_double)
        ELSE
            CALL abort('Model-file "' // TRIM(file) // '" has the wrong size')
        END IF
        CLOSE(lun)
    END SUBROUTINE

    SUBROUTINE init_temp()
        TYPE(s3d_grid) :: g
        TYPE(s3d_options) :: o
        INTEGER(8) :: n
        INTEGER :: p
        INTEGER(c_int), PARAMETER :: ids(7) = [S3D_RHOINV, S3D_LAMBDA, S3D_MU, S3D_A, &
                                               S3D_B, S3D_C, S3D_Q]
        REAL(c_double), ALLOCATABLE :: buf(:)
        REAL(c_double) :: courant

        ASSOCIATE (cf => configuration)
        g = s3d_grid(cf%nx, cf%ny, cf%nz, cf%lpd, cf%lat_min, cf%lat_max, cf%lon_min, &
                     cf%lon_max, cf%rad_min, cf%rad_max, cf%taper_width, cf%taper_slope, &
                     cf%dt, cf%nt)
        o%precision = MERGE(S3D_FP32, S3D_FP64, cf%precision == 'fp32')
        o%kahan_v = MERGE(1, 0, cf%kahan_v)
        o%kahan_sigma = MERGE(MERGE(2, 1, cf%kahan_sigma_bf16), 0, cf%kahan_sigma)
        o%bf16_static = MERGE(1, 0, cf%bf16_static)
        o%device = cf%device
        CALL check(s3d_create(g, o, ctx), 'create')
        n = INT(cf%nx, 8) * cf%ny * cf%nz * (cf%lpd + 1)**3
        ALLOCATE(buf(n))
        IF (cf%model_type == 1) THEN
            ! homogeneous model (src/models/homogeneous.f90): rho 3000,
            ! mu 7e10, lambda 1.1e11 (single-precision literals)
            buf = 1.0_c_double / REAL(3000.0, c_double)
            CALL check(s3d_set_model(ctx, S3D_RHOINV, buf), 'set_model')
            buf = REAL(1.1e11, c_double)
            CALL check(s3d_set_model(ctx, S3D_LAMBDA, buf), 'set_model')
            buf = REAL(7.0e10, c_double)
            CALL check(s3d_set_model(ctx, S3D_MU, buf), 'set_model')
        ELSE
            DO p = 1, 7
                IF (cf%file(p) == 'N/A') CYCLE
                CALL read_model_file(cf%file(p), n, buf)
                CALL check(s3d_set_model(ctx, ids(p), buf), 'set_model')
            END DO
        END IF
        DEALLOCATE(buf)
        CALL check(s3d_init(ctx), 'init')
        CALL check(s3d_cfl(ctx, courant), 'cfl')
        WRITE(*, '("Courant number:", F5.2)') courant
        END ASSOCIATE
    END SUBROUTINE

    !> One iteration of the forward loop body on the device: evolution,
    !! source injection and receiver sampling (src/ses3d-nt.F90:242-256).
    SUBROUTINE ses3d_evolution()
        CALL check(s3d_step(ctx, 1_c_int), 'step')
    END SUBROUTINE

    !> Peak-to-peak velocity (src/elastic_vars.F90:179-185).
    FUNCTION v_ptp_loc() RESULT(ptp)
        REAL(rk) :: ptp(2)
        CALL check(s3d_vrange(ctx, ptp), 'vrange')
    END FUNCTION

    SUBROUTINE finalize_evolution()
        IF (s3d_destroy(ctx) /= 0) CALL abort('destroy failed')
        ctx = c_null_ptr
    END SUBROUTINE

END MODULE evolution_mod
