!> Time evolution with the reference's interface (src/evolution.f90:41):
!! init_temp() prepares the step, ses3d_evolution() advances it once.
!! The device context lives here; model upload, geometry, mass matrix and the
!! derivative matrix are set up by init_temp (reference init_temp,
!! init_geometric_parameters and init_auxiliary_model_parameters).
MODULE evolution_mod
    USE, INTRINSIC :: ISO_C_BINDING
    USE ses3d_hip_c
    USE configuration_mod, ONLY : rk, configuration, configuration_typ, abort
    IMPLICIT NONE
    PRIVATE

    TYPE(c_ptr), PUBLIC :: ctx = c_null_ptr

    PUBLIC :: init_temp, ses3d_evolution, v_ptp_loc, finalize_evolution, &
              write_model_parameters, model_param

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

    !> Built-in 1-D model at the radii of the GLL points
    !! (make_model_parameters, src/model_paras.F90:206-343): v(:, ez, k) =
    !! rhoinv, lambda, mu, a, b, c. r runs from rad_max at the top element
    !! layer (ez = 0) to rad_min (src/geometric_paras.f90:153-161) and is
    !! computed here in FP64.
    SUBROUTINE builtin_profile(cf, v)
        USE gll_mod, ONLY : get_knots
        USE earth_models_mod, ONLY : builtin_model
        TYPE(configuration_typ), INTENT(IN) :: cf
        REAL(rk), INTENT(OUT) :: v(6, 0:cf%nz - 1, 0:cf%lpd)
        REAL(rk) :: knots(cf%lpd + 1), r
        INTEGER :: ez, k
        LOGICAL :: ok

        knots = 0.5_rk * (1.0_rk + get_knots(cf%lpd))
        DO k = 0, cf%lpd
            DO ez = 0, cf%nz - 1
                r = cf%rad_max + (ez + knots(k + 1)) * ((cf%rad_min - cf%rad_max) / cf%nz)
                IF (ez == cf%nz - 1 .AND. k == cf%lpd) r = cf%rad_min
                CALL builtin_model(cf%model_type, r, ez, v(1, ez, k), v(2, ez, k), &
                                   v(3, ez, k), v(4, ez, k), v(5, ez, k), v(6, ez, k), ok)
                IF (.NOT. ok) CALL abort('Unknown model_type')
                v(1, ez, k) = 1.0_rk / v(1, ez, k)
            END DO
        END DO
    END SUBROUTINE

    !> Expands a profile over (ez, k) to the reference's global layout
    !! (x, y, z elements, then i, j, k GLL indices; x fastest).
    SUBROUTINE expand_profile(cf, prof, buf)
        TYPE(configuration_typ), INTENT(IN) :: cf
        REAL(rk), INTENT(IN) :: prof(0:cf%nz - 1, 0:cf%lpd)
        REAL(c_double), INTENT(OUT) :: buf(:)
        INTEGER(8) :: plane, o
        INTEGER :: ez, k, q
        plane = INT(cf%nx, 8) * cf%ny
        o = 0
        DO k = 0, cf%lpd
            DO q = 1, (cf%lpd + 1)**2
                DO ez = 0, cf%nz - 1
                    buf(o + 1:o + plane) = prof(ez, k)
                    o = o + plane
                END DO
            END DO
        END DO
    END SUBROUTINE

    !> Uploads the built-in model selected by model_type.
    SUBROUTINE set_builtin_model(cf, buf)
        TYPE(configuration_typ), INTENT(IN) :: cf
        REAL(c_double), INTENT(INOUT) :: buf(:)
        INTEGER(c_int), PARAMETER :: ids(6) = [S3D_RHOINV, S3D_LAMBDA, S3D_MU, S3D_A, &
                                               S3D_B, S3D_C]
        REAL(rk) :: v(6, 0:cf%nz - 1, 0:cf%lpd)
        INTEGER :: p
        CALL builtin_profile(cf, v)
        DO p = 1, 6
            CALL expand_profile(cf, v(p, :, :), buf)
            CALL check(s3d_set_model(ctx, ids(p), buf), 'set_model')
        END DO
    END SUBROUTINE

    !> Workflow 'model' (src/ses3d-nt.F90:326-329, write_model_parameters at
    !! src/model_paras.F90:140-202): writes the built-in model to the
    !! parameter files named in the &model group, as raw native-endian
    !! REAL32 in the global layout. Q is written as HUGE (elastic), the
    !! value the built-in models leave it at. An existing file is replaced
    !! only with override = .TRUE.
    SUBROUTINE write_model_parameters()
        REAL(rk), ALLOCATABLE :: v(:, :, :)
        REAL(c_double), ALLOCATABLE :: buf(:)
        INTEGER(8) :: n
        INTEGER :: p, lun, ios
        LOGICAL :: there

        ASSOCIATE (cf => configuration)
        n = INT(cf%nx, 8) * cf%ny * cf%nz * (cf%lpd + 1)**3
        ALLOCATE(v(6, 0:cf%nz - 1, 0:cf%lpd), buf(n))
        CALL builtin_profile(cf, v)
        ! cf%file: rhoinv, lambda, mu, a, b, c, q -- the order of v
        DO p = 1, 7
            IF (cf%file(p) == 'N/A') CYCLE
            INQUIRE(FILE=cf%file(p), EXIST=there)
            IF (there .AND. .NOT. cf%override) &
                CALL abort('File "' // TRIM(cf%file(p)) // '" exists; set override = .TRUE.')
            IF (p <= 6) THEN
                CALL expand_profile(cf, v(p, :, :), buf)
            ELSE
                buf = REAL(HUGE(0.0), c_double)
            END IF
            OPEN(NEWUNIT=lun, FILE=cf%file(p), ACCESS='stream', FORM='unformatted', &
                 STATUS='replace', ACTION='write', IOSTAT=ios)
            IF (ios /= 0) CALL abort('Cannot write "' // TRIM(cf%file(p)) // '"')
            WRITE(lun) REAL(buf, 4)
            CLOSE(lun)
            WRITE(*, '(" writing: ", A)') TRIM(cf%file(p))
        END DO
        END ASSOCIATE
    END SUBROUTINE

    !> Model parameter p (1 rhoinv, 2 lambda, 3 mu, 4 a, 5 b, 6 c, 7 q) in
    !! the global 6-D layout, FP64, as init_temp uploads it: from the built-in
    !! model of model_type or from the parameter file. Without a file (and for
    !! q of a built-in model) a, b, c are 0 and q is HUGE (elastic).
    SUBROUTINE model_param(p, buf)
        INTEGER, INTENT(IN) :: p
        REAL(c_double), INTENT(OUT) :: buf(:)
        REAL(rk), ALLOCATABLE :: v(:, :, :)
        ASSOCIATE (cf => configuration)
        IF (cf%model_type /= -1 .AND. p <= 6) THEN
            ALLOCATE(v(6, 0:cf%nz - 1, 0:cf%lpd))
            CALL builtin_profile(cf, v)
            CALL expand_profile(cf, v(p, :, :), buf)
        ELSE IF (cf%model_type /= -1 .OR. cf%file(p) == 'N/A') THEN
            buf = MERGE(REAL(HUGE(0.0), c_double), 0.0_c_double, p == 7)
        ELSE
            CALL read_model_file(cf%file(p), SIZE(buf, KIND=8), buf)
        END IF
        END ASSOCIATE
    END SUBROUTINE

    SUBROUTINE init_temp()
        TYPE(s3d_grid) :: g
        TYPE(s3d_options) :: o
        INTEGER(8) :: n
        INTEGER :: p
        INTEGER(c_int), PARAMETER :: ids(7) = [S3D_RHOINV, S3D_LAMBDA, S3D_MU, S3D_A, &
                                               S3D_B, S3D_C, S3D_Q]
        REAL(c_double), ALLOCATABLE :: buf(:)
        REAL(c_double) :: courant, tile_ms
        INTEGER(c_int) :: tile

        ASSOCIATE (cf => configuration)
        g = s3d_grid(cf%nx, cf%ny, cf%nz, cf%lpd, cf%lat_min, cf%lat_max, cf%lon_min, &
                     cf%lon_max, cf%rad_min, cf%rad_max, cf%taper_width, cf%taper_slope, &
                     cf%dt, cf%nt)
        o%precision = MERGE(S3D_FP32, S3D_FP64, cf%precision == 'fp32')
        o%kahan_sigma = MERGE(MERGE(2, 1, cf%kahan_sigma_bf16), 0, cf%kahan_sigma)
        o%kahan_v = MERGE(MERGE(2, 1, cf%kahan_v_bf16), 0, cf%kahan_v)
        o%bf16_static = MERGE(1, 0, cf%bf16_static)
        o%device = cf%device
        o%state_buffers = cf%state_buffers
        o%memory_vars = cf%memory_vars
        CALL check(s3d_create(g, o, ctx), 'create')
        n = INT(cf%nx, 8) * cf%ny * cf%nz * (cf%lpd + 1)**3
        ALLOCATE(buf(n))
        IF (cf%model_type /= -1) THEN
            CALL set_builtin_model(cf, buf)
        ELSE
            DO p = 1, 7
                IF (cf%file(p) == 'N/A') CYCLE
                CALL read_model_file(cf%file(p), n, buf)
                CALL check(s3d_set_model(ctx, ids(p), buf), 'set_model')
            END DO
        END IF
        DEALLOCATE(buf)
        CALL check(s3d_init(ctx), 'init')
        IF (cf%precision == 'fp32' .AND. .NOT. cf%bf16_static) THEN
            IF (cf%tile == -1) THEN
                CALL check(s3d_autotune(ctx, tile, tile_ms), 'autotune')
            ELSE
                CALL check(s3d_set_tile(ctx, INT(cf%tile, c_int)), 'set_tile')
            END IF
        END IF
        CALL check(s3d_cfl(ctx, courant), 'cfl')
        WRITE(*, '("Courant number:", F5.2)') courant
        END ASSOCIATE
    END SUBROUTINE

    !> One iteration of the forward loop body on the device: evolution,
    !! source injection and receiver sampling (src/ses3d-nt.F90:242-256).
    SUBROUTINE ses3d_evolution()
        CALL check(s3d_step(ctx, 1_c_int), 'step')
    END SUBROUTINE

    !> Peak-to-peak velocity (src/elastic_vars.F90:179-185): the
    !! allocation-free reduction, or the allocating one on meshes beyond its
    !! fixed scratch size.
    FUNCTION v_ptp_loc() RESULT(ptp)
        REAL(rk) :: ptp(2)
        IF (s3d_vrange_noalloc(ctx, ptp) /= 0) CALL check(s3d_vrange(ctx, ptp), 'vrange')
    END FUNCTION

    SUBROUTINE finalize_evolution()
        IF (s3d_destroy(ctx) /= 0) CALL abort('destroy failed')
        ctx = c_null_ptr
    END SUBROUTINE

END MODULE evolution_mod
