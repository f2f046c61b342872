!> Sources and receivers with the reference's class interface
!! (source_cls%add, receiver_cls%record / %write;
!! src/config/source.f90:359-370, src/config/receiver.f90:136-192).
!!
!! On the GPU the moment-tensor injection and the receiver sampling run inside
!! the time step itself (s3d_step), so add() and record() only check that the
!! device has advanced to step it; write() copies the recorded traces back and
!! writes SAC files.
MODULE channels_mod
    USE, INTRINSIC :: ISO_C_BINDING
    USE ses3d_hip_c
    USE configuration_mod, ONLY : rk, configuration, source_entry, receiver_entry, abort, warn, &
                                  lower, max_attributes
    USE sac_io_mod, ONLY : write_sac, read_sac, sac_file_name
    IMPLICIT NONE
    PRIVATE

    TYPE, PUBLIC :: source_cls
        INTEGER :: ex = -1, ey = -1, ez = -1
    CONTAINS
        PROCEDURE :: add => source_add
    END TYPE

    TYPE, PUBLIC :: receiver_cls
        INTEGER :: id = -1
        TYPE(receiver_entry) :: info
        INTEGER :: nch = 0
        INTEGER(c_int) :: codes(max_attributes)
        CHARACTER(LEN=8) :: names(max_attributes)
    CONTAINS
        PROCEDURE :: record => receiver_record
        PROCEDURE :: write => receiver_write
    END TYPE

    PUBLIC :: parse_source_file, parse_receiver_file, count_receiver_channels

CONTAINS

    INTEGER(c_int) FUNCTION wavelet_code(name) RESULT(code)
        CHARACTER(LEN=*), INTENT(IN) :: name
        SELECT CASE (TRIM(name))
        CASE ('RICKER'); code = S3D_RICKER
        CASE ('GAUSS');  code = S3D_GAUSS
        CASE ('SIN**3'); code = S3D_SIN3
        CASE ('BRUNE');  code = S3D_BRUNE
        CASE ('DELTA');  code = S3D_DELTA
        CASE DEFAULT;    code = -1
        END SELECT
    END FUNCTION

    !> Read an ASCII source-time function, one value per line
    !! (reference read_stf, src/config/source.f90:662-687).
    SUBROUTINE read_stf(file, out)
        CHARACTER(LEN=*), INTENT(IN) :: file
        REAL(rk), INTENT(OUT) :: out(:)
        INTEGER :: lun, ios, i
        OPEN(NEWUNIT=lun, FILE=file, STATUS='old', ACTION='read', IOSTAT=ios)
        IF (ios /= 0) CALL abort('Source-time function file not found: ' // TRIM(file))
        out = 0.0_rk
        DO i = 1, SIZE(out)
            READ(lun, *, IOSTAT=ios) out(i)
            IF (ios /= 0) EXIT
        END DO
        CLOSE(lun)
    END SUBROUTINE

    !> A &source_sac trace (create_from_sac, src/config/source.f90:142-203):
    !! location from STLA/STLO/STDP, one channel named by KCMPNM, signature
    !! data * SCALE. On return stf6 (moment tensor, xx xy xz yy yz zz) or stf3
    !! (single force, x y z) holds the signature in its channel and zeros
    !! elsewhere; kind = 1 (moment), 2 (force) or 0 (unknown channel).
    SUBROUTINE sac_source(s, nt, stf6, stf3, kind)
        TYPE(source_entry), INTENT(INOUT) :: s
        INTEGER, INTENT(IN) :: nt
        REAL(c_double), INTENT(OUT) :: stf6(6 * nt), stf3(3 * nt)
        INTEGER, INTENT(OUT) :: kind
        REAL(8), ALLOCATABLE :: data(:)
        REAL(8) :: scale
        CHARACTER(LEN=8) :: ch
        INTEGER :: ios, c

        CALL read_sac(TRIM(s%file), data, s%lat, s%lon, s%depth, scale, ch, ios)
        IF (ios /= 0) CALL abort('Cannot read SAC-file "' // TRIM(s%file) // '"')
        IF (SIZE(data) < nt) CALL abort('SAC-trace too short!')
        stf6 = 0.0_c_double
        stf3 = 0.0_c_double
        ! channel names of src/config/channel_mod.f90:300-330
        SELECT CASE (TRIM(ch))
        CASE ('mxx'); kind = 1; c = 1
        CASE ('mxy', 'myx'); kind = 1; c = 2
        CASE ('mxz', 'mzx'); kind = 1; c = 3
        CASE ('myy'); kind = 1; c = 4
        CASE ('myz', 'mzy'); kind = 1; c = 5
        CASE ('mzz'); kind = 1; c = 6
        CASE ('fx'); kind = 2; c = 1
        CASE ('fy'); kind = 2; c = 2
        CASE ('fz'); kind = 2; c = 3
        CASE DEFAULT
            WRITE(*, '("WARNING: Unknown attribute ''", A, "'' in SAC source.")') TRIM(ch)
            kind = 0
            RETURN
        END SELECT
        IF (kind == 1) THEN
            stf6((c - 1) * nt + 1:c * nt) = data(1:nt) * scale
        ELSE
            stf3((c - 1) * nt + 1:c * nt) = data(1:nt) * scale
        END IF
    END SUBROUTINE

    !> Locate every source, build its signature and hand it to the device
    !! (src/config/source.f90:274-327, 449-536). At least one source must
    !! lie in the domain.
    !> Locate a point (lat, lon in degrees, depth in m) with s3d_locate and
    !! warn as the reference's point constructor does
    !! (src/config/point.f90:207-236): a point outside the domain is dropped,
    !! a point in the tapering boundaries (in_boundary, :686-712: within
    !! taper_width elements of the four sides or the bottom) is kept.
    LOGICAL FUNCTION located(ctx, lat, lon, depth, exyz, vm)
        TYPE(c_ptr), INTENT(IN) :: ctx
        REAL(c_double), INTENT(IN) :: lat, lon, depth
        INTEGER(c_int), INTENT(OUT) :: exyz(3)
        REAL(c_double), INTENT(OUT) :: vm(:)
        REAL(c_double), PARAMETER :: pi = 3.141592653589793238_c_double, &
            earth_radius = 6371000.0_c_double
        REAL(c_double) :: x, y, z, x0, x1, y0, y1, z0, z1, tw
        CHARACTER(LEN=256) :: msg
        WRITE(msg, '("Point (", G0, ", ", G0, ", ", G0, ")")') lat, lon, depth
        located = s3d_locate(ctx, lat, lon, depth, exyz, vm) == 0
        IF (.NOT. located) THEN
            CALL warn(TRIM(msg) // ' not in computational domain!')
            RETURN
        END IF
        ASSOCIATE (cf => configuration)
        x = (90.0_c_double - lat) * pi / 180.0_c_double
        y = lon * pi / 180.0_c_double
        z = earth_radius - depth
        x0 = (90.0_c_double - cf%lat_max) * pi / 180.0_c_double
        x1 = (90.0_c_double - cf%lat_min) * pi / 180.0_c_double
        y0 = cf%lon_min * pi / 180.0_c_double
        y1 = cf%lon_max * pi / 180.0_c_double
        z0 = cf%rad_min
        z1 = cf%rad_max
        tw = REAL(cf%taper_width, c_double)
        IF (x0 + tw * (x1 - x0) / cf%nx > x .OR. x > x1 - tw * (x1 - x0) / cf%nx .OR. &
            y0 + tw * (y1 - y0) / cf%ny > y .OR. y > y1 - tw * (y1 - y0) / cf%ny .OR. &
            z0 + tw * (z1 - z0) / cf%nz > z) &
            CALL warn(TRIM(msg) // ' located in tapering boundaries!')
        END ASSOCIATE
    END FUNCTION

    SUBROUTINE parse_source_file(ctx, sources)
        TYPE(c_ptr), INTENT(IN) :: ctx
        TYPE(source_cls), ALLOCATABLE, INTENT(OUT) :: sources(:)
        INTEGER :: i, c, nt, n, kind
        INTEGER(c_int) :: exyz(3), code
        REAL(c_double), ALLOCATABLE :: vm(:), stf(:), comp(:), comp3(:)
        TYPE(source_entry) :: s
        TYPE(source_cls), ALLOCATABLE :: tmp(:)

        nt = configuration%nt
        ALLOCATE(vm((configuration%lpd + 1)**3), stf(nt), tmp(SIZE(configuration%sources)))
        n = 0
        DO i = 1, SIZE(configuration%sources)
            s = configuration%sources(i)
            IF (s%sac) THEN
                ALLOCATE(comp(6 * nt), comp3(3 * nt))
                CALL sac_source(s, nt, comp, comp3, kind)
                IF (kind == 0) THEN
                    DEALLOCATE(comp, comp3)
                    CYCLE
                END IF
            END IF
            IF (.NOT. located(ctx, s%lat, s%lon, s%depth, exyz, vm)) THEN
                IF (ALLOCATED(comp)) DEALLOCATE(comp)
                IF (ALLOCATED(comp3)) DEALLOCATE(comp3)
                CYCLE
            END IF
            IF (s%sac) THEN
                IF (kind == 1) THEN
                    CALL check(s3d_add_source_mt(ctx, exyz, vm, comp), 'add_source_mt')
                ELSE
                    CALL check(s3d_add_source_sf(ctx, exyz, vm, comp3), 'add_source_sf')
                END IF
                DEALLOCATE(comp, comp3)
            ELSE
                code = wavelet_code(s%wavelet)
                IF (code >= 0) THEN
                    CALL check(s3d_wavelet(code, s%onset, s%width, configuration%dt, INT(nt, c_int), stf), &
                               'wavelet')
                ELSE
                    CALL read_stf(s%wavelet, stf)
                END IF
                IF (s%moment) THEN
                    ALLOCATE(comp(6 * nt))
                    DO c = 1, 6
                        comp((c - 1) * nt + 1:c * nt) = stf * s%mt(c)
                    END DO
                    CALL check(s3d_add_source_mt(ctx, exyz, vm, comp), 'add_source_mt')
                ELSE
                    ALLOCATE(comp(3 * nt))
                    DO c = 1, 3
                        comp((c - 1) * nt + 1:c * nt) = stf * s%dir(c)
                    END DO
                    CALL check(s3d_add_source_sf(ctx, exyz, vm, comp), 'add_source_sf')
                END IF
                DEALLOCATE(comp)
            END IF
            n = n + 1
            tmp(n) = source_cls(exyz(1), exyz(2), exyz(3))
        END DO
        IF (n < 1) CALL abort('No sources defined in config file. At least one source must be specified.')
        sources = tmp(1:n)
    END SUBROUTINE

    !> Channel code of a lower-case receiver attribute (the names of
    !! construct_channel_container_receiver, src/config/channel_mod.f90:270-336,
    !! and pressure); -1 when unknown.
    INTEGER(c_int) FUNCTION channel_code(name) RESULT(code)
        CHARACTER(LEN=*), INTENT(IN) :: name
        SELECT CASE (TRIM(name))
        CASE ('vx'); code = S3D_CH_VX
        CASE ('vy'); code = S3D_CH_VY
        CASE ('vz'); code = S3D_CH_VZ
        CASE ('n');  code = S3D_CH_N
        CASE ('e');  code = S3D_CH_E
        CASE ('z');  code = S3D_CH_Z
        CASE ('pressure'); code = S3D_CH_PRESSURE
        CASE ('rot_vx'); code = S3D_CH_ROT_VX
        CASE ('rot_vy'); code = S3D_CH_ROT_VY
        CASE ('rot_vz'); code = S3D_CH_ROT_VZ
        CASE ('mxx'); code = S3D_CH_MXX
        CASE ('mxy', 'myx'); code = S3D_CH_MXY
        CASE ('mxz', 'mzx'); code = S3D_CH_MXZ
        CASE ('myy'); code = S3D_CH_MYY
        CASE ('myz', 'mzy'); code = S3D_CH_MYZ
        CASE ('mzz'); code = S3D_CH_MZZ
        CASE ('fx'); code = S3D_CH_FX
        CASE ('fy'); code = S3D_CH_FY
        CASE ('fz'); code = S3D_CH_FZ
        CASE DEFAULT; code = -1
        END SELECT
    END FUNCTION

    !> Channels of a receiver entry as the reference builds them
    !! (construct_receiver_cls, src/config/receiver.f90:379-392): empty
    !! attributes are dropped, the others lower-cased; unknown ones are
    !! reported in one warning and dropped.
    SUBROUTINE receiver_channels(r)
        TYPE(receiver_cls), INTENT(INOUT) :: r
        INTEGER :: j
        INTEGER(c_int) :: code
        CHARACTER(LEN=8) :: a
        CHARACTER(LEN=:), ALLOCATABLE :: unknown
        r%nch = 0
        unknown = ''
        DO j = 1, SIZE(r%info%attributes)
            a = lower(ADJUSTL(r%info%attributes(j)))
            IF (LEN_TRIM(a) == 0) CYCLE
            code = channel_code(a)
            IF (code < 0) THEN
                unknown = unknown // ' ' // TRIM(a)
                CYCLE
            END IF
            r%nch = r%nch + 1
            r%names(r%nch) = a
            r%codes(r%nch) = code
        END DO
        IF (LEN(unknown) > 0) &
            CALL warn('Unknown attributes' // unknown // " in receiver '" // &
                      TRIM(r%info%network) // '.' // TRIM(r%info%station) // '.' // &
                      TRIM(r%info%location) // "'.")
    END SUBROUTINE

    !> Locate every receiver and register its channels on the device
    !! (src/config/receiver.f90:435-492).
    SUBROUTINE parse_receiver_file(ctx, receivers)
        TYPE(c_ptr), INTENT(IN) :: ctx
        TYPE(receiver_cls), ALLOCATABLE, INTENT(OUT) :: receivers(:)
        INTEGER :: i, n
        INTEGER(c_int) :: exyz(3), id
        REAL(c_double), ALLOCATABLE :: vm(:)
        TYPE(receiver_cls), ALLOCATABLE :: tmp(:)
        TYPE(receiver_cls) :: r

        ALLOCATE(vm((configuration%lpd + 1)**3), tmp(SIZE(configuration%receivers)))
        n = 0
        DO i = 1, SIZE(configuration%receivers)
            r%info = configuration%receivers(i)
            CALL receiver_channels(r)
            IF (r%nch == 0) CYCLE
            IF (.NOT. located(ctx, r%info%lat, r%info%lon, r%info%depth, exyz, vm)) CYCLE
            CALL check(s3d_add_receiver(ctx, exyz, vm, INT(r%nch, c_int), r%codes, id), &
                       'add_receiver')
            r%id = id
            n = n + 1
            tmp(n) = r
        END DO
        receivers = tmp(1:n)
    END SUBROUTINE

    !> Channels of every receiver of the configuration, without a device: the
    !! parse-only check of ses3d_nt (-c). Returns the receiver and channel counts.
    SUBROUTINE count_receiver_channels(nrec, nch)
        INTEGER, INTENT(OUT) :: nrec, nch
        TYPE(receiver_cls) :: r
        INTEGER :: i
        nrec = 0
        nch = 0
        DO i = 1, SIZE(configuration%receivers)
            r%info = configuration%receivers(i)
            CALL receiver_channels(r)
            IF (r%nch == 0) CYCLE
            nrec = nrec + 1
            nch = nch + r%nch
        END DO
    END SUBROUTINE

    SUBROUTINE source_add(self, it)
        CLASS(source_cls), INTENT(IN) :: self
        INTEGER, INTENT(IN) :: it
        ! injected inside the device step (elem_kernel); nothing to do here
        IF (it < 1 .OR. self%ex < -1) RETURN
    END SUBROUTINE

    SUBROUTINE receiver_record(self, it)
        CLASS(receiver_cls), INTENT(IN) :: self
        INTEGER, INTENT(IN) :: it
        ! sampled inside the device step (record_kernel); nothing to do here
        IF (it < 1 .OR. self%id < -1) RETURN
    END SUBROUTINE

    !> Write one SAC file per channel (src/config/receiver.f90:154-192).
    SUBROUTINE receiver_write(self, ctx)
        CLASS(receiver_cls), INTENT(IN) :: self
        TYPE(c_ptr), INTENT(IN) :: ctx
        REAL(c_double), ALLOCATABLE :: tr(:)
        INTEGER :: q, ios
        CHARACTER(LEN=:), ALLOCATABLE :: fn
        ALLOCATE(tr(configuration%nt))
        DO q = 1, self%nch
            CALL check(s3d_get_trace(ctx, INT(self%id, c_int), INT(q - 1, c_int), tr), 'get_trace')
            fn = TRIM(self%info%prefix) // sac_file_name(self%info%network, self%info%location, &
                                                         self%info%station, self%names(q))
            CALL write_sac(fn, tr, configuration%dt, self%info%lat, self%info%lon, &
                           self%info%depth, self%info%network, self%info%station, &
                           self%info%location, self%names(q), configuration%event_name, &
                           configuration%date_time, self%info%override, ios)
            IF (ios /= 0) WRITE(*, '("WARNING: could not write ", A)') fn
        END DO
    END SUBROUTINE

END MODULE channels_mod
