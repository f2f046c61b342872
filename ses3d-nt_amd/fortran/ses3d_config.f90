!> Configuration file reader: the reference's NAMELIST format
!! (src/parser/parser.f90:148-638) with its consistency checks
!! (src/config/config.F90:97-232), for one GPU holding the whole mesh.
!! Workflows 'forward' and 'model'; model_type -1 (parameter files) or one
!! of the built-in models 1, 2, 3, 11, 12, 13, 14, 15, 21.
!!
!! Groups: &grid, &model, &time, &general (exactly one each), &source_mt,
!! &source_sf, &source_sac, &receiver (any number, read in file order), and
!! the optional &gpu group of the GPU library: precision ('fp32' | 'fp64'), kahan_v,
!! kahan_v_bf16 (velocity compensation words in BF16, off by default),
!! kahan_sigma, kahan_sigma_bf16 (stress compensation words in BF16),
!! bf16_static (moduli, tau and dt/M stored in BF16), device, tile (element
!! tile of the FP32 step: -1 autotune, 0 default, 1, 2 or 4 elements per
!! workgroup), state_buffers (2: the step reads the state from one block and
!! writes the other, the default; 1: one block updated in place, half the
!! state memory), memory_vars (5: the deviatoric part of the memory
!! variables, which is all that enters the stress, five words per relaxation
!! mechanism, the default; 6: all six components, as the reference holds
!! them; no output of this program reads the memory variables).
MODULE configuration_mod
    USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : ERROR_UNIT, IOSTAT_END
    USE, INTRINSIC :: ISO_C_BINDING, ONLY : c_double, c_int
    IMPLICIT NONE

    PRIVATE
    INTEGER, PARAMETER, PUBLIC :: rk = c_double
    INTEGER, PARAMETER, PUBLIC :: fnl = 512
    INTEGER, PARAMETER, PUBLIC :: max_attributes = 30
    REAL(rk), PARAMETER :: default_real = -HUGE(1.0_rk)

    TYPE, PUBLIC :: source_entry
        LOGICAL :: moment = .TRUE.
        LOGICAL :: sac = .FALSE.                ! &source_sac: everything from the file
        CHARACTER(LEN=fnl) :: file = ' '
        REAL(rk) :: lat = 0.0_rk, lon = 0.0_rk, depth = 0.0_rk
        REAL(rk) :: onset = 0.0_rk, width = 0.0_rk
        REAL(rk) :: mt(6) = 0.0_rk
        REAL(rk) :: dir(3) = 0.0_rk
        CHARACTER(LEN=fnl) :: wavelet = ' '
    END TYPE

    TYPE, PUBLIC :: receiver_entry
        REAL(rk) :: lat, lon, depth
        CHARACTER(LEN=8) :: network, station, location
        CHARACTER(LEN=8) :: attributes(max_attributes)
        LOGICAL :: override
        CHARACTER(LEN=fnl) :: prefix
    END TYPE

    TYPE, PUBLIC :: output_entry
        INTEGER :: ts_start, ts_increment, ts_end
        CHARACTER(LEN=8) :: attributes(max_attributes)
        LOGICAL :: override
        CHARACTER(LEN=fnl) :: prefix
    END TYPE

    TYPE, PUBLIC :: configuration_typ
        ! &grid
        INTEGER :: nx, ny, nz, lpd, taper_width
        REAL(rk) :: taper_slope
        ! &model
        REAL(rk) :: lat_min, lat_max, lon_min, lon_max, rad_min, rad_max
        INTEGER :: model_type
        CHARACTER(LEN=fnl) :: file(7)   ! rhoinv, lambda, mu, a, b, c, q
        LOGICAL :: override
        ! &time
        REAL(rk) :: dt
        INTEGER :: nt, date_time(8)
        ! &general
        CHARACTER(LEN=16) :: workflow, event_name
        CHARACTER(LEN=fnl) :: log_file_dir
        ! &gpu
        CHARACTER(LEN=8) :: precision
        LOGICAL :: kahan_v, kahan_v_bf16, kahan_sigma, kahan_sigma_bf16, bf16_static
        INTEGER :: device, tile, state_buffers, memory_vars
        TYPE(source_entry), ALLOCATABLE :: sources(:)
        TYPE(receiver_entry), ALLOCATABLE :: receivers(:)
        TYPE(output_entry), ALLOCATABLE :: outputs(:)
        ! &output_netcdf groups: parsed and checked; the build has no netCDF
        ! library, so their steps only warn (as the reference built without
        ! -DNetCDF does, src/output_netcdf.F90:210-217)
        TYPE(output_entry), ALLOCATABLE :: netcdf(:)
        ! &grad (workflow 'adjoint'); prefix './'
        TYPE(output_entry) :: grad
    END TYPE

    TYPE(configuration_typ), PUBLIC :: configuration

    PUBLIC :: parse_conf_file, abort, warn, lower

CONTAINS

    SUBROUTINE abort(msg)
        CHARACTER(LEN=*), INTENT(IN) :: msg
        WRITE(ERROR_UNIT, '("ERROR: ", A)') msg
        ERROR STOP 1
    END SUBROUTINE

    !> Warning on the screen (reference WARN, src/misc/error_mod.F90:79-102).
    SUBROUTINE warn(msg)
        CHARACTER(LEN=*), INTENT(IN) :: msg
        WRITE(*, '("WARNING: ", A)') msg
    END SUBROUTINE

    INTEGER FUNCTION open_conf(file) RESULT(lun)
        CHARACTER(LEN=*), INTENT(IN) :: file
        INTEGER :: ios
        OPEN(NEWUNIT=lun, FILE=file, STATUS='old', ACTION='read', IOSTAT=ios)
        IF (ios /= 0) CALL abort('Config-File not found: ' // TRIM(file))
    END FUNCTION

    SUBROUTINE parse_conf_file(file)
        CHARACTER(LEN=*), INTENT(IN) :: file
        ! &grid
        INTEGER :: nx, ny, nz, lpd, taper_width
        REAL(rk) :: taper_slope
        NAMELIST /grid/ nx, ny, nz, lpd, taper_width, taper_slope
        ! &model
        REAL(rk) :: lat_max, lat_min, lon_max, lon_min, rad_max, rad_min
        INTEGER :: model_type
        CHARACTER(LEN=fnl) :: rhoinv, lambda, mu, a, b, c, q
        LOGICAL :: override
        NAMELIST /model/ lat_max, lat_min, lon_max, lon_min, rad_max, rad_min, &
            model_type, rhoinv, lambda, mu, a, b, c, q, override
        ! &time
        INTEGER :: nt, date_time(8)
        REAL(rk) :: dt
        NAMELIST /time/ dt, nt, date_time
        ! &general
        CHARACTER(LEN=16) :: event_name, workflow
        CHARACTER(LEN=fnl) :: log_file_dir
        NAMELIST /general/ log_file_dir, event_name, workflow
        ! &gpu
        CHARACTER(LEN=8) :: precision
        LOGICAL :: kahan_v, kahan_v_bf16, kahan_sigma, kahan_sigma_bf16, bf16_static
        INTEGER :: device, tile, state_buffers, memory_vars
        NAMELIST /gpu/ precision, kahan_v, kahan_v_bf16, kahan_sigma, kahan_sigma_bf16, bf16_static, device, tile, &
                       state_buffers, memory_vars
        INTEGER :: lun, ios, n
        TYPE(configuration_typ) :: cf

        lun = open_conf(file)
        ! ---- &grid (src/config/config.F90:111-128) ----
        nx = -1; ny = -1; nz = -1; lpd = -1; taper_width = -1; taper_slope = -1.0_rk
        n = count_groups(lun, 'grid')
        IF (n /= 1) CALL abort('Found ' // i2c(n) // ' &grid-groups in config-file!')
        READ(lun, NML=grid, IOSTAT=ios)
        IF (ios /= 0) CALL abort('Error while parsing &grid-group!')
        IF (nx < 4) CALL abort('Number of x-elements must be greater than 3!')
        IF (ny < 4) CALL abort('Number of y-elements must be greater than 3!')
        IF (nz < 4) CALL abort('Number of z-elements must be greater than 3!')
        IF (lpd < 2 .OR. lpd > 7) CALL abort('Lagrange polynomial degree must be between 2 and 7!')
        IF (ANY(2 * taper_width + 1 > [nx, ny, nz])) CALL abort('Relaxing boundaries are too wide!')
        cf%nx = nx; cf%ny = ny; cf%nz = nz; cf%lpd = lpd
        cf%taper_width = taper_width; cf%taper_slope = taper_slope

        ! ---- &general ----
        event_name = 'N/A'; workflow = ''; log_file_dir = 'N/A'
        n = count_groups(lun, 'general')
        IF (n /= 1) CALL abort('Found ' // i2c(n) // ' &general-groups in config-file!')
        READ(lun, NML=general, IOSTAT=ios)
        IF (ios /= 0) CALL abort('Error while parsing &general-group!')
        IF (TRIM(workflow) /= 'forward' .AND. TRIM(workflow) /= 'model' .AND. &
            TRIM(workflow) /= 'adjoint') CALL abort('Unknown workflow ' // TRIM(workflow))
        cf%workflow = workflow; cf%event_name = event_name; cf%log_file_dir = log_file_dir

        ! ---- &model (src/config/config.F90:160-189) ----
        lat_max = 0.0_rk; lat_min = 0.0_rk; lon_max = 0.0_rk; lon_min = 0.0_rk
        rad_max = 0.0_rk; rad_min = 0.0_rk; model_type = -1
        rhoinv = 'N/A'; lambda = 'N/A'; mu = 'N/A'; a = 'N/A'; b = 'N/A'; c = 'N/A'; q = 'N/A'
        override = .FALSE.
        n = count_groups(lun, 'model')
        IF (n /= 1) CALL abort('Found ' // i2c(n) // ' &model-groups in config-file!')
        READ(lun, NML=model, IOSTAT=ios)
        IF (ios /= 0) CALL abort('Error while parsing &model-group!')
        IF (lat_min > lat_max) CALL abort('lat_max < lat_min!')
        IF (lon_min > lon_max) CALL abort('lon_max < lon_min!')
        IF (rad_min > rad_max) CALL abort('rad_max < rad_min!')
        IF (ANY([lat_min, lat_max] < -90.0_rk) .OR. ANY([lat_min, lat_max] > 90.0_rk)) &
            CALL abort('Invalid latitude')
        IF (ANY([lon_min, lon_max] < -180.0_rk) .OR. ANY([lon_min, lon_max] > 360.0_rk)) &
            CALL abort('Invalid longitude')
        IF (ANY([rad_min, rad_max] < 1.0_rk)) CALL abort('Invalid Radius')
        IF (TRIM(workflow) == 'model') THEN
            ! src/config/model.f90:447-458
            IF (model_type == -1) &
                CALL abort("In workflow 'model' a model_type different from -1 must be specified!")
            IF (ALL([rhoinv, lambda, mu, a, b, c, q] == 'N/A')) &
                CALL abort("In workflow 'model' at least one model-parameter-filename must be specified!")
            IF (ALL(model_type /= [1, 2, 3, 11, 12, 13, 14, 15, 21])) &
                CALL abort('Unknown model_type ' // i2c(model_type))
        ELSE IF (model_type == -1) THEN
            IF (rhoinv == 'N/A') CALL abort('"rhoinv" model-file must be given!')
            IF (lambda == 'N/A') CALL abort('"lambda" model-file must be given!')
            IF (mu == 'N/A') CALL abort('"mu" model-file must be given!')
        ELSE
            ! src/config/model.f90:469-483
            IF (rhoinv /= 'N/A') CALL abort("'rhoinv' and a 'model_type' can't be specified at the same time")
            IF (lambda /= 'N/A') CALL abort("'lambda' and a 'model_type' can't be specified at the same time")
            IF (mu /= 'N/A') CALL abort("'mu' and a 'model_type' can't be specified at the same time")
            ! built-in models of src/model_paras.F90:219-336
            IF (ALL(model_type /= [1, 2, 3, 11, 12, 13, 14, 15, 21])) &
                CALL abort('Unknown model_type ' // i2c(model_type))
        END IF
        cf%lat_min = lat_min; cf%lat_max = lat_max; cf%lon_min = lon_min; cf%lon_max = lon_max
        cf%rad_min = rad_min; cf%rad_max = rad_max; cf%model_type = model_type
        cf%file = [CHARACTER(LEN=fnl) :: rhoinv, lambda, mu, a, b, c, q]
        cf%override = override

        ! ---- &time (src/config/config.F90:192-210) ----
        nt = -1; date_time = -1; dt = -1.0_rk
        n = count_groups(lun, 'time')
        IF (n /= 1) CALL abort('Found ' // i2c(n) // ' &time-groups in config-file!')
        READ(lun, NML=time, IOSTAT=ios)
        IF (ios /= 0) CALL abort('Error while parsing &time-group!')
        IF (2 > nt .OR. nt > 50000) CALL abort('Non-conforming nt value in &time-group!')
        IF (dt <= 0.0_rk) CALL abort('Non-conforming dt value in &time-group!')
        IF (ALL(date_time == -1)) CALL DATE_AND_TIME(VALUES=date_time)
        cf%nt = nt; cf%dt = dt; cf%date_time = date_time

        ! ---- &gpu (optional) ----
        precision = 'fp32'; kahan_v = .TRUE.; kahan_sigma = .TRUE.; device = 0
        kahan_sigma_bf16 = .TRUE.; bf16_static = .FALSE.; tile = -1; state_buffers = 2
        kahan_v_bf16 = .FALSE.
        memory_vars = 5
        n = count_groups(lun, 'gpu')
        IF (n > 1) CALL abort('Found ' // i2c(n) // ' &gpu-groups in config-file!')
        IF (n == 1) THEN
            READ(lun, NML=gpu, IOSTAT=ios)
            IF (ios /= 0) CALL abort('Error while parsing &gpu-group!')
        END IF
        IF (precision /= 'fp32' .AND. precision /= 'fp64') &
            CALL abort("precision must be 'fp32' or 'fp64'")
        cf%precision = precision; cf%kahan_v = kahan_v; cf%kahan_sigma = kahan_sigma
        cf%kahan_sigma_bf16 = kahan_sigma_bf16; cf%bf16_static = bf16_static
        cf%kahan_v_bf16 = kahan_v_bf16
        cf%device = device
        IF (ALL(tile /= [-1, 0, 1, 2, 4])) CALL abort('tile must be -1 (autotune), 0, 1, 2 or 4')
        cf%tile = tile
        IF (state_buffers /= 1 .AND. state_buffers /= 2) CALL abort('state_buffers must be 1 or 2')
        cf%state_buffers = state_buffers
        IF (memory_vars /= 5 .AND. memory_vars /= 6) CALL abort('memory_vars must be 5 or 6')
        cf%memory_vars = memory_vars

        CALL parse_sources(lun, cf)
        CALL parse_receivers(lun, cf)
        CALL parse_outputs(lun, cf, .FALSE.)
        CALL parse_outputs(lun, cf, .TRUE.)
        IF (TRIM(cf%workflow) == 'adjoint') CALL parse_grad(lun, cf)
        CLOSE(lun)
        configuration = cf
    END SUBROUTINE

    !> &source_mt, &source_sf and &source_sac groups
    !! (src/parser/parser.f90:265-411); the &source_sac entries follow the
    !! others, as in parse_source_file (src/config/source.f90:449-530).
    SUBROUTINE parse_sources(lun, cf)
        INTEGER, INTENT(IN) :: lun
        TYPE(configuration_typ), INTENT(INOUT) :: cf
        REAL(rk) :: lat, lon, depth, onset, width, moment_tensor(6), direction(3)
        CHARACTER(LEN=fnl) :: wavelet, file_name
        NAMELIST /source_mt/ lat, lon, depth, moment_tensor, wavelet, onset, width
        NAMELIST /source_sf/ lat, lon, depth, direction, wavelet, onset, width
        NAMELIST /source_sac/ file_name
        INTEGER :: nmt, nsf, nsac, i, ios
        TYPE(source_entry) :: s

        nmt = count_groups(lun, 'source_mt')
        nsf = count_groups(lun, 'source_sf')
        nsac = count_groups(lun, 'source_sac')
        ALLOCATE(cf%sources(nmt + nsf + nsac))
        REWIND(lun)
        DO i = 1, nmt
            lat = default_real; lon = default_real; depth = default_real
            onset = default_real; width = default_real
            moment_tensor = default_real; wavelet = ''
            READ(lun, NML=source_mt, IOSTAT=ios)
            IF (ios /= 0) CALL abort('Error while parsing &source_mt-group!')
            IF (ANY(moment_tensor == default_real)) CALL abort('Moment-tensor of dipol force invalid!')
            s%moment = .TRUE.
            s%lat = lat; s%lon = lon; s%depth = depth; s%onset = onset; s%width = width
            s%mt = moment_tensor; s%wavelet = wavelet
            cf%sources(i) = s
        END DO
        REWIND(lun)
        DO i = 1, nsf
            lat = default_real; lon = default_real; depth = default_real
            onset = default_real; width = default_real
            direction = default_real; wavelet = ''
            READ(lun, NML=source_sf, IOSTAT=ios)
            IF (ios /= 0) CALL abort('Error while parsing &source_sf-group!')
            IF (ANY(direction == default_real)) CALL abort('Direction of single force invalid!')
            s%moment = .FALSE.
            s%lat = lat; s%lon = lon; s%depth = depth; s%onset = onset; s%width = width
            s%dir = direction; s%wavelet = wavelet
            cf%sources(nmt + i) = s
        END DO
        REWIND(lun)
        DO i = 1, nsac
            file_name = ''
            READ(lun, NML=source_sac, IOSTAT=ios)
            IF (ios /= 0) CALL abort('Error while parsing &source_sac-group!')
            IF (LEN_TRIM(file_name) == 0) CALL abort('&source_sac: file_name missing!')
            s = source_entry()
            s%sac = .TRUE.
            s%file = file_name
            cf%sources(nmt + nsf + i) = s
        END DO
    END SUBROUTINE

    !> &receiver groups (src/parser/parser.f90:368-420)
    SUBROUTINE parse_receivers(lun, cf)
        INTEGER, INTENT(IN) :: lun
        TYPE(configuration_typ), INTENT(INOUT) :: cf
        REAL(rk) :: lat, lon, depth
        CHARACTER(LEN=8) :: network, station, location, attributes(max_attributes)
        LOGICAL :: override
        CHARACTER(LEN=fnl) :: prefix
        NAMELIST /receiver/ lat, lon, depth, network, station, location, &
                            attributes, override, prefix
        INTEGER :: n, i, ios

        n = count_groups(lun, 'receiver')
        ALLOCATE(cf%receivers(n))
        REWIND(lun)
        DO i = 1, n
            lat = default_real; lon = default_real; depth = default_real
            network = 'Ses3d-NT'; station = ''; location = ''; attributes = ''
            override = .FALSE.; prefix = './'
            READ(lun, NML=receiver, IOSTAT=ios)
            IF (ios /= 0) CALL abort('Error while parsing &receiver-group!')
            cf%receivers(i) = receiver_entry(lat, lon, depth, network, station, &
                                             location, attributes, override, prefix)
        END DO
    END SUBROUTINE

    !> &output_raw groups (netcdf false; src/parser/parser.f90:557-600) or
    !! &output_netcdf groups (netcdf true; :502-555), with the defaults and
    !! checks of the snapshot constructor (src/volume_snapshot.f90:70-105):
    !! timestep_end defaults to nt, timestep_start to 1, the prefix to './';
    !! attributes are lower-cased (get_field, :209).
    SUBROUTINE parse_outputs(lun, cf, netcdf)
        INTEGER, INTENT(IN) :: lun
        TYPE(configuration_typ), INTENT(INOUT) :: cf
        LOGICAL, INTENT(IN) :: netcdf
        INTEGER, PARAMETER :: unset = HUGE(1)
        INTEGER :: timestep_start, timestep_increment, timestep_end
        LOGICAL :: override
        CHARACTER(LEN=fnl) :: prefix
        CHARACTER(LEN=8) :: attributes(max_attributes)
        NAMELIST /output_raw/ timestep_start, timestep_increment, timestep_end, &
            attributes, override, prefix
        NAMELIST /output_netcdf/ timestep_start, timestep_increment, timestep_end, &
            attributes, override, prefix
        TYPE(output_entry), ALLOCATABLE :: out(:)
        INTEGER :: n, i, j, ios

        n = count_groups(lun, MERGE('output_netcdf', 'output_raw   ', netcdf))
        ALLOCATE(out(n))
        REWIND(lun)
        DO i = 1, n
            timestep_start = unset; timestep_increment = unset; timestep_end = unset
            override = .FALSE.; prefix = ''; attributes = ''
            IF (netcdf) THEN
                READ(lun, NML=output_netcdf, IOSTAT=ios)
                IF (ios /= 0) CALL abort('Error while parsing &output_netcdf-group!')
            ELSE
                READ(lun, NML=output_raw, IOSTAT=ios)
                IF (ios /= 0) CALL abort('Error while parsing &output_raw-group!')
            END IF
            IF (timestep_end == unset) timestep_end = cf%nt
            IF (timestep_start == unset) timestep_start = 1
            IF (LEN_TRIM(prefix) == 0) prefix = './'
            IF (timestep_start < 1) CALL abort('timestep_start out of range')
            IF (timestep_increment < 1 .OR. timestep_increment > cf%nt) &
                CALL abort('timestep_increment out of range')
            IF (timestep_end < timestep_start) CALL abort('timestep_end out of range')
            DO j = 1, max_attributes
                attributes(j) = lower(ADJUSTL(attributes(j)))
            END DO
            out(i) = output_entry(timestep_start, timestep_increment, timestep_end, &
                                  attributes, override, prefix)
        END DO
        IF (netcdf) THEN
            cf%netcdf = out
        ELSE
            cf%outputs = out
        END IF
    END SUBROUTINE

    !> The &grad group of workflow 'adjoint' (src/parser/parser.f90:462-497,
    !! src/SOURCE/ses3d_grad.F90:237-292): exactly one, with at least one
    !! attribute. timestep_increment (default -1, as in the reference) is the
    !! snapshot spacing |increment| and the signed factor increment*dt of the
    !! time integral (ses3d_grad.F90:159-213); 0 is refused.
    SUBROUTINE parse_grad(lun, cf)
        INTEGER, INTENT(IN) :: lun
        TYPE(configuration_typ), INTENT(INOUT) :: cf
        INTEGER :: timestep_start, timestep_increment, timestep_end
        CHARACTER(LEN=8) :: attributes(max_attributes)
        NAMELIST /grad/ timestep_start, timestep_increment, timestep_end, attributes
        INTEGER :: n, ios

        n = count_groups(lun, 'grad')
        IF (n /= 1) CALL abort('Found ' // i2c(n) // ' &grad-groups in config-file!')
        timestep_start = -1; timestep_increment = -1; timestep_end = -1; attributes = ''
        READ(lun, NML=grad, IOSTAT=ios)
        IF (ios /= 0) CALL abort('Error while parsing &grad-group!')
        IF (COUNT(attributes /= ' ') < 1) CALL abort('No gradients specified')
        IF (timestep_increment == 0) CALL abort('Non-conforming time steps in &grad-group!')
        cf%grad = output_entry(timestep_start, timestep_increment, timestep_end, &
                               attributes, .FALSE., './')
        REWIND(lun)
    END SUBROUTINE

    !> Number of occurrences of &name in the file; leaves the file rewound.
    INTEGER FUNCTION count_groups(lun, name) RESULT(n)
        INTEGER, INTENT(IN) :: lun
        CHARACTER(LEN=*), INTENT(IN) :: name
        CHARACTER(LEN=1024) :: line
        CHARACTER(LEN=64) :: tok
        INTEGER :: ios, p
        n = 0
        REWIND(lun)
        DO
            READ(lun, '(A)', IOSTAT=ios) line
            IF (ios /= 0) EXIT
            line = ADJUSTL(line)
            IF (line(1:1) /= '&') CYCLE
            tok = line(2:)
            p = SCAN(tok, ' ,/!')
            IF (p > 0) tok = tok(1:p-1)
            IF (lower(TRIM(tok)) == name) n = n + 1
        END DO
        REWIND(lun)
    END FUNCTION

    FUNCTION lower(s) RESULT(t)
        CHARACTER(LEN=*), INTENT(IN) :: s
        CHARACTER(LEN=LEN(s)) :: t
        INTEGER :: i
        t = s
        DO i = 1, LEN(s)
            IF (IACHAR(s(i:i)) >= 65 .AND. IACHAR(s(i:i)) <= 90) t(i:i) = ACHAR(IACHAR(s(i:i)) + 32)
        END DO
    END FUNCTION

    FUNCTION i2c(i) RESULT(s)
        INTEGER, INTENT(IN) :: i
        CHARACTER(LEN=:), ALLOCATABLE :: s
        CHARACTER(LEN=16) :: b
        WRITE(b, '(I0)') i
        s = TRIM(b)
    END FUNCTION

END MODULE configuration_mod
