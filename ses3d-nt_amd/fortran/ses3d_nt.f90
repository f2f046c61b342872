!> Ses3d-NT on one MI355X (reference main program, src/ses3d-nt.F90:30-359),
!! workflows 'forward', 'adjoint' and 'model'.
!!
!!   ses3d_nt <config-file>
!!   ses3d_nt -c <config-file>   parse and check the configuration only
!!                               (no device, no model or source files read)
!!
!! forward: reads the reference's NAMELIST config, uploads the model, runs nt
!! steps on the GPU and writes one SAC file per receiver channel.
!! adjoint: the forward loop with the sensitivity accumulation after each
!! step (src/ses3d-nt.F90:281-321; sources are the adjoint sources, the
!! forward velocity snapshots come from a forward run's &output_raw).
!! model: writes the built-in model of model_type to the parameter files
!! named in the &model group (no GPU needed).
PROGRAM ses3d_nt
    USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : OUTPUT_UNIT
    USE configuration_mod, ONLY : configuration, parse_conf_file, abort, rk
    USE evolution_mod, ONLY : ctx, init_temp, ses3d_evolution, v_ptp_loc, finalize_evolution, &
                              write_model_parameters
    USE channels_mod, ONLY : source_cls, receiver_cls, parse_source_file, parse_receiver_file, &
                             count_receiver_channels
    USE snapshots_mod, ONLY : write_snapshots, check_snapshot_attributes, write_netcdf, &
                              needs_displacement_gradient, curl_snapshot_due
    USE grad_mod, ONLY : ses3d_grad
    USE ses3d_hip_c, ONLY : check, s3d_enable_displacement_gradient, s3d_capture_curl_strain
    IMPLICIT NONE
    CHARACTER(LEN=512) :: conf_file_name
    TYPE(source_cls), ALLOCATABLE :: sources(:)
    TYPE(receiver_cls), ALLOCATABLE :: receivers(:)
    INTEGER :: it, i, every, nrec, nch
    REAL(rk) :: ptp(2)
    CHARACTER(LEN=8) :: flag
    LOGICAL :: check_only
    CHARACTER(LEN=*), PARAMETER :: fmt = '("Iteration ",I0,", ",I0,"%: ",G0," < v < ",G0)'

    IF (COMMAND_ARGUMENT_COUNT() < 1) CALL abort('Config-file missing!')
    CALL GET_COMMAND_ARGUMENT(1, flag)
    check_only = flag == '-c'
    IF (check_only) THEN
        IF (COMMAND_ARGUMENT_COUNT() < 2) CALL abort('Config-file missing!')
        CALL GET_COMMAND_ARGUMENT(2, conf_file_name)
    ELSE
        CALL GET_COMMAND_ARGUMENT(1, conf_file_name)
    END IF
    CALL parse_conf_file(TRIM(conf_file_name))
    IF (check_only) THEN
        CALL check_snapshot_attributes()
        CALL count_receiver_channels(nrec, nch)
        WRITE(OUTPUT_UNIT, '("Configuration OK: workflow ''", A, "'', ", I0, "x", I0, "x", I0, &
            &" elements, lpd ", I0, ", nt ", I0, ", ", I0, " sources, ", I0, " receivers with ", &
            &I0, " channels, ", I0, " raw and ", I0, " netCDF outputs")') &
            TRIM(configuration%workflow), configuration%nx, configuration%ny, configuration%nz, &
            configuration%lpd, configuration%nt, SIZE(configuration%sources), nrec, nch, &
            SIZE(configuration%outputs), SIZE(configuration%netcdf)
    ELSE IF (TRIM(configuration%workflow) == 'model') THEN
        CALL write_model_parameters()
        WRITE(OUTPUT_UNIT, '(A, /, "Workflow ''model'' successfully completed! ", /, A)') &
            REPEAT('=', 50), REPEAT('=', 50)
    ELSE
        CALL check_snapshot_attributes()
        CALL init_temp()
        IF (needs_displacement_gradient()) &
            CALL check(s3d_enable_displacement_gradient(ctx), 'enable_displacement_gradient')
        CALL parse_receiver_file(ctx, receivers)
        WRITE(OUTPUT_UNIT, '("Found ", I0, " receivers.")') SIZE(receivers)
        CALL parse_source_file(ctx, sources)
        WRITE(OUTPUT_UNIT, '("Found ", I0, " sources.")') SIZE(sources)
        WRITE(OUTPUT_UNIT, '(A, /, "Processing workflow ''", A, "'' of event ''", A, "''", /, A)') &
            REPEAT('=', 50), TRIM(configuration%workflow), TRIM(configuration%event_name), REPEAT('=', 50)

        every = MAX(1, configuration%nt / 10)
        DO it = 1, configuration%nt
            IF (curl_snapshot_due(it)) CALL check(s3d_capture_curl_strain(ctx), 'capture_curl_strain')
            CALL ses3d_evolution()
            DO i = 1, SIZE(sources)
                CALL sources(i)%add(it)
            END DO
            DO i = 1, SIZE(receivers)
                CALL receivers(i)%record(it)
            END DO
            CALL write_snapshots(ctx, it)
            CALL write_netcdf(it)
            IF (MOD(it, every) == 0 .OR. it == configuration%nt) THEN
                ptp = v_ptp_loc()
                WRITE(OUTPUT_UNIT, fmt) it, 100 * it / configuration%nt, ptp
            END IF
            IF (TRIM(configuration%workflow) == 'adjoint') CALL ses3d_grad(ctx, it)
        END DO

        DO i = 1, SIZE(receivers)
            CALL receivers(i)%write(ctx)
        END DO
        CALL finalize_evolution()
        WRITE(OUTPUT_UNIT, '(A, /, "Workflow ''", A, "'' successfully completed! ", /, A)') &
            REPEAT('=', 50), TRIM(configuration%workflow), REPEAT('=', 50)
    END IF
END PROGRAM ses3d_nt
