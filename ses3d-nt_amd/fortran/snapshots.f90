!> Raw volume snapshots (reference output_raw_cls, src/output_raw.f90:117-168,
!! and volume_snapshot_cls%process, src/volume_snapshot.f90:123-135): at
!! every step it with start <= it <= end and MOD(it - start, increment) == 0,
!! each attribute is written to prefix//attribute//'_'//it//'.vol' as the
!! global 6-D array, native endian, in the step precision (float32 for
!! 'fp32', float64 for 'fp64', like the reference's real_kind).
!! Attributes: vx vy vz sxx sxy sxz syy syz szz pressure taper, and the
!! sensitivity densities grad_rho grad_cs grad_cp grad_csh grad_csv
!! (src/volume_snapshot.f90:281-293), and the model: rho lambda mu vs/cs
!! vp/cp a b c q (:254-279), rebuilt on the host from the model files or the
!! built-in model, the displacement gradient exx exy exz eyy eyz ezz
!! (:220-236; integrated on the device when a snapshot asks for it) and the
!! curl fields rot_vx rot_vy rot_vz (:238-245; the strain-rate parts of the
!! step's start velocity are captured on the device before a step whose
!! snapshot asks for them).
MODULE snapshots_mod
    USE, INTRINSIC :: ISO_C_BINDING
    USE ses3d_hip_c
    USE configuration_mod, ONLY : configuration, output_entry, abort, warn
    USE evolution_mod, ONLY : model_param
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: write_snapshots, check_snapshot_attributes, write_netcdf, needs_displacement_gradient, &
              curl_snapshot_due

CONTAINS

    LOGICAL FUNCTION due(o, it)
        TYPE(output_entry), INTENT(IN) :: o
        INTEGER, INTENT(IN) :: it
        due = it >= o%ts_start .AND. it <= o%ts_end .AND. MOD(it - o%ts_start, o%ts_increment) == 0
    END FUNCTION

    !> &output_netcdf at step it: this build has no netCDF library, so a due
    !! output only warns, as the reference built without -DNetCDF does
    !! (write_netcdf, src/output_netcdf.F90:174-220).
    SUBROUTINE write_netcdf(it)
        INTEGER, INTENT(IN) :: it
        INTEGER :: i
        DO i = 1, SIZE(configuration%netcdf)
            IF (due(configuration%netcdf(i), it)) &
                CALL warn("Can't write netCDF file! Ses3d-NT is compiled without netCDF support.")
        END DO
    END SUBROUTINE

    SUBROUTINE check_snapshot_attributes()
        INTEGER :: i, j
        DO i = 1, SIZE(configuration%outputs)
            DO j = 1, SIZE(configuration%outputs(i)%attributes)
                SELECT CASE (TRIM(configuration%outputs(i)%attributes(j)))
                CASE ('', 'vx', 'vy', 'vz', 'sxx', 'sxy', 'sxz', 'syy', 'syz', 'szz', 'pressure', 'taper', &
                      'grad_rho', 'grad_cs', 'grad_cp', 'grad_csh', 'grad_csv', 'rho', 'lambda', 'mu', &
                      'vs', 'cs', 'vp', 'cp', 'a', 'b', 'c', 'q', &
                      'exx', 'exy', 'exz', 'eyy', 'eyz', 'ezz', 'rot_vx', 'rot_vy', 'rot_vz')
                CASE DEFAULT
                    CALL abort('Snapshot attribute not supported on the GPU path: ' // &
                               TRIM(configuration%outputs(i)%attributes(j)))
                END SELECT
            END DO
        END DO
    END SUBROUTINE

    !> Whether a snapshot asks for the displacement gradient (exx .. ezz),
    !! which the device then integrates from the first step on.
    LOGICAL FUNCTION needs_displacement_gradient()
        INTEGER :: i, j
        needs_displacement_gradient = .FALSE.
        DO i = 1, SIZE(configuration%outputs)
            DO j = 1, SIZE(configuration%outputs(i)%attributes)
                SELECT CASE (TRIM(configuration%outputs(i)%attributes(j)))
                CASE ('exx', 'exy', 'exz', 'eyy', 'eyz', 'ezz')
                    needs_displacement_gradient = .TRUE.
                END SELECT
            END DO
        END DO
    END FUNCTION

    !> Whether a snapshot due at step it asks for a curl field (rot_v*):
    !! the strain-rate parts of the step's start velocity are then captured
    !! before the step.
    LOGICAL FUNCTION curl_snapshot_due(it)
        INTEGER, INTENT(IN) :: it
        INTEGER :: i, j
        curl_snapshot_due = .FALSE.
        DO i = 1, SIZE(configuration%outputs)
            IF (.NOT. due(configuration%outputs(i), it)) CYCLE
            DO j = 1, SIZE(configuration%outputs(i)%attributes)
                SELECT CASE (TRIM(configuration%outputs(i)%attributes(j)))
                CASE ('rot_vx', 'rot_vy', 'rot_vz')
                    curl_snapshot_due = .TRUE.
                END SELECT
            END DO
        END DO
    END FUNCTION

    !> rho = 1/rhoinv where rhoinv > 0, else 0 (src/model_paras.F90:398-402)
    SUBROUTINE rho_of(f)
        REAL(c_double), INTENT(INOUT) :: f(:)
        WHERE (f > 0.0_c_double)
            f = 1.0_c_double / f
        ELSEWHERE
            f = 0.0_c_double
        END WHERE
    END SUBROUTINE

    SUBROUTINE get(ctx, id, f)
        TYPE(c_ptr), INTENT(IN) :: ctx
        INTEGER(c_int), INTENT(IN) :: id
        REAL(c_double), INTENT(OUT) :: f(:)
        CALL check(s3d_get_field(ctx, id, f), 'get_field')
    END SUBROUTINE

    SUBROUTINE write_snapshots(ctx, it)
        TYPE(c_ptr), INTENT(IN) :: ctx
        INTEGER, INTENT(IN) :: it
        INTEGER :: i, j, lun, ios
        INTEGER(8) :: n
        REAL(c_double), ALLOCATABLE :: f(:), g(:)
        CHARACTER(LEN=1024) :: fn
        CHARACTER(LEN=8) :: att
        CHARACTER(LEN=7) :: status

        n = INT(configuration%nx, 8) * configuration%ny * configuration%nz * (configuration%lpd + 1)**3
        DO i = 1, SIZE(configuration%outputs)
            IF (.NOT. due(configuration%outputs(i), it)) CYCLE
            IF (.NOT. ALLOCATED(f)) ALLOCATE(f(n))
            DO j = 1, SIZE(configuration%outputs(i)%attributes)
                att = configuration%outputs(i)%attributes(j)
                IF (LEN_TRIM(att) == 0) CYCLE
                SELECT CASE (TRIM(att))
                CASE ('vx');  CALL get(ctx, S3D_VX, f)
                CASE ('vy');  CALL get(ctx, S3D_VY, f)
                CASE ('vz');  CALL get(ctx, S3D_VZ, f)
                CASE ('sxx'); CALL get(ctx, S3D_SXX, f)
                CASE ('sxy'); CALL get(ctx, S3D_SXY, f)
                CASE ('sxz'); CALL get(ctx, S3D_SXZ, f)
                CASE ('syy'); CALL get(ctx, S3D_SYY, f)
                CASE ('syz'); CALL get(ctx, S3D_SYZ, f)
                CASE ('szz'); CALL get(ctx, S3D_SZZ, f)
                CASE ('taper'); CALL get(ctx, S3D_TAPER, f)
                ! displacement gradient, symmetric part (src/volume_snapshot.f90:220-236)
                CASE ('exx'); CALL get(ctx, S3D_EXX, f)
                CASE ('exy'); CALL get(ctx, S3D_EXY, f)
                CASE ('exz'); CALL get(ctx, S3D_EXZ, f)
                CASE ('eyy'); CALL get(ctx, S3D_EYY, f)
                CASE ('eyz'); CALL get(ctx, S3D_EYZ, f)
                CASE ('ezz'); CALL get(ctx, S3D_EZZ, f)
                ! curl of the velocity as the reference forms it (src/volume_snapshot.f90:238-245)
                CASE ('rot_vx'); CALL get(ctx, S3D_ROT_VX, f)
                CASE ('rot_vy'); CALL get(ctx, S3D_ROT_VY, f)
                CASE ('rot_vz'); CALL get(ctx, S3D_ROT_VZ, f)
                CASE ('rho')
                    CALL model_param(1, f)
                    CALL rho_of(f)
                CASE ('lambda'); CALL model_param(2, f)
                CASE ('mu'); CALL model_param(3, f)
                CASE ('a'); CALL model_param(4, f)
                CASE ('b'); CALL model_param(5, f)
                CASE ('c'); CALL model_param(6, f)
                CASE ('q'); CALL model_param(7, f)
                CASE ('vs', 'cs')
                    ! SQRT(mu / rho) (src/volume_snapshot.f90:260-261)
                    IF (.NOT. ALLOCATED(g)) ALLOCATE(g(n))
                    CALL model_param(1, g)
                    CALL rho_of(g)
                    CALL model_param(3, f)
                    f = SQRT(f / g)
                CASE ('vp', 'cp')
                    ! SQRT((lambda + 2 mu) / rho) (src/volume_snapshot.f90:263-264)
                    IF (.NOT. ALLOCATED(g)) ALLOCATE(g(n))
                    CALL model_param(2, f)
                    CALL model_param(3, g)
                    f = f + 2 * g
                    CALL model_param(1, g)
                    CALL rho_of(g)
                    f = SQRT(f / g)
                CASE ('grad_rho'); CALL get(ctx, S3D_GRAD_RHO, f)
                CASE ('grad_cs');  CALL get(ctx, S3D_GRAD_CS, f)
                CASE ('grad_cp');  CALL get(ctx, S3D_GRAD_CP, f)
                CASE ('grad_csh'); CALL get(ctx, S3D_GRAD_CSH, f)
                CASE ('grad_csv'); CALL get(ctx, S3D_GRAD_CSV, f)
                CASE ('pressure')
                    ! -(sxx + syy + szz) / 3 (src/config/channel_mod.f90:795)
                    IF (.NOT. ALLOCATED(g)) ALLOCATE(g(n))
                    CALL get(ctx, S3D_SXX, f)
                    CALL get(ctx, S3D_SYY, g)
                    f = f + g
                    CALL get(ctx, S3D_SZZ, g)
                    f = -(REAL(1.0 / 3.0, c_double) * (f + g))
                END SELECT
                WRITE(fn, '(A, A, "_", I0, ".vol")') TRIM(configuration%outputs(i)%prefix), TRIM(att), it
                status = MERGE('replace', 'new    ', configuration%outputs(i)%override)
                OPEN(NEWUNIT=lun, FILE=TRIM(fn), ACCESS='stream', FORM='unformatted', &
                     ACTION='write', STATUS=TRIM(status), IOSTAT=ios)
                IF (ios /= 0) CALL abort('Cannot write snapshot ' // TRIM(fn))
                IF (configuration%precision == 'fp32') THEN
                    WRITE(lun) REAL(f, 4)
                ELSE
                    WRITE(lun) f
                END IF
                CLOSE(lun)
            END DO
        END DO
    END SUBROUTINE

END MODULE snapshots_mod
