!> Sensitivity densities of the adjoint workflow (reference grad_mod,
!! src/SOURCE/ses3d_grad.F90): at adjoint step it, if the forward step
!! s = nt + 1 - it is a multiple of the &grad timestep_increment (the
!! reference's saving_vector, :159-174; a negative increment, such as the
!! default -1, selects the multiples of its magnitude and makes the factor
!! increment*dt negative, as there), the forward velocity snapshots
!! ./vx_<s>.vol, ./vy_<s>.vol, ./vz_<s>.vol (written by a forward run's
!! &output_raw, global 6-D layout, float32 or float64) are read and
!! grad_rho += increment*dt*(v_fw . v)/J is accumulated on the device
!! (:175-213). The densities are written through &output_raw attributes
!! grad_rho ... grad_csv.
MODULE grad_mod
    USE, INTRINSIC :: ISO_C_BINDING
    USE ses3d_hip_c
    USE configuration_mod, ONLY : configuration, abort
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: ses3d_grad

CONTAINS

    !> A snapshot file in the global 6-D layout, float64 or float32 by size
    !! (reference mpi_read_parallel, src/mpi_parallel_io.F90:61-341).
    SUBROUTINE read_snapshot(file, n, out)
        CHARACTER(LEN=*), INTENT(IN) :: file
        INTEGER(8), INTENT(IN) :: n
        REAL(c_double), INTENT(OUT) :: out(n)
        INTEGER :: lun, ios
        INTEGER(8) :: sz
        REAL(4), ALLOCATABLE :: f4(:)
        OPEN(NEWUNIT=lun, FILE=file, ACCESS='stream', FORM='unformatted', &
             STATUS='old', ACTION='read', IOSTAT=ios)
        IF (ios /= 0) CALL abort('File "' // TRIM(file) // '" does not exist!')
        INQUIRE(UNIT=lun, SIZE=sz)
        IF (sz == 8 * n) THEN
            READ(lun) out
        ELSE IF (sz == 4 * n) THEN
            ALLOCATE(f4(n))
            READ(lun) f4
            out = REAL(f4, c_double)
        ELSE
            CALL abort('File "' // TRIM(file) // '" has the wrong size')
        END IF
        CLOSE(lun)
    END SUBROUTINE

    SUBROUTINE ses3d_grad(ctx, it)
        TYPE(c_ptr), INTENT(IN) :: ctx
        INTEGER, INTENT(IN) :: it
        REAL(c_double), ALLOCATABLE :: fx(:), fy(:), fz(:)
        CHARACTER(LEN=1024) :: fn
        INTEGER(8) :: n
        INTEGER :: s, inc

        ASSOCIATE (cf => configuration)
        inc = cf%grad%ts_increment
        s = cf%nt + 1 - it
        IF (MOD(s, inc) /= 0) RETURN
        n = INT(cf%nx, 8) * cf%ny * cf%nz * (cf%lpd + 1)**3
        ALLOCATE(fx(n), fy(n), fz(n))
        WRITE(fn, '(A, A, "_", I0, ".vol")') TRIM(cf%grad%prefix), 'vx', s
        CALL read_snapshot(TRIM(fn), n, fx)
        WRITE(fn, '(A, A, "_", I0, ".vol")') TRIM(cf%grad%prefix), 'vy', s
        CALL read_snapshot(TRIM(fn), n, fy)
        WRITE(fn, '(A, A, "_", I0, ".vol")') TRIM(cf%grad%prefix), 'vz', s
        CALL read_snapshot(TRIM(fn), n, fz)
        CALL check(s3d_grad_add(ctx, fx, fy, fz, INT(inc, c_int)), 'grad_add')
        END ASSOCIATE
    END SUBROUTINE

END MODULE grad_mod
