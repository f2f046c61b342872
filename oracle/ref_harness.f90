! ref_harness.f90 -- driver for the reference's own pure modules.
!
! TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) against
! the reference sources where they lie: src/gll_mod.f90,
! src/interpolation_mod.f90, src/coordinates_mod.f90,
! src/misc/source_signature_mod.f90, src/misc/coordinate_utilities_mod.f90,
! src/models/ak135_f_mod.F90 with the reference's own test doubles
! src/examples/dummy_parameters.f90 and src/examples/dummy_error.f90 (the
! recipe of src/examples/Makefile:40-75). Those doubles fix real_kind =
! REAL32, so every value printed from these modules is the reference's
! single-precision result. src/models/prem_mod.F90 is built in its own
! STAND_ALONE mode (real_kind = 8) and printed with 17 digits.
!
! Reads commands from stdin, one per line, and prints results with
! 9 significant digits (enough to represent any REAL32 exactly):
!   gll P                      -> knots, weights, dl(i,j) (i fastest)
!   coords ELMS P MIN MAX      -> coordinates_1d(ELMS, P, MIN, MAX), elm fastest
!   lagrange P X               -> calc_lagrange_polynomial(X, get_knots(P))
!   vandermonde P X Y Z        -> create_3d_vandermonde_matrix, i fastest
!   wavelet NAME T ONSET WIDTH -> ricker / gauss / sin_3 / brune / delta
!   colat LAT                  -> deg2rad(lat2colat(LAT))
!   prem R NOCRUST NOOCEAN     -> prem_iso(R, ...): rho, vp, vs (REAL64)
!   premti R                   -> prem(R, ...): rho, vpv, vph, vsv, vsh, eta
!   ak135 X                    -> ak135_f(X, ...): rho, vp, vs
!   end
PROGRAM ref_harness
    USE parameters_mod, ONLY : real_kind
    USE gll_mod, ONLY : get_knots, get_weights, get_dlgll
    USE coordinates_mod, ONLY : theta_coordinates
    USE interp_mod, ONLY : calc_lagrange_polynomial, create_3d_vandermonde_matrix
    USE source_signature_mod, ONLY : ricker, gauss, sin_3, brune, delta
    USE coordinate_utilities_mod, ONLY : deg2rad, lat2colat
    USE prem_mod, ONLY : prem, prem_iso
    USE ak135_f_mod, ONLY : ak135_f
    IMPLICIT NONE
    CHARACTER(LEN=256) :: line
    CHARACTER(LEN=16) :: cmd, name
    INTEGER :: ios, p, elms, i, j
    REAL(real_kind) :: a, b, x, y, z, t, onset, width
    REAL(8) :: r8, m8(6)
    LOGICAL :: nocrust, noocean
    REAL(real_kind), ALLOCATABLE :: c2(:,:), l(:), lx(:), ly(:), lz(:), vm(:,:,:)

    DO
        READ(*, '(A)', IOSTAT=ios) line
        IF (ios /= 0) EXIT
        READ(line, *, IOSTAT=ios) cmd
        IF (ios /= 0) CYCLE
        SELECT CASE (TRIM(cmd))
        CASE ('gll')
            READ(line, *) cmd, p
            WRITE(*, '(A)') 'knots'
            WRITE(*, '(ES17.9E3)') get_knots(p)
            WRITE(*, '(A)') 'weights'
            WRITE(*, '(ES17.9E3)') get_weights(p)
            WRITE(*, '(A)') 'dl'
            DO j = 0, p
                DO i = 0, p
                    WRITE(*, '(ES17.9E3)') get_dlgll(p, i, j)
                END DO
            END DO
        CASE ('coords')
            READ(line, *) cmd, elms, p, a, b
            ALLOCATE(c2(elms, p+1))
            c2 = theta_coordinates(elms=elms, lpd=p, min=a, max=b)
            WRITE(*, '(A)') 'coords'
            DO j = 1, p+1
                DO i = 1, elms
                    WRITE(*, '(ES17.9E3)') c2(i, j)
                END DO
            END DO
            DEALLOCATE(c2)
        CASE ('lagrange')
            READ(line, *) cmd, p, x
            ALLOCATE(l(p+1))
            l = calc_lagrange_polynomial(x, get_knots(p))
            WRITE(*, '(A)') 'lagrange'
            WRITE(*, '(ES17.9E3)') l
            DEALLOCATE(l)
        CASE ('vandermonde')
            READ(line, *) cmd, p, x, y, z
            ALLOCATE(lx(p+1), ly(p+1), lz(p+1), vm(p+1, p+1, p+1))
            lx = calc_lagrange_polynomial(x, get_knots(p))
            ly = calc_lagrange_polynomial(y, get_knots(p))
            lz = calc_lagrange_polynomial(z, get_knots(p))
            vm = create_3d_vandermonde_matrix(lx, ly, lz)
            WRITE(*, '(A)') 'vandermonde'
            WRITE(*, '(ES17.9E3)') vm
            DEALLOCATE(lx, ly, lz, vm)
        CASE ('wavelet')
            READ(line, *) cmd, name, t, onset, width
            WRITE(*, '(A)') 'wavelet'
            SELECT CASE (TRIM(name))
            CASE ('RICKER')
                WRITE(*, '(ES17.9E3)') ricker(t, onset, width)
            CASE ('GAUSS')
                WRITE(*, '(ES17.9E3)') gauss(t, onset, width)
            CASE ('SIN**3')
                WRITE(*, '(ES17.9E3)') sin_3(t, onset, width)
            CASE ('BRUNE')
                WRITE(*, '(ES17.9E3)') brune(t, onset, width)
            CASE ('DELTA')
                WRITE(*, '(ES17.9E3)') delta(t, onset)
            END SELECT
        CASE ('colat')
            READ(line, *) cmd, a
            WRITE(*, '(A)') 'colat'
            WRITE(*, '(ES17.9E3)') deg2rad(lat2colat(a))
        CASE ('prem')
            READ(line, *) cmd, r8, nocrust, noocean
            CALL prem_iso(r8, rho=m8(1), vp=m8(2), vs=m8(3), nocrust=nocrust, noocean=noocean)
            WRITE(*, '(A)') 'prem'
            WRITE(*, '(ES25.17E3)') m8(1:3)
        CASE ('premti')
            READ(line, *) cmd, r8
            CALL prem(r8, rho=m8(1), vpv=m8(2), vph=m8(3), vsv=m8(4), vsh=m8(5), eta=m8(6))
            WRITE(*, '(A)') 'premti'
            WRITE(*, '(ES25.17E3)') m8
        CASE ('ak135')
            READ(line, *) cmd, x
            CALL ak135_f(x, rho=a, vp=b, vs=y)
            WRITE(*, '(A)') 'ak135'
            WRITE(*, '(ES17.9E3)') a, b, y
        CASE ('end')
            EXIT
        END SELECT
    END DO
END PROGRAM ref_harness
