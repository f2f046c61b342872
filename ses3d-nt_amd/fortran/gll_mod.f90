!> GLL basis with the reference's interface (src/gll_mod.f90:167-275):
!! get_knots(lpd), get_weights(lpd), get_dlgll(n, i, j).
!! Values come from the C library (s3d_gll), computed in FP64.
MODULE gll_mod
    USE, INTRINSIC :: ISO_C_BINDING, ONLY : c_double, c_int
    USE ses3d_hip_c, ONLY : s3d_gll
    IMPLICIT NONE

    PRIVATE
    INTEGER, PARAMETER :: rk = c_double

    PUBLIC :: get_knots, get_weights, get_dlgll

CONTAINS

    !> GLL collocation points of degree lpd (lpd+1 values).
    FUNCTION get_knots(lpd) RESULT(knots)
        INTEGER, INTENT(IN) :: lpd
        REAL(rk) :: knots(lpd+1)
        REAL(rk) :: w(8), d(64)
        IF (s3d_gll(INT(lpd, c_int), knots, w, d) /= 0) knots = 0.0_rk
    END FUNCTION

    !> GLL quadrature weights of degree lpd (lpd+1 values).
    FUNCTION get_weights(lpd) RESULT(weights)
        INTEGER, INTENT(IN) :: lpd
        REAL(rk) :: weights(lpd+1)
        REAL(rk) :: k(8), d(64)
        IF (s3d_gll(INT(lpd, c_int), k, weights, d) /= 0) weights = 0.0_rk
    END FUNCTION

    !> Derivative of the i-th Lagrange polynomial of degree n at GLL point j.
    FUNCTION get_dlgll(n, i, j) RESULT(y)
        INTEGER, INTENT(IN) :: n, i, j
        REAL(rk) :: y
        REAL(rk) :: k(8), w(8), d(64)
        y = 0.0_rk
        IF (s3d_gll(INT(n, c_int), k, w, d) /= 0) RETURN
        y = d(1 + i + (n + 1) * j)
    END FUNCTION

END MODULE gll_mod
