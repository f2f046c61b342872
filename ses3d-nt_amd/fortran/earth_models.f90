!> Built-in 1-D Earth models of the &model group's model_type
!! (src/model_paras.F90:206-342): PREM (Dziewonski & Anderson 1981), isotropic
!! with its crust / ocean variants (src/models/prem_mod.F90:90-250) and
!! transversely isotropic in the LVZ and LID (:260-330), and ak135-f
!! (Kennett, Engdahl & Buland 1995; Montagner & Kennett 1996), linearly
!! interpolated in radius (src/models/ak135_f_mod.F90:235-288).
!!
!! The coefficients and table entries are written as default-kind (single
!! precision) literals, as in the reference, so that an FP64 evaluation
!! carries the same promoted constants as the reference's FP64 build.
MODULE earth_models_mod
    USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : REAL64
    IMPLICIT NONE

    PRIVATE
    INTEGER, PARAMETER :: rk = REAL64

    ! PREM region boundaries [m]
    REAL(rk), PARAMETER :: R_EARTH = 6371000.0, R_OCEAN = 6368000.0, &
        R_MID_CRUST = 6356000.0, R_MOHO = 6346600.0, R_80 = 6291000.0, &
        R_220 = 6151000.0, R_400 = 5971000.0, R_600 = 5771000.0, &
        R_670 = 5701000.0, R_771 = 5600000.0, R_DDP = 3630000.0, &
        R_CMB = 3480000.0, R_ICB = 1221500.0

    ! ak135-f: radius [km], density [g/cm^3], vp, vs [km/s], surface to centre
    INTEGER, PARAMETER :: NAK = 145
    REAL(rk), PARAMETER :: AK(4, NAK) = RESHAPE([ &
         6371.00,   1.0200,   1.4500,  0.0000, &
         6368.00,   1.0200,   1.4500,  0.0000, &
         6368.00,   2.0000,   1.6500,  1.0000, &
         6367.70,   2.0000,   1.6500,  1.0000, &
         6367.70,   2.6000,   5.8000,  3.2000, &
         6361.00,   2.6000,   5.8000,  3.2000, &
         6361.00,   2.9200,   6.8000,  3.9000, &
         6353.00,   2.9200,   6.8000,  3.9000, &
         6353.00,   3.6410,   8.0355,  4.4839, &
         6328.00,   3.5801,   8.0379,  4.4856, &
         6291.00,   3.5020,   8.0400,  4.4800, &
         6291.00,   3.5020,   8.0450,  4.4900, &
         6251.00,   3.4268,   8.0505,  4.5000, &
         6251.00,   3.4268,   8.0505,  4.5000, &
         6206.00,   3.3711,   8.1750,  4.5090, &
         6161.00,   3.3243,   8.3007,  4.5184, &
         6161.00,   3.3243,   8.3007,  4.5184, &
         6111.00,   3.3663,   8.4822,  4.6094, &
         6061.00,   3.4110,   8.6650,  4.6964, &
         6011.00,   3.4577,   8.8476,  4.7832, &
         5961.00,   3.5068,   9.0302,  4.8702, &
         5961.00,   3.9317,   9.3601,  5.0806, &
         5911.00,   3.9273,   9.5280,  5.1864, &
         5861.00,   3.9233,   9.6962,  5.2922, &
         5811.00,   3.9218,   9.8640,  5.3989, &
         5761.00,   3.9206,  10.0320,  5.5047, &
         5711.00,   3.9201,  10.2000,  5.6104, &
         5711.00,   4.2387,  10.7909,  5.9607, &
         5661.00,   4.2986,  10.9222,  6.0898, &
         5611.00,   4.3565,  11.0553,  6.2100, &
         5561.50,   4.4118,  11.1355,  6.2424, &
         5512.00,   4.4650,  11.2228,  6.2799, &
         5462.50,   4.5162,  11.3068,  6.3164, &
         5413.00,   4.5654,  11.3897,  6.3519, &
         5363.50,   4.5926,  11.4704,  6.3860, &
         5314.00,   4.6198,  11.5493,  6.4182, &
         5264.50,   4.6467,  11.6265,  6.4514, &
         5215.00,   4.6735,  11.7020,  6.4822, &
         5165.50,   4.7001,  11.7768,  6.5131, &
         5116.00,   4.7266,  11.8491,  6.5431, &
         5066.50,   4.7528,  11.9208,  6.5728, &
         5017.00,   4.7790,  11.9891,  6.6009, &
         4967.50,   4.8050,  12.0571,  6.6285, &
         4918.00,   4.8307,  12.1247,  6.6554, &
         4868.50,   4.8562,  12.1912,  6.6813, &
         4819.00,   4.8817,  12.2558,  6.7070, &
         4769.50,   4.9069,  12.3181,  6.7323, &
         4720.00,   4.9321,  12.3813,  6.7579, &
         4670.50,   4.9570,  12.4427,  6.7820, &
         4621.00,   4.9817,  12.5030,  6.8056, &
         4571.50,   5.0062,  12.5638,  6.8289, &
         4522.00,   5.0306,  12.6226,  6.8517, &
         4472.50,   5.0548,  12.6807,  6.8743, &
         4423.00,   5.0789,  12.7384,  6.8972, &
         4373.50,   5.1027,  12.7956,  6.9194, &
         4324.00,   5.1264,  12.8524,  6.9416, &
         4274.50,   5.1499,  12.9093,  6.9625, &
         4225.00,   5.1732,  12.9663,  6.9852, &
         4175.50,   5.1963,  13.0226,  7.0069, &
         4126.00,   5.2192,  13.0786,  7.0286, &
         4076.50,   5.2420,  13.1337,  7.0504, &
         4027.00,   5.2646,  13.1895,  7.0722, &
         3977.50,   5.2870,  13.2465,  7.0932, &
         3928.00,   5.3092,  13.3017,  7.1144, &
         3878.50,   5.3313,  13.3584,  7.1368, &
         3829.00,   5.3531,  13.4156,  7.1584, &
         3779.50,   5.3748,  13.4741,  7.1804, &
         3731.00,   5.3962,  13.5311,  7.2031, &
         3681.00,   5.4176,  13.5899,  7.2253, &
         3631.00,   5.4387,  13.6498,  7.2485, &
         3631.00,   5.6934,  13.6498,  7.2485, &
         3581.33,   5.7196,  13.6533,  7.2593, &
         3531.67,   5.7458,  13.6570,  7.2700, &
         3479.50,   5.7721,  13.6601,  7.2817, &
         3479.50,   9.9145,   8.0000,  0.0000, &
         3431.67,   9.9942,   8.0382,  0.0000, &
         3381.34,  10.0722,   8.1283,  0.0000, &
         3331.01,  10.1485,   8.2213,  0.0000, &
         3280.68,  10.2233,   8.3122,  0.0000, &
         3230.34,  10.2964,   8.4001,  0.0000, &
         3180.01,  10.3679,   8.4861,  0.0000, &
         3129.68,  10.4378,   8.5692,  0.0000, &
         3079.35,  10.5062,   8.6496,  0.0000, &
         3029.02,  10.5731,   8.7283,  0.0000, &
         2978.69,  10.6385,   8.8036,  0.0000, &
         2928.36,  10.7023,   8.8761,  0.0000, &
         2878.03,  10.7647,   8.9461,  0.0000, &
         2827.70,  10.8257,   9.0138,  0.0000, &
         2777.36,  10.8852,   9.0792,  0.0000, &
         2727.03,  10.9434,   9.1426,  0.0000, &
         2676.70,  11.0001,   9.2042,  0.0000, &
         2626.37,  11.0555,   9.2634,  0.0000, &
         2576.04,  11.1095,   9.3205,  0.0000, &
         2525.71,  11.1623,   9.3760,  0.0000, &
         2475.38,  11.2137,   9.4297,  0.0000, &
         2425.05,  11.2639,   9.4814,  0.0000, &
         2374.72,  11.3127,   9.5306,  0.0000, &
         2324.38,  11.3604,   9.5777,  0.0000, &
         2274.05,  11.4069,   9.6232,  0.0000, &
         2223.72,  11.4521,   9.6673,  0.0000, &
         2173.39,  11.4962,   9.7100,  0.0000, &
         2123.06,  11.5391,   9.7513,  0.0000, &
         2072.73,  11.5809,   9.7914,  0.0000, &
         2022.40,  11.6216,   9.8304,  0.0000, &
         1972.07,  11.6612,   9.8682,  0.0000, &
         1921.74,  11.6998,   9.9051,  0.0000, &
         1871.40,  11.7373,   9.9410,  0.0000, &
         1821.07,  11.7737,   9.9761,  0.0000, &
         1770.74,  11.8092,  10.0103,  0.0000, &
         1720.41,  11.8437,  10.0439,  0.0000, &
         1670.08,  11.8772,  10.0768,  0.0000, &
         1619.75,  11.9098,  10.1095,  0.0000, &
         1569.42,  11.9414,  10.1415,  0.0000, &
         1519.09,  11.9722,  10.1739,  0.0000, &
         1468.76,  12.0001,  10.2049,  0.0000, &
         1418.42,  12.0311,  10.2329,  0.0000, &
         1368.09,  12.0593,  10.2565,  0.0000, &
         1317.76,  12.0867,  10.2745,  0.0000, &
         1267.43,  12.1133,  10.2854,  0.0000, &
         1217.50,  12.1391,  10.2890,  0.0000, &
         1217.50,  12.7037,  11.0427,  3.5043, &
         1166.39,  12.7289,  11.0585,  3.5187, &
         1115.68,  12.7530,  11.0718,  3.5314, &
         1064.96,  12.7760,  11.0850,  3.5435, &
         1014.25,  12.7980,  11.0983,  3.5551, &
          963.54,  12.8188,  11.1166,  3.5661, &
          912.83,  12.8387,  11.1316,  3.5765, &
          862.11,  12.8574,  11.1457,  3.5864, &
          811.40,  12.8751,  11.1590,  3.5957, &
          760.69,  12.8917,  11.1715,  3.6044, &
          709.98,  12.9072,  11.1832,  3.6126, &
          659.26,  12.9217,  11.1941,  3.6202, &
          608.55,  12.9351,  11.2041,  3.6272, &
          557.84,  12.9474,  11.2134,  3.6337, &
          507.13,  12.9586,  11.2219,  3.6396, &
          456.41,  12.9688,  11.2295,  3.6450, &
          405.70,  12.9779,  11.2364,  3.6498, &
          354.99,  12.9859,  11.2424,  3.6540, &
          304.28,  12.9929,  11.2477,  3.6577, &
          253.56,  12.9988,  11.2521,  3.6608, &
          202.85,  13.0036,  11.2557,  3.6633, &
          152.14,  13.0074,  11.2586,  3.6653, &
          101.43,  13.0100,  11.2606,  3.6667, &
           50.71,  13.0117,  11.2618,  3.6675, &
            0.00,  13.0122,  11.2622,  3.6678  &
        ], [4, NAK])

    PUBLIC :: prem_iso_at, prem_ti_at, ak135f_at, builtin_model

CONTAINS

    !> Isotropic PREM at radius r [m]: density [g/cm^3], vp, vs [km/s].
    !! nocrust: the crust and ocean take the LID values at the Moho;
    !! noocean: the ocean takes the middle-crust values.
    ELEMENTAL SUBROUTINE prem_iso_at(r, nocrust, noocean, rho, vp, vs)
        REAL(rk), INTENT(IN) :: r
        LOGICAL, INTENT(IN) :: nocrust, noocean
        REAL(rk), INTENT(OUT) :: rho, vp, vs
        REAL(rk) :: x
        x = ABS(r / R_EARTH)
        IF (r >= 0.0 .AND. r <= R_ICB) THEN                 ! inner core
            rho = 13.08850 - 8.83810*x**2
            vp  = 11.26220 - 6.36400*x**2
            vs  = 3.66780 - 4.44750*x**2
        ELSE IF (r > R_ICB .AND. r <= R_CMB) THEN           ! outer core
            rho = 12.58150 - 1.26380*x - 3.64260*x**2 - 5.52810*x**3
            vp  = 11.04870 - 4.03620*x + 4.80230*x**2 - 13.57320*x**3
            vs  = 0.0
        ELSE IF (r > R_CMB .AND. r <= R_DDP) THEN           ! D''
            rho = 7.95650-6.47610*x+5.52830*x*x-3.08070*x*x*x
            vp  = 15.38910-5.31810*x+5.52420*x*x-2.55140*x*x*x
            vs  = 6.92540+1.46720*x-2.08340*x*x+0.97830*x*x*x
        ELSE IF (r > R_DDP .AND. r <= R_771) THEN           ! lower mantle
            rho = 7.95650-6.47610*x+5.52830*x*x-3.08070*x*x*x
            vp  = 24.95200-40.46730*x+51.48320*x*x-26.64190*x*x*x
            vs  = 11.16710-13.78180*x+17.45750*x*x-9.27770*x*x*x
        ELSE IF (r > R_771 .AND. r <= R_670) THEN
            rho = 7.95650-6.47610*x+5.52830*x*x-3.08070*x*x*x
            vp  = 29.27660-23.60270*x+5.52420*x*x-2.55140*x*x*x
            vs  = 22.34590-17.24730*x-2.08340*x*x+0.97830*x*x*x
        ELSE IF (r > R_670 .AND. r <= R_600) THEN           ! transition zone
            rho = 5.31970-1.48360*x
            vp  = 19.09570-9.86720*x
            vs  = 9.98390-4.93240*x
        ELSE IF (r > R_600 .AND. r <= R_400) THEN
            rho = 11.24940-8.02980*x
            vp  = 39.70270-32.61660*x
            vs  = 22.35120-18.58560*x
        ELSE IF (r > R_400 .AND. r <= R_220) THEN
            rho = 7.10890-3.80450*x
            vp  = 20.39260-12.25690*x
            vs  = 8.94960-4.45970*x
        ELSE IF (r > R_220 .AND. r <= R_MOHO) THEN          ! LVZ and LID
            rho = 2.69100 + 0.69240*x
            vp  = 4.18750 + 3.93820*x
            vs  = 2.15190 + 2.34810*x
        ELSE IF (r > R_MOHO .AND. r <= R_MID_CRUST) THEN    ! lower crust
            rho = 2.90
            vp  = 6.80
            vs  = 3.90
        ELSE IF (r > R_MID_CRUST .AND. r <= R_OCEAN) THEN   ! upper crust
            rho = 2.60
            vp  = 5.80
            vs  = 3.20
        ELSE IF (r > R_OCEAN .AND. r <= R_EARTH) THEN       ! ocean
            rho = 1.020
            vp  = 1.450
            vs  = 0.0
        ELSE
            rho = 0.0
            vp  = 0.0
            vs  = 0.0
        END IF
        IF (noocean .AND. r > R_OCEAN) THEN
            rho = 2.60
            vp  = 5.80
            vs  = 3.20
        END IF
        IF (nocrust .AND. r > R_80) THEN
            rho = 2.69100+0.69240*(R_MOHO / R_EARTH)
            vp  = 4.18750+3.93820*(R_MOHO / R_EARTH)
            vs  = 2.15190+2.34810*(R_MOHO / R_EARTH)
        END IF
    END SUBROUTINE

    !> PREM with its transversely isotropic LVZ and LID at radius r [m]
    !! (elsewhere the isotropic model with the ocean replaced by crust).
    ELEMENTAL SUBROUTINE prem_ti_at(r, rho, vpv, vph, vsv, vsh, eta)
        REAL(rk), INTENT(IN) :: r
        REAL(rk), INTENT(OUT) :: rho, vpv, vph, vsv, vsh, eta
        REAL(rk) :: x
        x = ABS(r / R_EARTH)
        IF (r > R_220 .AND. r <= R_MOHO) THEN
            rho = 2.69100 + 0.69240*x
            vph = 3.59080 + 4.61720*x
            vpv = 0.83710 + 7.21800*x
            vsh = -1.08390 + 5.71760*x
            vsv = 5.85820 - 1.46780*x
            eta = 3.36870 - 2.47780*x
        ELSE
            CALL prem_iso_at(r, .FALSE., .TRUE., rho, vpv, vsv)
            vph = vpv
            vsh = vsv
            eta = 1.0
        END IF
    END SUBROUTINE

    !> ak135-f at radius x [km]: linear interpolation between the table
    !! nodes around x (nearest node first, ties to the shallower one); at a
    !! discontinuity the deeper node's values are returned.
    ELEMENTAL SUBROUTINE ak135f_at(x, rho, vp, vs)
        REAL(rk), INTENT(IN) :: x
        REAL(rk), INTENT(OUT) :: rho, vp, vs
        REAL(rk) :: d, best, w
        INTEGER :: n, q, lo, hi
        ! Above the surface the model is empty (the intent of
        ! src/models/ak135_f_mod.F90:243-249, whose lookup then continues
        ! with an index outside the table).
        IF (ABS(x) > AK(1, 1)) THEN
            rho = 0.0
            vp = 0.0
            vs = 0.0
            RETURN
        END IF
        n = 1
        best = ABS(AK(1, 1) - ABS(x))
        DO q = 2, NAK
            d = ABS(AK(1, q) - ABS(x))
            IF (d < best) THEN
                best = d
                n = q
            END IF
        END DO
        IF (AK(1, n) < ABS(x)) THEN
            lo = n
            hi = n - 1
        ELSE
            ! at the centre the last node is its own neighbour
            lo = MIN(n + 1, NAK)
            hi = n
        END IF
        IF (AK(1, lo) == AK(1, hi)) THEN
            w = 0.0
        ELSE
            w = (x - AK(1, lo)) / (AK(1, hi) - AK(1, lo))
        END IF
        rho = AK(2, lo) + (AK(2, hi) - AK(2, lo)) * w
        vp  = AK(3, lo) + (AK(3, hi) - AK(3, lo)) * w
        vs  = AK(4, lo) + (AK(4, hi) - AK(4, lo)) * w
    END SUBROUTINE

    !> Moduli of built-in model `model_type` at radius r [m] of element layer
    !! ez (0 = top), in the reference's parametrisation
    !! (src/model_paras.F90:219-336): rho [kg/m^3], lambda, mu and the
    !! anisotropy A, B, C [Pa]. ok = .FALSE. for an unknown model_type.
    ELEMENTAL SUBROUTINE builtin_model(model_type, r, ez, rho, lambda, mu, a, b, c, ok)
        INTEGER, INTENT(IN) :: model_type, ez
        REAL(rk), INTENT(IN) :: r
        REAL(rk), INTENT(OUT) :: rho, lambda, mu, a, b, c
        LOGICAL, INTENT(OUT) :: ok
        REAL(rk) :: vp, vs, vpv, vph, vsv, vsh, eta
        ok = .TRUE.
        a = 0.0
        b = 0.0
        c = 0.0
        SELECT CASE (model_type)
        CASE (1)                        ! homogeneous (src/models/homogeneous.f90)
            rho = 3000.0
            mu = 7.0e10
            lambda = 1.1e11
            RETURN
        CASE (11)                       ! PREM, anisotropic LVZ / LID
            CALL prem_ti_at(r, rho, vpv, vph, vsv, vsh, eta)
            rho = 1000.0 * rho
            vph = 1000.0 * vph
            vpv = 1000.0 * vpv
            vsh = 1000.0 * vsh
            vsv = 1000.0 * vsv
            mu = rho * vsh**2
            lambda = rho * vph**2 - 2.0 * mu
            a = rho * vph**2 - lambda - 2.0 * mu
            b = rho * vsv**2 - mu
            c = eta * (lambda + a) - lambda
            RETURN
        CASE (12)                       ! PREM isotropic, crust and ocean
            CALL prem_iso_at(r, .FALSE., .FALSE., rho, vp, vs)
        CASE (3, 13)                    ! PREM isotropic, crust, no ocean
            CALL prem_iso_at(r, .FALSE., .TRUE., rho, vp, vs)
        CASE (14)                       ! PREM isotropic, simplified crust
            CALL prem_iso_at(r, .TRUE., .TRUE., rho, vp, vs)
        CASE (2, 15)                    ! same, top element layer replaced
            CALL prem_iso_at(r, .TRUE., .TRUE., rho, vp, vs)
            IF (ez == 0) THEN
                rho = 3.00
                vp = 7.20
                vs = 4.10
            END IF
        CASE (21)                       ! ak135-f
            CALL ak135f_at(r / 1000.0, rho, vp, vs)
        CASE DEFAULT
            ok = .FALSE.
            rho = 0.0
            lambda = 0.0
            mu = 0.0
            RETURN
        END SELECT
        rho = 1000.0 * rho
        vp = 1000.0 * vp
        vs = 1000.0 * vs
        mu = rho * vs**2
        lambda = rho * vp**2 - 2.0 * mu
    END SUBROUTINE

END MODULE earth_models_mod
