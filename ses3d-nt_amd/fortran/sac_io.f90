!> Binary SAC traces (float32 data, 158-word header, native endian): the
!! writer with the fields the reference's receiver output sets
!! (src/sac_io/sac_io_mod.f90:178-222, 453-502) and its file name
!! NETWORK.LOCATION.STATION.CHANNEL.sac (auto_file_name,
!! src/sac_io/sac_io_mod.f90:407-416), and the reader of the header fields a
!! &source_sac trace carries (read_sac, src/sac_io/sac_io_mod.f90:229-288).
MODULE sac_io_mod
    USE, INTRINSIC :: ISO_FORTRAN_ENV, ONLY : REAL32, INT32
    IMPLICIT NONE
    PRIVATE
    PUBLIC :: write_sac, read_sac, sac_file_name

CONTAINS

    !> Day of the year (src/sac_io/sac_header_mod.f90:380-389)
    INTEGER FUNCTION iday(year, month, day)
        INTEGER, INTENT(IN) :: year, month, day
        iday = 3055 * (month + 2) / 100 - (month + 10) / 13 * 2 - 91 + &
               (1 - (MODULO(year, 4) + 3) / 4 + (MODULO(year, 100) + 99) / 100 - &
               (MODULO(year, 400) + 399) / 400) * (month + 10) / 13 + day
    END FUNCTION

    FUNCTION sac_file_name(network, location, station, channel) RESULT(fn)
        CHARACTER(LEN=*), INTENT(IN) :: network, location, station, channel
        CHARACTER(LEN=:), ALLOCATABLE :: fn
        fn = TRIM(network) // '.' // TRIM(location) // '.' // TRIM(station) // '.' // &
             TRIM(channel) // '.sac'
    END FUNCTION

    SUBROUTINE write_sac(file, data, dt, lat, lon, depth, network, station, &
                         location, channel, event, date_time, override, iostat)
        CHARACTER(LEN=*), INTENT(IN) :: file, network, station, location, channel, event
        REAL(8), INTENT(IN) :: data(:), dt, lat, lon, depth
        INTEGER, INTENT(IN) :: date_time(8)
        LOGICAL, INTENT(IN) :: override
        INTEGER, INTENT(OUT) :: iostat
        REAL(REAL32) :: f(70)
        INTEGER(INT32) :: iv(35), lv(5)
        CHARACTER(LEN=8) :: k(24)
        CHARACTER(LEN=16) :: kevnm
        INTEGER :: lun
        CHARACTER(LEN=7) :: status

        f = -12345.0_REAL32
        iv = -12345
        iv(7) = 6                                   ! NVHDR
        lv = [1, 0, 0, 0, 0]                        ! LEVEN, LPSPOL, LOVROK, LCALDA, -
        k = '-12345  '
        f(1) = REAL(dt, REAL32)                     ! DELTA
        f(4) = 1.0_REAL32                           ! SCALE
        f(6) = 0.0_REAL32                           ! B
        f(7) = REAL(dt * SIZE(data), REAL32)        ! E
        f(32) = REAL(lat, REAL32)                   ! STLA
        f(33) = REAL(lon, REAL32)                   ! STLO
        f(35) = REAL(depth, REAL32)                 ! STDP
        iv(1) = date_time(1)                        ! NZYEAR
        iv(2) = iday(date_time(1), date_time(2), date_time(3))   ! NZJDAY
        iv(3) = date_time(5)                        ! NZHOUR
        iv(4) = date_time(6)                        ! NZMIN
        iv(5) = date_time(7)                        ! NZSEC
        iv(6) = date_time(8)                        ! NZMSEC
        iv(10) = SIZE(data)                         ! NPTS
        iv(16) = 1                                  ! IFTYPE = ITIME
        iv(25) = 2                                  ! ISYNTH
        lv(3) = MERGE(1, 0, override)               ! LOVROK
        k(1) = station                              ! KSTNM
        kevnm = event
        k(2) = kevnm(1:8)                           ! KEVNM (16 characters)
        k(3) = kevnm(9:16)
        k(4) = location                             ! KHOLE
        k(21) = channel                             ! KCMPNM
        k(22) = network                             ! KNETWK
        status = MERGE('replace', 'new    ', override)
        OPEN(NEWUNIT=lun, FILE=file, ACCESS='stream', FORM='unformatted', &
             ACTION='write', STATUS=TRIM(status), IOSTAT=iostat)
        IF (iostat /= 0) RETURN
        WRITE(lun) f
        WRITE(lun) iv
        WRITE(lun) lv
        WRITE(lun) k
        WRITE(lun) REAL(data, REAL32)
        CLOSE(lun)
    END SUBROUTINE

    !> Reads the first data section of a SAC file and the header fields
    !! STLA, STLO, STDP, SCALE and KCMPNM. iostat /= 0 when the file cannot
    !! be read or is truncated.
    SUBROUTINE read_sac(file, data, lat, lon, depth, scale, channel, iostat)
        CHARACTER(LEN=*), INTENT(IN) :: file
        REAL(8), ALLOCATABLE, INTENT(OUT) :: data(:)
        REAL(8), INTENT(OUT) :: lat, lon, depth, scale
        CHARACTER(LEN=8), INTENT(OUT) :: channel
        INTEGER, INTENT(OUT) :: iostat
        REAL(REAL32) :: f(70)
        INTEGER(INT32) :: iv(35), lv(5)
        CHARACTER(LEN=8) :: k(24)
        REAL(REAL32), ALLOCATABLE :: d(:)
        INTEGER :: lun

        OPEN(NEWUNIT=lun, FILE=file, ACCESS='stream', FORM='unformatted', &
             ACTION='read', STATUS='old', IOSTAT=iostat)
        IF (iostat /= 0) RETURN
        READ(lun, IOSTAT=iostat) f, iv, lv, k
        IF (iostat == 0 .AND. iv(10) < 0) iostat = -1
        IF (iostat == 0) THEN
            ALLOCATE(d(iv(10)))                     ! NPTS
            READ(lun, IOSTAT=iostat) d
        END IF
        CLOSE(lun)
        IF (iostat /= 0) RETURN
        data = REAL(d, 8)
        scale = REAL(f(4), 8)                       ! SCALE
        lat = REAL(f(32), 8)                        ! STLA
        lon = REAL(f(33), 8)                        ! STLO
        depth = REAL(f(35), 8)                      ! STDP
        channel = ADJUSTL(k(21))                    ! KCMPNM
    END SUBROUTINE

END MODULE sac_io_mod
