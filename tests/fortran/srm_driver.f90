! Test driver of the Fortran module m_scarflib (modelled on the reference's
! examples/fortran/driver.f90 SRM and FIM blocks, without MPI).
!   srm_driver config        : initialize with keyword options, print the
!                              parameters the library holds, finalize
!   srm_driver run <file>    : one structured 3D SRM realization written raw
!                              to <file>, stats printed
!   srm_driver fim <file>    : the same grid by FIM (method left at its
!                              default, 0), padded, rescaled and tapered
!                              around one point of interest
PROGRAM srm_driver
  USE, INTRINSIC :: iso_c_binding
  USE m_scarflib
  IMPLICIT NONE
  INTERFACE
    INTEGER(c_int) FUNCTION c_get_config(out, n) BIND(c, name="scarf_get_config")
      IMPORT :: c_int, c_double
      REAL(c_double), DIMENSION(*), INTENT(OUT) :: out
      INTEGER(c_int), VALUE :: n
    END FUNCTION
    INTEGER(c_int) FUNCTION c_set_nharm(n) BIND(c, name="scarf_set_nharm")
      IMPORT :: c_int
      INTEGER(c_int), VALUE :: n
    END FUNCTION
  END INTERFACE
  CHARACTER(LEN=256) :: mode, fname
  INTEGER(f_int), DIMENSION(3) :: fs, fe
  REAL(f_real), DIMENSION(3) :: cl
  REAL(f_real), DIMENSION(3, 2) :: poi
  REAL(f_real), ALLOCATABLE, DIMENSION(:,:,:) :: field
  REAL(f_real), DIMENSION(8) :: stats
  REAL(c_double), DIMENSION(32) :: cfg
  INTEGER :: rc, u

  CALL get_command_argument(1, mode)
  fs = [1, 1, 1]
  fe = [40, 36, 20]
  cl = [2000._f_real, 2000._f_real, 2000._f_real]
  poi = RESHAPE([300._f_real, 400._f_real, 200._f_real, 900._f_real, 500._f_real, 600._f_real], [3, 2])

  IF (TRIM(mode) == 'config') THEN
    CALL scarf_initialize(fs, fe, 50._f_real, 0, cl, 0.05_f_real, method=1, hurst=0.25_f_real, mute=0._f_real, &
                          taper=500._f_real, rescale=1, alpha=30._f_real, nc=[-10._f_real, -20._f_real, -30._f_real])
    rc = c_get_config(cfg, 32)
    WRITE(*, '(32ES24.15)') cfg
    CALL scarf_finalize()
  ELSE IF (TRIM(mode) == 'fim') THEN
    CALL get_command_argument(2, fname)
    CALL scarf_initialize(fs, fe, 50._f_real, 0, cl, 0.05_f_real, hurst=0.25_f_real, poi=poi(:, 1:1), &
                          mute=100._f_real, taper=250._f_real, rescale=1, pad=1)
    ALLOCATE(field(fe(1), fe(2), fe(3)))
    CALL scarf_execute(1234, field, stats)
    IF (LEN_TRIM(scarf_last_error()) > 0) THEN
      WRITE(*, *) 'error: ', TRIM(scarf_last_error())
      STOP 1
    ENDIF
    CALL scarf_io(fe, field, TRIM(fname))
    WRITE(*, '(8ES24.15)') stats
    CALL scarf_finalize()
  ELSE
    CALL get_command_argument(2, fname)
    rc = c_set_nharm(384)
    CALL scarf_initialize(fs, fe, 50._f_real, 0, cl, 0.05_f_real, method=1, hurst=0.25_f_real, poi=poi, &
                          mute=100._f_real, taper=250._f_real, rescale=1)
    ALLOCATE(field(fe(1), fe(2), fe(3)))
    CALL scarf_execute(1234, field, stats)
    IF (LEN_TRIM(scarf_last_error()) > 0) THEN
      WRITE(*, *) 'error: ', TRIM(scarf_last_error())
      STOP 1
    ENDIF
    CALL scarf_io(fe, field, TRIM(fname))
    WRITE(*, '(8ES24.15)') stats
    CALL scarf_finalize()
  ENDIF
END PROGRAM srm_driver
