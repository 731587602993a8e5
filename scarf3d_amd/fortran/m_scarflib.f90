! m_scarflib: Fortran interface of the MI355X SCARF3D SRM library.
!
! Same module name and generic interfaces as the reference Fortran library
! (src/scarflib.f90:48-62): scarf_initialize (structured / unstructured,
! optional keyword arguments), scarf_execute (3D, 2D, point list),
! scarf_finalize and scarf_io. It is a thin iso_c_binding layer over the C ABI
! of include/scarf3d.h: struct scarf_opt carries what the C presence rules
! can express, and scarf_set_option applies every PRESENT optional argument
! with Fortran semantics afterwards (so e.g. mute = 0 is kept, as
! src/scarflib.f90:163-175 keeps it). Compile with -cpp, and -DDOUBLE_PREC
! for the double-precision library (libscarf3d.so).
MODULE m_scarflib

  USE, INTRINSIC :: iso_c_binding

  IMPLICIT NONE

  PRIVATE

  PUBLIC :: scarf_initialize, scarf_execute, scarf_finalize, scarf_io, f_real, f_int, scarf_last_error

#ifdef DOUBLE_PREC
  INTEGER, PARAMETER :: f_real = c_double
#else
  INTEGER, PARAMETER :: f_real = c_float
#endif
  INTEGER, PARAMETER :: f_int = c_int

  ! struct scarf_opt (include/scarf3d.h)
  TYPE, BIND(c) :: scarf_opt
    INTEGER(c_int) :: method
    REAL(f_real)   :: hurst, dh, ds
    TYPE(c_ptr)    :: poi
    INTEGER(c_int) :: npoi
    REAL(f_real)   :: mute, taper
    INTEGER(c_int) :: rescale, pad
    TYPE(c_ptr)    :: nc, fc
    REAL(f_real)   :: alpha, beta, gamma
  END TYPE scarf_opt

  INTERFACE scarf_initialize
    MODULE PROCEDURE initialize_structured, initialize_unstructured
  END INTERFACE

  INTERFACE scarf_execute
    MODULE PROCEDURE execute_structured_2d, execute_structured_3d, execute_unstructured
  END INTERFACE

  INTERFACE scarf_io
    MODULE PROCEDURE scarf_io_slice, scarf_io_one_2d, scarf_io_one_3d
  END INTERFACE

  INTERFACE
    SUBROUTINE c_opt_init(var) BIND(c, name="scarf_opt_init")
      IMPORT :: scarf_opt
      TYPE(scarf_opt), INTENT(INOUT) :: var
    END SUBROUTINE
    SUBROUTINE c_cart_initialize(nd, fs, fe, ds, acf, cl, sigma, var) BIND(c, name="scarf_cart_initialize")
      IMPORT :: c_int, f_real, scarf_opt
      INTEGER(c_int), VALUE :: nd, acf
      INTEGER(c_int), DIMENSION(*), INTENT(IN) :: fs, fe
      REAL(f_real), VALUE :: ds, sigma
      REAL(f_real), DIMENSION(*), INTENT(IN) :: cl
      TYPE(scarf_opt), INTENT(INOUT) :: var
    END SUBROUTINE
    SUBROUTINE c_nocart_initialize(nd, npts, x, y, z, dh, acf, cl, sigma, var) BIND(c, name="scarf_nocart_initialize")
      IMPORT :: c_int, c_ptr, f_real, scarf_opt
      INTEGER(c_int), VALUE :: nd, npts, acf
      TYPE(c_ptr), VALUE :: x, y, z
      REAL(f_real), VALUE :: dh, sigma
      REAL(f_real), DIMENSION(*), INTENT(IN) :: cl
      TYPE(scarf_opt), INTENT(INOUT) :: var
    END SUBROUTINE
    SUBROUTINE c_execute(seed, field, stats) BIND(c, name="scarf_execute")
      IMPORT :: c_int, c_ptr, f_real
      INTEGER(c_int), VALUE :: seed
      TYPE(c_ptr), VALUE :: field
      REAL(f_real), DIMENSION(*), INTENT(OUT) :: stats
    END SUBROUTINE
    SUBROUTINE c_finalize() BIND(c, name="scarf_finalize")
    END SUBROUTINE
    SUBROUTINE c_io_one(nd, npts, field, fname, nwriters) BIND(c, name="scarf_io_one")
      IMPORT :: c_int, c_ptr, c_char
      INTEGER(c_int), VALUE :: nd
      INTEGER(c_int), DIMENSION(*), INTENT(IN) :: npts
      TYPE(c_ptr), VALUE :: field, nwriters
      CHARACTER(kind=c_char), DIMENSION(*), INTENT(IN) :: fname
    END SUBROUTINE
    SUBROUTINE c_io_slice(npts, axis, plane, field, fname) BIND(c, name="scarf_io_slice")
      IMPORT :: c_int, c_ptr, c_char
      INTEGER(c_int), DIMENSION(*), INTENT(IN) :: npts
      CHARACTER(kind=c_char), DIMENSION(*), INTENT(IN) :: axis, fname
      INTEGER(c_int), VALUE :: plane
      TYPE(c_ptr), VALUE :: field
    END SUBROUTINE
    INTEGER(c_int) FUNCTION c_set_option(name, value) BIND(c, name="scarf_set_option")
      IMPORT :: c_int, c_char, c_double
      CHARACTER(kind=c_char), DIMENSION(*), INTENT(IN) :: name
      REAL(c_double), VALUE :: value
    END FUNCTION
    TYPE(c_ptr) FUNCTION c_last_error() BIND(c, name="scarf_last_error")
      IMPORT :: c_ptr
    END FUNCTION
  END INTERFACE

  ! points of interest handed to the library (it copies them at initialize)
  REAL(f_real), ALLOCATABLE, DIMENSION(:,:), TARGET, SAVE :: poi_buf

CONTAINS

  SUBROUTINE set(name, value)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(c_double),   INTENT(IN) :: value
    INTEGER(c_int) :: rc
    rc = c_set_option(TRIM(name) // c_null_char, value)
  END SUBROUTINE set

  SUBROUTINE apply_optionals(method, hurst, dh, ds, mute, taper, rescale, pad, nc, fc, alpha, beta, gamma)
    INTEGER(f_int),               OPTIONAL, INTENT(IN) :: method, rescale, pad
    REAL(f_real),                 OPTIONAL, INTENT(IN) :: hurst, dh, ds, mute, taper, alpha, beta, gamma
    REAL(f_real), DIMENSION(:),   OPTIONAL, INTENT(IN) :: nc, fc
    CHARACTER(LEN=3) :: key
    INTEGER :: i
    IF (PRESENT(method))  CALL set("method",  REAL(method, c_double))
    IF (PRESENT(hurst))   CALL set("hurst",   REAL(hurst, c_double))
    IF (PRESENT(dh))      CALL set("dh",      REAL(dh, c_double))
    IF (PRESENT(ds))      CALL set("ds",      REAL(ds, c_double))
    IF (PRESENT(mute))    CALL set("mute",    REAL(mute, c_double))
    IF (PRESENT(taper))   CALL set("taper",   REAL(taper, c_double))
    IF (PRESENT(rescale)) CALL set("rescale", REAL(rescale, c_double))
    IF (PRESENT(pad))     CALL set("pad",     REAL(pad, c_double))
    IF (PRESENT(alpha))   CALL set("alpha",   REAL(alpha, c_double))
    IF (PRESENT(beta))    CALL set("beta",    REAL(beta, c_double))
    IF (PRESENT(gamma))   CALL set("gamma",   REAL(gamma, c_double))
    IF (PRESENT(nc)) THEN
      DO i = 1, SIZE(nc)
        WRITE(key, '(A2,I1)') 'nc', i
        CALL set(key, REAL(nc(i), c_double))
      ENDDO
    ENDIF
    IF (PRESENT(fc)) THEN
      DO i = 1, SIZE(fc)
        WRITE(key, '(A2,I1)') 'fc', i
        CALL set(key, REAL(fc(i), c_double))
      ENDDO
    ENDIF
  END SUBROUTINE apply_optionals

  ! initialize_structured (src/scarflib.f90:97-189)
  SUBROUTINE initialize_structured(fs, fe, ds, acf, cl, sigma, method, hurst, dh, poi, mute, taper, rescale, pad, nc, fc, &
                                   alpha, beta, gamma)
    INTEGER(f_int), DIMENSION(:),                   INTENT(IN) :: fs, fe
    REAL(f_real),                                   INTENT(IN) :: ds
    INTEGER(f_int),                                 INTENT(IN) :: acf
    REAL(f_real),   DIMENSION(:),                   INTENT(IN) :: cl
    REAL(f_real),                                   INTENT(IN) :: sigma
    INTEGER(f_int),                       OPTIONAL, INTENT(IN) :: method, rescale, pad
    REAL(f_real),                         OPTIONAL, INTENT(IN) :: hurst, dh, mute, taper, alpha, beta, gamma
    REAL(f_real),   DIMENSION(:,:),       OPTIONAL, INTENT(IN) :: poi
    REAL(f_real),   DIMENSION(:),         OPTIONAL, INTENT(IN) :: nc, fc
    TYPE(scarf_opt) :: opt
    INTEGER(c_int), DIMENSION(SIZE(fs)) :: cfs, cfe

    CALL c_opt_init(opt)
    IF (PRESENT(poi)) THEN
      IF (SIZE(poi, 2) > 0) THEN
        IF (ALLOCATED(poi_buf)) DEALLOCATE(poi_buf)
        ALLOCATE(poi_buf(SIZE(poi, 1), SIZE(poi, 2)))
        poi_buf = poi
        opt%poi  = c_loc(poi_buf)
        opt%npoi = SIZE(poi, 2)
      ENDIF
    ENDIF
    cfs = fs
    cfe = fe
    CALL c_cart_initialize(SIZE(fs), cfs, cfe, ds, acf, cl, sigma, opt)
    CALL apply_optionals(method=method, hurst=hurst, dh=dh, mute=mute, taper=taper, rescale=rescale, pad=pad, nc=nc, &
                         fc=fc, alpha=alpha, beta=beta, gamma=gamma)
  END SUBROUTINE initialize_structured

  ! initialize_unstructured (src/scarflib.f90:195-296); x, y, z are borrowed until finalize
  SUBROUTINE initialize_unstructured(dh, acf, cl, sigma, x, y, z, method, hurst, ds, poi, mute, taper, rescale, pad, nc, &
                                     fc, alpha, beta, gamma)
    REAL(f_real),                                   INTENT(IN) :: dh
    INTEGER(f_int),                                 INTENT(IN) :: acf
    REAL(f_real),   DIMENSION(:),                   INTENT(IN) :: cl
    REAL(f_real),                                   INTENT(IN) :: sigma
    REAL(f_real),   DIMENSION(:), CONTIGUOUS, TARGET,           INTENT(IN) :: x, y
    REAL(f_real),   DIMENSION(:), CONTIGUOUS, TARGET, OPTIONAL, INTENT(IN) :: z
    INTEGER(f_int),                       OPTIONAL, INTENT(IN) :: method, rescale, pad
    REAL(f_real),                         OPTIONAL, INTENT(IN) :: hurst, ds, mute, taper, alpha, beta, gamma
    REAL(f_real),   DIMENSION(:,:),       OPTIONAL, INTENT(IN) :: poi
    REAL(f_real),   DIMENSION(:),         OPTIONAL, INTENT(IN) :: nc, fc
    TYPE(scarf_opt) :: opt
    TYPE(c_ptr) :: zp

    CALL c_opt_init(opt)
    IF (PRESENT(poi)) THEN
      IF (SIZE(poi, 2) > 0) THEN
        IF (ALLOCATED(poi_buf)) DEALLOCATE(poi_buf)
        ALLOCATE(poi_buf(SIZE(poi, 1), SIZE(poi, 2)))
        poi_buf = poi
        opt%poi  = c_loc(poi_buf)
        opt%npoi = SIZE(poi, 2)
      ENDIF
    ENDIF
    zp = c_null_ptr
    IF (PRESENT(z)) zp = c_loc(z)
    CALL c_nocart_initialize(SIZE(cl), SIZE(x), c_loc(x), c_loc(y), zp, dh, acf, cl, sigma, opt)
    CALL apply_optionals(method=method, hurst=hurst, ds=ds, mute=mute, taper=taper, rescale=rescale, pad=pad, nc=nc, &
                         fc=fc, alpha=alpha, beta=beta, gamma=gamma)
  END SUBROUTINE initialize_unstructured

  SUBROUTINE execute_structured_3d(seed, field, stats)
    INTEGER(f_int),                               INTENT(IN)  :: seed
    REAL(f_real),   DIMENSION(:,:,:), CONTIGUOUS, TARGET, INTENT(OUT) :: field
    REAL(f_real),   DIMENSION(8),                 INTENT(OUT) :: stats
    CALL c_execute(seed, c_loc(field), stats)
  END SUBROUTINE execute_structured_3d

  SUBROUTINE execute_structured_2d(seed, field, stats)
    INTEGER(f_int),                             INTENT(IN)  :: seed
    REAL(f_real),   DIMENSION(:,:), CONTIGUOUS, TARGET, INTENT(OUT) :: field
    REAL(f_real),   DIMENSION(8),               INTENT(OUT) :: stats
    CALL c_execute(seed, c_loc(field), stats)
  END SUBROUTINE execute_structured_2d

  SUBROUTINE execute_unstructured(seed, field, stats)
    INTEGER(f_int),                           INTENT(IN)  :: seed
    REAL(f_real),   DIMENSION(:), CONTIGUOUS, TARGET, INTENT(OUT) :: field
    REAL(f_real),   DIMENSION(8),             INTENT(OUT) :: stats
    CALL c_execute(seed, c_loc(field), stats)
  END SUBROUTINE execute_unstructured

  SUBROUTINE scarf_finalize()
    CALL c_finalize()
    IF (ALLOCATED(poi_buf)) DEALLOCATE(poi_buf)
  END SUBROUTINE scarf_finalize

  SUBROUTINE scarf_io_slice(n, axis, plane, field, filename)
    INTEGER(f_int),   DIMENSION(3),                         INTENT(IN) :: n
    CHARACTER(LEN=1),                                       INTENT(IN) :: axis
    INTEGER(f_int),                                         INTENT(IN) :: plane
    REAL(f_real),     DIMENSION(:,:,:), CONTIGUOUS, TARGET, INTENT(IN) :: field
    CHARACTER(LEN=*),                                       INTENT(IN) :: filename
    INTEGER(c_int), DIMENSION(3) :: cn
    cn = n
    CALL c_io_slice(cn, axis // c_null_char, plane, c_loc(field), TRIM(filename) // c_null_char)
  END SUBROUTINE scarf_io_slice

  SUBROUTINE scarf_io_one_3d(n, field, filename, nwriters)
    INTEGER(f_int),   DIMENSION(3),                         INTENT(IN) :: n
    REAL(f_real),     DIMENSION(:,:,:), CONTIGUOUS, TARGET, INTENT(IN) :: field
    CHARACTER(LEN=*),                                       INTENT(IN) :: filename
    INTEGER(f_int),                             OPTIONAL,   INTENT(IN) :: nwriters
    INTEGER(c_int), DIMENSION(3) :: cn
    INTEGER(c_int), TARGET :: nw
    cn = n
    nw = 1
    IF (PRESENT(nwriters)) nw = nwriters
    CALL c_io_one(3, cn, c_loc(field), TRIM(filename) // c_null_char, c_loc(nw))
  END SUBROUTINE scarf_io_one_3d

  SUBROUTINE scarf_io_one_2d(n, field, filename, nwriters)
    INTEGER(f_int),   DIMENSION(2),                       INTENT(IN) :: n
    REAL(f_real),     DIMENSION(:,:), CONTIGUOUS, TARGET, INTENT(IN) :: field
    CHARACTER(LEN=*),                                     INTENT(IN) :: filename
    INTEGER(f_int),                           OPTIONAL,   INTENT(IN) :: nwriters
    INTEGER(c_int), DIMENSION(2) :: cn
    INTEGER(c_int), TARGET :: nw
    cn = n
    nw = 1
    IF (PRESENT(nwriters)) nw = nwriters
    CALL c_io_one(2, cn, c_loc(field), TRIM(filename) // c_null_char, c_loc(nw))
  END SUBROUTINE scarf_io_one_2d

  ! last error message of the library ('' when the last call succeeded)
  FUNCTION scarf_last_error() RESULT(msg)
    CHARACTER(LEN=256) :: msg
    CHARACTER(kind=c_char), DIMENSION(:), POINTER :: p
    INTEGER :: i
    msg = ''
    CALL c_f_pointer(c_last_error(), p, [256])
    DO i = 1, 256
      IF (p(i) == c_null_char) EXIT
      msg(i:i) = p(i)
    ENDDO
  END FUNCTION scarf_last_error

END MODULE m_scarflib
