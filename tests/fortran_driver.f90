! Minimal Fortran caller of libwesthip through the iso_c_binding module:
! the calls a WEST build makes around its Wstat path (set-up, then the
! reference call surfaces).  Without a device, wh_init must fail cleanly.
PROGRAM fortran_driver
  USE west_hip
  IMPLICIT NONE
  INTEGER(C_INT) :: rc
  rc = wh_init(0_C_INT)
  IF (rc /= 0) THEN
     WRITE(*,'(A,I0,A,A)') 'wh_init rc=', rc, ' msg=', wh_last_error()
     STOP 0
  ENDIF
  WRITE(*,'(A)') 'wh_init ok'
  rc = wh_finalize()
END PROGRAM
