! Fortran caller of the q-point path through the iso_c_binding module: the
! k-grid registration a WEST build would make after wstat_setup (per k point
! the arrays of g2_kin/init_us_2 and the states, DFPT_kernel/dfpt_module.f90:144-209;
! per q point igq_q, sqvc and the k - q map, Tools/set_npwq.f90:50-85), then
! apply_operator at q point 2 (Wstat/apply_operator.f90:14 with iq) and the
! body of davidson_diago_k for q point 2 (Wstat/davidson_diago.f90:551-894).
! Input: kgs.bin (stream, written by tests/test_fortran_gpu.py); output: kout.bin.
PROGRAM fortran_qpoint
  USE west_hip
  IMPLICIT NONE
  INTEGER(C_INT) :: nr1, nr2, nr3, nks, npwx, nqs, npwqx, nbndval, nkb, m, nvec, rc, ik, iq, npw, nlog
  REAL(C_DOUBLE) :: omega, alpha_pv, tr2
  INTEGER(C_INT), ALLOCATABLE :: nl(:), ig(:), ikq(:), g0(:), conv(:)
  REAL(C_DOUBLE), ALLOCATABLE :: vrs(:), wk(:), dvan(:), g2kin(:), et(:), sqvc(:), ev(:)
  COMPLEX(C_DOUBLE_COMPLEX), ALLOCATABLE :: evc(:), vkb(:), dvg(:), dng(:)
  TYPE(wh_davidson_params) :: prm
  TYPE(wh_davidson_step) :: steps(42)
  INTEGER :: iu
  !
  OPEN(NEWUNIT=iu, FILE='kgs.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(iu) nr1, nr2, nr3, nks, npwx, nqs, npwqx, nbndval, nkb, m, nvec
  READ(iu) omega, alpha_pv, tr2
  ALLOCATE(vrs(nr1*nr2*nr3), wk(nks), dvan(MAX(1, nkb)))
  READ(iu) vrs, wk
  IF (nkb > 0) READ(iu) dvan(1:nkb)
  rc = wh_init(0_C_INT)
  IF (rc /= 0) CALL fail('wh_init')
  rc = wh_set_kgrid(nr1, nr2, nr3, omega, vrs, nks, wk, npwx, nqs, npwqx, nbndval, nkb, dvan, alpha_pv)
  IF (rc /= 0) CALL fail('wh_set_kgrid')
  rc = wh_set_params(250_C_INT, 0_C_INT)
  ALLOCATE(evc(npwx*nbndval), et(nbndval), vkb(MAX(1, npwx*nkb)), g2kin(npwx), nl(npwx))
  DO ik = 1, nks
     READ(iu) npw
     READ(iu) nl(1:npw), g2kin(1:npw), evc, et
     IF (nkb > 0) READ(iu) vkb(1:npwx*nkb)
     rc = wh_set_kpoint(ik, npw, nl, g2kin, evc, et, vkb)
     IF (rc /= 0) CALL fail('wh_set_kpoint')
  ENDDO
  ALLOCATE(ig(npwqx), sqvc(npwqx), ikq(nks), g0(3*nks))
  DEALLOCATE(nl, g2kin)
  ALLOCATE(nl(npwqx), g2kin(npwqx))
  DO iq = 1, nqs
     READ(iu) npw
     READ(iu) nl(1:npw), g2kin(1:npw), ig(1:npw), sqvc(1:npw), ikq, g0
     rc = wh_set_qpoint(iq, npw, nl, g2kin, ig, sqvc, ikq, g0)
     IF (rc /= 0) CALL fail('wh_set_qpoint')
  ENDDO
  ALLOCATE(dvg(npwqx*m), dng(npwqx*m))
  READ(iu) dvg
  CLOSE(iu)
  !
  rc = wh_apply_operator(m, dvg, dng, tr2, 2_C_INT)
  IF (rc /= 0) CALL fail('wh_apply_operator')
  prm%n_pdep_eigen = nvec
  prm%n_pdep_times = 4
  prm%n_pdep_maxiter = 40
  prm%trev_pdep = 1.0E-8_C_DOUBLE
  prm%trev_pdep_rel = 1.0E-7_C_DOUBLE
  prm%tr2_dfpt = 1.0E-18_C_DOUBLE
  ALLOCATE(ev(nvec), conv(nvec))
  rc = wh_davidson_diago_k(2_C_INT, prm, ev, conv, log=steps, log_capacity=42_C_INT, log_len=nlog)
  IF (rc < 0) CALL fail('wh_davidson_diago_k')
  !
  OPEN(NEWUNIT=iu, FILE='kout.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(iu) dng, ev, conv, nlog
  CLOSE(iu)
  rc = wh_finalize()
  WRITE(*,'(A)') 'fortran_qpoint ok'
CONTAINS
  SUBROUTINE fail(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    WRITE(*,'(A,A,A)') what, ' failed: ', wh_last_error()
    STOP 1
  END SUBROUTINE
END PROGRAM
