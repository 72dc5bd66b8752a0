! Fortran caller of the hot path through the iso_c_binding module, as a WEST
! build would make the calls: registration of the ground state (the module
! globals dfpt reads, DFPT_kernel/dfpt_module.f90:24-54), apply_operator
! (Wstat/apply_operator.f90:14), dfpt without its optional ierr array
! (DFPT_kernel/dfpt_module.f90:22) and the Davidson driver
! (Wstat/davidson_diago.f90:31) with its per-step ev/conv log.
! Input: gs.bin (stream, written by tests/test_fortran_gpu.py); output: out.bin.
PROGRAM fortran_hotpath
  USE west_hip
  IMPLICIT NONE
  INTEGER(C_INT) :: nr1, nr2, nr3, npw, npwx, gstart, nkb, nbndval, m, nvec, rc, nlog, i
  REAL(C_DOUBLE) :: omega, tpiba2, weight, alpha_pv, tr2
  INTEGER(C_INT), ALLOCATABLE :: nl(:), nlm(:), conv(:), conv_log(:)
  REAL(C_DOUBLE), ALLOCATABLE :: g2kin(:), vrs(:), dvan(:), et(:), occ(:), sqvc(:), ev(:), ev_log(:)
  COMPLEX(C_DOUBLE_COMPLEX), ALLOCATABLE :: vkb(:), evc(:), dvg(:), dng(:), dng2(:)
  TYPE(wh_davidson_params) :: prm
  TYPE(wh_davidson_step), ALLOCATABLE :: steps(:)
  INTEGER :: iu
  !
  OPEN(NEWUNIT=iu, FILE='gs.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(iu) nr1, nr2, nr3, npw, npwx, gstart, nkb, nbndval, m, nvec
  READ(iu) omega, tpiba2, weight, alpha_pv, tr2
  ALLOCATE(nl(npw), nlm(npw), g2kin(npwx), vrs(nr1*nr2*nr3), vkb(MAX(1, npwx*nkb)), dvan(MAX(1, nkb)))
  ALLOCATE(evc(npwx*nbndval), et(nbndval), occ(nbndval), sqvc(npwx), dvg(npwx*m), dng(npwx*m), dng2(npwx*m))
  READ(iu) nl, nlm, g2kin, vrs
  IF (nkb > 0) READ(iu) vkb(1:npwx*nkb), dvan(1:nkb)
  READ(iu) evc, et, occ, sqvc, dvg
  CLOSE(iu)
  !
  rc = wh_init(0_C_INT)
  IF (rc /= 0) CALL fail('wh_init')
  rc = wh_set_grid(nr1, nr2, nr3, npw, npwx, gstart, nl, nlm, omega, tpiba2, g2kin)
  IF (rc /= 0) CALL fail('wh_set_grid')
  rc = wh_set_params(250_C_INT, 0_C_INT)
  rc = wh_set_potential(vrs)
  IF (rc /= 0) CALL fail('wh_set_potential')
  rc = wh_set_nonlocal(nkb, vkb, dvan)
  IF (rc /= 0) CALL fail('wh_set_nonlocal')
  rc = wh_set_states(nbndval, evc, et, occ, weight, alpha_pv)
  IF (rc /= 0) CALL fail('wh_set_states')
  rc = wh_set_coulomb(sqvc)
  IF (rc /= 0) CALL fail('wh_set_coulomb')
  !
  rc = wh_apply_operator(m, dvg, dng, tr2, 1_C_INT)
  IF (rc /= 0) CALL fail('wh_apply_operator')
  rc = wh_dfpt(m, dvg, dng2, tr2, 1_C_INT)          ! ierr_per_pert absent
  IF (rc /= 0) CALL fail('wh_dfpt')
  !
  prm%n_pdep_eigen = nvec
  prm%n_pdep_times = 4
  prm%n_pdep_maxiter = 100
  prm%trev_pdep = 1.0E-6_C_DOUBLE
  prm%trev_pdep_rel = 1.0E-4_C_DOUBLE
  prm%tr2_dfpt = 1.0E-20_C_DOUBLE
  ALLOCATE(ev(nvec), conv(nvec), steps(102), ev_log(nvec*102), conv_log(nvec*102))
  rc = wh_davidson_diago(prm, ev, conv, log=steps, log_capacity=102_C_INT, log_len=nlog, &
       & ev_log=ev_log, conv_log=conv_log)
  IF (rc < 0) CALL fail('wh_davidson_diago')
  !
  OPEN(NEWUNIT=iu, FILE='out.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(iu) dng, dng2, nlog, ev, conv
  WRITE(iu) ev_log(1:nvec*nlog), conv_log(1:nvec*nlog)
  WRITE(iu) (steps(i)%dav_iter, i = 1, nlog), (steps(i)%dfpt_dim, i = 1, nlog), (steps(i)%sternop_ncalls, i = 1, nlog)
  CLOSE(iu)
  rc = wh_finalize()
  WRITE(*,'(A)') 'fortran_hotpath ok'
CONTAINS
  SUBROUTINE fail(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    WRITE(*,'(A,A,A)') what, ' failed: ', wh_last_error()
    STOP 1
  END SUBROUTINE
END PROGRAM
