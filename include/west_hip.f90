!
! west_hip.f90 -- iso_c_binding interface to libwesthip.so (include/west_hip.h).
!
! A WEST build that links libwesthip.so USEs this module and calls the
! wh_* functions where it calls apply_operator (Wstat/apply_operator.f90:14),
! dfpt (DFPT_kernel/dfpt_module.f90:22), linsolve_sternheimer_m_wfcts
! (DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:14) and davidson_diago
! (Wstat/davidson_diago.f90:31).  Argument order and meaning are the
! reference's; scalars are passed by value, arrays by reference.
!
MODULE west_hip
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  !
  TYPE, BIND(C) :: wh_davidson_params
     INTEGER(C_INT) :: n_pdep_eigen
     INTEGER(C_INT) :: n_pdep_times
     INTEGER(C_INT) :: n_pdep_maxiter
     REAL(C_DOUBLE) :: trev_pdep
     REAL(C_DOUBLE) :: trev_pdep_rel
     REAL(C_DOUBLE) :: tr2_dfpt
  END TYPE
  !
  TYPE, BIND(C) :: wh_davidson_step
     INTEGER(C_INT) :: dav_iter
     INTEGER(C_INT) :: notcnv
     INTEGER(C_INT) :: dfpt_dim
     INTEGER(C_INT) :: diago_dim
     INTEGER(C_LONG) :: sternop_ncalls
     REAL(C_DOUBLE) :: dfpt_tr2
     REAL(C_DOUBLE) :: time_elap
     REAL(C_DOUBLE) :: time_iter
  END TYPE
  !
  INTERFACE
     !
     INTEGER(C_INT) FUNCTION wh_init(device) BIND(C, NAME="wh_init")
       IMPORT :: C_INT
       INTEGER(C_INT), VALUE :: device
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_finalize() BIND(C, NAME="wh_finalize")
       IMPORT :: C_INT
     END FUNCTION
     !
     TYPE(C_PTR) FUNCTION wh_last_error_c() BIND(C, NAME="wh_last_error")
       IMPORT :: C_PTR
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_grid(nr1, nr2, nr3, npw, npwx, gstart, nl, nlm, omega, tpiba2, g2kin) &
          & BIND(C, NAME="wh_set_grid")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: nr1, nr2, nr3, npw, npwx, gstart
       INTEGER(C_INT), INTENT(IN) :: nl(*), nlm(*)
       REAL(C_DOUBLE), VALUE :: omega, tpiba2
       REAL(C_DOUBLE), INTENT(IN) :: g2kin(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_potential(vrs) BIND(C, NAME="wh_set_potential")
       IMPORT :: C_INT, C_DOUBLE
       REAL(C_DOUBLE), INTENT(IN) :: vrs(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_nonlocal(nkb, vkb, dvan) BIND(C, NAME="wh_set_nonlocal")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: nkb
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: vkb(*)
       REAL(C_DOUBLE), INTENT(IN) :: dvan(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_states(nbndval, evc, et, occ, weight, alpha_pv) BIND(C, NAME="wh_set_states")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: nbndval
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: evc(*)
       REAL(C_DOUBLE), INTENT(IN) :: et(*), occ(*)
       REAL(C_DOUBLE), VALUE :: weight, alpha_pv
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_coulomb(sqvc) BIND(C, NAME="wh_set_coulomb")
       IMPORT :: C_INT, C_DOUBLE
       REAL(C_DOUBLE), INTENT(IN) :: sqvc(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_params(n_dfpt_maxiter, l_kinetic_only) BIND(C, NAME="wh_set_params")
       IMPORT :: C_INT
       INTEGER(C_INT), VALUE :: n_dfpt_maxiter, l_kinetic_only
     END FUNCTION
     !
     ! dir: TRIM(wstat_restart_dir)//C_NULL_CHAR
     INTEGER(C_INT) FUNCTION wh_set_davidson_restart(dir, n_steps_write_restart, restart_from) &
         & BIND(C, NAME="wh_set_davidson_restart")
       IMPORT :: C_INT, C_CHAR
       CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: dir
       INTEGER(C_INT), VALUE :: n_steps_write_restart, restart_from
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_system_create(handle) BIND(C, NAME="wh_system_create")
       IMPORT :: C_INT
       INTEGER(C_INT), INTENT(OUT) :: handle
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_system_select(handle) BIND(C, NAME="wh_system_select")
       IMPORT :: C_INT
       INTEGER(C_INT), VALUE :: handle
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_system_destroy(handle) BIND(C, NAME="wh_system_destroy")
       IMPORT :: C_INT
       INTEGER(C_INT), VALUE :: handle
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_set_precision(mode, refine_delta) BIND(C, NAME="wh_set_precision")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: mode
       REAL(C_DOUBLE), VALUE :: refine_delta
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_apply_sternheimerop_to_m_wfcs(nbndval, psi, hpsi, e, alpha, m) &
          & BIND(C, NAME="wh_apply_sternheimerop_to_m_wfcs")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: nbndval, m
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: psi(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: hpsi(*)
       REAL(C_DOUBLE), INTENT(IN) :: e(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: alpha
     END FUNCTION
     !
     ! alpha_s(NLSTEPS, NRHS), beta_s(NLSTEPS-1, NRHS), q_s(npwx, NRHS, NLSTEPS), bnorm(NRHS)
     INTEGER(C_INT) FUNCTION wh_solve_deflated_lanczos_w_full_ortho(nbnd_to_deflate, nrhs, nlsteps, b, alpha_s, &
          & beta_s, q_s, bnorm) BIND(C, NAME="wh_solve_deflated_lanczos_w_full_ortho")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: nbnd_to_deflate, nrhs, nlsteps
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: b(*)
       REAL(C_DOUBLE), INTENT(OUT) :: alpha_s(*), bnorm(*)
       REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: beta_s(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT), OPTIONAL :: q_s(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_glbrak_gamma(na, a, nb, b, c) BIND(C, NAME="wh_glbrak_gamma")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: na, nb
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: a(*), b(*)
       REAL(C_DOUBLE), INTENT(OUT) :: c(*)
     END FUNCTION
     !
     ! no reference counterpart: the operator without device allocation (device arrays only)
     INTEGER(C_INT) FUNCTION wh_apply_sternheimerop_to_m_wfcs_noalloc(nbndval, psi, hpsi, e, alpha, m) &
          & BIND(C, NAME="wh_apply_sternheimerop_to_m_wfcs_noalloc")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: nbndval, m
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: psi(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: hpsi(*)
       REAL(C_DOUBLE), INTENT(IN) :: e(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: alpha
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_linsolve_sternheimer_m_wfcts(nbndval, m, b, x, e, eprec, tr2, ierr) &
          & BIND(C, NAME="wh_linsolve_sternheimer_m_wfcts")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: nbndval, m
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(INOUT) :: b(*), x(*)
       REAL(C_DOUBLE), INTENT(IN) :: e(*), eprec(*)
       REAL(C_DOUBLE), VALUE :: tr2
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: ierr
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_dfpt(m, dvg, dng, tr2, iq, ierr_per_pert) BIND(C, NAME="wh_dfpt")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: m, iq
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: dvg(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: dng(*)
       REAL(C_DOUBLE), VALUE :: tr2
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: ierr_per_pert(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_apply_operator(m, dvg, dng, tr2, iq) BIND(C, NAME="wh_apply_operator")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: m, iq
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: dvg(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT) :: dng(*)
       REAL(C_DOUBLE), VALUE :: tr2
     END FUNCTION
     !
     ! ev_log(n_pdep_eigen, log_capacity), conv_log(n_pdep_eigen, log_capacity):
     ! the ev and conv of each step, as output_ev_and_time records them
     ! q points (Wstat away from Gamma): dfpt_module.f90 non-Gamma branches,
     ! davidson_diago_k (Wstat/davidson_diago.f90:441-905)
     INTEGER(C_INT) FUNCTION wh_set_kgrid(nr1, nr2, nr3, omega, vrs, nks, wk, npwx, nqs, npwqx, nbndval, nkb, &
          & dvan, alpha_pv) BIND(C, NAME="wh_set_kgrid")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: nr1, nr2, nr3, nks, npwx, nqs, npwqx, nbndval, nkb
       REAL(C_DOUBLE), VALUE :: omega, alpha_pv
       REAL(C_DOUBLE), INTENT(IN) :: vrs(*), wk(*)
       REAL(C_DOUBLE), INTENT(IN), OPTIONAL :: dvan(*)
     END FUNCTION

     INTEGER(C_INT) FUNCTION wh_set_kpoint(ik, npw, nl, g2kin, evc, et, vkb) BIND(C, NAME="wh_set_kpoint")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX
       INTEGER(C_INT), VALUE :: ik, npw
       INTEGER(C_INT), INTENT(IN) :: nl(*)
       REAL(C_DOUBLE), INTENT(IN) :: g2kin(*), et(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN) :: evc(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(IN), OPTIONAL :: vkb(*)
     END FUNCTION

     INTEGER(C_INT) FUNCTION wh_set_qpoint(iq, npwq, nl, gq2kin, ig_global, sqvc, ikq, g0) &
          & BIND(C, NAME="wh_set_qpoint")
       IMPORT :: C_INT, C_DOUBLE
       INTEGER(C_INT), VALUE :: iq, npwq
       INTEGER(C_INT), INTENT(IN) :: nl(*), ig_global(*), ikq(*), g0(*)
       REAL(C_DOUBLE), INTENT(IN) :: gq2kin(*)
       REAL(C_DOUBLE), INTENT(IN), OPTIONAL :: sqvc(*)
     END FUNCTION

     INTEGER(C_INT) FUNCTION wh_set_current_k(ik) BIND(C, NAME="wh_set_current_k")
       IMPORT :: C_INT
       INTEGER(C_INT), VALUE :: ik
     END FUNCTION

     INTEGER(C_INT) FUNCTION wh_davidson_diago_k(iq, params, ev, conv, dvg_out, log, log_capacity, log_len, &
          & ev_log, conv_log) BIND(C, NAME="wh_davidson_diago_k")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX, wh_davidson_params, wh_davidson_step
       INTEGER(C_INT), VALUE :: iq
       TYPE(wh_davidson_params), INTENT(IN) :: params
       REAL(C_DOUBLE), INTENT(OUT) :: ev(*)
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: conv(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT), OPTIONAL :: dvg_out(*)
       TYPE(wh_davidson_step), INTENT(OUT), OPTIONAL :: log(*)
       INTEGER(C_INT), VALUE :: log_capacity
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: log_len
       REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: ev_log(*)
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: conv_log(*)
     END FUNCTION

     INTEGER(C_INT) FUNCTION wh_davidson_diago(params, ev, conv, dvg_out, log, log_capacity, log_len, &
          & ev_log, conv_log) BIND(C, NAME="wh_davidson_diago")
       IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX, wh_davidson_params, wh_davidson_step
       TYPE(wh_davidson_params), INTENT(IN) :: params
       REAL(C_DOUBLE), INTENT(OUT) :: ev(*)
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: conv(*)
       COMPLEX(C_DOUBLE_COMPLEX), INTENT(OUT), OPTIONAL :: dvg_out(*)
       TYPE(wh_davidson_step), INTENT(OUT), OPTIONAL :: log(*)
       INTEGER(C_INT), VALUE :: log_capacity
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: log_len
       REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: ev_log(*)
       INTEGER(C_INT), INTENT(OUT), OPTIONAL :: conv_log(*)
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_get_counters(stern_ncalls, cg_iters) BIND(C, NAME="wh_get_counters")
       IMPORT :: C_INT, C_LONG
       INTEGER(C_LONG), INTENT(OUT), OPTIONAL :: stern_ncalls, cg_iters
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_get_operator_columns(band_applications) BIND(C, NAME="wh_get_operator_columns")
       IMPORT :: C_INT, C_LONG
       INTEGER(C_LONG), INTENT(OUT) :: band_applications
     END FUNCTION
     !
     INTEGER(C_INT) FUNCTION wh_get_precision_counters(fp32_applies, refinements, fp64_fallbacks) &
          & BIND(C, NAME="wh_get_precision_counters")
       IMPORT :: C_INT, C_LONG
       INTEGER(C_LONG), INTENT(OUT) :: fp32_applies, refinements, fp64_fallbacks
     END FUNCTION
     !
  END INTERFACE
  !
CONTAINS
  !
  FUNCTION wh_last_error() RESULT(msg)
    ! message of the last failing call, as a Fortran string
    CHARACTER(LEN=:), ALLOCATABLE :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: c(:)
    TYPE(C_PTR) :: p
    INTEGER :: n
    p = wh_last_error_c()
    CALL C_F_POINTER(p, c, [4096])
    n = 0
    DO WHILE (n < 4096)
       IF (c(n + 1) == C_NULL_CHAR) EXIT
       n = n + 1
    ENDDO
    ALLOCATE(CHARACTER(LEN=n) :: msg)
    msg = TRANSFER(c(1:n), msg)
  END FUNCTION
  !
END MODULE west_hip
