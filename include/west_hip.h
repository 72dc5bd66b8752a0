/*
 * west_hip.h -- C ABI of the MI355X Wstat dielectric-response library
 * (libwesthip.so).
 *
 * It replaces, behind the reference's own call surfaces, the Gamma-point
 * hot path of WEST's Wstat:
 *
 *   davidson_diago_gamma  (Wstat/davidson_diago.f90:31-437)
 *     -> apply_operator   (Wstat/apply_operator.f90:14-69)
 *       -> dfpt           (DFPT_kernel/dfpt_module.f90:22-533)
 *         -> linsolve_sternheimer_m_wfcts (DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:14-353)
 *           -> apply_sternheimerop_to_m_wfcs (DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:14-84)
 *
 * Conventions
 *  - Arrays follow Fortran: column major, leading dimension first.  A
 *    "complex" array is COMPLEX(DP), i.e. interleaved (re, im) doubles; the
 *    prototypes use double* for it.  Integers are C int; scalars by value.
 *  - Every array argument may be host memory or device memory of the
 *    library's device; the library detects which.  Calls are synchronous:
 *    outputs are complete when the call returns.
 *  - The caller owns every array; registered ground-state data is copied at
 *    wh_set_* and no caller pointer is retained.
 *  - Return value: 0 = success; > 0 = number of unconverged Sternheimer
 *    systems (a warning, like ierr in the reference); < 0 = error, with the
 *    message in wh_last_error().
 *  - One process drives one device.  Several host threads may call the
 *    library at once when each has selected its own system
 *    (wh_system_select); two threads on one system race.
 *  - Gamma point only (iq must be 1), nspin = 1, integer occupations,
 *    norm-conserving pseudopotentials with a diagonal D per projector.
 */
#ifndef WEST_HIP_H
#define WEST_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- lifecycle ------------------------------------------------------- */

/* Select the device and create the library's stream.  Replaces the device
 * set-up of west_gpu_start (Modules/west_gpu.f90:115-140). */
int wh_init(int device);
/* Release all device memory.  Replaces west_gpu's deallocations
 * (Modules/west_gpu.f90:194-227). */
int wh_finalize(void);
/* Message of the last failing call (errore in the reference). */
const char* wh_last_error(void);

/* ---- ground state: the module globals the reference reads ------------- */

/* FFT grid and wave-function sphere: dffts%nr1..3, npw, npwx, gstart,
 * dffts%nl / dffts%nlm (1-based, as in QE), omega, tpiba2, g2kin
 * (DFPT_kernel/dfpt_module.f90:27-35, FFT_kernel/fft_at_gamma.f90:59-63,
 *  wvfct%g2kin used by precondition_m_wfcts.f90:41 and set_eprec.f90:46). */
int wh_set_grid(int nr1, int nr2, int nr3, int npw, int npwx, int gstart, const int* nl, const int* nlm,
                double omega, double tpiba2, const double* g2kin);

/* Local potential vrs(nr1*nr2*nr3) in Ry (scf%vrs used by QE vloc_psi). */
int wh_set_potential(const double* vrs);

/* Kleinman-Bylander projectors vkb(npwx, nkb) (complex) and their diagonal
 * coefficients dvan(nkb) in Ry (uspp%vkb from init_us_2,
 * DFPT_kernel/dfpt_module.f90:157-161; deeq used by QE add_vuspsi). */
int wh_set_nonlocal(int nkb, const double* vkb, const double* dvan);

/* Occupied states: evc(npwx, nbndval) (complex), et(nbndval), occ(nbndval),
 * the k-point weight (k_grid%weight, dfpt_module.f90:486) and alpha_pv
 * (westcom alphapv_dfpt, Tools/get_alpha_pv.f90:35-57). */
int wh_set_states(int nbndval, const double* evc, const double* et, const double* occ, double weight,
                  double alpha_pv);

/* Coulomb factor sqvc(npwqx) (pot3D%sqvc, Coulomb_kernel/class_coulomb.f90:139-176). */
int wh_set_coulomb(const double* sqvc);

/* n_dfpt_maxiter (westcom, default 250) and l_kinetic_only
 * (apply_sternheimerop_to_m_wfcs.f90:49-50). */
int wh_set_params(int n_dfpt_maxiter, int l_kinetic_only);

/* Arithmetic of the Sternheimer operator and solve inside
 * apply_sternheimerop_to_m_wfcs, linsolve_sternheimer_m_wfcts, dfpt and
 * apply_operator. mode 0: FP64 throughout (the drop-in, same arithmetic as
 * DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:107-332).
 * modes 1 and 2: the operator runs in FP32 (FFTs, vectors; projector GEMMs on
 * the matrix cores in split BF16 with FP32 accumulation: 3 bf16 products per
 * product in mode 1, about 16 bits; 6 in mode 2, about FP32). The CG iterates
 * in FP32 on an increment of the FP64 solution (its first residual formed with
 * the FP32 operator); the FP64 residual b - A x is recomputed whenever the
 * FP32 rho has dropped by refine_delta (<= 0: 1e-6 in mode 1, 1e-8 in mode 2)
 * since the last recomputation, or after 40 FP32 iterations, and a band
 * converges only when that FP64 rho is below tr2 (the reference's test).
 * Bands whose FP64 residual twice drops less than tenfold between
 * recomputations, bands still unconverged after n_dfpt_maxiter/2 FP32
 * iterations and the last 1/128 of a large batch are finished by the FP64 CG.
 * The reference has no such mode; call surfaces and convergence test are
 * unchanged. Modes 1-2
 * need a cubic FFT grid of a size with compiled kernels (18, 20, 24, 30, 32,
 * 36, 40, 45, 48, 50, 54, 60, 64, 72, 75, 80, 90, 96, 100, 108). */
int wh_set_precision(int mode, double refine_delta);

/* ---- several systems ------------------------------------------------- */

/* Independent ground states (systems) in one process on one GPU. wh_init
 * creates system 0; wh_system_create adds one (own HIP stream, workspaces and
 * counters) and returns its handle; wh_system_select makes it the target of
 * every following call from the calling host thread (each thread starts on
 * system 0). Systems selected in different threads run concurrently. No
 * counterpart in the reference (WEST runs one system per process); it serves
 * batches of small, unevenly sized systems that one system alone cannot fill
 * the GPU with. */
int wh_system_create(int* handle);
int wh_system_select(int handle);
int wh_system_destroy(int handle);

/* ---- the hot path --------------------------------------------------- */

/* hpsi = H psi - e psi + alpha P_v psi for m bands (psi, hpsi: (npwx, m)
 * complex; alpha: one complex scalar, by reference).
 * Replaces apply_sternheimerop_to_m_wfcs (DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:14). */
int wh_apply_sternheimerop_to_m_wfcs(int nbndval, const double* psi, double* hpsi, const double* e,
                                     const double* alpha, int m);

/* Preconditioned CG for (H - e_i + alpha_pv P_v) x_i = b_i, i = 1..m.
 * b (npwx, m) complex is destroyed, x is the in/out guess; *ierr = number
 * of unconverged columns after n_dfpt_maxiter iterations.
 * Replaces linsolve_sternheimer_m_wfcts (DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:14). */
int wh_linsolve_sternheimer_m_wfcts(int nbndval, int m, double* b, double* x, const double* e, const double* eprec,
                                    double tr2, int* ierr);

/* dng = chi0 dvg for m perturbations (dvg, dng: (npwqx, m) complex); tr2 < 0
 * selects the lite path (no Sternheimer solve).  ierr_per_pert (m, may be
 * NULL) receives the unconverged count of each perturbation.
 * Replaces dfpt (DFPT_kernel/dfpt_module.f90:22). */
int wh_dfpt(int m, const double* dvg, double* dng, double tr2, int iq, int* ierr_per_pert);

/* dng = sqvc * chi0 (sqvc * dvg).  Replaces apply_operator (Wstat/apply_operator.f90:14). */
int wh_apply_operator(int m, const double* dvg, double* dng, double tr2, int iq);

/* Davidson for the n_pdep_eigen lowest eigenpairs of sqvc chi0 sqvc.
 * Inputs are the wstat_control values of westcom (Modules/westcom.f90:118-152);
 * the start vectors are the library's seeded random vectors (do_randomize,
 * Wstat/davidson_diago.f90:1121-1182).  Outputs: ev(n_pdep_eigen),
 * conv(n_pdep_eigen) (1 = converged), dvg_out(npwqx, n_pdep_eigen) complex
 * (may be NULL), and one log record per Davidson step (may be NULL; at most
 * log_capacity are kept, *log_len receives how many), with the step's
 * eigenvalues and convergence flags in ev_log(n_pdep_eigen, log_capacity)
 * and conv_log(n_pdep_eigen, log_capacity) (each may be NULL): the fields of
 * exec.davitr written by output_ev_and_time (davidson_diago.f90:1310-1329).
 * Replaces davidson_diago_gamma (Wstat/davidson_diago.f90:31). */
typedef struct {
  int n_pdep_eigen;
  int n_pdep_times;
  int n_pdep_maxiter;
  double trev_pdep;
  double trev_pdep_rel;
  double tr2_dfpt;
} wh_davidson_params;

typedef struct {            /* exec.davitr[] of output_ev_and_time (davidson_diago.f90:1310-1329) */
  int dav_iter;
  int notcnv;
  int dfpt_dim;
  int diago_dim;
  long sternop_ncalls;
  double dfpt_tr2;
  double time_elap;
  double time_iter;
} wh_davidson_step;

int wh_davidson_diago(const wh_davidson_params* params, double* ev, int* conv, double* dvg_out,
                      wh_davidson_step* log, int log_capacity, int* log_len, double* ev_log, int* conv_log);

/* Restart files of wh_davidson_diago, in the reference's format
 * (Wstat/davidson_restart.f90:33-222, 360-410, 413-462, 515-760; the inputs
 * wstat_restart_dir, n_steps_write_restart and wstat_calculation 'R' of
 * Modules/westcom.f90).  With n_steps_write_restart > 0 the Davidson writes
 * <dir>/summary.json (dav_iter, notcnv, nbase, conv, ev, ew), <dir>/hr_vr.dat
 * (hr then vr, nvecx x nvecx, Fortran unformatted sequential records) and the
 * basis <dir>/V%06d.dat, N%06d.dat (PDEP vector files: 32-integer header,
 * npw complex coefficients in the caller's G order) after the lite pass and
 * the second pass when it is 1 (davidson_diago.f90:198,249), then every
 * n_steps_write_restart steps (421-422), and removes them on convergence
 * (376).  restart_from = 1 starts the next wh_davidson_diago (that call only;
 * the ones after it start from scratch, as davidson_diago.f90:613 does) from
 * those files instead of random vectors (158).  dir may be NULL when both
 * are 0. */
int wh_set_davidson_restart(const char* wstat_restart_dir, int n_steps_write_restart, int restart_from);

/* ---- Wfreq: chi(i omega) by deflated Lanczos chains ------------------ */

/* NLSTEPS Lanczos steps of H restricted to the complement of the first
 * nbnd_to_deflate registered states (h_psi followed by P_c), for NRHS start
 * vectors b(npwx, NRHS) at once, with full re-orthogonalisation (modified
 * Gram-Schmidt against all earlier vectors of the chain).  Outputs
 * alpha_s(NLSTEPS, NRHS), beta_s(NLSTEPS-1, NRHS), the basis
 * q_s(npwx, NRHS, NLSTEPS) (may be NULL) and bnorm(NRHS) = |b|.
 * nbnd_to_deflate may not exceed the registered nbndval.  FP64.
 * Replaces solve_deflated_lanczos_w_full_ortho
 * (Wfreq/solve_deflated_lanczos_w_full_ortho.f90:14). */
int wh_solve_deflated_lanczos_w_full_ortho(int nbnd_to_deflate, int nrhs, int nlsteps, const double* b,
                                           double* alpha_s, double* beta_s, double* q_s, double* bnorm);

/* c(na, nb) = 2 Re <a_i|b_j> - Re a_i(G=0) Re b_j(G=0) over the registered
 * sphere, for a(npwx, na), b(npwx, nb) complex: the Gamma inner-product
 * matrix (Hamiltonian_kernel/glbrak.f90:14-43) of get_brak_hyper_parallel
 * (Wfreq/get_brak_hyper_parallel.f90:14), on the matrix cores. */
int wh_glbrak_gamma(int na, const double* a, int nb, const double* b, double* c);

/* ---- q points: Wstat away from Gamma ------------------------------- */

/* Ground state on a k grid (the gamma_only = .FALSE. inputs of dfpt,
 * DFPT_kernel/dfpt_module.f90:163-209: QE's wk, npwx, the k-point states and
 * projectors, nbnd_occ, alphapv_dfpt; the q grid of Tools/class_bz_grid.f90:71-131
 * and igq_q of Tools/set_npwq.f90:50-85).  After wh_set_kgrid the system's
 * operator, CG and chi0 entry points (wh_apply_sternheimerop_to_m_wfcs,
 * wh_linsolve_sternheimer_m_wfcts, wh_dfpt, wh_apply_operator) act at k and q
 * points: vectors are (npwx, m) on the |k+G| sphere of the current k point
 * (wh_set_current_k) or (npwqx, m) on the |q+G| sphere of q point iq; iq of
 * wh_dfpt / wh_apply_operator selects the q point (1..nqs).  wh_set_grid
 * returns the system to Gamma.
 *   wh_set_kgrid: FFT grid, cell volume, vrs (nr1*nr2*nr3, first index
 *     fastest), nks k points with weights wk (sum 2 for nspin = 1), row length
 *     npwx of k-point vectors, nqs q points with row length npwqx, occupied
 *     bands, nkb projectors with D = dvan (Ry), alpha_pv.
 *   wh_set_kpoint(ik): npw plane waves k+G with 1-based grid index nl
 *     (dffts%nl(igk_k)), g2kin = |k+G|^2 (Ry), evc (npwx, nbndval) and et
 *     (nbndval) at k, vkb (npwx, nkb) at k (init_us_2).
 *   wh_set_qpoint(iq): npwq vectors q+G (nl, |q+G|^2 in Ry, 1-based global G
 *     index ig_global, the counter of the random start vectors), sqvc (npwq,
 *     may be NULL: then only wh_dfpt), and for each k the 1-based index ikq
 *     of k - q on the k grid with g0 (3, nks) integer crystal components,
 *     k - q = k_ikq + g0 (findp, class_bz_grid.f90:166-214). */
int wh_set_kgrid(int nr1, int nr2, int nr3, double omega, const double* vrs, int nks, const double* wk, int npwx,
                 int nqs, int npwqx, int nbndval, int nkb, const double* dvan, double alpha_pv);
int wh_set_kpoint(int ik, int npw, const int* nl, const double* g2kin, const double* evc, const double* et,
                  const double* vkb);
int wh_set_qpoint(int iq, int npwq, const int* nl, const double* gq2kin, const int* ig_global, const double* sqvc,
                  const int* ikq, const int* g0);
/* k point (1..nks) of the operator and CG entry points: QE's current_k,
 * set by dfpt before h_psi (dfpt_module.f90:149). */
int wh_set_current_k(int ik);

/* One q point of the Davidson at q points (the QPOINTS_LOOP body of
 * davidson_diago_k, Wstat/davidson_diago.f90:551-894): start vectors of
 * do_randomize_q (1186-1250, the counter generator of wh_davidson_diago drawn
 * per global G, divided by |q+G|^2 + 1 in Ry), complex Gram-Schmidt and
 * subspace products, rocSOLVER zheevd.  Arguments as wh_davidson_diago;
 * dvg_out is (npwqx, n_pdep_eigen) on the q sphere.  No restart files. */
int wh_davidson_diago_k(int iq, const wh_davidson_params* params, double* ev, int* conv, double* dvg_out,
                        wh_davidson_step* log, int log_capacity, int* log_len, double* ev_log, int* conv_log);

/* ---- allocation-free operator -------------------------------------- */

/* wh_apply_sternheimerop_to_m_wfcs with no device allocation (and no variant
 * timing) inside the call: for callers that cannot allocate -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  psi, hpsi (npwx, m) and e (m) must be device
 * arrays; alpha is one complex scalar (host or device).  Every workspace
 * comes from a fixed 64 MiB device array per GPU (held until the process
 * ends); a call whose workspaces do not fit (about 2 m npw complex values plus
 * the projection and FFT buffers) fails with an error and launches nothing.
 * FP64 only (precision mode 0).  Calls that overlap in time on one GPU share
 * that array and race: serialise them.  Synchronous, like the other entry
 * points. */
int wh_apply_sternheimerop_to_m_wfcs_noalloc(int nbndval, const double* psi, double* hpsi, const double* e,
                                             const double* alpha, int m);

/* ---- counters -------------------------------------------------------- */

/* Operator calls so far as the reference's 'stern' clock counts them (one
 * call of apply_sternheimerop_to_m_wfcs per perturbation's Sternheimer solve
 * per CG step, DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:127,251;
 * test-suite/test001/ref/wstat.json:820-826: 3611 calls for 171
 * perturbations), whatever the batching here; and CG iterations. */
int wh_get_counters(long* stern_ncalls, long* cg_iters);

/* Operator applications so far counted in band columns (an application to m
 * bands counts m), all precisions: the work measure of the throughput. */
int wh_get_operator_columns(long* band_applications);

/* Of the operator applications: those made in FP32; the band residuals
 * recomputed in FP64 by the mixed-precision solve; the bands it handed to the
 * FP64 CG. */
int wh_get_precision_counters(long* fp32_applies, long* refinements, long* fp64_fallbacks);

/* Per-kernel device timing with HIP events on the library stream.
 * wh_profile(1) enables and clears, wh_profile(0) disables.
 * wh_profile_query(name, &ms, &count) returns the accumulated time (ms) and
 * launch count of one kernel class ("zinv", "plane_vloc", "zfwd",
 * "plane_density", "plane_to_real", "gemm_bra", "gemm_ket", "cg_prec",
 * "cg_dir", "cg_update"). */
int wh_profile(int enable);
int wh_profile_query(const char* name, double* ms, long* count);

#ifdef __cplusplus
}
#endif

#endif /* WEST_HIP_H */
