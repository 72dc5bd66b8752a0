// Library state: the ground state registered through the C ABI, device
// workspaces, the stream, counters and the optional per-kernel timer.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "wh_common.h"
#include "wh_fft.h"

namespace wh {

// Growable device buffer (allocated once, reused across calls).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t n) {
    if (n > bytes) {
      // work queued on the device may still use the old block
      if (p) {
        WH_HIP(hipDeviceSynchronize());
        WH_HIP(hipFree(p));
      }
      p = nullptr;
      bytes = 0;
      WH_HIP(hipMalloc(&p, n));
      bytes = n;
    }
    return p;
  }
  template <typename T>
  T* as(size_t count) {
    return static_cast<T*>(get(sizeof(T) * std::max<size_t>(count, 1)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// Pinned host staging (hipHostMalloc) for the small per-iteration copies of
// the CG (band lists, coefficients, rho): asynchronous DMA instead of staged
// pageable copies.  Same growth rule as DevBuf.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t n) {
    if (n > bytes) {
      if (p) {
        WH_HIP(hipDeviceSynchronize());
        WH_HIP(hipHostFree(p));
      }
      p = nullptr;
      bytes = 0;
      WH_HIP(hipHostMalloc(&p, n, hipHostMallocDefault));
      bytes = n;
    }
    return p;
  }
  template <typename T>
  T* as(size_t count) {
    return static_cast<T*>(get(sizeof(T) * std::max<size_t>(count, 1)));
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct Profiler {
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<Pending> pending;
  std::map<std::string, std::pair<double, long>> acc;
  hipEvent_t take();
  void flush();
  void clear();
};

struct KPath;   // the q/k-point path's state (wh_kpoint.hip)

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev_caller = nullptr;   // orders the stream after the caller's null-stream work
  // grid / sphere
  bool have_grid = false;
  int nr1 = 0, nr2 = 0, nr3 = 0, nnr = 0, npw = 0, npwx = 0, gstart = 2;
  double omega = 0.0, tpiba2 = 0.0;
  StickPlan plan;
  int* d_perm = nullptr;
  double* d_g2kin = nullptr;
  std::vector<double> g2kin_s;   // stick order (host)
  // potential
  bool have_pot = false;
  double* d_vrs = nullptr;
  // states and projectors (stick order, rows of npw complex)
  bool have_states = false;
  int nbndval = 0, nkb = 0;
  std::vector<z64> evc_s, vkb_s;
  std::vector<double> et, occ, dvan, eprec;
  double weight = 2.0, alpha_pv = 0.0;
  z64* d_q = nullptr;            // [evc ; vkb], (nbndval + nkb) rows
  int q_rows = 0;
  double* d_qscale = nullptr;    // per Q row: alpha (evc rows) or dvan (vkb rows)
  double qscale_alpha = 0.0;     // alpha currently stored in d_qscale
  double* d_occ = nullptr;
  double* d_eprec = nullptr;
  bool have_sqvc = false;
  double* d_sqvc = nullptr;      // stick order
  std::vector<double> sqvc_s;
  // precomputed sticks of the occupied pairs (dfpt right-hand side)
  DevBuf evc_sticks;
  bool evc_sticks_valid = false;
  // params
  int n_dfpt_maxiter = 250;
  // Davidson restart files (wstat_restart_dir, n_steps_write_restart and a
  // restart run, Modules/westcom.f90; Wstat/davidson_restart.f90)
  std::string restart_dir;
  int n_steps_write_restart = 0;
  int restart_from = 0;
  int l_kinetic_only = 0;
  size_t sticks_budget = size_t(4) << 30;   // bytes for transform intermediates
  size_t vectors_budget = size_t(64) << 30;  // bytes for the CG vectors of one dfpt chunk
  // workspaces
  DevBuf w_items, w_items2, w_items_id, w_sticks_a, w_sticks_b, w_cmat;
  DevBuf w_x, w_t, w_g, w_h, w_hold, w_bc, w_tc, w_b, w_dvr, w_in, w_out, w_dvin, w_dng;
  DevBuf w_l2i, w_beta, w_eu, w_rho, w_rhob, w_ecol, w_epcol, w_dscr;
  HostBuf h_l2i, h_beta, h_eu, h_rho;   // pinned staging of the FP64 CG loop
  // the FP64 CG's per-iteration upload, one copy: [beta (nc) | e (nc) | list (nc ints)]
  DevBuf w_cgpk;
  DevBuf lz_q, lz_r, lz_sc;   // Lanczos basis, residual and scalars (wh_lanczos.hip)
  DevBuf w_conv;               // device-driven CG: flags, count and record
  HostBuf h_cgrec;             // per-iteration (active count, calls), pinned
  HostBuf h_cgpk;
  HostBuf hm_l2i, hm_beta, hm_eu, hm_rho, hm_eu64, hm_u;   // and of the mixed-precision loop
  // mixed precision: FP32 operator with split-BF16 projectors, FP64 refinement
  int precision = 0;              // 0: FP64, 1: mixed (3-product split BF16), 2: mixed (6-product)
  double refine_delta = 1e-4;     // refine when rho < refine_delta * rho of the last refinement
  bool mixed_valid = false;       // FP32 copies below match the registered state
  long ld32 = 0;                  // FP32 vector row length in complex numbers (npw rounded up to 16)
  long nq_pad = 0;                // Q rows rounded up to 32
  DevBuf m_vrs32, m_q32, m_qp[3], m_qtp[3];   // Q in FP32; its bf16 pieces; pieces of Q^T
  DevBuf m_g, m_h, m_hold, m_dx, m_bc, m_tc, m_cm, m_work, m_u2i, m_rho64, m_eu64;
  DevBuf m_fb, m_fx, m_fe, m_fep;   // bands finished by the FP64 CG (right sides, solutions, e, eprec)
  DevBuf dav[9];                    // Davidson basis, products and subspace arrays
  // counters
  long nstern = 0;     // operator applications, in band columns (all precisions)
  long nstern_calls = 0;   // operator calls as the reference counts them: one per perturbation's linsolve per application
  long nstern32 = 0;   // of which FP32
  long nrefine = 0;    // FP64 refinement residuals (band count)
  long nfallback = 0;  // bands the mixed solve handed to the FP64 CG
  long cg_iters = 0;
  Profiler prof;
  // k-point ground state (wh_set_kgrid): when present, the operator, CG and
  // chi0 entry points act at k points / q points instead of Gamma
  KPath* kpath = nullptr;

  long sticks_bytes() const { return plan.sticks_elems() * static_cast<long>(sizeof(z64)); }
  int items_capacity() const;   // transform items per launch chunk
};

Context& ctx();
// Make the current system's stream wait for the work the caller queued on the
// null (default) stream, e.g. the kernels that produced a device array it
// passes in; neither the host nor other streams wait.
void wait_for_caller(Context& c);
// the q/k-point path (wh_kpoint.hip): state release, and the C-ABI hot-path
// entries when a k grid is registered in the current system
void free_kpath(KPath* k);
bool kpath_on();
int kpath_apply_sternheimerop(int nbndval, const double* psi, double* hpsi, const double* e, const double* alpha,
                              int m);
int kpath_linsolve(int nbndval, int m, double* b, double* x, const double* e, const double* eprec, double tr2,
                   int* ierr);
int kpath_dfpt(int m, const double* dvg, double* dng, double tr2, int iq, int* ierr_pert, bool coulomb);
// takes this device's address of the allocation-free operator's scratch array (wh_noalloc.hip)
void noalloc_init(int device);
void require_ready(bool need_states);
// true for device (or managed) memory, false for host memory
bool on_device_mem(const void* p);

// timed launch: the statement runs between two events when profiling is on
#define WH_TIMED(NAME, STMT)                                                       \
  do {                                                                             \
    ::wh::Context& c_ = ::wh::ctx();                                               \
    if (c_.prof.on) {                                                              \
      hipEvent_t a_ = c_.prof.take(), b_ = c_.prof.take();                         \
      WH_HIP(hipEventRecord(a_, c_.stream));                                       \
      STMT;                                                                        \
      WH_HIP(hipEventRecord(b_, c_.stream));                                       \
      c_.prof.pending.push_back({NAME, a_, b_});                                   \
    } else {                                                                       \
      STMT;                                                                        \
    }                                                                              \
  } while (0)

// operations on device data in stick order (rows of npw complex, ld = npw)
// Buffers handed to stern_apply by a caller that must not allocate: FFT item
// lists (npair and chunk items), two stick buffers of chunk items, the
// projection coefficients (nq x m), the bra product's split-K partial sums
// and its variant, and the projector row scales with alpha in the evc rows.
struct SternWs {
  FftItem* items;
  FftItem* ident;
  z64* sa;
  z64* sb;
  int chunk;
  double* cmat;
  double* partials;
  int bra_variant;
  const double* qscale;
};
// d_nc (device int, may be null): m is a capacity and only the first *d_nc
// columns are active (the device-driven CG); inactive columns are skipped.
void stern_apply(const z64* x, z64* t, const double* d_e, int m, double alpha, const SternWs* ws = nullptr,
                 const int* d_nc = nullptr);
// maxiter <= 0: n_dfpt_maxiter
int linsolve_dev(int m, z64* b, z64* x, const std::vector<double>& e_col, const double* d_e_col,
                 const double* d_eprec_col, double tr2, std::vector<int>* ierr_col, int maxiter = 0, int group = 0,
                 const int* col_id = nullptr);
int dfpt_dev(int m, const z64* dvg, z64* dng, double tr2, const double* out_weight, std::vector<int>* ierr_pert);
void glbrak_q(int row0, int nq, const z64* f, int m, double* cmat);
void stern_apply32(const z32* x, z32* t, const double* d_e, int m, double alpha);
int linsolve_mixed(int m, z64* b, z64* x, const std::vector<double>& e_col, const double* d_e_col,
                   const double* d_eprec_col, double tr2, std::vector<int>* ierr_col, int group = 0);
// The reference's 'stern' clock counts calls of apply_sternheimerop, one per
// perturbation's linsolve per CG iteration (DFPT_kernel/dfpt_module.f90:356,
// linsolve_sternheimer_m_wfcts.f90:127,251).  Columns here are
// perturbation-major, `group` bands per perturbation (group <= 0: one
// perturbation); this adds the perturbations among the n columns listed in
// ascending order in cols (null: columns 0..n-1), each column mapped through
// col_id when given.
void count_calls(int n, const int* cols, int group, const int* col_id = nullptr);   // cmat (nq x m) = <Q_row|f_col>

}  // namespace wh
