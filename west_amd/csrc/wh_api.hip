// C ABI of libwesthip: ground-state registration, the Sternheimer operator,
// the CG solver and chi0 (dfpt / apply_operator) on one MI355X.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>

#include "../../include/west_hip.h"
#include "wh_context.h"
#include "wh_fft_tpl.h"
#include "wh_fft_reg.h"
#include "wh_gemm.h"
#include "wh_gemm_bf16.h"
#include "wh_vec.h"

namespace wh {

// ---------------------------------------------------------------------------
// error state
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// systems: each registered ground state lives in its own Context (own stream,
// workspaces, counters); system 0 is created by wh_init. A host thread works
// on the system it selected last (system 0 by default), so independent
// systems can run concurrently from several threads on one GPU.
// ---------------------------------------------------------------------------
static std::mutex g_sys_mutex;
static std::vector<Context*> g_sys;   // index = handle; null after destroy
static int g_device = -1;
static thread_local int t_current = 0;

Context& ctx() {
  Context* c = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_sys_mutex);
    if (g_sys.empty()) throw Error{-1, "wh_init has not been called"};
    if (t_current >= 0 && t_current < static_cast<int>(g_sys.size())) c = g_sys[t_current];
  }
  if (!c) throw Error{-1, "the selected system does not exist"};
  return *c;
}

// true when this process holds one system: a stream capture must then not
// meet another thread's device-wide synchronisation
static bool single_system() {
  std::lock_guard<std::mutex> lk(g_sys_mutex);
  int n = 0;
  for (Context* c : g_sys) n += c != nullptr;
  return n == 1;
}

static Context* new_context(int device) {
  Context* c = new Context();
  c->device = device;
  WH_HIP(hipSetDevice(device));
  WH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  WH_HIP(hipEventCreateWithFlags(&c->ev_caller, hipEventDisableTiming));
  return c;
}

static void free_context(Context* cp) {
  Context& c = *cp;
  (void)hipStreamSynchronize(c.stream);
  if (c.ev_caller) (void)hipEventDestroy(c.ev_caller);
  c.prof.flush();
  for (auto e : c.prof.pool) (void)hipEventDestroy(e);
  c.plan.release();
  for (void* p : {static_cast<void*>(c.d_perm), static_cast<void*>(c.d_g2kin), static_cast<void*>(c.d_vrs),
                  static_cast<void*>(c.d_q), static_cast<void*>(c.d_qscale), static_cast<void*>(c.d_occ),
                  static_cast<void*>(c.d_eprec), static_cast<void*>(c.d_sqvc)})
    if (p) (void)hipFree(p);
  for (DevBuf* b : {&c.evc_sticks, &c.w_items, &c.w_items2, &c.w_items_id, &c.w_sticks_a, &c.w_sticks_b,
                    &c.w_cmat, &c.w_x, &c.w_t, &c.w_g, &c.w_h, &c.w_hold, &c.w_bc, &c.w_tc, &c.w_b, &c.w_dvr,
                    &c.w_in, &c.w_out, &c.w_dvin, &c.w_dng, &c.w_l2i, &c.w_beta, &c.w_eu, &c.w_rho, &c.w_ecol,
                    &c.w_epcol, &c.w_dscr, &c.m_vrs32, &c.m_q32, &c.m_qp[0], &c.m_qp[1], &c.m_qp[2], &c.m_qtp[0],
                    &c.m_qtp[1], &c.m_qtp[2], &c.m_g, &c.m_h, &c.m_hold, &c.m_dx, &c.m_bc, &c.m_tc, &c.m_cm,
                    &c.m_work, &c.m_u2i, &c.m_rho64, &c.m_eu64, &c.m_fb, &c.m_fx, &c.m_fe, &c.m_fep,
                    &c.w_rhob, &c.w_cgpk, &c.lz_q, &c.lz_r, &c.lz_sc, &c.w_conv})
    b->release();
  for (HostBuf* b : {&c.h_l2i, &c.h_beta, &c.h_eu, &c.h_rho, &c.hm_l2i, &c.hm_beta, &c.hm_eu, &c.hm_rho, &c.hm_eu64,
                     &c.hm_u, &c.h_cgpk, &c.h_cgrec})
    b->release();
  for (DevBuf& b : c.dav) b.release();
  free_kpath(c.kpath);
  c.kpath = nullptr;
  (void)hipStreamDestroy(c.stream);
  delete cp;
}

void require_ready(bool need_states) {
  Context& c = ctx();
  WH_REQUIRE(c.have_grid, "wh_set_grid has not been called");
  WH_REQUIRE(c.l_kinetic_only || c.have_pot, "wh_set_potential has not been called");
  if (need_states) WH_REQUIRE(c.have_states, "wh_set_states has not been called");
}

int Context::items_capacity() const {
  const long per = 2L * sticks_bytes();
  long cap = static_cast<long>(sticks_budget) / std::max<long>(per, 1);
  // items are the grid's y dimension in the transform launches: at most 65535
  return static_cast<int>(std::max<long>(1, std::min<long>(cap, 65535L)));
}

// ---------------------------------------------------------------------------
// profiler
// ---------------------------------------------------------------------------
hipEvent_t Profiler::take() {
  if (pool.empty()) {
    hipEvent_t e;
    WH_HIP(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = pool.back();
  pool.pop_back();
  return e;
}

void Profiler::flush() {
  for (auto& p : pending) {
    WH_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    WH_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& slot = acc[p.name];
    slot.first += ms;
    slot.second += 1;
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
}

void Profiler::clear() {
  flush();
  acc.clear();
}

// ---------------------------------------------------------------------------
// host/device staging of caller arrays
// ---------------------------------------------------------------------------
static bool on_device(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

bool on_device_mem(const void* p) { return on_device(p); }

void wait_for_caller(Context& c) {
  WH_HIP(hipEventRecord(c.ev_caller, nullptr));
  WH_HIP(hipStreamWaitEvent(c.stream, c.ev_caller, 0));
}

// device view of an input array (copied if it is host memory)
template <typename T>
static const T* dev_in(const T* p, size_t count, DevBuf& tmp) {
  if (on_device(p)) {
    // the library's stream does not synchronise with the null stream by
    // itself: order it after the caller's null-stream work (a device array
    // produced on another stream must be complete when the call is made)
    wait_for_caller(ctx());
    return p;
  }
  T* d = tmp.as<T>(count);
  WH_HIP(hipMemcpyAsync(d, p, sizeof(T) * count, hipMemcpyHostToDevice, ctx().stream));
  return d;
}

// ---------------------------------------------------------------------------
// projectors
// ---------------------------------------------------------------------------
static void rebuild_q() {
  Context& c = ctx();
  c.mixed_valid = false;
  const int rows = c.nbndval + c.nkb;
  if (c.d_q) WH_HIP(hipFree(c.d_q));
  if (c.d_qscale) WH_HIP(hipFree(c.d_qscale));
  c.d_q = nullptr;
  c.d_qscale = nullptr;
  c.q_rows = rows;
  if (rows == 0) return;
  std::vector<z64> q(static_cast<size_t>(rows) * c.npw);
  if (c.nbndval) std::memcpy(q.data(), c.evc_s.data(), sizeof(z64) * c.nbndval * c.npw);
  if (c.nkb) std::memcpy(q.data() + static_cast<size_t>(c.nbndval) * c.npw, c.vkb_s.data(), sizeof(z64) * c.nkb * c.npw);
  WH_HIP(hipMalloc(&c.d_q, sizeof(z64) * q.size()));
  WH_HIP(hipMemcpy(c.d_q, q.data(), sizeof(z64) * q.size(), hipMemcpyHostToDevice));
  std::vector<double> sc(rows);
  for (int i = 0; i < c.nbndval; ++i) sc[i] = c.alpha_pv;
  for (int i = 0; i < c.nkb; ++i) sc[c.nbndval + i] = c.dvan[i];
  c.qscale_alpha = c.alpha_pv;
  WH_HIP(hipMalloc(&c.d_qscale, sizeof(double) * rows));
  WH_HIP(hipMemcpy(c.d_qscale, sc.data(), sizeof(double) * rows, hipMemcpyHostToDevice));
}

static void set_qscale_alpha(double alpha) {
  Context& c = ctx();
  if (c.nbndval == 0 || alpha == c.qscale_alpha) return;
  std::vector<double> sc(c.nbndval, alpha);
  WH_HIP(hipMemcpy(c.d_qscale, sc.data(), sizeof(double) * c.nbndval, hipMemcpyHostToDevice));
  c.qscale_alpha = alpha;
}

// cmat (nq x m, col major) = glbrak_gamma(Q[row0:row0+nq], f[0:m])
//   = 2 Q_r^T f_r - Re Q(G=0) Re f(G=0)     (Hamiltonian_kernel/glbrak.f90:38-39)
void glbrak_q(int row0, int nq, const z64* f, int m, double* cmat) {
  Context& c = ctx();
  const long k2 = 2L * c.npw;
  const double* q = reinterpret_cast<const double*>(c.d_q) + static_cast<long>(row0) * k2;
  const double* fr = reinterpret_cast<const double*>(f);
  WH_TIMED("gemm_bra", dgemm_bra(nq, m, static_cast<int>(k2), 2.0, q, k2, fr, k2, c.gstart == 2 ? 1.0 : 0.0,
                                 nullptr, cmat, nq, c.stream));
}

// ---------------------------------------------------------------------------
// Sternheimer operator: t = H x - e x + alpha P_v x
//   (DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:49-76; QE h_psi_ at Gamma:
//    kinetic + vloc_psi + calbec/add_vuspsi)
// ---------------------------------------------------------------------------
void stern_apply(const z64* x, z64* t, const double* d_e, int m, double alpha, const SternWs* ws,
                 const int* d_nc) {
  Context& c = ctx();
  if (m <= 0) return;
  hipStream_t s = c.stream;
  const int npw = c.npw;
  const long ld = npw;
  if (c.l_kinetic_only) {
    WH_TIMED("kinetic", kinetic_minus_e(m, npw, x, t, ld, c.d_g2kin, d_e, s));
  } else {
    const int npair = (m + 1) / 2;
    const int cap = ws ? ws->chunk : c.items_capacity();
    const int chunk = std::min(cap, npair);
    FftItem* items = ws ? ws->items : c.w_items.as<FftItem>(npair);
    FftItem* ident = nullptr;
    if (d_nc) {
      // items of the inactive tail are switched off on the device
      fill_pairs_dev(items, npair, d_nc, s);
      ident = c.w_items_id.as<FftItem>(npair);
      fill_ident_dev(ident, npair, chunk, d_nc, s);
    } else {
      fill_pairs(items, m, 0, s);
      ident = ws ? ws->ident : c.w_items_id.as<FftItem>(chunk);
      fill_ident(ident, chunk, s);
    }
    z64* sa = ws ? ws->sa : c.w_sticks_a.as<z64>(static_cast<size_t>(chunk) * c.plan.sticks_elems());
    z64* sb = ws ? ws->sb : c.w_sticks_b.as<z64>(static_cast<size_t>(chunk) * c.plan.sticks_elems());
    ZfwdArgs za{ZFWD_PAIR_STERN, 1.0, c.d_g2kin, d_e, nullptr, 0};
    for (int q0 = 0; q0 < npair; q0 += chunk) {
      const int nq = std::min(chunk, npair - q0);
      WH_TIMED("zinv", launch_zinv(c.plan, items + q0, nq, x, ld, sa, ZINV_PAIR, s));
      WH_TIMED("plane_vloc", launch_plane_vloc(c.plan, d_nc ? ident + q0 : ident, nq, sa, sb, c.d_vrs, 0,
                                               1.0 / c.nnr, s));
      WH_TIMED("zfwd", launch_zfwd(c.plan, items + q0, nq, sb, x, ld, t, ld, za, s));
    }
  }
  // projector part: rows [evc ; vkb] (or evc only for the kinetic-only operator)
  const int nq = c.l_kinetic_only ? c.nbndval : c.nbndval + c.nkb;
  if (nq > 0) {
    const long k2 = 2L * npw;
    const double* q = reinterpret_cast<const double*>(c.d_q);
    const double* xr = reinterpret_cast<const double*>(x);
    if (ws) {
      WH_TIMED("gemm_bra", dgemm_bra_into(nq, m, static_cast<int>(k2), 2.0, q, k2, xr, k2, c.gstart == 2 ? 1.0 : 0.0,
                                          ws->qscale, ws->cmat, nq, ws->partials, ws->bra_variant, s));
      WH_TIMED("gemm_ket", dgemm_untimed(false, static_cast<int>(k2), m, nq, 1.0, q, k2, ws->cmat, nq, 1.0,
                                         reinterpret_cast<double*>(t), k2, s));
    } else {
      set_qscale_alpha(alpha);
      double* cm = c.w_cmat.as<double>(static_cast<size_t>(nq) * m);
      WH_TIMED("gemm_bra", dgemm_bra(nq, m, static_cast<int>(k2), 2.0, q, k2, xr, k2, c.gstart == 2 ? 1.0 : 0.0,
                                     c.d_qscale, cm, nq, s, d_nc));
      WH_TIMED("gemm_ket", dgemm(false, static_cast<int>(k2), m, nq, 1.0, q, k2, cm, nq, 1.0,
                                 reinterpret_cast<double*>(t), k2, s, d_nc));
    }
  }
  if (!d_nc) c.nstern += m;   // device-driven callers add their counts themselves
}

void count_calls(int n, const int* cols, int group, const int* col_id) {
  Context& c = ctx();
  long calls = 0;
  int last = -1;
  for (int k = 0; k < n; ++k) {
    int col = cols ? cols[k] : k;
    if (col_id) col = col_id[col];
    const int p = group > 0 ? col / group : 0;
    if (p != last) ++calls;
    last = p;
  }
  c.nstern_calls += calls;
}

// ---------------------------------------------------------------------------
// Preconditioned CG  (DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:107-332)
// ---------------------------------------------------------------------------
int linsolve_dev(int m, z64* b, z64* x, const std::vector<double>& e_col, const double* d_e_col,
                 const double* d_eprec_col, double tr2, std::vector<int>* ierr_col, int maxiter, int group,
                 const int* col_id) {
  (void)e_col;
  Context& c = ctx();
  hipStream_t s = c.stream;
  const int npw = c.npw;
  const long ld = npw;
  const size_t vec = static_cast<size_t>(m) * npw;
  z64* g = c.w_g.as<z64>(vec);
  z64* h = c.w_h.as<z64>(vec);
  z64* hold = c.w_hold.as<z64>(vec);
  z64* bc = c.w_bc.as<z64>(vec);
  z64* tc = c.w_tc.as<z64>(vec);
  // the iteration's active list and coefficients, built on the device
  int* d_l2i = c.w_l2i.as<int>(m);
  double* d_beta = c.w_beta.as<double>(m);
  double* d_eu = c.w_eu.as<double>(m);
  double* d_rhob = c.w_rhob.as<double>(m);   // rho by band: prec_rho first, then the update kernel
  double* d_rhoold = c.w_rho.as<double>(m);
  int* d_conv = c.w_conv.as<int>(static_cast<size_t>(m) + 4);
  int* d_nc = d_conv + m;       // active count of the current iteration
  int* d_rec = d_conv + m + 1;  // (active count, perturbations among them)
  const RhoOut rho_out{c.d_g2kin, d_eprec_col, d_rhob};
  const int itmax = maxiter > 0 ? maxiter : c.n_dfpt_maxiter;
  // one record per iteration, read by the host two iterations late
  int* rec = c.h_cgrec.as<int>(2 * static_cast<size_t>(itmax + 1));
  WH_HIP(hipMemsetAsync(d_conv, 0, sizeof(int) * (static_cast<size_t>(m) + 4), s));
  // step 1: g = A x - b, and rho of every band
  stern_apply(x, g, d_e_col, m, c.alpha_pv);
  count_calls(m, nullptr, group, col_id);
  axpby(static_cast<long>(vec), -1.0, b, 1.0, g, s);
  iota(d_l2i, m, s);
  WH_TIMED("cg_prec", prec_rho(m, npw, c.gstart, d_l2i, g, ld, c.d_g2kin, d_eprec_col, d_rhob, s, true));
  // The loop (linsolve_sternheimer_m_wfcts.f90:146-330) runs on the device:
  // each iteration starts with cg_select (convergence test, compaction,
  // beta), and its kernels are launched for a capacity -- the active count
  // the host last read, rounded up to a power of two -- and skip the columns
  // beyond the device's count.  The host never waits for the iteration it
  // launches: it reads the active count of the iteration two back (already
  // finished while the next one is queued), to size the launches and to stop
  // once a count of zero is seen; iterations launched after the last band
  // converged do nothing.
  // Narrow batches (every product a streaming kernel, launch-bound) replay
  // an iteration captured in a HIP graph the second time the same capacity
  // and h/hold assignment come up; WH_GRAPH=0 keeps every iteration eager;
  // so do the profiler's timed mode and a process holding several systems
  // (their threads' device-wide synchronisations may not meet a capture).
  const int nqp = c.l_kinetic_only ? c.nbndval : c.nbndval + c.nkb;
  const char* genv = std::getenv("WH_GRAPH");
  const bool graphs = !(genv && genv[0] == '0') && !c.prof.on && nqp <= 32 && m <= 256 && single_system();
  struct CgGraph {
    int cap;
    const void* h;
    hipGraphExec_t exec;
  };
  std::vector<CgGraph> cg_graphs;
  std::vector<std::pair<int, const void*>> eager_seen;
  hipEvent_t ev[3];
  for (hipEvent_t& e : ev) WH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int last = 0;   // iterations launched
  int cap = m;
  for (int iter = 1; iter <= itmax; ++iter) {
    if (iter >= 3) {
      WH_HIP(hipEventSynchronize(ev[(iter - 2) % 3]));
      const int nlag = rec[2 * (iter - 2)];
      if (nlag == 0) break;
      int p2 = 1;
      while (p2 < nlag) p2 <<= 1;
      cap = std::min(m, p2);
    }
    const int first = iter == 1 ? 1 : 0;
    auto body = [&]() {
      WH_TIMED("cg_select", cg_select(m, tr2, first, group, d_rhob, d_rhoold, d_conv, d_e_col, d_l2i, d_beta, d_eu,
                                      d_nc, d_rec, s));
      WH_TIMED("cg_dir", direction(cap, npw, d_l2i, d_beta, g, c.d_g2kin, d_eprec_col, h, hold, bc, ld, s, d_nc));
      stern_apply(bc, tc, d_eu, cap, c.alpha_pv, nullptr, d_nc);
      WH_TIMED("cg_update", cg_update(cap, npw, c.gstart, d_l2i, h, g, tc, x, ld, nullptr, s, &rho_out, d_nc));
    };
    const CgGraph* cached = nullptr;
    for (const CgGraph& gr : cg_graphs)
      if (gr.cap == cap && gr.h == h) cached = &gr;
    const std::pair<int, const void*> key{cap, h};
    if (cached && !first) {
      WH_HIP(hipGraphLaunch(cached->exec, s));
    } else if (graphs && !first &&
               std::find(eager_seen.begin(), eager_seen.end(), key) != eager_seen.end()) {
      hipGraph_t graph = nullptr;
      WH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      body();
      WH_HIP(hipStreamEndCapture(s, &graph));
      hipGraphExec_t exec = nullptr;
      WH_HIP(hipGraphInstantiateWithFlags(&exec, graph, 0));
      WH_HIP(hipGraphDestroy(graph));
      cg_graphs.push_back({cap, h, exec});
      WH_HIP(hipGraphLaunch(exec, s));
    } else {
      body();
      if (!first) eager_seen.push_back(key);
    }
    WH_HIP(hipMemcpyAsync(rec + 2 * iter, d_rec, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    WH_HIP(hipEventRecord(ev[iter % 3], s));
    last = iter;
    std::swap(h, hold);   // hold <- h  (linsolve 319)
  }
  WH_HIP(hipStreamSynchronize(s));
  for (hipEvent_t e : ev) WH_HIP(hipEventDestroy(e));
  for (const CgGraph& gr : cg_graphs) WH_HIP(hipGraphExecDestroy(gr.exec));
  // counters: an iteration counts when it had active bands
  for (int it = 1; it <= last; ++it) {
    const int n = rec[2 * it];
    if (n > 0) {
      ++c.cg_iters;
      c.nstern += n;
      c.nstern_calls += rec[2 * it + 1];
    }
  }
  // ierr: the bands still active when the iterations ran out (linsolve 332)
  std::vector<int> conv(m);
  WH_HIP(hipMemcpy(conv.data(), d_conv, sizeof(int) * m, hipMemcpyDeviceToHost));
  int nc = 0;
  if (ierr_col) ierr_col->assign(m, 0);
  for (int ib = 0; ib < m; ++ib)
    if (!conv[ib]) {
      ++nc;
      if (ierr_col) (*ierr_col)[ib] = 1;
    }
  return nc;
}

// ---------------------------------------------------------------------------
// Mixed precision: the CG iterates in FP32 (FFTs, vectors, split-BF16
// projector GEMMs with FP32 accumulation) on an increment dx of the FP64
// iterate x; the residual b - A x is recomputed with the FP64 operator
// whenever the FP32 residual has dropped by refine_delta (in rho) since the
// last recomputation, or below tr2. Convergence is decided on that FP64
// residual only (rho_64 < tr2), so a converged band meets the FP64 criterion
// of linsolve_sternheimer_m_wfcts.f90:190-205.
// ---------------------------------------------------------------------------
static long round_up(long a, long b) { return (a + b - 1) / b * b; }

static void prepare_mixed() {
  Context& c = ctx();
  if (c.mixed_valid) return;
  WH_REQUIRE(c.plan.tpl32 != nullptr,
             "mixed precision needs a cubic FFT grid of a compiled size (18, 20, 24, 30, 32, 36, 40, 45, 48, 50, 54, 60, "
             "64, 72, 75, 80, 90, 96, 100, 108)");
  hipStream_t s = c.stream;
  c.ld32 = round_up(c.npw, 16);
  const long lda = 2 * c.ld32;
  float* v32 = c.m_vrs32.as<float>(c.nnr);
  if (c.have_pot) to_f32(1, c.nnr, c.d_vrs, c.nnr, v32, c.nnr, s);
  const int nq = c.q_rows;
  c.nq_pad = round_up(std::max(nq, 1), 32);
  if (nq > 0) {
    float* q32 = c.m_q32.as<float>(static_cast<size_t>(nq) * lda);
    WH_HIP(hipMemsetAsync(q32, 0, sizeof(float) * nq * lda, s));
    to_f32(nq, 2 * c.npw, reinterpret_cast<const double*>(c.d_q), 2L * c.npw, q32, lda, s);
    const size_t qb = sizeof(uint16_t) * nq * lda;
    void* const qp[3] = {c.m_qp[0].get(qb), c.m_qp[1].get(qb), c.m_qp[2].get(qb)};
    split_bf16(nq, 2 * c.npw, q32, lda, qp, lda, s);
    const size_t tb = sizeof(uint16_t) * 2 * c.npw * c.nq_pad;
    void* const qtp[3] = {c.m_qtp[0].get(tb), c.m_qtp[1].get(tb), c.m_qtp[2].get(tb)};
    split_bf16_t(nq, 2 * c.npw, q32, lda, qtp, c.nq_pad, s);
  }
  WH_HIP(hipStreamSynchronize(s));
  c.mixed_valid = true;
}

// t = (H - e + alpha P_v) x in FP32; x, t rows of ld32 complex
void stern_apply32(const z32* x, z32* t, const double* d_e, int m, double alpha) {
  Context& c = ctx();
  if (m <= 0) return;
  hipStream_t s = c.stream;
  const int npw = c.npw;
  const long l32 = c.ld32;
  if (c.l_kinetic_only) {
    WH_TIMED("kinetic", kinetic_minus_e(m, npw, x, t, l32, c.d_g2kin, d_e, s));
  } else {
    const TplKernels32* K = c.plan.tpl32;
    const int npair = (m + 1) / 2;
    const int cap = c.items_capacity();
    FftItem* items = c.w_items.as<FftItem>(npair);
    fill_pairs(items, m, 0, s);
    FftItem* ident = c.w_items_id.as<FftItem>(std::min(cap, npair));
    fill_ident(ident, std::min(cap, npair), s);
    const int chunk = std::min(cap, npair);
    z32* sa = c.w_sticks_a.as<z32>(static_cast<size_t>(chunk) * c.plan.sticks_elems());
    z32* sb = c.w_sticks_b.as<z32>(static_cast<size_t>(chunk) * c.plan.sticks_elems());
    const float* v32 = static_cast<const float*>(c.m_vrs32.p);
    ZfwdArgs za{ZFWD_PAIR_STERN, 1.0, c.d_g2kin, d_e, nullptr, 0};
    // register-line z kernels where the grid has them (WH_ZREG=0: the LDS-pass ones)
    static const bool zreg_env = [] {
      const char* e = std::getenv("WH_ZREG");
      return !(e && e[0] == '0');
    }();
    const bool zreg = zreg_env && zreg_available(c.plan.dev);
    for (int q0 = 0; q0 < npair; q0 += chunk) {
      const int nq = std::min(chunk, npair - q0);
      if (zreg)
        WH_TIMED("zinv32", zinv_reg(c.plan.dev, items + q0, nq, x, l32, x, l32, sa, ZINV_PAIR, s));
      else
        WH_TIMED("zinv32", K->zinv(c.plan.dev, items + q0, nq, x, l32, x, l32, sa, ZINV_PAIR, s));
      // variants 1..max_occ: LDS-pass kernel sized for that many workgroups
      // per CU; above: register-line kernels.  Untimed small launches keep
      // the LDS-pass kernel (the 6-H2 Davidson, mixed: 277 ms against 284 ms
      // with the first register-line variant)
      const int nvar = K->max_occ + plane_reg_variants(c.plan.dev, true);
      WH_TIMED("plane_vloc32", plane_occ_launch(
                                   c.plan.vloc32_occ, nvar, nq,
                                   [&](int v) {
                                     if (v <= K->max_occ)
                                       K->vloc(c.plan.dev, ident, nq, sa, sb, v32, 0, 1.0 / c.nnr, v, s);
                                     else
                                       plane_reg_vloc(v - K->max_occ - 1, c.plan.dev, ident, nq, sa, sb, v32, 0,
                                                      1.0 / c.nnr, s);
                                   },
                                   s));
      if (zreg)
        WH_TIMED("zfwd32", zfwd_reg(c.plan.dev, items + q0, nq, sb, x, l32, t, l32, za, s));
      else
        WH_TIMED("zfwd32", K->zfwd(c.plan.dev, items + q0, nq, sb, x, l32, t, l32, za, s));
      WH_HIP(hipGetLastError());
    }
  }
  const int nq = c.l_kinetic_only ? c.nbndval : c.nbndval + c.nkb;
  if (nq > 0) {
    set_qscale_alpha(alpha);
    const int K2 = 2 * npw;
    const long lda = 2 * l32;
    const int terms = c.precision == 2 ? 6 : 3;
    const size_t wf = gemm_bra_bf16_work(nq, m, K2);
    float* work = c.m_work.as<float>(wf);
    float* cm = c.m_cm.as<float>(static_cast<size_t>(c.nq_pad) * m);
    const void* const qp[3] = {c.m_qp[0].p, c.m_qp[1].p, c.m_qp[2].p};
    const void* const qtp[3] = {c.m_qtp[0].p, c.m_qtp[1].p, c.m_qtp[2].p};
    WH_TIMED("gemm_bra32", gemm_bra_bf16(terms, nq, m, K2, qp, lda, static_cast<const float*>(c.m_q32.p),
                                         reinterpret_cast<const float*>(x), lda, c.gstart == 2 ? 1.f : 0.f,
                                         c.d_qscale, cm, c.nq_pad, work, wf, s));
    WH_TIMED("gemm_ket32", gemm_tn_bf16(terms, K2, m, nq, 1.f, qtp, c.nq_pad, cm, c.nq_pad, 1.f,
                                        reinterpret_cast<float*>(t), 1L, lda, s));
  }
  c.nstern += m;
  c.nstern32 += m;
}

int linsolve_mixed(int m, z64* b, z64* x, const std::vector<double>& e_col, const double* d_e_col,
                   const double* d_eprec_col, double tr2, std::vector<int>* ierr_col, int group) {
  Context& c = ctx();
  prepare_mixed();
  hipStream_t s = c.stream;
  const int npw = c.npw;
  const long ld = npw;
  const long l32 = c.ld32;
  const size_t v32 = static_cast<size_t>(m) * l32;
  z32* g = c.m_g.as<z32>(v32);
  z32* h = c.m_h.as<z32>(v32);
  z32* hold = c.m_hold.as<z32>(v32);
  z32* dx = c.m_dx.as<z32>(v32);
  z32* bc = c.m_bc.as<z32>(v32);
  z32* tc = c.m_tc.as<z32>(v32);
  z64* xc = c.w_bc.as<z64>(static_cast<size_t>(m) * npw);
  z64* tc64 = c.w_tc.as<z64>(static_cast<size_t>(m) * npw);
  int* d_u2i = c.m_u2i.as<int>(m);
  int* d_l2i0 = c.w_l2i.as<int>(m);   // the first residual's list
  // each iteration's active list, beta and e go to the device in one copy,
  // laid out for that iteration's nc: [beta (nc) | e (nc) | list (nc ints)]
  double* d_pk = c.w_cgpk.as<double>(static_cast<size_t>(5) * m / 2 + 1);
  double* pk = c.h_cgpk.as<double>(static_cast<size_t>(5) * m / 2 + 1);
  double* d_rho = c.w_rho.as<double>(m);
  double* d_rho64 = c.m_rho64.as<double>(m);
  double* d_eu64 = c.m_eu64.as<double>(m);
  // rho of each band's FP32 g after an update, by band (fused into the update kernel)
  double* d_rhob = c.w_rhob.as<double>(m);
  const RhoOut rho_out{c.d_g2kin, d_eprec_col, d_rhob};
  WH_HIP(hipMemsetAsync(dx, 0, sizeof(z32) * v32, s));

  std::vector<char> is_conv(m, 0), to_fp64(m, 0);
  std::vector<int> list, stall(m, 0), since(m, 0);
  std::vector<double> rho(m, 0.0), rhoold(m, 0.0), rho_ref(m, 0.0);
  // pinned staging of the per-iteration copies
  int* l2i = c.hm_l2i.as<int>(m);
  int* uh = c.hm_u.as<int>(m);
  double* tmp = c.hm_rho.as<double>(m);
  double* eu64 = c.hm_eu64.as<double>(m);
  list.reserve(m);
  // FP32 iterations after which a refinement is forced
  const int kforce = 40;

  // x += dx; r = A x - b in FP64; g = r in FP32; rho_ref = <M^-1 r|r>; converged if rho_ref < tr2.
  // A band whose FP64 residual twice drops less than tenfold between
  // refinements (the FP32 operator is not accurate enough for its
  // conditioning) is finished by the FP64 CG instead.
  auto refine = [&](const std::vector<int>& u, bool first) {
    const int nu = static_cast<int>(u.size());
    for (int k = 0; k < nu; ++k) {
      eu64[k] = e_col[u[k]];
      uh[k] = u[k];
    }
    WH_HIP(hipMemcpyAsync(d_u2i, uh, sizeof(int) * nu, hipMemcpyHostToDevice, s));
    WH_HIP(hipMemcpyAsync(d_eu64, eu64, sizeof(double) * nu, hipMemcpyHostToDevice, s));
    WH_TIMED("refine_fold", fold_increment(nu, npw, d_u2i, x, ld, dx, l32, xc, s));
    stern_apply(xc, tc64, d_eu64, nu, c.alpha_pv);
    count_calls(nu, u.data(), group);
    WH_TIMED("refine_residual", true_residual(nu, npw, c.gstart, d_u2i, tc64, b, ld, c.d_g2kin, d_eprec_col, g, l32,
                                              d_rho64, s));
    WH_HIP(hipMemcpyAsync(tmp, d_rho64, sizeof(double) * nu, hipMemcpyDeviceToHost, s));
    WH_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < nu; ++k) {
      const int ib = u[k];
      const double prev = rho_ref[ib];
      rho_ref[ib] = tmp[k];
      since[ib] = 0;
      if (tmp[k] < tr2) {
        is_conv[ib] = 1;
      } else if (!first && prev > 100.0 * tr2 && tmp[k] > 0.1 * prev) {
        if (++stall[ib] >= 2) to_fp64[ib] = 1;
      }
    }
    c.nrefine += nu;
  };
  // h = M^-1 g and rho for the bands in idx (device list d_idx)
  auto precond = [&](const std::vector<int>& idx, int* d_idx) {
    const int n = static_cast<int>(idx.size());
    std::copy(idx.begin(), idx.end(), uh);
    WH_HIP(hipMemcpyAsync(d_idx, uh, sizeof(int) * n, hipMemcpyHostToDevice, s));
    WH_TIMED("cg_prec32", prec_rho(n, npw, c.gstart, d_idx, g, l32, c.d_g2kin, d_eprec_col, d_rho, s));
    WH_HIP(hipMemcpyAsync(tmp, d_rho, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    WH_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < n; ++k) rho[idx[k]] = tmp[k];
  };

  // initial residual with the FP32 operator, g = A x0 - b; the first FP64
  // residual is formed at the first refinement
  list.resize(m);
  std::iota(list.begin(), list.end(), 0);
  to_f32(m, 2 * npw, reinterpret_cast<const double*>(x), 2L * npw, reinterpret_cast<float*>(bc), 2 * l32, s);
  stern_apply32(bc, tc, d_e_col, m, c.alpha_pv);
  count_calls(m, nullptr, group);
  WH_TIMED("refine_residual", residual32(m, npw, tc, l32, b, ld, g, s));
  precond(list, d_l2i0);
  for (int ib = 0; ib < m; ++ib) rho_ref[ib] = rho[ib];
  {
    // an FP32 residual already below tr2 is confirmed in FP64 before the band is let go
    std::vector<int> small;
    for (int ib = 0; ib < m; ++ib)
      if (rho[ib] < tr2) small.push_back(ib);
    if (!small.empty()) {
      refine(small, true);
      std::vector<int> again;
      for (int ib : small)
        if (!is_conv[ib]) again.push_back(ib);
      if (!again.empty()) precond(again, d_u2i);
    }
  }
  std::vector<int> act;
  for (int ib = 0; ib < m; ++ib)
    if (!is_conv[ib]) act.push_back(ib);
  for (int iter = 1; iter <= c.n_dfpt_maxiter && !act.empty(); ++iter) {
    if (iter > 1) {
      // every band in act was updated in the previous iteration
      WH_HIP(hipMemcpyAsync(tmp, d_rhob, sizeof(double) * m, hipMemcpyDeviceToHost, s));
      WH_HIP(hipStreamSynchronize(s));
      for (int ib : act) rho[ib] = tmp[ib];
    }
    ++c.cg_iters;
    list.clear();
    for (int ib : act)
      if (rho[ib] < tr2 || rho[ib] < c.refine_delta * rho_ref[ib] || ++since[ib] >= kforce) list.push_back(ib);
    if (!list.empty()) {
      refine(list, false);
      std::vector<int> again;
      for (int ib : list)
        if (!is_conv[ib] && !to_fp64[ib]) again.push_back(ib);
      if (!again.empty()) precond(again, d_u2i);
    }
    int nc = 0;
    for (int ib : act)
      if (!is_conv[ib] && !to_fp64[ib]) l2i[nc++] = ib;
    act.assign(l2i, l2i + nc);
    if (nc == 0) break;
    // the last few slow bands (their iterations are latency-bound, and the
    // FP64 CG restarted from the current FP64 iterate needs fewer of them),
    // and every band still unconverged after half of n_dfpt_maxiter, go to
    // the FP64 CG
    if ((m >= 256 && nc <= m / 128 && iter >= 10) || iter >= std::max(1, c.n_dfpt_maxiter / 2)) {
      refine(act, true);
      for (int ib : act)
        if (!is_conv[ib]) to_fp64[ib] = 1;
      act.clear();
      break;
    }
    int* pl = reinterpret_cast<int*>(pk + 2 * nc);
    for (int lb = 0; lb < nc; ++lb) {
      const int ib = l2i[lb];
      pk[lb] = (iter != 1) ? rho[ib] / rhoold[ib] : 0.0;   // beta
      pk[nc + lb] = e_col[ib];
      pl[lb] = ib;
    }
    WH_HIP(hipMemcpyAsync(d_pk, pk, (2 * sizeof(double) + sizeof(int)) * nc, hipMemcpyHostToDevice, s));
    const double* d_beta = d_pk;
    const double* d_eu = d_pk + nc;
    const int* d_l2i = reinterpret_cast<const int*>(d_pk + 2 * nc);
    WH_TIMED("cg_dir32", direction(nc, npw, d_l2i, d_beta, g, c.d_g2kin, d_eprec_col, h, hold, bc, l32, s));
    stern_apply32(bc, tc, d_eu, nc, c.alpha_pv);
    count_calls(nc, l2i, group);
    WH_TIMED("cg_update32", cg_update(nc, npw, c.gstart, d_l2i, h, g, tc, dx, l32, nullptr, s, &rho_out));
    std::swap(h, hold);
    for (int lb = 0; lb < nc; ++lb) rhoold[l2i[lb]] = rho[l2i[lb]];
    // the host lists sent above are next modified after the synchronisation
    // that follows the next iteration's rho read, which orders those copies
  }
  WH_HIP(hipStreamSynchronize(s));
  // bands stopped by n_dfpt_maxiter keep their last FP32 increment
  if (!act.empty()) {
    WH_HIP(hipMemcpyAsync(d_u2i, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, s));
    fold_increment(static_cast<int>(act.size()), npw, d_u2i, x, ld, dx, l32, xc, s);
    WH_HIP(hipStreamSynchronize(s));
  }
  if (ierr_col) ierr_col->assign(m, 0);
  for (int ib : act)
    if (ierr_col) (*ierr_col)[ib] = 1;
  int nonconv = static_cast<int>(act.size());
  // bands handed to the FP64 CG, restarted from their FP64 iterate
  std::vector<int> fl;
  for (int ib = 0; ib < m; ++ib)
    if (to_fp64[ib] && !is_conv[ib]) fl.push_back(ib);
  if (!fl.empty()) {
    const int nf = static_cast<int>(fl.size());
    z64* fb = c.m_fb.as<z64>(static_cast<size_t>(nf) * npw);
    z64* fx = c.m_fx.as<z64>(static_cast<size_t>(nf) * npw);
    double* fe = c.m_fe.as<double>(nf);
    double* fep = c.m_fep.as<double>(nf);
    std::vector<double> e_f(nf);
    for (int k = 0; k < nf; ++k) e_f[k] = e_col[fl[k]];
    WH_HIP(hipMemcpyAsync(d_u2i, fl.data(), sizeof(int) * nf, hipMemcpyHostToDevice, s));
    WH_HIP(hipMemcpyAsync(fe, e_f.data(), sizeof(double) * nf, hipMemcpyHostToDevice, s));
    copy_rows(nf, npw, d_u2i, b, ld, fb, ld, false, s);
    copy_rows(nf, npw, d_u2i, x, ld, fx, ld, false, s);
    gather_real(nf, d_u2i, d_eprec_col, fep, s);
    WH_HIP(hipStreamSynchronize(s));
    c.nfallback += nf;
    std::vector<int> ierr_f;
    // a full n_dfpt_maxiter budget: these are the hardest bands, restarted
    nonconv += linsolve_dev(nf, fb, fx, e_f, fe, fep, tr2, &ierr_f, c.n_dfpt_maxiter, group, fl.data());
    WH_HIP(hipMemcpyAsync(d_u2i, fl.data(), sizeof(int) * nf, hipMemcpyHostToDevice, s));
    copy_rows(nf, npw, d_u2i, fx, ld, x, ld, true, s);
    WH_HIP(hipStreamSynchronize(s));
    if (ierr_col)
      for (int k = 0; k < nf; ++k) (*ierr_col)[fl[k]] = ierr_f[k];
  }
  return nonconv;
}

// ---------------------------------------------------------------------------
// chi0 at Gamma  (DFPT_kernel/dfpt_module.f90:139-503, Gamma branches)
//   dvg: m perturbations (rows, stick order); dng: m rows out.
//   out_weight: per-G factor applied to dng (sqvc for apply_operator) or null.
// ---------------------------------------------------------------------------
static int dfpt_chunk(int m, const z64* dvg, z64* dng, double tr2, const double* out_weight,
                      std::vector<int>* ierr_pert);

// Perturbations are processed in chunks whose Sternheimer work vectors fit
// the vector budget (seven npw-vectors per (perturbation, band) pair).
int dfpt_dev(int m, const z64* dvg, z64* dng, double tr2, const double* out_weight, std::vector<int>* ierr_pert) {
  Context& c = ctx();
  require_ready(true);
  if (ierr_pert) ierr_pert->assign(std::max(m, 0), 0);
  if (m <= 0) return 0;
  const double pair_bytes = 7.0 * 16.0 * c.npw;
  const long max_pairs = std::max<long>(1, static_cast<long>(static_cast<double>(c.vectors_budget) / pair_bytes));
  const int chunk = static_cast<int>(std::max<long>(1, std::min<long>(m, max_pairs / std::max(c.nbndval, 1))));
  int nonconv = 0;
  const long ld = c.npw;
  for (int p0 = 0; p0 < m; p0 += chunk) {
    const int n = std::min(chunk, m - p0);
    std::vector<int> ie;
    nonconv += dfpt_chunk(n, dvg + p0 * ld, dng + p0 * ld, tr2, out_weight, ierr_pert ? &ie : nullptr);
    if (ierr_pert)
      for (int k = 0; k < n; ++k) (*ierr_pert)[p0 + k] = ie[k];
  }
  return nonconv;
}

static int dfpt_chunk(int m, const z64* dvg, z64* dng, double tr2, const double* out_weight,
                      std::vector<int>* ierr_pert) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  const int npw = c.npw;
  const long ld = npw;
  const int nv = c.nbndval;
  const int nevp = (nv + 1) / 2;
  const long sl = c.plan.sticks_elems();
  const int cap = c.items_capacity();
  if (ierr_pert) ierr_pert->assign(m, 0);
  if (m <= 0) return 0;

  // occupied-pair sticks, once per ground state
  if (!c.evc_sticks_valid) {
    z64* es = c.evc_sticks.as<z64>(static_cast<size_t>(nevp) * sl);
    FftItem* it = c.w_items2.as<FftItem>(nevp);
    fill_pairs(it, nv, 0, s);
    WH_TIMED("zinv", launch_zinv(c.plan, it, nevp, c.d_q, ld, es, ZINV_PAIR, s));
    c.evc_sticks_valid = true;
  }
  // 1. dv(r) of every perturbation (single_invfft_gamma, dfpt_module.f90:229-235)
  double* dvr = c.w_dvr.as<double>(static_cast<size_t>(m) * c.nnr);
  {
    const int npp = (m + 1) / 2;
    FftItem* it = c.w_items.as<FftItem>(npp);
    fill_pairs(it, m, 0, s);
    const int chunk = std::min(cap, npp);
    z64* sa = c.w_sticks_a.as<z64>(static_cast<size_t>(chunk) * sl);
    for (int q0 = 0; q0 < npp; q0 += chunk) {
      const int nq = std::min(chunk, npp - q0);
      WH_TIMED("zinv", launch_zinv(c.plan, it + q0, nq, dvg, ld, sa, ZINV_PAIR_REALG0, s));
      WH_TIMED("plane_to_real", launch_plane_to_real(c.plan, it + q0, nq, sa, dvr, c.nnr, s));
    }
  }
  // 2. right-hand side dvpsi = P[psi_v dv]  (dfpt_module.f90:245-278)
  const long M = static_cast<long>(m) * nv;
  z64* dvpsi = c.w_b.as<z64>(static_cast<size_t>(M) * npw);
  z64* dpsi = c.w_x.as<z64>(static_cast<size_t>(M) * npw);
  {
    const long nitem = static_cast<long>(m) * nevp;
    std::vector<FftItem> hp(nitem), hz(nitem);
    for (int p = 0; p < m; ++p)
      for (int q = 0; q < nevp; ++q) {
        const long i = static_cast<long>(p) * nevp + q;
        hp[i] = FftItem{q, -1, p, 0};
        hz[i] = FftItem{p * nv + 2 * q, (2 * q + 1 < nv) ? p * nv + 2 * q + 1 : -1, -1, 0};
      }
    FftItem* dp = c.w_items.as<FftItem>(2 * nitem);
    FftItem* dz = dp + nitem;
    WH_HIP(hipMemcpyAsync(dp, hp.data(), sizeof(FftItem) * nitem, hipMemcpyHostToDevice, s));
    WH_HIP(hipMemcpyAsync(dz, hz.data(), sizeof(FftItem) * nitem, hipMemcpyHostToDevice, s));
    const int chunk = static_cast<int>(std::min<long>(cap, nitem));
    z64* sb = c.w_sticks_b.as<z64>(static_cast<size_t>(chunk) * sl);
    ZfwdArgs za{ZFWD_PAIR_STORE, 1.0, c.d_g2kin, nullptr, nullptr, 0};
    for (long i0 = 0; i0 < nitem; i0 += chunk) {
      const int ni = static_cast<int>(std::min<long>(chunk, nitem - i0));
      WH_TIMED("plane_vloc", launch_plane_vloc(c.plan, dp + i0, ni, static_cast<const z64*>(c.evc_sticks.p), sb,
                                               dvr, c.nnr, 1.0 / c.nnr, s));
      WH_TIMED("zfwd", launch_zfwd(c.plan, dz + i0, ni, sb, nullptr, 0, dvpsi, ld, za, s));
    }
    WH_HIP(hipStreamSynchronize(s));   // host item vectors go out of scope
  }
  // 3. dvpsi <- -P_c dvpsi = -dvpsi + evc <evc|dvpsi>  (apply_alpha_pc, alpha = -1)
  if (nv > 0) {
    double* cm = c.w_cmat.as<double>(static_cast<size_t>(nv) * M);
    glbrak_q(0, nv, dvpsi, static_cast<int>(M), cm);
    const long k2 = 2L * npw;
    WH_TIMED("gemm_ket", dgemm(false, static_cast<int>(k2), static_cast<int>(M), nv, 1.0,
                               reinterpret_cast<const double*>(c.d_q), k2, cm, nv, -1.0,
                               reinterpret_cast<double*>(dvpsi), k2, s));
  }
  // 4. dpsi = M^-1 dvpsi (initial guess, dfpt_module.f90:348)
  WH_TIMED("precondition", precondition(static_cast<int>(M), npw, dvpsi, dpsi, ld, c.d_g2kin, c.d_eprec, nv, s));
  // 5. Sternheimer solve (dfpt_module.f90:350-361)
  int nonconv = 0;
  if (tr2 >= 0.0) {
    std::vector<double> e_col(M), ep_col(M);
    for (long i = 0; i < M; ++i) {
      e_col[i] = c.et[i % nv];
      ep_col[i] = c.eprec[i % nv];
    }
    double* de = c.w_ecol.as<double>(M);
    double* dep = c.w_epcol.as<double>(M);
    WH_HIP(hipMemcpyAsync(de, e_col.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
    WH_HIP(hipMemcpyAsync(dep, ep_col.data(), sizeof(double) * M, hipMemcpyHostToDevice, s));
    std::vector<int> ierr_col;
    nonconv = c.precision != 0
                  ? linsolve_mixed(static_cast<int>(M), dvpsi, dpsi, e_col, de, dep, tr2, &ierr_col, nv)
                  : linsolve_dev(static_cast<int>(M), dvpsi, dpsi, e_col, de, dep, tr2, &ierr_col, 0, nv);
    if (ierr_pert)
      for (long i = 0; i < M; ++i) (*ierr_pert)[i / nv] += ierr_col[i];
  }
  // 6. response density and dng (dfpt_module.f90:403-503)
  {
    const int pc = std::max(1, std::min(m, (2 * cap) / std::max(nv, 1)));
    z64* sa = c.w_sticks_a.as<z64>(static_cast<size_t>(pc) * nv * sl);
    z64* sb = c.w_sticks_b.as<z64>(static_cast<size_t>(pc) * sl);
    std::vector<FftItem> hi(static_cast<size_t>(pc) * nv), hz(pc);
    FftItem* di = c.w_items.as<FftItem>(static_cast<size_t>(pc) * nv + pc);
    FftItem* dz = di + static_cast<size_t>(pc) * nv;
    // fwfft normalisation 1/N times the 2 w_k / omega of dfpt_module.f90:486
    ZfwdArgs za{ZFWD_SINGLE, 2.0 * c.weight / (c.omega * c.nnr), c.d_g2kin, nullptr, out_weight,
                c.gstart == 2 ? 1 : 0};
    for (int p0 = 0; p0 < m; p0 += pc) {
      const int np_ = std::min(pc, m - p0);
      for (int p = 0; p < np_; ++p) {
        for (int v = 0; v < nv; ++v) hi[static_cast<size_t>(p) * nv + v] = FftItem{v, (p0 + p) * nv + v, -1, 0};
        hz[p] = FftItem{p0 + p, -1, -1, 0};
      }
      WH_HIP(hipMemcpyAsync(di, hi.data(), sizeof(FftItem) * np_ * nv, hipMemcpyHostToDevice, s));
      WH_HIP(hipMemcpyAsync(dz, hz.data(), sizeof(FftItem) * np_, hipMemcpyHostToDevice, s));
      WH_TIMED("zinv", launch_zinv(c.plan, di, np_ * nv, c.d_q, ld, dpsi, ld, sa, ZINV_PAIR, s));
      const long dsc = plane_density_scratch(c.plan);
      double* dscr = dsc > 0 ? c.w_dscr.as<double>(static_cast<size_t>(dsc) * np_) : nullptr;
      WH_TIMED("plane_density", launch_plane_density(c.plan, np_, nv, sa, c.d_occ, sb, dscr, s));
      WH_TIMED("zfwd", launch_zfwd(c.plan, dz, np_, sb, nullptr, 0, dng, ld, za, s));
      WH_HIP(hipStreamSynchronize(s));
    }
  }
  return nonconv;
}

}  // namespace wh

// ===========================================================================
// extern "C" entry points
// ===========================================================================
using namespace wh;

#define WH_API_BEGIN try {
#define WH_API_END                                                  \
  }                                                                 \
  catch (const ::wh::Error& e) {                                    \
    ::wh::set_error(e.msg);                                         \
    return e.code < 0 ? e.code : -1;                                \
  }                                                                 \
  catch (const std::exception& e) {                                 \
    ::wh::set_error(e.what());                                      \
    return -3;                                                      \
  }

static void finish_call() {
  Context& c = ctx();
  WH_HIP(hipStreamSynchronize(c.stream));
  if (c.prof.on) c.prof.flush();
}

extern "C" {

const char* wh_last_error(void) { return last_error(); }

int wh_init(int device) {
  WH_API_BEGIN
  int n = 0;
  WH_HIP(hipGetDeviceCount(&n));
  WH_REQUIRE(device >= 0 && device < n, "device index out of range");
  std::lock_guard<std::mutex> lk(g_sys_mutex);
  if (!g_sys.empty()) {
    WH_REQUIRE(g_device == device, "the library is already initialised on another device");
    return 0;
  }
  g_sys.push_back(new_context(device));
  g_device = device;
  noalloc_init(device);
  return 0;
  WH_API_END
}

int wh_finalize(void) {
  WH_API_BEGIN
  std::lock_guard<std::mutex> lk(g_sys_mutex);
  for (Context* c : g_sys)
    if (c) free_context(c);
  g_sys.clear();
  g_device = -1;
  t_current = 0;
  return 0;
  WH_API_END
}

int wh_system_create(int* handle) {
  WH_API_BEGIN
  WH_REQUIRE(handle != nullptr, "handle is null");
  std::lock_guard<std::mutex> lk(g_sys_mutex);
  WH_REQUIRE(!g_sys.empty(), "wh_init has not been called");
  g_sys.push_back(new_context(g_device));
  *handle = static_cast<int>(g_sys.size()) - 1;
  return 0;
  WH_API_END
}

int wh_system_select(int handle) {
  WH_API_BEGIN
  std::lock_guard<std::mutex> lk(g_sys_mutex);
  WH_REQUIRE(handle >= 0 && handle < static_cast<int>(g_sys.size()) && g_sys[handle],
             "no such system");
  t_current = handle;
  WH_HIP(hipSetDevice(g_device));
  return 0;
  WH_API_END
}

int wh_system_destroy(int handle) {
  WH_API_BEGIN
  std::lock_guard<std::mutex> lk(g_sys_mutex);
  WH_REQUIRE(handle > 0 && handle < static_cast<int>(g_sys.size()) && g_sys[handle],
             "no such system (system 0 ends with wh_finalize)");
  free_context(g_sys[handle]);
  g_sys[handle] = nullptr;
  return 0;
  WH_API_END
}

int wh_set_grid(int nr1, int nr2, int nr3, int npw, int npwx, int gstart, const int* nl, const int* nlm,
                double omega, double tpiba2, const double* g2kin) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(nr1 > 0 && nr2 > 0 && nr3 > 0, "FFT dimensions must be positive");
  WH_REQUIRE(npw > 0 && npwx >= npw, "need 0 < npw <= npwx");
  WH_REQUIRE(gstart == 2, "Gamma point: gstart must be 2 (G = 0 is the first plane wave)");
  WH_REQUIRE(nl && nlm && g2kin, "nl, nlm and g2kin are required");
  std::vector<int> hnl(npw), hnlm(npw);
  std::vector<double> hg2(npw);
  WH_HIP(hipMemcpy(hnl.data(), nl, sizeof(int) * npw, hipMemcpyDefault));
  WH_HIP(hipMemcpy(hnlm.data(), nlm, sizeof(int) * npw, hipMemcpyDefault));
  WH_HIP(hipMemcpy(hg2.data(), g2kin, sizeof(double) * npw, hipMemcpyDefault));
  c.plan.build(nr1, nr2, nr3, npw, hnl.data(), hnlm.data());
  WH_REQUIRE(c.plan.perm[0] == 0, "G = 0 must be the first plane wave (gstart = 2)");
  c.nr1 = nr1; c.nr2 = nr2; c.nr3 = nr3;
  c.nnr = nr1 * nr2 * nr3;
  c.npw = npw; c.npwx = npwx; c.gstart = gstart;
  c.omega = omega; c.tpiba2 = tpiba2;
  c.g2kin_s.resize(npw);
  for (int s = 0; s < npw; ++s) c.g2kin_s[s] = hg2[c.plan.perm[s]];
  if (c.d_perm) WH_HIP(hipFree(c.d_perm));
  if (c.d_g2kin) WH_HIP(hipFree(c.d_g2kin));
  WH_HIP(hipMalloc(&c.d_perm, sizeof(int) * npw));
  WH_HIP(hipMalloc(&c.d_g2kin, sizeof(double) * npw));
  WH_HIP(hipMemcpy(c.d_perm, c.plan.perm.data(), sizeof(int) * npw, hipMemcpyHostToDevice));
  WH_HIP(hipMemcpy(c.d_g2kin, c.g2kin_s.data(), sizeof(double) * npw, hipMemcpyHostToDevice));
  c.have_grid = true;
  // a Gamma grid replaces a k grid registered before in this system
  free_kpath(c.kpath);
  c.kpath = nullptr;
  c.mixed_valid = false;
  c.have_pot = false;
  c.have_states = false;
  c.have_sqvc = false;
  c.nkb = 0;
  c.nbndval = 0;
  c.evc_sticks_valid = false;
  return 0;
  WH_API_END
}

int wh_set_potential(const double* vrs) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(c.have_grid, "wh_set_grid first");
  if (c.d_vrs) WH_HIP(hipFree(c.d_vrs));
  WH_HIP(hipMalloc(&c.d_vrs, sizeof(double) * c.nnr));
  WH_HIP(hipMemcpy(c.d_vrs, vrs, sizeof(double) * c.nnr, hipMemcpyDefault));
  c.have_pot = true;
  c.mixed_valid = false;
  return 0;
  WH_API_END
}

static std::vector<z64> to_sticks(const double* a, int ncol, int npwx) {
  Context& c = ctx();
  std::vector<z64> host(static_cast<size_t>(ncol) * npwx);
  if (ncol) WH_HIP(hipMemcpy(host.data(), a, sizeof(z64) * host.size(), hipMemcpyDefault));
  std::vector<z64> out(static_cast<size_t>(ncol) * c.npw);
  for (int j = 0; j < ncol; ++j)
    for (int s = 0; s < c.npw; ++s) out[static_cast<size_t>(j) * c.npw + s] = host[static_cast<size_t>(j) * npwx + c.plan.perm[s]];
  return out;
}

int wh_set_nonlocal(int nkb, const double* vkb, const double* dvan) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(c.have_grid, "wh_set_grid first");
  WH_REQUIRE(nkb >= 0, "nkb must be >= 0");
  c.nkb = nkb;
  c.vkb_s = to_sticks(vkb, nkb, c.npwx);
  // a real projector has a real G = 0 coefficient; the Gamma algebra relies on it
  for (int j = 0; j < nkb; ++j) c.vkb_s[static_cast<size_t>(j) * c.npw].y = 0.0;
  c.dvan.assign(nkb, 0.0);
  if (nkb) WH_HIP(hipMemcpy(c.dvan.data(), dvan, sizeof(double) * nkb, hipMemcpyDefault));
  rebuild_q();
  return 0;
  WH_API_END
}

int wh_set_states(int nbndval, const double* evc, const double* et, const double* occ, double weight,
                  double alpha_pv) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(c.have_grid, "wh_set_grid first");
  WH_REQUIRE(nbndval > 0, "nbndval must be positive");
  c.nbndval = nbndval;
  c.evc_s = to_sticks(evc, nbndval, c.npwx);
  for (int j = 0; j < nbndval; ++j) c.evc_s[static_cast<size_t>(j) * c.npw].y = 0.0;
  c.et.assign(nbndval, 0.0);
  c.occ.assign(nbndval, 1.0);
  WH_HIP(hipMemcpy(c.et.data(), et, sizeof(double) * nbndval, hipMemcpyDefault));
  if (occ) WH_HIP(hipMemcpy(c.occ.data(), occ, sizeof(double) * nbndval, hipMemcpyDefault));
  c.weight = weight;
  c.alpha_pv = alpha_pv;
  rebuild_q();
  if (c.d_occ) WH_HIP(hipFree(c.d_occ));
  if (c.d_eprec) WH_HIP(hipFree(c.d_eprec));
  WH_HIP(hipMalloc(&c.d_occ, sizeof(double) * nbndval));
  WH_HIP(hipMalloc(&c.d_eprec, sizeof(double) * nbndval));
  WH_HIP(hipMemcpy(c.d_occ, c.occ.data(), sizeof(double) * nbndval, hipMemcpyHostToDevice));
  eprec_of(nbndval, c.npw, c.d_q, c.npw, c.d_g2kin, c.d_eprec, c.stream);   // set_eprec (dfpt_module.f90:211)
  c.eprec.assign(nbndval, 0.0);
  WH_HIP(hipMemcpyAsync(c.eprec.data(), c.d_eprec, sizeof(double) * nbndval, hipMemcpyDeviceToHost, c.stream));
  WH_HIP(hipStreamSynchronize(c.stream));
  c.have_states = true;
  c.evc_sticks_valid = false;
  return 0;
  WH_API_END
}

int wh_set_coulomb(const double* sqvc) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(c.have_grid, "wh_set_grid first");
  std::vector<double> h(c.npwx);
  WH_HIP(hipMemcpy(h.data(), sqvc, sizeof(double) * c.npwx, hipMemcpyDefault));
  c.sqvc_s.resize(c.npw);
  for (int s = 0; s < c.npw; ++s) c.sqvc_s[s] = h[c.plan.perm[s]];
  if (c.d_sqvc) WH_HIP(hipFree(c.d_sqvc));
  WH_HIP(hipMalloc(&c.d_sqvc, sizeof(double) * c.npw));
  WH_HIP(hipMemcpy(c.d_sqvc, c.sqvc_s.data(), sizeof(double) * c.npw, hipMemcpyHostToDevice));
  c.have_sqvc = true;
  return 0;
  WH_API_END
}

int wh_set_params(int n_dfpt_maxiter, int l_kinetic_only) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(n_dfpt_maxiter >= 1, "n_dfpt_maxiter must be >= 1");
  c.n_dfpt_maxiter = n_dfpt_maxiter;
  c.l_kinetic_only = l_kinetic_only ? 1 : 0;
  return 0;
  WH_API_END
}

int wh_set_davidson_restart(const char* wstat_restart_dir, int n_steps_write_restart, int restart_from) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(n_steps_write_restart >= 0, "n_steps_write_restart must be >= 0");
  WH_REQUIRE(!(n_steps_write_restart > 0 || restart_from) || (wstat_restart_dir && wstat_restart_dir[0]),
             "a restart directory is required");
  c.restart_dir = wstat_restart_dir ? wstat_restart_dir : "";
  c.n_steps_write_restart = n_steps_write_restart;
  c.restart_from = restart_from ? 1 : 0;
  return 0;
  WH_API_END
}

int wh_apply_sternheimerop_to_m_wfcs(int nbndval, const double* psi, double* hpsi, const double* e,
                                     const double* alpha, int m) {
  WH_API_BEGIN
  if (kpath_on()) return kpath_apply_sternheimerop(nbndval, psi, hpsi, e, alpha, m);
  Context& c = ctx();
  // without registered states (nbndval = 0) this is H - e, used to find the states themselves
  require_ready(false);
  WH_REQUIRE(nbndval == c.nbndval, "nbndval differs from the registered states");
  WH_REQUIRE(m >= 0, "m must be >= 0");
  if (m == 0) return 0;
  hipStream_t s = c.stream;
  double a[2] = {0.0, 0.0};
  WH_HIP(hipMemcpy(a, alpha, sizeof(double) * 2, hipMemcpyDefault));
  const size_t cin = static_cast<size_t>(m) * c.npwx;
  const z64* pin = dev_in(reinterpret_cast<const z64*>(psi), cin, c.w_in);
  const double* pe = dev_in(e, m, c.w_ecol);
  z64* x = c.w_x.as<z64>(static_cast<size_t>(m) * c.npw);
  z64* t = c.w_t.as<z64>(static_cast<size_t>(m) * c.npw);
  gather_sticks(m, c.npw, c.d_perm, pin, c.npwx, x, c.npw, nullptr, s);
  if (c.precision != 0) {
    // the FP32 operator of the mixed-precision solve, on FP64 data
    prepare_mixed();
    const long l32 = c.ld32;
    z32* x32 = c.m_bc.as<z32>(static_cast<size_t>(m) * l32);
    z32* t32 = c.m_tc.as<z32>(static_cast<size_t>(m) * l32);
    to_f32(m, 2 * c.npw, reinterpret_cast<const double*>(x), 2L * c.npw, reinterpret_cast<float*>(x32), 2 * l32, s);
    stern_apply32(x32, t32, pe, m, a[0]);
    to_f64(m, 2 * c.npw, reinterpret_cast<const float*>(t32), 2 * l32, reinterpret_cast<double*>(t), 2L * c.npw, s);
  } else {
    stern_apply(x, t, pe, m, a[0]);
  }
  ++c.nstern_calls;
  const bool dout = on_device(hpsi);
  z64* o = dout ? reinterpret_cast<z64*>(hpsi) : c.w_out.as<z64>(cin);
  scatter_caller(m, c.npw, c.npwx, c.d_perm, t, c.npw, o, c.npwx, s);
  if (!dout) WH_HIP(hipMemcpyAsync(hpsi, o, sizeof(z64) * cin, hipMemcpyDeviceToHost, s));
  finish_call();
  return 0;
  WH_API_END
}

int wh_linsolve_sternheimer_m_wfcts(int nbndval, int m, double* b, double* x, const double* e, const double* eprec,
                                    double tr2, int* ierr) {
  WH_API_BEGIN
  if (kpath_on()) return kpath_linsolve(nbndval, m, b, x, e, eprec, tr2, ierr);
  Context& c = ctx();
  require_ready(true);
  WH_REQUIRE(nbndval == c.nbndval, "nbndval differs from the registered states");
  WH_REQUIRE(m >= 0, "m must be >= 0");
  if (ierr) *ierr = 0;
  if (m == 0) return 0;
  hipStream_t s = c.stream;
  const size_t cin = static_cast<size_t>(m) * c.npwx;
  std::vector<double> e_col(m), ep(m);
  WH_HIP(hipMemcpy(e_col.data(), e, sizeof(double) * m, hipMemcpyDefault));
  WH_HIP(hipMemcpy(ep.data(), eprec, sizeof(double) * m, hipMemcpyDefault));
  double* de = c.w_ecol.as<double>(m);
  double* dep = c.w_epcol.as<double>(m);
  WH_HIP(hipMemcpyAsync(de, e_col.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
  WH_HIP(hipMemcpyAsync(dep, ep.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
  z64* db = c.w_b.as<z64>(static_cast<size_t>(m) * c.npw);
  z64* dx = c.w_x.as<z64>(static_cast<size_t>(m) * c.npw);
  const z64* bin = dev_in(reinterpret_cast<const z64*>(b), cin, c.w_in);
  gather_sticks(m, c.npw, c.d_perm, bin, c.npwx, db, c.npw, nullptr, s);
  const z64* xin = dev_in(reinterpret_cast<const z64*>(x), cin, c.w_out);
  gather_sticks(m, c.npw, c.d_perm, xin, c.npwx, dx, c.npw, nullptr, s);
  const int nc = c.precision != 0 ? linsolve_mixed(m, db, dx, e_col, de, dep, tr2, nullptr)
                                  : linsolve_dev(m, db, dx, e_col, de, dep, tr2, nullptr);
  const bool dout = on_device(x);
  z64* o = dout ? reinterpret_cast<z64*>(x) : c.w_out.as<z64>(cin);
  scatter_caller(m, c.npw, c.npwx, c.d_perm, dx, c.npw, o, c.npwx, s);
  if (!dout) WH_HIP(hipMemcpyAsync(x, o, sizeof(z64) * cin, hipMemcpyDeviceToHost, s));
  finish_call();
  if (ierr) *ierr = nc;
  return nc;
  WH_API_END
}

static int dfpt_entry(int m, const double* dvg, double* dng, double tr2, int iq, int* ierr_pert, bool coulomb) {
  if (kpath_on()) return kpath_dfpt(m, dvg, dng, tr2, iq, ierr_pert, coulomb);
  Context& c = ctx();
  require_ready(true);
  WH_REQUIRE(iq == 1, "at Gamma iq must be 1 (q points need a k grid: wh_set_kgrid)");
  WH_REQUIRE(m >= 0, "m must be >= 0");
  if (coulomb) WH_REQUIRE(c.have_sqvc, "wh_set_coulomb has not been called");
  hipStream_t s = c.stream;
  const size_t cin = static_cast<size_t>(m) * c.npwx;
  if (m == 0) return 0;
  const z64* din = dev_in(reinterpret_cast<const z64*>(dvg), cin, c.w_in);
  z64* dvk = c.w_dvin.as<z64>(static_cast<size_t>(m) * c.npw);
  gather_sticks(m, c.npw, c.d_perm, din, c.npwx, dvk, c.npw, coulomb ? c.d_sqvc : nullptr, s);
  z64* dn = c.w_dng.as<z64>(static_cast<size_t>(m) * c.npw);
  std::vector<int> ierr;
  const int nonconv = dfpt_dev(m, dvk, dn, tr2, coulomb ? c.d_sqvc : nullptr, &ierr);
  const bool dout = on_device(dng);
  z64* o = dout ? reinterpret_cast<z64*>(dng) : c.w_out.as<z64>(cin);
  scatter_caller(m, c.npw, c.npwx, c.d_perm, dn, c.npw, o, c.npwx, s);
  if (!dout) WH_HIP(hipMemcpyAsync(dng, o, sizeof(z64) * cin, hipMemcpyDeviceToHost, s));
  finish_call();
  if (ierr_pert) std::copy(ierr.begin(), ierr.end(), ierr_pert);
  return nonconv;
}

int wh_dfpt(int m, const double* dvg, double* dng, double tr2, int iq, int* ierr_per_pert) {
  WH_API_BEGIN
  return dfpt_entry(m, dvg, dng, tr2, iq, ierr_per_pert, false);
  WH_API_END
}

int wh_apply_operator(int m, const double* dvg, double* dng, double tr2, int iq) {
  WH_API_BEGIN
  return dfpt_entry(m, dvg, dng, tr2, iq, nullptr, true);
  WH_API_END
}

int wh_set_precision(int mode, double refine_delta) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_REQUIRE(mode >= 0 && mode <= 2, "precision mode must be 0 (FP64), 1 or 2 (mixed)");
  WH_REQUIRE(!(refine_delta >= 1.0), "refine_delta must be < 1 (<= 0 selects the default 1e-4)");
  c.precision = mode;
  // defaults measured on the 64-atom Si batch: the 6-product operator (about
  // FP32 accurate) carries eight decades of rho between FP64 residuals, the
  // 3-product one six
  c.refine_delta = refine_delta > 0.0 ? refine_delta : (mode == 2 ? 1e-8 : 1e-6);
  if (mode != 0 && c.have_grid)
    WH_REQUIRE(c.plan.tpl32 != nullptr, "mixed precision needs a cubic FFT grid of a compiled size");
  return 0;
  WH_API_END
}

int wh_get_precision_counters(long* fp32_applies, long* refinements, long* fp64_fallbacks) {
  WH_API_BEGIN
  Context& c = ctx();
  if (fp32_applies) *fp32_applies = c.nstern32;
  if (refinements) *refinements = c.nrefine;
  if (fp64_fallbacks) *fp64_fallbacks = c.nfallback;
  return 0;
  WH_API_END
}

int wh_get_counters(long* stern_ncalls, long* cg_iters) {
  WH_API_BEGIN
  Context& c = ctx();
  if (stern_ncalls) *stern_ncalls = c.nstern_calls;
  if (cg_iters) *cg_iters = c.cg_iters;
  return 0;
  WH_API_END
}

int wh_get_operator_columns(long* band_applications) {
  WH_API_BEGIN
  Context& c = ctx();
  if (band_applications) *band_applications = c.nstern;
  return 0;
  WH_API_END
}

int wh_profile(int enable) {
  WH_API_BEGIN
  Context& c = ctx();
  WH_HIP(hipStreamSynchronize(c.stream));
  c.prof.clear();
  c.prof.on = enable != 0;
  return 0;
  WH_API_END
}

int wh_profile_query(const char* name, double* ms, long* count) {
  WH_API_BEGIN
  Context& c = ctx();
  c.prof.flush();
  auto it = c.prof.acc.find(name ? name : "");
  if (ms) *ms = it == c.prof.acc.end() ? 0.0 : it->second.first;
  if (count) *count = it == c.prof.acc.end() ? 0 : it->second.second;
  return 0;
  WH_API_END
}

}  // extern "C"
