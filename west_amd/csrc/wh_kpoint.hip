// Wstat at q points on the device (FP64): the Sternheimer operator and CG at
// a k point, chi0(q) dv, and the complex Davidson of one q point.
//
// Follows the non-Gamma branches of DFPT_kernel/dfpt_module.f90:22-533,
// DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:14-382 (full-sphere products,
// 183-205 and 277-305), DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:14-84,
// Hamiltonian_kernel/apply_alpha_pc_to_m_wfcs.f90 and apply_alpha_pv_to_m_wfcs.f90
// (complex branches, 56-78), and Wstat/davidson_diago.f90:441-905
// (davidson_diago_k) with the complex branch of do_mgs (909-1117) and
// do_randomize_q (1186-1250).
//
// Layout: vectors keep the caller's order (rows of npwx or npwqx complex).
// Away from Gamma a wave function has no Hermitian partner, so there is no
// two-band packing and no half sphere: each band is one complex transform on
// the full FFT grid (GridFft, three batched passes of LDS line transforms).
// The complex projections run on the FP64 matrix-core GEMMs of the Gamma path
// through the real view of the arrays: with Q2 = [evc, i evc, vkb, i vkb]
// (2 npwx x 2 nq reals), Q2^T f = [Re; Im] of Q^H f and Q2 [Re X; Im X] = Q X,
// so each complex bra or ket with a fixed left operand is one real GEMM.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/west_hip.h"
#include "wh_context.h"
#include "wh_fft.h"
#include "wh_gemm.h"
#include "wh_vec.h"

namespace wh {

struct KPointDev {
  int npw = 0;
  int* nl = nullptr;          // 0-based grid index of each k + G
  double* g2kin = nullptr;    // |k+G|^2 (Ry)
  z64* q2 = nullptr;          // (npwx, 2 nq): evc, i evc, vkb, i vkb
  std::vector<double> et;     // nbndval
  std::vector<double> eprec;  // nbndval: 1.35 sum |evc|^2 |k+G|^2 (Tools/set_eprec.f90:40-56)
  double* d_et = nullptr;     // the same two on the device
  double* d_eprec = nullptr;
  bool set = false;
};

struct QPointDev {
  int npwq = 0;
  int* nl = nullptr;          // 0-based grid index of each q + G
  double* gq2 = nullptr;      // |q+G|^2 (Ry)
  int* ig = nullptr;          // 1-based global G index (start-vector counter)
  double* sqvc = nullptr;
  bool is_gamma = false;      // q = 0: q + G = 0 is the first vector
  std::vector<int> ikq;       // 0-based index of k - q for each k
  std::vector<int> g0;        // (3, nks): k - q = k_ikq + g0, crystal integers
  KlQpoint* d_kq = nullptr;   // the same two on the device, per k point
  bool set = false;
};

struct KPath {
  int nks = 0, nqs = 0, npwx = 0, npwqx = 0, nbndval = 0, nkb = 0;
  double omega = 0.0, alpha_pv = 0.0;
  std::vector<double> wk, dvan;
  double* vrs = nullptr;
  GridFft fft;
  std::vector<KPointDev> k;
  std::vector<QPointDev> q;
  int current_k = 0;
  double* qscale2 = nullptr;  // (2 nq): alpha (evc and i evc rows), D (vkb and i vkb rows)
  double qscale_alpha = 0.0;
  bool qscale_set = false;
  size_t grid_budget = size_t(2) << 30;   // bytes of grids per batch
  std::vector<void*> allocs;
  DevBuf g_a, g_b, g_c, g_aux;                       // grids
  DevBuf w_x, w_t, w_g, w_h, w_hold, w_bc, w_tc;     // CG vectors
  DevBuf w_c2, w_l2i, w_rho, w_rhoold, w_conv, w_nonconv, w_beta, w_eu, w_ecol, w_epcol;
  DevBuf w_dv, w_dn, w_in, w_out, w_in2;
  DevBuf dav[8];
  hipEvent_t cg_ev[4] = {nullptr, nullptr, nullptr, nullptr};   // lagged reads of the CG loop
  KlKpoint* ktab = nullptr;   // device: each k point's sphere and projectors (batched CG)
  double* sck = nullptr;      // device: 2 w_k / (Omega N_r) per k point (densities at q)
  DevBuf w_dk;                // densities per (k point, perturbation) before the sum over k
};

// The k-point CG reads an iteration's active count this many iterations late:
// its iterations are short (tens of microseconds), so the queue keeps two
// of them ahead of a host that wakes from an event wait.
constexpr int kCgLag = 3;
static_assert(sizeof(KPath::cg_ev) / sizeof(hipEvent_t) == kCgLag + 1, "one event per iteration in flight");

void free_kpath(KPath* k) {
  if (!k) return;
  for (void* p : k->allocs) (void)hipFree(p);
  k->fft.release();
  for (DevBuf* b : {&k->g_a, &k->g_b, &k->g_c, &k->g_aux, &k->w_x, &k->w_t, &k->w_g, &k->w_h, &k->w_hold, &k->w_bc,
                    &k->w_tc, &k->w_c2, &k->w_l2i, &k->w_rho, &k->w_rhoold, &k->w_conv, &k->w_nonconv, &k->w_beta,
                    &k->w_eu, &k->w_ecol, &k->w_epcol, &k->w_dv, &k->w_dn, &k->w_in, &k->w_out, &k->w_in2})
    b->release();
  for (DevBuf& b : k->dav) b.release();
  k->w_dk.release();
  for (hipEvent_t e : k->cg_ev)
    if (e) (void)hipEventDestroy(e);
  delete k;
}

bool kpath_on() { return ctx().kpath != nullptr; }

namespace {

constexpr int NT = 256;

template <typename T>
T* upload(KPath& K, const T* h, size_t n) {
  T* d = nullptr;
  WH_HIP(hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)));
  if (n) WH_HIP(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  K.allocs.push_back(d);
  return d;
}

void release(KPath& K, void* p) {
  if (!p) return;
  auto it = std::find(K.allocs.begin(), K.allocs.end(), p);
  if (it != K.allocs.end()) K.allocs.erase(it);
  WH_HIP(hipFree(p));
}

KPath& kp() {
  Context& c = ctx();
  WH_REQUIRE(c.kpath != nullptr, "no k grid registered (wh_set_kgrid)");
  return *c.kpath;
}

inline unsigned blocks(long n) { return static_cast<unsigned>(std::min<long>((n + NT - 1) / NT, 1L << 20)); }

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NT / 64; ++i) s += red[i];
  return s;
}

// grid[b][nl[ig]] = psi[b][ig]   (grids zeroed before)
__global__ void k_scatter(int m, int npw, const int* __restrict__ nl, const z64* __restrict__ psi, long ld,
                          z64* __restrict__ grid, long nnr) {
  const long n = static_cast<long>(m) * npw;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const int b = static_cast<int>(t / npw);
    const int ig = static_cast<int>(t - static_cast<long>(b) * npw);
    grid[b * nnr + nl[ig]] = psi[b * ld + ig];
  }
}

// out[b][ig] = scale grid[b][nl[ig]] (+ (g2[ig] - e[b]) in[b][ig]) for ig < npw,
// 0 for npw <= ig < ld; with acc, out += scale grid[b][nl[ig]] instead.
// grid may be null (kinetic term only).
__global__ void k_gather(int m, int npw, long ld, const int* __restrict__ nl, const z64* __restrict__ grid, long nnr,
                         double scale, const z64* __restrict__ in, const double* __restrict__ g2,
                         const double* __restrict__ e, int acc, z64* __restrict__ out) {
  const long n = static_cast<long>(m) * ld;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const int b = static_cast<int>(t / ld);
    const int ig = static_cast<int>(t - static_cast<long>(b) * ld);
    if (ig >= npw) {
      if (!acc) out[t] = mk(0.0, 0.0);
      continue;
    }
    z64 v = grid ? cscale(grid[b * nnr + nl[ig]], scale) : mk(0.0, 0.0);
    if (in) v = v + cscale(in[t], g2[ig] - e[b]);
    out[t] = acc ? out[t] + v : v;
  }
}

__global__ void k_mul_real(int m, long nnr, const double* __restrict__ v, z64* __restrict__ g) {
  const long n = m * nnr;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT)
    g[t] = cscale(g[t], v[t % nnr]);
}

// grid[v][r] *= exp(-i g0 . r) at r = (i/n1, j/n2, l/n3) in crystal units
// (Tools/types_bz_grid.f90:71-113); the phase is taken from integer numerators
__global__ void k_phase(int nb, int n1, int n2, int n3, int g1, int g2, int g3, z64* __restrict__ g) {
  const long nnr = static_cast<long>(n1) * n2 * n3;
  const long n = nb * nnr;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const long r = t % nnr;
    const int i = static_cast<int>(r % n1);
    const int j = static_cast<int>((r / n1) % n2);
    const int l = static_cast<int>(r / (static_cast<long>(n1) * n2));
    const long a1 = ((static_cast<long>(g1) * i) % n1 + n1) % n1;
    const long a2 = ((static_cast<long>(g2) * j) % n2 + n2) % n2;
    const long a3 = ((static_cast<long>(g3) * l) % n3 + n3) % n3;
    const double x = 2.0 * (static_cast<double>(a1) / n1 + static_cast<double>(a2) / n2 + static_cast<double>(a3) / n3);
    double sn, cs;
    sincospi(x, &sn, &cs);
    g[t] = cmul(g[t], mk(cs, -sn));
  }
}

// out[p nb + v][r] = psik[v][r] aux[p][r]   (dfpt_module.f90:295)
__global__ void k_rhs(int nb, int mp, long nnr, const z64* __restrict__ psik, const z64* __restrict__ aux,
                      z64* __restrict__ out) {
  const long n = static_cast<long>(mp) * nb * nnr;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const long col = t / nnr;
    const long r = t - col * nnr;
    const int p = static_cast<int>(col / nb);
    const int v = static_cast<int>(col - static_cast<long>(p) * nb);
    out[t] = cmul(psik[v * nnr + r], aux[p * nnr + r]);
  }
}

// out[p][r] = sum_v conj(psik[v][r]) dpsi[p nb + v][r]   (dfpt_module.f90:441)
__global__ void k_density(int nb, int mp, long nnr, const z64* __restrict__ psik, const z64* __restrict__ dpsi,
                          z64* __restrict__ out) {
  const long n = static_cast<long>(mp) * nnr;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const int p = static_cast<int>(t / nnr);
    const long r = t - static_cast<long>(p) * nnr;
    z64 s = mk(0.0, 0.0);
    for (int v = 0; v < nb; ++v) s = s + cmul(conjg(psik[v * nnr + r]), dpsi[(static_cast<long>(p) * nb + v) * nnr + r]);
    out[t] = s;
  }
}

// a[b][ig] *= w[ig] for ig < n
__global__ void k_scale_rows(int m, int n, long ld, const double* __restrict__ w, z64* __restrict__ a) {
  const long tot = static_cast<long>(m) * n;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < tot; t += static_cast<long>(gridDim.x) * NT) {
    const int b = static_cast<int>(t / n);
    const int ig = static_cast<int>(t - static_cast<long>(b) * n);
    a[b * ld + ig] = cscale(a[b * ld + ig], w[ig]);
  }
}

// pf = f / max(1, g2kin / eprec[col])  (precondition_m_wfcts.f90:35-51)
__global__ void k_precond(int m, int npw, long ld, const double* __restrict__ g2, const double* __restrict__ ep,
                          const z64* __restrict__ f, z64* __restrict__ pf) {
  const long n = static_cast<long>(m) * ld;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const int b = static_cast<int>(t / ld);
    const int ig = static_cast<int>(t - static_cast<long>(b) * ld);
    pf[t] = ig < npw ? cscale(f[t], 1.0 / fmax(1.0, g2[ig] / ep[b])) : mk(0.0, 0.0);
  }
}

// ---- CG at a k point (linsolve_sternheimer_m_wfcts.f90, non-Gamma products)

// The iteration's kernels are launched for a capacity of listed columns and
// skip those at or beyond the device's active count *nc (a count set by
// cg_select on the device; nc null: all launched columns).

// h = g / max(1, g2kin/eprec) on the listed columns, rho[band] = Re <h|g> (150-160, 183-205)
__global__ __launch_bounds__(NT) void k_cg_prec_rho(const int* __restrict__ l2i, const int* __restrict__ nc,
                                                    int npw, long ld, const double* __restrict__ g2,
                                                    const double* __restrict__ ep, const z64* __restrict__ g,
                                                    z64* __restrict__ h, double* __restrict__ rho) {
  __shared__ double red[NT / 64];
  if (nc && static_cast<int>(blockIdx.x) >= *nc) return;   // whole workgroup
  const int col = l2i[blockIdx.x];
  const z64* gc = g + col * ld;
  z64* hc = h + col * ld;
  const double e = ep[col];
  double s = 0.0;
  for (int ig = threadIdx.x; ig < ld; ig += NT) {
    if (ig < npw) {
      const z64 hv = cscale(gc[ig], 1.0 / fmax(1.0, g2[ig] / e));
      hc[ig] = hv;
      s += hv.x * gc[ig].x + hv.y * gc[ig].y;
    } else {
      hc[ig] = mk(0.0, 0.0);
    }
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) rho[col] = s;
}

// h = -h + beta hold on the listed columns; bc[lbnd] = h   (226-240)
__global__ void k_cg_dir(int cap, const int* __restrict__ nc, const int* __restrict__ l2i, long ld,
                         const double* __restrict__ beta, const double* __restrict__ e, z64* __restrict__ h,
                         const z64* __restrict__ hold, z64* __restrict__ bc, double* __restrict__ eu) {
  const long n = static_cast<long>(min(cap, *nc)) * ld;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT) {
    const int lb = static_cast<int>(t / ld);
    const long ig = t - static_cast<long>(lb) * ld;
    const int col = l2i[lb];
    const z64 v = cscale(h[col * ld + ig], -1.0) + cscale(hold[col * ld + ig], beta[lb]);
    h[col * ld + ig] = v;
    bc[t] = v;
    if (ig == 0) eu[lb] = e[col];
  }
}

// One workgroup per listed column: a = Re <h|g>, c = Re <h|t> (277-305);
// x -= a/c h, g -= a/c t, hold = h (313-321); then the next iteration's
// h = g / max(1, g2kin/eprec) and rho[band] = Re <h|g> (150-205).
__global__ __launch_bounds__(NT) void k_cg_step(const int* __restrict__ l2i, const int* __restrict__ nc, int npw,
                                                long ld, const double* __restrict__ g2,
                                                const double* __restrict__ ep, z64* __restrict__ x,
                                                z64* __restrict__ g, z64* __restrict__ h,
                                                const z64* __restrict__ tc, z64* __restrict__ hold,
                                                double* __restrict__ rho) {
  __shared__ double red[NT / 64];
  const int lb = blockIdx.x;
  if (lb >= *nc) return;   // whole workgroup
  const int col = l2i[lb];
  z64* xc = x + col * ld;
  z64* gc = g + col * ld;
  z64* hc = h + col * ld;
  z64* ho = hold + col * ld;
  const z64* tt = tc + lb * ld;
  double sa = 0.0, sc = 0.0;
  for (int ig = threadIdx.x; ig < npw; ig += NT) {
    sa += hc[ig].x * gc[ig].x + hc[ig].y * gc[ig].y;
    sc += hc[ig].x * tt[ig].x + hc[ig].y * tt[ig].y;
  }
  sa = block_sum(sa, red);
  sc = block_sum(sc, red);
  const double st = sa / sc;
  const double e = ep[col];
  double sr = 0.0;
  for (int ig = threadIdx.x; ig < ld; ig += NT) {
    const z64 hv = hc[ig];
    const z64 gn = gc[ig] - cscale(tt[ig], st);
    xc[ig] = xc[ig] - cscale(hv, st);
    gc[ig] = gn;
    ho[ig] = hv;
    if (ig < npw) {
      const z64 hn = cscale(gn, 1.0 / fmax(1.0, g2[ig] / e));
      hc[ig] = hn;
      sr += hn.x * gn.x + hn.y * gn.y;
    } else {
      hc[ig] = mk(0.0, 0.0);
    }
  }
  sr = block_sum(sr, red);
  if (threadIdx.x == 0) rho[col] = sr;
}

// nonconv[(j / group) % nper] += !conv[j]  (group <= 0: one total)
__global__ void k_count_nonconv(int m, int group, int nper, const int* __restrict__ conv, int* __restrict__ nonconv) {
  const int j = blockIdx.x * NT + threadIdx.x;
  if (j < m && !conv[j]) atomicAdd(nonconv + (group > 0 ? (j / group) % nper : 0), 1);
}

// a[j * ld] = 0, j < m
__global__ void k_zero_first(int m, long ld, z64* __restrict__ a) {
  const int j = blockIdx.x * NT + threadIdx.x;
  if (j < m) a[j * ld] = mk(0.0, 0.0);
}

// per-column energies of `cols` columns, nb bands per perturbation
__global__ void k_band_cols(int cols, int nb, const double* __restrict__ et, const double* __restrict__ ep,
                            double* __restrict__ ec, double* __restrict__ epc) {
  const int j = blockIdx.x * NT + threadIdx.x;
  if (j < cols) {
    ec[j] = et[j % nb];
    epc[j] = ep[j % nb];
  }
}

__global__ void k_sub(long n, const z64* __restrict__ b, z64* __restrict__ g) {
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * NT)
    g[t] = g[t] - b[t];
}

// ---- Davidson helpers

// do_randomize_q (davidson_diago.f90:1230-1242): vector ivec, plane wave with
// global index ig: (u0 cos 2pi u1 + i u0 sin 2pi u1) / (|q+G|^2 + 1), 0 at q+G = 0
__global__ void k_randomize_q(int npwq, int ivec0, int nvec, const int* __restrict__ ig,
                              const double* __restrict__ gq2, z64* __restrict__ amat, long ld) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (s >= npwq || r >= nvec) return;
  double u0, u1;
  start_uniforms(static_cast<unsigned long long>(ivec0 + r), static_cast<unsigned long long>(ig[s]), u0, u1);
  z64 v = mk(0.0, 0.0);
  if (gq2[s] >= 1e-8) {
    const double arg = 2.0 * 3.14159265358979323846 * u1;
    const double den = gq2[s] + 1.0;
    v = mk(u0 * cos(arg) / den, u0 * sin(arg) / den);
  }
  amat[static_cast<long>(r) * ld + s] = v;
}

// v <- v / sqrt(sum |v|^2)   (do_mgs, 1005-1025, complex branch)
__global__ __launch_bounds__(NT) void k_znormalize(int n, z64* v) {
  __shared__ double red[NT / 64];
  double s = 0.0;
  for (int q = threadIdx.x; q < n; q += NT) s += v[q].x * v[q].x + v[q].y * v[q].y;
  s = block_sum(s, red);
  const double f = 1.0 / sqrt(s);
  for (int q = threadIdx.x; q < n; q += NT) v[q] = cscale(v[q], f);
}

// w[:, j] = b[:, j], w[:, nb + j] = -i b[:, j]   (rows < n)
__global__ void k_rot_copy(int nb, int n, long ld, const z64* __restrict__ b, z64* __restrict__ w) {
  const long tot = static_cast<long>(nb) * n;
  for (long t = blockIdx.x * static_cast<long>(NT) + threadIdx.x; t < tot; t += static_cast<long>(gridDim.x) * NT) {
    const int j = static_cast<int>(t / n);
    const int ig = static_cast<int>(t - static_cast<long>(j) * n);
    const z64 v = b[j * ld + ig];
    w[j * ld + ig] = v;
    w[(static_cast<long>(nb) + j) * ld + ig] = mk(v.y, -v.x);
  }
}

// y += alpha i t   (rows < n)
__global__ void k_add_i(int nb, int n, long ld, double alpha, const z64* __restrict__ t, z64* __restrict__ y) {
  const long tot = static_cast<long>(nb) * n;
  for (long q = blockIdx.x * static_cast<long>(NT) + threadIdx.x; q < tot; q += static_cast<long>(gridDim.x) * NT) {
    const int j = static_cast<int>(q / n);
    const long o = j * ld + (q - static_cast<long>(j) * n);
    y[o] = mk(y[o].x - alpha * t[o].y, y[o].y + alpha * t[o].x);
  }
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// device view of a caller array (copied when it is host memory)
template <typename T>
const T* dev_in(const T* p, size_t count, DevBuf& tmp) {
  if (on_device_mem(p)) {
    wait_for_caller(ctx());
    return p;
  }
  T* d = tmp.as<T>(count);
  WH_HIP(hipMemcpyAsync(d, p, sizeof(T) * count, hipMemcpyHostToDevice, ctx().stream));
  return d;
}

// columns of grids that fit the grid budget
int grid_cols(const KPath& K) {
  const long per = K.fft.nnr() * static_cast<long>(sizeof(z64));
  return static_cast<int>(std::max<long>(1, std::min<long>(static_cast<long>(K.grid_budget) / per, 1L << 16)));
}

void zero_grids(z64* g, long count, hipStream_t s) { WH_HIP(hipMemsetAsync(g, 0, sizeof(z64) * count, s)); }

void set_qscale2(KPath& K, double alpha) {
  const int nq = K.nbndval + K.nkb;
  if (nq == 0 || (K.qscale_set && alpha == K.qscale_alpha)) return;
  std::vector<double> sc(2 * static_cast<size_t>(nq));
  for (int i = 0; i < K.nbndval; ++i) sc[i] = sc[K.nbndval + i] = alpha;
  for (int i = 0; i < K.nkb; ++i) sc[2 * K.nbndval + i] = sc[2 * K.nbndval + K.nkb + i] = K.dvan[i];
  if (!K.qscale2) K.qscale2 = upload(K, sc.data(), sc.size());
  else WH_HIP(hipMemcpy(K.qscale2, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice));
  K.qscale_alpha = alpha;
  K.qscale_set = true;
}

}  // namespace

// ---------------------------------------------------------------------------
// t = H_k x - e x + alpha P_v x  (apply_sternheimerop_to_m_wfcs.f90:49-76 at k:
// QE h_psi_ = |k+G|^2 + vloc_psi_k + calbec/add_vuspsi, then alpha P_v)
// ---------------------------------------------------------------------------
// Small grids keep the local term in LDS (WH_KLDS=0: the full-grid
// transforms through HBM)
static bool klds_enabled() {
  static const bool on = [] {
    const char* ev = std::getenv("WH_KLDS");
    return !(ev && ev[0] == '0');
  }();
  return on;
}

// d_nc (device int, may be null): only the first min(m, *d_nc) columns are
// used (a CG launch sized for a capacity); the caller counts the applications.
static void stern_k(KPath& K, int ik, const z64* x, z64* t, const double* d_e, int m, double alpha,
                    const int* d_nc = nullptr) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  const KPointDev& P = K.k[ik];
  const long ld = K.npwx;
  const long nnr = K.fft.nnr();
  const int gc = grid_cols(K);
  // small grids: the whole local term in one launch with the grid in LDS
  const bool lds = klds_enabled() && !c.l_kinetic_only && kvloc_lds_available(K.fft.n1, K.fft.n2, K.fft.n3);
  if (lds)
    WH_TIMED("kvloc_lds", kvloc_lds(K.fft.n1, K.fft.n2, K.fft.n3, m, P.npw, ld, P.nl, x, K.vrs, P.g2kin, d_e, t, s,
                                    d_nc));
  for (int c0 = 0; c0 < (lds ? 0 : m); c0 += gc) {
    const int mc = std::min(gc, m - c0);
    const z64* xc = x + c0 * ld;
    z64* tc = t + c0 * ld;
    z64* g = nullptr;
    if (!c.l_kinetic_only) {
      g = K.g_a.as<z64>(static_cast<size_t>(mc) * nnr);
      zero_grids(g, mc * nnr, s);
      k_scatter<<<blocks(static_cast<long>(mc) * P.npw), NT, 0, s>>>(mc, P.npw, P.nl, xc, ld, g, nnr);
      WH_TIMED("kfft", launch_grid_fft(K.fft, g, mc, +1, s));
      k_mul_real<<<blocks(mc * nnr), NT, 0, s>>>(mc, nnr, K.vrs, g);
      WH_TIMED("kfft", launch_grid_fft(K.fft, g, mc, -1, s));
    }
    k_gather<<<blocks(mc * ld), NT, 0, s>>>(mc, P.npw, ld, P.nl, g, nnr, 1.0 / static_cast<double>(nnr), xc,
                                            P.g2kin, d_e + c0, 0, tc);
    WH_HIP(hipGetLastError());
  }
  // alpha P_v and the KB term: t += Q (S (Q^H x)) as two real GEMMs on Q2
  const int nq = c.l_kinetic_only ? K.nbndval : K.nbndval + K.nkb;
  if (nq > 0 && m > 0) {
    set_qscale2(K, alpha);
    double* c2 = K.w_c2.as<double>(static_cast<size_t>(2) * nq * m);
    const double* q2 = reinterpret_cast<const double*>(P.q2);
    WH_TIMED("gemm_bra", dgemm_bra(2 * nq, m, 2 * P.npw, 1.0, q2, 2 * ld, reinterpret_cast<const double*>(x), 2 * ld,
                                   0.0, K.qscale2, c2, 2 * nq, s, d_nc));
    WH_TIMED("gemm_ket", dgemm(false, 2 * P.npw, m, 2 * nq, 1.0, q2, 2 * ld, c2, 2 * nq, 1.0,
                               reinterpret_cast<double*>(t), 2 * ld, s, d_nc));
  }
  if (!d_nc) c.nstern += m;
}

// f <- -f + evc (evc^H f)   (apply_alpha_pc_to_m_wfcs with alpha = -1, dfpt_module.f90:346)
static void minus_pc_k(KPath& K, int ik, z64* f, int m) {
  Context& c = ctx();
  const KPointDev& P = K.k[ik];
  const long ld = K.npwx;
  const int nb = K.nbndval;
  if (nb == 0 || m == 0) return;
  double* c2 = K.w_c2.as<double>(static_cast<size_t>(2) * nb * m);
  const double* q2 = reinterpret_cast<const double*>(P.q2);
  WH_TIMED("gemm_bra", dgemm_bra(2 * nb, m, 2 * P.npw, 1.0, q2, 2 * ld, reinterpret_cast<const double*>(f), 2 * ld,
                                 0.0, nullptr, c2, 2 * nb, c.stream));
  WH_TIMED("gemm_ket", dgemm(false, 2 * P.npw, m, 2 * nb, 1.0, q2, 2 * ld, c2, 2 * nb, -1.0,
                             reinterpret_cast<double*>(f), 2 * ld, c.stream));
}

// PCG for (H_k - e_i + alpha_pv P_v) x_i = b_i, i < m; b is destroyed.
// Columns are perturbation-major, `group` bands per perturbation (call count).
// The loop (linsolve_sternheimer_m_wfcts.f90:146-330) runs on the device:
// cg_select does the convergence test, the compaction of the active list and
// beta at the start of each iteration, and the iteration's kernels are
// launched for a capacity (the active count the host last read, rounded up
// to a power of two) and skip the columns beyond the device's count.  The
// host reads the active count of the iteration kCgLag back (finished while
// the later ones are queued) and stops once it is zero; iterations launched
// after the last band converged do nothing.  Nothing waits for the end: the
// unconverged bands are added on the device to nonconv[(band / group) % nper]
// (group <= 0: nonconv[0]), each iteration's (active count, perturbations)
// goes to rec[2 it] (pinned, 2 (n_dfpt_maxiter + 1) ints), and the caller,
// once the stream is done, passes rec and the returned number of iterations
// launched to cg_k_count.
// nkb = 1: the cols columns belong to k point ik.  nkb = nks: columns
// [k cols, (k + 1) cols) belong to k point k, all in one solve (the fused
// iteration only, kbatch_ok).
static int linsolve_k(KPath& K, int ik, int nkb, int cols, z64* b, z64* x, const double* d_e, const double* d_ep,
                      double tr2, int group, int nper, int* rec, int* nonconv) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  const int m = nkb * cols;
  WH_REQUIRE(nkb == 1 || nkb == K.nks, "linsolve_k: one k point or all of them");
  const KPointDev& P = K.k[nkb == 1 ? ik : 0];
  const long ld = K.npwx;
  const size_t vec = static_cast<size_t>(m) * ld;
  z64* g = K.w_g.as<z64>(vec);
  z64* h = K.w_h.as<z64>(vec);
  z64* hold = K.w_hold.as<z64>(vec);
  z64* bc = K.w_bc.as<z64>(vec);
  z64* tc = K.w_tc.as<z64>(vec);
  int* d_l2i = K.w_l2i.as<int>(m);
  double* d_rhob = K.w_rho.as<double>(m);       // rho by band
  double* d_rhoold = K.w_rhoold.as<double>(m);
  double* d_beta = K.w_beta.as<double>(m);
  double* d_eu = K.w_eu.as<double>(m);
  int* d_conv = K.w_conv.as<int>(static_cast<size_t>(m) + 4);
  int* d_nc = d_conv + m;       // active count of the current iteration
  int* d_rec = d_conv + m + 1;  // (active count, perturbations among them)
  if (!K.cg_ev[0])
    for (hipEvent_t& e : K.cg_ev) WH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  const int itmax = c.n_dfpt_maxiter;
  WH_HIP(hipMemsetAsync(hold, 0, sizeof(z64) * vec, s));
  WH_HIP(hipMemsetAsync(d_conv, 0, sizeof(int) * (static_cast<size_t>(m) + 4), s));
  // small grids with few projector rows: the iteration after cg_select is
  // one workgroup per column (cg_iter_lds) instead of five launches
  const int nq2 = 2 * (K.nbndval + K.nkb);
  const bool fused = klds_enabled() && !c.l_kinetic_only &&
                     cg_iter_lds_available(K.fft.n1, K.fft.n2, K.fft.n3, nq2);
  WH_REQUIRE(fused || nkb == 1, "linsolve_k: several k points need the fused iteration");
  if (fused) set_qscale2(K, K.alpha_pv);
  // step 1 (127-138): g = A x - b, and h = M^-1 g, rho of every band
  if (nkb > 1) {   // every k point's columns in one launch
    WH_TIMED("cg_iter_lds", cg_first_lds(K.fft.n1, K.fft.n2, K.fft.n3, ld, K.vrs, nq2, K.qscale2, m, d_e, d_ep, x,
                                         b, g, h, d_rhob, K.ktab, cols, s));
    c.nstern += m;
  } else {
    iota(d_l2i, m, s);
    stern_k(K, ik, x, g, d_e, m, K.alpha_pv);
    k_sub<<<blocks(static_cast<long>(vec)), NT, 0, s>>>(static_cast<long>(vec), b, g);
    k_cg_prec_rho<<<m, NT, 0, s>>>(d_l2i, nullptr, P.npw, ld, P.g2kin, d_ep, g, h, d_rhob);
    WH_HIP(hipGetLastError());
  }
  count_calls(m, nullptr, group);
  int last = 0;   // iterations launched
  int cap = m;
  for (int it = 1; it <= itmax; ++it) {
    if (it > kCgLag) {
      WH_HIP(hipEventSynchronize(K.cg_ev[(it - kCgLag) % (kCgLag + 1)]));
      const int nlag = rec[2 * (it - kCgLag)];
      if (nlag == 0) break;
      int p2 = 1;
      while (p2 < nlag) p2 <<= 1;
      cap = std::min(m, p2);
    }
    // convergence test, compaction, beta (208-224)
    cg_select(m, tr2, it == 1 ? 1 : 0, group, d_rhob, d_rhoold, d_conv, d_e, d_l2i, d_beta, d_eu, d_nc, d_rec, s);
    if (fused) {
      WH_TIMED("cg_iter_lds", cg_iter_lds(K.fft.n1, K.fft.n2, K.fft.n3, P.npw, ld, P.nl, K.vrs, P.g2kin, nq2,
                                          reinterpret_cast<const double*>(P.q2), K.qscale2, cap, d_l2i, d_nc,
                                          d_beta, d_eu, d_ep, h, hold, x, g, bc, tc, d_rhob, s,
                                          nkb > 1 ? K.ktab : nullptr, cols));
    } else {
      k_cg_dir<<<blocks(static_cast<long>(cap) * ld), NT, 0, s>>>(cap, d_nc, d_l2i, ld, d_beta, d_e, h, hold, bc,
                                                                  d_eu);
      stern_k(K, ik, bc, tc, d_eu, cap, K.alpha_pv, d_nc);
      k_cg_step<<<cap, NT, 0, s>>>(d_l2i, d_nc, P.npw, ld, P.g2kin, d_ep, x, g, h, tc, hold, d_rhob);
      WH_HIP(hipGetLastError());
    }
    WH_HIP(hipMemcpyAsync(rec + 2 * it, d_rec, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    WH_HIP(hipEventRecord(K.cg_ev[it % (kCgLag + 1)], s));
    last = it;
  }
  // the bands still active when the iterations ran out (332)
  k_count_nonconv<<<blocks(m), NT, 0, s>>>(m, group, nper, d_conv, nonconv);
  WH_HIP(hipGetLastError());
  return last;
}

// counters of a finished linsolve_k: an iteration counts when it had active bands
static void cg_k_count(const int* rec, int last) {
  Context& c = ctx();
  for (int it = 1; it <= last; ++it) {
    const int n = rec[2 * it];
    if (n > 0) {
      ++c.cg_iters;
      c.nstern += n;
      c.nstern_calls += rec[2 * it + 1];
    }
  }
}

// Whether dfpt at q can take every k point at once for m perturbations: the
// fused CG iteration applies (small grid, at most 32 projector rows in the
// real view) and the solve's vectors (seven of nks m nbndval columns), the
// states at k - q and density scratch grids and the per-k densities fit the
// grid budget.
static bool kbatch_ok(const KPath& K, int m) {
  const char* ev = std::getenv("WH_KBATCH");   // read per call: 0 keeps the k points one after another
  if (ev && ev[0] == '0') return false;
  const Context& c = ctx();
  const long all = static_cast<long>(K.nks) * m * K.nbndval;
  const long grids = static_cast<long>(K.nks) * (K.nbndval + m);   // states at k - q, density scratch
  const long bytes = static_cast<long>(sizeof(z64)) * (7 * all * K.npwx + grids * K.fft.nnr() +
                                                       static_cast<long>(K.nks) * m * K.npwqx);
  return K.nks > 1 && klds_enabled() && !c.l_kinetic_only && all <= (1L << 24) &&
         cg_iter_lds_available(K.fft.n1, K.fft.n2, K.fft.n3, 2 * (K.nbndval + K.nkb)) &&
         bytes <= static_cast<long>(K.grid_budget);
}

// dng (npwqx, m) = chi0(q) dvg for q point iq (0-based)  (dfpt_module.f90:163-505)
static int dfpt_q(KPath& K, int iq, int m, const z64* dvg, z64* dng, double tr2, std::vector<int>* ierr_pert) {
  Context& c = ctx();
  hipStream_t s = c.stream;
  const QPointDev& Q = K.q[iq];
  WH_REQUIRE(Q.set, "this q point has not been registered (wh_set_qpoint)");
  for (int ik = 0; ik < K.nks; ++ik) WH_REQUIRE(K.k[ik].set, "a k point has not been registered (wh_set_kpoint)");
  const long ld = K.npwx;
  const long ldq = K.npwqx;
  const long nnr = K.fft.nnr();
  const int nb = K.nbndval;
  WH_HIP(hipMemsetAsync(dng, 0, sizeof(z64) * m * ldq, s));
  if (ierr_pert) ierr_pert->assign(m, 0);
  if (m == 0 || nb == 0) return 0;
  // perturbations per chunk: their band columns' grids within the budget
  const int mp = std::max(1, std::min(m, grid_cols(K) / std::max(nb, 1)));
  // per solve: its CG records (pinned) and iterations launched, read at the end
  const int nsolve = K.nks * ((m + mp - 1) / mp);
  const size_t rec_len = 2 * (static_cast<size_t>(c.n_dfpt_maxiter) + 1);
  int* recs = c.h_cgrec.as<int>(rec_len * nsolve);
  std::vector<int> launched;
  int* d_nonconv = K.w_nonconv.as<int>(m);   // unconverged bands per perturbation
  WH_HIP(hipMemsetAsync(d_nonconv, 0, sizeof(int) * m, s));
  // the perturbations in real space (237)
  z64* aux = K.g_aux.as<z64>(static_cast<size_t>(m) * nnr);
  zero_grids(aux, m * nnr, s);
  k_scatter<<<blocks(static_cast<long>(m) * Q.npwq), NT, 0, s>>>(m, Q.npwq, Q.nl, dvg, ldq, aux, nnr);
  WH_TIMED("kfft", launch_grid_fft(K.fft, aux, m, +1, s));
  // states at k - q in real space with the phase exp(-i g0.r) (288, 295)
  auto states_kmq = [&](int ik) {
    const KPointDev& PQ = K.k[Q.ikq[ik]];
    z64* psik = K.g_b.as<z64>(static_cast<size_t>(nb) * nnr);
    zero_grids(psik, nb * nnr, s);
    k_scatter<<<blocks(static_cast<long>(nb) * PQ.npw), NT, 0, s>>>(nb, PQ.npw, PQ.nl, PQ.q2, ld, psik, nnr);
    WH_TIMED("kfft", launch_grid_fft(K.fft, psik, nb, +1, s));
    k_phase<<<blocks(nb * nnr), NT, 0, s>>>(nb, K.fft.n1, K.fft.n2, K.fft.n3, Q.g0[3 * ik], Q.g0[3 * ik + 1],
                                           Q.g0[3 * ik + 2], psik);
    WH_HIP(hipGetLastError());
    return psik;
  };
  // perturbations p0 .. p0 + mc - 1 at k point ik: per column e at k - q
  // (222) and eprec from the states at k (211); dvpsi = -P_c fwfft(psi_{k-q}
  // phase dv) on the k sphere (295-302, 346) and dpsi = M^-1 dvpsi (348)
  auto rhs = [&](int ik, int p0, int mc, const z64* psik, z64* dvpsi, z64* dpsi, double* d_e, double* d_ep) {
    const KPointDev& P = K.k[ik];
    const KPointDev& PQ = K.k[Q.ikq[ik]];
    const int cols = mc * nb;
    k_band_cols<<<blocks(cols), NT, 0, s>>>(cols, nb, PQ.d_et, P.d_eprec, d_e, d_ep);
    z64* R = K.g_a.as<z64>(static_cast<size_t>(cols) * nnr);
    k_rhs<<<blocks(static_cast<long>(cols) * nnr), NT, 0, s>>>(nb, mc, nnr, psik, aux + p0 * nnr, R);
    WH_TIMED("kfft", launch_grid_fft(K.fft, R, cols, -1, s));
    k_gather<<<blocks(static_cast<long>(cols) * ld), NT, 0, s>>>(cols, P.npw, ld, P.nl, R, nnr,
                                                                 1.0 / static_cast<double>(nnr), nullptr, nullptr,
                                                                 nullptr, 0, dvpsi);
    WH_HIP(hipGetLastError());
    minus_pc_k(K, ik, dvpsi, cols);
    k_precond<<<blocks(static_cast<long>(cols) * ld), NT, 0, s>>>(cols, P.npw, ld, P.g2kin, d_ep, dvpsi, dpsi);
    WH_HIP(hipGetLastError());
  };
  // response density sum_v conj(psi_{k-q} phase) dpsi (433-441) and
  // dng += 2 wk / Omega fwfft(aux) on the q sphere (480-487)
  auto density = [&](int ik, int p0, int mc, const z64* psik, const z64* dpsi) {
    const KPointDev& P = K.k[ik];
    const int cols = mc * nb;
    z64* Rd = K.g_a.as<z64>(static_cast<size_t>(cols) * nnr);
    zero_grids(Rd, static_cast<long>(cols) * nnr, s);
    k_scatter<<<blocks(static_cast<long>(cols) * P.npw), NT, 0, s>>>(cols, P.npw, P.nl, dpsi, ld, Rd, nnr);
    WH_TIMED("kfft", launch_grid_fft(K.fft, Rd, cols, +1, s));
    z64* D = K.g_c.as<z64>(static_cast<size_t>(mc) * nnr);
    k_density<<<blocks(static_cast<long>(mc) * nnr), NT, 0, s>>>(nb, mc, nnr, psik, Rd, D);
    WH_TIMED("kfft", launch_grid_fft(K.fft, D, mc, -1, s));
    const double sc = 2.0 * K.wk[ik] / K.omega / static_cast<double>(nnr);
    k_gather<<<blocks(static_cast<long>(mc) * Q.npwq), NT, 0, s>>>(mc, Q.npwq, ldq, Q.nl, D, nnr, sc, nullptr,
                                                                   nullptr, nullptr, 1, dng + p0 * ldq);
    WH_HIP(hipGetLastError());
  };
  if (tr2 >= 0.0 && mp == m && kbatch_ok(K, m)) {
    // every k point at once (350-361): the states at k - q and the
    // right-hand sides of all k points on LDS grids, -P_c and M^-1 per k
    // point, one CG whose iterations take every k point's columns, then the
    // densities on LDS grids and their sum over k in k order
    const int n1 = K.fft.n1, n2 = K.fft.n2, n3 = K.fft.n3;
    const int cols = m * nb;
    const size_t all = static_cast<size_t>(K.nks) * cols;
    z64* psik = K.g_b.as<z64>(static_cast<size_t>(K.nks) * nb * nnr);
    z64* dvpsi = K.w_dv.as<z64>(all * ld);
    z64* dpsi = K.w_x.as<z64>(all * ld);
    double* d_e = K.w_ecol.as<double>(all);
    double* d_ep = K.w_epcol.as<double>(all);
    WH_TIMED("kgrid_lds", states_kmq_lds(n1, n2, n3, K.nks, nb, ld, K.ktab, Q.d_kq, psik, s));
    WH_TIMED("kgrid_lds", rhs_lds(n1, n2, n3, K.nks, nb, m, ld, K.ktab, psik, aux, dvpsi, s));
    pc_precond_all(K.nks, nb, m, ld, K.ktab, Q.d_kq, dvpsi, dpsi, d_e, d_ep, s);
    launched.push_back(linsolve_k(K, 0, K.nks, cols, dvpsi, dpsi, d_e, d_ep, tr2, nb, m, recs, d_nonconv));
    if (!K.sck) {
      std::vector<double> sk(K.nks);
      for (int ik = 0; ik < K.nks; ++ik) sk[ik] = 2.0 * K.wk[ik] / K.omega / static_cast<double>(nnr);
      K.sck = upload(K, sk.data(), sk.size());
    }
    z64* dsc = K.g_c.as<z64>(static_cast<size_t>(K.nks) * m * nnr);
    z64* dk = K.w_dk.as<z64>(static_cast<size_t>(K.nks) * m * ldq);
    WH_TIMED("kgrid_lds", density_lds(n1, n2, n3, K.nks, nb, m, ld, K.ktab, psik, dpsi, dsc, Q.npwq, ldq, Q.nl,
                                      K.sck, dk, dng, s));
  } else {
    for (int ik = 0; ik < K.nks; ++ik) {
      const z64* psik = states_kmq(ik);
      for (int p0 = 0; p0 < m; p0 += mp) {
        const int mc = std::min(mp, m - p0);
        const int cols = mc * nb;
        double* d_e = K.w_ecol.as<double>(cols);
        double* d_ep = K.w_epcol.as<double>(cols);
        z64* dvpsi = K.w_dv.as<z64>(static_cast<size_t>(cols) * ld);
        z64* dpsi = K.w_x.as<z64>(static_cast<size_t>(cols) * ld);
        rhs(ik, p0, mc, psik, dvpsi, dpsi, d_e, d_ep);
        if (tr2 >= 0.0) {                                                // (350-361)
          int* rec = recs + rec_len * launched.size();
          launched.push_back(linsolve_k(K, ik, 1, cols, dvpsi, dpsi, d_e, d_ep, tr2, nb, mc, rec,
                                        d_nonconv + p0));
        }
        density(ik, p0, mc, psik, dpsi);
      }
    }
  }
  if (Q.is_gamma) {   // (505): the perturbation acts on the body only
    k_zero_first<<<blocks(m), NT, 0, s>>>(m, ldq, dng);
    WH_HIP(hipGetLastError());
  }
  std::vector<int> nc_pert(m);
  WH_HIP(hipMemcpyAsync(nc_pert.data(), d_nonconv, sizeof(int) * m, hipMemcpyDeviceToHost, s));
  WH_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < launched.size(); ++i) cg_k_count(recs + rec_len * i, launched[i]);
  int nonconv = 0;
  for (int p = 0; p < m; ++p) {
    nonconv += nc_pert[p];
    if (ierr_pert) (*ierr_pert)[p] = nc_pert[p];
  }
  return nonconv;
}

// dng = sqvc chi0(q) (sqvc dvg)   (Wstat/apply_operator.f90:49-67)
static int apply_operator_q(KPath& K, int iq, int m, const z64* dvg, z64* dng, double tr2) {
  Context& c = ctx();
  const QPointDev& Q = K.q[iq];
  WH_REQUIRE(Q.sqvc != nullptr, "this q point has no Coulomb factor (sqvc)");
  const long ldq = K.npwqx;
  z64* dv = K.w_dn.as<z64>(static_cast<size_t>(m) * ldq);
  WH_HIP(hipMemcpyAsync(dv, dvg, sizeof(z64) * m * ldq, hipMemcpyDeviceToDevice, c.stream));
  k_scale_rows<<<blocks(static_cast<long>(m) * Q.npwq), NT, 0, c.stream>>>(m, Q.npwq, ldq, Q.sqvc, dv);
  const int nonconv = dfpt_q(K, iq, m, dv, dng, tr2, nullptr);
  k_scale_rows<<<blocks(static_cast<long>(m) * Q.npwq), NT, 0, c.stream>>>(m, Q.npwq, ldq, Q.sqvc, dng);
  WH_HIP(hipGetLastError());
  return nonconv;
}

// ---------------------------------------------------------------------------
// C-ABI pieces reached through the Gamma entry points when a k grid is set
// ---------------------------------------------------------------------------
static void finish() {
  Context& c = ctx();
  WH_HIP(hipStreamSynchronize(c.stream));
  if (c.prof.on) c.prof.flush();
}

int kpath_apply_sternheimerop(int nbndval, const double* psi, double* hpsi, const double* e, const double* alpha,
                              int m) {
  Context& c = ctx();
  KPath& K = kp();
  WH_REQUIRE(nbndval == K.nbndval, "nbndval differs from the registered states");
  WH_REQUIRE(m >= 0, "m must be >= 0");
  const int ik = K.current_k;
  WH_REQUIRE(K.k[ik].set, "the current k point has not been registered (wh_set_kpoint)");
  if (m == 0) return 0;
  double a[2] = {0.0, 0.0};
  WH_HIP(hipMemcpy(a, alpha, sizeof(double) * 2, hipMemcpyDefault));
  const size_t cin = static_cast<size_t>(m) * K.npwx;
  const z64* pin = dev_in(reinterpret_cast<const z64*>(psi), cin, K.w_in);
  const double* pe = dev_in(e, m, K.w_ecol);
  z64* t = K.w_t.as<z64>(cin);
  stern_k(K, ik, pin, t, pe, m, a[0]);
  ++c.nstern_calls;
  WH_HIP(hipMemcpyAsync(hpsi, t, sizeof(z64) * cin, hipMemcpyDefault, c.stream));
  finish();
  return 0;
}

int kpath_linsolve(int nbndval, int m, double* b, double* x, const double* e, const double* eprec, double tr2,
                   int* ierr) {
  Context& c = ctx();
  KPath& K = kp();
  WH_REQUIRE(nbndval == K.nbndval, "nbndval differs from the registered states");
  WH_REQUIRE(m >= 0, "m must be >= 0");
  const int ik = K.current_k;
  WH_REQUIRE(K.k[ik].set, "the current k point has not been registered (wh_set_kpoint)");
  if (ierr) *ierr = 0;
  if (m == 0) return 0;
  hipStream_t s = c.stream;
  const size_t cin = static_cast<size_t>(m) * K.npwx;
  z64* db = K.w_dv.as<z64>(cin);
  z64* dx = K.w_x.as<z64>(cin);
  WH_HIP(hipMemcpyAsync(db, b, sizeof(z64) * cin, hipMemcpyDefault, s));
  WH_HIP(hipMemcpyAsync(dx, x, sizeof(z64) * cin, hipMemcpyDefault, s));
  double* de = K.w_ecol.as<double>(m);
  double* dep = K.w_epcol.as<double>(m);
  WH_HIP(hipMemcpyAsync(de, e, sizeof(double) * m, hipMemcpyDefault, s));
  WH_HIP(hipMemcpyAsync(dep, eprec, sizeof(double) * m, hipMemcpyDefault, s));
  int* d_nonconv = K.w_nonconv.as<int>(1);
  WH_HIP(hipMemsetAsync(d_nonconv, 0, sizeof(int), s));
  int* rec = c.h_cgrec.as<int>(2 * (static_cast<size_t>(c.n_dfpt_maxiter) + 1));
  const int last = linsolve_k(K, ik, 1, m, db, dx, de, dep, tr2, 0, 1, rec, d_nonconv);
  int nc = 0;
  WH_HIP(hipMemcpyAsync(&nc, d_nonconv, sizeof(int), hipMemcpyDeviceToHost, s));
  WH_HIP(hipStreamSynchronize(s));
  cg_k_count(rec, last);
  WH_HIP(hipMemcpyAsync(x, dx, sizeof(z64) * cin, hipMemcpyDefault, s));
  finish();
  if (ierr) *ierr = nc;
  return nc;
}

int kpath_dfpt(int m, const double* dvg, double* dng, double tr2, int iq, int* ierr_pert, bool coulomb) {
  Context& c = ctx();
  KPath& K = kp();
  WH_REQUIRE(iq >= 1 && iq <= K.nqs, "iq out of range (1..number of q points)");
  WH_REQUIRE(m >= 0, "m must be >= 0");
  if (m == 0) return 0;
  const size_t cin = static_cast<size_t>(m) * K.npwqx;
  const z64* din = dev_in(reinterpret_cast<const z64*>(dvg), cin, K.w_in2);
  z64* dn = K.w_out.as<z64>(cin);
  std::vector<int> ierr;
  const int nonconv = coulomb ? apply_operator_q(K, iq - 1, m, din, dn, tr2)
                              : dfpt_q(K, iq - 1, m, din, dn, tr2, &ierr);
  WH_HIP(hipMemcpyAsync(dng, dn, sizeof(z64) * cin, hipMemcpyDefault, c.stream));
  finish();
  if (ierr_pert && !coulomb) std::copy(ierr.begin(), ierr.end(), ierr_pert);
  return nonconv;
}

// ---------------------------------------------------------------------------
// Davidson at one q point (davidson_diago_k, davidson_diago.f90:551-894)
// ---------------------------------------------------------------------------
namespace {

using zc = std::complex<double>;

struct ZSolver {
  rocblas_handle h = nullptr;
  ~ZSolver() {
    if (h) rocblas_destroy_handle(h);
  }
};

rocblas_handle zsolver() {
  static thread_local ZSolver s;
  if (!s.h) WH_REQUIRE(rocblas_create_handle(&s.h) == rocblas_status_success, "rocblas_create_handle failed");
  WH_REQUIRE(rocblas_set_stream(s.h, ctx().stream) == rocblas_status_success, "rocblas_set_stream failed");
  return s.h;
}

// out (na x 2 nb, column major, ld na) = [Re | Im] of A^H B over n rows:
// A^T [B, -i B] in the real view
void zbra(KPath& K, const z64* A, int na, const z64* B, int nb, int n, long ld, double* out) {
  Context& c = ctx();
  if (na <= 0 || nb <= 0) return;
  z64* W = K.dav[3].as<z64>(static_cast<size_t>(2) * nb * ld);
  k_rot_copy<<<blocks(static_cast<long>(nb) * n), NT, 0, c.stream>>>(nb, n, ld, B, W);
  WH_HIP(hipGetLastError());
  WH_TIMED("gemm_bra", dgemm_bra(na, 2 * nb, 2 * n, 1.0, reinterpret_cast<const double*>(A), 2 * ld,
                                 reinterpret_cast<const double*>(W), 2 * ld, 0.0, nullptr, out, na, c.stream));
}

// Y = alpha A (Xr + i Xi) + beta Y over n rows (X: na x nb, ld na)
void zket(KPath& K, const z64* A, int na, const double* Xr, const double* Xi, int nb, int n, long ld, double alpha,
          double beta, z64* Y) {
  Context& c = ctx();
  if (nb <= 0) return;
  WH_TIMED("gemm_ket", dgemm(false, 2 * n, nb, na, alpha, reinterpret_cast<const double*>(A), 2 * ld, Xr, na, beta,
                             reinterpret_cast<double*>(Y), 2 * ld, c.stream));
  z64* T = K.dav[4].as<z64>(static_cast<size_t>(nb) * ld);
  WH_TIMED("gemm_ket", dgemm(false, 2 * n, nb, na, 1.0, reinterpret_cast<const double*>(A), 2 * ld, Xi, na, 0.0,
                             reinterpret_cast<double*>(T), 2 * ld, c.stream));
  k_add_i<<<blocks(static_cast<long>(nb) * n), NT, 0, c.stream>>>(nb, n, ld, alpha, T, Y);
  WH_HIP(hipGetLastError());
}

// do_mgs (complex branch): columns [ms, me) against [0, ms) (two classical
// passes as GEMMs), then modified Gram-Schmidt inside the block
void do_mgs_k(KPath& K, z64* amat, int ms, int me, int n, long ld) {
  Context& c = ctx();
  const int nnew = me - ms;
  if (nnew <= 0) return;
  if (ms > 0) {
    double* cm = K.dav[5].as<double>(static_cast<size_t>(ms) * 2 * nnew);
    for (int pass = 0; pass < 2; ++pass) {
      zbra(K, amat, ms, amat + ms * ld, nnew, n, ld, cm);
      zket(K, amat, ms, cm, cm + static_cast<size_t>(ms) * nnew, nnew, n, ld, -1.0, 1.0, amat + ms * ld);
    }
  }
  double* cm = K.dav[5].as<double>(static_cast<size_t>(2) * std::max(nnew, 1));
  for (int k = ms; k < me; ++k) {
    z64* vk = amat + k * ld;
    k_znormalize<<<1, NT, 0, c.stream>>>(n, vk);
    WH_HIP(hipGetLastError());
    const int nr = me - k - 1;
    if (nr <= 0) continue;
    zbra(K, vk, 1, vk + ld, nr, n, ld, cm);
    zket(K, vk, 1, cm, cm + nr, nr, n, ld, -1.0, 1.0, vk + ld);
  }
}

}  // namespace

}  // namespace wh

using namespace wh;

#define WH_K_BEGIN try {
#define WH_K_END                            \
  }                                         \
  catch (const ::wh::Error& e) {            \
    ::wh::set_error(e.msg);                 \
    return e.code < 0 ? e.code : -1;        \
  }                                         \
  catch (const std::exception& e) {         \
    ::wh::set_error(e.what());              \
    return -3;                              \
  }

extern "C" {

int wh_set_kgrid(int nr1, int nr2, int nr3, double omega, const double* vrs, int nks, const double* wk, int npwx,
                 int nqs, int npwqx, int nbndval, int nkb, const double* dvan, double alpha_pv) {
  WH_K_BEGIN
  Context& c = ctx();
  WH_REQUIRE(nr1 > 0 && nr2 > 0 && nr3 > 0, "FFT dimensions must be positive");
  WH_REQUIRE(nks >= 1 && nqs >= 1, "need at least one k point and one q point");
  WH_REQUIRE(npwx > 0 && npwqx > 0, "npwx and npwqx must be positive");
  WH_REQUIRE(nbndval >= 0 && nkb >= 0, "nbndval and nkb must be >= 0");
  WH_REQUIRE(omega > 0.0, "omega must be positive");
  WH_REQUIRE(vrs != nullptr && wk != nullptr, "vrs and wk are required");
  WH_REQUIRE(nkb == 0 || dvan != nullptr, "dvan is required when nkb > 0");
  WH_HIP(hipStreamSynchronize(c.stream));
  free_kpath(c.kpath);
  c.kpath = nullptr;
  KPath* K = new KPath();
  try {
    K->nks = nks;
    K->nqs = nqs;
    K->npwx = npwx;
    K->npwqx = npwqx;
    K->nbndval = nbndval;
    K->nkb = nkb;
    K->omega = omega;
    K->alpha_pv = alpha_pv;
    K->wk.assign(wk, wk + nks);
    if (nkb) K->dvan.assign(dvan, dvan + nkb);
    K->fft.build(nr1, nr2, nr3);
    const long nnr = K->fft.nnr();
    std::vector<double> v(nnr);
    WH_HIP(hipMemcpy(v.data(), vrs, sizeof(double) * nnr, hipMemcpyDefault));
    K->vrs = upload(*K, v.data(), v.size());
    K->k.resize(nks);
    K->q.resize(nqs);
  } catch (...) {
    free_kpath(K);
    throw;
  }
  c.kpath = K;
  return 0;
  WH_K_END
}

int wh_set_kpoint(int ik, int npw, const int* nl, const double* g2kin, const double* evc, const double* et,
                  const double* vkb) {
  WH_K_BEGIN
  KPath& K = kp();
  WH_REQUIRE(ik >= 1 && ik <= K.nks, "ik out of range (1..nks)");
  WH_REQUIRE(npw > 0 && npw <= K.npwx, "need 0 < npw <= npwx");
  WH_REQUIRE(nl && g2kin, "nl and g2kin are required");
  WH_REQUIRE(K.nbndval == 0 || (evc && et), "evc and et are required");
  WH_REQUIRE(K.nkb == 0 || vkb, "vkb is required when nkb > 0");
  WH_HIP(hipStreamSynchronize(ctx().stream));
  KPointDev& P = K.k[ik - 1];
  for (void* p : {static_cast<void*>(P.nl), static_cast<void*>(P.g2kin), static_cast<void*>(P.q2),
                  static_cast<void*>(P.d_et), static_cast<void*>(P.d_eprec)})
    release(K, p);
  const long nnr = K.fft.nnr();
  std::vector<int> hnl(npw);
  std::vector<double> hg2(npw);
  WH_HIP(hipMemcpy(hnl.data(), nl, sizeof(int) * npw, hipMemcpyDefault));
  WH_HIP(hipMemcpy(hg2.data(), g2kin, sizeof(double) * npw, hipMemcpyDefault));
  for (int& x : hnl) {
    WH_REQUIRE(x >= 1 && x <= nnr, "nl index outside the FFT grid");
    x -= 1;
  }
  const int nb = K.nbndval;
  const int nq = nb + K.nkb;
  const size_t ld = K.npwx;
  std::vector<z64> q2(ld * 2 * std::max(nq, 1), mk(0.0, 0.0));
  P.et.assign(nb, 0.0);
  P.eprec.assign(nb, 0.0);
  if (nb) {
    std::vector<z64> h(ld * nb);
    WH_HIP(hipMemcpy(h.data(), evc, sizeof(z64) * h.size(), hipMemcpyDefault));
    WH_HIP(hipMemcpy(P.et.data(), et, sizeof(double) * nb, hipMemcpyDefault));
    for (int v = 0; v < nb; ++v) {
      double ep = 0.0;
      for (int ig = 0; ig < npw; ++ig) {
        const z64 a = h[v * ld + ig];
        q2[v * ld + ig] = a;
        q2[(nb + v) * ld + ig] = mk(-a.y, a.x);
        ep += (a.x * a.x + a.y * a.y) * hg2[ig];
      }
      P.eprec[v] = 1.35 * ep;
    }
  }
  if (K.nkb) {
    std::vector<z64> h(ld * K.nkb);
    WH_HIP(hipMemcpy(h.data(), vkb, sizeof(z64) * h.size(), hipMemcpyDefault));
    for (int j = 0; j < K.nkb; ++j)
      for (int ig = 0; ig < npw; ++ig) {
        const z64 a = h[j * ld + ig];
        q2[(2 * nb + j) * ld + ig] = a;
        q2[(2 * nb + K.nkb + j) * ld + ig] = mk(-a.y, a.x);
      }
  }
  P.npw = npw;
  P.nl = upload(K, hnl.data(), hnl.size());
  P.g2kin = upload(K, hg2.data(), hg2.size());
  P.q2 = upload(K, q2.data(), q2.size());
  P.d_et = upload(K, P.et.data(), P.et.size());
  P.d_eprec = upload(K, P.eprec.data(), P.eprec.size());
  if (!K.ktab) {
    const std::vector<KlKpoint> none(K.nks, KlKpoint{0, nullptr, nullptr, nullptr, nullptr, nullptr});
    K.ktab = upload(K, none.data(), none.size());
  }
  const KlKpoint kt{npw, P.nl, P.g2kin, reinterpret_cast<const double*>(P.q2), P.d_et, P.d_eprec};
  WH_HIP(hipMemcpy(K.ktab + (ik - 1), &kt, sizeof(kt), hipMemcpyHostToDevice));
  P.set = true;
  return 0;
  WH_K_END
}

int wh_set_qpoint(int iq, int npwq, const int* nl, const double* gq2kin, const int* ig_global, const double* sqvc,
                  const int* ikq, const int* g0) {
  WH_K_BEGIN
  KPath& K = kp();
  WH_REQUIRE(iq >= 1 && iq <= K.nqs, "iq out of range (1..number of q points)");
  WH_REQUIRE(npwq > 0 && npwq <= K.npwqx, "need 0 < npwq <= npwqx");
  WH_REQUIRE(nl && gq2kin && ig_global && ikq && g0, "nl, gq2kin, ig_global, ikq and g0 are required");
  WH_HIP(hipStreamSynchronize(ctx().stream));
  QPointDev& Q = K.q[iq - 1];
  for (void* p : {static_cast<void*>(Q.nl), static_cast<void*>(Q.gq2), static_cast<void*>(Q.ig),
                  static_cast<void*>(Q.sqvc), static_cast<void*>(Q.d_kq)})
    release(K, p);
  Q.sqvc = nullptr;
  const long nnr = K.fft.nnr();
  std::vector<int> hnl(npwq), hig(npwq);
  std::vector<double> hg2(npwq);
  WH_HIP(hipMemcpy(hnl.data(), nl, sizeof(int) * npwq, hipMemcpyDefault));
  WH_HIP(hipMemcpy(hig.data(), ig_global, sizeof(int) * npwq, hipMemcpyDefault));
  WH_HIP(hipMemcpy(hg2.data(), gq2kin, sizeof(double) * npwq, hipMemcpyDefault));
  for (int& x : hnl) {
    WH_REQUIRE(x >= 1 && x <= nnr, "nl index outside the FFT grid");
    x -= 1;
  }
  Q.ikq.assign(K.nks, 0);
  Q.g0.assign(3 * static_cast<size_t>(K.nks), 0);
  WH_HIP(hipMemcpy(Q.ikq.data(), ikq, sizeof(int) * K.nks, hipMemcpyDefault));
  WH_HIP(hipMemcpy(Q.g0.data(), g0, sizeof(int) * 3 * K.nks, hipMemcpyDefault));
  for (int& x : Q.ikq) {
    WH_REQUIRE(x >= 1 && x <= K.nks, "ikq out of range (1..nks)");
    x -= 1;
  }
  // q = 0 exactly when some q + G vanishes; it is then the first (smallest) one
  Q.is_gamma = hg2[0] < 1e-8;
  for (int i = 1; i < npwq; ++i) WH_REQUIRE(hg2[i] >= 1e-8, "only the first q + G may vanish");
  Q.npwq = npwq;
  Q.nl = upload(K, hnl.data(), hnl.size());
  Q.gq2 = upload(K, hg2.data(), hg2.size());
  Q.ig = upload(K, hig.data(), hig.size());
  std::vector<KlQpoint> kq(K.nks);
  for (int k = 0; k < K.nks; ++k) kq[k] = KlQpoint{Q.ikq[k], Q.g0[3 * k], Q.g0[3 * k + 1], Q.g0[3 * k + 2]};
  Q.d_kq = upload(K, kq.data(), kq.size());
  if (sqvc) {
    std::vector<double> hs(npwq);
    WH_HIP(hipMemcpy(hs.data(), sqvc, sizeof(double) * npwq, hipMemcpyDefault));
    Q.sqvc = upload(K, hs.data(), hs.size());
  }
  Q.set = true;
  return 0;
  WH_K_END
}

int wh_set_current_k(int ik) {
  WH_K_BEGIN
  KPath& K = kp();
  WH_REQUIRE(ik >= 1 && ik <= K.nks, "ik out of range (1..nks)");
  K.current_k = ik - 1;
  return 0;
  WH_K_END
}

int wh_davidson_diago_k(int iq, const wh_davidson_params* prm, double* ev_out, int* conv_out, double* dvg_out,
                        wh_davidson_step* log, int log_capacity, int* log_len, double* ev_log, int* conv_log) {
  WH_K_BEGIN
  Context& c = ctx();
  KPath& K = kp();
  WH_REQUIRE(prm != nullptr && ev_out != nullptr, "params and ev are required");
  WH_REQUIRE(iq >= 1 && iq <= K.nqs, "iq out of range (1..number of q points)");
  const QPointDev& Q = K.q[iq - 1];
  WH_REQUIRE(Q.set && Q.sqvc, "this q point has not been registered with its Coulomb factor");
  const int nvec = prm->n_pdep_eigen;
  const int nvecx = prm->n_pdep_eigen * prm->n_pdep_times;
  WH_REQUIRE(nvec >= 1, "n_pdep_eigen must be >= 1");
  WH_REQUIRE(nvec <= nvecx / 2, "nvecx is too small");   // (davidson_diago.f90:539)
  WH_REQUIRE(nvecx <= Q.npwq, "the basis is larger than the q sphere");
  hipStream_t s = c.stream;
  const int n = Q.npwq;
  const long ld = K.npwqx;
  const double t0 = now_s();
  double tit = t0;
  long ns0 = c.nstern_calls;
  z64* dvg = K.dav[0].as<z64>(static_cast<size_t>(nvecx) * ld);
  z64* dng = K.dav[1].as<z64>(static_cast<size_t>(nvecx) * ld);
  z64* tmp = K.dav[2].as<z64>(static_cast<size_t>(nvecx) * ld);
  WH_HIP(hipMemsetAsync(dvg, 0, sizeof(z64) * nvecx * ld, s));
  WH_HIP(hipMemsetAsync(dng, 0, sizeof(z64) * nvecx * ld, s));
  std::vector<zc> hr(static_cast<size_t>(nvecx) * nvecx, 0.0), vr(static_cast<size_t>(nvecx) * nvecx, 0.0);
  std::vector<double> ew(nvecx, 0.0), ev(nvec, 0.0);
  std::vector<char> conv(nvec, 0);
  int nbase = nvec, notcnv = nvec, nlog = 0;
  auto apply_op = [&](int j0, int cnt, double tr2) {
    if (cnt > 0) apply_operator_q(K, iq - 1, cnt, dvg + j0 * ld, dng + j0 * ld, tr2);
  };
  // hr(0:g_e, l2s:l2e) = dvg^H dng (complex build_hr)
  auto build_hr = [&](int l2s, int l2e, int g_e) {
    const int nc = l2e - l2s;
    double* d = K.dav[6].as<double>(static_cast<size_t>(g_e) * 2 * nc);
    zbra(K, dvg, g_e, dng + l2s * ld, nc, n, ld, d);
    std::vector<double> h(static_cast<size_t>(g_e) * 2 * nc);
    WH_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    WH_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nc; ++j) {
      zc* col = &hr[static_cast<size_t>(l2s + j) * nvecx];
      std::fill(col, col + nvecx, zc(0.0, 0.0));
      for (int i = 0; i < g_e; ++i)
        col[i] = zc(h[static_cast<size_t>(j) * g_e + i], h[static_cast<size_t>(nc + j) * g_e + i]);
    }
  };
  // serial_diagox_zhe: eigenpairs of hr(0:nb, 0:nb), upper triangle
  auto diagox = [&](int nb) {
    std::vector<zc> a(static_cast<size_t>(nb) * nb);
    for (int j = 0; j < nb; ++j)
      std::copy(&hr[static_cast<size_t>(j) * nvecx], &hr[static_cast<size_t>(j) * nvecx] + nb, &a[static_cast<size_t>(j) * nb]);
    rocblas_double_complex* A = K.dav[7].as<rocblas_double_complex>(a.size() + 4 * static_cast<size_t>(nb) + 2);
    double* D = reinterpret_cast<double*>(A + a.size());
    double* E = D + nb;
    rocblas_int* info = reinterpret_cast<rocblas_int*>(E + nb);
    WH_HIP(hipMemcpyAsync(A, a.data(), sizeof(zc) * a.size(), hipMemcpyHostToDevice, s));
    WH_REQUIRE(rocsolver_zheevd(zsolver(), rocblas_evect_original, rocblas_fill_upper, nb, A, nb, D, E, info) ==
                   rocblas_status_success, "rocsolver_zheevd failed");
    std::vector<double> w(nb);
    rocblas_int hinfo = 0;
    WH_HIP(hipMemcpyAsync(a.data(), A, sizeof(zc) * a.size(), hipMemcpyDeviceToHost, s));
    WH_HIP(hipMemcpyAsync(w.data(), D, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    WH_HIP(hipMemcpyAsync(&hinfo, info, sizeof(rocblas_int), hipMemcpyDeviceToHost, s));
    WH_HIP(hipStreamSynchronize(s));
    WH_REQUIRE(hinfo == 0, "zheevd did not converge");
    for (int j = 0; j < nvec; ++j) {
      ew[j] = w[j];
      std::copy(&a[static_cast<size_t>(j) * nb], &a[static_cast<size_t>(j) * nb] + nb, &vr[static_cast<size_t>(j) * nvecx]);
    }
  };
  // coefficients W (nb x ncol) as [Re | Im] device arrays
  auto upload_coef = [&](const std::vector<zc>& w, int nb, int ncol, double* dst) {
    std::vector<double> h(static_cast<size_t>(2) * nb * ncol);
    for (int j = 0; j < ncol; ++j)
      for (int i = 0; i < nb; ++i) {
        h[static_cast<size_t>(j) * nb + i] = w[static_cast<size_t>(j) * nb + i].real();
        h[static_cast<size_t>(ncol + j) * nb + i] = w[static_cast<size_t>(j) * nb + i].imag();
      }
    WH_HIP(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
  };
  auto refresh = [&](z64* X, int nb, int ncol) {   // X[0:ncol] <- X[0:nb] vr(0:nb, 0:ncol)
    std::vector<zc> w(static_cast<size_t>(nb) * ncol);
    for (int j = 0; j < ncol; ++j)
      std::copy(&vr[static_cast<size_t>(j) * nvecx], &vr[static_cast<size_t>(j) * nvecx] + nb, &w[static_cast<size_t>(j) * nb]);
    double* W = K.dav[6].as<double>(static_cast<size_t>(2) * nb * ncol);
    upload_coef(w, nb, ncol, W);
    zket(K, X, nb, W, W + static_cast<size_t>(nb) * ncol, ncol, n, ld, 1.0, 0.0, tmp);
    WH_HIP(hipMemcpyAsync(X, tmp, sizeof(z64) * ncol * ld, hipMemcpyDeviceToDevice, s));
    if (nvecx > ncol) WH_HIP(hipMemsetAsync(X + ncol * ld, 0, sizeof(z64) * (nvecx - ncol) * ld, s));
    WH_HIP(hipStreamSynchronize(s));
  };
  auto record = [&](int dav_iter, double tr2, int dfpt_dim, int diago_dim) {
    if (log && nlog < log_capacity) {
      wh_davidson_step& r = log[nlog];
      r.dav_iter = dav_iter;
      r.notcnv = notcnv;
      r.dfpt_dim = dfpt_dim;
      r.diago_dim = diago_dim;
      r.sternop_ncalls = c.nstern_calls - ns0;
      r.dfpt_tr2 = tr2;
      r.time_elap = now_s() - t0;
      r.time_iter = now_s() - tit;
    }
    if (nlog < log_capacity) {
      if (ev_log) std::memcpy(ev_log + static_cast<size_t>(nlog) * nvec, ev.data(), sizeof(double) * nvec);
      if (conv_log)
        for (int i = 0; i < nvec; ++i) conv_log[static_cast<size_t>(nlog) * nvec + i] = conv[i] ? 1 : 0;
    }
    ++nlog;
  };
  // ---- from scratch (627-660)
  k_randomize_q<<<dim3(ceil_div(n, NT), nvec), NT, 0, s>>>(n, 1, nvec, Q.ig, Q.gq2, dvg, ld);
  WH_HIP(hipGetLastError());
  do_mgs_k(K, dvg, 0, nvec, n, ld);
  apply_op(0, nvec, -1.0);
  // ---- second starting pass (671-718)
  tit = now_s();
  ns0 = c.nstern_calls;
  WH_HIP(hipMemcpyAsync(dvg, dng, sizeof(z64) * nvec * ld, hipMemcpyDeviceToDevice, s));
  do_mgs_k(K, dvg, 0, nvec, n, ld);
  double tr2 = std::min(1.0e-4, 1.0e8 * prm->tr2_dfpt);
  apply_op(0, nvec, tr2);
  build_hr(0, nvec, nvec);
  diagox(nbase);
  for (int i = 0; i < nvec; ++i) ev[i] = ew[i];
  record(0, tr2, nbase, nbase);
  int dav_iter = 0;
  bool converged = false;
  for (int kter = 1; kter <= prm->n_pdep_maxiter; ++kter) {   // (723-892)
    tit = now_s();
    ns0 = c.nstern_calls;
    ++dav_iter;
    const int dfpt_dim = notcnv;
    const int diago_dim = nbase + notcnv;
    int np = 0;
    for (int i = 0; i < nvec; ++i) {
      if (!conv[i]) {
        std::copy(&vr[static_cast<size_t>(i) * nvecx], &vr[static_cast<size_t>(i) * nvecx] + nvecx,
                  &vr[static_cast<size_t>(nbase + np) * nvecx]);
        ew[nbase + np] = ev[i];
        ++np;
      }
    }
    // update_with_vr_distr: new_j = sum_i vr(i,j) (dng_i - ew_j dvg_i)
    {
      std::vector<zc> w1(static_cast<size_t>(nbase) * notcnv), w2(w1.size());
      for (int j = 0; j < notcnv; ++j)
        for (int i = 0; i < nbase; ++i) {
          const zc v = vr[static_cast<size_t>(nbase + j) * nvecx + i];
          w1[static_cast<size_t>(j) * nbase + i] = v;
          w2[static_cast<size_t>(j) * nbase + i] = -ew[nbase + j] * v;
        }
      double* W1 = K.dav[6].as<double>(static_cast<size_t>(4) * nbase * notcnv);
      double* W2 = W1 + static_cast<size_t>(2) * nbase * notcnv;
      upload_coef(w1, nbase, notcnv, W1);
      upload_coef(w2, nbase, notcnv, W2);
      z64* newv = dvg + nbase * ld;
      zket(K, dng, nbase, W1, W1 + static_cast<size_t>(nbase) * notcnv, notcnv, n, ld, 1.0, 0.0, newv);
      zket(K, dvg, nbase, W2, W2 + static_cast<size_t>(nbase) * notcnv, notcnv, n, ld, 1.0, 1.0, newv);
      WH_HIP(hipStreamSynchronize(s));
    }
    do_mgs_k(K, dvg, nbase, nbase + notcnv, n, ld);
    tr2 = prm->tr2_dfpt;
    apply_op(nbase, notcnv, tr2);
    build_hr(nbase, nbase + notcnv, nbase + notcnv);
    nbase += notcnv;
    diagox(nbase);
    notcnv = 0;
    for (int i = 0; i < nvec; ++i) {
      conv[i] = (std::fabs((ew[i] - ev[i]) / ev[i]) < prm->trev_pdep_rel) && (std::fabs(ew[i] - ev[i]) < prm->trev_pdep);
      if (!conv[i]) ++notcnv;
    }
    for (int i = 0; i < nvec; ++i) ev[i] = ew[i];
    record(dav_iter, tr2, dfpt_dim, diago_dim);
    if (notcnv == 0 || nbase + notcnv > nvecx || kter == prm->n_pdep_maxiter) {
      if (notcnv == 0) {
        refresh(dvg, nbase, nvec);
        converged = true;
        break;
      }
      if (kter == prm->n_pdep_maxiter) break;
      refresh(dvg, nbase, nvec);   // (866-880)
      refresh(dng, nbase, nvec);
      nbase = nvec;
      std::fill(hr.begin(), hr.end(), zc(0.0, 0.0));
      std::fill(vr.begin(), vr.end(), zc(0.0, 0.0));
      for (int i = 0; i < nbase; ++i) {
        hr[static_cast<size_t>(i) * nvecx + i] = ev[i];
        vr[static_cast<size_t>(i) * nvecx + i] = 1.0;
      }
    }
  }
  (void)converged;
  std::memcpy(ev_out, ev.data(), sizeof(double) * nvec);
  if (conv_out)
    for (int i = 0; i < nvec; ++i) conv_out[i] = conv[i] ? 1 : 0;
  if (dvg_out) WH_HIP(hipMemcpyAsync(dvg_out, dvg, sizeof(z64) * nvec * ld, hipMemcpyDefault, s));
  WH_HIP(hipStreamSynchronize(s));
  if (c.prof.on) c.prof.flush();
  if (log_len) *log_len = std::min(nlog, log_capacity);
  return notcnv;
  WH_K_END
}

}  // extern "C"
