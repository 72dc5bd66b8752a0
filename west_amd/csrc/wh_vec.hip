// Vector kernels of the Sternheimer CG and the dfpt data movement (FP64).
//
// Vectors are rows of npw complex numbers in stick order (G = 0 first).
// Gamma inner products follow the reference: 2 Re sum a* b - Re a(G=0) Re b(G=0)
// (DFPT_kernel/linsolve_sternheimer_m_wfcts.f90:162-181, 253-276).
#include "wh_vec.h"

namespace wh {

namespace {

constexpr int NT = 256;

// sum over the block (BT threads): wavefront reduction, then across the
// wavefronts in a fixed order
template <int BT = NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < BT / 64; ++i) s += red[i];
  return s;
}

__device__ __forceinline__ double gamma_dot_part(const z64* a, const z64* b, int npw) {
  double s = 0.0;
  for (int g = threadIdx.x; g < npw; g += NT) s += a[g].x * b[g].x + a[g].y * b[g].y;
  return s;
}

// rho[lb] = <M^-1 g|g> with M^-1 g = g[ib] / max(1, g2kin/eprec[ib])   (linsolve 150-181);
// M^-1 g itself is formed again by k_direction (one read of g here instead of
// a read of g and a write and re-read of M^-1 g).
// V is the vector precision; the inner products are summed in FP64.
template <typename V>
__global__ __launch_bounds__(NT) void k_prec_rho(int npw, int gstart, const int* l2i, const cplx<V>* g, long ld,
                                                 const double* g2kin, const double* eprec, double* rho,
                                                 bool by_band, const int* d_nc = nullptr) {
  __shared__ double red[NT / 64];
  const int lb = blockIdx.x;
  if (d_nc && lb >= *d_nc) return;   // uniform per workgroup
  const int ib = l2i[lb];
  const cplx<V>* gi = g + static_cast<long>(ib) * ld;
  const double ep = eprec[ib];
  double s = 0.0, r0 = 0.0;
  for (int q = threadIdx.x; q < npw; q += NT) {
    const V f = V(1.0 / fmax(1.0, g2kin[q] / ep));
    const cplx<V> gv = gi[q];
    const cplx<V> hv = mk<V>(gv.x * f, gv.y * f);
    s += static_cast<double>(hv.x) * gv.x + static_cast<double>(hv.y) * gv.y;
    if (q == 0) r0 = static_cast<double>(hv.x) * gv.x;
  }
  s = block_sum(s, red);
  r0 = block_sum(r0, red);
  if (threadIdx.x == 0) rho[by_band ? ib : lb] = 2.0 * s - (gstart == 2 ? r0 : 0.0);
}

// h[ib] = -M^-1 g[ib] + beta[lb] hold[ib];  bc[lb] = h[ib]   (linsolve 150-157, 226-240)
template <typename V>
__global__ __launch_bounds__(NT) void k_direction(int npw, const int* l2i, const double* beta, const cplx<V>* g,
                                                  const double* g2kin, const double* eprec, cplx<V>* h,
                                                  const cplx<V>* hold, cplx<V>* bc, long ld, long ldc,
                                                  const int* d_nc) {
  const int lb = blockIdx.y;
  if (d_nc && lb >= *d_nc) return;
  const int ib = l2i[lb];
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= npw) return;
  const V bt = V(beta[lb]);
  const long o = static_cast<long>(ib) * ld + q;
  const V f = V(1.0 / fmax(1.0, g2kin[q] / eprec[ib]));
  const cplx<V> gv = g[o];
  cplx<V> v = mk<V>(-(gv.x * f), -(gv.y * f));
  if (bt != V(0)) {
    const cplx<V> w = hold[o];
    v = mk<V>(v.x + bt * w.x, v.y + bt * w.y);
  }
  h[o] = v;
  bc[static_cast<long>(lb) * ldc + q] = v;
}

// a = <h|g>, c = <h|t>; x -= (a/c) h; g -= (a/c) t   (linsolve 253-321)
template <typename V>
__global__ __launch_bounds__(NT) void k_cg_update(int npw, int gstart, const int* l2i, const cplx<V>* h, cplx<V>* g,
                                                  const cplx<V>* tc, cplx<V>* x, long ld, long ldc, double* ac_out,
                                                  const int* d_nc) {
  __shared__ double red[NT / 64];
  const int lb = blockIdx.x;
  if (d_nc && lb >= *d_nc) return;
  const int ib = l2i[lb];
  const cplx<V>* hi = h + static_cast<long>(ib) * ld;
  cplx<V>* gi = g + static_cast<long>(ib) * ld;
  const cplx<V>* ti = tc + static_cast<long>(lb) * ldc;
  cplx<V>* xi = x + static_cast<long>(ib) * ld;
  double sa = 0.0, sc = 0.0;
  for (int q = threadIdx.x; q < npw; q += NT) {
    const cplx<V> hv = hi[q];
    const cplx<V> gv = gi[q];
    const cplx<V> tv = ti[q];
    sa += static_cast<double>(hv.x) * gv.x + static_cast<double>(hv.y) * gv.y;
    sc += static_cast<double>(hv.x) * tv.x + static_cast<double>(hv.y) * tv.y;
  }
  sa = block_sum(sa, red);
  sc = block_sum(sc, red);
  double a = 2.0 * sa, c = 2.0 * sc;
  if (gstart == 2) {
    a -= static_cast<double>(hi[0].x) * gi[0].x;
    c -= static_cast<double>(hi[0].x) * ti[0].x;
  }
  __syncthreads();   // every thread has read gi[0] before it is updated
  const V step = V(a / c);
  if (threadIdx.x == 0 && ac_out) {
    ac_out[2 * lb] = a;
    ac_out[2 * lb + 1] = c;
  }
  for (int q = threadIdx.x; q < npw; q += NT) {
    const cplx<V> hv = hi[q];
    const cplx<V> tv = ti[q];
    cplx<V> xv = xi[q];
    cplx<V> gv = gi[q];
    xi[q] = mk<V>(xv.x - step * hv.x, xv.y - step * hv.y);
    gi[q] = mk<V>(gv.x - step * tv.x, gv.y - step * tv.y);
  }
}

// The same update with the band's h and t (and g when CG_ is set) held in
// registers between the inner products and the update (one read of each
// instead of two): 1024 threads, up to E elements each.
// With RHO the band's next rho = <M^-1 g|g> (as k_prec_rho forms it) is summed
// from the updated g while it is in registers and stored at rho[ib].
constexpr int NTU = 1024;
template <typename V, int E, bool CG_, bool RHO>
__global__ __launch_bounds__(NTU) void k_cg_update_reg(int npw, int gstart, const int* l2i, const cplx<V>* h,
                                                      cplx<V>* g, const cplx<V>* tc, cplx<V>* x, long ld, long ldc,
                                                      const double* g2kin, const double* eprec, double* rho,
                                                      const int* d_nc) {
  __shared__ double red[NTU / 64];
  const int lb = blockIdx.x;
  if (d_nc && lb >= *d_nc) return;
  const int ib = l2i[lb];
  const cplx<V>* hi = h + static_cast<long>(ib) * ld;
  cplx<V>* gi = g + static_cast<long>(ib) * ld;
  const cplx<V>* ti = tc + static_cast<long>(lb) * ldc;
  cplx<V>* xi = x + static_cast<long>(ib) * ld;
  cplx<V> hv[E], gv[CG_ ? E : 1], tv[E];
  double sa = 0.0, sc = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int q = threadIdx.x + e * NTU;
    if (q < npw) {
      hv[e] = hi[q];
      const cplx<V> gq = gi[q];
      if constexpr (CG_) gv[e] = gq;
      tv[e] = ti[q];
      sa += static_cast<double>(hv[e].x) * gq.x + static_cast<double>(hv[e].y) * gq.y;
      sc += static_cast<double>(hv[e].x) * tv[e].x + static_cast<double>(hv[e].y) * tv[e].y;
    }
  }
  sa = block_sum<NTU>(sa, red);
  sc = block_sum<NTU>(sc, red);
  double a = 2.0 * sa, c = 2.0 * sc;
  if (gstart == 2) {
    // element 0 belongs to thread 0; the block sums above end with a barrier
    a -= static_cast<double>(hi[0].x) * gi[0].x;
    c -= static_cast<double>(hi[0].x) * ti[0].x;
  }
  __syncthreads();   // every thread has read gi[0] before it is updated
  const V step = V(a / c);
  const double ep = RHO ? eprec[ib] : 1.0;
  double sr = 0.0, r0 = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int q = threadIdx.x + e * NTU;
    if (q < npw) {
      const cplx<V> xv = xi[q];
      cplx<V> gq;
      if constexpr (CG_) gq = gv[e];
      else gq = gi[q];
      xi[q] = mk<V>(xv.x - step * hv[e].x, xv.y - step * hv[e].y);
      const cplx<V> gn = mk<V>(gq.x - step * tv[e].x, gq.y - step * tv[e].y);
      gi[q] = gn;
      if constexpr (RHO) {
        const V f = V(1.0 / fmax(1.0, g2kin[q] / ep));
        const cplx<V> pv = mk<V>(gn.x * f, gn.y * f);
        const double t = static_cast<double>(pv.x) * gn.x + static_cast<double>(pv.y) * gn.y;
        sr += t;
        if (q == 0) r0 = static_cast<double>(pv.x) * gn.x;
      }
    }
  }
  if constexpr (RHO) {
    sr = block_sum<NTU>(sr, red);
    // element 0 is thread 0's
    if (threadIdx.x == 0) rho[ib] = 2.0 * sr - (gstart == 2 ? r0 : 0.0);
  }
}

// out[r][s] = in[r][perm[s]] * (w ? w[s] : 1)     caller order -> stick order
__global__ void k_gather(int m, int npw, const int* perm, const z64* in, long ldin, z64* out, long ldout,
                         const double* w) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (s >= npw || r >= m) return;
  z64 v = in[static_cast<long>(r) * ldin + perm[s]];
  if (w) v = cscale(v, w[s]);
  out[static_cast<long>(r) * ldout + s] = v;
}

// out[r][perm[s]] = in[r][s]; out[r][npw..ldout) = 0        stick order -> caller order
__global__ void k_scatter(int m, int npw, int npwx, const int* perm, const z64* in, long ldin, z64* out,
                          long ldout) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (r >= m) return;
  if (s < npw) out[static_cast<long>(r) * ldout + perm[s]] = in[static_cast<long>(r) * ldin + s];
  else if (s < npwx) out[static_cast<long>(r) * ldout + s] = mk(0.0, 0.0);
}

// pf[r] = f[r] / max(1, g2kin / eprec[r % period])
__global__ void k_precondition(int m, int npw, const z64* f, z64* pf, long ld, const double* g2kin,
                               const double* eprec, int period) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (s >= npw || r >= m) return;
  const double ep = eprec[r % period];
  const double fac = 1.0 / fmax(1.0, g2kin[s] / ep);
  pf[static_cast<long>(r) * ld + s] = cscale(f[static_cast<long>(r) * ld + s], fac);
}

// t[r] = (g2kin - e[r]) * x[r]      (k_psi - e psi)
template <typename V>
__global__ void k_kinetic(int m, int npw, const cplx<V>* x, cplx<V>* t, long ld, const double* g2kin,
                          const double* e) {
  const int r = blockIdx.y;
  const int s = blockIdx.x * NT + threadIdx.x;
  if (s >= npw || r >= m) return;
  t[static_cast<long>(r) * ld + s] = cscale(x[static_cast<long>(r) * ld + s], V(g2kin[s] - e[r]));
}

// x[ib] += dx[ib] (FP32 increment folded into the FP64 iterate); dx[ib] = 0; xc[u] = x[ib]
__global__ void k_fold(int npw, const int* u2i, z64* x, long ldx, z32* dx, long ld32, z64* xc) {
  const int u = blockIdx.y;
  const int ib = u2i[u];
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= npw) return;
  const z32 d = dx[static_cast<long>(ib) * ld32 + q];
  const long o = static_cast<long>(ib) * ldx + q;
  const z64 v = mk(x[o].x + static_cast<double>(d.x), x[o].y + static_cast<double>(d.y));
  x[o] = v;
  dx[static_cast<long>(ib) * ld32 + q] = mk(0.f, 0.f);
  xc[static_cast<long>(u) * ldx + q] = v;
}

// FP64 residual of the refinement step: r = (A x)[u] - b[ib];
// rho[u] = <M^-1 r|r> (Gamma product); the FP32 iteration restarts from g[ib] = r.
__global__ __launch_bounds__(NT) void k_true_residual(int npw, int gstart, const int* u2i, const z64* tc,
                                                      const z64* b, long ldx, const double* g2kin,
                                                      const double* eprec, z32* g, long ld32, double* rho) {
  __shared__ double red[NT / 64];
  const int u = blockIdx.x;
  const int ib = u2i[u];
  const z64* ti = tc + static_cast<long>(u) * ldx;
  const z64* bi = b + static_cast<long>(ib) * ldx;
  z32* gi = g + static_cast<long>(ib) * ld32;
  const double ep = eprec[ib];
  double s = 0.0, r0 = 0.0;
  for (int q = threadIdx.x; q < npw; q += NT) {
    const z64 r = ti[q] - bi[q];
    const double f = 1.0 / fmax(1.0, g2kin[q] / ep);
    s += f * (r.x * r.x + r.y * r.y);
    if (q == 0) r0 = f * r.x * r.x;
    gi[q] = mk(static_cast<float>(r.x), static_cast<float>(r.y));
  }
  s = block_sum(s, red);
  r0 = block_sum(r0, red);
  if (threadIdx.x == 0) rho[u] = 2.0 * s - (gstart == 2 ? r0 : 0.0);
}

__global__ void k_residual32(int npw, const z32* t, long ld32, const z64* b, long ldb, z32* g) {
  const int r = blockIdx.y;
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= npw) return;
  const z32 tv = t[static_cast<long>(r) * ld32 + q];
  const z64 bv = b[static_cast<long>(r) * ldb + q];
  g[static_cast<long>(r) * ld32 + q] = mk(static_cast<float>(static_cast<double>(tv.x) - bv.x),
                                          static_cast<float>(static_cast<double>(tv.y) - bv.y));
}

__global__ void k_to_f64(int cols, const float* in, long ld_in, double* out, long ld_out) {
  const int r = blockIdx.y;
  const long c = static_cast<long>(blockIdx.x) * NT + threadIdx.x;
  if (c >= cols) return;
  out[static_cast<long>(r) * ld_out + c] = static_cast<double>(in[static_cast<long>(r) * ld_in + c]);
}

__global__ void k_copy_rows(int npw, const int* list, const z64* in, long ld_in, z64* out, long ld_out, int scatter) {
  const int u = blockIdx.y;
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= npw) return;
  const long src = static_cast<long>(scatter ? u : list[u]) * ld_in + q;
  const long dst = static_cast<long>(scatter ? list[u] : u) * ld_out + q;
  out[dst] = in[src];
}

__global__ void k_gather_real(int n, const int* list, const double* in, double* out) {
  const int u = blockIdx.x * NT + threadIdx.x;
  if (u < n) out[u] = in[list[u]];
}

__global__ void k_to_f32(int cols, const double* in, long ld_in, float* out, long ld_out) {
  const int r = blockIdx.y;
  const long c = static_cast<long>(blockIdx.x) * NT + threadIdx.x;
  if (c >= cols) return;
  out[static_cast<long>(r) * ld_out + c] = static_cast<float>(in[static_cast<long>(r) * ld_in + c]);
}

// y = a*x + b*y over n complex
__global__ void k_axpby(long n, double a, const z64* x, double b, z64* y) {
  const long i = static_cast<long>(blockIdx.x) * NT + threadIdx.x;
  if (i >= n) return;
  const z64 xv = x[i];
  const z64 yv = y[i];
  y[i] = mk(a * xv.x + b * yv.x, a * xv.y + b * yv.y);
}

__global__ void k_fill_pairs(FftItem* items, int n, int m, int row0) {
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= n) return;
  FftItem it;
  it.a = row0 + 2 * q;
  it.b = (2 * q + 1 < m) ? row0 + 2 * q + 1 : -1;
  it.pot = -1;
  it.pad = 0;
  items[q] = it;
}

__global__ void k_fill_pairs_dev(FftItem* items, int cap, const int* nc) {
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= cap) return;
  const int m = *nc;
  FftItem it;
  it.a = (2 * q < m) ? 2 * q : -1;
  it.b = (2 * q + 1 < m) ? 2 * q + 1 : -1;
  it.pot = -1;
  it.pad = 0;
  items[q] = it;
}

__global__ void k_fill_ident_dev(FftItem* items, int cap, int chunk, const int* nc) {
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= cap) return;
  const int npair = (*nc + 1) / 2;
  FftItem it;
  it.a = q < npair ? q % chunk : -1;
  it.b = -1;
  it.pot = -1;
  it.pad = 0;
  items[q] = it;
}

__global__ void k_iota(int* v, int n) {
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q < n) v[q] = q;
}

// Convergence test, compaction and coefficients of one CG iteration on the
// device (linsolve_sternheimer_m_wfcts.f90:190-240): one workgroup walks the
// bands in order, so the active list keeps the ascending band order of
// map_lbnd2ibnd.
constexpr int NTS_SEL = 1024;
__global__ __launch_bounds__(NTS_SEL) void k_cg_select(int m, double tr2, int first, int group,
                                                       const double* __restrict__ rhob, double* __restrict__ rhoold,
                                                       int* __restrict__ conv, const double* __restrict__ e_col,
                                                       int* __restrict__ l2i, double* __restrict__ beta,
                                                       double* __restrict__ eu, int* __restrict__ nc_out,
                                                       int* __restrict__ rec) {
  __shared__ int wcount[NTS_SEL / 64];
  __shared__ int base_s;
  __shared__ int calls_w[NTS_SEL / 64];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  if (tid == 0) base_s = 0;
  __syncthreads();
  for (int c0 = 0; c0 < m; c0 += NTS_SEL) {
    const int ib = c0 + tid;
    bool act = false;
    double r = 0.0;
    if (ib < m && !conv[ib]) {
      r = rhob[ib];
      if (r < tr2) conv[ib] = 1;
      else act = true;
    }
    const unsigned long long mask = __ballot(act);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[w] = __popcll(mask);
    __syncthreads();
    int off = base_s;
    for (int k = 0; k < w; ++k) off += wcount[k];
    if (act) {
      const int pos = off + before;
      l2i[pos] = ib;
      beta[pos] = first ? 0.0 : r / rhoold[ib];
      rhoold[ib] = r;
      eu[pos] = e_col[ib];
    }
    __syncthreads();
    if (tid == 0) {
      int tot = 0;
      for (int k = 0; k < NTS_SEL / 64; ++k) tot += wcount[k];
      base_s += tot;
    }
    __syncthreads();
  }
  const int nc = base_s;
  // perturbations among the active bands (the list is ascending)
  int calls = 0;
  for (int pos = tid; pos < nc; pos += NTS_SEL) {
    const int gp = group > 0 ? l2i[pos] / group : 0;
    const int gq = pos > 0 ? (group > 0 ? l2i[pos - 1] / group : 0) : -1;
    calls += gp != gq;
  }
  for (int o = 32; o > 0; o >>= 1) calls += __shfl_xor(calls, o, 64);
  if (lane == 0) calls_w[w] = calls;
  __syncthreads();
  if (tid == 0) {
    int tot = 0;
    for (int k = 0; k < NTS_SEL / 64; ++k) tot += calls_w[k];
    *nc_out = nc;
    rec[0] = nc;
    rec[1] = tot;
  }
}

__global__ void k_fill_ident(FftItem* items, int n) {
  const int q = blockIdx.x * NT + threadIdx.x;
  if (q >= n) return;
  FftItem it;
  it.a = q;
  it.b = -1;
  it.pot = -1;
  it.pad = 0;
  items[q] = it;
}

// eprec[v] = 1.35 sum_{g<npw} |evc_v(g)|^2 g2kin(g)   (Tools/set_eprec.f90:40-51)
__global__ __launch_bounds__(NT) void k_eprec(int npw, const z64* evc, long ld, const double* g2kin, double* eprec) {
  __shared__ double red[NT / 64];
  const int v = blockIdx.x;
  const z64* e = evc + static_cast<long>(v) * ld;
  double s = 0.0;
  for (int q = threadIdx.x; q < npw; q += NT) s += (e[q].x * e[q].x + e[q].y * e[q].y) * g2kin[q];
  s = block_sum(s, red);
  if (threadIdx.x == 0) eprec[v] = 1.35 * s;
}

}  // namespace

void prec_rho(int nc, int npw, int gstart, const int* l2i, const z64* g, long ld, const double* g2kin,
              const double* eprec, double* rho, hipStream_t s, bool by_band) {
  if (nc <= 0) return;
  k_prec_rho<double><<<nc, NT, 0, s>>>(npw, gstart, l2i, g, ld, g2kin, eprec, rho, by_band);
  WH_HIP(hipGetLastError());
}

void prec_rho(int nc, int npw, int gstart, const int* l2i, const z32* g, long ld, const double* g2kin,
              const double* eprec, double* rho, hipStream_t s, bool by_band) {
  if (nc <= 0) return;
  k_prec_rho<float><<<nc, NT, 0, s>>>(npw, gstart, l2i, g, ld, g2kin, eprec, rho, by_band);
  WH_HIP(hipGetLastError());
}

void direction(int nc, int npw, const int* l2i, const double* beta, const z64* g, const double* g2kin,
               const double* eprec, z64* h, const z64* hold, z64* bc, long ld, hipStream_t s, const int* d_nc) {
  if (nc <= 0) return;
  k_direction<double><<<dim3(ceil_div(npw, NT), nc), NT, 0, s>>>(npw, l2i, beta, g, g2kin, eprec, h, hold, bc, ld,
                                                                  ld, d_nc);
  WH_HIP(hipGetLastError());
}

void direction(int nc, int npw, const int* l2i, const double* beta, const z32* g, const double* g2kin,
               const double* eprec, z32* h, const z32* hold, z32* bc, long ld, hipStream_t s) {
  if (nc <= 0) return;
  k_direction<float><<<dim3(ceil_div(npw, NT), nc), NT, 0, s>>>(npw, l2i, beta, g, g2kin, eprec, h, hold, bc, ld,
                                                                 ld, nullptr);
  WH_HIP(hipGetLastError());
}

template <typename V, bool RHO>
static void cg_update_reg(int nc, int npw, int gstart, const int* l2i, const cplx<V>* h, cplx<V>* g,
                          const cplx<V>* tc, cplx<V>* x, long ld, const RhoOut& r, hipStream_t s, const int* d_nc) {
  // FP64 at 9 elements per thread keeps h and t only (g too would spill)
  if (npw <= 4 * NTU)
    k_cg_update_reg<V, 4, true, RHO><<<nc, NTU, 0, s>>>(npw, gstart, l2i, h, g, tc, x, ld, ld, r.g2kin, r.eprec,
                                                         r.rho, d_nc);
  else
    k_cg_update_reg<V, 9, (sizeof(V) == 4), RHO><<<nc, NTU, 0, s>>>(npw, gstart, l2i, h, g, tc, x, ld, ld, r.g2kin,
                                                                    r.eprec, r.rho, d_nc);
}

template <typename V>
static void cg_update_any(int nc, int npw, int gstart, const int* l2i, const cplx<V>* h, cplx<V>* g,
                          const cplx<V>* tc, cplx<V>* x, long ld, double* ac_out, const RhoOut* r, hipStream_t s,
                          const int* d_nc) {
  if (nc <= 0) return;
  if (!ac_out && npw <= 9 * NTU) {
    if (r) cg_update_reg<V, true>(nc, npw, gstart, l2i, h, g, tc, x, ld, *r, s, d_nc);
    else cg_update_reg<V, false>(nc, npw, gstart, l2i, h, g, tc, x, ld, RhoOut{}, s, d_nc);
  } else {
    k_cg_update<V><<<nc, NT, 0, s>>>(npw, gstart, l2i, h, g, tc, x, ld, ld, ac_out, d_nc);
    if (r) k_prec_rho<V><<<nc, NT, 0, s>>>(npw, gstart, l2i, g, ld, r->g2kin, r->eprec, r->rho, true, d_nc);
  }
  WH_HIP(hipGetLastError());
}

void cg_update(int nc, int npw, int gstart, const int* l2i, const z64* h, z64* g, const z64* tc, z64* x, long ld,
               double* ac_out, hipStream_t s, const RhoOut* rho, const int* d_nc) {
  cg_update_any<double>(nc, npw, gstart, l2i, h, g, tc, x, ld, ac_out, rho, s, d_nc);
}

void cg_update(int nc, int npw, int gstart, const int* l2i, const z32* h, z32* g, const z32* tc, z32* x, long ld,
               double* ac_out, hipStream_t s, const RhoOut* rho) {
  cg_update_any<float>(nc, npw, gstart, l2i, h, g, tc, x, ld, ac_out, rho, s, nullptr);
}

void gather_sticks(int m, int npw, const int* perm, const z64* in, long ldin, z64* out, long ldout, const double* w,
                   hipStream_t s) {
  if (m <= 0) return;
  k_gather<<<dim3(ceil_div(npw, NT), m), NT, 0, s>>>(m, npw, perm, in, ldin, out, ldout, w);
  WH_HIP(hipGetLastError());
}

void scatter_caller(int m, int npw, int npwx, const int* perm, const z64* in, long ldin, z64* out, long ldout,
                    hipStream_t s) {
  if (m <= 0) return;
  k_scatter<<<dim3(ceil_div(npwx, NT), m), NT, 0, s>>>(m, npw, npwx, perm, in, ldin, out, ldout);
  WH_HIP(hipGetLastError());
}

void precondition(int m, int npw, const z64* f, z64* pf, long ld, const double* g2kin, const double* eprec,
                  int period, hipStream_t s) {
  if (m <= 0) return;
  k_precondition<<<dim3(ceil_div(npw, NT), m), NT, 0, s>>>(m, npw, f, pf, ld, g2kin, eprec, period);
  WH_HIP(hipGetLastError());
}

void kinetic_minus_e(int m, int npw, const z64* x, z64* t, long ld, const double* g2kin, const double* e,
                     hipStream_t s) {
  if (m <= 0) return;
  k_kinetic<double><<<dim3(ceil_div(npw, NT), m), NT, 0, s>>>(m, npw, x, t, ld, g2kin, e);
  WH_HIP(hipGetLastError());
}

void kinetic_minus_e(int m, int npw, const z32* x, z32* t, long ld, const double* g2kin, const double* e,
                     hipStream_t s) {
  if (m <= 0) return;
  k_kinetic<float><<<dim3(ceil_div(npw, NT), m), NT, 0, s>>>(m, npw, x, t, ld, g2kin, e);
  WH_HIP(hipGetLastError());
}

void fold_increment(int nu, int npw, const int* u2i, z64* x, long ldx, z32* dx, long ld32, z64* xc, hipStream_t s) {
  if (nu <= 0) return;
  k_fold<<<dim3(ceil_div(npw, NT), nu), NT, 0, s>>>(npw, u2i, x, ldx, dx, ld32, xc);
  WH_HIP(hipGetLastError());
}

void true_residual(int nu, int npw, int gstart, const int* u2i, const z64* tc, const z64* b, long ldx,
                   const double* g2kin, const double* eprec, z32* g, long ld32, double* rho, hipStream_t s) {
  if (nu <= 0) return;
  k_true_residual<<<nu, NT, 0, s>>>(npw, gstart, u2i, tc, b, ldx, g2kin, eprec, g, ld32, rho);
  WH_HIP(hipGetLastError());
}

void residual32(int m, int npw, const z32* t, long ld32, const z64* b, long ldb, z32* g, hipStream_t s) {
  if (m <= 0) return;
  k_residual32<<<dim3(ceil_div(npw, NT), m), NT, 0, s>>>(npw, t, ld32, b, ldb, g);
  WH_HIP(hipGetLastError());
}

void to_f64(int rows, int cols, const float* in, long ld_in, double* out, long ld_out, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  k_to_f64<<<dim3(ceil_div(cols, NT), rows), NT, 0, s>>>(cols, in, ld_in, out, ld_out);
  WH_HIP(hipGetLastError());
}

void copy_rows(int n, int npw, const int* list, const z64* in, long ld_in, z64* out, long ld_out, bool scatter,
               hipStream_t s) {
  if (n <= 0) return;
  k_copy_rows<<<dim3(ceil_div(npw, NT), n), NT, 0, s>>>(npw, list, in, ld_in, out, ld_out, scatter ? 1 : 0);
  WH_HIP(hipGetLastError());
}

void gather_real(int n, const int* list, const double* in, double* out, hipStream_t s) {
  if (n <= 0) return;
  k_gather_real<<<ceil_div(n, NT), NT, 0, s>>>(n, list, in, out);
  WH_HIP(hipGetLastError());
}

void to_f32(int rows, int cols, const double* in, long ld_in, float* out, long ld_out, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  k_to_f32<<<dim3(ceil_div(cols, NT), rows), NT, 0, s>>>(cols, in, ld_in, out, ld_out);
  WH_HIP(hipGetLastError());
}

void axpby(long n, double a, const z64* x, double b, z64* y, hipStream_t s) {
  if (n <= 0) return;
  k_axpby<<<ceil_div(n, NT), NT, 0, s>>>(n, a, x, b, y);
  WH_HIP(hipGetLastError());
}

void fill_pairs(FftItem* items, int m, int row0, hipStream_t s) {
  const int n = (m + 1) / 2;
  if (n <= 0) return;
  k_fill_pairs<<<ceil_div(n, NT), NT, 0, s>>>(items, n, m, row0);
  WH_HIP(hipGetLastError());
}

void fill_pairs_dev(FftItem* items, int cap, const int* nc, hipStream_t s) {
  if (cap <= 0) return;
  k_fill_pairs_dev<<<ceil_div(cap, NT), NT, 0, s>>>(items, cap, nc);
  WH_HIP(hipGetLastError());
}

void fill_ident_dev(FftItem* items, int cap, int chunk, const int* nc, hipStream_t s) {
  if (cap <= 0) return;
  k_fill_ident_dev<<<ceil_div(cap, NT), NT, 0, s>>>(items, cap, chunk, nc);
  WH_HIP(hipGetLastError());
}

void iota(int* v, int n, hipStream_t s) {
  if (n <= 0) return;
  k_iota<<<ceil_div(n, NT), NT, 0, s>>>(v, n);
  WH_HIP(hipGetLastError());
}

void cg_select(int m, double tr2, int first, int group, const double* rhob, double* rhoold, int* conv,
               const double* e_col, int* l2i, double* beta, double* eu, int* nc, int* rec, hipStream_t s) {
  k_cg_select<<<1, NTS_SEL, 0, s>>>(m, tr2, first, group, rhob, rhoold, conv, e_col, l2i, beta, eu, nc, rec);
  WH_HIP(hipGetLastError());
}

void fill_ident(FftItem* items, int n, hipStream_t s) {
  if (n <= 0) return;
  k_fill_ident<<<ceil_div(n, NT), NT, 0, s>>>(items, n);
  WH_HIP(hipGetLastError());
}

void eprec_of(int nb, int npw, const z64* evc, long ld, const double* g2kin, double* eprec, hipStream_t s) {
  if (nb <= 0) return;
  k_eprec<<<nb, NT, 0, s>>>(npw, evc, ld, g2kin, eprec);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
