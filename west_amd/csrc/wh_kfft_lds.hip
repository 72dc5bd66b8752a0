// The q-point path on small full grids, one workgroup per band, column or
// perturbation with the whole grid in LDS:
// * the local-potential term of the Sternheimer operator at a k point,
//     t = (1/N_r) gather(F^-1 V(r) F scatter(x)) + (|k+G|^2 - e) x
//   (the same result as k_scatter, the two GridFft transforms, k_mul_real and
//   k_gather, which make six launches and six HBM round trips of the grid);
// * one CG iteration of the Sternheimer solve per column (direction, the
//   operator with its projectors, CG step, next preconditioned residual), and
//   its first step, over the columns of one k point or of every k point;
// * the states at k - q, the right-hand sides and the response densities of
//   dfpt at q for every k point, and -P_c with M^-1 on the right-hand sides.
// Each 1D pass gives one line to one thread, transformed in registers with a
// compile-time radix plan for its length; the operator's a3 pass applies V(r)
// between its inverse and forward transforms.  The LDS grid has an odd row
// pitch, so lines read in lock step by neighbouring threads fall on distinct
// banks.
#include <algorithm>

#include "wh_fft.h"
#include "wh_fft_dev.h"

namespace wh {

namespace {

constexpr int KL_NT = 512;

// 2,3,5-smooth line lengths with a compiled transform (up to 24: a line of
// 24 complex values and its radix passes fit the registers of 2 waves per SIMD)
#define WH_KL_SIZES(X) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24)

constexpr bool kl_size(int n) {
#define WH_KL_CASE(NN) n == NN ||
  return WH_KL_SIZES(WH_KL_CASE) false;
#undef WH_KL_CASE
}

template <int NL, int DIR>
__device__ __forceinline__ void kl_line(z64* p, int stride) {
  z64 v[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) v[i] = p[i * stride];
  fft_line<double, NL, DIR>(v);
#pragma unroll
  for (int i = 0; i < NL; ++i) p[i * stride] = v[i];
}

// inverse, times V(r) (vr: the line's potential, stride vs), forward
template <int NL>
__device__ __forceinline__ void kl_line_vloc(z64* p, int stride, const double* vr, long vs) {
  z64 v[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) v[i] = p[i * stride];
  fft_line<double, NL, +1>(v);
#pragma unroll
  for (int i = 0; i < NL; ++i) v[i] = cscale(v[i], vr[i * vs]);
  fft_line<double, NL, -1>(v);
#pragma unroll
  for (int i = 0; i < NL; ++i) p[i * stride] = v[i];
}

template <int DIR>
__device__ void kl_pass(int n, z64* p, int stride) {
  switch (n) {
#define WH_KL_CASE(NN)          \
  case NN:                      \
    kl_line<NN, DIR>(p, stride); \
    break;
    WH_KL_SIZES(WH_KL_CASE)
#undef WH_KL_CASE
    default: break;
  }
}

__device__ void kl_pass_vloc(int n, z64* p, int stride, const double* vr, long vs) {
  switch (n) {
#define WH_KL_CASE(NN)                    \
  case NN:                                \
    kl_line_vloc<NN>(p, stride, vr, vs);  \
    break;
    WH_KL_SIZES(WH_KL_CASE)
#undef WH_KL_CASE
    default: break;
  }
}

// LDS index of grid point (i1, i2, i3): pitch p1 (odd) along a1
__device__ __forceinline__ int kl_at(int i1, int i2, int i3, int p1, int n2) { return i1 + p1 * (i2 + n2 * i3); }

// The local term on the LDS grid: inverse passes along a1 and a2, the a3
// pass with V(r) between its inverse and forward transforms, forward passes
// along a2 and a1 (the caller's barrier before, this function's after).
__device__ __forceinline__ void kl_local(z64* grid, int n1, int n2, int n3, int p1, const double* __restrict__ vrs) {
  const int tid = threadIdx.x;
  for (int l = tid; l < n2 * n3; l += KL_NT) kl_pass<+1>(n1, grid + p1 * l, 1);
  __syncthreads();
  for (int l = tid; l < n1 * n3; l += KL_NT) {
    const int i1 = l % n1, i3 = l / n1;
    kl_pass<+1>(n2, grid + kl_at(i1, 0, i3, p1, n2), p1);
  }
  __syncthreads();
  for (int l = tid; l < n1 * n2; l += KL_NT) {
    const int i1 = l % n1, i2 = l / n1;
    kl_pass_vloc(n3, grid + kl_at(i1, i2, 0, p1, n2), p1 * n2, vrs + i1 + static_cast<long>(n1) * i2,
                 static_cast<long>(n1) * n2);
  }
  __syncthreads();
  for (int l = tid; l < n1 * n3; l += KL_NT) {
    const int i1 = l % n1, i3 = l / n1;
    kl_pass<-1>(n2, grid + kl_at(i1, 0, i3, p1, n2), p1);
  }
  __syncthreads();
  for (int l = tid; l < n2 * n3; l += KL_NT) kl_pass<-1>(n1, grid + p1 * l, 1);
  __syncthreads();
}

// 3D transform of the LDS grid (DIR +1 inverse, -1 forward, unnormalised),
// one line per thread in each of the three passes (barrier after each)
template <int DIR>
__device__ __forceinline__ void kl_transform(z64* grid, int n1, int n2, int n3, int p1) {
  const int tid = threadIdx.x;
  for (int l = tid; l < n2 * n3; l += KL_NT) kl_pass<DIR>(n1, grid + p1 * l, 1);
  __syncthreads();
  for (int l = tid; l < n1 * n3; l += KL_NT) {
    const int i1 = l % n1, i3 = l / n1;
    kl_pass<DIR>(n2, grid + kl_at(i1, 0, i3, p1, n2), p1);
  }
  __syncthreads();
  for (int l = tid; l < n1 * n2; l += KL_NT) {
    const int i1 = l % n1, i2 = l / n1;
    kl_pass<DIR>(n3, grid + kl_at(i1, i2, 0, p1, n2), p1 * n2);
  }
  __syncthreads();
}

// LDS index of the grid point with linear index r (r = i1 + n1 (i2 + n2 i3))
__device__ __forceinline__ int kl_lin(long r, int n1, int p1) {
  return static_cast<int>(r % n1) + p1 * static_cast<int>(r / n1);
}

__global__ __launch_bounds__(KL_NT) void k_kvloc_lds(int n1, int n2, int n3, int p1, int npw, long ld,
                                                   const int* __restrict__ nl, const z64* __restrict__ x,
                                                   const double* __restrict__ vrs, const double* __restrict__ g2,
                                                   const double* __restrict__ e, z64* __restrict__ t,
                                                   const int* __restrict__ d_m) {
  extern __shared__ char smem_raw[];
  z64* grid = reinterpret_cast<z64*>(smem_raw);
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  if (d_m && b >= *d_m) return;   // a band the device switched off (whole workgroup)
  const int np = p1 * n2 * n3;
  const z64* xb = x + b * ld;
  for (int i = tid; i < np; i += KL_NT) grid[i] = mk(0.0, 0.0);
  __syncthreads();
  for (int ig = tid; ig < npw; ig += KL_NT) {
    const int r = nl[ig];
    const int i1 = r % n1;
    grid[i1 + p1 * (r / n1)] = xb[ig];
  }
  __syncthreads();
  kl_local(grid, n1, n2, n3, p1, vrs);
  const double sc = 1.0 / (static_cast<double>(n1) * n2 * n3);
  const double eb = e[b];
  z64* tb = t + b * ld;
  for (int ig = tid; ig < ld; ig += KL_NT) {
    if (ig >= npw) {
      tb[ig] = mk(0.0, 0.0);
      continue;
    }
    const int r = nl[ig];
    const z64 w = grid[r % n1 + p1 * (r / n1)];
    tb[ig] = cscale(w, sc) + cscale(xb[ig], g2[ig] - eb);
  }
}

// sum over the workgroup of NV values per thread (NV <= KL_NQ2); every
// thread gets the sums in out[] (LDS)
constexpr int KL_NQ2 = 32;   // projector rows (real view) the fused CG iteration takes
__device__ __forceinline__ void kl_block_sums(double* v, int nv, double (*red)[KL_NT / 64], double* out) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < KL_NQ2; ++i) {
    if (i < nv) {
      double x = v[i];
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
      if (lane == 0) red[i][w] = x;
    }
  }
  __syncthreads();
  if (threadIdx.x < nv) {
    double x = 0.0;
#pragma unroll
    for (int q = 0; q < KL_NT / 64; ++q) x += red[threadIdx.x][q];
    out[threadIdx.x] = x;
  }
  __syncthreads();
}

// One CG iteration of the Sternheimer solve at a k point for listed column
// lb < *nc (band col = l2i[lb]), one workgroup per column, the grid in LDS:
//   h = -h + beta hold, bc = h                                (226-240)
//   t = (H_k - e + alpha P_v) h: local term on the LDS grid, |k+G|^2 - e,
//       and t += Q2 (s * (Q2^T h)) in the real view of the projectors  (251)
//   a = Re <h|g>, c = Re <h|t>; x -= a/c h, g -= a/c t, hold = h  (253-321)
//   h = g / max(1, g2kin/eprec), rho[col] = Re <h|g> (the next 150-205)
// q2: real view of the projector block (2 ld rows, nq2 columns), qs: its
// row scales.  Each thread keeps the plane waves ig = tid + k KL_NT in every
// phase, so the values it writes to bc and tc it reads back itself.
struct KlCgArgs {
  int n1, n2, n3, p1, npw;
  long ld;
  const int* nl;
  const double* vrs;
  const double* g2;
  int nq2;
  const double* q2;
  const double* qs;
  const int* l2i;
  const int* nc;
  const double* beta;
  const double* eu;
  const double* ep;
  z64* h;
  z64* hold;
  z64* x;
  z64* g;
  z64* bc;
  z64* tc;
  double* rho;
  const KlKpoint* ktab;
  int cols_per_k;
  const z64* b;   // FIRST: the right-hand sides
};

// FIRST: the CG's first step instead (127-205), every launched column in
// order (col = lb, eu = e by column): the operator acts on x, then
// g = A x - b, h = g / max(1, g2kin/eprec), rho[col] = Re <h|g>.
template <bool FIRST>
__global__ __launch_bounds__(KL_NT) void k_cg_iter_lds(KlCgArgs a) {
  extern __shared__ char smem_raw[];
  __shared__ double red[KL_NQ2][KL_NT / 64];
  __shared__ double sums[KL_NQ2];
  __shared__ double c2s[KL_NQ2];
  z64* grid = reinterpret_cast<z64*>(smem_raw);
  const int tid = threadIdx.x;
  const int lb = blockIdx.x;
  if (!FIRST && lb >= *a.nc) return;   // whole workgroup
  const int col = FIRST ? lb : a.l2i[lb];
  const long ld = a.ld;
  int npw = a.npw;
  const int* nl = a.nl;
  const double* g2 = a.g2;
  const double* q2 = a.q2;
  if (a.ktab) {   // the column's k point
    const KlKpoint kp = a.ktab[col / a.cols_per_k];
    npw = kp.npw;
    nl = kp.nl;
    g2 = kp.g2;
    q2 = kp.q2;
  }
  const int nq2 = a.nq2;
  const long ldr = 2 * ld;   // real view
  z64* hc = a.h + col * ld;
  z64* xc = a.x + col * ld;
  z64* gc = a.g + col * ld;
  z64* ho = FIRST ? nullptr : a.hold + col * ld;
  z64* bb = FIRST ? nullptr : a.bc + lb * ld;
  z64* tt = FIRST ? nullptr : a.tc + lb * ld;
  const int np = a.p1 * a.n2 * a.n3;
  for (int i = tid; i < np; i += KL_NT) grid[i] = mk(0.0, 0.0);
  __syncthreads();
  // new direction (FIRST: x), scattered into the grid; bra sums of the projectors
  const double bt = FIRST ? 0.0 : a.beta[lb];
  double acc[KL_NQ2];
#pragma unroll
  for (int i = 0; i < KL_NQ2; ++i) acc[i] = 0.0;
  for (int ig = tid; ig < ld; ig += KL_NT) {
    z64 v;
    if constexpr (FIRST) {
      v = xc[ig];
    } else {
      v = cscale(hc[ig], -1.0) + cscale(ho[ig], bt);
      hc[ig] = v;
      bb[ig] = v;
    }
    if (ig < npw) {
      const int r = nl[ig];
      grid[r % a.n1 + a.p1 * (r / a.n1)] = v;
#pragma unroll
      for (int i = 0; i < KL_NQ2; ++i)
        if (i < nq2) acc[i] += q2[i * ldr + 2 * ig] * v.x + q2[i * ldr + 2 * ig + 1] * v.y;
    }
  }
  kl_block_sums(acc, nq2, red, sums);
  if (tid < nq2) c2s[tid] = a.qs[tid] * sums[tid];
  // (kl_block_sums ends with a barrier: the grid is complete)
  kl_local(grid, a.n1, a.n2, a.n3, a.p1, a.vrs);
  // t: local + kinetic + projectors; dots of the new direction
  const double sc = 1.0 / (static_cast<double>(a.n1) * a.n2 * a.n3);
  const double eb = a.eu[lb];
  if constexpr (FIRST) {
    const z64* bcol = a.b + col * ld;
    const double e = a.ep[col];
    double rr[1] = {0.0};
    for (int ig = tid; ig < ld; ig += KL_NT) {
      if (ig >= npw) {
        gc[ig] = mk(0.0, 0.0) - bcol[ig];
        hc[ig] = mk(0.0, 0.0);
        continue;
      }
      const int r = nl[ig];
      const z64 v = xc[ig];
      z64 t = cscale(grid[r % a.n1 + a.p1 * (r / a.n1)], sc) + cscale(v, g2[ig] - eb);
      double kr = 0.0, ki = 0.0;
#pragma unroll
      for (int i = 0; i < KL_NQ2; ++i)
        if (i < nq2) {
          kr += q2[i * ldr + 2 * ig] * c2s[i];
          ki += q2[i * ldr + 2 * ig + 1] * c2s[i];
        }
      t = t + mk(kr, ki);
      const z64 gn = t - bcol[ig];
      gc[ig] = gn;
      const z64 hn = cscale(gn, 1.0 / fmax(1.0, g2[ig] / e));
      hc[ig] = hn;
      rr[0] += hn.x * gn.x + hn.y * gn.y;
    }
    kl_block_sums(rr, 1, red, sums);
    if (tid == 0) a.rho[col] = sums[0];
    return;
  } else {
  double dots[2] = {0.0, 0.0};
  for (int ig = tid; ig < ld; ig += KL_NT) {
    if (ig >= npw) {
      tt[ig] = mk(0.0, 0.0);
      continue;
    }
    const int r = nl[ig];
    const z64 v = bb[ig];
    z64 t = cscale(grid[r % a.n1 + a.p1 * (r / a.n1)], sc) + cscale(v, g2[ig] - eb);
    double kr = 0.0, ki = 0.0;
#pragma unroll
    for (int i = 0; i < KL_NQ2; ++i)
      if (i < nq2) {
        kr += q2[i * ldr + 2 * ig] * c2s[i];
        ki += q2[i * ldr + 2 * ig + 1] * c2s[i];
      }
    t = t + mk(kr, ki);
    tt[ig] = t;
    dots[0] += v.x * gc[ig].x + v.y * gc[ig].y;
    dots[1] += v.x * t.x + v.y * t.y;
  }
  kl_block_sums(dots, 2, red, sums);
  const double st = sums[0] / sums[1];
  const double e = a.ep[col];
  double rr[1] = {0.0};
  for (int ig = tid; ig < ld; ig += KL_NT) {
    const z64 v = bb[ig];
    const z64 gn = gc[ig] - cscale(tt[ig], st);
    xc[ig] = xc[ig] - cscale(v, st);
    gc[ig] = gn;
    ho[ig] = v;
    if (ig < npw) {
      const z64 hn = cscale(gn, 1.0 / fmax(1.0, g2[ig] / e));
      hc[ig] = hn;
      rr[0] += hn.x * gn.x + hn.y * gn.y;
    } else {
      hc[ig] = mk(0.0, 0.0);
    }
  }
  kl_block_sums(rr, 1, red, sums);
  if (tid == 0) a.rho[col] = sums[0];
  }
}

// -P_c and M^-1 on the right-hand side columns of every k point (no grid),
// one workgroup per column j (k point j / cols, band j % nb):
//   f <- evc (evc^H f) - f on the k sphere, x = f / max(1, g2kin/eprec), 0 beyond npw
__global__ __launch_bounds__(KL_NT) void k_pc_precond(int nb, int cols, long ld, const KlKpoint* __restrict__ ktab,
                                                    const KlQpoint* __restrict__ qtab, z64* __restrict__ f,
                                                    z64* __restrict__ x, double* __restrict__ e,
                                                    double* __restrict__ ep) {
  __shared__ double red[KL_NQ2][KL_NT / 64];
  __shared__ double sums[KL_NQ2];
  const int tid = threadIdx.x;
  const int j = blockIdx.x;
  const int kq = j / cols;
  const int v = (j - kq * cols) % nb;
  const KlKpoint pk = ktab[kq];
  const double epj = pk.eprec[v];
  if (tid == 0) {
    e[j] = ktab[qtab[kq].ikq].et[v];   // energies at k - q (dfpt_module.f90:222)
    ep[j] = epj;
  }
  const int nr = 2 * nb;   // evc and i evc in the real view
  const long ldr = 2 * ld;
  z64* fc = f + static_cast<long>(j) * ld;
  z64* xc = x + static_cast<long>(j) * ld;
  double acc[KL_NQ2];
#pragma unroll
  for (int i = 0; i < KL_NQ2; ++i) acc[i] = 0.0;
  for (int ig = tid; ig < pk.npw; ig += KL_NT) {
    const z64 fv = fc[ig];
#pragma unroll
    for (int i = 0; i < KL_NQ2; ++i)
      if (i < nr) acc[i] += pk.q2[i * ldr + 2 * ig] * fv.x + pk.q2[i * ldr + 2 * ig + 1] * fv.y;
  }
  kl_block_sums(acc, nr, red, sums);
  for (int ig = tid; ig < ld; ig += KL_NT) {
    if (ig >= pk.npw) {
      xc[ig] = mk(0.0, 0.0);
      continue;
    }
    double kr = 0.0, ki = 0.0;
#pragma unroll
    for (int i = 0; i < KL_NQ2; ++i)
      if (i < nr) {
        kr += pk.q2[i * ldr + 2 * ig] * sums[i];
        ki += pk.q2[i * ldr + 2 * ig + 1] * sums[i];
      }
    const z64 fn = mk(kr, ki) - fc[ig];
    fc[ig] = fn;
    xc[ig] = cscale(fn, 1.0 / fmax(1.0, pk.g2[ig] / epj));
  }
}

// States at k - q in real space with the phase exp(-i g0.r) for every k
// point (dfpt_module.f90:288-295), workgroup kq * nb + v:
//   psik[kq][v](r) = F^-1[evc_v at k - q](r) exp(-i g0.r)
// The phase is taken from integer numerators (Tools/types_bz_grid.f90:71-113).
__global__ __launch_bounds__(KL_NT) void k_states_lds(int n1, int n2, int n3, int p1, int nb, long ld,
                                                    const KlKpoint* __restrict__ ktab,
                                                    const KlQpoint* __restrict__ qtab, z64* __restrict__ psik) {
  extern __shared__ char smem_raw[];
  z64* grid = reinterpret_cast<z64*>(smem_raw);
  const int tid = threadIdx.x;
  const int kq = blockIdx.x / nb;
  const int v = blockIdx.x - kq * nb;
  const KlQpoint qk = qtab[kq];
  const KlKpoint pk = ktab[qk.ikq];
  const z64* evc = reinterpret_cast<const z64*>(pk.q2) + static_cast<long>(v) * ld;
  const int np = p1 * n2 * n3;
  for (int i = tid; i < np; i += KL_NT) grid[i] = mk(0.0, 0.0);
  __syncthreads();
  for (int ig = tid; ig < pk.npw; ig += KL_NT) grid[kl_lin(pk.nl[ig], n1, p1)] = evc[ig];
  __syncthreads();
  kl_transform<+1>(grid, n1, n2, n3, p1);
  const long nnr = static_cast<long>(n1) * n2 * n3;
  z64* out = psik + static_cast<long>(blockIdx.x) * nnr;
  for (long r = tid; r < nnr; r += KL_NT) {
    const int i = static_cast<int>(r % n1);
    const int j = static_cast<int>((r / n1) % n2);
    const int l = static_cast<int>(r / (static_cast<long>(n1) * n2));
    const long a1 = ((static_cast<long>(qk.g1) * i) % n1 + n1) % n1;
    const long a2 = ((static_cast<long>(qk.g2) * j) % n2 + n2) % n2;
    const long a3 = ((static_cast<long>(qk.g3) * l) % n3 + n3) % n3;
    const double x = 2.0 * (static_cast<double>(a1) / n1 + static_cast<double>(a2) / n2 + static_cast<double>(a3) / n3);
    double sn, cs;
    sincospi(x, &sn, &cs);
    out[r] = cmul(grid[kl_lin(r, n1, p1)], mk(cs, -sn));
  }
}

// Right-hand sides on the k spheres (dfpt_module.f90:295-302), workgroup
// kq * cols + c with c = p nb + v:
//   out[kq cols + c][ig] = (1/N_r) F[psik[kq][v] aux[p]](nl_kq[ig]), 0 for npw <= ig < ld
__global__ __launch_bounds__(KL_NT) void k_rhs_lds(int n1, int n2, int n3, int p1, int nb, int cols, long ld,
                                                 const KlKpoint* __restrict__ ktab, const z64* __restrict__ psik,
                                                 const z64* __restrict__ aux, z64* __restrict__ out) {
  extern __shared__ char smem_raw[];
  z64* grid = reinterpret_cast<z64*>(smem_raw);
  const int tid = threadIdx.x;
  const int kq = blockIdx.x / cols;
  const int c = blockIdx.x - kq * cols;
  const int p = c / nb;
  const int v = c - p * nb;
  const long nnr = static_cast<long>(n1) * n2 * n3;
  const z64* ps = psik + (static_cast<long>(kq) * nb + v) * nnr;
  const z64* av = aux + static_cast<long>(p) * nnr;
  for (long r = tid; r < nnr; r += KL_NT) grid[kl_lin(r, n1, p1)] = cmul(ps[r], av[r]);
  __syncthreads();
  kl_transform<-1>(grid, n1, n2, n3, p1);
  const KlKpoint pk = ktab[kq];
  const double sc = 1.0 / static_cast<double>(nnr);
  z64* o = out + static_cast<long>(blockIdx.x) * ld;
  for (int ig = tid; ig < ld; ig += KL_NT)
    o[ig] = ig < pk.npw ? cscale(grid[kl_lin(pk.nl[ig], n1, p1)], sc) : mk(0.0, 0.0);
}

// Response density of perturbation p at k point kq on the q sphere
// (dfpt_module.f90:433-441, 480-487), workgroup kq * m + p:
//   D(r) = sum_v conj(psik[kq][v](r)) F^-1[scatter_kq(dpsi[kq cols + p nb + v])](r)
//   out[kq m + p][igq] = sck[kq] F[D](nlq[igq]) for igq < npwq
// D is kept in dsc (one grid per workgroup) between the bands.
__global__ __launch_bounds__(KL_NT) void k_density_lds(int n1, int n2, int n3, int p1, int nb, int m, long ld,
                                                     const KlKpoint* __restrict__ ktab, const z64* __restrict__ psik,
                                                     const z64* __restrict__ dpsi, z64* __restrict__ dsc, int npwq,
                                                     long ldq, const int* __restrict__ nlq,
                                                     const double* __restrict__ sck, z64* __restrict__ out) {
  extern __shared__ char smem_raw[];
  z64* grid = reinterpret_cast<z64*>(smem_raw);
  const int tid = threadIdx.x;
  const int kq = blockIdx.x / m;
  const int p = blockIdx.x - kq * m;
  const KlKpoint pk = ktab[kq];
  const long nnr = static_cast<long>(n1) * n2 * n3;
  const int np = p1 * n2 * n3;
  z64* D = dsc + static_cast<long>(blockIdx.x) * nnr;
  for (int v = 0; v < nb; ++v) {
    for (int i = tid; i < np; i += KL_NT) grid[i] = mk(0.0, 0.0);
    __syncthreads();
    const z64* dp = dpsi + (static_cast<long>(kq) * m * nb + static_cast<long>(p) * nb + v) * ld;
    for (int ig = tid; ig < pk.npw; ig += KL_NT) grid[kl_lin(pk.nl[ig], n1, p1)] = dp[ig];
    __syncthreads();
    kl_transform<+1>(grid, n1, n2, n3, p1);
    const z64* ps = psik + (static_cast<long>(kq) * nb + v) * nnr;
    for (long r = tid; r < nnr; r += KL_NT) {
      const z64 t = cmul(conjg(ps[r]), grid[kl_lin(r, n1, p1)]);
      D[r] = v == 0 ? mk(0.0, 0.0) + t : D[r] + t;
    }
    __syncthreads();
  }
  // each thread wrote the D points it now loads
  for (long r = tid; r < nnr; r += KL_NT) grid[kl_lin(r, n1, p1)] = D[r];
  __syncthreads();
  kl_transform<-1>(grid, n1, n2, n3, p1);
  const double sc = sck[kq];
  z64* o = out + static_cast<long>(blockIdx.x) * ldq;
  for (int ig = tid; ig < npwq; ig += KL_NT) o[ig] = cscale(grid[kl_lin(nlq[ig], n1, p1)], sc);
}

// dng[p][igq] += sum_kq dk[kq m + p][igq] (kq in order), igq < npwq
__global__ void k_sum_k(int nks, int m, int npwq, long ldq, const z64* __restrict__ dk, z64* __restrict__ dng) {
  const long n = static_cast<long>(m) * npwq;
  for (long t = blockIdx.x * static_cast<long>(KL_NT) + threadIdx.x; t < n;
       t += static_cast<long>(gridDim.x) * KL_NT) {
    const int p = static_cast<int>(t / npwq);
    const int ig = static_cast<int>(t - static_cast<long>(p) * npwq);
    z64 acc = dng[p * ldq + ig];
    for (int kq = 0; kq < nks; ++kq) acc = acc + dk[(static_cast<long>(kq) * m + p) * ldq + ig];
    dng[p * ldq + ig] = acc;
  }
}

constexpr size_t kKlMaxSmem = 160u * 1024u;
// LDS the fused CG iteration uses besides the grid (its reduction arrays)
constexpr size_t kKlCgStatic = sizeof(double) * (KL_NQ2 * (KL_NT / 64) + 2 * KL_NQ2);

size_t kl_smem(int n1, int n2, int n3) { return sizeof(z64) * static_cast<size_t>(n1 | 1) * n2 * n3; }

}  // namespace

bool kvloc_lds_available(int n1, int n2, int n3) {
  return kl_size(n1) && kl_size(n2) && kl_size(n3) && kl_smem(n1, n2, n3) <= kKlMaxSmem;
}

bool cg_iter_lds_available(int n1, int n2, int n3, int nq2) {
  return kl_size(n1) && kl_size(n2) && kl_size(n3) && nq2 >= 0 && nq2 <= KL_NQ2 &&
         kl_smem(n1, n2, n3) + kKlCgStatic <= kKlMaxSmem;
}

namespace {
template <typename Kern>
void kl_set_smem(Kern k) {
  WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(kKlMaxSmem)));
}
bool kl_q_attrs() {   // once per process (thread-safe static initialisation)
  static const bool set = [] {
    kl_set_smem(k_states_lds);
    kl_set_smem(k_rhs_lds);
    kl_set_smem(k_density_lds);
    return true;
  }();
  return set;
}
}  // namespace

void states_kmq_lds(int n1, int n2, int n3, int nks, int nb, long ld, const KlKpoint* ktab, const KlQpoint* qtab,
                    z64* psik, hipStream_t s) {
  WH_REQUIRE(kvloc_lds_available(n1, n2, n3), "states_kmq_lds: grid not supported");
  if (nks <= 0 || nb <= 0) return;
  (void)kl_q_attrs();
  k_states_lds<<<nks * nb, KL_NT, kl_smem(n1, n2, n3), s>>>(n1, n2, n3, n1 | 1, nb, ld, ktab, qtab, psik);
  WH_HIP(hipGetLastError());
}

void rhs_lds(int n1, int n2, int n3, int nks, int nb, int m, long ld, const KlKpoint* ktab, const z64* psik,
             const z64* aux, z64* out, hipStream_t s) {
  WH_REQUIRE(kvloc_lds_available(n1, n2, n3), "rhs_lds: grid not supported");
  if (nks <= 0 || nb <= 0 || m <= 0) return;
  (void)kl_q_attrs();
  k_rhs_lds<<<nks * m * nb, KL_NT, kl_smem(n1, n2, n3), s>>>(n1, n2, n3, n1 | 1, nb, m * nb, ld, ktab, psik, aux,
                                                             out);
  WH_HIP(hipGetLastError());
}

void density_lds(int n1, int n2, int n3, int nks, int nb, int m, long ld, const KlKpoint* ktab, const z64* psik,
                 const z64* dpsi, z64* dsc, int npwq, long ldq, const int* nlq, const double* sck, z64* dk, z64* dng,
                 hipStream_t s) {
  WH_REQUIRE(kvloc_lds_available(n1, n2, n3), "density_lds: grid not supported");
  if (nks <= 0 || nb <= 0 || m <= 0) return;
  (void)kl_q_attrs();
  k_density_lds<<<nks * m, KL_NT, kl_smem(n1, n2, n3), s>>>(n1, n2, n3, n1 | 1, nb, m, ld, ktab, psik, dpsi, dsc,
                                                            npwq, ldq, nlq, sck, dk);
  WH_HIP(hipGetLastError());
  const long n = static_cast<long>(m) * npwq;
  const unsigned nblk = static_cast<unsigned>(std::min<long>((n + KL_NT - 1) / KL_NT, 1L << 16));
  k_sum_k<<<nblk, KL_NT, 0, s>>>(nks, m, npwq, ldq, dk, dng);
  WH_HIP(hipGetLastError());
}

void cg_iter_lds(int n1, int n2, int n3, int npw, long ld, const int* nl, const double* vrs, const double* g2,
                 int nq2, const double* q2, const double* qs, int cap, const int* l2i, const int* nc,
                 const double* beta, const double* eu, const double* ep, z64* h, z64* hold, z64* x, z64* g, z64* bc,
                 z64* tc, double* rho, hipStream_t s, const KlKpoint* ktab, int cols_per_k) {
  WH_REQUIRE(cg_iter_lds_available(n1, n2, n3, nq2), "cg_iter_lds: grid or projector count not supported");
  WH_REQUIRE(!ktab || cols_per_k > 0, "cg_iter_lds: cols_per_k must be > 0");
  if (cap <= 0) return;
  const size_t dyn = kl_smem(n1, n2, n3);
  static const bool attr_set = [] {   // once per process (thread-safe static initialisation)
    WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cg_iter_lds<false>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(kKlMaxSmem - kKlCgStatic)));
    return true;
  }();
  (void)attr_set;
  const KlCgArgs a{n1, n2, n3, n1 | 1, npw, ld, nl, vrs, g2, nq2, q2, qs, l2i, nc, beta, eu, ep,
                   h, hold, x, g, bc, tc, rho, ktab, cols_per_k, nullptr};
  k_cg_iter_lds<false><<<cap, KL_NT, dyn, s>>>(a);
  WH_HIP(hipGetLastError());
}

void cg_first_lds(int n1, int n2, int n3, long ld, const double* vrs, int nq2, const double* qs, int m,
                  const double* e, const double* ep, const z64* x, const z64* b, z64* g, z64* h, double* rho,
                  const KlKpoint* ktab, int cols_per_k, hipStream_t s) {
  WH_REQUIRE(cg_iter_lds_available(n1, n2, n3, nq2), "cg_first_lds: grid or projector count not supported");
  WH_REQUIRE(ktab && cols_per_k > 0, "cg_first_lds: a k-point table and cols_per_k > 0 are required");
  if (m <= 0) return;
  static const bool attr_set = [] {   // once per process (thread-safe static initialisation)
    WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cg_iter_lds<true>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(kKlMaxSmem - kKlCgStatic)));
    return true;
  }();
  (void)attr_set;
  const KlCgArgs a{n1, n2, n3, n1 | 1, 0, ld, nullptr, vrs, nullptr, nq2, nullptr, qs, nullptr, nullptr, nullptr,
                   e, ep, h, nullptr, const_cast<z64*>(x), g, nullptr, nullptr, rho, ktab, cols_per_k, b};
  k_cg_iter_lds<true><<<m, KL_NT, kl_smem(n1, n2, n3), s>>>(a);
  WH_HIP(hipGetLastError());
}

void pc_precond_all(int nks, int nb, int m, long ld, const KlKpoint* ktab, const KlQpoint* qtab, z64* f, z64* x,
                    double* e, double* ep, hipStream_t s) {
  WH_REQUIRE(nb > 0 && 2 * nb <= KL_NQ2, "pc_precond_all: 2 nbndval must be <= 32");
  if (nks <= 0 || m <= 0) return;
  k_pc_precond<<<nks * m * nb, KL_NT, 0, s>>>(nb, m * nb, ld, ktab, qtab, f, x, e, ep);
  WH_HIP(hipGetLastError());
}

void kvloc_lds(int n1, int n2, int n3, int m, int npw, long ld, const int* nl, const z64* x, const double* vrs,
               const double* g2, const double* e, z64* t, hipStream_t s, const int* d_m) {
  WH_REQUIRE(kvloc_lds_available(n1, n2, n3), "kvloc_lds: grid not supported");
  if (m <= 0) return;
  static const bool attr_set = [] {   // once per process (thread-safe static initialisation)
    WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_kvloc_lds), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(kKlMaxSmem)));
    return true;
  }();
  (void)attr_set;
  k_kvloc_lds<<<m, KL_NT, kl_smem(n1, n2, n3), s>>>(n1, n2, n3, n1 | 1, npw, ld, nl, x, vrs, g2, e, t, d_m);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
