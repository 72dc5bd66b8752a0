// Gamma-point FFT kernels on the stick layout (see wh_fft.h).
//
// Reference behaviour reproduced (FFT_kernel/fft_at_gamma.f90):
//   double_invfft_gamma 55-68 : b(nl) = a1 + i a2, then b(nlm) = conj(a1 - i a2), invfft
//   double_fwfft_gamma  98-118: fwfft, fp = (b(nl)+b(nlm))/2, fm = (b(nl)-b(nlm))/2,
//                               b1 = (Re fp, Im fm), b2 = (Im fp, -Re fm)
//   single_fwfft_gamma 185-201: b1 = fwfft(a)(nl)
// with QE's conventions: invfft = sum_G c(G) exp(+iG.r) (unnormalised),
// fwfft = (1/N) sum_r f(r) exp(-iG.r).
//
// 1D transforms are mixed-radix (2, 3, 4, 5) Stockham passes over lines held
// in LDS; each pass stages its butterflies in registers between two
// workgroup barriers, so it runs in place.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>

#include "wh_fft.h"
#include "wh_fft_reg.h"
#include "wh_fft_q4.h"
#include "wh_fft_tpl.h"

namespace wh {

namespace {

constexpr double kC3 = -0.5;
constexpr double kS3 = 0.86602540378443864676;
constexpr double kC51 = 0.30901699437494742410;
constexpr double kC52 = -0.80901699437494742410;
constexpr double kS51 = 0.95105651629515357212;
constexpr double kS52 = 0.58778525229247312917;

template <typename T, int R>
__device__ __forceinline__ void dft_small(cplx<T>* v, int dir) {
  if constexpr (R == 2) {
    cplx<T> a = v[0] + v[1];
    cplx<T> b = v[0] - v[1];
    v[0] = a;
    v[1] = b;
  } else if constexpr (R == 3) {
    cplx<T> t1 = v[1] + v[2];
    cplx<T> t2 = v[1] - v[2];
    cplx<T> m = v[0] + cscale(t1, T(kC3));
    cplx<T> u = mul_i(cscale(t2, T(kS3)), dir);
    v[0] = v[0] + t1;
    v[1] = m + u;
    v[2] = m - u;
  } else if constexpr (R == 4) {
    cplx<T> t0 = v[0] + v[2];
    cplx<T> t1 = v[0] - v[2];
    cplx<T> t2 = v[1] + v[3];
    cplx<T> t3 = v[1] - v[3];
    cplx<T> u = mul_i(t3, dir);
    v[0] = t0 + t2;
    v[2] = t0 - t2;
    v[1] = t1 + u;
    v[3] = t1 - u;
  } else {
    static_assert(R == 5, "radix");
    cplx<T> t1 = v[1] + v[4];
    cplx<T> t2 = v[2] + v[3];
    cplx<T> t3 = v[1] - v[4];
    cplx<T> t4 = v[2] - v[3];
    cplx<T> m1 = v[0] + cscale(t1, T(kC51)) + cscale(t2, T(kC52));
    cplx<T> m2 = v[0] + cscale(t1, T(kC52)) + cscale(t2, T(kC51));
    cplx<T> n1 = cscale(t3, T(kS51)) + cscale(t4, T(kS52));
    cplx<T> n2 = cscale(t3, T(kS52)) - cscale(t4, T(kS51));
    cplx<T> u1 = mul_i(n1, dir);
    cplx<T> u2 = mul_i(n2, dir);
    v[0] = v[0] + t1 + t2;
    v[1] = m1 + u1;
    v[4] = m1 - u1;
    v[2] = m2 + u2;
    v[3] = m2 - u2;
  }
}

// One Stockham pass of radix R over nlines lines of length n (line stride
// lstride) in LDS.  p = product of the radices of the previous passes.
// Requires nlines*n <= MAXE*NT.  Ends with a barrier.
template <typename T, int R, int NT, int MAXE>
__device__ __forceinline__ void fft_pass(cplx<T>* buf, int nlines, int lstride, int n, int p,
                                         const cplx<T>* tw, int dir) {
  constexpr int NB = (MAXE + R - 1) / R;
  const int nbl = n / R;
  const int nbf = nlines * nbl;
  const int step = n / (p * R);
  cplx<T> v[NB][R];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int bf = threadIdx.x + t * NT;
    if (bf < nbf) {
      const int line = bf / nbl;
      const int i = bf - line * nbl;
      const cplx<T>* src = buf + line * lstride + i;
#pragma unroll
      for (int j = 0; j < R; ++j) v[t][j] = src[j * nbl];
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int bf = threadIdx.x + t * NT;
    if (bf < nbf) {
      const int line = bf / nbl;
      const int i = bf - line * nbl;
      const int k = i % p;
      if (k != 0) {
#pragma unroll
        for (int j = 1; j < R; ++j) {
          cplx<T> w = tw[j * k * step];
          if (dir > 0) w.y = -w.y;
          v[t][j] = cmul(v[t][j], w);
        }
      }
      dft_small<T, R>(v[t], dir);
      cplx<T>* dst = buf + line * lstride + (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) dst[j * p] = v[t][j];
    }
  }
  __syncthreads();
}

// Full 1D FFT of every line; dir = +1: exp(+i) (QE invfft), -1: exp(-i).
// Caller must have synchronised the data.  Ends with a barrier.
template <typename T, int NT, int MAXE>
__device__ void lds_fft(cplx<T>* buf, int nlines, int lstride, const FftAxis& ax, const cplx<T>* tw,
                        int dir) {
  int p = 1;
  for (int s = 0; s < ax.nstage; ++s) {
    const int r = ax.radix[s];
    if (r == 4)
      fft_pass<T, 4, NT, MAXE>(buf, nlines, lstride, ax.n, p, tw, dir);
    else if (r == 2)
      fft_pass<T, 2, NT, MAXE>(buf, nlines, lstride, ax.n, p, tw, dir);
    else if (r == 3)
      fft_pass<T, 3, NT, MAXE>(buf, nlines, lstride, ax.n, p, tw, dir);
    else
      fft_pass<T, 5, NT, MAXE>(buf, nlines, lstride, ax.n, p, tw, dir);
    p *= r;
  }
}

__device__ __forceinline__ int wrap(int k, int n) {
  int z = k % n;
  return z < 0 ? z + n : z;
}

constexpr int kNT = 256;
constexpr int kCPW = 8;   // stored columns per z-kernel workgroup

// --------------------------------------------------------------------------
// sphere -> sticks
// --------------------------------------------------------------------------
template <int MAXE>
__global__ __launch_bounds__(kNT) void k_zinv(StickPlanDev P, const FftItem* items, const z64* psiA,
                                              long ldA, const z64* psiB, long ldB, z64* sticks,
                                              int mode) {
  extern __shared__ char smem_raw[];
  z64* tw = reinterpret_cast<z64*>(smem_raw);
  const int n3 = P.n3;
  z64* lines = tw + n3;
  const int it = blockIdx.y;
  const int c0 = blockIdx.x * kCPW;
  const int c1 = min(c0 + kCPW, P.ncs);
  const int ncl = c1 - c0;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const z64* A = psiA + static_cast<long>(item.a) * ldA;
  const z64* B = item.b >= 0 ? psiB + static_cast<long>(item.b) * ldB : nullptr;
  const z64* tws = reinterpret_cast<const z64*>(P.tw3);
  for (int t = threadIdx.x; t < n3; t += kNT) tw[t] = tws[t];
  for (int t = threadIdx.x; t < 2 * ncl * n3; t += kNT) lines[t] = mk(0.0, 0.0);
  __syncthreads();
  const int s0 = P.col_off[c0];
  const int s1 = P.col_off[c1];
  for (int s = s0 + threadIdx.x; s < s1; s += kNT) {
    const int lc = P.pt_col[s] - c0;
    const int k = P.pt_k[s];
    const z64 a1 = A[s];
    const z64 a2 = B ? B[s] : mk(0.0, 0.0);
    lines[(2 * lc) * n3 + wrap(k, n3)] = mk(a1.x - a2.y, a1.y + a2.x);
  }
  __syncthreads();
  for (int s = s0 + threadIdx.x; s < s1; s += kNT) {
    const int lc = P.pt_col[s] - c0;
    const int k = P.pt_k[s];
    const z64 a1 = A[s];
    const z64 a2 = B ? B[s] : mk(0.0, 0.0);
    z64 v = mk(a1.x + a2.y, a2.x - a1.y);
    if (mode == ZINV_PAIR_REALG0 && s == 0) v = mk(a1.x, a2.x);
    const int line = (c0 + lc == 0) ? 0 : 2 * lc + 1;
    lines[line * n3 + wrap(-k, n3)] = v;
  }
  __syncthreads();
  lds_fft<double, kNT, MAXE>(lines, 2 * ncl, n3, P.ax3, tw, +1);
  z64* out = sticks + static_cast<long>(it) * n3 * P.nct;
  for (int t = threadIdx.x; t < n3 * 2 * ncl; t += kNT) {
    const int z = t / (2 * ncl);
    const int r = t - z * 2 * ncl;
    const bool own = r < ncl;
    const int lc = own ? r : r - ncl;
    const int c = c0 + lc;
    if (!own && c == 0) continue;
    const int col = own ? c : c + P.ncs - 1;
    out[static_cast<long>(z) * P.nct + col] = lines[(2 * lc + (own ? 0 : 1)) * n3 + z];
  }
}

// --------------------------------------------------------------------------
// full complex grids (k points): batches of grids in QE layout (first index
// fastest).  A workgroup transforms a block of lines in LDS; along the
// second and third axes the block is nrun consecutive positions of the
// first index, so that the loads and stores are contiguous runs.
// --------------------------------------------------------------------------
constexpr int kGridE = 16;   // lines x length per workgroup <= kGridE * kNT

// lines along the first axis: contiguous, lpb lines per workgroup
__global__ __launch_bounds__(kNT) void k_grid_fft_x(z64* data, long nlines, int n, FftAxis ax, const double* tws,
                                                    int dir, int lpb) {
  extern __shared__ char smem_raw[];
  z64* tw = reinterpret_cast<z64*>(smem_raw);
  z64* buf = tw + n;
  const long l0 = static_cast<long>(blockIdx.x) * lpb;
  const int nl = static_cast<int>(min(static_cast<long>(lpb), nlines - l0));
  for (int t = threadIdx.x; t < n; t += kNT) tw[t] = reinterpret_cast<const z64*>(tws)[t];
  z64* src = data + l0 * n;
  for (int t = threadIdx.x; t < nl * n; t += kNT) buf[t] = src[t];
  __syncthreads();
  lds_fft<double, kNT, kGridE>(buf, nl, n, ax, tw, dir);
  for (int t = threadIdx.x; t < nl * n; t += kNT) src[t] = buf[t];
}

// lines of n points at stride `stride`; blockIdx.x = o selects the line
// family at offset (o / inner) * outer_stride + (o % inner) * inner_stride,
// blockIdx.y the run of nrun lines starting at first-axis position x0
__global__ __launch_bounds__(kNT) void k_grid_fft_s(z64* data, int n, long stride, int n1, int nrun, int inner,
                                                    long inner_stride, long outer_stride, FftAxis ax,
                                                    const double* tws, int dir) {
  extern __shared__ char smem_raw[];
  z64* tw = reinterpret_cast<z64*>(smem_raw);
  z64* buf = tw + n;
  const int ls = n + 1;   // LDS line stride
  const int x0 = blockIdx.y * nrun;
  const int nl = min(nrun, n1 - x0);
  const long o = blockIdx.x;
  z64* base = data + (o / inner) * outer_stride + (o % inner) * inner_stride + x0;
  for (int t = threadIdx.x; t < n; t += kNT) tw[t] = reinterpret_cast<const z64*>(tws)[t];
  for (int t = threadIdx.x; t < nl * n; t += kNT) {
    const int i = t / nl;
    const int l = t - i * nl;
    buf[l * ls + i] = base[l + i * stride];
  }
  __syncthreads();
  lds_fft<double, kNT, kGridE>(buf, nl, ls, ax, tw, dir);
  for (int t = threadIdx.x; t < nl * n; t += kNT) {
    const int i = t / nl;
    const int l = t - i * nl;
    base[l + i * stride] = buf[l * ls + i];
  }
}

// --------------------------------------------------------------------------
// per-plane 2D part: shared helpers
// --------------------------------------------------------------------------
struct PlaneSmem {
  z64* tw1;
  z64* tw2;
  z64* pl;   // nrows x n2
  z64* x;    // njc x n1
};

__device__ __forceinline__ PlaneSmem plane_smem(const StickPlanDev& P, int njc) {
  extern __shared__ char smem_raw[];
  PlaneSmem s;
  s.tw1 = reinterpret_cast<z64*>(smem_raw);
  s.tw2 = s.tw1 + P.n1;
  s.pl = s.tw2 + P.n2;
  s.x = s.pl + P.nrows * P.n2;
  return s;
}

__device__ __forceinline__ void plane_load(const StickPlanDev& P, PlaneSmem& S, const z64* src) {
  const z64* t1 = reinterpret_cast<const z64*>(P.tw1);
  const z64* t2 = reinterpret_cast<const z64*>(P.tw2);
  for (int t = threadIdx.x; t < P.n1; t += kNT) S.tw1[t] = t1[t];
  for (int t = threadIdx.x; t < P.n2; t += kNT) S.tw2[t] = t2[t];
  for (int t = threadIdx.x; t < P.nrows * P.n2; t += kNT) S.pl[t] = mk(0.0, 0.0);
  __syncthreads();
  for (int c = threadIdx.x; c < P.nct; c += kNT) S.pl[P.col_pos[c]] = src[c];
  __syncthreads();
}

// rows (|i| <= imax) of j-columns j0..j0+nj-1 -> x lines (zero padded to n1)
__device__ __forceinline__ void rows_to_x(const StickPlanDev& P, PlaneSmem& S, int j0, int nj) {
  for (int t = threadIdx.x; t < nj * P.n1; t += kNT) S.x[t] = mk(0.0, 0.0);
  __syncthreads();
  for (int t = threadIdx.x; t < P.nrows * nj; t += kNT) {
    const int r = t / nj;
    const int jj = t - r * nj;
    S.x[jj * P.n1 + wrap(r - P.imax, P.n1)] = S.pl[r * P.n2 + j0 + jj];
  }
  __syncthreads();
}

__device__ __forceinline__ void x_to_rows(const StickPlanDev& P, PlaneSmem& S, int j0, int nj) {
  for (int t = threadIdx.x; t < P.nrows * nj; t += kNT) {
    const int r = t / nj;
    const int jj = t - r * nj;
    S.pl[r * P.n2 + j0 + jj] = S.x[jj * P.n1 + wrap(r - P.imax, P.n1)];
  }
  __syncthreads();
}

// --------------------------------------------------------------------------
// sticks -> (V(r) * .) -> sticks
// --------------------------------------------------------------------------
template <int MAXE>
__global__ __launch_bounds__(kNT) void k_plane_vloc(StickPlanDev P, const FftItem* items, const z64* sin,
                                                    z64* sout, const double* vpot, long ldpot,
                                                    double scale, int njc) {
  PlaneSmem S = plane_smem(P, njc);
  const int z = blockIdx.x;
  const int it = blockIdx.y;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const long sl = static_cast<long>(P.n3) * P.nct;
  plane_load(P, S, sin + static_cast<long>(item.a) * sl + static_cast<long>(z) * P.nct);
  lds_fft<double, kNT, MAXE>(S.pl, P.nrows, P.n2, P.ax2, S.tw2, +1);
  const double* V = vpot + (item.pot >= 0 ? static_cast<long>(item.pot) * ldpot : 0L) +
                    static_cast<long>(z) * P.n1 * P.n2;
  for (int j0 = 0; j0 < P.n2; j0 += njc) {
    const int nj = min(njc, P.n2 - j0);
    rows_to_x(P, S, j0, nj);
    lds_fft<double, kNT, MAXE>(S.x, nj, P.n1, P.ax1, S.tw1, +1);
    for (int t = threadIdx.x; t < nj * P.n1; t += kNT) {
      const int jj = t / P.n1;
      const int x = t - jj * P.n1;
      S.x[t] = cscale(S.x[t], V[(j0 + jj) * P.n1 + x] * scale);
    }
    __syncthreads();
    lds_fft<double, kNT, MAXE>(S.x, nj, P.n1, P.ax1, S.tw1, -1);
    x_to_rows(P, S, j0, nj);
  }
  lds_fft<double, kNT, MAXE>(S.pl, P.nrows, P.n2, P.ax2, S.tw2, -1);
  z64* dst = sout + static_cast<long>(it) * sl + static_cast<long>(z) * P.nct;
  for (int c = threadIdx.x; c < P.nct; c += kNT) dst[c] = S.pl[P.col_pos[c]];
}

// --------------------------------------------------------------------------
// sticks -> real-space grid: Re -> row a, Im -> row b (if b >= 0)
// --------------------------------------------------------------------------
template <int MAXE>
__global__ __launch_bounds__(kNT) void k_plane_to_real(StickPlanDev P, const FftItem* items,
                                                       const z64* sin, double* out, long ldout, int njc) {
  PlaneSmem S = plane_smem(P, njc);
  const int z = blockIdx.x;
  const int it = blockIdx.y;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const long sl = static_cast<long>(P.n3) * P.nct;
  plane_load(P, S, sin + static_cast<long>(it) * sl + static_cast<long>(z) * P.nct);
  lds_fft<double, kNT, MAXE>(S.pl, P.nrows, P.n2, P.ax2, S.tw2, +1);
  double* oa = out + static_cast<long>(item.a) * ldout + static_cast<long>(z) * P.n1 * P.n2;
  double* ob = item.b >= 0 ? out + static_cast<long>(item.b) * ldout + static_cast<long>(z) * P.n1 * P.n2
                           : nullptr;
  for (int j0 = 0; j0 < P.n2; j0 += njc) {
    const int nj = min(njc, P.n2 - j0);
    rows_to_x(P, S, j0, nj);
    lds_fft<double, kNT, MAXE>(S.x, nj, P.n1, P.ax1, S.tw1, +1);
    for (int t = threadIdx.x; t < nj * P.n1; t += kNT) {
      const int jj = t / P.n1;
      const int x = t - jj * P.n1;
      oa[(j0 + jj) * P.n1 + x] = S.x[t].x;
      if (ob) ob[(j0 + jj) * P.n1 + x] = S.x[t].y;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
// response density: dn(r) = sum_v occ_v Re(psic_v) Im(psic_v), psic_v from the
// sticks of (psi_v + i dpsi_v); forward transform of dn back to sticks.
// sticks_in slot = p*nband + v; sticks_out slot = p.
// --------------------------------------------------------------------------
template <int MAXE>
__global__ __launch_bounds__(kNT) void k_plane_density(StickPlanDev P, int nband, const z64* sin,
                                                       const double* occ, z64* sout, int njc) {
  PlaneSmem S = plane_smem(P, njc);
  double* acc = reinterpret_cast<double*>(S.x + njc * P.n1);
  const int z = blockIdx.x;
  const int p = blockIdx.y;
  const long sl = static_cast<long>(P.n3) * P.nct;
  const int n12 = P.n1 * P.n2;
  for (int t = threadIdx.x; t < n12; t += kNT) acc[t] = 0.0;
  for (int v = 0; v < nband; ++v) {
    plane_load(P, S, sin + (static_cast<long>(p) * nband + v) * sl + static_cast<long>(z) * P.nct);
    lds_fft<double, kNT, MAXE>(S.pl, P.nrows, P.n2, P.ax2, S.tw2, +1);
    const double o = occ[v];
    for (int j0 = 0; j0 < P.n2; j0 += njc) {
      const int nj = min(njc, P.n2 - j0);
      rows_to_x(P, S, j0, nj);
      lds_fft<double, kNT, MAXE>(S.x, nj, P.n1, P.ax1, S.tw1, +1);
      for (int t = threadIdx.x; t < nj * P.n1; t += kNT) {
        const int jj = t / P.n1;
        const int x = t - jj * P.n1;
        acc[(j0 + jj) * P.n1 + x] += o * S.x[t].x * S.x[t].y;
      }
      __syncthreads();
    }
  }
  // forward: x then y, keeping |i| <= imax and the stick columns
  for (int j0 = 0; j0 < P.n2; j0 += njc) {
    const int nj = min(njc, P.n2 - j0);
    for (int t = threadIdx.x; t < nj * P.n1; t += kNT) {
      const int jj = t / P.n1;
      const int x = t - jj * P.n1;
      S.x[t] = mk(acc[(j0 + jj) * P.n1 + x], 0.0);
    }
    __syncthreads();
    lds_fft<double, kNT, MAXE>(S.x, nj, P.n1, P.ax1, S.tw1, -1);
    x_to_rows(P, S, j0, nj);
  }
  lds_fft<double, kNT, MAXE>(S.pl, P.nrows, P.n2, P.ax2, S.tw2, -1);
  z64* dst = sout + static_cast<long>(p) * sl + static_cast<long>(z) * P.nct;
  for (int c = threadIdx.x; c < P.nct; c += kNT) dst[c] = S.pl[P.col_pos[c]];
}

// --------------------------------------------------------------------------
// sticks -> sphere
// --------------------------------------------------------------------------
template <int MAXE>
__global__ __launch_bounds__(kNT) void k_zfwd(StickPlanDev P, const FftItem* items, const z64* sticks,
                                              const z64* in, long ldin, z64* out, long ldout, ZfwdArgs A) {
  extern __shared__ char smem_raw[];
  z64* tw = reinterpret_cast<z64*>(smem_raw);
  const int n3 = P.n3;
  z64* lines = tw + n3;
  const int it = blockIdx.y;
  const int c0 = blockIdx.x * kCPW;
  const int c1 = min(c0 + kCPW, P.ncs);
  const int ncl = c1 - c0;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const z64* tws = reinterpret_cast<const z64*>(P.tw3);
  for (int t = threadIdx.x; t < n3; t += kNT) tw[t] = tws[t];
  const z64* src = sticks + static_cast<long>(it) * n3 * P.nct;
  const bool single = A.mode == ZFWD_SINGLE;
  for (int t = threadIdx.x; t < n3 * 2 * ncl; t += kNT) {
    const int z = t / (2 * ncl);
    const int r = t - z * 2 * ncl;
    const bool own = r < ncl;
    const int lc = own ? r : r - ncl;
    const int c = c0 + lc;
    z64 v = mk(0.0, 0.0);
    if (own || (c != 0 && !single)) {
      const int col = own ? c : c + P.ncs - 1;
      v = src[static_cast<long>(z) * P.nct + col];
    }
    lines[(2 * lc + (own ? 0 : 1)) * n3 + z] = v;
  }
  __syncthreads();
  lds_fft<double, kNT, MAXE>(lines, 2 * ncl, n3, P.ax3, tw, -1);
  const int s0 = P.col_off[c0];
  const int s1 = P.col_off[c1];
  z64* oa = out + static_cast<long>(item.a) * ldout;
  z64* ob = item.b >= 0 ? out + static_cast<long>(item.b) * ldout : nullptr;
  for (int s = s0 + threadIdx.x; s < s1; s += kNT) {
    const int lc = P.pt_col[s] - c0;
    const int k = P.pt_k[s];
    const z64 F = lines[(2 * lc) * n3 + wrap(k, n3)];
    if (single) {
      double w = A.scale * (A.weight ? A.weight[s] : 1.0);
      z64 v = cscale(F, w);
      if (A.zero_g0 && s == 0) v = mk(0.0, 0.0);
      oa[s] = v;
      continue;
    }
    const int ml = (c0 + lc == 0) ? 0 : 2 * lc + 1;
    const z64 Fm = lines[ml * n3 + wrap(-k, n3)];
    const z64 fp = cscale(F + Fm, 0.5);
    const z64 fm = cscale(F - Fm, 0.5);
    const z64 b1 = mk(fp.x, fm.y);
    const z64 b2 = mk(fp.y, -fm.x);
    if (A.mode == ZFWD_PAIR_STORE) {
      oa[s] = cscale(b1, A.scale);
      if (ob) ob[s] = cscale(b2, A.scale);
    } else {
      const double g2 = A.g2kin[s];
      const z64 ia = in[static_cast<long>(item.a) * ldin + s];
      oa[s] = b1 + cscale(ia, g2 - A.e[item.a]);
      if (ob) {
        const z64 ib = in[static_cast<long>(item.b) * ldin + s];
        ob[s] = b2 + cscale(ib, g2 - A.e[item.b]);
      }
    }
  }
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
FftAxis make_axis(int n) {
  FftAxis a{};
  a.n = n;
  int r = n;
  std::vector<int> f;
  while (r % 4 == 0) { f.push_back(4); r /= 4; }
  while (r % 2 == 0) { f.push_back(2); r /= 2; }
  while (r % 3 == 0) { f.push_back(3); r /= 3; }
  while (r % 5 == 0) { f.push_back(5); r /= 5; }
  WH_REQUIRE(r == 1, "FFT dimension must factor into 2, 3 and 5");
  WH_REQUIRE(static_cast<int>(f.size()) <= kMaxStages, "FFT dimension has too many factors");
  a.nstage = static_cast<int>(f.size());
  for (int i = 0; i < a.nstage; ++i) a.radix[i] = f[i];
  return a;
}

double* upload_twiddles(int n, std::vector<void*>& allocs) {
  std::vector<double> h(2 * n);
  for (int m = 0; m < n; ++m) {
    long double ang = -2.0L * 3.141592653589793238462643383279502884L * m / n;
    h[2 * m] = static_cast<double>(cosl(ang));
    h[2 * m + 1] = static_cast<double>(sinl(ang));
  }
  double* d = nullptr;
  WH_HIP(hipMalloc(&d, sizeof(double) * 2 * n));
  WH_HIP(hipMemcpy(d, h.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  allocs.push_back(d);
  return d;
}

template <typename V>
const V* upload(const std::vector<V>& h, std::vector<void*>& allocs) {
  V* d = nullptr;
  const size_t bytes = sizeof(V) * std::max<size_t>(h.size(), 1);
  WH_HIP(hipMalloc(&d, bytes));
  if (!h.empty()) WH_HIP(hipMemcpy(d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice));
  allocs.push_back(d);
  return d;
}

int pick_maxe(long elems) {
  for (int e : {8, 16, 24, 32})
    if (elems <= static_cast<long>(e) * kNT) return e;
  throw Error{-1, "FFT plane too large for one workgroup"};
}

int plane_njc(const StickPlanDev& P) {
  // x-line chunk: bounded by the register staging of the y pass
  int nj = std::max(1, (P.nrows * P.n2) / P.n1);
  nj = std::min(nj, P.n2);
  nj = std::min(nj, 32);
  return nj;
}

size_t plane_smem_bytes(const StickPlanDev& P, int njc, bool density) {
  size_t b = sizeof(z64) * (P.n1 + P.n2 + static_cast<size_t>(P.nrows) * P.n2 + static_cast<size_t>(njc) * P.n1);
  if (density) b += sizeof(double) * static_cast<size_t>(P.n1) * P.n2;
  return b;
}

template <typename K>
void set_smem(K kernel, size_t bytes) {
  static std::map<const void*, size_t> done;
  static std::mutex done_mutex;
  std::lock_guard<std::mutex> lk(done_mutex);
  auto key = reinterpret_cast<const void*>(kernel);
  auto itr = done.find(key);
  if (itr != done.end() && itr->second >= bytes) return;
  WH_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
  done[key] = bytes;
}

}  // namespace

void StickPlan::release() {
  for (void* p : allocations) (void)hipFree(p);
  allocations.clear();
}

void StickPlan::build(int nr1, int nr2, int nr3, int npw, const int* nl, const int* nlm) {
  release();
  WH_REQUIRE(npw > 0, "npw must be positive");
  const long nnr = static_cast<long>(nr1) * nr2 * nr3;
  std::vector<int> mi(npw), mj(npw), mk_(npw);
  auto center = [](int v, int n) { return v > n / 2 ? v - n : v; };
  for (int g = 0; g < npw; ++g) {
    const long idx = static_cast<long>(nl[g]) - 1;
    WH_REQUIRE(idx >= 0 && idx < nnr, "nl index outside the FFT grid");
    const int i = static_cast<int>(idx % nr1);
    const int j = static_cast<int>((idx / nr1) % nr2);
    const int k = static_cast<int>(idx / (static_cast<long>(nr1) * nr2));
    mi[g] = center(i, nr1);
    mj[g] = center(j, nr2);
    mk_[g] = center(k, nr3);
    const long im = static_cast<long>(nlm[g]) - 1;
    const long expect = ((-mi[g] % nr1 + nr1) % nr1) + static_cast<long>(nr1) *
                        (((-mj[g] % nr2 + nr2) % nr2) + static_cast<long>(nr2) * ((-mk_[g] % nr3 + nr3) % nr3));
    WH_REQUIRE(im == expect, "nlm is not the index of -G");
    WH_REQUIRE(2 * std::abs(mi[g]) < nr1 && 2 * std::abs(mj[g]) < nr2 && 2 * std::abs(mk_[g]) < nr3,
               "G-sphere does not fit the FFT grid");
    const bool half = mi[g] > 0 || (mi[g] == 0 && mj[g] > 0) || (mi[g] == 0 && mj[g] == 0 && mk_[g] >= 0);
    WH_REQUIRE(half, "G vectors must lie in the Gamma half space (i>0 | i=0,j>0 | i=j=0,k>=0)");
  }
  // group by stored column, sorted by (i, j); k ascending inside
  std::vector<int> order(npw);
  for (int g = 0; g < npw; ++g) order[g] = g;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (mi[a] != mi[b]) return mi[a] < mi[b];
    if (mj[a] != mj[b]) return mj[a] < mj[b];
    return mk_[a] < mk_[b];
  });
  WH_REQUIRE(mi[order[0]] == 0 && mj[order[0]] == 0 && mk_[order[0]] == 0,
             "the G-sphere must contain G = 0 (Gamma point, gstart = 2)");
  col_off.clear(); col_kmin.clear(); col_i.clear(); col_j.clear();
  pt_col.assign(npw, 0); pt_k.assign(npw, 0); perm.assign(npw, 0);
  int imax = 0;
  for (int s = 0; s < npw; ++s) {
    const int g = order[s];
    const bool newcol = (s == 0) || mi[g] != mi[order[s - 1]] || mj[g] != mj[order[s - 1]];
    if (newcol) {
      col_off.push_back(s);
      col_kmin.push_back(mk_[g]);
      col_i.push_back(mi[g]);
      col_j.push_back(mj[g]);
    } else {
      WH_REQUIRE(mk_[g] == mk_[order[s - 1]] + 1, "G-sphere column is not contiguous along a3");
    }
    pt_col[s] = static_cast<int>(col_off.size()) - 1;
    pt_k[s] = mk_[g];
    perm[s] = g;
    imax = std::max(imax, std::abs(mi[g]));
  }
  const int ncs = static_cast<int>(col_off.size());
  col_off.push_back(npw);
  const int nct = 2 * ncs - 1;
  const int nrows = 2 * imax + 1;
  WH_REQUIRE(nrows <= nr1, "sphere rows exceed nr1");
  col_pos.assign(nct, 0);
  for (int c = 0; c < ncs; ++c) {
    col_pos[c] = (col_i[c] + imax) * nr2 + ((col_j[c] % nr2 + nr2) % nr2);
    if (c > 0) col_pos[c + ncs - 1] = (-col_i[c] + imax) * nr2 + (((-col_j[c]) % nr2 + nr2) % nr2);
  }
  for (int c = 0; c < nct; ++c) WH_REQUIRE(col_pos[c] >= 0 && col_pos[c] < nrows * nr2, "column position");
  StickPlanDev& d = dev;
  d.n1 = nr1; d.n2 = nr2; d.n3 = nr3;
  d.npw = npw; d.ncs = ncs; d.nct = nct; d.imax = imax; d.nrows = nrows;
  d.kext = 0;
  for (int s = 0; s < npw; ++s) d.kext = std::max(d.kext, std::abs(pt_k[s]));
  d.jext = 0;
  for (int c = 0; c < ncs; ++c) d.jext = std::max(d.jext, std::abs(col_j[c]));
  d.zblk_pts = 0;
  for (int c0 = 0; c0 < ncs; c0 += kZregCols)
    d.zblk_pts = std::max(d.zblk_pts, col_off[std::min(c0 + kZregCols, ncs)] - col_off[c0]);
  d.ax1 = make_axis(nr1);
  d.ax2 = make_axis(nr2);
  d.ax3 = make_axis(nr3);
  d.col_off = upload(col_off, allocations);
  d.col_kmin = upload(col_kmin, allocations);
  d.col_pos = upload(col_pos, allocations);
  d.pt_col = upload(pt_col, allocations);
  d.pt_k = upload(pt_k, allocations);
  d.tw1 = upload_twiddles(nr1, allocations);
  d.tw2 = upload_twiddles(nr2, allocations);
  d.tw3 = upload_twiddles(nr3, allocations);
  // the workgroup-level FFTs must fit their register staging
  pick_maxe(static_cast<long>(nrows) * nr2);
  pick_maxe(static_cast<long>(2 * kCPW) * nr3);
  // kernels specialised on the grid size, unless WH_FFT_GENERIC is set
  const char* gen = std::getenv("WH_FFT_GENERIC");
  tpl = (gen && gen[0] == '1') ? nullptr : tpl_kernels_f64(dev);
  tpl32 = tpl_kernels_f32(dev);
  // plane-kernel variant: WH_PLANE_OCC (1..9) fixes it, else timed on the first large launch
  const char* occ_env = std::getenv("WH_PLANE_OCC");
  const int occ = (occ_env && occ_env[0] >= '1' && occ_env[0] <= '9') ? occ_env[0] - '0' : 0;
  vloc_occ = occ;
  vloc32_occ = occ;
}

void launch_zinv(const StickPlan& p, const FftItem* d_items, int nitems, const z64* psiA, long ldA,
                 const z64* psiB, long ldB, z64* sticks, int mode, hipStream_t s) {
  if (nitems <= 0) return;
  const StickPlanDev& P = p.dev;
  if (p.tpl) {
    p.tpl->zinv(P, d_items, nitems, psiA, ldA, psiB, ldB, sticks, mode, s);
    WH_HIP(hipGetLastError());
    return;
  }
  const int maxe = pick_maxe(2L * kCPW * P.n3);
  const size_t smem = sizeof(z64) * (P.n3 + 2 * kCPW * P.n3);
  dim3 grid(ceil_div(P.ncs, kCPW), nitems);
  switch (maxe) {
    case 8: set_smem(k_zinv<8>, smem); k_zinv<8><<<grid, kNT, smem, s>>>(P, d_items, psiA, ldA, psiB, ldB, sticks, mode); break;
    case 16: set_smem(k_zinv<16>, smem); k_zinv<16><<<grid, kNT, smem, s>>>(P, d_items, psiA, ldA, psiB, ldB, sticks, mode); break;
    default: set_smem(k_zinv<32>, smem); k_zinv<32><<<grid, kNT, smem, s>>>(P, d_items, psiA, ldA, psiB, ldB, sticks, mode); break;
  }
  WH_HIP(hipGetLastError());
}

void launch_zinv(const StickPlan& p, const FftItem* d_items, int nitems, const z64* psi, long ldpsi,
                 z64* sticks, int mode, hipStream_t s) {
  launch_zinv(p, d_items, nitems, psi, ldpsi, psi, ldpsi, sticks, mode, s);
}

#define WH_PLANE_DISPATCH(KERNEL, SMEM, ...)                                                   \
  switch (maxe) {                                                                              \
    case 8: set_smem(KERNEL<8>, SMEM); KERNEL<8><<<grid, kNT, SMEM, s>>>(__VA_ARGS__); break;   \
    case 16: set_smem(KERNEL<16>, SMEM); KERNEL<16><<<grid, kNT, SMEM, s>>>(__VA_ARGS__); break;\
    case 24: set_smem(KERNEL<24>, SMEM); KERNEL<24><<<grid, kNT, SMEM, s>>>(__VA_ARGS__); break;\
    default: set_smem(KERNEL<32>, SMEM); KERNEL<32><<<grid, kNT, SMEM, s>>>(__VA_ARGS__); break;\
  }

void launch_plane_vloc(const StickPlan& p, const FftItem* d_items, int nitems, const z64* sin, z64* sout,
                       const double* vpot, long ldpot, double scale, hipStream_t s) {
  if (nitems <= 0) return;
  const StickPlanDev& P = p.dev;
  if (p.tpl) {
    WH_REQUIRE(static_cast<const void*>(sin) != static_cast<const void*>(sout), "plane_vloc: input and output alias");
    // variants 1..max_occ: LDS-pass kernel sized for that many workgroups per
    // CU; above: register-line kernels (2, 3, 4 planes per workgroup), then
    // the quad-lane kernels.  The untimed small launches take the first
    // quad-lane kernel where the grid has one (6-H2 Davidson, 60^3: 0.226 s
    // against 0.244 s with the register-line kernel), else the first
    // register-line kernel (235 ms against 240 ms with the LDS-pass kernel;
    // 24 clusters, concurrent: 1.35 s against 1.42 s)
    const int max_occ = p.tpl->max_occ;
    const int nreg = plane_reg_variants(P, false);
    const int nq4 = plane_q4_variants(P);
    const int untimed = nq4 > 0 ? max_occ + nreg + 1 : (nreg > 0 ? max_occ + 1 : 1);
    plane_occ_launch(
        p.vloc_occ, max_occ + nreg + nq4, nitems,
        [&](int v) {
          if (v <= max_occ)
            p.tpl->vloc(P, d_items, nitems, sin, sout, vpot, ldpot, scale, v, s);
          else if (v <= max_occ + nreg)
            plane_reg_vloc(v - max_occ - 1, P, d_items, nitems, sin, sout, vpot, ldpot, scale, s);
          else
            plane_q4_vloc(v - max_occ - nreg - 1, P, d_items, nitems, sin, sout, vpot, ldpot, scale, s);
        },
        s, untimed);
    WH_HIP(hipGetLastError());
    return;
  }
  const int njc = plane_njc(P);
  const int maxe = pick_maxe(std::max<long>(static_cast<long>(P.nrows) * P.n2, static_cast<long>(njc) * P.n1));
  const size_t smem = plane_smem_bytes(P, njc, false);
  dim3 grid(P.n3, nitems);
  WH_PLANE_DISPATCH(k_plane_vloc, smem, P, d_items, sin, sout, vpot, ldpot, scale, njc);
  WH_HIP(hipGetLastError());
}

void launch_plane_to_real(const StickPlan& p, const FftItem* d_items, int nitems, const z64* sin,
                          double* out, long ldout, hipStream_t s) {
  if (nitems <= 0) return;
  const StickPlanDev& P = p.dev;
  if (p.tpl) {
    p.tpl->to_real(P, d_items, nitems, sin, out, ldout, s);
    WH_HIP(hipGetLastError());
    return;
  }
  const int njc = plane_njc(P);
  const int maxe = pick_maxe(std::max<long>(static_cast<long>(P.nrows) * P.n2, static_cast<long>(njc) * P.n1));
  const size_t smem = plane_smem_bytes(P, njc, false);
  dim3 grid(P.n3, nitems);
  WH_PLANE_DISPATCH(k_plane_to_real, smem, P, d_items, sin, out, ldout, njc);
  WH_HIP(hipGetLastError());
}

long plane_density_scratch(const StickPlan& p) {
  return (p.tpl && p.tpl->density_scratch > 0) ? p.tpl->density_scratch : 0;
}

void launch_plane_density(const StickPlan& p, int npert, int nband, const z64* sin, const double* occ,
                          z64* sout, double* scratch, hipStream_t s) {
  if (npert <= 0) return;
  const StickPlanDev& P = p.dev;
  if (p.tpl && p.tpl->density_scratch >= 0) {
    p.tpl->density(P, npert, nband, sin, occ, sout, scratch, s);
    WH_HIP(hipGetLastError());
    return;
  }
  const int njc = plane_njc(P);
  const int maxe = pick_maxe(std::max<long>(static_cast<long>(P.nrows) * P.n2, static_cast<long>(njc) * P.n1));
  const size_t smem = plane_smem_bytes(P, njc, true);
  WH_REQUIRE(smem <= 160u * 1024u, "response-density plane does not fit in LDS for this grid in FP64");
  dim3 grid(P.n3, npert);
  WH_PLANE_DISPATCH(k_plane_density, smem, P, nband, sin, occ, sout, njc);
  WH_HIP(hipGetLastError());
}

void launch_zfwd(const StickPlan& p, const FftItem* d_items, int nitems, const z64* sticks, const z64* in,
                 long ldin, z64* out, long ldout, const ZfwdArgs& args, hipStream_t s) {
  if (nitems <= 0) return;
  const StickPlanDev& P = p.dev;
  if (p.tpl) {
    p.tpl->zfwd(P, d_items, nitems, sticks, in, ldin, out, ldout, args, s);
    WH_HIP(hipGetLastError());
    return;
  }
  const int maxe = pick_maxe(2L * kCPW * P.n3);
  const size_t smem = sizeof(z64) * (P.n3 + 2 * kCPW * P.n3);
  dim3 grid(ceil_div(P.ncs, kCPW), nitems);
  switch (maxe) {
    case 8: set_smem(k_zfwd<8>, smem); k_zfwd<8><<<grid, kNT, smem, s>>>(P, d_items, sticks, in, ldin, out, ldout, args); break;
    case 16: set_smem(k_zfwd<16>, smem); k_zfwd<16><<<grid, kNT, smem, s>>>(P, d_items, sticks, in, ldin, out, ldout, args); break;
    default: set_smem(k_zfwd<32>, smem); k_zfwd<32><<<grid, kNT, smem, s>>>(P, d_items, sticks, in, ldin, out, ldout, args); break;
  }
  WH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// full-grid transforms
// ---------------------------------------------------------------------------
void GridFft::build(int nr1, int nr2, int nr3) {
  release();
  WH_REQUIRE(nr1 > 0 && nr2 > 0 && nr3 > 0, "grid dimensions must be positive");
  WH_REQUIRE(std::max({nr1, nr2, nr3}) <= kGridE * kNT, "grid dimension too large for the grid transforms");
  n1 = nr1;
  n2 = nr2;
  n3 = nr3;
  ax1 = make_axis(nr1);
  ax2 = make_axis(nr2);
  ax3 = make_axis(nr3);
  tw1 = upload_twiddles(nr1, allocations);
  tw2 = upload_twiddles(nr2, allocations);
  tw3 = upload_twiddles(nr3, allocations);
}

void GridFft::release() {
  for (void* p : allocations) (void)hipFree(p);
  allocations.clear();
  n1 = n2 = n3 = 0;
}

void launch_grid_fft(const GridFft& g, z64* data, int batch, int dir, hipStream_t s) {
  if (batch <= 0) return;
  WH_REQUIRE(g.n1 > 0, "grid transforms used before GridFft::build");
  const long nnr = static_cast<long>(g.n1) * g.n2 * g.n3;
  {
    const int lpb = std::max(1, kGridE * kNT / g.n1);
    const long nlines = static_cast<long>(g.n2) * g.n3 * batch;
    const size_t sm = sizeof(z64) * (g.n1 + static_cast<size_t>(lpb) * g.n1);
    set_smem(k_grid_fft_x, sm);
    k_grid_fft_x<<<static_cast<unsigned>((nlines + lpb - 1) / lpb), kNT, sm, s>>>(data, nlines, g.n1, g.ax1, g.tw1,
                                                                                   dir, lpb);
    WH_HIP(hipGetLastError());
  }
  auto strided = [&](int n, long stride, int inner, long inner_stride, long outer_stride, long nfam,
                     const FftAxis& ax, const double* tw) {
    const int nrun = std::max(1, std::min(g.n1, kGridE * kNT / n));
    const size_t sm = sizeof(z64) * (n + static_cast<size_t>(nrun) * (n + 1));
    set_smem(k_grid_fft_s, sm);
    WH_REQUIRE(nfam < (1L << 31), "too many grid lines for one launch");
    k_grid_fft_s<<<dim3(static_cast<unsigned>(nfam), ceil_div(g.n1, nrun)), kNT, sm, s>>>(
        data, n, stride, g.n1, nrun, inner, inner_stride, outer_stride, ax, tw, dir);
    WH_HIP(hipGetLastError());
  };
  const long n12 = static_cast<long>(g.n1) * g.n2;
  // second axis: one family per (grid, third-axis plane), planes contiguous
  strided(g.n2, g.n1, 1 << 30, n12, 0, static_cast<long>(g.n3) * batch, g.ax2, g.tw2);
  // third axis: one family per (grid, second-axis row)
  strided(g.n3, n12, g.n2, g.n1, nnr, static_cast<long>(g.n2) * batch, g.ax3, g.tw3);
}

}  // namespace wh
