// Gamma-point FFT kernels with compile-time line length (cubic grids).
//
// Same algorithm and data layout as the runtime-size kernels of wh_fft.hip
// (stick layout, zinv -> plane -> zfwd), specialised on N = nr1 = nr2 = nr3:
//  * each 1D transform is a Stockham sequence of large radices chosen at
//    compile time (N = 72 = 9 x 8: two LDS passes instead of four), with
//    composite radices 6, 8, 9, 10, 12 done in registers as two-level
//    Cooley-Tukey butterflies with constant twiddles;
//  * all index arithmetic is by compile-time constants;
//  * in the plane kernel the zero padding of the a1 transform is folded into
//    the first pass (loads from the |i| <= imax rows), the real-space
//    potential multiply into the last inverse pass, and the |i| <= imax
//    selection into the last forward pass, so the padded lines are never
//    materialised and no separate scatter/gather sweeps remain.
#include <map>
#include <mutex>

#include "wh_fft.h"
#include "wh_fft_dev.h"
#include "wh_fft_tpl.h"

namespace wh {

namespace {

constexpr int NT = 256;
// stored columns per workgroup of the z-kernels: one z row of a workgroup's
// columns is one (zfwd) or two (zinv) 128-byte lines of the sticks buffer
// (measured best on the 72^3 grid in both precisions)
template <typename T, bool INV>
constexpr int cpw() {
  return (INV ? 256 : 128) / static_cast<int>(sizeof(cplx<T>));
}

// ---------------------------------------------------------------------------
// one Stockham pass over nlines lines of length N: loads through ld(line, idx),
// stores through st(line, idx, value); a barrier separates the two phases and
// follows the stores.  MAXL bounds nlines (register staging).
// ---------------------------------------------------------------------------
template <typename T, int N, int S, int MAXL, int BT, int DIR, bool REV = false, class Ld, class St>
__device__ __forceinline__ void pass(int nlines, const cplx<T>* tw, Ld ld, St st) {
  using PL = Plan<N, REV>;
  constexpr int R = PL::radix(S);
  constexpr int P = PL::stride(S);
  constexpr int NBL = N / R;
  constexpr int NB = (MAXL * NBL + BT - 1) / BT;
  const int nbf = nlines * NBL;
  cplx<T> v[NB][R];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int bf = threadIdx.x + t * BT;
    if (bf < nbf) {
      const int line = bf / NBL;
      const int i = bf - line * NBL;
#pragma unroll
      for (int j = 0; j < R; ++j) v[t][j] = ld(line, i + j * NBL);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int bf = threadIdx.x + t * BT;
    if (bf < nbf) {
      const int line = bf / NBL;
      const int i = bf - line * NBL;
      const int k = (P == 1) ? 0 : i % P;
      if (P > 1 && k != 0) {
        const cplx<T>* wk = tw + PL::twoff(S) + k * PL::tws(S);
#pragma unroll
        for (int j = 1; j < R; ++j) {
          cplx<T> w = wk[j - 1];
          if (DIR > 0) w.y = -w.y;
          v[t][j] = cmul(v[t][j], w);
        }
      }
      dft<T, R, DIR>(v[t]);
      const int base = (i - k) * R + k;
#pragma unroll
      for (int j = 0; j < R; ++j) st(line, base + j * P, v[t][j]);
    }
  }
  __syncthreads();
}

// full transform: pass 0 loads through first(), the last pass stores through
// last(), the passes in between work in place on buf (line stride N)
template <typename T, int N, int S, int MAXL, int BT, int DIR, class First, class Last, int LS = N>
__device__ __forceinline__ void passes(int nlines, cplx<T>* buf, const cplx<T>* tw, First first, Last last) {
  constexpr int NS = Plan<N>::nst;
  auto in_buf = [buf](int line, int idx) { return buf[line * LS + idx]; };
  auto to_buf = [buf](int line, int idx, cplx<T> v) { buf[line * LS + idx] = v; };
  if constexpr (NS == 1) {
    pass<T, N, 0, MAXL, BT, DIR>(nlines, tw, first, last);
  } else if constexpr (S == 0) {
    pass<T, N, 0, MAXL, BT, DIR>(nlines, tw, first, to_buf);
    passes<T, N, 1, MAXL, BT, DIR, First, Last, LS>(nlines, buf, tw, first, last);
  } else if constexpr (S == NS - 1) {
    pass<T, N, S, MAXL, BT, DIR>(nlines, tw, in_buf, last);
  } else {
    pass<T, N, S, MAXL, BT, DIR>(nlines, tw, in_buf, to_buf);
    passes<T, N, S + 1, MAXL, BT, DIR, First, Last, LS>(nlines, buf, tw, first, last);
  }
}

// LS: distance between lines in buf (N, or N + 1 for the plane rows, whose
// column reads would otherwise fall on one or two LDS bank groups)
template <typename T, int N, int MAXL, int BT, int DIR, int LS = N>
__device__ __forceinline__ void fft_inplace(int nlines, cplx<T>* buf, const cplx<T>* tw) {
  auto in_buf = [buf](int line, int idx) { return buf[line * LS + idx]; };
  auto to_buf = [buf](int line, int idx, cplx<T> v) { buf[line * LS + idx] = v; };
  passes<T, N, 0, MAXL, BT, DIR, decltype(in_buf), decltype(to_buf), LS>(nlines, buf, tw, in_buf, to_buf);
}

// pass twiddle tables of both plans (Plan<N, REV>::twoff) from the forward
// roots g[m] = exp(-2 pi i m / N)
template <typename T, int N, bool REV>
__device__ __forceinline__ void load_tw_plan(cplx<T>* tw, const double* g) {
  using PL = Plan<N, REV>;
#pragma unroll
  for (int s = 1; s < PL::nst; ++s) {
    const int R = PL::radix(s);
    const int step = N / (PL::stride(s) * R);
    const int n = PL::stride(s) * (R - 1);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      const int k = t / (R - 1);
      const int j = t - k * (R - 1) + 1;
      const int m = (j * k * step) % N;
      tw[PL::twoff(s) + k * PL::tws(s) + j - 1] = mk<T>(T(g[2 * m]), T(g[2 * m + 1]));
    }
  }
}

template <typename T, int N>
__device__ __forceinline__ void load_tw(cplx<T>* tw, const double* g) {
  load_tw_plan<T, N, false>(tw, g);
  load_tw_plan<T, N, true>(tw, g);
}

// The turn between an inverse transform (forward-order plan) and a forward
// transform (reversed plan) along the same lines: the inverse's last pass,
// the pointwise op mid(line, position, value), and the reversed plan's first
// pass, in registers.  Loads through ld like a pass; stores pass-0 outputs of
// the reversed plan through st.
template <typename T, int N, int MAXL, int BT, class Ld, class Mid, class St>
__device__ __forceinline__ void pass_turn(int nlines, const cplx<T>* tw, Ld ld, Mid mid, St st) {
  using PF = Plan<N, false>;
  constexpr int S = PF::nst - 1;
  constexpr int R = PF::radix(S);
  constexpr int P = PF::stride(S);
  constexpr int NBL = N / R;
  static_assert(NBL == P, "last pass: one butterfly per residue");
  static_assert(Plan<N, true>::radix(0) == R, "reversed plan starts with the last radix");
  constexpr int NB = (MAXL * NBL + BT - 1) / BT;
  const int nbf = nlines * NBL;
  cplx<T> v[NB][R];
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int bf = threadIdx.x + t * BT;
    if (bf < nbf) {
      const int line = bf / NBL;
      const int i = bf - line * NBL;
#pragma unroll
      for (int j = 0; j < R; ++j) v[t][j] = ld(line, i + j * NBL);
    }
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NB; ++t) {
    const int bf = threadIdx.x + t * BT;
    if (bf < nbf) {
      const int line = bf / NBL;
      const int i = bf - line * NBL;   // = k of the last pass (i < P)
      if (P > 1 && i != 0) {
        const cplx<T>* wk = tw + PF::twoff(S) + i * PF::tws(S);
#pragma unroll
        for (int j = 1; j < R; ++j) {
          cplx<T> w = wk[j - 1];
          w.y = -w.y;   // inverse: exp(+i ...)
          v[t][j] = cmul(v[t][j], w);
        }
      }
      dft<T, R, +1>(v[t]);
#pragma unroll
      for (int j = 0; j < R; ++j) v[t][j] = mid(line, i + j * P, v[t][j]);
      dft<T, R, -1>(v[t]);
#pragma unroll
      for (int j = 0; j < R; ++j) st(line, i * R + j, v[t][j]);
    }
  }
  __syncthreads();
}

// threads of the per-plane kernels (two waves per SIMD)
constexpr int NTP = 512;

// maximum number of |i| <= imax rows: the wave-function sphere spans at most
// about half of the grid along a1
template <int N>
constexpr int max_rows() {
  return N / 2 + 2;
}
// row stride of the plane buffer (one element of padding: column accesses
// stride over all LDS banks)
template <int N>
constexpr int prow() {
  return N + 1;
}
// x-line chunk of the plane kernel when `occ` workgroups share the 160 KB of
// LDS of a CU (more resident workgroups hide the barriers of the LDS passes)
template <typename T, int N>
constexpr int x_budget(bool density, int occ) {
  const int budget = 160 * 1024 / occ - 2048;
  const int fixed = static_cast<int>(sizeof(cplx<T>)) * (Plan<N>::twsize() + max_rows<N>() * prow<N>()) +
                    (density ? static_cast<int>(sizeof(T)) * N * N : 0);
  return (budget - fixed) / (static_cast<int>(sizeof(cplx<T>)) * N);
}
// x lines of the potential kernel: one element of padding after every 8, so
// the stride-8 stores of a radix-8 pass spread over the LDS banks
template <int N>
constexpr int xline() {
  return N + N / 8 + 1;
}
__device__ __forceinline__ int xpad(int x) { return x + (x >> 3); }

template <typename T, int N>
constexpr int vloc_budget(int occ) {
  const int budget = 160 * 1024 / occ - 2048;
  const int fixed = static_cast<int>(sizeof(cplx<T>)) * (Plan<N>::twsize() + max_rows<N>() * prow<N>());
  return (budget - fixed) / (static_cast<int>(sizeof(cplx<T>)) * xline<N>());
}
template <typename T, int N>
constexpr int vloc_chunk(int occ) {
  int jc = vloc_budget<T, N>(occ);
  if (jc > N) jc = N;
  if (jc < 1) return 1;
  const int nchunk = (N + jc - 1) / jc;
  return (N + nchunk - 1) / nchunk;
}
// occ workgroups per CU leave room for at least 4 x-lines per chunk
template <typename T, int N>
constexpr bool occ_fits(int occ) {
  return occ == 1 || vloc_budget<T, N>(occ) >= 4;
}
template <typename T, int N>
constexpr int x_chunk(bool density, int occ = 1) {
  int jc = x_budget<T, N>(density, occ);
  if (jc > N) jc = N;
  if (jc < 1) return 1;
  const int nchunk = (N + jc - 1) / jc;   // equal chunks
  return (N + nchunk - 1) / nchunk;
}

// ---------------------------------------------------------------------------
// sphere -> sticks
// ---------------------------------------------------------------------------
// V: precision of the sphere vectors (double, or float for the FP32 operator)
template <typename T, int N, typename V>
__global__ __launch_bounds__(NT) void k_zinv_t(StickPlanDev P, const FftItem* items, const cplx<V>* psiA, long ldA,
                                               const cplx<V>* psiB, long ldB, cplx<T>* sticks, int mode) {
  constexpr int CPW = cpw<T, true>();
  __shared__ cplx<T> tw[Plan<N>::twsize()];
  constexpr int LS = N + 1;   // line stride: column-order accesses spread over the LDS banks
  __shared__ cplx<T> lines[2 * CPW * LS];
  const int it = blockIdx.y;
  const int c0 = blockIdx.x * CPW;
  const int c1 = min(c0 + CPW, P.ncs);
  const int ncl = c1 - c0;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const cplx<V>* A = psiA + static_cast<long>(item.a) * ldA;
  const cplx<V>* B = item.b >= 0 ? psiB + static_cast<long>(item.b) * ldB : nullptr;
  load_tw<T, N>(tw, P.tw3);
  for (int t = threadIdx.x; t < 2 * ncl * LS; t += NT) lines[t] = mk<T>(T(0), T(0));
  __syncthreads();
  const int s0 = P.col_off[c0];
  const int s1 = P.col_off[c1];
  for (int s = s0 + threadIdx.x; s < s1; s += NT) {
    const int lc = P.pt_col[s] - c0;
    int z = P.pt_k[s] % N;
    if (z < 0) z += N;
    const cplx<V> a1 = A[s];
    const cplx<V> a2 = B ? B[s] : mk<V>(V(0), V(0));
    lines[(2 * lc) * LS + z] = mk<T>(T(a1.x - a2.y), T(a1.y + a2.x));
  }
  __syncthreads();
  for (int s = s0 + threadIdx.x; s < s1; s += NT) {
    const int lc = P.pt_col[s] - c0;
    int zm = (-P.pt_k[s]) % N;
    if (zm < 0) zm += N;
    const cplx<V> a1 = A[s];
    const cplx<V> a2 = B ? B[s] : mk<V>(V(0), V(0));
    cplx<V> v = mk<V>(a1.x + a2.y, a2.x - a1.y);
    if (mode == ZINV_PAIR_REALG0 && s == 0) v = mk<V>(a1.x, a2.x);
    const int line = (c0 + lc == 0) ? 0 : 2 * lc + 1;
    lines[line * LS + zm] = mk<T>(T(v.x), T(v.y));
  }
  __syncthreads();
  fft_inplace<T, N, 2 * CPW, NT, +1, LS>(2 * ncl, lines, tw);
  cplx<T>* out = sticks + static_cast<long>(it) * N * P.nct;
  for (int t = threadIdx.x; t < N * 2 * ncl; t += NT) {
    const int z = t / (2 * ncl);
    const int r = t - z * 2 * ncl;
    const bool own = r < ncl;
    const int lc = own ? r : r - ncl;
    const int c = c0 + lc;
    if (!own && c == 0) continue;
    const int col = own ? c : c + P.ncs - 1;
    out[static_cast<long>(z) * P.nct + col] = lines[(2 * lc + (own ? 0 : 1)) * LS + z];
  }
}

// ---------------------------------------------------------------------------
// sticks -> sphere
// ---------------------------------------------------------------------------
template <typename T, int N, typename V>
__global__ __launch_bounds__(NT) void k_zfwd_t(StickPlanDev P, const FftItem* items, const cplx<T>* sticks,
                                               const cplx<V>* in, long ldin, cplx<V>* out, long ldout, ZfwdArgs A) {
  constexpr int CPW = cpw<T, false>();
  __shared__ cplx<T> tw[Plan<N>::twsize()];
  constexpr int LS = N + 1;   // line stride: column-order accesses spread over the LDS banks
  __shared__ cplx<T> lines[2 * CPW * LS];
  const int it = blockIdx.y;
  const int c0 = blockIdx.x * CPW;
  const int c1 = min(c0 + CPW, P.ncs);
  const int ncl = c1 - c0;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  load_tw<T, N>(tw, P.tw3);
  const cplx<T>* src = sticks + static_cast<long>(it) * N * P.nct;
  const bool single = A.mode == ZFWD_SINGLE;
  const int nl = single ? ncl : 2 * ncl;   // SINGLE needs no mirror lines
  for (int t = threadIdx.x; t < N * nl; t += NT) {
    const int z = t / nl;
    const int r = t - z * nl;
    const bool own = r < ncl;
    const int lc = own ? r : r - ncl;
    const int c = c0 + lc;
    cplx<T> v = mk<T>(T(0), T(0));
    if (own || c != 0) {
      const int col = own ? c : c + P.ncs - 1;
      v = src[static_cast<long>(z) * P.nct + col];
    }
    lines[(single ? lc : 2 * lc + (own ? 0 : 1)) * LS + z] = v;
  }
  __syncthreads();
  fft_inplace<T, N, 2 * CPW, NT, -1, LS>(nl, lines, tw);
  const int s0 = P.col_off[c0];
  const int s1 = P.col_off[c1];
  cplx<V>* oa = out + static_cast<long>(item.a) * ldout;
  cplx<V>* ob = item.b >= 0 ? out + static_cast<long>(item.b) * ldout : nullptr;
  for (int s = s0 + threadIdx.x; s < s1; s += NT) {
    const int lc = P.pt_col[s] - c0;
    const int k = P.pt_k[s];
    int z = k % N;
    if (z < 0) z += N;
    if (single) {
      const cplx<T> F = lines[lc * LS + z];
      const double w = A.scale * (A.weight ? A.weight[s] : 1.0);
      cplx<V> v = mk<V>(V(static_cast<double>(F.x) * w), V(static_cast<double>(F.y) * w));
      if (A.zero_g0 && s == 0) v = mk<V>(V(0), V(0));
      oa[s] = v;
      continue;
    }
    int zm = (-k) % N;
    if (zm < 0) zm += N;
    const cplx<T> Ft = lines[(2 * lc) * LS + z];
    const int ml = (c0 + lc == 0) ? 0 : 2 * lc + 1;
    const cplx<T> Fmt = lines[ml * LS + zm];
    const cplx<V> F = mk<V>(V(Ft.x), V(Ft.y));
    const cplx<V> Fm = mk<V>(V(Fmt.x), V(Fmt.y));
    const cplx<V> fp = cscale(F + Fm, V(0.5));
    const cplx<V> fm = cscale(F - Fm, V(0.5));
    const cplx<V> b1 = mk<V>(fp.x, fm.y);
    const cplx<V> b2 = mk<V>(fp.y, -fm.x);
    if (A.mode == ZFWD_PAIR_STORE) {
      oa[s] = cscale(b1, V(A.scale));
      if (ob) ob[s] = cscale(b2, V(A.scale));
    } else {
      const double g2 = A.g2kin[s];
      const cplx<V> ia = in[static_cast<long>(item.a) * ldin + s];
      oa[s] = b1 + cscale(ia, V(g2 - A.e[item.a]));
      if (ob) {
        const cplx<V> ib = in[static_cast<long>(item.b) * ldin + s];
        ob[s] = b2 + cscale(ib, V(g2 - A.e[item.b]));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// per-plane 2D part
// ---------------------------------------------------------------------------
template <typename T, int N>
struct PlaneLds {
  cplx<T>* tw;
  cplx<T>* pl;   // max_rows x N
  cplx<T>* x;    // JC x N
};

template <typename T, int N>
__device__ __forceinline__ PlaneLds<T, N> plane_lds() {
  extern __shared__ char smem_raw[];
  PlaneLds<T, N> s;
  s.tw = reinterpret_cast<cplx<T>*>(smem_raw);
  s.pl = s.tw + Plan<N>::twsize();
  s.x = s.pl + max_rows<N>() * prow<N>();
  return s;
}

// plane-buffer index of a column position r * N + j (col_pos)
template <int N>
__device__ __forceinline__ int pl_pos(int pos) {
  const int r = pos / N;
  return pos + r * (prow<N>() - N);
}

template <typename T, int N>
__device__ __forceinline__ void plane_fill(const StickPlanDev& P, PlaneLds<T, N>& S, const cplx<T>* src) {
  for (int t = threadIdx.x; t < P.nrows * prow<N>(); t += NTP) S.pl[t] = mk<T>(T(0), T(0));
  __syncthreads();
  for (int c = threadIdx.x; c < P.nct; c += NTP) S.pl[pl_pos<N>(P.col_pos[c])] = src[c];
  __syncthreads();
}

// row of the sphere holding a1-frequency x, or -1
__device__ __forceinline__ int row_of(int x, int N, int imax) {
  if (x <= imax) return x + imax;
  if (x >= N - imax) return x - N + imax;
  return -1;
}

// PT: storage type of the potential (double, or float for the FP32 operator);
// OCC: workgroups per CU the LDS footprint is sized for
template <typename T, int N, typename PT, int OCC>
__global__ __launch_bounds__(NTP) __attribute__((amdgpu_waves_per_eu(2 * OCC))) void k_plane_vloc_t(StickPlanDev P, const FftItem* items, const cplx<T>* sin,
                                                     cplx<T>* sout, const PT* vpot, long ldpot, double scale) {
  constexpr int JC = vloc_chunk<T, N>(OCC);
  constexpr int XL = xline<N>();
  constexpr int MR = max_rows<N>();
  PlaneLds<T, N> S = plane_lds<T, N>();
  const int z = blockIdx.x;
  const int it = blockIdx.y;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const long sl = static_cast<long>(N) * P.nct;
  const int imax = P.imax;
  load_tw<T, N>(S.tw, P.tw1);
  plane_fill<T, N>(P, S, sin + static_cast<long>(item.a) * sl + static_cast<long>(z) * P.nct);
  fft_inplace<T, N, MR, NTP, +1, prow<N>()>(P.nrows, S.pl, S.tw);
  const PT* V = vpot + (item.pot >= 0 ? static_cast<long>(item.pot) * ldpot : 0L) + static_cast<long>(z) * N * N;
  const T sc = T(scale);
  cplx<T>* pl = S.pl;
  cplx<T>* xb = S.x;
  for (int j0 = 0; j0 < N; j0 += JC) {
    const int nj = min(JC, N - j0);
    // inverse along a1: zero-padded rows in, V(r) applied on the way out
    auto from_rows = [=](int jj, int x) {
      const int r = row_of(x, N, imax);
      return r >= 0 ? pl[r * prow<N>() + j0 + jj] : mk<T>(T(0), T(0));
    };
    auto in_x = [=](int jj, int x) { return xb[jj * XL + xpad(x)]; };
    auto to_x = [=](int jj, int x, cplx<T> v) { xb[jj * XL + xpad(x)] = v; };
    auto times_v = [=](int jj, int x, cplx<T> v) { return cscale(v, T(V[(j0 + jj) * N + x]) * sc); };
    // forward along a1 (reversed radix order): only |i| <= imax frequencies go back to the rows
    auto to_rows = [=](int jj, int x, cplx<T> v) {
      const int r = row_of(x, N, imax);
      if (r >= 0) pl[r * prow<N>() + j0 + jj] = v;
    };
    if constexpr (Plan<N>::nst == 2) {
      pass<T, N, 0, JC, NTP, +1>(nj, S.tw, from_rows, to_x);
      // inverse's last pass, V(r), forward's first pass in registers
      pass_turn<T, N, JC, NTP>(nj, S.tw, in_x, times_v, to_x);
      pass<T, N, 1, JC, NTP, -1, true>(nj, S.tw, in_x, to_rows);
    } else {
      // other plans: inverse with V(r) applied on the way out, then the forward
      auto store_v = [=](int jj, int x, cplx<T> v) { xb[jj * XL + x] = times_v(jj, x, v); };
      auto in_xl = [=](int jj, int x) { return xb[jj * XL + x]; };
      passes<T, N, 0, JC, NTP, +1, decltype(from_rows), decltype(store_v), XL>(nj, xb, S.tw, from_rows, store_v);
      passes<T, N, 0, JC, NTP, -1, decltype(in_xl), decltype(to_rows), XL>(nj, xb, S.tw, in_xl, to_rows);
    }
  }
  fft_inplace<T, N, MR, NTP, -1, prow<N>()>(P.nrows, S.pl, S.tw);
  cplx<T>* dst = sout + static_cast<long>(it) * sl + static_cast<long>(z) * P.nct;
  for (int c = threadIdx.x; c < P.nct; c += NTP) dst[c] = S.pl[pl_pos<N>(P.col_pos[c])];
}

template <typename T, int N>
__global__ __launch_bounds__(NTP) void k_plane_to_real_t(StickPlanDev P, const FftItem* items, const cplx<T>* sin,
                                                        double* out, long ldout) {
  constexpr int JC = x_chunk<T, N>(false);
  constexpr int MR = max_rows<N>();
  PlaneLds<T, N> S = plane_lds<T, N>();
  const int z = blockIdx.x;
  const int it = blockIdx.y;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const long sl = static_cast<long>(N) * P.nct;
  const int imax = P.imax;
  load_tw<T, N>(S.tw, P.tw1);
  plane_fill<T, N>(P, S, sin + static_cast<long>(it) * sl + static_cast<long>(z) * P.nct);
  fft_inplace<T, N, MR, NTP, +1, prow<N>()>(P.nrows, S.pl, S.tw);
  double* oa = out + static_cast<long>(item.a) * ldout + static_cast<long>(z) * N * N;
  double* ob = item.b >= 0 ? out + static_cast<long>(item.b) * ldout + static_cast<long>(z) * N * N : nullptr;
  cplx<T>* pl = S.pl;
  for (int j0 = 0; j0 < N; j0 += JC) {
    const int nj = min(JC, N - j0);
    auto from_rows = [=](int jj, int x) {
      const int r = row_of(x, N, imax);
      return r >= 0 ? pl[r * prow<N>() + j0 + jj] : mk<T>(T(0), T(0));
    };
    auto to_out = [=](int jj, int x, cplx<T> v) {
      oa[(j0 + jj) * N + x] = static_cast<double>(v.x);
      if (ob) ob[(j0 + jj) * N + x] = static_cast<double>(v.y);
    };
    passes<T, N, 0, JC, NTP, +1>(nj, S.x, S.tw, from_rows, to_out);
  }
}

template <typename T, int N>
__global__ __launch_bounds__(NTP) void k_plane_density_t(StickPlanDev P, int nband, const cplx<T>* sin,
                                                        const double* occ, cplx<T>* sout) {
  constexpr int JC = x_chunk<T, N>(true);
  constexpr int MR = max_rows<N>();
  PlaneLds<T, N> S = plane_lds<T, N>();
  T* acc = reinterpret_cast<T*>(S.x + JC * N);
  const int z = blockIdx.x;
  const int p = blockIdx.y;
  const long sl = static_cast<long>(N) * P.nct;
  const int imax = P.imax;
  load_tw<T, N>(S.tw, P.tw1);
  for (int t = threadIdx.x; t < N * N; t += NTP) acc[t] = T(0);
  cplx<T>* pl = S.pl;
  for (int v = 0; v < nband; ++v) {
    plane_fill<T, N>(P, S, sin + (static_cast<long>(p) * nband + v) * sl + static_cast<long>(z) * P.nct);
    fft_inplace<T, N, MR, NTP, +1, prow<N>()>(P.nrows, S.pl, S.tw);
    const T o = T(occ[v]);
    for (int j0 = 0; j0 < N; j0 += JC) {
      const int nj = min(JC, N - j0);
      auto from_rows = [=](int jj, int x) {
        const int r = row_of(x, N, imax);
        return r >= 0 ? pl[r * prow<N>() + j0 + jj] : mk<T>(T(0), T(0));
      };
      // each (jj, x) is produced by exactly one thread: no race on acc
      auto accumulate = [=](int jj, int x, cplx<T> w) { acc[(j0 + jj) * N + x] += o * w.x * w.y; };
      passes<T, N, 0, JC, NTP, +1>(nj, S.x, S.tw, from_rows, accumulate);
    }
  }
  // forward of the real density: a1 then a2, keeping |i| <= imax and the columns
  for (int j0 = 0; j0 < N; j0 += JC) {
    const int nj = min(JC, N - j0);
    auto from_acc = [=](int jj, int x) { return mk<T>(acc[(j0 + jj) * N + x], T(0)); };
    auto to_rows = [=](int jj, int x, cplx<T> w) {
      const int r = row_of(x, N, imax);
      if (r >= 0) pl[r * prow<N>() + j0 + jj] = w;
    };
    passes<T, N, 0, JC, NTP, -1>(nj, S.x, S.tw, from_acc, to_rows);
  }
  fft_inplace<T, N, MR, NTP, -1, prow<N>()>(P.nrows, S.pl, S.tw);
  cplx<T>* dst = sout + static_cast<long>(p) * sl + static_cast<long>(z) * P.nct;
  for (int c = threadIdx.x; c < P.nct; c += NTP) dst[c] = S.pl[pl_pos<N>(P.col_pos[c])];
}

// Response density for grids whose N x N FP64 accumulator does not fit in LDS
// beside the plane (N = 100, 108 in FP64): the x range is covered in two
// halves, each a pass over all bands with an N x N/2 accumulator; the first
// half is parked in global scratch (N x N/2 reals per plane, written and read
// back by the same workgroup) until the forward transforms.
template <typename T, int N>
constexpr int xh_chunk() {
  const int budget = 160 * 1024 - 2048;
  const int fixed = static_cast<int>(sizeof(cplx<T>)) * (Plan<N>::twsize() + max_rows<N>() * prow<N>()) +
                    static_cast<int>(sizeof(T)) * N * (N / 2);
  int jc = (budget - fixed) / (static_cast<int>(sizeof(cplx<T>)) * N);
  if (jc > N) jc = N;
  if (jc < 1) return 0;
  const int nchunk = (N + jc - 1) / jc;
  return (N + nchunk - 1) / nchunk;
}

template <typename T, int N>
constexpr size_t big_density_smem() {
  return sizeof(cplx<T>) * (static_cast<size_t>(Plan<N>::twsize()) + static_cast<size_t>(max_rows<N>()) * prow<N>() +
                            static_cast<size_t>(xh_chunk<T, N>()) * N) +
         sizeof(T) * static_cast<size_t>(N) * (N / 2);
}

// LDS of k_plane_density_t (whole N x N accumulator)
template <typename T, int N>
constexpr bool density_fits_lds() {
  return sizeof(cplx<T>) * (static_cast<size_t>(Plan<N>::twsize()) + static_cast<size_t>(max_rows<N>()) * prow<N>() +
                            static_cast<size_t>(x_chunk<T, N>(true)) * N) +
             sizeof(T) * static_cast<size_t>(N) * N <=
         160u * 1024u;
}

template <typename T, int N>
__global__ __launch_bounds__(NTP) void k_plane_density_big_t(StickPlanDev P, int nband, const cplx<T>* sin,
                                                            const double* occ, cplx<T>* sout, T* scratch) {
  constexpr int JC = xh_chunk<T, N>();
  constexpr int MR = max_rows<N>();
  constexpr int H = N / 2;
  static_assert(JC >= 1 && N % 2 == 0, "big density kernel: even N with room for one x line");
  PlaneLds<T, N> S = plane_lds<T, N>();
  T* acc = reinterpret_cast<T*>(S.x + JC * N);
  const int z = blockIdx.x;
  const int p = blockIdx.y;
  const long sl = static_cast<long>(N) * P.nct;
  const int imax = P.imax;
  load_tw<T, N>(S.tw, P.tw1);
  T* park = scratch + (static_cast<long>(p) * N + z) * N * H;   // [j][x < H]
  cplx<T>* pl = S.pl;
  for (int h = 0; h < 2; ++h) {
    for (int t = threadIdx.x; t < N * H; t += NTP) acc[t] = T(0);
    for (int v = 0; v < nband; ++v) {
      plane_fill<T, N>(P, S, sin + (static_cast<long>(p) * nband + v) * sl + static_cast<long>(z) * P.nct);
      fft_inplace<T, N, MR, NTP, +1, prow<N>()>(P.nrows, S.pl, S.tw);
      const T o = T(occ[v]);
      for (int j0 = 0; j0 < N; j0 += JC) {
        const int nj = min(JC, N - j0);
        auto from_rows = [=](int jj, int x) {
          const int r = row_of(x, N, imax);
          return r >= 0 ? pl[r * prow<N>() + j0 + jj] : mk<T>(T(0), T(0));
        };
        // each (jj, x) is produced by exactly one thread: no race on acc
        auto accumulate = [=](int jj, int x, cplx<T> w) {
          const int xl = x - h * H;
          if (xl >= 0 && xl < H) acc[(j0 + jj) * H + xl] += o * w.x * w.y;
        };
        passes<T, N, 0, JC, NTP, +1>(nj, S.x, S.tw, from_rows, accumulate);
      }
    }
    if (h == 0) {
      for (int t = threadIdx.x; t < N * H; t += NTP) park[t] = acc[t];
      __threadfence_block();
      __syncthreads();
    }
  }
  // forward of the real density: a1 then a2, keeping |i| <= imax and the columns
  for (int j0 = 0; j0 < N; j0 += JC) {
    const int nj = min(JC, N - j0);
    auto from_acc = [=](int jj, int x) {
      const T v = x < H ? park[(j0 + jj) * H + x] : acc[(j0 + jj) * H + x - H];
      return mk<T>(v, T(0));
    };
    auto to_rows = [=](int jj, int x, cplx<T> w) {
      const int r = row_of(x, N, imax);
      if (r >= 0) pl[r * prow<N>() + j0 + jj] = w;
    };
    passes<T, N, 0, JC, NTP, -1>(nj, S.x, S.tw, from_acc, to_rows);
  }
  fft_inplace<T, N, MR, NTP, -1, prow<N>()>(P.nrows, S.pl, S.tw);
  cplx<T>* dst = sout + static_cast<long>(p) * sl + static_cast<long>(z) * P.nct;
  for (int c = threadIdx.x; c < P.nct; c += NTP) dst[c] = S.pl[pl_pos<N>(P.col_pos[c])];
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <typename T, int N>
size_t plane_smem(bool density, int occ = 1) {
  const int jc = occ == 1 ? x_chunk<T, N>(density, 1) : occ == 2 ? x_chunk<T, N>(density, 2) : x_chunk<T, N>(density, 3);
  size_t b = sizeof(cplx<T>) * (static_cast<size_t>(Plan<N>::twsize()) + static_cast<size_t>(max_rows<N>()) * prow<N>() +
                                static_cast<size_t>(jc) * N);
  if (density) b += sizeof(T) * static_cast<size_t>(N) * N;
  return b;
}

template <class K>
void allow_smem(K kernel, size_t bytes) {
  static std::map<const void*, size_t> done;
  static std::mutex done_mutex;
  std::lock_guard<std::mutex> lk(done_mutex);
  auto key = reinterpret_cast<const void*>(kernel);
  auto itr = done.find(key);
  if (itr != done.end() && itr->second >= bytes) return;
  WH_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
  done[key] = bytes;
}

template <typename T, int N, int OCC>
constexpr size_t vloc_smem() {
  return sizeof(cplx<T>) * (static_cast<size_t>(Plan<N>::twsize()) + static_cast<size_t>(max_rows<N>()) * prow<N>() +
                            static_cast<size_t>(vloc_chunk<T, N>(OCC)) * xline<N>());
}

template <typename T, int N, typename PT, int OCC>
void vloc_occ(const StickPlanDev& P, const FftItem* it, int n, const void* si, void* so, const PT* v, long ldv,
              double scale, hipStream_t s) {
  constexpr size_t sm = vloc_smem<T, N, OCC>();
  static_assert(sm <= 160u * 1024u, "plane kernel exceeds LDS");
  allow_smem(k_plane_vloc_t<T, N, PT, OCC>, sm);
  k_plane_vloc_t<T, N, PT, OCC><<<dim3(N, n), NTP, sm, s>>>(P, it, static_cast<const cplx<T>*>(si),
                                                            static_cast<cplx<T>*>(so), v, ldv, scale);
}

// occ: requested workgroups per CU, lowered to the largest that fits (<= 3)
template <typename T, int N, typename PT>
void vloc_any(const StickPlanDev& P, const FftItem* it, int n, const void* si, void* so, const PT* v, long ldv,
              double scale, int occ, hipStream_t s) {
  if constexpr (occ_fits<T, N>(3)) {
    if (occ >= 3) return vloc_occ<T, N, PT, 3>(P, it, n, si, so, v, ldv, scale, s);
  }
  if constexpr (occ_fits<T, N>(2)) {
    if (occ >= 2) return vloc_occ<T, N, PT, 2>(P, it, n, si, so, v, ldv, scale, s);
  }
  vloc_occ<T, N, PT, 1>(P, it, n, si, so, v, ldv, scale, s);
}

template <typename T, int N>
constexpr int max_occ() {
  return occ_fits<T, N>(3) ? 3 : occ_fits<T, N>(2) ? 2 : 1;
}

template <typename T, int N>
struct Launch {
  static void zinv(const StickPlanDev& P, const FftItem* it, int n, const z64* a, long lda, const z64* b, long ldb,
                   void* st, int mode, hipStream_t s) {
    k_zinv_t<T, N, double><<<dim3(ceil_div(P.ncs, cpw<T, true>()), n), NT, 0, s>>>(P, it, a, lda, b, ldb, static_cast<cplx<T>*>(st), mode);
  }
  static void zfwd(const StickPlanDev& P, const FftItem* it, int n, const void* st, const z64* in, long ldin, z64* out,
                   long ldout, const ZfwdArgs& A, hipStream_t s) {
    k_zfwd_t<T, N, double><<<dim3(ceil_div(P.ncs, cpw<T, false>()), n), NT, 0, s>>>(P, it, static_cast<const cplx<T>*>(st), in, ldin, out,
                                                                 ldout, A);
  }
  static void vloc(const StickPlanDev& P, const FftItem* it, int n, const void* si, void* so, const double* v, long ldv,
                   double scale, int occ, hipStream_t s) {
    vloc_any<T, N, double>(P, it, n, si, so, v, ldv, scale, occ, s);
  }
  static void to_real(const StickPlanDev& P, const FftItem* it, int n, const void* si, double* out, long ldout,
                      hipStream_t s) {
    const size_t sm = plane_smem<T, N>(false);
    allow_smem(k_plane_to_real_t<T, N>, sm);
    k_plane_to_real_t<T, N><<<dim3(N, n), NTP, sm, s>>>(P, it, static_cast<const cplx<T>*>(si), out, ldout);
  }
  static void density(const StickPlanDev& P, int np, int nb, const void* si, const double* occ, void* so,
                      void* scratch, hipStream_t s) {
    if constexpr (density_fits_lds<T, N>()) {
      const size_t sm = plane_smem<T, N>(true);
      allow_smem(k_plane_density_t<T, N>, sm);
      k_plane_density_t<T, N><<<dim3(N, np), NTP, sm, s>>>(P, nb, static_cast<const cplx<T>*>(si), occ,
                                                          static_cast<cplx<T>*>(so));
      return;
    }
    if constexpr (!density_fits_lds<T, N>() && N % 2 == 0 && xh_chunk<T, N>() >= 1) {
      WH_REQUIRE(scratch != nullptr, "response density: scratch buffer required for this grid");
      constexpr size_t smb = big_density_smem<T, N>();
      static_assert(smb <= 160u * 1024u, "big density kernel exceeds LDS");
      allow_smem(k_plane_density_big_t<T, N>, smb);
      k_plane_density_big_t<T, N><<<dim3(N, np), NTP, smb, s>>>(P, nb, static_cast<const cplx<T>*>(si), occ,
                                                               static_cast<cplx<T>*>(so), static_cast<T*>(scratch));
    } else {
      WH_REQUIRE(false, "response-density plane does not fit in LDS for this grid");
    }
  }
};

// FP32 operator: FP32 sticks, vectors and potential
template <int N>
struct Launch32 {
  static void zinv(const StickPlanDev& P, const FftItem* it, int n, const z32* a, long lda, const z32* b, long ldb,
                   void* st, int mode, hipStream_t s) {
    k_zinv_t<float, N, float><<<dim3(ceil_div(P.ncs, cpw<float, true>()), n), NT, 0, s>>>(P, it, a, lda, b, ldb,
                                                                           static_cast<z32*>(st), mode);
  }
  static void zfwd(const StickPlanDev& P, const FftItem* it, int n, const void* st, const z32* in, long ldin, z32* out,
                   long ldout, const ZfwdArgs& A, hipStream_t s) {
    k_zfwd_t<float, N, float><<<dim3(ceil_div(P.ncs, cpw<float, false>()), n), NT, 0, s>>>(P, it, static_cast<const z32*>(st), in,
                                                                           ldin, out, ldout, A);
  }
  static void vloc(const StickPlanDev& P, const FftItem* it, int n, const void* si, void* so, const float* v, long ldv,
                   double scale, int occ, hipStream_t s) {
    vloc_any<float, N, float>(P, it, n, si, so, v, ldv, scale, occ, s);
  }
};

// scratch reals per perturbation the density launcher needs (0: the plane fits in LDS;
// -1: no kernel for this size)
template <typename T, int N>
long density_scratch() {
  if constexpr (density_fits_lds<T, N>()) {
    return 0;
  } else if constexpr (N % 2 == 0 && xh_chunk<T, N>() >= 1) {
    return static_cast<long>(N) * N * (N / 2);
  } else {
    return -1;
  }
}

template <typename T>
const TplKernels* table_for(int n) {
#define WH_TPL_CASE(NN)                                                                                       \
  case NN: {                                                                                                  \
    static const TplKernels k{&Launch<T, NN>::zinv, &Launch<T, NN>::zfwd, &Launch<T, NN>::vloc,             \
                              &Launch<T, NN>::to_real, &Launch<T, NN>::density, density_scratch<T, NN>(), \
                              max_occ<T, NN>()};                                                              \
    return &k;                                                                                                \
  }
  switch (n) {
    WH_TPL_CASE(18)
    WH_TPL_CASE(20)
    WH_TPL_CASE(24)
    WH_TPL_CASE(30)
    WH_TPL_CASE(32)
    WH_TPL_CASE(36)
    WH_TPL_CASE(40)
    WH_TPL_CASE(45)
    WH_TPL_CASE(48)
    WH_TPL_CASE(50)
    WH_TPL_CASE(54)
    WH_TPL_CASE(60)
    WH_TPL_CASE(64)
    WH_TPL_CASE(72)
    WH_TPL_CASE(75)
    WH_TPL_CASE(80)
    WH_TPL_CASE(90)
    WH_TPL_CASE(96)
    WH_TPL_CASE(100)
    WH_TPL_CASE(108)
    default:
      return nullptr;
  }
#undef WH_TPL_CASE
}

const TplKernels32* table32_for(int n) {
#define WH_TPL32_CASE(NN)                                                                  \
  case NN: {                                                                               \
    static const TplKernels32 k{&Launch32<NN>::zinv, &Launch32<NN>::zfwd, &Launch32<NN>::vloc, max_occ<float, NN>()}; \
    return &k;                                                                             \
  }
  switch (n) {
    WH_TPL32_CASE(18)
    WH_TPL32_CASE(20)
    WH_TPL32_CASE(24)
    WH_TPL32_CASE(30)
    WH_TPL32_CASE(32)
    WH_TPL32_CASE(36)
    WH_TPL32_CASE(40)
    WH_TPL32_CASE(45)
    WH_TPL32_CASE(48)
    WH_TPL32_CASE(50)
    WH_TPL32_CASE(54)
    WH_TPL32_CASE(60)
    WH_TPL32_CASE(64)
    WH_TPL32_CASE(72)
    WH_TPL32_CASE(75)
    WH_TPL32_CASE(80)
    WH_TPL32_CASE(90)
    WH_TPL32_CASE(96)
    WH_TPL32_CASE(100)
    WH_TPL32_CASE(108)
    default:
      return nullptr;
  }
#undef WH_TPL32_CASE
}

}  // namespace

const TplKernels32* tpl_kernels_f32(const StickPlanDev& P) {
  if (P.n1 != P.n2 || P.n2 != P.n3) return nullptr;
  if (P.nrows > P.n1 / 2 + 2) return nullptr;
  return table32_for(P.n1);
}

const TplKernels* tpl_kernels_f64(const StickPlanDev& P) {
  if (P.n1 != P.n2 || P.n2 != P.n3) return nullptr;
  if (P.nrows > P.n1 / 2 + 2) return nullptr;
  return table_for<double>(P.n1);
}

}  // namespace wh
