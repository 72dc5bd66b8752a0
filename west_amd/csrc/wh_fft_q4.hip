// FP64 plane kernel of the local-potential transform with each line held by
// a quad of lanes.
//
// Same step and data layout as k_plane_vloc_t (one a3-plane of sticks:
// inverse along a2 on the |i| <= imax rows, inverse along a1, times V(r),
// forward along a1 keeping |i| <= imax, forward along a2).  An N-point line
// (N = 4 M) is split over the four lanes of a quad, lane l holding the M
// points l + 4 m:
//   inverse: a length-M DFT in registers over m, the twiddles
//            exp(+2 pi i l k1 / N), then the length-4 DFT across the quad
//            (two butterfly stages exchanging data with DPP quad
//            permutations), leaving lane l with the points k1 + M rev(l)
//            (rev: the 2-bit reversal);
//   forward: the same steps in the opposite order with conjugate factors,
//            from that layout back to l + 4 m.
// So an FP64 line needs 4 M complex values per quad instead of one thread
// holding all N (which fits one wave per SIMD only), the transforms need no
// LDS passes and no barriers, and LDS carries only the rows of the planes
// between the a2 and the a1 phases.  A workgroup takes PZ planes of one item.
//
// Compiled with -fno-signed-zeros -ffinite-math-only (like the register-line
// kernels): the arithmetic on inputs known to be zero (|i| > IMX) folds away.
#include <map>
#include <mutex>

#include "wh_fft.h"
#include "wh_fft_dev.h"
#include "wh_fft_q4.h"

namespace wh {

namespace {

// compile-time bound on the sphere's |i|, |j| (as for the register-line kernels)
template <int N>
constexpr int q4_imx() {
  return (N + 3) / 4;
}

// value of lane l ^ 1 (X = 1) or l ^ 2 (X = 2) of the quad
template <int X>
__device__ __forceinline__ double quad_xor(double v) {
  constexpr int ctrl = X == 1 ? 0xB1 : 0x4E;   // quad_perm [1,0,3,2] / [2,3,0,1]
  const long long b = __double_as_longlong(v);
  // every source lane exists (a permutation within the quad): no old value
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), ctrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), ctrl, 0xF, 0xF, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// butterfly with the partner lane l ^ X: v <- p + v (bit clear), p - v (bit set)
template <int X>
__device__ __forceinline__ z64 quad_bfly(z64 v, double s) {
  const double px = quad_xor<X>(v.x);
  const double py = quad_xor<X>(v.y);
  return mk(fma(s, v.x, px), fma(s, v.y, py));
}

// the length-4 DFT across the quad; INV: natural input, bit-reversed output,
// exp(+i); otherwise bit-reversed input, natural output, exp(-i)
template <int M, bool INV>
__device__ __forceinline__ void quad_dft4(z64* v, int l) {
  const double s1 = (l & 1) ? -1.0 : 1.0;
  const double s2 = (l & 2) ? -1.0 : 1.0;
  // lane 3's factor between the stages: +i (inverse) or -i (forward)
  const double c = l == 3 ? 0.0 : 1.0;
  const double d = l == 3 ? (INV ? 1.0 : -1.0) : 0.0;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    z64 u = v[k];
    u = INV ? quad_bfly<2>(u, s2) : quad_bfly<1>(u, s1);
    u = mk(c * u.x - d * u.y, c * u.y + d * u.x);
    u = INV ? quad_bfly<1>(u, s1) : quad_bfly<2>(u, s2);
    v[k] = u;
  }
}

// v[k] *= exp(DIR 2 pi i l k / N), the factors read from the workgroup's
// table tw[l M + k] = exp(+2 pi i l k / N) (four addresses per wave: the
// reads are broadcasts)
template <int M, int DIR>
__device__ __forceinline__ void quad_twiddle(z64* v, const z64* tw) {
#pragma unroll
  for (int k = 1; k < M; ++k) {
    const z64 t = tw[k];
    v[k] = cmul(v[k], DIR > 0 ? t : mk(t.x, -t.y));
  }
}

__device__ __forceinline__ int rev2(int l) { return ((l & 1) << 1) | (l >> 1); }

// swap of a double between the two halves of the wave (X = 32: lanes l and
// l ^ 32 exchange the value of a in the upper half with b in the lower half)
// or between neighbouring rows of 16 lanes (X = 16); afterwards every lane
// holds in a the value of the pair's lower lane and in b that of its upper
// lane (lower / upper: lane bit log2 X clear / set)
template <int X>
__device__ __forceinline__ void xswap(double& a, double& b) {
  const long long ia = __double_as_longlong(a), ib = __double_as_longlong(b);
  const unsigned alo = static_cast<unsigned>(ia), ahi = static_cast<unsigned>(ia >> 32);
  const unsigned blo = static_cast<unsigned>(ib), bhi = static_cast<unsigned>(ib >> 32);
  unsigned rlo0, rlo1, rhi0, rhi1;
  if constexpr (X == 32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    rlo0 = lo[0]; rlo1 = lo[1]; rhi0 = hi[0]; rhi1 = hi[1];
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    rlo0 = lo[0]; rlo1 = lo[1]; rhi0 = hi[0]; rhi1 = hi[1];
  }
  a = __longlong_as_double((static_cast<long long>(rhi0) << 32) | rlo0);
  b = __longlong_as_double((static_cast<long long>(rhi1) << 32) | rlo1);
}

template <int X>
__device__ __forceinline__ void xswap(z64& a, z64& b) {
  xswap<X>(a.x, b.x);
  xswap<X>(a.y, b.y);
}

// Where the four lanes of a line sit and how they combine.
// Quad layout (ROWL false): lanes 4 g + l (l = 0..3), the length-4 DFT by
// butterflies with DPP-exchanged partners; the transformed point k1 + M rev(l)
// is in v[k1].
// Row layout (ROWL true): lanes 16 w + g + 16 l of wave w (l = b0 + 2 b1 from
// lane bits 4 and 5); each stage of the length-4 DFT first swaps registers
// between the lanes of a pair (permlane swaps: one instruction moves a dword
// each way), which turns the lane bit into a register index, and then does
// the butterfly in registers, where the factor i is a free exchange of real
// and imaginary parts.  With H = M / 2 and Q = H / 2:
//   bit 5 (b1): registers (e, e + H) pair, e < H; afterwards register c H + j
//     holds the c-th output (c = 0, 1) of element j + H b1;
//   bit 4 (b0): registers (c H + j, c H + j + Q) pair, j < Q; afterwards the
//     two hold the points of element j + Q b0 + H b1 with k2 = c and c + 2;
//     for odd H the registers 2Q and H + 2Q pair with each other and hold
//     element 2Q + H b1 with k2 = b0 + 2 c' (c' = 0 for register 2Q, 1 for
//     H + 2Q), the factor i selected per lane;
// the transformed point x = k1 + M k2 of register r is pos<r>().
template <int M, bool ROWL>
struct Lane4 {
  static constexpr int H = M / 2;
  static constexpr int Q = H / 2;
  static constexpr bool ODD = (H % 2) != 0;
  static_assert(!ROWL || M % 2 == 0, "row layout needs an even M");
  int l;      // this lane's quarter of the line (0..3)
  int off;    // position offset of the paired registers
  int off2;   // position offset of the odd-H registers (row layout)

  __device__ explicit Lane4(int tid) {
    if constexpr (ROWL) {
      l = (tid >> 4) & 3;
      const int b0 = l & 1, b1 = l >> 1;
      off = Q * b0 + H * b1;
      off2 = H * b1 + M * b0;
    } else {
      l = tid & 3;
      off = M * rev2(l);
      off2 = off;
    }
  }
  static __device__ int line(int tid) { return ROWL ? (tid >> 6) * 16 + (tid & 15) : tid >> 2; }

  template <int R>
  static constexpr int pos_const() {
    if constexpr (!ROWL) {
      return R;
    } else {
      constexpr int c = R / H, j = R % H;
      if constexpr (j < 2 * Q) return j % Q + M * (c + 2 * (j >= Q ? 1 : 0));
      else return 2 * Q + 2 * M * c;
    }
  }
  template <int R>
  __device__ int pos() const {
    if constexpr (ROWL && ODD && R % H == 2 * Q) return pos_const<R>() + off2;
    else return pos_const<R>() + off;
  }

  // length-4 DFT across the line's lanes: inverse (exp(+)) from lane l's
  // register k1 to the points pos<r>; forward (exp(-)) back
  __device__ void inv(z64* v) const {
    if constexpr (!ROWL) {
      quad_dft4<M, true>(v, l);
    } else {
      sfor<H>([&](auto E) {
        constexpr int e = decltype(E)::value;
        xswap<32>(v[e], v[e + H]);
        const z64 t = v[e];
        v[e] = t + v[e + H];
        v[e + H] = t - v[e + H];
      });
      sfor<2 * Q>([&](auto E) {
        constexpr int c = decltype(E)::value / Q, j = decltype(E)::value % Q;
        constexpr int a = c * H + j, b = a + Q;
        xswap<16>(v[a], v[b]);
        const z64 t = c == 0 ? v[b] : mk(-v[b].y, v[b].x);
        const z64 u = v[a];
        v[a] = u + t;
        v[b] = u - t;
      });
      if constexpr (ODD) {
        constexpr int a = 2 * Q, b = H + 2 * Q;
        xswap<16>(v[a], v[b]);
        const bool rot = (l & 1) != 0;
        const z64 t = mk(rot ? -v[b].y : v[b].x, rot ? v[b].x : v[b].y);
        const z64 u = v[a];
        v[a] = u + t;
        v[b] = u - t;
      }
    }
  }
  __device__ void fwd(z64* v) const {
    if constexpr (!ROWL) {
      quad_dft4<M, false>(v, l);
    } else {
      sfor<2 * Q>([&](auto E) {
        constexpr int c = decltype(E)::value / Q, j = decltype(E)::value % Q;
        constexpr int a = c * H + j, b = a + Q;
        const z64 d = v[a] - v[b];
        v[a] = v[a] + v[b];
        v[b] = c == 0 ? d : mk(d.y, -d.x);
        xswap<16>(v[a], v[b]);
      });
      if constexpr (ODD) {
        constexpr int a = 2 * Q, b = H + 2 * Q;
        const z64 d = v[a] - v[b];
        v[a] = v[a] + v[b];
        const bool rot = (l & 1) != 0;
        v[b] = mk(rot ? d.y : d.x, rot ? -d.x : d.y);
        xswap<16>(v[a], v[b]);
      }
      sfor<H>([&](auto E) {
        constexpr int e = decltype(E)::value;
        const z64 d = v[e] - v[e + H];
        v[e] = v[e] + v[e + H];
        v[e + H] = d;
        xswap<32>(v[e], v[e + H]);
      });
    }
  }
};

// LDS rows of PZ planes: NR = 2 IMX + 1 rows (row IMX + i for a1 frequency i),
// stride RS = N + 1 (odd: row-parallel accesses spread over the banks)
template <int N, int PZ, int NTQ, int WPE, bool ROWL>
__global__ __launch_bounds__(NTQ) __attribute__((amdgpu_waves_per_eu(WPE))) void k_plane_vloc_q4(StickPlanDev P, const FftItem* items, const z64* sin,
                                                      z64* sout, const double* vpot, long ldpot, double scale) {
  constexpr int M = N / 4;
  constexpr int IMX = q4_imx<N>();
  constexpr int NR = 2 * IMX + 1;
  constexpr int RS = N + 1;
  constexpr int PE = NR * RS;
  static_assert(N % 4 == 0, "lines split over a quad need N = 4 M");
  extern __shared__ char smem_raw[];
  z64* pl = reinterpret_cast<z64*>(smem_raw);   // [PZ][NR][RS]
  const int nrows = P.nrows;
  const int r0 = IMX - P.imax;
  const int nct = P.nct;
  const int z0 = blockIdx.x * PZ;
  const int nz = min(PZ, N - z0);
  const FftItem item = items[blockIdx.y];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const long sl = static_cast<long>(N) * nct;
  const int tid = threadIdx.x;
  const Lane4<N / 4, ROWL> ln(tid);
  const int l = ln.l;
  const int quad = Lane4<N / 4, ROWL>::line(tid);
  constexpr int NQ = NTQ / 4;
  static_assert(!ROWL || NTQ % 64 == 0, "row layout needs whole waves");

  // this lane's twiddle factors after the planes
  z64* tw = pl + PZ * PE;
  const z64* twl = tw + l * M;
  if (tid < 4 * M) {
    double sn, cs;
    sincospi(2.0 * ((tid / M) * (tid % M)) / N, &sn, &cs);
    tw[tid] = mk(cs, sn);
  }
  // sticks of planes z0 .. z0 + nz - 1 into zeroed rows (loading a batch of
  // elements per thread before the rows are cleared measured slower)
  const z64* src = sin + static_cast<long>(item.a) * sl + static_cast<long>(z0) * nct;
  const int ntot = nz * nct;
  for (int t = tid; t < nz * PE; t += NTQ) pl[t] = mk(0.0, 0.0);
  __syncthreads();
  for (int t = tid; t < ntot; t += NTQ) {
    const int p = t / nct;
    const int pos = P.col_pos[t - p * nct];
    const int r = pos / N;
    pl[p * PE + (r0 + r) * RS + (pos - r * N)] = src[t];
  }
  __syncthreads();

  // a2 lines (rows of the sphere), inverse: points j = l + 4 m in, k1 + M rev(l) out
  for (int q = quad; q < nz * nrows; q += NQ) {
    const int p = q / nrows;
    z64* row = pl + p * PE + (r0 + q - p * nrows) * RS;
    z64 v[M];
    sfor<M>([&](auto Mi) {
      constexpr int m = decltype(Mi)::value;
      constexpr int lo = 4 * m, hi = 4 * m + 3;
      if constexpr (lo > IMX && hi < N - IMX) v[m] = mk(0.0, 0.0);   // outside every sphere column
      else v[m] = row[l + 4 * m];
    });
    fft_line<double, M, +1>(v);
    quad_twiddle<M, +1>(v, twl);
    ln.inv(v);
    sfor<M>([&](auto Ki) { row[ln.template pos<decltype(Ki)::value>()] = v[decltype(Ki)::value]; });
  }
  __syncthreads();

  // a1 lines: inverse from the |i| <= IMX rows, V(r), forward back to them
  {
    const double sc = scale;
    const double* V0 = vpot + (item.pot >= 0 ? static_cast<long>(item.pot) * ldpot : 0L);
    for (int q = quad; q < nz * N; q += NQ) {
      const int p = q / N;
      const int y = q - p * N;
      z64* col = pl + p * PE + IMX * RS + y;   // row of frequency i = 0
      z64 v[M];
      sfor<M>([&](auto Mi) {
        constexpr int m = decltype(Mi)::value;
        constexpr int lo = 4 * m, hi = 4 * m + 3;
        if constexpr (hi <= IMX) {
          v[m] = col[(l + 4 * m) * RS];
        } else if constexpr (lo >= N - IMX) {
          v[m] = col[(l + 4 * m - N) * RS];
        } else if constexpr (lo > IMX && hi < N - IMX) {
          v[m] = mk(0.0, 0.0);
        } else {
          const int i = l + 4 * m;
          v[m] = i <= IMX ? col[i * RS] : (i >= N - IMX ? col[(i - N) * RS] : mk(0.0, 0.0));
        }
      });
      fft_line<double, M, +1>(v);
      quad_twiddle<M, +1>(v, twl);
      ln.inv(v);
      // register k holds the point x = pos<k>; V(r) is loaded here, not
      // ahead of the inverse transform (its 2 M registers would be live
      // through it)
      __builtin_amdgcn_sched_barrier(0);
      const double* vr = V0 + (static_cast<long>(z0 + p) * N + y) * N;
      sfor<M>([&](auto Ki) {
        constexpr int k = decltype(Ki)::value;
        v[k] = cscale(v[k], vr[ln.template pos<k>()] * sc);
      });
      ln.fwd(v);
      quad_twiddle<M, -1>(v, twl);
      fft_line<double, M, -1>(v);
      // lane l holds frequencies i = l + 4 m; only |i| <= IMX are rows
      sfor<M>([&](auto Mi) {
        constexpr int m = decltype(Mi)::value;
        constexpr int lo = 4 * m, hi = 4 * m + 3;
        if constexpr (hi <= IMX) {
          col[(l + 4 * m) * RS] = v[m];
        } else if constexpr (lo >= N - IMX) {
          col[(l + 4 * m - N) * RS] = v[m];
        } else if constexpr (!(lo > IMX && hi < N - IMX)) {
          const int i = l + 4 * m;
          if (i <= IMX) col[i * RS] = v[m];
          else if (i >= N - IMX) col[(i - N) * RS] = v[m];
        }
      });
    }
  }
  __syncthreads();

  // a2 lines, forward: points k1 + M rev(l) in, j = l + 4 m out (|j| <= IMX kept)
  for (int q = quad; q < nz * nrows; q += NQ) {
    const int p = q / nrows;
    z64* row = pl + p * PE + (r0 + q - p * nrows) * RS;
    z64 v[M];
    sfor<M>([&](auto Ki) { v[decltype(Ki)::value] = row[ln.template pos<decltype(Ki)::value>()]; });
    ln.fwd(v);
    quad_twiddle<M, -1>(v, twl);
    fft_line<double, M, -1>(v);
    // every lane of the line has read its points of the row (one wave: the
    // loads above complete before the exchanges that depend on them)
    sfor<M>([&](auto Mi) {
      constexpr int m = decltype(Mi)::value;
      constexpr int lo = 4 * m, hi = 4 * m + 3;
      if constexpr (!(lo > IMX && hi < N - IMX)) row[l + 4 * m] = v[m];
    });
  }
  __syncthreads();

  z64* dst = sout + static_cast<long>(blockIdx.y) * sl + static_cast<long>(z0) * nct;
  for (int t = tid; t < ntot; t += NTQ) {
    const int p = t / nct;
    const int pos = P.col_pos[t - p * nct];
    const int r = pos / N;
    dst[t] = pl[p * PE + (r0 + r) * RS + (pos - r * N)];
  }
}

// LDS bytes: the planes' rows and the twiddle table
size_t q4_smem(int n, int pz) {
  const int imx = (n + 3) / 4;
  return sizeof(z64) * (static_cast<size_t>(pz) * (2 * imx + 1) * (n + 1) + n);
}

constexpr size_t kQ4MaxSmem = 160u * 1024u;

template <int N, int PZ, int NTQ, int WPE, bool ROWL>
void launch_q4(const StickPlanDev& P, const FftItem* it, int n, const z64* si, z64* so, const double* v, long ldv,
               double scale, hipStream_t s) {
  static std::mutex mu;
  static bool set = false;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!set) {
      WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_plane_vloc_q4<N, PZ, NTQ, WPE, ROWL>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(kQ4MaxSmem)));
      set = true;
    }
  }
  const size_t sm = q4_smem(N, PZ);
  k_plane_vloc_q4<N, PZ, NTQ, WPE, ROWL><<<dim3(ceil_div(N, PZ), n), NTQ, sm, s>>>(P, it, si, so, v, ldv, scale);
}

// 3 planes per workgroup: 768 threads in the quad layout, 1024 in the row
// layout (4 waves per SIMD, 128 VGPRs)
constexpr int kQ4Variants = 2;

}  // namespace

int plane_q4_variants(const StickPlanDev& P) {
  if (P.n1 != P.n2 || P.n1 != P.n3) return 0;
  if (P.n1 != 72) return 0;
  if (P.imax > (P.n1 + 3) / 4 || q4_smem(P.n1, 3) > kQ4MaxSmem) return 0;
  return kQ4Variants;
}

void plane_q4_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z64* sin, z64* sout,
                   const double* vpot, long ldpot, double scale, hipStream_t s) {
  WH_REQUIRE(v >= 0 && v < plane_q4_variants(P), "plane_q4_vloc: no such kernel for this grid");
  WH_REQUIRE(static_cast<const void*>(sin) != static_cast<const void*>(sout), "plane_q4_vloc: input and output alias");
  switch (v) {
    case 0: return launch_q4<72, 3, 768, 3, false>(P, items, n, sin, sout, vpot, ldpot, scale, s);
    default: return launch_q4<72, 3, 1024, 4, true>(P, items, n, sin, sout, vpot, ldpot, scale, s);
  }
}

}  // namespace wh
