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

// LDS rows of PZ planes: NR = 2 IMX + 1 rows (row IMX + i for a1 frequency i),
// stride RS = N + 1 (odd: row-parallel accesses spread over the banks)
// IMX: compile-time bound on the sphere's |i| and |j| (at most N/4; the
// rows beyond the actual imax stay zero)
template <int N, int IMX, int PZ, int NTQ, int WPE>
__global__ __launch_bounds__(NTQ) __attribute__((amdgpu_waves_per_eu(WPE))) void k_plane_vloc_q4(StickPlanDev P, const FftItem* items, const z64* sin,
                                                      z64* sout, const double* vpot, long ldpot, double scale) {
  constexpr int M = N / 4;
  static_assert(IMX <= q4_imx<N>(), "the pruned lines need |i|, |j| <= N/4");
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
  const int l = tid & 3;
  const int quad = tid >> 2;
  constexpr int NQ = NTQ / 4;
  // after the inverse length-4 DFT register k holds the point k + off
  const int off = M * rev2(l);

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
    quad_dft4<M, true>(v, l);
    sfor<M>([&](auto Ki) { row[decltype(Ki)::value + off] = v[decltype(Ki)::value]; });
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
      quad_dft4<M, true>(v, l);
      // register k holds the point x = k + off; V(r) is loaded here, not
      // ahead of the inverse transform (its 2 M registers would be live
      // through it)
      __builtin_amdgcn_sched_barrier(0);
      const double* vr = V0 + (static_cast<long>(z0 + p) * N + y) * N;
      sfor<M>([&](auto Ki) {
        constexpr int k = decltype(Ki)::value;
        v[k] = cscale(v[k], vr[k + off] * sc);
      });
      quad_dft4<M, false>(v, l);
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
    sfor<M>([&](auto Ki) { v[decltype(Ki)::value] = row[decltype(Ki)::value + off]; });
    quad_dft4<M, false>(v, l);
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
size_t q4_smem(int n, int imx, int pz) {
  return sizeof(z64) * (static_cast<size_t>(pz) * (2 * imx + 1) * (n + 1) + n);
}

constexpr size_t kQ4MaxSmem = 160u * 1024u;

template <int N, int IMX, int PZ, int NTQ, int WPE>
void launch_q4(const StickPlanDev& P, const FftItem* it, int n, const z64* si, z64* so, const double* v, long ldv,
               double scale, hipStream_t s) {
  static std::mutex mu;
  static bool set = false;
  const size_t sm = q4_smem(N, IMX, PZ);
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!set) {
      WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_plane_vloc_q4<N, IMX, PZ, NTQ, WPE>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(sm)));
      set = true;
    }
  }
  k_plane_vloc_q4<N, IMX, PZ, NTQ, WPE><<<dim3(ceil_div(N, PZ), n), NTQ, sm, s>>>(P, it, si, so, v, ldv, scale);
}

// variants (sphere bound, planes per workgroup, threads): 0: (18, 3, 768),
// one workgroup of 130 KB per CU, one a1 column per quad and plane in the
// first pass; 1: (16, 3, 768), 4 fewer LDS rows per plane and more zeros
// pruned, where the sphere spans |i|, |j| <= 16 (the Si64 cell).  Two
// planes of 16-bounded rows (78 KB: two workgroups of 384 threads per CU)
// measured 2.38 ms against 1.85 ms per 1,024 items and are not offered.
bool q4_tight(const StickPlanDev& P) { return P.imax <= 16 && P.jext <= 16; }

}  // namespace

int plane_q4_variants(const StickPlanDev& P) {
  if (P.n1 != P.n2 || P.n1 != P.n3) return 0;
  if (P.n1 != 72 && P.n1 != 60 && P.n1 != 100) return 0;
  // the a1 and a2 transforms take |i| and |j| beyond N/4 as known zeros
  const int imx = (P.n1 + 3) / 4;
  if (P.imax > imx || P.jext > imx || q4_smem(P.n1, imx, P.n1 == 100 ? 1 : 3) > kQ4MaxSmem) return 0;
  return (P.n1 == 72 && q4_tight(P)) ? 2 : 1;
}

void plane_q4_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z64* sin, z64* sout,
                   const double* vpot, long ldpot, double scale, hipStream_t s) {
  WH_REQUIRE(v >= 0 && v < plane_q4_variants(P), "plane_q4_vloc: no such kernel for this grid");
  WH_REQUIRE(static_cast<const void*>(sin) != static_cast<const void*>(sout), "plane_q4_vloc: input and output alias");
  if (P.n1 == 60) launch_q4<60, 15, 3, 768, 3>(P, items, n, sin, sout, vpot, ldpot, scale, s);
  else if (P.n1 == 100) launch_q4<100, 25, 1, 448, 2>(P, items, n, sin, sout, vpot, ldpot, scale, s);
  else if (v == 0) launch_q4<72, 18, 3, 768, 3>(P, items, n, sin, sout, vpot, ldpot, scale, s);
  else launch_q4<72, 16, 3, 768, 3>(P, items, n, sin, sout, vpot, ldpot, scale, s);
}

}  // namespace wh
