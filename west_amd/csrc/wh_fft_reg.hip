// Plane kernel of the local-potential transform with register-resident lines.
//
// Same step and data layout as k_plane_vloc_t of wh_fft_tpl.hip (one a3-plane
// of sticks: inverse FFT along a2 on the |i| <= imax rows, inverse FFT along
// a1, times V(r), forward along a1 keeping |i| <= imax, forward along a2), but
// every 1D transform is done whole by one thread with its N points in
// registers: all radix passes, twiddles (compile-time constants) and the
// V(r) multiply happen without LDS traffic, and LDS only carries the two
// transposes between the a2 and the a1 lines.  A workgroup takes PZ planes
// of one item so that its PZ * N a1-lines fill whole wavefronts.
//
// The transforms know at compile time which of their inputs are zero
// (|i| > IMX along a1, |j| > IMX along a2: the sphere bound) and which
// outputs are dropped; this file is
// compiled with -fno-signed-zeros -ffinite-math-only so that the arithmetic
// on those known zeros folds away.  Neither flag changes any finite result.
//
// The FP64 plane kernels are built from this file in a translation unit of
// their own (wh_fft_reg64.hip defines WH_REG_FP64), so that the two sets of
// instantiations compile in parallel.
#include <map>
#include <mutex>
#include <utility>

#include "wh_fft.h"
#include "wh_fft_dev.h"
#include "wh_fft_reg.h"

namespace wh {

namespace {

// compile-time bound on imax the a1 pruning assumes (the smooth grid of a
// norm-conserving calculation has N >= 4 imax + 1)
template <int N>
constexpr int imx() {
  return (N + 3) / 4;
}

template <int N, int PZ>
constexpr int nt_reg() {
  return (PZ * N + 63) / 64 * 64;
}

// PT: storage type of the potential; PZ: planes per workgroup.
// LDS holds PZ planes of NR = 2 IMX + 1 rows (row IMX + i for a1 frequency i,
// |i| <= IMX; rows beyond the sphere's |i| <= imax stay zero), so that the a1
// lines load and store their rows at compile-time offsets without branches.
template <typename T, int N, typename PT, int PZ>
// (FP64: the registers of one wave per SIMD are allowed for)
__global__ __launch_bounds__((nt_reg<N, PZ>())) __attribute__((amdgpu_waves_per_eu((sizeof(T) == 8 ? 1 : 2)))) void k_plane_vloc_reg(
    StickPlanDev P, const FftItem* items, const cplx<T>* sin, cplx<T>* sout, const PT* vpot, long ldpot,
    double scale) {
  constexpr int RS = N + 1;   // row stride: odd, so row-parallel accesses spread over all banks
  constexpr int IMX = imx<N>();
  constexpr int NR = 2 * IMX + 1;
  constexpr int PE = NR * RS;   // elements per plane
  constexpr int NTR = nt_reg<N, PZ>();
  extern __shared__ char smem_raw[];
  cplx<T>* pl = reinterpret_cast<cplx<T>*>(smem_raw);   // [PZ][NR][RS]
  const int nrows = P.nrows;
  const int r0 = IMX - P.imax;   // first row of the sphere
  const int nct = P.nct;
  const int z0 = blockIdx.x * PZ;
  const int nz = min(PZ, N - z0);
  const FftItem item = items[blockIdx.y];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const long sl = static_cast<long>(N) * nct;
  const int tid = threadIdx.x;

  // sticks of planes z0 .. z0 + nz - 1 (contiguous) into zeroed rows
  for (int t = tid; t < nz * PE; t += NTR) pl[t] = mk<T>(T(0), T(0));
  __syncthreads();
  {
    const cplx<T>* src = sin + static_cast<long>(item.a) * sl + static_cast<long>(z0) * nct;
    for (int t = tid; t < nz * nct; t += NTR) {
      const int p = t / nct;
      const int pos = P.col_pos[t - p * nct];
      const int r = pos / N;
      pl[p * PE + (r0 + r) * RS + (pos - r * N)] = src[t];
    }
  }
  __syncthreads();

  // a2 lines (rows of the sphere): inverse
  for (int t = tid; t < nz * nrows; t += NTR) {
    const int p = t / nrows;
    cplx<T>* row = pl + p * PE + (r0 + t - p * nrows) * RS;
    cplx<T> v[N];
    // |j| > IMX: known zeros (no stored column there)
    sfor<N>([&](auto Y) {
      constexpr int y = decltype(Y)::value;
      if constexpr (y > IMX && y < N - IMX) v[y] = mk<T>(T(0), T(0));
      else v[y] = row[y];
    });
    fft_line<T, N, +1>(v);
    sfor<N>([&](auto Y) { row[decltype(Y)::value] = v[decltype(Y)::value]; });
  }
  __syncthreads();

  // a1 lines: inverse from the |i| <= IMX rows, V(r), forward back to them
  {
    const T sc = T(scale);
    const PT* V0 = vpot + (item.pot >= 0 ? static_cast<long>(item.pot) * ldpot : 0L);
    for (int t = tid; t < nz * N; t += NTR) {
      const int p = t / N;
      const int y = t - p * N;
      cplx<T>* col = pl + p * PE + IMX * RS + y;   // row of i = 0
      cplx<T> v[N];
      sfor<N>([&](auto X) {
        constexpr int x = decltype(X)::value;
        if constexpr (x <= IMX) {
          v[x] = col[x * RS];
        } else if constexpr (x >= N - IMX) {
          v[x] = col[-(N - x) * RS];
        } else {
          v[x] = mk<T>(T(0), T(0));
        }
      });
      fft_line<T, N, +1>(v);
      const PT* vr = V0 + (static_cast<long>(z0 + p) * N + y) * N;
      // keep the potential loads after the inverse transform (register pressure)
      __builtin_amdgcn_sched_barrier(0);
      sfor<N>([&](auto X) {
        constexpr int x = decltype(X)::value;
        v[x] = cscale(v[x], T(vr[x]) * sc);
      });
      // reversed radix order: of its outputs only |i| <= IMX are kept, which
      // prunes more of its last pass (and spills nothing at N = 72)
      fft_line<T, N, -1, true>(v);
      sfor<N>([&](auto X) {
        constexpr int x = decltype(X)::value;
        if constexpr (x <= IMX) {
          col[x * RS] = v[x];
        } else if constexpr (x >= N - IMX) {
          col[-(N - x) * RS] = v[x];
        }
      });
    }
  }
  __syncthreads();

  // a2 lines: forward
  for (int t = tid; t < nz * nrows; t += NTR) {
    const int p = t / nrows;
    cplx<T>* row = pl + p * PE + (r0 + t - p * nrows) * RS;
    cplx<T> v[N];
    sfor<N>([&](auto Y) { v[decltype(Y)::value] = row[decltype(Y)::value]; });
    fft_line<T, N, -1>(v);
    // only |j| <= IMX reach a stick (the others are dropped)
    sfor<N>([&](auto Y) {
      constexpr int y = decltype(Y)::value;
      if constexpr (!(y > IMX && y < N - IMX)) row[y] = v[y];
    });
  }
  __syncthreads();

  cplx<T>* dst = sout + static_cast<long>(blockIdx.y) * sl + static_cast<long>(z0) * nct;
  for (int t = tid; t < nz * nct; t += NTR) {
    const int p = t / nct;
    const int pos = P.col_pos[t - p * nct];
    const int r = pos / N;
    dst[t] = pl[p * PE + (r0 + r) * RS + (pos - r * N)];
  }
}

// planes per workgroup of variant v
constexpr int kPz[] = {2, 3, 4};
constexpr int kNumPz = 3;

size_t reg_smem(const StickPlanDev& P, int pz, size_t elem) {
  const int nr = 2 * ((P.n1 + 3) / 4) + 1;
  return elem * static_cast<size_t>(pz) * nr * (P.n1 + 1);
}

template <class K>
void allow_smem_reg(K kernel, size_t bytes) {
  static std::map<const void*, size_t> done;
  static std::mutex done_mutex;
  std::lock_guard<std::mutex> lk(done_mutex);
  auto key = reinterpret_cast<const void*>(kernel);
  auto itr = done.find(key);
  if (itr != done.end() && itr->second >= bytes) return;
  WH_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
  done[key] = bytes;
}

template <typename T, int N, typename PT, int PZ>
void launch_pz(const StickPlanDev& P, const FftItem* it, int n, const cplx<T>* si, cplx<T>* so, const PT* v, long ldv,
               double scale, hipStream_t s) {
  const size_t sm = reg_smem(P, PZ, sizeof(cplx<T>));
  allow_smem_reg(k_plane_vloc_reg<T, N, PT, PZ>, sm);
  k_plane_vloc_reg<T, N, PT, PZ><<<dim3(ceil_div(N, PZ), n), nt_reg<N, PZ>(), sm, s>>>(P, it, si, so, v, ldv, scale);
}

template <typename T, int N, typename PT>
void launch_n(int var, const StickPlanDev& P, const FftItem* it, int n, const cplx<T>* si, cplx<T>* so, const PT* v,
              long ldv, double scale, hipStream_t s) {
  switch (kPz[var]) {
    case 2: return launch_pz<T, N, PT, 2>(P, it, n, si, so, v, ldv, scale, s);
    case 3: return launch_pz<T, N, PT, 3>(P, it, n, si, so, v, ldv, scale, s);
    default: return launch_pz<T, N, PT, 4>(P, it, n, si, so, v, ldv, scale, s);
  }
}

// ---------------------------------------------------------------------------
// z transforms with register-resident lines (FP32 operator).  A workgroup
// takes kZregCols stored columns of one item: thread t < kZregCols holds the
// a3 line of stored column c0 + t, thread kZregCols + t the line of its mirror
// (none for column 0, whose line holds both halves).  Same layout and
// packing as k_zinv_t / k_zfwd_t of wh_fft_tpl.hip.
// ---------------------------------------------------------------------------
constexpr int ZNT = 2 * kZregCols;

// sphere -> sticks.  The block's points are staged in LDS with coalesced
// loads; each thread then builds its line at compile-time positions (the
// line's nonzeros lie at |k| <= IMX) and transforms it in registers.
template <int N>
__global__ __launch_bounds__(ZNT) __attribute__((amdgpu_waves_per_eu(2))) void k_zinv_reg(
    StickPlanDev P, const FftItem* items, const z32* psiA, long ldA, const z32* psiB, long ldB, z32* sticks, int mode) {
  constexpr int KX = imx<N>();
  extern __shared__ char smem_raw[];
  z32* pa = reinterpret_cast<z32*>(smem_raw);
  z32* pb = pa + P.zblk_pts;
  const int it = blockIdx.y;
  const int c0 = blockIdx.x * kZregCols;
  const int c1 = min(c0 + kZregCols, P.ncs);
  const int S0 = P.col_off[c0];
  const int npt = P.col_off[c1] - S0;
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  const z32* A = psiA + static_cast<long>(item.a) * ldA + S0;
  const z32* B = item.b >= 0 ? psiB + static_cast<long>(item.b) * ldB + S0 : nullptr;
  {
    // all of a thread's loads in flight at once (a column block holds at most
    // kZregCols (2 IMX + 1) points)
    constexpr int U = (kZregCols * (2 * KX + 1) + ZNT - 1) / ZNT;
    z32 ra[U], rb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = threadIdx.x + u * ZNT;
      if (t < npt) {
        ra[u] = A[t];
        rb[u] = B ? B[t] : mk<float>(0.f, 0.f);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = threadIdx.x + u * ZNT;
      if (t < npt) {
        pa[t] = ra[u];
        pb[t] = rb[u];
      }
    }
  }
  __syncthreads();
  const bool own = threadIdx.x < kZregCols;
  const int c = c0 + (own ? threadIdx.x : threadIdx.x - kZregCols);
  if (c >= c1 || (!own && c == 0)) return;   // no barrier below
  const int s0 = P.col_off[c] - S0;
  const int kmin = P.col_kmin[c];
  const int kmax = kmin + (P.col_off[c + 1] - P.col_off[c]) - 1;
  const bool gets_mirror = !own || c == 0;   // this line holds mirror values
  z32 v[N];
  sfor<N>([&](auto Z) {
    constexpr int z = decltype(Z)::value;
    if constexpr (z > KX && z < N - KX) {
      v[z] = mk<float>(0.f, 0.f);
    } else {
      constexpr int ko = z <= KX ? z : z - N;    // point whose own value lands at z
      constexpr int km = z <= KX ? -z : N - z;   // point whose mirror value lands at z
      const bool okm = gets_mirror && km >= kmin && km <= kmax;
      const bool oko = own && ko >= kmin && ko <= kmax;
      z32 val = mk<float>(0.f, 0.f);
      if (okm || oko) {
        const int q = s0 + (okm ? km : ko) - kmin;
        const z32 a1 = pa[q];
        const z32 a2 = pb[q];
        if (!okm) val = mk<float>(a1.x - a2.y, a1.y + a2.x);        // b(nl) = a1 + i a2
        else if (mode == ZINV_PAIR_REALG0 && c == 0 && km == 0) val = mk<float>(a1.x, a2.x);
        else val = mk<float>(a1.x + a2.y, a2.x - a1.y);              // b(nlm) = conj(a1 - i a2), last
      }
      v[z] = val;
    }
  });
  fft_line<float, N, +1>(v);
  z32* out = sticks + static_cast<long>(it) * N * P.nct + (own ? c : c + P.ncs - 1);
  sfor<N>([&](auto Z) { out[static_cast<long>(decltype(Z)::value) * P.nct] = v[decltype(Z)::value]; });
}

// sticks -> sphere (pair modes).  Each line is loaded (coalesced along the
// stored columns), transformed in registers, and its |k| <= IMX outputs go to
// LDS; the fp/fm unpack and epilogue then run over the block's points.
template <int N>
__global__ __launch_bounds__(ZNT) __attribute__((amdgpu_waves_per_eu(2))) void k_zfwd_reg(
    StickPlanDev P, const FftItem* items, const z32* sticks, const z32* in, long ldin, z32* out, long ldout,
    ZfwdArgs A) {
  constexpr int KX = imx<N>();
  constexpr int NJ = 2 * KX + 1;   // kept positions per line: j = k (k >= 0), KX - k (k < 0)
  extern __shared__ char smem_raw[];
  z32* X = reinterpret_cast<z32*>(smem_raw);   // [2 kZregCols lines][NJ | 1]
  constexpr int XS = NJ | 1;
  const int it = blockIdx.y;
  const int c0 = blockIdx.x * kZregCols;
  const int c1 = min(c0 + kZregCols, P.ncs);
  const FftItem item = items[it];
  if (item.a < 0) return;   // inactive item (uniform per workgroup)
  {
    const bool own = threadIdx.x < kZregCols;
    const int lc = own ? threadIdx.x : threadIdx.x - kZregCols;
    const int c = c0 + lc;
    if (c < c1 && (own || c != 0)) {
      const z32* src = sticks + static_cast<long>(it) * N * P.nct + (own ? c : c + P.ncs - 1);
      z32 v[N];
      sfor<N>([&](auto Z) { v[decltype(Z)::value] = src[static_cast<long>(decltype(Z)::value) * P.nct]; });
      fft_line<float, N, -1>(v);
      z32* row = X + (2 * lc + (own ? 0 : 1)) * XS;
      sfor<N>([&](auto Z) {
        constexpr int z = decltype(Z)::value;
        if constexpr (z <= KX) row[z] = v[z];
        else if constexpr (z >= N - KX) row[KX + N - z] = v[z];
      });
    }
  }
  __syncthreads();
  const int s0 = P.col_off[c0];
  const int s1 = P.col_off[c1];
  z32* oa = out + static_cast<long>(item.a) * ldout;
  z32* ob = item.b >= 0 ? out + static_cast<long>(item.b) * ldout : nullptr;
  // the block's points, U per thread with their loads issued together
  constexpr int U = (kZregCols * (2 * KX + 1) + ZNT - 1) / ZNT;
  const bool stern = A.mode == ZFWD_PAIR_STERN;
  const z32* ina = in + static_cast<long>(item.a) * ldin;
  const z32* inb = item.b >= 0 ? in + static_cast<long>(item.b) * ldin : nullptr;
  const double ea = stern ? A.e[item.a] : 0.0;
  const double eb = (stern && ob) ? A.e[item.b] : 0.0;
  int pc[U], pk[U];
  double g2[U];
  z32 ia[U], ib[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = s0 + threadIdx.x + u * ZNT;
    if (s < s1) {
      pc[u] = P.pt_col[s];
      pk[u] = P.pt_k[s];
      if (stern) {
        g2[u] = A.g2kin[s];
        ia[u] = ina[s];
        if (inb) ib[u] = inb[s];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int s = s0 + threadIdx.x + u * ZNT;
    if (s < s1) {
      const int lc = pc[u] - c0;
      const int k = pk[u];
      const int jz = k >= 0 ? k : KX - k;    // position of k in its own line
      const int jm = k <= 0 ? -k : KX + k;   // position of -k in the mirror line
      const z32 F = X[(2 * lc) * XS + jz];
      const int ml = (c0 + lc == 0) ? 0 : 2 * lc + 1;
      const z32 Fm = X[ml * XS + jm];
      const z32 fp = cscale(F + Fm, 0.5f);
      const z32 fm = cscale(F - Fm, 0.5f);
      const z32 b1 = mk<float>(fp.x, fm.y);
      const z32 b2 = mk<float>(fp.y, -fm.x);
      if (!stern) {
        oa[s] = cscale(b1, float(A.scale));
        if (ob) ob[s] = cscale(b2, float(A.scale));
      } else {
        oa[s] = b1 + cscale(ia[u], float(g2[u] - ea));
        if (ob) ob[s] = b2 + cscale(ib[u], float(g2[u] - eb));
      }
    }
  }
}

template <int N>
void zinv_n(const StickPlanDev& P, const FftItem* it, int n, const z32* a, long lda, const z32* b, long ldb,
            z32* st, int mode, hipStream_t s) {
  const size_t sm = 2 * sizeof(z32) * static_cast<size_t>(P.zblk_pts);
  allow_smem_reg(k_zinv_reg<N>, sm);
  k_zinv_reg<N><<<dim3(ceil_div(P.ncs, kZregCols), n), ZNT, sm, s>>>(P, it, a, lda, b, ldb, st, mode);
}

template <int N>
void zfwd_n(const StickPlanDev& P, const FftItem* it, int n, const z32* st, const z32* in, long ldin, z32* out,
            long ldout, const ZfwdArgs& A, hipStream_t s) {
  constexpr int XS = (2 * imx<N>() + 1) | 1;
  const size_t sm = sizeof(z32) * 2 * kZregCols * XS;
  allow_smem_reg(k_zfwd_reg<N>, sm);
  k_zfwd_reg<N><<<dim3(ceil_div(P.ncs, kZregCols), n), ZNT, sm, s>>>(P, it, st, in, ldin, out, ldout, A);
}

// compiled line lengths: FP32, and FP64 up to 64 points.  FP64 lines of 36
// points or fewer fit the vector registers at 2-4 waves per SIMD; longer
// ones also take accumulation registers and run one wave per SIMD, which
// still beats the LDS-pass kernel up to 60 points (by 31 % at 60^3) but not
// at 72 (5.4 against 5.0 ms per 2,206 planes), so 72 and up are not built.
#define WH_REG_SIZES(X) X(18) X(20) X(24) X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75) X(80)
#define WH_REG64_SIZES(X) X(18) X(20) X(24) X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64)

bool compiled(int n, bool fp32) {
#define WH_REG_CASE(NN) \
  case NN:              \
    return true;
  if (!fp32) {
    switch (n) {
      WH_REG64_SIZES(WH_REG_CASE)
      default: return false;
    }
  }
  switch (n) {
    WH_REG_SIZES(WH_REG_CASE)
    default: return false;
  }
#undef WH_REG_CASE
}

#ifdef WH_REG_FP64

void dispatch64(int var, const StickPlanDev& P, const FftItem* it, int n, const z64* si, z64* so, const double* v,
                long ldv, double scale, hipStream_t s) {
  switch (P.n1) {
#define WH_REG_CASE(NN) \
  case NN:              \
    return launch_n<double, NN, double>(var, P, it, n, si, so, v, ldv, scale, s);
    WH_REG64_SIZES(WH_REG_CASE)
#undef WH_REG_CASE
    default: WH_REQUIRE(false, "plane_reg_vloc: no FP64 kernel for this grid");
  }
}

}  // namespace

void plane_reg_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z64* sin, z64* sout,
                    const double* vpot, long ldpot, double scale, hipStream_t s) {
  WH_REQUIRE(v >= 0 && v < plane_reg_variants(P, false), "plane_reg_vloc: no such variant");
  WH_REQUIRE(static_cast<const void*>(sin) != static_cast<const void*>(sout), "plane_reg_vloc: input and output alias");
  dispatch64(v, P, items, n, sin, sout, vpot, ldpot, scale, s);
}

#else

void dispatch32(int var, const StickPlanDev& P, const FftItem* it, int n, const z32* si, z32* so, const float* v,
                long ldv, double scale, hipStream_t s) {
  switch (P.n1) {
#define WH_REG_CASE(NN) \
  case NN:              \
    return launch_n<float, NN, float>(var, P, it, n, si, so, v, ldv, scale, s);
    WH_REG_SIZES(WH_REG_CASE)
#undef WH_REG_CASE
    default: WH_REQUIRE(false, "plane_reg_vloc: no kernel for this grid");
  }
}

void dispatch_zinv(const StickPlanDev& P, const FftItem* it, int n, const z32* a, long lda, const z32* b, long ldb,
                   z32* st, int mode, hipStream_t s) {
  switch (P.n3) {
#define WH_REG_CASE(NN) \
  case NN:              \
    return zinv_n<NN>(P, it, n, a, lda, b, ldb, st, mode, s);
    WH_REG_SIZES(WH_REG_CASE)
#undef WH_REG_CASE
    default: WH_REQUIRE(false, "zinv_reg: no kernel for this grid");
  }
}

void dispatch_zfwd(const StickPlanDev& P, const FftItem* it, int n, const z32* st, const z32* in, long ldin, z32* out,
                   long ldout, const ZfwdArgs& A, hipStream_t s) {
  switch (P.n3) {
#define WH_REG_CASE(NN) \
  case NN:              \
    return zfwd_n<NN>(P, it, n, st, in, ldin, out, ldout, A, s);
    WH_REG_SIZES(WH_REG_CASE)
#undef WH_REG_CASE
    default: WH_REQUIRE(false, "zfwd_reg: no kernel for this grid");
  }
}

}  // namespace

bool zreg_available(const StickPlanDev& P) {
  return P.n1 == P.n2 && P.n1 == P.n3 && compiled(P.n3, true) && P.kext <= (P.n3 + 3) / 4 &&
         2 * sizeof(z32) * static_cast<size_t>(P.zblk_pts) <= 80u * 1024u;
}

void zinv_reg(const StickPlanDev& P, const FftItem* items, int n, const z32* psiA, long ldA, const z32* psiB,
              long ldB, z32* sticks, int mode, hipStream_t s) {
  WH_REQUIRE(zreg_available(P), "zinv_reg: not available for this grid");
  if (n <= 0) return;
  dispatch_zinv(P, items, n, psiA, ldA, psiB, ldB, sticks, mode, s);
}

void zfwd_reg(const StickPlanDev& P, const FftItem* items, int n, const z32* sticks, const z32* in, long ldin,
              z32* out, long ldout, const ZfwdArgs& A, hipStream_t s) {
  WH_REQUIRE(zreg_available(P), "zfwd_reg: not available for this grid");
  WH_REQUIRE(A.mode == ZFWD_PAIR_STORE || A.mode == ZFWD_PAIR_STERN, "zfwd_reg: pair modes only");
  if (n <= 0) return;
  dispatch_zfwd(P, items, n, sticks, in, ldin, out, ldout, A, s);
}

int plane_reg_variants(const StickPlanDev& P, bool fp32) {
  if (P.n1 != P.n2 || P.n1 != P.n3 || !compiled(P.n1, fp32)) return 0;
  // the a1 and a2 transforms take |i| and |j| beyond N/4 as known zeros
  if (P.imax > (P.n1 + 3) / 4 || P.jext > (P.n2 + 3) / 4) return 0;
  const size_t elem = fp32 ? sizeof(z32) : sizeof(z64);
  int nv = 0;
  for (int v = 0; v < kNumPz; ++v)
    if (reg_smem(P, kPz[v], elem) <= 160u * 1024u) nv = v + 1;
  return nv;
}

void plane_reg_vloc(int v, const StickPlanDev& P, const FftItem* items, int n, const z32* sin, z32* sout,
                    const float* vpot, long ldpot, double scale, hipStream_t s) {
  WH_REQUIRE(v >= 0 && v < plane_reg_variants(P, true), "plane_reg_vloc: no such variant");
  WH_REQUIRE(static_cast<const void*>(sin) != static_cast<const void*>(sout), "plane_reg_vloc: input and output alias");
  dispatch32(v, P, items, n, sin, sout, vpot, ldpot, scale, s);
}

#endif  // WH_REG_FP64

}  // namespace wh
