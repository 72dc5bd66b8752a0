// FP64 GEMMs on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
//   C (M x N) = alpha * op(A) * B + beta * C,  all column major,
//   op(A) = A^T when TA (A stored K x M), A otherwise (A stored M x K).
//
// These carry the Gamma-point projections of the reference, which are DGEMMs
// on the real view of the complex arrays (Hamiltonian_kernel/glbrak.f90:38-39,
// apply_alpha_pc_to_m_wfcs.f90:59, apply_alpha_pv_to_m_wfcs.f90:71) and the
// Davidson subspace products (Wstat/wstat_tools.f90 build_hr/update/refresh).
//
// Tile 128 x 128 x 16 per 256-thread workgroup: four wavefronts in a 2 x 2
// arrangement, each owning 64 x 64 = 4 x 4 MFMA 16x16 blocks, with the next
// K tile prefetched into registers while the current one is multiplied out
// of a double-buffered LDS tile.  The projection ("bra") products have a
// small M (occupied bands + projectors) and a long K (2 npw), so they split K
// over blockIdx.z; the partial products are summed in a fixed order by the
// epilogue kernel, which also applies the G = 0 correction of glbrak_gamma
// (its DGER) and the per-row scale, so results do not depend on scheduling.
//
// Operand layout of the 16x16x4 f64 instruction (checked on the device):
// lane l holds A[l%16][l/16] and B[l/16][l%16]; result register r of lane l
// holds D[l/16 + 4r][l%16].  The kernel passes the B block as the
// instruction's A operand and the A block as its B operand, so it forms
// D = (A B)^T: the 16 lanes of a quarter wave then hold 16 consecutive rows
// of C, i.e. consecutive addresses of the column-major result, and the
// epilogue's loads and stores of C are whole 128-byte runs (the ket products
// read and write a C of 2 npw rows and up to thousands of columns).
#include "wh_gemm.h"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>
#include <functional>

namespace wh {

namespace {

constexpr int BM = 128;
constexpr int BK = 16;
constexpr int NT = 256;
constexpr int PAD = 2;

typedef double v4d __attribute__((ext_vector_type(4)));

// Tile loads, BK * XT / 256 elements per thread, mapped so that each load
// instruction of a wavefront reads contiguous memory:
//  * x-contiguous operand (element (k, x) at P[k*ld + x]: A of the plain
//    product, XT = 128): x = tid % 128, k = tid / 128 + 2q  (512 bytes per
//    instruction);
//  * k-contiguous operand (element (k, x) at P[x*ld + k]: A of the bra
//    product, B of both): k = tid % 16, x = tid / 16 + 16q  (four 128-byte runs).
// CHK: zero-fill outside [0, X) x [k0, kend) (edge tiles only).
template <bool KCONT, bool CHK, int XT>
__device__ __forceinline__ void load_tile(double (&r)[BK * XT / NT], const double* __restrict__ P, long ld, int X,
                                          int x0, int k0, int kend) {
  const int tid = threadIdx.x;
  if (KCONT) {
    const int gk = k0 + (tid & 15);
#pragma unroll
    for (int q = 0; q < BK * XT / NT; ++q) {
      const int gx = x0 + (tid >> 4) + 16 * q;
      if (CHK) r[q] = (gx < X && gk < kend) ? P[static_cast<long>(gx) * ld + gk] : 0.0;
      else r[q] = P[static_cast<long>(gx) * ld + gk];
    }
  } else {
    static_assert(KCONT || XT == 128, "x-contiguous tiles are 128 wide");
    const int gx = x0 + (tid & 127);
#pragma unroll
    for (int q = 0; q < BK * XT / NT; ++q) {
      const int gk = k0 + (tid >> 7) + 2 * q;
      if (CHK) r[q] = (gx < X && gk < kend) ? P[static_cast<long>(gk) * ld + gx] : 0.0;
      else r[q] = P[static_cast<long>(gk) * ld + gx];
    }
  }
}

// the LDS tile is [k][x] for both operands
template <bool KCONT, int XT>
__device__ __forceinline__ void store_tile(double (*S)[XT + PAD], const double (&r)[BK * XT / NT]) {
  const int tid = threadIdx.x;
  if (KCONT) {
#pragma unroll
    for (int q = 0; q < BK * XT / NT; ++q) S[tid & 15][(tid >> 4) + 16 * q] = r[q];
  } else {
#pragma unroll
    for (int q = 0; q < BK * XT / NT; ++q) S[(tid >> 7) + 2 * q][tid & 127] = r[q];
  }
}

// Tiles (bx, by, kz): gx x gy output tiles times gz K chunks, k range
// [kz*kchunk, min(K, (kz+1)*kchunk)); gz > 1: raw partial sums go to
// C + kz*pstride (alpha, beta ignored).  The grid is 1D and its blocks are
// dispatched round-robin over the 8 XCDs (block L on XCD L % 8); tile
// W = (L % 8) * ceil(tiles / 8) + L / 8 puts consecutive tiles, which share
// their B panel (same by, kz) or A panel (same bx, kz), on one XCD and its L2.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_blocks(int tiles) { return (tiles + kXcds - 1) / kXcds * kXcds; }

// WN: 16-column MFMA blocks per wavefront along N, i.e. tiles of
// BM x (32 WN) (WN = 4: 128 x 128; narrower tiles give more workgroups to
// products with few columns).
template <bool TA, int WN>
__global__ __launch_bounds__(NT, 2) void k_dgemm(int M, int N, int K, int kchunk, double alpha,
                                              const double* __restrict__ A, long lda, const double* __restrict__ B,
                                              long ldb, double beta, double* __restrict__ C, long ldc, long pstride,
                                              int gx, int gy, int gz, const int* __restrict__ d_n) {
  constexpr int TN = 32 * WN;
  if (d_n) {
    // columns the device left active: the tiles of the active columns are
    // renumbered over the whole grid, so they still spread over all XCDs
    N = min(N, *d_n);
    gy = min(gy, (N + TN - 1) / TN);
  }
  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][TN + PAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  const int tiles = gx * gy * gz;
  // (with fewer active tiles than the grid was sized for, only the first
  // xcd_blocks(tiles) workgroups take part, so each tile has exactly one)
  if (static_cast<int>(blockIdx.x) >= xcd_blocks(tiles)) return;
  const int W = static_cast<int>(blockIdx.x % kXcds) * (xcd_blocks(tiles) / kXcds) + static_cast<int>(blockIdx.x / kXcds);
  if (W >= tiles) return;   // uniform per workgroup
  const int m0 = (W % gx) * BM;
  const int n0 = ((W / gx) % gy) * TN;
  if (n0 >= N) return;      // uniform per workgroup
  const int kz = W / (gx * gy);
  const int kbeg = kz * kchunk;
  const int kend = min(K, kbeg + kchunk);
  // interior tiles load without bounds checks (all but a short last K step)
  const bool interior = m0 + BM <= M && n0 + TN <= N;
  const bool partial = gz > 1;
  // C = alpha AB + beta C on an interior tile: the C loads of a row block are
  // issued while the previous block is combined and stored
  const bool rmw = !partial && beta != 0.0 && interior;
  double* Cz = C + (partial ? static_cast<long>(kz) * pstride : 0L);
  const int row0 = m0 + wm * 64 + (lane & 15);
  const int col0 = n0 + wn * (16 * WN) + (lane >> 4);
  double cv[4 * WN];
  auto load_c = [&](int i) {
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) cv[4 * j + r] = Cz[static_cast<long>(col0 + j * 16 + 4 * r) * ldc + row0 + i * 16];
  };
  v4d acc[4][WN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  const int ntile = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  double ra[BK * BM / NT], rb[BK * TN / NT];
  auto load = [&](int k0) {
    if (interior && k0 + BK <= kend) {
      load_tile<TA, false, BM>(ra, A, lda, M, m0, k0, kend);
      load_tile<true, false, TN>(rb, B, ldb, N, n0, k0, kend);
    } else {
      load_tile<TA, true, BM>(ra, A, lda, M, m0, k0, kend);
      load_tile<true, true, TN>(rb, B, ldb, N, n0, k0, kend);
    }
  };
  if (ntile > 0) {
    load(kbeg);
    store_tile<TA, BM>(As[0], ra);
    store_tile<true, TN>(Bs[0], rb);
  }
  __syncthreads();
  for (int t = 0; t < ntile; ++t) {
    const int cur = t & 1;
    if (t + 1 < ntile) load(kbeg + (t + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a[4], b[WN];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[cur][kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < WN; ++j) b[j] = Bs[cur][kr][wn * (16 * WN) + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    if (t + 1 < ntile) {
      store_tile<TA, BM>(As[cur ^ 1], ra);
      store_tile<true, TN>(Bs[cur ^ 1], rb);
    }
    __syncthreads();
  }
  // block (i, j) holds the transposed product: register r of lane l is
  // C(row i*16 + l%16, column j*16 + l/16 + 4r), so lanes 0..15 (and each
  // further quarter wave) touch 16 consecutive rows of one column of C
  if (rmw) {
    // row block i is combined in its accumulators, the next block's C loads
    // are issued, then block i is stored
    load_c(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] = alpha * acc[i][j][r] + beta * cv[4 * j + r];
      if (i + 1 < 4) load_c(i + 1);
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) Cz[static_cast<long>(col0 + j * 16 + 4 * r) * ldc + row0 + i * 16] = acc[i][j][r];
    }
  } else if (partial || beta == 0.0) {
    const double f = partial ? 1.0 : alpha;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = row0 + i * 16, col = col0 + j * 16 + 4 * r;
          if (interior || (row < M && col < N)) Cz[static_cast<long>(col) * ldc + row] = f * acc[i][j][r];
        }
  } else {
    // edge tile of C = alpha AB + beta C
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = row0 + i * 16;
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = col0 + j * 16 + 4 * r;
          cv[4 * j + r] = (row < M && col < N) ? Cz[static_cast<long>(col) * ldc + row] : 0.0;
        }
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = col0 + j * 16 + 4 * r;
          if (row < M && col < N) Cz[static_cast<long>(col) * ldc + row] = alpha * acc[i][j][r] + beta * cv[4 * j + r];
        }
    }
  }
}

// Bra product with few rows (M <= MB, e.g. the bands of a small molecule with
// no projectors): one workgroup per (column j, K chunk z) streams B's column
// once and keeps the M running sums per thread (A's rows come through the
// caches); partial sums P_z(i, j) at P + z * pstride + j * ldc + i.  With
// FIN (one K chunk) the workgroup writes the result itself, as k_reduce_g0
// would: P(i, j) = s_i * (alpha * sum - g0 * A[i*lda] * B[j*ldb]).
constexpr int NTS = 256;
struct BraFin {
  double alpha, g0;
  const double* s;
};
template <int MB, bool FIN>
__global__ __launch_bounds__(NTS) void k_bra_skinny(int M, int K, int kchunk, const double* __restrict__ A, long lda,
                                                   const double* __restrict__ B, long ldb, double* __restrict__ P,
                                                   long ldc, long pstride, const int* __restrict__ d_n, BraFin fin) {
  __shared__ double red[MB][NTS / 64];
  const int j = blockIdx.x;
  if (d_n && j >= *d_n) return;
  const int z = blockIdx.y;
  const int kb = z * kchunk;
  const int ke = min(K, kb + kchunk);
  const double* Bj = B + static_cast<long>(j) * ldb;
  double acc[MB];
#pragma unroll
  for (int i = 0; i < MB; ++i) acc[i] = 0.0;
  for (int k = kb + threadIdx.x; k < ke; k += NTS) {
    const double b = Bj[k];
#pragma unroll
    for (int i = 0; i < MB; ++i)
      if (i < M) acc[i] += A[static_cast<long>(i) * lda + k] * b;
  }
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    if (i < M) {
      double v = acc[i];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[i][w] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < M) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NTS / 64; ++q) v += red[threadIdx.x][q];
    if (FIN) {
      const double acc = 0.0 + v;
      v = fin.alpha * acc - fin.g0 * A[static_cast<long>(threadIdx.x) * lda] * Bj[0];
      if (fin.s) v *= fin.s[threadIdx.x];
    }
    P[static_cast<long>(z) * pstride + static_cast<long>(j) * ldc + threadIdx.x] = v;
  }
}

// C = alpha A B + beta C (column major, A: M x K, B: K x N) for small K
// (the ket product of a molecule with a few bands and no projectors): one
// thread per element of C, B's column in LDS.
template <int KB>
__global__ __launch_bounds__(NTS) void k_ket_skinny(int M, int K, double alpha, const double* __restrict__ A, long lda,
                                                   const double* __restrict__ B, long ldb, double beta,
                                                   double* __restrict__ C, long ldc, const int* __restrict__ d_n) {
  __shared__ double bj[KB];
  const int j = blockIdx.y;
  if (d_n && j >= *d_n) return;
  if (threadIdx.x < K) bj[threadIdx.x] = B[static_cast<long>(j) * ldb + threadIdx.x];
  __syncthreads();
  const int r = blockIdx.x * NTS + threadIdx.x;
  if (r >= M) return;
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < KB; ++i)
    if (i < K) acc += A[static_cast<long>(i) * lda + r] * bj[i];
  double* c = C + static_cast<long>(j) * ldc + r;
  *c = (beta == 0.0) ? alpha * acc : alpha * acc + beta * (*c);
}

// C(i, j) = s_i * (alpha * sum_z P_z(i, j) - g0 * a(1, i) * b(1, j))
//   P_z: split partials (ldc, stride pstride); a(1, i) = A[i*lda], b(1, j) = B[j*ldb]
__global__ void k_reduce_g0(int M, int N, int nsplit, const double* P, long pstride, double alpha, double g0,
                            const double* A, long lda, const double* B, long ldb, const double* s, double* C,
                            long ldc, const int* d_n) {
  if (d_n) N = min(N, *d_n);
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(M) * N) return;
  const int i = static_cast<int>(idx % M);
  const int j = static_cast<int>(idx / M);
  const long o = static_cast<long>(j) * ldc + i;
  double acc = 0.0;
  for (int z = 0; z < nsplit; ++z) acc += P[static_cast<long>(z) * pstride + o];
  double v = alpha * acc - g0 * A[static_cast<long>(i) * lda] * B[static_cast<long>(j) * ldb];
  if (s) v *= s[i];
  C[o] = v;
}

// C(i, j) = scale_i * (C(i, j) - g0 * A[i*lda] * B[j*ldb])   (scale may be null)
__global__ void k_g0_scale(int M, int N, double g0, const double* A, long lda, const double* B, long ldb,
                           const double* s, double* C, long ldc) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(M) * N) return;
  const int i = static_cast<int>(idx % M);
  const int j = static_cast<int>(idx / M);
  double v = C[static_cast<long>(j) * ldc + i] - g0 * A[static_cast<long>(i) * lda] * B[static_cast<long>(j) * ldb];
  if (s) v *= s[i];
  C[static_cast<long>(j) * ldc + i] = v;
}

// Grows on demand; before a larger buffer replaces the old one, the streams
// that used it (this host thread's calls only) are drained, so no other
// system's stream is waited for.
struct Scratch {
  double* p = nullptr;
  size_t n = 0;
  std::vector<hipStream_t> users;
  double* get(size_t count, hipStream_t s) {
    if (count > n) {
      if (p) {
        for (hipStream_t u : users) WH_HIP(hipStreamSynchronize(u));
        WH_HIP(hipFree(p));
      }
      users.clear();
      p = nullptr;
      WH_HIP(hipMalloc(&p, sizeof(double) * count));
      n = count;
    }
    if (std::find(users.begin(), users.end(), s) == users.end()) users.push_back(s);
    return p;
  }
};

// split-K partial sums, per host thread
Scratch& partials() {
  static thread_local Scratch s;
  return s;
}

}  // namespace

template <bool TA, int WN>
static void launch_tiles(int M, int N, int K, int kchunk, int gz, double alpha, const double* A, long lda,
                         const double* B, long ldb, double beta, double* C, long ldc, long pstride, hipStream_t s,
                         const int* d_n) {
  const int gx = ceil_div(M, BM), gy = ceil_div(N, 32 * WN);
  k_dgemm<TA, WN><<<xcd_blocks(gx * gy * gz), NT, 0, s>>>(M, N, K, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                                          pstride, gx, gy, gz, d_n);
  WH_HIP(hipGetLastError());
}

// tile width variants: WN = 4, 2, 1 (128-, 64- and 32-column tiles)
template <bool TA>
static void launch_wn(int v, int M, int N, int K, int kchunk, int gz, double alpha, const double* A, long lda,
                      const double* B, long ldb, double beta, double* C, long ldc, long pstride, hipStream_t s,
                      const int* d_n = nullptr) {
  if (v == 0) launch_tiles<TA, 4>(M, N, K, kchunk, gz, alpha, A, lda, B, ldb, beta, C, ldc, pstride, s, d_n);
  else if (v == 1) launch_tiles<TA, 2>(M, N, K, kchunk, gz, alpha, A, lda, B, ldb, beta, C, ldc, pstride, s, d_n);
  else launch_tiles<TA, 1>(M, N, K, kchunk, gz, alpha, A, lda, B, ldb, beta, C, ldc, pstride, s, d_n);
}
constexpr int kWnVariants = 3;
inline int wn_of(int v) { return 4 >> v; }

// The variant of a product shape is chosen once, by timing every variant on
// the first call's own arguments (two rounds; the first warms the code up);
// results never depend on the choice beyond rounding of a fixed order.
using VariantKey = std::tuple<int, int, int, int, int>;
static std::map<VariantKey, int>& chosen_variants() {
  static std::map<VariantKey, int> m;
  return m;
}
static std::mutex g_variant_mutex;

static int time_variants(int nv, const std::function<void(int)>& run, hipStream_t s, float* t) {
  hipEvent_t a, b;
  WH_HIP(hipEventCreate(&a));
  WH_HIP(hipEventCreate(&b));
  for (int rep = 0; rep < 2; ++rep)
    for (int v = 0; v < nv; ++v) {
      WH_HIP(hipEventRecord(a, s));
      run(v);
      WH_HIP(hipEventRecord(b, s));
      WH_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      WH_HIP(hipEventElapsedTime(&ms, a, b));
      if (rep == 1) t[v] = ms;
    }
  WH_HIP(hipEventDestroy(a));
  WH_HIP(hipEventDestroy(b));
  int best = 0;
  for (int v = 1; v < nv; ++v)
    if (t[v] < t[best]) best = v;
  return best;
}

static int log2_bucket(int n) {
  int lg = 0;
  while ((2 << lg) <= n) ++lg;
  return lg;
}

void dgemm(bool transa, int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
           double beta, double* C, long ldc, hipStream_t s, const int* d_n) {
  if (M <= 0 || N <= 0) return;
  if (!transa && K > 0 && K <= 32 && N <= 256) {
    // inner dimension of a few bands: a streaming kernel (the matrix-core tiles would be nearly empty)
    dim3 grid(ceil_div(M, NTS), N);
    if (K <= 8) k_ket_skinny<8><<<grid, NTS, 0, s>>>(M, K, alpha, A, lda, B, ldb, beta, C, ldc, d_n);
    else if (K <= 16) k_ket_skinny<16><<<grid, NTS, 0, s>>>(M, K, alpha, A, lda, B, ldb, beta, C, ldc, d_n);
    else k_ket_skinny<32><<<grid, NTS, 0, s>>>(M, K, alpha, A, lda, B, ldb, beta, C, ldc, d_n);
    WH_HIP(hipGetLastError());
    return;
  }
  const int K0 = std::max(K, 0), kc = std::max(K, 1);
  auto run = [&](int v, double* Cv) {
    if (transa) launch_wn<true>(v, M, N, K0, kc, 1, alpha, A, lda, B, ldb, beta, Cv, ldc, 0, s, d_n);
    else launch_wn<false>(v, M, N, K0, kc, 1, alpha, A, lda, B, ldb, beta, Cv, ldc, 0, s, d_n);
  };
  // tile width, timed once per (op, M, log2 K, log2 N)
  const VariantKey key{0, transa ? 1 : 0, M, log2_bucket(std::max(K, 1)), log2_bucket(N)};
  std::unique_lock<std::mutex> lk(g_variant_mutex);
  auto it = chosen_variants().find(key);
  if (it == chosen_variants().end()) {
    // timed on a scratch copy of C's shape: C += AB must run once only
    double* tmp = nullptr;
    WH_HIP(hipMalloc(&tmp, sizeof(double) * static_cast<size_t>(ldc) * N));
    WH_HIP(hipMemsetAsync(tmp, 0, sizeof(double) * static_cast<size_t>(ldc) * N, s));
    float t[kWnVariants] = {};
    const int best = time_variants(kWnVariants, [&](int v) { run(v, tmp); }, s, t);
    WH_HIP(hipStreamSynchronize(s));
    WH_HIP(hipFree(tmp));
    if (tune_log())
      std::fprintf(stderr, "[west_hip] dgemm%s M=%d N=%d K=%d: tiles 128x128 %.3f 128x64 %.3f 128x32 %.3f ms -> %d\n",
                   transa ? "_t" : "", M, N, K, t[0], t[1], t[2], best);
    it = chosen_variants().emplace(key, best).first;
  }
  const int pick = it->second;
  lk.unlock();
  run(pick, C);
}

// v = 3 * d + w: K split aimed at 512 << d workgroups (512: two per compute
// unit), tile width variant w.  Returns the split depth and K chunk.
static void bra_split(int M, int N, int K, int v, int& split, int& kchunk) {
  const int w = v % kWnVariants;
  const int wg_target = 512 << (v / kWnVariants);
  const int tiles = ceil_div(M, BM) * ceil_div(N, 32 * wn_of(w));
  // enough workgroups for the compute units (few output tiles, as in the
  // narrow products of a partial batch, need a deep split), each K chunk at
  // least 16 steps long
  split = std::max(1, std::min(64, ceil_div(wg_target, tiles)));
  split = std::max(1, std::min(split, K / (16 * BK)));
  kchunk = ceil_div(ceil_div(std::max(K, 1), split), BK) * BK;
  split = ceil_div(std::max(K, 1), kchunk);
}

// P: room for the split-K partial sums (null: the per-thread scratch)
static void dgemm_bra_own(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                          double g0, const double* scale, double* C, long ldc, hipStream_t s, int v = 0,
                          double* P = nullptr, const int* d_n = nullptr) {
  int split = 1, kchunk = 1;
  bra_split(M, N, K, v, split, kchunk);
  const long pstride = static_cast<long>(ldc) * N;
  if (!P) P = partials().get(static_cast<size_t>(split) * pstride, s);
  launch_wn<true>(v % kWnVariants, M, N, K, kchunk, split, 1.0, A, lda, B, ldb, 0.0, P, ldc, pstride, s, d_n);
  const long n = static_cast<long>(M) * N;
  k_reduce_g0<<<ceil_div(n, 256), 256, 0, s>>>(M, N, split, P, pstride, alpha, g0, A, lda, B, ldb, scale, C, ldc,
                                               d_n);
  WH_HIP(hipGetLastError());
}

static void skinny_split(int N, int K, int& split, int& kchunk) {
  // about 2048 workgroups, K chunks of at least 1024
  split = std::max(1, std::min(ceil_div(2048, N), K / 1024));
  kchunk = ceil_div(std::max(K, 1), split);
  split = ceil_div(std::max(K, 1), kchunk);
}

static void dgemm_bra_skinny(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                             double g0, const double* scale, double* C, long ldc, hipStream_t s, double* P = nullptr,
                             const int* d_n = nullptr) {
  int split = 1, kchunk = 1;
  skinny_split(N, K, split, kchunk);
  const long pstride = static_cast<long>(ldc) * N;
  const BraFin fin{alpha, g0, scale};
  dim3 grid(N, split);
  if (split == 1) {   // one K chunk: the workgroups write C themselves
    if (M <= 8) k_bra_skinny<8, true><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, C, ldc, pstride, d_n, fin);
    else if (M <= 16)
      k_bra_skinny<16, true><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, C, ldc, pstride, d_n, fin);
    else k_bra_skinny<32, true><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, C, ldc, pstride, d_n, fin);
    WH_HIP(hipGetLastError());
    return;
  }
  if (!P) P = partials().get(static_cast<size_t>(split) * pstride, s);
  if (M <= 8) k_bra_skinny<8, false><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, P, ldc, pstride, d_n, fin);
  else if (M <= 16)
    k_bra_skinny<16, false><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, P, ldc, pstride, d_n, fin);
  else k_bra_skinny<32, false><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, P, ldc, pstride, d_n, fin);
  WH_HIP(hipGetLastError());
  const long n = static_cast<long>(M) * N;
  k_reduce_g0<<<ceil_div(n, 256), 256, 0, s>>>(M, N, split, P, pstride, alpha, g0, A, lda, B, ldb, scale, C, ldc,
                                               d_n);
  WH_HIP(hipGetLastError());
}

void dgemm_bra(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb, double g0,
               const double* scale, double* C, long ldc, hipStream_t s, const int* d_n) {
  if (M <= 0 || N <= 0) return;
  // a few rows and columns: the matrix-core tiles (128 x 128) would be almost
  // empty (with many columns the tiles' reuse of A wins again)
  if (M <= 32 && N <= 256 && K > 0)
    return dgemm_bra_skinny(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s, nullptr, d_n);
  if (K <= 0 || N < 32) return dgemm_bra_own(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s, 0, nullptr, d_n);
  // variants: K split depth (3) x tile width (3), timed once per (M, K, log2 N);
  // C is written (not accumulated), so the timing runs on the call's own output
  constexpr int nv = 3 * kWnVariants;
  auto run = [&](int v) { dgemm_bra_own(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s, v, nullptr, d_n); };
  const VariantKey key{1, 1, M, K, log2_bucket(N)};
  std::unique_lock<std::mutex> lk(g_variant_mutex);
  auto it = chosen_variants().find(key);
  if (it == chosen_variants().end()) {
    float t[nv] = {};
    const int best = time_variants(nv, run, s, t);
    if (tune_log()) {
      std::fprintf(stderr, "[west_hip] dgemm_bra M=%d N=%d K=%d:", M, N, K);
      for (int v = 0; v < nv; ++v) std::fprintf(stderr, " split%d/tile128x%d %.3f", 512 << (v / kWnVariants),
                                                32 * wn_of(v % kWnVariants), t[v]);
      std::fprintf(stderr, " ms -> %d\n", best);
    }
    it = chosen_variants().emplace(key, best).first;
  }
  const int pick = it->second;
  lk.unlock();
  // recomputed with the chosen variant, so the result does not depend on the timing order
  run(pick);
}

// ---- products for callers that neither allocate nor time -------------------
// The variant an earlier timed call chose for the shape, else a fixed one.
static int chosen_or(const VariantKey& key, int dflt) {
  std::lock_guard<std::mutex> lk(g_variant_mutex);
  auto it = chosen_variants().find(key);
  return it != chosen_variants().end() ? it->second : dflt;
}

static int bra_untimed_variant(int M, int N, int K) {
  // default: K split aimed at 2048 workgroups on 128 x 64 tiles (the most
  // frequent winner of the timing)
  return (K <= 0 || N < 32) ? 0 : chosen_or(VariantKey{1, 1, M, K, log2_bucket(N)}, 7);
}

long bra_partials_doubles(int M, int N, int K, long ldc, int* variant) {
  *variant = 0;
  if (M <= 0 || N <= 0) return 0;
  int split = 1, kchunk = 1;
  if (M <= 32 && N <= 256 && K > 0) {
    skinny_split(N, K, split, kchunk);
  } else {
    *variant = bra_untimed_variant(M, N, K);
    bra_split(M, N, K, *variant, split, kchunk);
  }
  return static_cast<long>(split) * ldc * N;
}

void dgemm_bra_into(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                    double g0, const double* scale, double* C, long ldc, double* P, int variant, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  if (M <= 32 && N <= 256 && K > 0) return dgemm_bra_skinny(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s, P);
  dgemm_bra_own(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s, variant, P);
}

void dgemm_untimed(bool transa, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                   long ldb, double beta, double* C, long ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  if (!transa && K > 0 && K <= 32 && N <= 256) return dgemm(transa, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
  // default tile width: the widest that still gives two workgroups per compute unit
  int dflt = kWnVariants - 1;
  for (int v = 0; v < kWnVariants; ++v)
    if (static_cast<long>(ceil_div(M, BM)) * ceil_div(N, 32 * wn_of(v)) >= 512) {
      dflt = v;
      break;
    }
  const int v = chosen_or(VariantKey{0, transa ? 1 : 0, M, log2_bucket(std::max(K, 1)), log2_bucket(N)}, dflt);
  const int K0 = std::max(K, 0), kc = std::max(K, 1);
  if (transa) launch_wn<true>(v, M, N, K0, kc, 1, alpha, A, lda, B, ldb, beta, C, ldc, 0, s);
  else launch_wn<false>(v, M, N, K0, kc, 1, alpha, A, lda, B, ldb, beta, C, ldc, 0, s);
}

void g0_fix_scale(int M, int N, double g0, const double* A, long lda, const double* B, long ldb,
                  const double* scale, double* C, long ldc, hipStream_t s) {
  const long n = static_cast<long>(M) * N;
  if (n <= 0) return;
  k_g0_scale<<<ceil_div(n, 256), 256, 0, s>>>(M, N, g0, A, lda, B, ldb, scale, C, ldc);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
