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
// holds D[l/16 + 4r][l%16].
#include "wh_gemm.h"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>

namespace wh {

// Backend of the FP64 products. Plain products (dgemm) use rocBLAS, measured
// about twice as fast as the kernel below for the ket shape (M = 2 npw,
// K = nq) on MI355X. The bra products (small M = nq, long K = 2 npw, fused
// G = 0 correction and row scale) use whichever of the split-K kernel of this
// file (three K-split depths) and rocBLAS + k_g0_scale was fastest when first
// timed for that (M, K) and the power-of-two bucket of N, for N >= 512 columns
// (at 8192 columns the kernel won at nq = 384, rocBLAS at nq = 1296); smaller
// calls use the kernel.
// WH_GEMM=own or WH_GEMM=rocblas puts every product on one backend.
enum class Backend { Default, Own, Rocblas };
static Backend backend() {
  static const Backend v = [] {
    const char* e = std::getenv("WH_GEMM");
    if (e && std::strcmp(e, "rocblas") == 0) return Backend::Rocblas;
    if (e && std::strcmp(e, "own") == 0) return Backend::Own;
    return Backend::Default;
  }();
  return v;
}

// one rocBLAS handle per host thread (handles are not shared between threads)
struct BlasHandle {
  rocblas_handle h = nullptr;
  ~BlasHandle() {
    if (h) rocblas_destroy_handle(h);
  }
};
static rocblas_handle blas(hipStream_t s) {
  static thread_local BlasHandle bh;
  if (!bh.h) WH_REQUIRE(rocblas_create_handle(&bh.h) == rocblas_status_success, "rocblas_create_handle failed");
  WH_REQUIRE(rocblas_set_stream(bh.h, s) == rocblas_status_success, "rocblas_set_stream failed");
  return bh.h;
}

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int NT = 256;
constexpr int PAD = 2;

typedef double v4d __attribute__((ext_vector_type(4)));

template <bool TA>
__device__ __forceinline__ void load_a(double (&ra)[8], const double* __restrict__ A, long lda, int M, int m0, int k0,
                                       int kend) {
  const int tid = threadIdx.x;
  if (TA) {   // element (k, m) at A[m*lda + k]: 128 m x 16 k, 8 consecutive k per thread
    const int m = tid >> 1;
    const int kq = (tid & 1) * 8;
    const int gm = m0 + m;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int gk = k0 + kq + q;
      ra[q] = (gm < M && gk < kend) ? A[static_cast<long>(gm) * lda + gk] : 0.0;
    }
  } else {    // element (m, k) at A[k*lda + m]: 16 k x 128 m, 8 consecutive m per thread
    const int k = tid >> 4;
    const int mq = (tid & 15) * 8;
    const int gk = k0 + k;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int gm = m0 + mq + q;
      ra[q] = (gm < M && gk < kend) ? A[static_cast<long>(gk) * lda + gm] : 0.0;
    }
  }
}

template <bool TA>
__device__ __forceinline__ void store_a(double (*As)[BM + PAD], const double (&ra)[8]) {
  const int tid = threadIdx.x;
  if (TA) {
    const int m = tid >> 1;
    const int kq = (tid & 1) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) As[kq + q][m] = ra[q];
  } else {
    const int k = tid >> 4;
    const int mq = (tid & 15) * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) As[k][mq + q] = ra[q];
  }
}

__device__ __forceinline__ void load_b(double (&rb)[8], const double* __restrict__ B, long ldb, int N, int n0, int k0,
                                       int kend) {
  const int tid = threadIdx.x;
  const int n = tid >> 1;
  const int kq = (tid & 1) * 8;
  const int gn = n0 + n;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int gk = k0 + kq + q;
    rb[q] = (gn < N && gk < kend) ? B[static_cast<long>(gn) * ldb + gk] : 0.0;
  }
}

__device__ __forceinline__ void store_b(double (*Bs)[BN + PAD], const double (&rb)[8]) {
  const int tid = threadIdx.x;
  const int n = tid >> 1;
  const int kq = (tid & 1) * 8;
#pragma unroll
  for (int q = 0; q < 8; ++q) Bs[kq + q][n] = rb[q];
}

// Tiles (bx, by, kz): gx x gy output tiles times gz K chunks, k range
// [kz*kchunk, min(K, (kz+1)*kchunk)); gz > 1: raw partial sums go to
// C + kz*pstride (alpha, beta ignored).  The grid is 1D and its blocks are
// dispatched round-robin over the 8 XCDs (block L on XCD L % 8); tile
// W = (L % 8) * ceil(tiles / 8) + L / 8 puts consecutive tiles, which share
// their B panel (same by, kz) or A panel (same bx, kz), on one XCD and its L2.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_blocks(int tiles) { return (tiles + kXcds - 1) / kXcds * kXcds; }

template <bool TA>
__global__ __launch_bounds__(NT, 2) void k_dgemm(int M, int N, int K, int kchunk, double alpha,
                                              const double* __restrict__ A, long lda, const double* __restrict__ B,
                                              long ldb, double beta, double* __restrict__ C, long ldc, long pstride,
                                              int gx, int gy, int gz) {
  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][BN + PAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  const int tiles = gx * gy * gz;
  const int W = static_cast<int>(blockIdx.x % kXcds) * (xcd_blocks(tiles) / kXcds) + static_cast<int>(blockIdx.x / kXcds);
  if (W >= tiles) return;   // uniform per workgroup
  const int m0 = (W % gx) * BM;
  const int n0 = ((W / gx) % gy) * BN;
  const int kz = W / (gx * gy);
  const int kbeg = kz * kchunk;
  const int kend = min(K, kbeg + kchunk);
  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  const int ntile = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  double ra[8], rb[8];
  if (ntile > 0) {
    load_a<TA>(ra, A, lda, M, m0, kbeg, kend);
    load_b(rb, B, ldb, N, n0, kbeg, kend);
    store_a<TA>(As[0], ra);
    store_b(Bs[0], rb);
  }
  __syncthreads();
  for (int t = 0; t < ntile; ++t) {
    const int cur = t & 1;
    if (t + 1 < ntile) {
      load_a<TA>(ra, A, lda, M, m0, kbeg + (t + 1) * BK, kend);
      load_b(rb, B, ldb, N, n0, kbeg + (t + 1) * BK, kend);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[cur][kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[cur][kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (t + 1 < ntile) {
      store_a<TA>(As[cur ^ 1], ra);
      store_b(Bs[cur ^ 1], rb);
    }
    __syncthreads();
  }
  const bool partial = gz > 1;
  double* Cz = C + (partial ? static_cast<long>(kz) * pstride : 0L);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + wn * 64 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double* c = Cz + static_cast<long>(col) * ldc + row;
          if (partial) {
            *c = acc[i][j][r];
          } else {
            const double v = alpha * acc[i][j][r];
            *c = (beta == 0.0) ? v : v + beta * (*c);
          }
        }
      }
    }
}

// Bra product with few rows (M <= MB, e.g. the bands of a small molecule with
// no projectors): one workgroup per (column j, K chunk z) streams B's column
// once and keeps the M running sums per thread (A's rows come through the
// caches); partial sums P_z(i, j) at P + z * pstride + j * ldc + i.
constexpr int NTS = 256;
template <int MB>
__global__ __launch_bounds__(NTS) void k_bra_skinny(int M, int K, int kchunk, const double* __restrict__ A, long lda,
                                                   const double* __restrict__ B, long ldb, double* __restrict__ P,
                                                   long ldc, long pstride) {
  __shared__ double red[MB][NTS / 64];
  const int j = blockIdx.x;
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
    P[static_cast<long>(z) * pstride + static_cast<long>(j) * ldc + threadIdx.x] = v;
  }
}

// C = alpha A B + beta C (column major, A: M x K, B: K x N) for small K
// (the ket product of a molecule with a few bands and no projectors): one
// thread per element of C, B's column in LDS.
template <int KB>
__global__ __launch_bounds__(NTS) void k_ket_skinny(int M, int K, double alpha, const double* __restrict__ A, long lda,
                                                   const double* __restrict__ B, long ldb, double beta,
                                                   double* __restrict__ C, long ldc) {
  __shared__ double bj[KB];
  const int j = blockIdx.y;
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
                            long ldc) {
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

struct Scratch {
  double* p = nullptr;
  size_t n = 0;
  double* get(size_t count) {
    if (count > n) {
      if (p) {
        WH_HIP(hipDeviceSynchronize());
        WH_HIP(hipFree(p));
      }
      p = nullptr;
      WH_HIP(hipMalloc(&p, sizeof(double) * count));
      n = count;
    }
    return p;
  }
};

// split-K partial sums, per host thread
Scratch& partials() {
  static thread_local Scratch s;
  return s;
}

}  // namespace

void dgemm(bool transa, int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
           double beta, double* C, long ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  if (!transa && K > 0 && K <= 32 && N <= 256 && backend() != Backend::Rocblas) {
    // inner dimension of a few bands: a streaming kernel (the matrix-core tiles would be nearly empty)
    dim3 grid(ceil_div(M, NTS), N);
    if (K <= 8) k_ket_skinny<8><<<grid, NTS, 0, s>>>(M, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else if (K <= 16) k_ket_skinny<16><<<grid, NTS, 0, s>>>(M, K, alpha, A, lda, B, ldb, beta, C, ldc);
    else k_ket_skinny<32><<<grid, NTS, 0, s>>>(M, K, alpha, A, lda, B, ldb, beta, C, ldc);
    WH_HIP(hipGetLastError());
    return;
  }
  if (backend() != Backend::Own && K > 0) {
    WH_REQUIRE(rocblas_dgemm(blas(s), transa ? rocblas_operation_transpose : rocblas_operation_none,
                             rocblas_operation_none, M, N, K, &alpha, A, lda, B, ldb, &beta, C, ldc) ==
                   rocblas_status_success,
               "rocblas_dgemm failed");
    return;
  }
  const int gx = ceil_div(M, BM), gy = ceil_div(N, BN);
  const int nb = xcd_blocks(gx * gy);
  const int kc = std::max(K, 1);
  if (transa)
    k_dgemm<true><<<nb, NT, 0, s>>>(M, N, std::max(K, 0), kc, alpha, A, lda, B, ldb, beta, C, ldc, 0, gx, gy, 1);
  else
    k_dgemm<false><<<nb, NT, 0, s>>>(M, N, std::max(K, 0), kc, alpha, A, lda, B, ldb, beta, C, ldc, 0, gx, gy, 1);
  WH_HIP(hipGetLastError());
}

// wg_target: workgroups the K split aims at (512: two per compute unit)
static void dgemm_bra_own(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                          double g0, const double* scale, double* C, long ldc, hipStream_t s, int wg_target = 512) {
  const int tiles = ceil_div(M, BM) * ceil_div(N, BN);
  // enough workgroups for the compute units (few output tiles, as in the
  // narrow products of a partial batch, need a deep split), each K chunk at
  // least 16 steps long
  int split = std::max(1, std::min(64, ceil_div(wg_target, tiles)));
  split = std::max(1, std::min(split, K / (16 * BK)));
  const int kchunk = ceil_div(ceil_div(std::max(K, 1), split), BK) * BK;
  split = ceil_div(std::max(K, 1), kchunk);
  const long pstride = static_cast<long>(ldc) * N;
  double* P = partials().get(static_cast<size_t>(split) * pstride);
  const int gx = ceil_div(M, BM), gy = ceil_div(N, BN);
  k_dgemm<true><<<xcd_blocks(gx * gy * split), NT, 0, s>>>(M, N, K, kchunk, 1.0, A, lda, B, ldb, 0.0, P, ldc,
                                                           pstride, gx, gy, split);
  WH_HIP(hipGetLastError());
  const long n = static_cast<long>(M) * N;
  k_reduce_g0<<<ceil_div(n, 256), 256, 0, s>>>(M, N, split, P, pstride, alpha, g0, A, lda, B, ldb, scale, C, ldc);
  WH_HIP(hipGetLastError());
}

static void dgemm_bra_skinny(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                             double g0, const double* scale, double* C, long ldc, hipStream_t s) {
  // about 2048 workgroups, K chunks of at least 1024
  int split = std::max(1, std::min(ceil_div(2048, N), K / 1024));
  const int kchunk = ceil_div(std::max(K, 1), split);
  split = ceil_div(std::max(K, 1), kchunk);
  const long pstride = static_cast<long>(ldc) * N;
  double* P = partials().get(static_cast<size_t>(split) * pstride);
  dim3 grid(N, split);
  if (M <= 8) k_bra_skinny<8><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, P, ldc, pstride);
  else if (M <= 16) k_bra_skinny<16><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, P, ldc, pstride);
  else k_bra_skinny<32><<<grid, NTS, 0, s>>>(M, K, kchunk, A, lda, B, ldb, P, ldc, pstride);
  WH_HIP(hipGetLastError());
  const long n = static_cast<long>(M) * N;
  k_reduce_g0<<<ceil_div(n, 256), 256, 0, s>>>(M, N, split, P, pstride, alpha, g0, A, lda, B, ldb, scale, C, ldc);
  WH_HIP(hipGetLastError());
}

static void dgemm_bra_lib(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                          double g0, const double* scale, double* C, long ldc, hipStream_t s) {
  const double zero = 0.0;
  WH_REQUIRE(rocblas_dgemm(blas(s), rocblas_operation_transpose, rocblas_operation_none, M, N, K, &alpha, A, lda, B,
                           ldb, &zero, C, ldc) == rocblas_status_success,
             "rocblas_dgemm failed");
  g0_fix_scale(M, N, g0, A, lda, B, ldb, scale, C, ldc, s);
}

void dgemm_bra(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb, double g0,
               const double* scale, double* C, long ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  const Backend be = backend();
  if (be == Backend::Rocblas && K > 0) return dgemm_bra_lib(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s);
  // a few rows and columns: the matrix-core tiles (128 x 128) would be almost
  // empty (with many columns the tiles' reuse of A wins again)
  if (M <= 32 && N <= 256 && K > 0) return dgemm_bra_skinny(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s);
  if (be == Backend::Own || K <= 0) return dgemm_bra_own(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s);
  // variants: 0..2 the kernel with its K split aimed at 512, 1024, 2048 workgroups; 3 rocBLAS
  static std::map<std::tuple<int, int, int>, int> chosen;   // (M, K, log2 N) -> variant
  static std::mutex chosen_mutex;
  std::unique_lock<std::mutex> lk(chosen_mutex);
  int lg = 0;
  while ((2 << lg) <= N) ++lg;
  const auto key = std::make_tuple(M, K, lg);
  auto run = [&](int v) {
    if (v == 3) dgemm_bra_lib(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s);
    else dgemm_bra_own(M, N, K, alpha, A, lda, B, ldb, g0, scale, C, ldc, s, 512 << v);
  };
  auto it = chosen.find(key);
  if (it == chosen.end() && N >= 32) {
    // time every variant on this call's arguments (a first round warms them up)
    hipEvent_t a, b;
    WH_HIP(hipEventCreate(&a));
    WH_HIP(hipEventCreate(&b));
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    for (int rep = 0; rep < 2; ++rep)
      for (int v = 0; v < 4; ++v) {
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
    for (int v = 1; v < 4; ++v)
      if (t[v] < t[best]) best = v;
    if (tune_log())
      std::fprintf(stderr, "[west_hip] dgemm_bra M=%d N=%d K=%d: own512 %.3f own1024 %.3f own2048 %.3f rocblas %.3f ms -> %d\n",
                   M, N, K, t[0], t[1], t[2], t[3], best);
    it = chosen.emplace(key, best).first;
  }
  const int pick = it != chosen.end() ? it->second : 0;
  lk.unlock();
  // recomputed with the chosen variant, so the result does not depend on the timing order
  run(pick);
}

void g0_fix_scale(int M, int N, double g0, const double* A, long lda, const double* B, long ldb,
                  const double* scale, double* C, long ldc, hipStream_t s) {
  const long n = static_cast<long>(M) * N;
  if (n <= 0) return;
  k_g0_scale<<<ceil_div(n, 256), 256, 0, s>>>(M, N, g0, A, lda, B, ldb, scale, C, ldc);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
