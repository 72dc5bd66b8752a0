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
// Tile 64 x 64 x 16 per 256-thread workgroup: four wavefronts in a 2 x 2
// arrangement, each owning 32 x 32 = 2 x 2 MFMA 16x16 blocks.  Operand
// layout of the 16x16x4 f64 instruction (checked on the device): lane l
// holds A[l%16][l/16] and B[l/16][l%16]; result register r of lane l holds
// D[l/16 + 4r][l%16].
#include "wh_gemm.h"

namespace wh {

namespace {

constexpr int BM = 64;
constexpr int BN = 64;
constexpr int BK = 16;
constexpr int NT = 256;

typedef double v4d __attribute__((ext_vector_type(4)));

template <bool TA>
__global__ __launch_bounds__(NT) void k_dgemm(int M, int N, int K, double alpha, const double* __restrict__ A,
                                              long lda, const double* __restrict__ B, long ldb, double beta,
                                              double* __restrict__ C, long ldc) {
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1;
  const int wn = wave >> 1;
  const int m0 = blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < K; k0 += BK) {
    // A tile -> As[k][m]
    if (TA) {
      const int m = tid >> 2;
      const int kq = (tid & 3) * 4;
      const int gm = m0 + m;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gk = k0 + kq + q;
        As[kq + q][m] = (gm < M && gk < K) ? A[static_cast<long>(gm) * lda + gk] : 0.0;
      }
    } else {
      const int k = tid >> 4;
      const int mq = (tid & 15) * 4;
      const int gk = k0 + k;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gm = m0 + mq + q;
        As[k][mq + q] = (gm < M && gk < K) ? A[static_cast<long>(gk) * lda + gm] : 0.0;
      }
    }
    {
      const int n = tid >> 2;
      const int kq = (tid & 3) * 4;
      const int gn = n0 + n;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gk = k0 + kq + q;
        Bs[kq + q][n] = (gn < N && gk < K) ? B[static_cast<long>(gn) * ldb + gk] : 0.0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double* c = C + static_cast<long>(col) * ldc + row;
          const double v = alpha * acc[i][j][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
    }
}

// c(i, j) = s_i * (c(i, j) - g0 * a(1, i) * b(1, j)), column-major c (ldc),
// a(1, i) = A[i*lda], b(1, j) = B[j*ldb]   (the DGER of glbrak_gamma + row scaling)
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

}  // namespace

void dgemm(bool transa, int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
           double beta, double* C, long ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 grid(ceil_div(M, BM), ceil_div(N, BN));
  if (K <= 0) {
    // C = beta*C only: a zero-size product still scales C (BLAS semantics)
    if (transa)
      k_dgemm<true><<<grid, NT, 0, s>>>(M, N, 0, alpha, A, lda, B, ldb, beta, C, ldc);
    else
      k_dgemm<false><<<grid, NT, 0, s>>>(M, N, 0, alpha, A, lda, B, ldb, beta, C, ldc);
  } else if (transa) {
    k_dgemm<true><<<grid, NT, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  } else {
    k_dgemm<false><<<grid, NT, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  WH_HIP(hipGetLastError());
}

void g0_fix_scale(int M, int N, double g0, const double* A, long lda, const double* B, long ldb,
                  const double* scale, double* C, long ldc, hipStream_t s) {
  const long n = static_cast<long>(M) * N;
  if (n <= 0) return;
  k_g0_scale<<<ceil_div(n, 256), 256, 0, s>>>(M, N, g0, A, lda, B, ldb, scale, C, ldc);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
