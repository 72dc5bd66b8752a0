// Split-BF16 GEMMs on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Used by the FP32 Sternheimer operator for its projector terms
// (calbec / add_vuspsi and the alpha_pv P_v term,
//  DFPT_kernel/apply_sternheimerop_to_m_wfcs.f90:57-76,
//  Hamiltonian_kernel/apply_alpha_pv_to_m_wfcs.f90:56-78).
//
// Both operands of the "TN" product C[i][j] = sum_k A[i][k] B[j][k] keep k
// contiguous, which is the layout the bf16 MFMA wants per lane: lane l holds
// 8 consecutive k of row (l % 16) starting at k = 8 (l / 16), for A and for B;
// the 16x16 result sits at row 4 (l / 16) + r, column l % 16, register r.
//
// Tiles: 128 x 128 x 32 per 256-thread workgroup, each wavefront a 64 x 64
// block of 4 x 4 MFMA tiles. A arrives split into bf16 pieces; B is FP32 and
// is split while it is staged into LDS.
#include <cstdint>

#include "wh_gemm_bf16.h"

namespace wh {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 32;
constexpr int LDP = BK + 8;   // bf16 per LDS row (16-byte pad against bank conflicts)

// TERMS = 3: x = hi + lo, products hi*hi + hi*lo + lo*hi (about 16 bits);
// TERMS = 6: x = hi + mid + lo, products of total order <= 2 in the pieces'
// scale (hi*hi, hi*mid, mid*hi, hi*lo, mid*mid, lo*hi: about FP32 accuracy).
// bf16 pieces of the A operand (hi, mid, lo), passed by value
struct PiecePtrs {
  const __bf16* p[3];
};

template <int TERMS>
struct Pieces {
  static constexpr int n = TERMS == 3 ? 2 : 3;
};

template <int TERMS>
__device__ __forceinline__ void split4(f4 v, bf4 (&p)[Pieces<TERMS>::n]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    float r = v[e];
#pragma unroll
    for (int q = 0; q < Pieces<TERMS>::n; ++q) {
      const __bf16 h = static_cast<__bf16>(r);
      p[q][e] = h;
      r -= static_cast<float>(h);
    }
  }
}

// The grid is 1D over gx x gy x gz tiles (gz: K chunks) in an XCD-aware
// order: blocks are dispatched round-robin over the 8 XCDs (block L on XCD
// L % 8), and block L runs tile W = (L % 8) ceil(T / 8) + L / 8, so
// consecutive tiles, which share an A or B panel, share an XCD and its L2.
constexpr int kXcds = 8;
inline int xcd_blocks(int tiles) { return (tiles + kXcds - 1) / kXcds * kXcds; }

template <int TERMS, bool PARTIAL>
__global__ __launch_bounds__(NT) void k_gemm_tn(int M, int N, int K, int kchunk, PiecePtrs ap, long lda, const float* __restrict__ b, long ldb, float alpha,
                                                float beta, float* out, long ci, long cj, long pstride, int vec_out,
                                                int gx, int gy, int gz) {
  constexpr int NP = Pieces<TERMS>::n;
  __shared__ __attribute__((aligned(16))) __bf16 sa[NP][BM * LDP];
  __shared__ __attribute__((aligned(16))) __bf16 sb[NP][BN * LDP];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w & 1;
  const int wn = w >> 1;
  const int tiles = gx * gy * gz;
  const int W = static_cast<int>(blockIdx.x % kXcds) * ((tiles + kXcds - 1) / kXcds) + static_cast<int>(blockIdx.x / kXcds);
  if (W >= tiles) return;   // uniform per workgroup
  const int bz = W / (gx * gy);
  const int i0 = (W % gx) * BM;
  const int j0 = ((W / gx) % gy) * BN;
  const int kb = bz * kchunk;
  const int ke = min(K, kb + kchunk);
  const __bf16* apc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) apc[q] = ap.p[q];

  // global -> register staging
  bf8 ra[2][NP];
  f4 rb[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int g = t + u * NT;
      const int r = g >> 2;
      const int kk = k0 + 8 * (g & 3);
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        bf8 v = {};
        if (i0 + r < M) {
          v = *reinterpret_cast<const bf8*>(apc[q] + static_cast<long>(i0 + r) * lda + kk);
          if (kk + 8 > K) {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (kk + e >= K) v[e] = static_cast<__bf16>(0.0f);
          }
        }
        ra[u][q] = v;
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = t + u * NT;
      const int r = f >> 3;
      const int kk = k0 + 4 * (f & 7);
      f4 v = {0.f, 0.f, 0.f, 0.f};
      if (j0 + r < N) {
        v = *reinterpret_cast<const f4*>(b + static_cast<long>(j0 + r) * ldb + kk);
        if (kk + 4 > K) {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (kk + e >= K) v[e] = 0.f;
        }
      }
      rb[u] = v;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int g = t + u * NT;
      const int o = (g >> 2) * LDP + 8 * (g & 3);
#pragma unroll
      for (int q = 0; q < NP; ++q) *reinterpret_cast<bf8*>(&sa[q][o]) = ra[u][q];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = t + u * NT;
      const int o = (f >> 3) * LDP + 4 * (f & 7);
      bf4 p[NP];
      split4<TERMS>(rb[u], p);
#pragma unroll
      for (int q = 0; q < NP; ++q) *reinterpret_cast<bf4*>(&sb[q][o]) = p[q];
    }
  };

  f4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = f4{0.f, 0.f, 0.f, 0.f};

  const int lr = lane & 15;
  const int lk = 8 * (lane >> 4);
  if (kb < ke) load(kb);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    __syncthreads();   // the previous step has finished reading LDS
    stage();
    __syncthreads();
    if (k0 + BK < ke) load(k0 + BK);
    bf8 fb[4][NP];
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int o = (wn * 64 + y * 16 + lr) * LDP + lk;
#pragma unroll
      for (int q = 0; q < NP; ++q) fb[y][q] = *reinterpret_cast<const bf8*>(&sb[q][o]);
    }
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int o = (wm * 64 + x * 16 + lr) * LDP + lk;
      bf8 fa[NP];
#pragma unroll
      for (int q = 0; q < NP; ++q) fa[q] = *reinterpret_cast<const bf8*>(&sa[q][o]);
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        // smallest products first
        if constexpr (TERMS == 6) {
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[2], fb[y][0], acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1], fb[y][1], acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][2], acc[x][y], 0, 0, 0);
        }
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[1], fb[y][0], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][1], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0], fb[y][0], acc[x][y], 0, 0, 0);
      }
    }
  }

  float* o0 = out + (PARTIAL ? static_cast<long>(bz) * pstride : 0L);
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int ib = i0 + wm * 64 + x * 16 + 4 * (lane >> 4);
      const int j = j0 + wn * 64 + y * 16 + lr;
      if (j >= N) continue;
      float* p = o0 + static_cast<long>(ib) * ci + static_cast<long>(j) * cj;
      if (vec_out && ib + 3 < M) {
        f4 v = acc[x][y];
        if (!PARTIAL) {
          v = v * alpha;
          if (beta != 0.f) v += beta * *reinterpret_cast<const f4*>(p);
        }
        *reinterpret_cast<f4*>(p) = v;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (ib + r >= M) break;
          float* q = p + static_cast<long>(r) * ci;
          float v = acc[x][y][r];
          if (!PARTIAL) v = alpha * v + (beta != 0.f ? beta * *q : 0.f);
          *q = v;
        }
      }
    }
}

// c[j*ldc + i] = scale_i (2 sum_z P[z][j][i] - g0 a0[i] b[j][0]); 0 in the row padding
__global__ void k_bra_reduce(int M, int N, int splits, const float* P, const float* a0, long lda0, const float* b,
                             long ldb, float g0, const double* scale, float* c, long ldc) {
  const long idx = static_cast<long>(blockIdx.x) * NT + threadIdx.x;
  if (idx >= ldc * N) return;
  const int i = static_cast<int>(idx % ldc);
  const int j = static_cast<int>(idx / ldc);
  if (i >= M) {
    c[idx] = 0.f;
    return;
  }
  double s = 0.0;
  for (int z = 0; z < splits; ++z) s += P[(static_cast<long>(z) * N + j) * M + i];
  double v = 2.0 * s;
  if (g0 != 0.f) v -= static_cast<double>(g0) * a0[static_cast<long>(i) * lda0] * b[static_cast<long>(j) * ldb];
  if (scale) v *= scale[i];
  c[idx] = static_cast<float>(v);
}

// out_q[r][c] = piece q of in[r][c] (q = 0, 1, 2: hi, mid, lo); 0 in columns [cols, ld_out)
__global__ void k_split(int rows, int cols, const float* in, long ld_in, __bf16* p0, __bf16* p1, __bf16* p2,
                        long ld_out) {
  const int r = blockIdx.y;
  const long c = static_cast<long>(blockIdx.x) * NT + threadIdx.x;
  if (c >= ld_out) return;
  float v = c < cols ? in[static_cast<long>(r) * ld_in + c] : 0.f;
  const long o = static_cast<long>(r) * ld_out + c;
  __bf16* p[3] = {p0, p1, p2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const __bf16 h = static_cast<__bf16>(v);
    p[q][o] = h;
    v -= static_cast<float>(h);
  }
}

// out_q[c][r] = piece q of in[r][c]; r in [rows, ld_out) -> 0
__global__ void k_split_t(int rows, int cols, const float* in, long ld_in, __bf16* p0, __bf16* p1, __bf16* p2,
                          long ld_out) {
  const int c = blockIdx.y;
  const long r = static_cast<long>(blockIdx.x) * NT + threadIdx.x;
  if (r >= ld_out || c >= cols) return;
  float v = r < rows ? in[r * ld_in + c] : 0.f;
  const long o = static_cast<long>(c) * ld_out + r;
  __bf16* p[3] = {p0, p1, p2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const __bf16 h = static_cast<__bf16>(v);
    p[q][o] = h;
    v -= static_cast<float>(h);
  }
}

struct SplitK {
  int splits, kchunk;
};

SplitK plan_split(int M, int N, int K) {
  const long tiles = static_cast<long>(ceil_div(M, BM)) * ceil_div(N, BN);
  const int ksteps = ceil_div(K, BK);
  int splits = static_cast<int>(std::min<long>(32, std::max<long>(1, (1024 + tiles - 1) / tiles)));
  splits = std::max(1, std::min(splits, ksteps / 4));
  const int kchunk = ceil_div(ksteps, splits) * BK;
  return SplitK{ceil_div(K, kchunk), kchunk};
}

void check_ld(long lda, long ldb, int K) {
  const long kr = static_cast<long>(ceil_div(K, BK)) * BK;
  WH_REQUIRE(lda % 32 == 0 && ldb % 32 == 0 && lda >= kr && ldb >= kr,
             "gemm_tn_bf16: leading dimensions must be multiples of 32 covering K rounded up to 32");
}

}  // namespace

void split_bf16(int rows, int cols, const float* in, long ld_in, void* const p[3], long ld_out, hipStream_t s) {
  if (rows <= 0 || ld_out <= 0) return;
  k_split<<<dim3(ceil_div(ld_out, NT), rows), NT, 0, s>>>(rows, cols, in, ld_in, static_cast<__bf16*>(p[0]),
                                                         static_cast<__bf16*>(p[1]), static_cast<__bf16*>(p[2]), ld_out);
  WH_HIP(hipGetLastError());
}

void split_bf16_t(int rows, int cols, const float* in, long ld_in, void* const p[3], long ld_out, hipStream_t s) {
  if (cols <= 0 || ld_out <= 0) return;
  k_split_t<<<dim3(ceil_div(ld_out, NT), cols), NT, 0, s>>>(rows, cols, in, ld_in, static_cast<__bf16*>(p[0]),
                                                           static_cast<__bf16*>(p[1]), static_cast<__bf16*>(p[2]),
                                                           ld_out);
  WH_HIP(hipGetLastError());
}

namespace {

template <int TERMS, bool PARTIAL>
void launch_tn(dim3 grid, int M, int N, int K, int kchunk, const void* const ap[3], long lda, const float* b, long ldb,
               float alpha, float beta, float* out, long ci, long cj, long pstride, int vec, hipStream_t s) {
  PiecePtrs pp{{static_cast<const __bf16*>(ap[0]), static_cast<const __bf16*>(ap[1]),
                static_cast<const __bf16*>(ap[2])}};
  const int gx = static_cast<int>(grid.x), gy = static_cast<int>(grid.y), gz = static_cast<int>(grid.z);
  k_gemm_tn<TERMS, PARTIAL><<<xcd_blocks(gx * gy * gz), NT, 0, s>>>(M, N, K, kchunk, pp, lda, b, ldb, alpha, beta, out,
                                                                   ci, cj, pstride, vec, gx, gy, gz);
}

}  // namespace

void gemm_tn_bf16(int terms, int M, int N, int K, float alpha, const void* const ap[3], long lda, const float* b,
                  long ldb, float beta, float* out, long ci, long cj, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  WH_REQUIRE(terms == 3 || terms == 6, "gemm_tn_bf16: terms must be 3 or 6");
  check_ld(lda, ldb, K);
  const int vec = (ci == 1 && cj % 4 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0) ? 1 : 0;
  dim3 grid(ceil_div(M, BM), ceil_div(N, BN), 1);
  const int kchunk = ceil_div(std::max(K, 1), BK) * BK;
  if (terms == 3)
    launch_tn<3, false>(grid, M, N, K, kchunk, ap, lda, b, ldb, alpha, beta, out, ci, cj, 0L, vec, s);
  else
    launch_tn<6, false>(grid, M, N, K, kchunk, ap, lda, b, ldb, alpha, beta, out, ci, cj, 0L, vec, s);
  WH_HIP(hipGetLastError());
}

size_t gemm_bra_bf16_work(int M, int N, int K) {
  if (M <= 0 || N <= 0) return 0;
  return static_cast<size_t>(plan_split(M, N, K).splits) * M * N;
}

void gemm_bra_bf16(int terms, int M, int N, int K, const void* const ap[3], long lda, const float* a0, const float* b,
                   long ldb, float g0, const double* scale, float* c, long ldc, float* work, size_t work_floats,
                   hipStream_t s) {
  if (N <= 0) return;
  WH_REQUIRE(terms == 3 || terms == 6, "gemm_bra_bf16: terms must be 3 or 6");
  WH_REQUIRE(ldc >= M && M > 0, "gemm_bra_bf16: need 0 < M <= ldc");
  check_ld(lda, ldb, K);
  const SplitK sp = plan_split(M, N, K);
  WH_REQUIRE(work_floats >= static_cast<size_t>(sp.splits) * M * N, "gemm_bra_bf16: workspace too small");
  const int vec = (M % 4 == 0 && reinterpret_cast<uintptr_t>(work) % 16 == 0) ? 1 : 0;
  dim3 grid(ceil_div(M, BM), ceil_div(N, BN), sp.splits);
  if (terms == 3)
    launch_tn<3, true>(grid, M, N, K, sp.kchunk, ap, lda, b, ldb, 1.f, 0.f, work, 1L, static_cast<long>(M),
                       static_cast<long>(M) * N, vec, s);
  else
    launch_tn<6, true>(grid, M, N, K, sp.kchunk, ap, lda, b, ldb, 1.f, 0.f, work, 1L, static_cast<long>(M),
                       static_cast<long>(M) * N, vec, s);
  WH_HIP(hipGetLastError());
  k_bra_reduce<<<ceil_div(ldc * N, NT), NT, 0, s>>>(M, N, sp.splits, work, a0, lda, b, ldb, g0, scale, c, ldc);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
