// Split-BF16 matrix-core GEMMs with FP32 accumulation (declarations).
//
// A real FP32 operand x is carried as bf16 pieces, x ~ hi + mid + lo with
// hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid). With terms = 3
// a product is summed as hi*hi + hi*mid + mid*hi (about 16 significant bits);
// with terms = 6 the three products of the next order are added (about the
// 24 bits of FP32). All products run on v_mfma_f32_16x16x32_bf16.
#pragma once

#include "wh_common.h"

namespace wh {

// pieces of a (rows x cols) FP32 matrix (leading dimension ld_in) into three
// (rows x ld_out) bf16 arrays p[0..2], zero in columns [cols, ld_out)
void split_bf16(int rows, int cols, const float* in, long ld_in, void* const p[3], long ld_out, hipStream_t s);

// same for the transpose: p_q[c][r] = piece q of in[r][c], rows of ld_out >= rows, zero padded
void split_bf16_t(int rows, int cols, const float* in, long ld_in, void* const p[3], long ld_out, hipStream_t s);

// out[i*ci + j*cj] = alpha * sum_{k<K} A[i][k] B[j][k] + beta * out[...]
//   A: bf16 pieces (rows of lda); B: FP32 rows of ldb;
//   lda and ldb are multiples of 32 and >= K rounded up to 32 (loads stay in the row).
void gemm_tn_bf16(int terms, int M, int N, int K, float alpha, const void* const a[3], long lda, const float* b,
                  long ldb, float beta, float* out, long ci, long cj, hipStream_t s);

// Gamma bra product of the FP32 path:
//   C[j*ldc + i] = scale_i * (2 * sum_k A[i][k] B[j][k] - g0 * a0[i*lda] * B[j][0])   for i < M,
//   C[j*ldc + i] = 0 for M <= i < ldc,
// a0: the FP32 matrix the pieces were cut from. The K sum is split over
// workgroups into FP32 partials (work) and summed in a fixed order in FP64.
void gemm_bra_bf16(int terms, int M, int N, int K, const void* const a[3], long lda, const float* a0, const float* b,
                   long ldb, float g0, const double* scale, float* c, long ldc, float* work, size_t work_floats,
                   hipStream_t s);

// FP32 partial-sum workspace (floats) gemm_bra_bf16 needs for this shape
size_t gemm_bra_bf16_work(int M, int N, int K);

}  // namespace wh
