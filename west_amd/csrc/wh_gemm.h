// FP64 matrix-core GEMMs (declarations).
#pragma once

#include "wh_common.h"

namespace wh {

// C (M x N) = alpha * op(A) * B + beta * C, column major; op(A) = A^T if transa.
// d_n (device int, may be null): only the first min(N, *d_n) columns of B and
// C are used (launches sized for a capacity whose tail the device switched off).
void dgemm(bool transa, int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
           double beta, double* C, long ldc, hipStream_t s, const int* d_n = nullptr);

// Gamma bra product: C(i, j) = scale_i * (alpha * (A^T B)(i, j) - g0 * A[i*lda] * B[j*ldb])
// (A stored K x M); K is split across workgroups and summed in a fixed order.
void dgemm_bra(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb, double g0,
               const double* scale, double* C, long ldc, hipStream_t s, const int* d_n = nullptr);

// C(i, j) = scale_i * (C(i, j) - g0 * A[i*lda] * B[j*ldb])   (scale may be null)
void g0_fix_scale(int M, int N, double g0, const double* A, long lda, const double* B, long ldb,
                  const double* scale, double* C, long ldc, hipStream_t s);

// The same products for callers that must neither allocate device memory
// nor time variants (the variant an earlier timed call chose for the shape,
// else a fixed one).  bra_partials_doubles picks the bra product's variant
// and returns how many doubles of split-K partial sums it needs;
// dgemm_bra_into runs that variant with its partial sums in P.
long bra_partials_doubles(int M, int N, int K, long ldc, int* variant);
void dgemm_bra_into(int M, int N, int K, double alpha, const double* A, long lda, const double* B, long ldb,
                    double g0, const double* scale, double* C, long ldc, double* P, int variant, hipStream_t s);
void dgemm_untimed(bool transa, int M, int N, int K, double alpha, const double* A, long lda, const double* B,
                   long ldb, double beta, double* C, long ldc, hipStream_t s);

}  // namespace wh
