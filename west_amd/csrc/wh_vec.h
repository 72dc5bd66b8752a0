// Vector kernels of the Sternheimer CG and dfpt data movement (declarations).
#pragma once

#include "wh_common.h"
#include "wh_fft.h"

namespace wh {

void prec_rho(int nc, int npw, int gstart, const int* l2i, const z64* g, z64* h, long ld, const double* g2kin,
              const double* eprec, double* rho, hipStream_t s);
void direction(int nc, int npw, const int* l2i, const double* beta, z64* h, const z64* hold, z64* bc, long ld,
               hipStream_t s);
void cg_update(int nc, int npw, int gstart, const int* l2i, const z64* h, z64* g, const z64* tc, z64* x, long ld,
               double* ac_out, hipStream_t s);
void gather_sticks(int m, int npw, const int* perm, const z64* in, long ldin, z64* out, long ldout, const double* w,
                   hipStream_t s);
void scatter_caller(int m, int npw, int npwx, const int* perm, const z64* in, long ldin, z64* out, long ldout,
                    hipStream_t s);
void precondition(int m, int npw, const z64* f, z64* pf, long ld, const double* g2kin, const double* eprec,
                  int period, hipStream_t s);
void kinetic_minus_e(int m, int npw, const z64* x, z64* t, long ld, const double* g2kin, const double* e,
                     hipStream_t s);
void axpby(long n, double a, const z64* x, double b, z64* y, hipStream_t s);
void fill_pairs(FftItem* items, int m, int row0, hipStream_t s);
void fill_ident(FftItem* items, int n, hipStream_t s);
void eprec_of(int nb, int npw, const z64* evc, long ld, const double* g2kin, double* eprec, hipStream_t s);

}  // namespace wh
