// Vector kernels of the Sternheimer CG and dfpt data movement (declarations).
#pragma once

#include "wh_common.h"
#include "wh_fft.h"

namespace wh {

// rho[lb] (rho[l2i[lb]] with by_band) = <M^-1 g|g> for the columns l2i[0..nc)
void prec_rho(int nc, int npw, int gstart, const int* l2i, const z64* g, long ld, const double* g2kin,
              const double* eprec, double* rho, hipStream_t s, bool by_band = false);
void prec_rho(int nc, int npw, int gstart, const int* l2i, const z32* g, long ld, const double* g2kin,
              const double* eprec, double* rho, hipStream_t s, bool by_band = false);
// optional output of cg_update: rho[ib] = <M^-1 g|g> of every updated band's new g
struct RhoOut {
  const double* g2kin = nullptr;
  const double* eprec = nullptr;
  double* rho = nullptr;
};
// h = -M^-1 g + beta hold, bc = h (compact)
// d_nc (device int, may be null): only the first *d_nc of the nc listed columns
// are processed (launches sized for a capacity; the device decides the count)
void direction(int nc, int npw, const int* l2i, const double* beta, const z64* g, const double* g2kin,
               const double* eprec, z64* h, const z64* hold, z64* bc, long ld, hipStream_t s,
               const int* d_nc = nullptr);
void direction(int nc, int npw, const int* l2i, const double* beta, const z32* g, const double* g2kin,
               const double* eprec, z32* h, const z32* hold, z32* bc, long ld, hipStream_t s);
void cg_update(int nc, int npw, int gstart, const int* l2i, const z32* h, z32* g, const z32* tc, z32* x, long ld,
               double* ac_out, hipStream_t s, const RhoOut* rho = nullptr);
void kinetic_minus_e(int m, int npw, const z32* x, z32* t, long ld, const double* g2kin, const double* e,
                     hipStream_t s);
// mixed-precision refinement (u2i: bands of the update list)
//   x[ib] += dx[ib]; dx[ib] = 0; xc[u] = x[ib]
void fold_increment(int nu, int npw, const int* u2i, z64* x, long ldx, z32* dx, long ld32, z64* xc, hipStream_t s);
//   r = tc[u] - b[ib]; rho[u] = <M^-1 r|r>; g[ib] = r (FP32)
void true_residual(int nu, int npw, int gstart, const int* u2i, const z64* tc, const z64* b, long ldx,
                   const double* g2kin, const double* eprec, z32* g, long ld32, double* rho, hipStream_t s);
// g[r] = t[r] - b[r] (FP32 operator output minus the FP64 right-hand side), rows 0..m-1
void residual32(int m, int npw, const z32* t, long ld32, const z64* b, long ldb, z32* g, hipStream_t s);
// out[r][c] = in[r][c] for c < cols, FP64 -> FP32, rows of ld_out (pad left untouched)
void to_f32(int rows, int cols, const double* in, long ld_in, float* out, long ld_out, hipStream_t s);
void to_f64(int rows, int cols, const float* in, long ld_in, double* out, long ld_out, hipStream_t s);
// rows by list: out[u] = in[list[u]] (scatter = false) or out[list[u]] = in[u] (scatter = true)
void copy_rows(int n, int npw, const int* list, const z64* in, long ld_in, z64* out, long ld_out, bool scatter,
               hipStream_t s);
void gather_real(int n, const int* list, const double* in, double* out, hipStream_t s);
void cg_update(int nc, int npw, int gstart, const int* l2i, const z64* h, z64* g, const z64* tc, z64* x, long ld,
               double* ac_out, hipStream_t s, const RhoOut* rho = nullptr, const int* d_nc = nullptr);
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
// device-driven CG (linsolve_sternheimer_m_wfcts.f90:146-222 on the device):
// one workgroup marks bands whose rho < tr2 as converged, compacts the rest
// into l2i (ascending band order) with beta = rho / rhoold (0 on the first
// iteration) and e, sets rhoold = rho for them, and writes the active count
// to *nc and rec[0], and to rec[1] the perturbations among them (groups of
// `group` consecutive bands; group <= 0: one)
void cg_select(int m, double tr2, int first, int group, const double* rhob, double* rhoold, int* conv,
               const double* e_col, int* l2i, double* beta, double* eu, int* nc, int* rec, hipStream_t s);
// pair items (2i, 2i+1) of the first *nc rows, inactive (a = -1) beyond, for cap items
void fill_pairs_dev(FftItem* items, int cap, const int* nc, hipStream_t s);
// plane items of chunked launches: item q -> sticks slot q % chunk while q
// is below the active pair count of *nc rows, inactive beyond (cap items)
void fill_ident_dev(FftItem* items, int cap, int chunk, const int* nc, hipStream_t s);
void iota(int* v, int n, hipStream_t s);
void fill_ident(FftItem* items, int n, hipStream_t s);
void eprec_of(int nb, int npw, const z64* evc, long ld, const double* g2kin, double* eprec, hipStream_t s);

}  // namespace wh
