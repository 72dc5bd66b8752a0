// Gamma-point 3D FFTs on the G-sphere "stick" layout (declarations).
//
// The wave-function sphere is stored in stick order: columns of constant
// Miller (i, j) in the Gamma half plane (i > 0, or i == 0 and j >= 0), column
// (0,0) first, k ascending inside a column, so that G = 0 is element 0.
// Every column c has a mirror column (-i, -j) that is never stored; it is
// generated from the stored one by Hermitian symmetry.
//
// A transform pair (inverse, pointwise real multiply, forward) of two real
// bands packed as a1 + i a2 runs as three kernels:
//   zinv  : sphere -> sticks    (1D inverse FFT along a3 for stored + mirror columns)
//   plane : sticks -> sticks    (per a3-plane: inverse FFT along a2 on |i| <= imax rows,
//                                inverse FFT along a1, pointwise op, forward FFT along a1
//                                keeping |i| <= imax, forward FFT along a2 keeping columns)
//   zfwd  : sticks -> sphere    (1D forward FFT along a3, fp/fm unpacking, epilogue)
// so the full real-space grid never goes to HBM for the local-potential term.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "wh_common.h"

namespace wh {

constexpr int kMaxStages = 8;

struct FftAxis {
  int n;
  int nstage;
  int radix[kMaxStages];
};

struct StickPlanDev {
  int n1, n2, n3;
  int npw;        // points in the stored half sphere
  int ncs;        // stored columns
  int nct;        // all columns (2*ncs - 1)
  int imax;       // max |i| over columns
  int nrows;      // 2*imax + 1
  int kext;       // max |k| over the sphere
  int jext;       // max |j| over the stored columns
  int zblk_pts;   // most points in kZregCols consecutive stored columns
  const int* col_off;   // (ncs+1)
  const int* col_kmin;  // (ncs)
  const int* col_pos;   // (nct) (i+imax)*n2 + (j mod n2)
  const int* pt_col;    // (npw) stored column of each point
  const int* pt_k;      // (npw) Miller k of each point
  FftAxis ax1, ax2, ax3;
  const double* tw1;    // (2*n1) forward roots exp(-2 pi i m/n1), interleaved
  const double* tw2;
  const double* tw3;
};

// stored columns per workgroup of the register-line z kernels
constexpr int kZregCols = 128;

struct TplKernels;
struct TplKernels32;

struct StickPlan {
  StickPlanDev dev{};
  const TplKernels* tpl = nullptr;   // size-specialised kernels for this grid, if compiled
  const TplKernels32* tpl32 = nullptr;   // FP32 operator kernels for this grid, if compiled
  // workgroups per CU the size-specialised plane kernel's LDS is sized for
  // (FP64 / FP32 operator); 0 until timed on the first large launch
  mutable int vloc_occ = 0;
  mutable int vloc32_occ = 0;
  std::vector<int> col_off, col_kmin, col_pos, pt_col, pt_k;
  std::vector<int> col_i, col_j;   // stored columns' Miller (i, j)
  std::vector<int> perm;           // stick position -> caller index (0-based)
  std::vector<void*> allocations;
  void build(int nr1, int nr2, int nr3, int npw, const int* nl_1based, const int* nlm_1based);
  void release();
  long sticks_elems() const { return static_cast<long>(dev.n3) * dev.nct; }
};

// Variant of the plane kernel for a launch of nitems items: a fixed choice
// (WH_PLANE_OCC or an earlier timing) or, for the first launch of at least
// 256 items, the fastest of 1..max_occ timed on that launch itself (the
// variants compute the same transform; the launch's output is the last
// variant's).  Smaller launches before any timing run variant `untimed`.
// run(v) enqueues one launch.
template <class Run>
void plane_occ_launch(int& chosen, int max_occ, int nitems, Run run, hipStream_t s, int untimed = 1) {
  if (chosen > 0 || max_occ <= 1 || nitems < 256) {
    run(chosen > 0 ? std::min(chosen, max_occ) : std::min(std::max(untimed, 1), max_occ));
    return;
  }
  hipEvent_t a, b;
  WH_HIP(hipEventCreate(&a));
  WH_HIP(hipEventCreate(&b));
  // warm-up: each variant's first launch also loads its code
  for (int occ = 1; occ <= max_occ; ++occ) run(occ);
  float best = 0.f;
  int pick = 1;
  for (int occ = 1; occ <= max_occ; ++occ) {
    WH_HIP(hipEventRecord(a, s));
    run(occ);
    WH_HIP(hipEventRecord(b, s));
    WH_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    WH_HIP(hipEventElapsedTime(&ms, a, b));
    if (tune_log()) std::fprintf(stderr, "[west_hip] plane variant %d: %.3f ms for %d items\n", occ, ms, nitems);
    if (occ == 1 || ms < best) {
      best = ms;
      pick = occ;
    }
  }
  WH_HIP(hipEventDestroy(a));
  WH_HIP(hipEventDestroy(b));
  chosen = pick;
}

// One transform item: bands a and b (row indices into a vector block; b < 0 -> zero),
// pot = index of the real-space potential (per-item multiply) or -1 for the shared one.
// a < 0 marks an inactive item: the kernels skip it (launches sized for a
// capacity whose tail the device has switched off).
struct FftItem {
  int a, b, pot, pad;
};

enum ZinvMode : int {
  ZINV_PAIR = 0,      // b(nl) = a1 + i a2, b(nlm) = conj(a1 - i a2) (nlm written last)
  ZINV_PAIR_REALG0 = 1  // same, but G = 0 takes (Re a1, Re a2): exact real parts of two single transforms
};

enum ZfwdMode : int {
  ZFWD_PAIR_STORE = 0,   // out_a = scale*b1, out_b = scale*b2
  ZFWD_PAIR_STERN = 1,   // out = b1 + (g2kin - e) * in   (local potential + kinetic - e)
  ZFWD_SINGLE = 2        // out_a = scale * F(nl) * w(G), w = sqvc or 1 ; G = 0 set to zero if requested
};

struct ZfwdArgs {
  int mode;
  double scale;
  const double* g2kin;    // stick order
  const double* e;        // per-row eigenvalue (STERN), indexed by the out row
  const double* weight;   // per-G weight (SINGLE), may be null
  int zero_g0;
};

// Full complex grids of the k-point path (QE layout, first index fastest):
// dir = +1 is QE's invfft (unnormalised exp(+iGr)), dir = -1 the forward sum
// without the 1/N (the caller scales).  A batch is `batch` grids back to back.
struct GridFft {
  int n1 = 0, n2 = 0, n3 = 0;
  FftAxis ax1{}, ax2{}, ax3{};
  const double* tw1 = nullptr;
  const double* tw2 = nullptr;
  const double* tw3 = nullptr;
  std::vector<void*> allocations;
  void build(int nr1, int nr2, int nr3);
  void release();
  long nnr() const { return static_cast<long>(n1) * n2 * n3; }
};
void launch_grid_fft(const GridFft& g, z64* data, int batch, int dir, hipStream_t s);
// t[b] = (1/N_r) nl-gather(F^-1 V(r) F nl-scatter(x[b])) + (g2 - e[b]) x[b] for m bands (rows of
// length ld, npw points, zero beyond) in one launch with the grid in LDS, where every grid
// dimension has a compiled line transform and the grid fits the LDS (wh_kfft_lds.hip);
// d_m (device int, may be null): only the first min(m, *d_m) bands
bool kvloc_lds_available(int n1, int n2, int n3);
void kvloc_lds(int n1, int n2, int n3, int m, int npw, long ld, const int* nl, const z64* x, const double* vrs,
               const double* g2, const double* e, z64* t, hipStream_t s, const int* d_m = nullptr);
// One CG iteration of the k-point Sternheimer solve per listed column (lb <
// *nc of cap launched; band l2i[lb]) in one workgroup with the grid in LDS:
// direction h = -h + beta hold (bc = h), t = (H_k - eu + projectors) h with
// the local term in LDS and the projector block q2 (real view, 2 ld rows,
// nq2 <= 32 columns, row scales qs), the CG step on x, g, hold, and the next
// preconditioned residual h and rho[band] (wh_kfft_lds.hip)
// Columns of several k points in one launch: ktab (device, may be null) holds
// each k point's sphere and projectors, and band column col belongs to k point
// col / cols_per_k (ktab null: npw, nl, g2 and q2 for every column).
struct KlKpoint {
  int npw;
  const int* nl;
  const double* g2;
  const double* q2;      // real view: evc, i evc, vkb, i vkb (2 ld rows each)
  const double* et;      // nbndval energies
  const double* eprec;   // nbndval preconditioner energies
};
bool cg_iter_lds_available(int n1, int n2, int n3, int nq2);
// The grids of dfpt at q on LDS grids for every k point in one launch each
// (kvloc_lds_available grids; qtab: per k point its k - q index and g0):
// states_kmq_lds: psik[k][v] = F^-1[evc_v at k - q] exp(-i g0.r) (nks nb grids);
// rhs_lds: out[k m nb + p nb + v] = (1/N_r) F[psik[k][v] aux[p]] on the k sphere;
// density_lds: dng[p] += sum_k sck[k] F[sum_v conj(psik[k][v]) F^-1[dpsi[k m nb + p nb + v]]]
// on the q sphere (dsc: nks m grids of scratch, dk: nks m columns of ldq).
struct KlQpoint {
  int ikq, g1, g2, g3;
};
void states_kmq_lds(int n1, int n2, int n3, int nks, int nb, long ld, const KlKpoint* ktab, const KlQpoint* qtab,
                    z64* psik, hipStream_t s);
// For the nks m nb right-hand side columns f (k point k = col / (m nb), band
// v = col % nb; 2 nb <= 32): e[col] = et(k - q)[v], ep[col] = eprec(k)[v],
// f <- -f + evc (evc^H f) on the k sphere (apply_alpha_pc with alpha = -1,
// dfpt_module.f90:346) and x = f / max(1, |k+G|^2 / ep) (348)
void pc_precond_all(int nks, int nb, int m, long ld, const KlKpoint* ktab, const KlQpoint* qtab, z64* f, z64* x,
                    double* e, double* ep, hipStream_t s);
// The CG's first step for m = nks cols columns of cg_iter_lds's kind
// (linsolve_sternheimer_m_wfcts.f90:127-205): g = (H_k - e + alpha P_v) x - b
// with the local term in LDS, h = g / max(1, |k+G|^2 / eprec), rho[col] = Re <h|g>
void cg_first_lds(int n1, int n2, int n3, long ld, const double* vrs, int nq2, const double* qs, int m,
                  const double* e, const double* ep, const z64* x, const z64* b, z64* g, z64* h, double* rho,
                  const KlKpoint* ktab, int cols_per_k, hipStream_t s);
void rhs_lds(int n1, int n2, int n3, int nks, int nb, int m, long ld, const KlKpoint* ktab, const z64* psik,
             const z64* aux, z64* out, hipStream_t s);
void density_lds(int n1, int n2, int n3, int nks, int nb, int m, long ld, const KlKpoint* ktab, const z64* psik,
                 const z64* dpsi, z64* dsc, int npwq, long ldq, const int* nlq, const double* sck, z64* dk, z64* dng,
                 hipStream_t s);
void cg_iter_lds(int n1, int n2, int n3, int npw, long ld, const int* nl, const double* vrs, const double* g2,
                 int nq2, const double* q2, const double* qs, int cap, const int* l2i, const int* nc,
                 const double* beta, const double* eu, const double* ep, z64* h, z64* hold, z64* x, z64* g, z64* bc,
                 z64* tc, double* rho, hipStream_t s, const KlKpoint* ktab = nullptr, int cols_per_k = 1);

// kernel launchers (double precision)
void launch_zinv(const StickPlan& p, const FftItem* d_items, int nitems, const z64* psi, long ldpsi,
                 z64* sticks, int mode, hipStream_t s);
// bands a from psiA, bands b from psiB
void launch_zinv(const StickPlan& p, const FftItem* d_items, int nitems, const z64* psiA, long ldA,
                 const z64* psiB, long ldB, z64* sticks, int mode, hipStream_t s);
void launch_plane_vloc(const StickPlan& p, const FftItem* d_items, int nitems, const z64* sin, z64* sout,
                       const double* vpot, long ldpot, double scale, hipStream_t s);
void launch_plane_to_real(const StickPlan& p, const FftItem* d_items, int nitems, const z64* sin,
                          double* out, long ldout, hipStream_t s);
// scratch reals per perturbation launch_plane_density needs for this grid (0: none)
long plane_density_scratch(const StickPlan& p);
void launch_plane_density(const StickPlan& p, int npert, int nband, const z64* sin, const double* occ,
                          z64* sout, double* scratch, hipStream_t s);
void launch_zfwd(const StickPlan& p, const FftItem* d_items, int nitems, const z64* sticks,
                 const z64* in, long ldin, z64* out, long ldout, const ZfwdArgs& args, hipStream_t s);

}  // namespace wh
