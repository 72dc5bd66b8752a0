// Local-potential term of the Sternheimer operator at a k point on a small
// full grid, in one workgroup per band with the whole grid in LDS:
//   t = (1/N_r) gather(F^-1 V(r) F scatter(x)) + (|k+G|^2 - e) x
// (the same result as k_scatter, the two GridFft transforms, k_mul_real and
// k_gather of the q-point path, which make six launches and six HBM round
// trips of the grid per band).  Each 1D pass gives one line to one thread,
// transformed in registers with a compile-time radix plan for its length;
// the a3 pass applies V(r) between its inverse and forward transforms.  The
// LDS grid has an odd row pitch, so lines read in lock step by neighbouring
// threads fall on distinct banks.
#include <algorithm>

#include "wh_fft.h"
#include "wh_fft_dev.h"

namespace wh {

namespace {

constexpr int KL_NT = 512;

// 2,3,5-smooth line lengths with a compiled transform (up to 24: a line of
// 24 complex values and its radix passes fit the registers of 2 waves per SIMD)
#define WH_KL_SIZES(X) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24)

constexpr bool kl_size(int n) {
#define WH_KL_CASE(NN) n == NN ||
  return WH_KL_SIZES(WH_KL_CASE) false;
#undef WH_KL_CASE
}

template <int NL, int DIR>
__device__ __forceinline__ void kl_line(z64* p, int stride) {
  z64 v[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) v[i] = p[i * stride];
  fft_line<double, NL, DIR>(v);
#pragma unroll
  for (int i = 0; i < NL; ++i) p[i * stride] = v[i];
}

// inverse, times V(r) (vr: the line's potential, stride vs), forward
template <int NL>
__device__ __forceinline__ void kl_line_vloc(z64* p, int stride, const double* vr, long vs) {
  z64 v[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) v[i] = p[i * stride];
  fft_line<double, NL, +1>(v);
#pragma unroll
  for (int i = 0; i < NL; ++i) v[i] = cscale(v[i], vr[i * vs]);
  fft_line<double, NL, -1>(v);
#pragma unroll
  for (int i = 0; i < NL; ++i) p[i * stride] = v[i];
}

template <int DIR>
__device__ void kl_pass(int n, z64* p, int stride) {
  switch (n) {
#define WH_KL_CASE(NN)          \
  case NN:                      \
    kl_line<NN, DIR>(p, stride); \
    break;
    WH_KL_SIZES(WH_KL_CASE)
#undef WH_KL_CASE
    default: break;
  }
}

__device__ void kl_pass_vloc(int n, z64* p, int stride, const double* vr, long vs) {
  switch (n) {
#define WH_KL_CASE(NN)                    \
  case NN:                                \
    kl_line_vloc<NN>(p, stride, vr, vs);  \
    break;
    WH_KL_SIZES(WH_KL_CASE)
#undef WH_KL_CASE
    default: break;
  }
}

// LDS index of grid point (i1, i2, i3): pitch p1 (odd) along a1
__device__ __forceinline__ int kl_at(int i1, int i2, int i3, int p1, int n2) { return i1 + p1 * (i2 + n2 * i3); }

__global__ __launch_bounds__(KL_NT) void k_kvloc_lds(int n1, int n2, int n3, int p1, int npw, long ld,
                                                   const int* __restrict__ nl, const z64* __restrict__ x,
                                                   const double* __restrict__ vrs, const double* __restrict__ g2,
                                                   const double* __restrict__ e, z64* __restrict__ t,
                                                   const int* __restrict__ d_m) {
  extern __shared__ char smem_raw[];
  z64* grid = reinterpret_cast<z64*>(smem_raw);
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  if (d_m && b >= *d_m) return;   // a band the device switched off (whole workgroup)
  const int np = p1 * n2 * n3;
  const z64* xb = x + b * ld;
  for (int i = tid; i < np; i += KL_NT) grid[i] = mk(0.0, 0.0);
  __syncthreads();
  for (int ig = tid; ig < npw; ig += KL_NT) {
    const int r = nl[ig];
    const int i1 = r % n1;
    grid[i1 + p1 * (r / n1)] = xb[ig];
  }
  __syncthreads();
  // inverse along a1 and a2
  for (int l = tid; l < n2 * n3; l += KL_NT) kl_pass<+1>(n1, grid + p1 * l, 1);
  __syncthreads();
  for (int l = tid; l < n1 * n3; l += KL_NT) {
    const int i1 = l % n1, i3 = l / n1;
    kl_pass<+1>(n2, grid + kl_at(i1, 0, i3, p1, n2), p1);
  }
  __syncthreads();
  // a3: inverse, V(r), forward
  for (int l = tid; l < n1 * n2; l += KL_NT) {
    const int i1 = l % n1, i2 = l / n1;
    kl_pass_vloc(n3, grid + kl_at(i1, i2, 0, p1, n2), p1 * n2, vrs + i1 + static_cast<long>(n1) * i2,
                 static_cast<long>(n1) * n2);
  }
  __syncthreads();
  // forward along a2 and a1
  for (int l = tid; l < n1 * n3; l += KL_NT) {
    const int i1 = l % n1, i3 = l / n1;
    kl_pass<-1>(n2, grid + kl_at(i1, 0, i3, p1, n2), p1);
  }
  __syncthreads();
  for (int l = tid; l < n2 * n3; l += KL_NT) kl_pass<-1>(n1, grid + p1 * l, 1);
  __syncthreads();
  const double sc = 1.0 / (static_cast<double>(n1) * n2 * n3);
  const double eb = e[b];
  z64* tb = t + b * ld;
  for (int ig = tid; ig < ld; ig += KL_NT) {
    if (ig >= npw) {
      tb[ig] = mk(0.0, 0.0);
      continue;
    }
    const int r = nl[ig];
    const z64 w = grid[r % n1 + p1 * (r / n1)];
    tb[ig] = cscale(w, sc) + cscale(xb[ig], g2[ig] - eb);
  }
}

constexpr size_t kKlMaxSmem = 160u * 1024u;

size_t kl_smem(int n1, int n2, int n3) { return sizeof(z64) * static_cast<size_t>(n1 | 1) * n2 * n3; }

}  // namespace

bool kvloc_lds_available(int n1, int n2, int n3) {
  return kl_size(n1) && kl_size(n2) && kl_size(n3) && kl_smem(n1, n2, n3) <= kKlMaxSmem;
}

void kvloc_lds(int n1, int n2, int n3, int m, int npw, long ld, const int* nl, const z64* x, const double* vrs,
               const double* g2, const double* e, z64* t, hipStream_t s, const int* d_m) {
  WH_REQUIRE(kvloc_lds_available(n1, n2, n3), "kvloc_lds: grid not supported");
  if (m <= 0) return;
  static const bool attr_set = [] {   // once per process (thread-safe static initialisation)
    WH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_kvloc_lds), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(kKlMaxSmem)));
    return true;
  }();
  (void)attr_set;
  k_kvloc_lds<<<m, KL_NT, kl_smem(n1, n2, n3), s>>>(n1, n2, n3, n1 | 1, npw, ld, nl, x, vrs, g2, e, t, d_m);
  WH_HIP(hipGetLastError());
}

}  // namespace wh
