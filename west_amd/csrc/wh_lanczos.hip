// Wfreq's deflated Lanczos with full re-orthogonalisation at Gamma
// (Wfreq/solve_deflated_lanczos_w_full_ortho.f90:14-357, gamma_only branches)
// and the Gamma inner-product matrix of get_brak_hyper_parallel
// (Wfreq/get_brak_hyper_parallel.f90:14-143, Hamiltonian_kernel/glbrak.f90).
//
// All NRHS chains run together: each Lanczos step is one application of H to
// the NRHS current vectors (the Sternheimer operator's FFT and projector
// kernels with e = 0 and alpha = 0), one deflation P_c = 1 - sum_v |psi_v><psi_v|
// as a bra and a ket product, then per-column kernels for the three-term
// recurrence, the modified Gram-Schmidt sweep over the earlier vectors and
// the norm.  alpha, beta and the basis stay on the device until the end.
#include <algorithm>

#include "../../include/west_hip.h"
#include "wh_context.h"
#include "wh_gemm.h"
#include "wh_vec.h"

namespace wh {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ double block_sum_l(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < NT / 64; ++k) t += red[k];
  __syncthreads();
  return t;
}

// Gamma inner product of columns: 2 Re sum a* b - Re a(0) Re b(0) (gstart == 2)
__device__ __forceinline__ double gamma_dot(const z64* a, const z64* b, int npw, int gstart, double* red) {
  double s = 0.0;
  for (int g = threadIdx.x; g < npw; g += NT) s += a[g].x * b[g].x + a[g].y * b[g].y;
  s = block_sum_l(s, red);
  return 2.0 * s - (gstart == 2 ? a[0].x * b[0].x : 0.0);
}

// beta[ip] = sqrt(<r|r>); also into out[ip] when given (solve...ortho.f90:76-112, 290-339)
__global__ __launch_bounds__(NT) void k_lan_norm(int npw, int gstart, const z64* r, long ld, double* beta,
                                                 double* out) {
  __shared__ double red[NT / 64];
  const int ip = blockIdx.x;
  const z64* ri = r + static_cast<long>(ip) * ld;
  const double b = sqrt(gamma_dot(ri, ri, npw, gstart, red));
  if (threadIdx.x == 0) {
    beta[ip] = b;
    if (out) out[ip] = b;
  }
}

// q = r / beta  (solve...ortho.f90:118-124)
__global__ void k_lan_scale(int npw, const z64* r, const double* beta, z64* q, long ld) {
  const int ip = blockIdx.y;
  const int g = blockIdx.x * NT + threadIdx.x;
  if (g >= npw) return;
  const double inv = 1.0 / beta[ip];
  const z64 v = r[static_cast<long>(ip) * ld + g];
  q[static_cast<long>(ip) * ld + g] = mk<double>(v.x * inv, v.y * inv);
}

// r -= beta_prev q_prev (il > 1);  alpha = <q|r>;  r -= alpha q   (solve...ortho.f90:145-221)
__global__ __launch_bounds__(NT) void k_lan_alpha(int npw, int gstart, z64* r, const z64* q, const z64* qprev,
                                                  const double* beta_prev, double* alpha_out, long ld) {
  __shared__ double red[NT / 64];
  const int ip = blockIdx.x;
  const long o = static_cast<long>(ip) * ld;
  z64* ri = r + o;
  const z64* qi = q + o;
  double s = 0.0;
  if (qprev) {
    const double bp = beta_prev[ip];
    const z64* pi = qprev + o;
    for (int g = threadIdx.x; g < npw; g += NT) {
      z64 v = ri[g];
      v.x -= pi[g].x * bp;
      v.y -= pi[g].y * bp;
      ri[g] = v;
      s += qi[g].x * v.x + qi[g].y * v.y;
    }
  } else {
    for (int g = threadIdx.x; g < npw; g += NT) s += qi[g].x * ri[g].x + qi[g].y * ri[g].y;
  }
  s = block_sum_l(s, red);
  // element 0 was written by thread 0 itself
  const double a = 2.0 * s - (gstart == 2 ? qi[0].x * ri[0].x : 0.0);
  __syncthreads();
  if (threadIdx.x == 0) alpha_out[ip] = a;
  for (int g = threadIdx.x; g < npw; g += NT) {
    z64 v = ri[g];
    v.x -= qi[g].x * a;
    v.y -= qi[g].y * a;
    ri[g] = v;
  }
}

// one modified Gram-Schmidt step: t = <q_a|r>; r -= t q_a   (solve...ortho.f90:225-287)
__global__ __launch_bounds__(NT) void k_lan_mgs(int npw, int gstart, z64* r, const z64* qa, long ld) {
  __shared__ double red[NT / 64];
  const int ip = blockIdx.x;
  const long o = static_cast<long>(ip) * ld;
  z64* ri = r + o;
  const z64* ai = qa + o;
  const double t = gamma_dot(ai, ri, npw, gstart, red);
  __syncthreads();
  for (int g = threadIdx.x; g < npw; g += NT) {
    z64 v = ri[g];
    v.x -= ai[g].x * t;
    v.y -= ai[g].y * t;
    ri[g] = v;
  }
}

}  // namespace

}  // namespace wh

using namespace wh;

extern "C" int wh_solve_deflated_lanczos_w_full_ortho(int nbnd_to_deflate, int nrhs, int nlsteps, const double* b,
                                                       double* alpha_s, double* beta_s, double* q_s,
                                                       double* bnorm) {
  try {
    Context& c = ctx();
    require_ready(true);
    WH_REQUIRE(nrhs >= 0 && nlsteps >= 1, "nrhs must be >= 0 and nlsteps >= 1");
    WH_REQUIRE(nbnd_to_deflate >= 0 && nbnd_to_deflate <= c.nbndval,
               "nbnd_to_deflate must not exceed the registered states (nbndval)");
    WH_REQUIRE(!c.l_kinetic_only, "the Lanczos chains need the full Hamiltonian (l_kinetic_only is set)");
    WH_REQUIRE(c.precision == 0, "the Lanczos chains run in FP64 (precision mode 0)");
    if (nrhs == 0) return 0;
    hipStream_t s = c.stream;
    const int npw = c.npw;
    const long ld = npw;
    const long k2 = 2L * npw;
    const size_t blk = static_cast<size_t>(nrhs) * npw;
    // basis (nlsteps blocks of nrhs columns, stick order), residual, scalars
    z64* Q = c.lz_q.as<z64>(static_cast<size_t>(nlsteps) * blk);
    z64* R = c.lz_r.as<z64>(blk);
    double* sc = c.lz_sc.as<double>(static_cast<size_t>(2 * nlsteps + 3) * nrhs);
    double* d_alpha = sc;                                   // nlsteps x nrhs
    double* d_beta_s = sc + static_cast<size_t>(nlsteps) * nrhs;   // nlsteps x nrhs (last unused)
    double* d_beta = d_beta_s + static_cast<size_t>(nlsteps) * nrhs;
    double* d_bnorm = d_beta + nrhs;
    double* d_e0 = d_bnorm + nrhs;
    WH_HIP(hipMemsetAsync(d_e0, 0, sizeof(double) * nrhs, s));
    {
      const size_t cin = static_cast<size_t>(nrhs) * c.npwx;
      const z64* bin = nullptr;
      if (on_device_mem(b)) {
        wait_for_caller(c);
        bin = reinterpret_cast<const z64*>(b);
      } else {
        z64* t = c.w_in.as<z64>(cin);
        WH_HIP(hipMemcpyAsync(t, b, sizeof(z64) * cin, hipMemcpyHostToDevice, s));
        bin = t;
      }
      gather_sticks(nrhs, npw, c.d_perm, bin, c.npwx, R, ld, nullptr, s);
    }
    const double* q = reinterpret_cast<const double*>(c.d_q);
    double* cm = c.w_cmat.as<double>(static_cast<size_t>(std::max(nbnd_to_deflate, 1)) * nrhs);
    // r -> q & beta (76-112)
    k_lan_norm<<<nrhs, NT, 0, s>>>(npw, c.gstart, R, ld, d_beta, d_bnorm);
    WH_HIP(hipGetLastError());
    for (int il = 0; il < nlsteps; ++il) {
      z64* qi = Q + static_cast<size_t>(il) * blk;
      k_lan_scale<<<dim3(ceil_div(npw, NT), nrhs), NT, 0, s>>>(npw, R, d_beta, qi, ld);
      WH_HIP(hipGetLastError());
      // r = H q  (h_psi_: the operator with e = 0, alpha = 0), then r = P_c r (144-145)
      stern_apply(qi, R, d_e0, nrhs, 0.0);
      if (nbnd_to_deflate > 0) {
        glbrak_q(0, nbnd_to_deflate, R, nrhs, cm);
        WH_TIMED("gemm_ket", dgemm(false, static_cast<int>(k2), nrhs, nbnd_to_deflate, -1.0, q, k2, cm,
                                   nbnd_to_deflate, 1.0, reinterpret_cast<double*>(R), k2, s));
      }
      k_lan_alpha<<<nrhs, NT, 0, s>>>(npw, c.gstart, R, qi, il > 0 ? qi - blk : nullptr, d_beta,
                                     d_alpha + static_cast<size_t>(il) * nrhs, ld);
      WH_HIP(hipGetLastError());
      // full re-orthogonalisation against q_1 .. q_{il-2} (225-287), one vector at a time
      for (int ia = 0; ia + 1 < il; ++ia) {
        k_lan_mgs<<<nrhs, NT, 0, s>>>(npw, c.gstart, R, Q + static_cast<size_t>(ia) * blk, ld);
        WH_HIP(hipGetLastError());
      }
      k_lan_norm<<<nrhs, NT, 0, s>>>(npw, c.gstart, R, ld, d_beta, d_beta_s + static_cast<size_t>(il) * nrhs);
      WH_HIP(hipGetLastError());
    }
    // outputs in the reference's layouts: alpha_s(NLSTEPS, NRHS), beta_s(NLSTEPS-1, NRHS),
    // q_s(npwx, NRHS, NLSTEPS), bnorm(NRHS)
    // [alpha | beta_s | beta | bnorm], nrhs each row
    std::vector<double> h(static_cast<size_t>(2 * nlsteps + 2) * nrhs);
    WH_HIP(hipMemcpyAsync(h.data(), sc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    if (q_s) {
      const size_t cin = static_cast<size_t>(nrhs) * c.npwx;
      const bool dout = on_device_mem(q_s);
      for (int il = 0; il < nlsteps; ++il) {
        z64* dst = reinterpret_cast<z64*>(q_s) + static_cast<size_t>(il) * cin;
        z64* o = dout ? dst : c.w_out.as<z64>(cin);
        scatter_caller(nrhs, npw, c.npwx, c.d_perm, Q + static_cast<size_t>(il) * blk, ld, o, c.npwx, s);
        if (!dout) {
          WH_HIP(hipMemcpyAsync(dst, o, sizeof(z64) * cin, hipMemcpyDeviceToHost, s));
          WH_HIP(hipStreamSynchronize(s));
        }
      }
    }
    WH_HIP(hipStreamSynchronize(s));
    for (int ip = 0; ip < nrhs; ++ip) {
      for (int il = 0; il < nlsteps; ++il)
        alpha_s[static_cast<size_t>(ip) * nlsteps + il] = h[static_cast<size_t>(il) * nrhs + ip];
      if (beta_s)
        for (int il = 0; il + 1 < nlsteps; ++il)
          beta_s[static_cast<size_t>(ip) * (nlsteps - 1) + il] =
              h[static_cast<size_t>(nlsteps + il) * nrhs + ip];
      if (bnorm) bnorm[ip] = h[static_cast<size_t>(2 * nlsteps + 1) * nrhs + ip];
    }
    if (c.prof.on) c.prof.flush();
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code < 0 ? e.code : -1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -3;
  }
}

// c(na, nb) = 2 Re <a_i|b_j> - Re a_i(G=0) Re b_j(G=0) for a(npwx, na), b(npwx, nb)
extern "C" int wh_glbrak_gamma(int na, const double* a, int nb, const double* b, double* c_out) {
  try {
    Context& c = ctx();
    require_ready(false);
    WH_REQUIRE(na >= 0 && nb >= 0, "na and nb must be >= 0");
    if (na == 0 || nb == 0) return 0;
    hipStream_t s = c.stream;
    const int npw = c.npw;
    const long ld = npw;
    auto stage = [&](const double* p, int n, DevBuf& tmp, DevBuf& st) {
      const size_t cin = static_cast<size_t>(n) * c.npwx;
      const z64* in = nullptr;
      if (on_device_mem(p)) {
        wait_for_caller(c);
        in = reinterpret_cast<const z64*>(p);
      } else {
        z64* t = tmp.as<z64>(cin);
        WH_HIP(hipMemcpyAsync(t, p, sizeof(z64) * cin, hipMemcpyHostToDevice, s));
        in = t;
      }
      z64* o = st.as<z64>(static_cast<size_t>(n) * npw);
      gather_sticks(n, npw, c.d_perm, in, c.npwx, o, ld, nullptr, s);
      return o;
    };
    const z64* A = stage(a, na, c.w_in, c.lz_q);
    const z64* B = stage(b, nb, c.w_out, c.lz_r);
    double* d = c.w_cmat.as<double>(static_cast<size_t>(na) * nb);
    WH_TIMED("gemm_bra", dgemm_bra(na, nb, static_cast<int>(2 * npw), 2.0, reinterpret_cast<const double*>(A), 2L * npw,
                                   reinterpret_cast<const double*>(B), 2L * npw, c.gstart == 2 ? 1.0 : 0.0, nullptr,
                                   d, na, s));
    WH_HIP(hipMemcpyAsync(c_out, d, sizeof(double) * na * nb, hipMemcpyDefault, s));
    WH_HIP(hipStreamSynchronize(s));
    return 0;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code < 0 ? e.code : -1;
  } catch (const std::exception& e) {
    set_error(e.what());
    return -3;
  }
}
