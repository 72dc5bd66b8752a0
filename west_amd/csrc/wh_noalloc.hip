// The Sternheimer operator for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): every workspace of the operator -- the stick-order copies of
// psi and hpsi, the FFT item lists and stick buffers, the projection
// coefficients, the bra product's split-K partial sums and the projector row
// scales -- is carved from one file-scope device array of fixed size, and
// the products run the variants an earlier timed call chose (or fixed ones),
// so nothing is timed either.
//
// The array is per GPU (each device holds its own copy until the process
// ends); its device address is taken once per device by noalloc_init (called
// from wh_init).  Calls on one GPU share it: calls that overlap in time on one
// GPU (two host threads, two systems) race and must be serialised by the
// caller.  Every buffer taken from it is written before it is read in each
// call, so no state carries over from the previous call.
#include <algorithm>
#include <array>
#include <cstdio>
#include <mutex>

#include "../../include/west_hip.h"
#include "wh_context.h"
#include "wh_fft.h"
#include "wh_gemm.h"
#include "wh_vec.h"

namespace wh {

constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static std::array<unsigned char*, kMaxDevices> g_scratch_ptr{};

// Device address of this device's scratch array (the host-side symbol is a
// stand-in, not a device address).  Called with the device current.
void noalloc_init(int device) {
  WH_REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  WH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  WH_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_ptr[device] = static_cast<unsigned char*>(p);
}

static unsigned char* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  WH_REQUIRE(device >= 0 && device < kMaxDevices && g_scratch_ptr[device],
             "the scratch array of this device has no address (wh_init not called)");
  return g_scratch_ptr[device];
}

// projector row scales: alpha for the evc rows, D (from the registered
// scales) for the vkb rows
__global__ void k_qscale(int nq, int nbndval, double alpha, const double* __restrict__ reg, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq) out[i] = i < nbndval ? alpha : reg[i];
}

static bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// byte layout of one call inside the scratch array
struct Carve {
  size_t off = 0;
  template <typename T>
  size_t take(size_t count) {
    const size_t at = off;
    off += (sizeof(T) * std::max<size_t>(count, 1) + 255) / 256 * 256;
    return at;
  }
};

}  // namespace wh

using namespace wh;

extern "C" int wh_apply_sternheimerop_to_m_wfcs_noalloc(int nbndval, const double* psi, double* hpsi,
                                                         const double* e, const double* alpha, int m) {
  try {
    Context& c = ctx();
    require_ready(true);
    WH_REQUIRE(nbndval == c.nbndval, "nbndval differs from the registered states");
    WH_REQUIRE(m >= 0, "m must be >= 0");
    WH_REQUIRE(c.precision == 0, "the allocation-free operator is FP64 only (precision mode 0)");
    if (m == 0) return 0;
    WH_REQUIRE(is_device_ptr(psi) && is_device_ptr(hpsi) && is_device_ptr(e),
               "psi, hpsi and e must be device arrays (host arrays would need staging buffers)");
    hipStream_t s = c.stream;
    double a[2] = {0.0, 0.0};
    WH_HIP(hipMemcpy(a, alpha, sizeof(double) * 2, hipMemcpyDefault));
    const int npw = c.npw;
    const int npair = (m + 1) / 2;
    const int nq = c.l_kinetic_only ? c.nbndval : c.nbndval + c.nkb;
    const long sticks = c.plan.sticks_elems();
    // fixed part of the layout
    Carve cv;
    const size_t o_x = cv.take<z64>(static_cast<size_t>(m) * npw);
    const size_t o_t = cv.take<z64>(static_cast<size_t>(m) * npw);
    const size_t o_items = cv.take<FftItem>(npair);
    const size_t o_qs = cv.take<double>(nq);
    const size_t o_cm = cv.take<double>(static_cast<size_t>(nq) * m);
    int bra_variant = 0;
    const long np = nq > 0 ? bra_partials_doubles(nq, m, 2 * npw, nq, &bra_variant) : 0;
    const size_t o_p = cv.take<double>(static_cast<size_t>(np));
    // FFT items per chunk: what fits, below the plane kernel's timing threshold (256 items)
    const size_t item_bytes = 2 * ((sizeof(z64) * sticks + 255) / 256 * 256) + 256;
    const size_t left = cv.off < kScratchBytes ? kScratchBytes - cv.off : 0;
    const int chunk = c.l_kinetic_only ? 1 : static_cast<int>(std::min<size_t>({left / item_bytes, 255,
                                                                               static_cast<size_t>(npair)}));
    if (chunk < 1 || cv.off > kScratchBytes) {
      char msg[256];
      std::snprintf(msg, sizeof msg,
                    "m = %d bands need more than the %zu-byte scratch of the allocation-free operator "
                    "(%zu bytes before the FFT buffers); use wh_apply_sternheimerop_to_m_wfcs or fewer bands",
                    m, kScratchBytes, cv.off);
      throw Error{-1, msg};
    }
    const size_t o_id = cv.take<FftItem>(chunk);
    const size_t o_sa = cv.take<z64>(static_cast<size_t>(chunk) * sticks);
    const size_t o_sb = cv.take<z64>(static_cast<size_t>(chunk) * sticks);
    WH_REQUIRE(cv.off <= kScratchBytes, "scratch layout exceeds the array");
    unsigned char* base = scratch_ptr(c.device);
    z64* x = reinterpret_cast<z64*>(base + o_x);
    z64* t = reinterpret_cast<z64*>(base + o_t);
    double* qs = reinterpret_cast<double*>(base + o_qs);
    SternWs ws{reinterpret_cast<FftItem*>(base + o_items), reinterpret_cast<FftItem*>(base + o_id),
               reinterpret_cast<z64*>(base + o_sa), reinterpret_cast<z64*>(base + o_sb), chunk,
               reinterpret_cast<double*>(base + o_cm), reinterpret_cast<double*>(base + o_p), bra_variant, qs};
    // the profiler's event pool grows on use: off for this call
    const bool prof = c.prof.on;
    c.prof.on = false;
    try {
      if (nq > 0) {
        k_qscale<<<ceil_div(nq, 256), 256, 0, s>>>(nq, c.nbndval, a[0], c.d_qscale, qs);
        WH_HIP(hipGetLastError());
      }
      gather_sticks(m, npw, c.d_perm, reinterpret_cast<const z64*>(psi), c.npwx, x, npw, nullptr, s);
      stern_apply(x, t, e, m, a[0], &ws);
      ++c.nstern_calls;
      scatter_caller(m, npw, c.npwx, c.d_perm, t, npw, reinterpret_cast<z64*>(hpsi), c.npwx, s);
      WH_HIP(hipStreamSynchronize(s));
    } catch (...) {
      c.prof.on = prof;
      throw;
    }
    c.prof.on = prof;
    return 0;
  } catch (const Error& err) {
    set_error(err.msg);
    return err.code < 0 ? err.code : -1;
  } catch (const std::exception& err) {
    set_error(err.what());
    return -3;
  }
}
