// Common definitions for the MI355X Wstat library (gfx950, HIP).
#pragma once

#include <cstdlib>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace wh {

// ---------------------------------------------------------------------------
// error handling: every entry point returns 0 / >0 (warning count) / <0 and
// records a message retrievable through wh_last_error().
// ---------------------------------------------------------------------------
struct Error {
  int code;
  std::string msg;
};

void set_error(const std::string& msg);
const char* last_error();

#define WH_HIP(expr)                                                                \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      throw ::wh::Error{-2, std::string(#expr) + ": " + hipGetErrorString(e_)};     \
    }                                                                               \
  } while (0)

#define WH_REQUIRE(cond, text)                                                      \
  do {                                                                              \
    if (!(cond)) throw ::wh::Error{-1, std::string(text)};                          \
  } while (0)

// ---------------------------------------------------------------------------
// complex arithmetic on plain structs (interleaved re/im, like Fortran COMPLEX)
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return mk<T>(a.x * s, a.y * s);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conjg(cplx<T> a) {
  return mk<T>(a.x, -a.y);
}
// multiply by +i (dir = +1) or -i (dir = -1)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> mul_i(cplx<T> a, int dir) {
  return dir > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

using z64 = cplx<double>;
using z32 = cplx<float>;

// WH_TUNE_LOG=1: print the timings behind every timed choice to stderr
inline bool tune_log() {
  static const bool on = [] {
    const char* e = std::getenv("WH_TUNE_LOG");
    return e != nullptr && e[0] == '1';
  }();
  return on;
}

inline int ceil_div(long a, long b) { return static_cast<int>((a + b - 1) / b); }

// Start vectors of the Davidson drivers: SplitMix64's finaliser, and the two
// uniforms in [0, 1) of basis vector ivec (1-based) at plane wave ig (1-based
// counter): u_k = mix64(seed ^ (ivec << 32) ^ (2 ig + k)) / 2^64, 53 bits.
constexpr unsigned long long kRngSeed = 0x5EED0F1E57ULL;
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
__device__ __forceinline__ void start_uniforms(unsigned long long ivec, unsigned long long ig, double& u0,
                                               double& u1) {
  const unsigned long long base = kRngSeed ^ (ivec << 32);
  u0 = static_cast<double>(mix64(base ^ (2ULL * ig)) >> 11) * (1.0 / 9007199254740992.0);
  u1 = static_cast<double>(mix64(base ^ (2ULL * ig + 1ULL)) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace wh
