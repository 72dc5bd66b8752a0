// Device helpers shared by the compile-time-size FFT kernels: compile-time
// trigonometry, small DFT butterflies in registers and radix plans.
#pragma once

#include <utility>

#include "wh_common.h"

namespace wh {
namespace {

// ---------------------------------------------------------------------------
// compile-time trigonometry (Taylor series after reduction to [-pi/4, pi/4])
// ---------------------------------------------------------------------------
constexpr long double kPi = 3.141592653589793238462643383279502884L;

constexpr long double taylor_sin(long double x) {
  long double t = x, s = x;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / ((2 * k) * (2 * k + 1));
    s += t;
  }
  return s;
}
constexpr long double taylor_cos(long double x) {
  long double t = 1.0L, s = 1.0L;
  for (int k = 1; k < 14; ++k) {
    t *= -x * x / ((2 * k - 1) * (2 * k));
    s += t;
  }
  return s;
}
// cos and sin of 2 pi m / n for integers
constexpr long double cos2pi(int m, int n) {
  m %= n;
  if (m < 0) m += n;
  // reduce by octants: angle = 2 pi m / n
  const long double a = 2.0L * kPi * m / n;
  if (8 * m <= n) return taylor_cos(a);
  if (8 * m <= 3 * n) return taylor_sin(kPi / 2 - a);
  if (8 * m <= 5 * n) return -taylor_cos(kPi - a);
  if (8 * m <= 7 * n) return taylor_sin(a - 3 * kPi / 2);
  return taylor_cos(2 * kPi - a);
}
// sin(2 pi m / n) = cos(2 pi (4m - n) / (4n))
constexpr long double sin2pi(int m, int n) { return cos2pi(4 * m - n, 4 * n); }

// (long double is the 64-bit double on the device target: a few ulp of 1)
constexpr bool near(long double a, long double b) { return (a - b) < 5e-16L && (b - a) < 5e-16L; }
static_assert(near(cos2pi(1, 4), 0.0L) && near(cos2pi(1, 3), -0.5L) && near(cos2pi(1, 6), 0.5L), "cos");
static_assert(near(cos2pi(1, 8), 0.70710678118654752440L) && near(cos2pi(3, 8), -0.70710678118654752440L), "cos");
static_assert(near(cos2pi(5, 8), -0.70710678118654752440L) && near(cos2pi(7, 8), 0.70710678118654752440L), "cos");
static_assert(near(sin2pi(1, 3), 0.86602540378443864676L) && near(sin2pi(2, 3), -0.86602540378443864676L), "sin");
static_assert(near(sin2pi(5, 6), -0.86602540378443864676L) && near(sin2pi(1, 12), 0.5L), "sin");
static_assert(near(sin2pi(7, 12), -0.5L) && near(cos2pi(2, 5), -0.80901699437494742410L), "sin/cos");

// ---------------------------------------------------------------------------
// small DFTs in registers; dir = -1: exp(-i), +1: exp(+i)
// ---------------------------------------------------------------------------
template <typename T, int R, int DIR>
__device__ __forceinline__ void dft(cplx<T>* v);

template <typename T>
__device__ __forceinline__ void dft2(cplx<T>* v) {
  const cplx<T> a = v[0] + v[1];
  v[1] = v[0] - v[1];
  v[0] = a;
}
template <typename T, int DIR>
__device__ __forceinline__ void dft3(cplx<T>* v) {
  const cplx<T> t1 = v[1] + v[2];
  const cplx<T> t2 = v[1] - v[2];
  const cplx<T> m = v[0] + cscale(t1, T(-0.5));
  const cplx<T> u = mul_i(cscale(t2, T(0.86602540378443864676)), DIR);
  v[0] = v[0] + t1;
  v[1] = m + u;
  v[2] = m - u;
}
template <typename T, int DIR>
__device__ __forceinline__ void dft4(cplx<T>* v) {
  const cplx<T> t0 = v[0] + v[2];
  const cplx<T> t1 = v[0] - v[2];
  const cplx<T> t2 = v[1] + v[3];
  const cplx<T> t3 = v[1] - v[3];
  const cplx<T> u = mul_i(t3, DIR);
  v[0] = t0 + t2;
  v[2] = t0 - t2;
  v[1] = t1 + u;
  v[3] = t1 - u;
}
template <typename T, int DIR>
__device__ __forceinline__ void dft5(cplx<T>* v) {
  constexpr T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
  constexpr T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
  const cplx<T> t1 = v[1] + v[4];
  const cplx<T> t2 = v[2] + v[3];
  const cplx<T> t3 = v[1] - v[4];
  const cplx<T> t4 = v[2] - v[3];
  const cplx<T> m1 = v[0] + cscale(t1, c1) + cscale(t2, c2);
  const cplx<T> m2 = v[0] + cscale(t1, c2) + cscale(t2, c1);
  const cplx<T> u1 = mul_i(cscale(t3, s1) + cscale(t4, s2), DIR);
  const cplx<T> u2 = mul_i(cscale(t3, s2) - cscale(t4, s1), DIR);
  v[0] = v[0] + t1 + t2;
  v[1] = m1 + u1;
  v[4] = m1 - u1;
  v[2] = m2 + u2;
  v[3] = m2 - u2;
}

// two-level Cooley-Tukey: R = A*B, j = j1 + A j2, k = k2 + B k1
// y * (c + i s) for constant c, s (folded at compile time after unrolling):
// quarter turns are swaps, eighth turns two adds and two multiplies
template <typename T>
__device__ __forceinline__ cplx<T> twiddle_mul(cplx<T> y, T c, T s) {
  const T tol = T(1e-6);
  if (fabs(s) < tol) return c > T(0) ? y : mk<T>(-y.x, -y.y);
  if (fabs(c) < tol) return s > T(0) ? mk<T>(-y.y, y.x) : mk<T>(y.y, -y.x);
  if (fabs(c - s) < tol) return mk<T>(c * (y.x - y.y), c * (y.x + y.y));
  if (fabs(c + s) < tol) return mk<T>(c * (y.x + y.y), c * (y.y - y.x));
  return cmul(y, mk<T>(c, s));
}

template <typename T, int A, int B, int DIR>
__device__ __forceinline__ void dft_ct(cplx<T>* v) {
  constexpr int R = A * B;
  cplx<T> y[A][B];
#pragma unroll
  for (int j1 = 0; j1 < A; ++j1)
#pragma unroll
    for (int j2 = 0; j2 < B; ++j2) y[j1][j2] = v[j1 + A * j2];
#pragma unroll
  for (int j1 = 0; j1 < A; ++j1) dft<T, B, DIR>(y[j1]);
#pragma unroll
  for (int j1 = 1; j1 < A; ++j1)
#pragma unroll
    for (int k2 = 1; k2 < B; ++k2) {
      const T c = T(cos2pi(j1 * k2, R));
      const T s = T(sin2pi(j1 * k2, R));
      // exp(DIR * 2 pi i j1 k2 / R)
      y[j1][k2] = twiddle_mul(y[j1][k2], c, DIR > 0 ? s : -s);
    }
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) {
    cplx<T> col[A];
#pragma unroll
    for (int j1 = 0; j1 < A; ++j1) col[j1] = y[j1][k2];
    dft<T, A, DIR>(col);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = col[k1];
  }
}

template <typename T, int R, int DIR>
__device__ __forceinline__ void dft(cplx<T>* v) {
  if constexpr (R == 2) dft2(v);
  else if constexpr (R == 3) dft3<T, DIR>(v);
  else if constexpr (R == 4) dft4<T, DIR>(v);
  else if constexpr (R == 5) dft5<T, DIR>(v);
  else if constexpr (R == 6) dft_ct<T, 2, 3, DIR>(v);
  else if constexpr (R == 8) dft_ct<T, 2, 4, DIR>(v);
  else if constexpr (R == 9) dft_ct<T, 3, 3, DIR>(v);
  else if constexpr (R == 10) dft_ct<T, 2, 5, DIR>(v);
  else {
    static_assert(R == 12, "unsupported radix");
    dft_ct<T, 4, 3, DIR>(v);
  }
}

// ---------------------------------------------------------------------------
// compile-time radix plan
// ---------------------------------------------------------------------------
struct RadixPlan {
  int n;
  int r[8];
};

constexpr RadixPlan make_plan(int N) {
  RadixPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
  constexpr int pref[9] = {9, 8, 12, 10, 6, 5, 4, 3, 2};
  int m = N;
  while (m > 1 && p.n < 8) {
    int pick = 0;
    for (int i = 0; i < 9; ++i)
      if (m % pref[i] == 0) {
        pick = pref[i];
        break;
      }
    if (pick == 0) return RadixPlan{-1, {0, 0, 0, 0, 0, 0, 0, 0}};
    p.r[p.n++] = pick;
    m /= pick;
  }
  return p;
}

// REV: the radices in reverse order. A forward transform with the reversed
// plan can start from the output of an inverse transform's last pass without
// an LDS round trip: that pass leaves butterfly i holding positions
// i + j N/R, exactly what butterfly i of the reversed plan's first pass reads.
template <int N, bool REV = false>
struct Plan {
  static constexpr RadixPlan P = make_plan(N);
  static_assert(P.n > 0, "N must factor into 2, 3, 5");
  static constexpr int nst = P.n;
  static constexpr int radix(int s) { return REV ? P.r[P.n - 1 - s] : P.r[s]; }
  static constexpr int stride(int s) {   // product of the radices of passes before s
    int p = 1;
    for (int i = 0; i < s; ++i) p *= radix(i);
    return p;
  }
  // twiddles of pass s >= 1 are stored [k][j] (k < stride(s), 1 <= j < radix(s))
  // from twoff(s) (the reversed plan's tables follow the forward plan's) with
  // row length tws(s) = radix(s) - 1 rounded up to odd: the R - 1 roots one
  // butterfly needs are adjacent, and neighbouring butterflies read an odd
  // number of elements apart, so their reads fall on distinct LDS banks
  static constexpr int tws(int s) { return (radix(s) - 1) | 1; }
  static constexpr int twtotal() {
    int o = 0;
    for (int i = 1; i < nst; ++i) o += stride(i) * tws(i);
    return o;
  }
  static constexpr int twoff(int s) {
    int o = REV ? Plan<N, false>::twtotal() : 0;
    for (int i = 1; i < s; ++i) o += stride(i) * tws(i);
    return o;
  }
  static constexpr int twsize() {
    const int t = Plan<N, false>::twtotal() + Plan<N, true>::twtotal();
    return t > 0 ? t : 1;
  }
};

template <int... I, class F>
__device__ __forceinline__ void sfor_seq(std::integer_sequence<int, I...>, F&& f) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, i>) for i = 0 .. N-1, unrolled at compile time
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_seq(std::make_integer_sequence<int, N>{}, static_cast<F&&>(f));
}

// In-register transform of v[0..N): the Stockham passes of Plan<N, REV> with
// every index and twiddle exp(DIR 2 pi i j k / (P R)) a compile-time constant.
template <typename T, int N, int DIR, bool REV = false, int S = 0>
__device__ __forceinline__ void fft_line(cplx<T>* v) {
  using PL = Plan<N, REV>;
  if constexpr (S < PL::nst) {
    constexpr int R = PL::radix(S);
    constexpr int P = PL::stride(S);
    constexpr int NBL = N / R;
    cplx<T> w[N];
    sfor<NBL>([&](auto I) {
      constexpr int i = decltype(I)::value;
      constexpr int k = i % P;
      cplx<T> b[R];
      sfor<R>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr (k != 0 && j != 0) {
          constexpr T c = T(cos2pi(j * k, P * R));
          constexpr T s = T(DIR * sin2pi(j * k, P * R));
          b[j] = twiddle_mul(v[i + j * NBL], c, s);
        } else {
          b[j] = v[i + j * NBL];
        }
      });
      dft<T, R, DIR>(b);
      sfor<R>([&](auto J) {
        constexpr int j = decltype(J)::value;
        w[(i - k) * R + k + j * P] = b[j];
      });
    });
    sfor<N>([&](auto I) { v[decltype(I)::value] = w[decltype(I)::value]; });
    fft_line<T, N, DIR, REV, S + 1>(v);
  }
}

}  // namespace
}  // namespace wh
