// afft_device.h -- device-side building blocks of the gfx950 FFT kernels.
//
// Complex values are (x = re, y = im) pairs.  Every transform computed here is
// the forward DFT, X[k] = sum_j x[j] exp(-2 pi i jk/L); the inverse is obtained
// by swapping re and im on load and on store, which is the identity afft uses
// for its inverse plans (common.hpp:355-378, fft_internal.hpp:166-198):
//   IDFT(x) = swap(DFT(swap(x))).
//
// Radix-R butterflies (R = 2..32) are written as a compile-time Cooley-Tukey
// split R = P*Q with the P-point DFTs first (stride-Q inputs), the W_R^(n2*k1)
// twiddles between, and the Q-point DFTs last.  Trivial twiddles (powers of
// -i and the (+-1 +-i)/sqrt2 family) are specialised at compile time so no
// multiplications by exact 0 or 1 are emitted.  afft's own radix-4 DIT
// butterflies are onedim.hpp:55-133; the algebra is the textbook one and the
// register layout here is independent of theirs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace afgpu {

typedef long long idx_t;

// Every compile-time-loop body must inline, or its register arrays end up in
// scratch memory when many tiles share one kernel.
#define AFGPU_INLINE __attribute__((always_inline))

template <typename T>
struct cx {
  T x, y;
};

// Two-element vector types, for 8 / 16-byte accesses through the
// non-temporal builtins (which take scalar or vector element types only).
template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct Vec2<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> swapri(cx<T> a) { return {a.y, a.x}; }
template <typename T>
__device__ __forceinline__ cx<T> conjg(cx<T> a) { return {a.x, -a.y}; }
// a * (-i)
template <typename T>
__device__ __forceinline__ cx<T> mul_neg_i(cx<T> a) { return {a.y, -a.x}; }

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> w) {
  return {fma(a.x, w.x, -(a.y * w.y)), fma(a.x, w.y, a.y * w.x)};
}
// a * conj(w)
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> w) {
  return {fma(a.x, w.x, a.y * w.y), fma(a.y, w.x, -(a.x * w.y))};
}

// ---- compile-time loops ---------------------------------------------------
template <typename F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// cos(2 pi q / 64) for q = 0..16 (first quadrant), correctly rounded doubles.
__device__ constexpr double kCos64[17] = {
    1.0,
    0.9951847266721968862448,
    0.9807852804032304491262,
    0.9569403357322088649358,
    0.9238795325112867561282,
    0.8819212643483550297128,
    0.8314696123025452370788,
    0.7730104533627369608109,
    0.7071067811865475244008,
    0.6343932841636454982152,
    0.5555702330196022247428,
    0.4713967368259976485564,
    0.3826834323650897717285,
    0.2902846772544623676362,
    0.1950903220161282678483,
    0.0980171403295606019942,
    0.0};

__host__ __device__ constexpr double cos64(int q) {
  q &= 63;
  return q <= 16 ? kCos64[q] : q <= 32 ? -kCos64[32 - q] : q <= 48 ? -kCos64[q - 32] : kCos64[64 - q];
}
__host__ __device__ constexpr double sin64(int q) { return cos64(16 - q); }

// a * W_R^M with W_R = exp(-2 pi i / R), R a power of two <= 64.
template <typename T, int R, int M>
__device__ __forceinline__ cx<T> twc(cx<T> a) {
  constexpr int m = ((M % R) + R) % R;
  constexpr int q = m * (64 / R);
  constexpr T s = T(0.7071067811865475244008);
  if constexpr (q == 0) {
    return a;
  } else if constexpr (q == 16) {
    return {a.y, -a.x};
  } else if constexpr (q == 32) {
    return {-a.x, -a.y};
  } else if constexpr (q == 48) {
    return {-a.y, a.x};
  } else if constexpr (q == 8) {
    return {(a.x + a.y) * s, (a.y - a.x) * s};
  } else if constexpr (q == 24) {
    return {(a.y - a.x) * s, -(a.x + a.y) * s};
  } else if constexpr (q == 40) {
    return {-(a.x + a.y) * s, (a.x - a.y) * s};
  } else if constexpr (q == 56) {
    return {(a.x - a.y) * s, (a.x + a.y) * s};
  } else {
    constexpr T c = T(cos64(q));
    constexpr T sn = T(-sin64(q));
    return cmul(a, cx<T>{c, sn});
  }
}

// W_N^m = exp(-2 pi i m / N) in FP32 for N = 2^lg (3 <= lg <= 27, any 32-bit m)
// without a table.  The index is reduced exactly in integers: quadrant q =
// m / (N/4), position r in the quadrant, folded to u = min(r, N/4 - r) <= N/8
// so that the angle x = 2 pi u / N lies in [0, pi/4].  sin x and cos x come
// from their Taylor series through x^9 and x^10 (the first omitted terms are
// below 2e-9 there); the fold swaps them, and the quadrant rotates the pair
// by q quarter turns.  Over every m of N = 2^20 this is within 9.3e-8 of the
// exact value (rms 2.7e-8), as accurate as the two-level table product; the
// same holds up to N = 2^27 (u <= 2^24 stays exact in a float).
__device__ __forceinline__ cx<float> tw_pow2f(unsigned m, int lg) {
  const unsigned q = m >> (lg - 2);
  const unsigned r = m & ((1u << (lg - 2)) - 1u);
  const unsigned e8 = 1u << (lg - 3);
  const bool fold = r > e8;
  const unsigned u = fold ? 2u * e8 - r : r;
  const float step = 6.28318530717958647692f * __int_as_float((127 - lg) << 23);  // 2 pi / N
  const float x = (float)u * step;
  const float x2 = x * x;
  const float s =
      fmaf(x * x2, fmaf(x2, fmaf(x2, fmaf(x2, 1.0f / 362880.0f, -1.0f / 5040.0f), 1.0f / 120.0f), -1.0f / 6.0f), x);
  const float c = fmaf(
      x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, -1.0f / 3628800.0f, 1.0f / 40320.0f), -1.0f / 720.0f), 1.0f / 24.0f),
               -0.5f),
      1.0f);
  const float cq = fold ? s : c, sq = fold ? c : s;  // cos and sin of the angle within the quadrant
  float a = (q & 1u) ? -sq : cq, b = (q & 1u) ? cq : sq;
  if (q & 2u) {
    a = -a;
    b = -b;
  }
  return {a, -b};
}

// W_2h^k (k < h, h a power of two) of a real pass: computed in FP32 while
// 2h <= 2^27, read from the table (W_2h^k at entry k) otherwise.
template <typename T>
__device__ __forceinline__ cx<T> real_tw_pow2(const cx<T>* __restrict__ tab, idx_t k, idx_t h) {
  if constexpr (std::is_same_v<T, float>) {
    const int lg = __builtin_ctzll((unsigned long long)h) + 1;
    if (lg >= 3 && lg <= 27) return tw_pow2f((unsigned)k, lg);
  }
  return tab[k];
}

// In-place, natural-order forward DFT of R points held in registers.
template <typename T, int R>
struct Dft {
  static __device__ __forceinline__ void run(cx<T>* v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
      cx<T> a = v[0], b = v[1];
      v[0] = a + b;
      v[1] = a - b;
    } else if constexpr (R == 4) {
      cx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
      cx<T> t2 = v[1] + v[3], t3 = mul_neg_i(v[1] - v[3]);
      v[0] = t0 + t2;
      v[2] = t0 - t2;
      v[1] = t1 + t3;
      v[3] = t1 - t3;
    } else {
      constexpr int P = R >= 16 ? 4 : 2;
      constexpr int Q = R / P;
      static_for<Q>([&](auto n2c) AFGPU_INLINE {
        constexpr int n2 = decltype(n2c)::value;
        cx<T> a[P];
        static_for<P>([&](auto n1c) AFGPU_INLINE { a[decltype(n1c)::value] = v[Q * decltype(n1c)::value + n2]; });
        Dft<T, P>::run(a);
        static_for<P>([&](auto k1c) AFGPU_INLINE {
          constexpr int k1 = decltype(k1c)::value;
          v[Q * k1 + n2] = twc<T, R, n2 * k1>(a[k1]);
        });
      });
      cx<T> out[R];
      static_for<P>([&](auto k1c) AFGPU_INLINE {
        constexpr int k1 = decltype(k1c)::value;
        cx<T> b[Q];
        static_for<Q>([&](auto n2c) AFGPU_INLINE { b[decltype(n2c)::value] = v[Q * k1 + decltype(n2c)::value]; });
        Dft<T, Q>::run(b);
        static_for<Q>([&](auto k2c) AFGPU_INLINE { out[k1 + P * decltype(k2c)::value] = b[decltype(k2c)::value]; });
      });
      static_for<R>([&](auto i) AFGPU_INLINE { v[decltype(i)::value] = out[decltype(i)::value]; });
    }
  }
};

// ---- per-length stage schedule ----------------------------------------------
// L = R0 * R1 * ... ; every thread holds E = max radix elements; each stage of
// radix R performs E/R butterflies per thread.  TPT = L / E threads per
// transform.  SV selects an alternative schedule for the same length (SV = 1:
// fewer, larger radices with more elements per thread -- one LDS exchange less,
// half the threads per transform).
template <int L, int SV = 0>
struct Sched;
#define AFGPU_SCHED(LEN, SVV, EE, NSTAGE, ...)                     \
  template <>                                                      \
  struct Sched<LEN, SVV> {                                         \
    static constexpr int E = EE;                                   \
    static constexpr int NST = NSTAGE;                             \
    static constexpr int R[NSTAGE] = {__VA_ARGS__};                \
  };
AFGPU_SCHED(1, 0, 1, 1, 1)
AFGPU_SCHED(2, 0, 2, 1, 2)
AFGPU_SCHED(4, 0, 4, 1, 4)
AFGPU_SCHED(8, 0, 8, 1, 8)
AFGPU_SCHED(16, 0, 16, 1, 16)
AFGPU_SCHED(32, 0, 8, 2, 8, 4)
AFGPU_SCHED(64, 0, 8, 2, 8, 8)
AFGPU_SCHED(128, 0, 16, 2, 16, 8)
AFGPU_SCHED(256, 0, 16, 2, 16, 16)
AFGPU_SCHED(512, 0, 8, 3, 8, 8, 8)
AFGPU_SCHED(1024, 0, 16, 3, 16, 16, 4)
AFGPU_SCHED(2048, 0, 16, 3, 16, 16, 8)
AFGPU_SCHED(4096, 0, 16, 3, 16, 16, 16)
AFGPU_SCHED(8192, 0, 16, 4, 16, 16, 16, 2)
AFGPU_SCHED(16384, 0, 16, 4, 16, 16, 16, 4)
AFGPU_SCHED(32768, 0, 32, 3, 32, 32, 32)
AFGPU_SCHED(512, 1, 32, 2, 32, 16)
AFGPU_SCHED(1024, 1, 32, 2, 32, 32)
AFGPU_SCHED(2048, 1, 32, 3, 32, 32, 2)
#undef AFGPU_SCHED

template <int L, int SV, int S>
struct StageNs {  // product of the radices before stage S
  static constexpr int value = StageNs<L, SV, S - 1>::value * Sched<L, SV>::R[S - 1];
};
template <int L, int SV>
struct StageNs<L, SV, 0> {
  static constexpr int value = 1;
};

// Offset (in complex entries) of stage S's twiddles in the per-length stage
// table.  Stage S (S >= 1) has Ns * (R - 1) entries laid out [k][r - 1] with
// value W_{Ns*R}^{r*k}.
template <int L, int SV, int S>
struct StageTwOff {
  static constexpr int value = StageTwOff<L, SV, S - 1>::value +
                               (S - 1 >= 1 ? StageNs<L, SV, S - 1>::value * (Sched<L, SV>::R[S - 1] - 1) : 0);
};
template <int L, int SV>
struct StageTwOff<L, SV, 0> {
  static constexpr int value = 0;
};

// LDS padding: one spare slot every 16 entries keeps the stride-R butterfly
// writes of the early stages off a single bank group.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
template <int L>
struct LdsLen {
  static constexpr int value = L + (L >> 4) + 1;
};

}  // namespace afgpu
