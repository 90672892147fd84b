// Common device helpers for the gfx950 KDK Schroedinger-Poisson kernels:
// a plain complex type, compile-time DFT twiddles, FP64 phase range
// reduction and small reduction utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "qj_launch.h"

namespace qj {

// ---------------------------------------------------------------------------
// complex value (interleaved re/im, 8 B for float, 16 B for double)
// ---------------------------------------------------------------------------
template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T r, i;
};

template <class T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <class T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.r - b.r, a.i - b.i}; }
// Fused multiply-add written out, so that every kernel rounds a complex
// product the same way whatever the compiler's contraction choices in its
// context (the fused and split schedules of a sweep then agree bitwise).
template <class T>
__host__ __device__ __forceinline__ T fmad(T a, T b, T c) {
  if constexpr (std::is_same<T, float>::value)
    return __builtin_fmaf(a, b, c);
  else
    return __builtin_fma(a, b, c);
}
template <class T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {fmad(a.r, b.r, -(a.i * b.i)), fmad(a.r, b.i, a.i * b.r)};
}
template <class T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
  return {fmad(a.r, b.r, a.i * b.i), fmad(a.i, b.r, -(a.r * b.i))};
}
template <class T>
__device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return {a.r, -a.i}; }
template <class T>
__device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return {a.r * s, a.i * s}; }

// ---------------------------------------------------------------------------
// compile-time loop
// ---------------------------------------------------------------------------
template <int I, int E, class F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  if constexpr (I < E) {
    f(std::integral_constant<int, I>{});
    sfor<I + 1, E>(f);
  }
}

// cos(pi*k/32), k = 0..16 (the first quadrant of the 64-point circle).
constexpr double kCosQ[17] = {
    1.0, 0.9951847266721969, 0.9807852804032304, 0.9569403357322088,
    0.9238795325112867, 0.881921264348355, 0.8314696123025452, 0.773010453362737,
    0.7071067811865476, 0.6343932841636455, 0.5555702330196023, 0.4713967368259978,
    0.38268343236508984, 0.29028467725446233, 0.19509032201612833, 0.09801714032956077,
    0.0};

// cos(2*pi*k/64) for any integer k, from the quadrant table by symmetry.
constexpr double cos64(int k) {
  k = ((k % 64) + 64) % 64;
  if (k <= 16) return kCosQ[k];
  if (k <= 32) return -kCosQ[32 - k];
  if (k <= 48) return -kCosQ[k - 32];
  return kCosQ[64 - k];
}
constexpr double sin64(int k) { return cos64(k - 16); }

// v * exp(DIR * 2*pi*i*M/E), M and E compile-time (E divides 64).
// DIR = -1 is the forward (numpy) sign convention.
template <int M, int E, int DIR, class T>
__device__ __forceinline__ cpx<T> twc(cpx<T> v) {
  constexpr int m = ((M % E) + E) % E;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (2 * m == E) {
    return {-v.r, -v.i};
  } else if constexpr (4 * m == E) {  // exp(DIR*i*pi/2) = DIR*i
    if constexpr (DIR < 0) return {v.i, -v.r};
    else return {-v.i, v.r};
  } else if constexpr (4 * m == 3 * E) {  // exp(DIR*i*3pi/2) = -DIR*i
    if constexpr (DIR < 0) return {-v.i, v.r};
    else return {v.i, -v.r};
  } else {
    static_assert(64 % E == 0, "register DFT size must divide 64");
    constexpr T c = (T)cos64(m * (64 / E));
    constexpr T s = (T)(DIR * sin64(m * (64 / E)));
    return {v.r * c - v.i * s, v.r * s + v.i * c};
  }
}

// ---------------------------------------------------------------------------
// FP64 phase handling.  Cody-Waite reduction of an FP64 phase to [-pi, pi]
// with a two-part 2*pi constant and FMA (the product n*2pi_hi is formed
// exactly inside the fused multiply-add), then sine/cosine in T.
// ---------------------------------------------------------------------------
constexpr double kTwoPiHi = 6.283185307179586232e+00;   // nearest double to 2*pi
constexpr double kTwoPiLo = 2.449293598294706e-16;     // 2*pi - kTwoPiHi
constexpr double kInvTwoPi = 0.15915494309189535;

__device__ __forceinline__ double reduce_2pi(double th) {
  double n = rint(th * kInvTwoPi);
  double r = fma(-n, kTwoPiHi, th);
  return fma(-n, kTwoPiLo, r);
}

// (-1)^k / n! for the Taylor coefficients below
constexpr double taylor_coef(int k, int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return ((k & 1) ? -1.0 : 1.0) / f;
}

// sin r and cos r for |r| <= 1/2 from their Taylor polynomials in r^2 by
// Horner's rule, sin r = r (1 + r^2 (s1 + r^2 (s2 + ... s_K))) with
// s_k = (-1)^k / (2k+1)! and cos r = 1 + r^2 (c1 + ... c_K) with
// c_k = (-1)^k / (2k)!.  K terms leave a truncation error below
// 0.5^(2K+2) / (2K+2)!: under half an ulp of T for K = 4 (float) and K = 7
// (double).  One FMA per term and function.
template <class T, int K>
__host__ __device__ __forceinline__ void sincos_small(T r, T& s, T& c) {
  const T r2 = r * r;
  constexpr T sK = (T)taylor_coef(K, 2 * K + 1), cK = (T)taylor_coef(K, 2 * K);
  T ps = sK, pc = cK;
  sfor<1, K>([&](auto kk) {
    constexpr int k = K - (int)decltype(kk)::value;  // K-1 .. 1
    constexpr T sk = (T)taylor_coef(k, 2 * k + 1), ck = (T)taylor_coef(k, 2 * k);
    ps = fmad(r2, ps, sk);
    pc = fmad(r2, pc, ck);
  });
  ps = fmad(r2, ps, (T)1);
  pc = fmad(r2, pc, (T)1);
  s = r * ps;
  c = pc;
}

// exp(-i*theta) with an FP64 argument: range reduction in FP64, then the
// short series for the common small reduced phase, else the library sincos.
template <class T>
__device__ __forceinline__ cpx<T> expmi(double th) {
  double r = reduce_2pi(th);
  if (fabs(r) <= 0.5) {
    T s, c;
    sincos_small<T, std::is_same<T, float>::value ? 4 : 7>((T)r, s, c);
    return {c, -s};
  }
  if constexpr (std::is_same<T, float>::value) {
    float s, c;
    sincosf((float)r, &s, &c);
    return {c, -s};
  } else {
    double s, c;
    sincos(r, &s, &c);
    return {c, -s};
  }
}

// ---------------------------------------------------------------------------
// atomics on non-negative IEEE values (bit patterns order like the values)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_max_nonneg(float* addr, float v) {
  atomicMax(reinterpret_cast<unsigned int*>(addr), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr),
            (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void atomic_min_nonneg(double* addr, double v) {
  atomicMin(reinterpret_cast<unsigned long long*>(addr),
            (unsigned long long)__double_as_longlong(v));
}

// max(a, b) that keeps a NaN of b (a NaN already in a stays as well)
__device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// Signal speed cs + |v + h a| of one gas cell (quantum-fluid-accretion.py:
// 696-699, isothermal c_eff = cs): the kicked velocity in T with explicit
// FMAs, the norm in FP64.  Every kernel that evaluates it calls this, so
// they all give the same bits.
template <class T>
__device__ __forceinline__ double gas_signal_sq(T u, T v, T w, T ax, T ay, T az, T h) {
  const double du = fmad(ax, h, u), dv = fmad(ay, h, v), dw = fmad(az, h, w);
  return __builtin_fma(dw, dw, __builtin_fma(dv, dv, du * du));
}
template <class T>
__device__ __forceinline__ double gas_signal(T u, T v, T w, T ax, T ay, T az, T h, double cs) {
  return cs + sqrt(gas_signal_sq<T>(u, v, w, ax, ay, az, h));
}
template <class T>
__device__ __forceinline__ double gas_signal(T u, T v, T w, double cs) {
  const double du = u, dv = v, dw = w;
  return cs + sqrt(__builtin_fma(dw, dw, __builtin_fma(dv, dv, du * du)));
}

// wavefront (64-lane) reductions
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Slots of the device-side diagnostics word array (doubles).
enum DiagSlot {
  kDiagAMax = 0,     // running max |a_grid| (quantum-stars.py:260,327)
  kDiagDtS = 1,      // min |dx / vel| (quantum-stars.py:328)
  kDiagMass = 2,     // FP64 sum of the last Poisson source over cells
  kDiagNaN = 3,      // non-zero once a NaN/Inf was seen in the source
  kDiagCount = 8
};

}  // namespace qj
