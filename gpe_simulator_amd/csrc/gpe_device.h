// Device-side building blocks shared by every kernel of libgpe_hip.so:
// complex arithmetic, compile-time trigonometry, static loops and the
// in-register DFT codelets used by the pencil FFT passes.
//
// Written for gfx950 (CDNA4): 64-wide wavefronts, FP64 at full vector rate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace gpe {

// ---------------------------------------------------------------- complex --
template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator+(cpx<R> a, cpx<R> b) {
  return cpx<R>{a.x + b.x, a.y + b.y};
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> operator-(cpx<R> a, cpx<R> b) {
  return cpx<R>{a.x - b.x, a.y - b.y};
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return cpx<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__host__ __device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) {
  return cpx<R>{a.x * s, a.y * s};
}
// Multiply by S*i with S = +1 or -1 known at compile time.
template <int S, typename R>
__host__ __device__ __forceinline__ cpx<R> mul_i(cpx<R> a) {
  if constexpr (S > 0) {
    return cpx<R>{-a.y, a.x};
  } else {
    return cpx<R>{a.y, -a.x};
  }
}

// FP64 product and sum, each rounded on its own and never fused into an FMA
// with the operation that feeds or consumes it, as the plain C expression
// a*b + c evaluates without contraction.  hipcc contracts by default, and
// __dmul_rn / __dadd_rn are plain * and + in HIP's headers, so they fuse as
// well once inlined; the pragma drops the contract flag of these operations.
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

// a * w with the product formed in FP64 and rounded once to R: for factors
// that are themselves computed per point in FP64 (the kinetic phase) and for
// the generic transform's FP64 twiddle table.
template <typename R>
__device__ __forceinline__ cpx<R> cmul_d(cpx<R> a, cpx<double> w) {
  const double x = a.x, y = a.y;
  return cpx<R>{R(x * w.x - y * w.y), R(x * w.y + y * w.x)};
}
template <typename R>
__device__ __forceinline__ cpx<R> cscale_d(cpx<R> a, double s) {
  return cpx<R>{R(double(a.x) * s), R(double(a.y) * s)};
}

// A constant complex factor (c, s) = c + i s in the field's arithmetic.
// complex128: FP64.  complex64: each component as a float pair hi + lo, lo
// the rounding error of hi, so the constant is carried to ~2^-48 and the
// product needs float instructions only (one multiply and three FMAs per
// component, against two conversions in, two FP64 operations and one
// conversion out).  The product's rounding depends on the data, so, as with
// FP64 constants, no rounding of the constant repeats identically every step.
template <typename R>
struct cconst {
  R c, s;
};
template <>
struct alignas(16) cconst<float> {
  float ch, sh, cl, sl;
};

template <typename R>
__host__ __device__ constexpr cconst<R> make_cconst(double c, double s) {
  if constexpr (std::is_same_v<R, float>) {
    const float ch = float(c), sh = float(s);
    return cconst<float>{ch, sh, float(c - double(ch)), float(s - double(sh))};
  } else {
    return cconst<double>{c, s};
  }
}
__device__ __forceinline__ cconst<float> conj_c(cconst<float> w) { return {w.ch, -w.sh, w.cl, -w.sl}; }
__device__ __forceinline__ cconst<double> conj_c(cconst<double> w) { return {w.c, -w.s}; }

__device__ __forceinline__ cpx<float> cmul_c(cpx<float> a, cconst<float> w) {
  return cpx<float>{fmaf(a.x, w.ch, fmaf(-a.y, w.sh, fmaf(a.x, w.cl, -a.y * w.sl))),
                    fmaf(a.x, w.sh, fmaf(a.y, w.ch, fmaf(a.x, w.sl, a.y * w.cl)))};
}
__device__ __forceinline__ cpx<double> cmul_c(cpx<double> a, cconst<double> w) {
  return cpx<double>{a.x * w.c - a.y * w.s, a.x * w.s + a.y * w.c};
}

// a * c for a real factor c held as the real part of a cconst (its hi + lo
// pair for complex64): the imaginary-time kinetic factor.
__device__ __forceinline__ cpx<float> scale_re(cpx<float> a, cconst<float> w) {
  return cpx<float>{fmaf(a.x, w.ch, a.x * w.cl), fmaf(a.y, w.ch, a.y * w.cl)};
}
__device__ __forceinline__ cpx<double> scale_re(cpx<double> a, cconst<double> w) {
  return cpx<double>{a.x * w.c, a.y * w.c};
}

// A real constant in the same representation (float hi + lo for complex64).
template <typename R>
struct rconst {
  R v;
};
template <>
struct rconst<float> {
  float hi, lo;
};
template <typename R>
__host__ __device__ constexpr rconst<R> make_rconst(double h) {
  if constexpr (std::is_same_v<R, float>) {
    const float hi = float(h);
    return rconst<float>{hi, float(h - double(hi))};
  } else {
    return rconst<double>{h};
  }
}
__device__ __forceinline__ float scale_r(float x, rconst<float> h) { return fmaf(x, h.hi, x * h.lo); }
__device__ __forceinline__ double scale_r(double x, rconst<double> h) { return x * h.v; }
template <typename R>
__device__ __forceinline__ cpx<R> scale_r(cpx<R> a, rconst<R> h) {
  return cpx<R>{scale_r(a.x, h), scale_r(a.y, h)};
}
// a * h for a compile-time real constant h.
template <typename R>
__device__ __forceinline__ R scale_c(R x, double h) {
  return scale_r(x, make_rconst<R>(h));
}

// ------------------------------------------------------------ static loop --
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// --------------------------------------------------- compile-time sin/cos --
// Taylor series on |a| <= pi/4 after an exact integer octant reduction of the
// angle 2*pi*r/m.  Used only in constant expressions (codelet twiddles).
constexpr double ce_sin_small(double a) {
  double term = a, sum = a;
  for (int n = 1; n < 15; ++n) {
    term *= -a * a / double((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos_small(double a) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 15; ++n) {
    term *= -a * a / double((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
constexpr double kQuarterPi = 0.78539816339744830961566084581987572;

// cos(2*pi*r/m) (want_sin == false) or sin(2*pi*r/m) (want_sin == true).
constexpr double ce_trig2pi(long r, long m, bool want_sin) {
  long rr = r % m;
  if (rr < 0) rr += m;
  const long u = 8 * rr;        // angle = (pi/4) * u / m
  const long oct = u / m;       // 0..7
  const long f = u % m;         // remainder within the octant
  const double a = kQuarterPi * double(f) / double(m);
  const double b = kQuarterPi * double(m - f) / double(m);
  double s = 0.0, c = 0.0;
  switch (oct) {
    case 0: s = ce_sin_small(a); c = ce_cos_small(a); break;
    case 1: s = ce_cos_small(b); c = ce_sin_small(b); break;
    case 2: s = ce_cos_small(a); c = -ce_sin_small(a); break;
    case 3: s = ce_sin_small(b); c = -ce_cos_small(b); break;
    case 4: s = -ce_sin_small(a); c = -ce_cos_small(a); break;
    case 5: s = -ce_cos_small(b); c = -ce_sin_small(b); break;
    case 6: s = -ce_cos_small(a); c = ce_sin_small(a); break;
    default: s = -ce_sin_small(b); c = ce_cos_small(b); break;
  }
  return want_sin ? s : c;
}

constexpr int ce_first_factor(int m) {
  if (m % 4 == 0 && m != 4) return 4;
  if (m % 2 == 0 && m != 2) return 2;
  if (m % 3 == 0 && m != 3) return 3;
  if (m % 5 == 0 && m != 5) return 5;
  if (m % 7 == 0 && m != 7) return 7;
  return m;  // prime (or a base case)
}

// a *= W_m^r with W_m = exp(DIR * 2*pi*i / m); DIR = -1 forward, +1 inverse.
template <int r, int m, int DIR, typename R>
__device__ __forceinline__ void twiddle(cpx<R>& a) {
  constexpr int rr = ((r % m) + m) % m;
  if constexpr (rr == 0) {
    return;
  } else if constexpr (2 * rr == m) {
    a = cpx<R>{-a.x, -a.y};
  } else if constexpr (4 * rr == m) {
    a = mul_i<DIR>(a);
  } else if constexpr (4 * rr == 3 * m) {
    a = mul_i<-DIR>(a);
  } else if constexpr ((8 * rr) % m == 0) {
    // odd multiples of pi/4: (c + DIR*i*s) with |c| = |s| = sqrt(1/2)
    constexpr double c = ce_trig2pi(rr, m, false);
    constexpr double s = DIR * ce_trig2pi(rr, m, true);
    constexpr double h = 0.70710678118654752440084436210484904;
    const R p = a.x, q = a.y;
    if constexpr (c > 0 && s > 0) {
      a = cpx<R>{scale_c(p - q, h), scale_c(p + q, h)};
    } else if constexpr (c > 0 && s < 0) {
      a = cpx<R>{scale_c(p + q, h), scale_c(q - p, h)};
    } else if constexpr (c < 0 && s > 0) {
      a = cpx<R>{-scale_c(p + q, h), scale_c(p - q, h)};
    } else {
      a = cpx<R>{scale_c(q - p, h), -scale_c(p + q, h)};
    }
  } else {
    constexpr double c = ce_trig2pi(rr, m, false);
    constexpr double s = DIR * ce_trig2pi(rr, m, true);
    constexpr cconst<R> w = make_cconst<R>(c, s);
    a = cmul_c(a, w);
  }
}

// ---------------------------------------------------------------- codelets --
// In-place DFT of length M held in registers, natural order in and out:
//   v[k] <- sum_n v[n] * exp(DIR * 2*pi*i*n*k / M)
template <int M, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R>* v);

template <int M, int DIR, typename R>
__device__ __forceinline__ void dft_prime(cpx<R>* v) {
  cpx<R> out[M];
  static_for<0, M>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    cpx<R> acc = v[0];
    static_for<1, M>([&](auto nc) {
      constexpr int n = decltype(nc)::value;
      cpx<R> term = v[n];
      twiddle<(n * k) % M, M, DIR>(term);
      acc = acc + term;
    });
    out[k] = acc;
  });
  static_for<0, M>([&](auto kc) { v[decltype(kc)::value] = out[decltype(kc)::value]; });
}

template <int M, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R>* v) {
  if constexpr (M == 1) {
    return;
  } else if constexpr (M == 2) {
    const cpx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (M == 3) {
    constexpr double h = 0.86602540378443864676372317075293618;  // sqrt(3)/2
    const cpx<R> a0 = v[0], t = v[1] + v[2], d = v[1] - v[2];
    const cpx<R> m = cpx<R>{a0.x - R(0.5) * t.x, a0.y - R(0.5) * t.y};
    const cpx<R> r = mul_i<DIR>(cpx<R>{scale_c(d.x, h), scale_c(d.y, h)});
    v[0] = a0 + t;
    v[1] = m + r;
    v[2] = m - r;
  } else if constexpr (M == 4) {
    const cpx<R> s02 = v[0] + v[2], d02 = v[0] - v[2];
    const cpx<R> s13 = v[1] + v[3], d13 = mul_i<DIR>(v[1] - v[3]);
    v[0] = s02 + s13;
    v[2] = s02 - s13;
    v[1] = d02 + d13;
    v[3] = d02 - d13;
  } else if constexpr (ce_first_factor(M) == M) {
    dft_prime<M, DIR, R>(v);
  } else {
    // Cooley-Tukey M = P*Q:  n = Q*n1 + n2,  k = k1 + P*k2.
    constexpr int P = ce_first_factor(M);
    constexpr int Q = M / P;
    cpx<R> a[Q][P];
    static_for<0, Q>([&](auto n2c) {
      constexpr int n2 = decltype(n2c)::value;
      static_for<0, P>([&](auto n1c) {
        constexpr int n1 = decltype(n1c)::value;
        a[n2][n1] = v[Q * n1 + n2];
      });
      dft<P, DIR, R>(a[n2]);
      static_for<1, P>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        twiddle<n2 * k1, M, DIR>(a[n2][k1]);
      });
    });
    static_for<0, P>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      cpx<R> b[Q];
      static_for<0, Q>([&](auto n2c) { b[decltype(n2c)::value] = a[decltype(n2c)::value][k1]; });
      dft<Q, DIR, R>(b);
      static_for<0, Q>([&](auto k2c) {
        constexpr int k2 = decltype(k2c)::value;
        v[k1 + P * k2] = b[k2];
      });
    });
  }
}

// ---------------------------------------------------- global memory access --
// Pointers read from the device-resident solve descriptor are generic; these
// casts let the compiler emit global_load/global_store instead of flat ones.
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  typedef const __attribute__((address_space(1))) T* gptr;
  return *(gptr)(p);
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, T v) {
  typedef __attribute__((address_space(1))) T* gptr;
  *(gptr)(p) = v;
}
template <typename R>
__device__ __forceinline__ cpx<R> gload(const cpx<R>* p) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  typedef const __attribute__((address_space(1))) v2* gptr;
  const v2 v = *(gptr)(p);
  return cpx<R>{v.x, v.y};
}
template <typename R>
__device__ __forceinline__ void gstore(cpx<R>* p, cpx<R> c) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  typedef __attribute__((address_space(1))) v2* gptr;
  v2 v;
  v.x = c.x;
  v.y = c.y;
  *(gptr)(p) = v;
}

// ----------------------------------------------------------------- sincos --
// sin and cos of the local-step phase, with no library call on any path (a
// call in the per-point loop makes every value live across it spill to
// scratch memory, even when the call never runs).
//
//   |x| < 0.78     Taylor polynomials directly (split-step phases are usually
//                  this small);
//   |x| < 2^40     Cody-Waite reduction by pi/2 in FP64, then the polynomials
//                  and the quadrant fix-up;
//   otherwise      NaN (also for non-finite x): a phase of 1e12 rad per half
//                  step has no meaningful value in either precision.
//
// The reduction uses a three-part split of pi/2 in FMA form: each FMA forms
// its product q * p_i exactly, so r = x - q*pi/2 carries an absolute error
// near 1e-16 while q = rint(x * 2/pi) is exact.  Below 2^40 the rounding of
// x * 2/pi moves q by at most one only when r is within 1.2e-4 of +-pi/4, so
// |r| stays inside the polynomials' accurate range.  The quadrant is q mod 4,
// taken in FP64 (q may exceed the int range).
__device__ __forceinline__ double reduce_half_pi(double x, int* quad) {
  const double q = rint(x * 0.63661977236758134308);
  double r = fma(q, -1.5707963267948966, x);
  r = fma(q, -6.123233995736766e-17, r);
  r = fma(q, 1.4973849048591698e-33, r);
  *quad = int(q - 4.0 * floor(q * 0.25));
  return r;
}
constexpr double kPhaseLimit = 1099511627776.0;   // 2^40

// Polynomials on |r| <= pi/4 (Taylor to r^9 / r^10 in float, r^19 / r^18 in
// double).
__device__ __forceinline__ void sincos_poly(float r, float* s, float* c) {
  const float r2 = r * r;
  float ps = fmaf(r2, float(1.0 / 362880.0), float(-1.0 / 5040.0));
  ps = fmaf(r2, ps, float(1.0 / 120.0));
  ps = fmaf(r2, ps, float(-1.0 / 6.0));
  *s = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, float(-1.0 / 3628800.0), float(1.0 / 40320.0));
  pc = fmaf(r2, pc, float(-1.0 / 720.0));
  pc = fmaf(r2, pc, float(1.0 / 24.0));
  pc = fmaf(r2, pc, -0.5f);
  *c = fmaf(r2, pc, 1.0f);
}

// Sign and order of (sin, cos) of r + quad * pi/2.
template <typename R>
__device__ __forceinline__ void quadrant(int quad, R sn, R cs, R* s, R* c) {
  const R a = (quad & 1) ? cs : sn;
  const R b = (quad & 1) ? sn : cs;
  *s = (quad & 2) ? -a : a;
  *c = ((quad + 1) & 2) ? -b : b;
}

__device__ __forceinline__ void phase_sincos(float x, float* s, float* c) {
  if (fabsf(x) < 0.78f) {
    sincos_poly(x, s, c);
    return;
  }
  if (!(fabsf(x) < float(kPhaseLimit))) {
    *s = *c = __builtin_nanf("");
    return;
  }
  int quad;
  const float r = float(reduce_half_pi(double(x), &quad));
  float sn, cs;
  sincos_poly(r, &sn, &cs);
  quadrant(quad, sn, cs, s, c);
}

__device__ __forceinline__ void sincos_poly(double r, double* s, double* c) {
  const double r2 = r * r;
  double ps = fma(r2, -1.0 / 121645100408832000.0, 1.0 / 355687428096000.0);
  ps = fma(r2, ps, -1.0 / 1307674368000.0);
  ps = fma(r2, ps, 1.0 / 6227020800.0);
  ps = fma(r2, ps, -1.0 / 39916800.0);
  ps = fma(r2, ps, 1.0 / 362880.0);
  ps = fma(r2, ps, -1.0 / 5040.0);
  ps = fma(r2, ps, 1.0 / 120.0);
  ps = fma(r2, ps, -1.0 / 6.0);
  *s = fma(r * r2, ps, r);
  double pc = fma(r2, 1.0 / 6402373705728000.0, -1.0 / 20922789888000.0);
  pc = fma(r2, pc, 1.0 / 87178291200.0);
  pc = fma(r2, pc, -1.0 / 479001600.0);
  pc = fma(r2, pc, 1.0 / 3628800.0);
  pc = fma(r2, pc, -1.0 / 40320.0);
  pc = fma(r2, pc, 1.0 / 720.0);
  pc = fma(r2, pc, -1.0 / 24.0);
  pc = fma(r2, pc, 0.5);
  *c = fma(r2, -pc, 1.0);
}

__device__ __forceinline__ void phase_sincos(double x, double* s, double* c) {
  if (fabs(x) < 0.78) {
    sincos_poly(x, s, c);
    return;
  }
  if (!(fabs(x) < kPhaseLimit)) {
    *s = *c = __builtin_nan("");
    return;
  }
  int quad;
  const double r = reduce_half_pi(x, &quad);
  double sn, cs;
  sincos_poly(r, &sn, &cs);
  quadrant(quad, sn, cs, s, c);
}

// ------------------------------------------------------ wave-local exchange --
// Orders one wave's LDS accesses across its lanes without a workgroup barrier.
// A wave issues its LDS instructions in program order and the LDS serves them
// in that order, so every lane's writes before this point are seen by every
// lane's reads after it once the compiler keeps the accesses on their side of
// it (the wavefront-scope fences; no instruction is emitted for them).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------- reductions --
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide deterministic sum of NQ doubles; result valid in thread 0.
// `scratch` must hold NQ * (blockDim.x / 64) doubles.
template <int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double* scratch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) scratch[q * nw + wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += scratch[q * nw + w];
      v[q] = s;
    }
  }
}

__device__ __forceinline__ double coherent_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace gpe
