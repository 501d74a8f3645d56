// Pencil passes of the split-step propagator.
//
// A 3D field psi[i][j][k] (C order, z = k fastest, n^3 points) is transformed
// one axis at a time.  A "pass" is one read + one write sweep of the field in
// which every block owns a tile of whole pencils along one axis, runs a 1D FFT
// (or FFT -> pointwise -> inverse FFT) on them and fuses the pointwise physics:
//
//   Z passes (contiguous axis): local half-steps exp(-i(V + beta|psi|^2)dt/2),
//                               renormalisation scale, FP64 norm / energy sums;
//   Y passes (stride n):        plain forward / inverse FFT;
//   X pass  (stride n^2):       FFT_x -> kinetic factor -> inverse FFT_x.
//
// Two pencil transforms share the same pass code:
//   FastTF<R, N, TP, B, W>: N = TP * M with TP threads per pencil (B pencils
//     per block, register budget for W waves per SIMD); each thread
//     holds M values at positions t + TP*m; a length-M codelet in registers,
//     an FP64-derived twiddle, one LDS exchange and TP-point codelets.  The
//     output comes back in the same register layout, so FFT -> pointwise ->
//     inverse FFT needs no extra shuffles.
//   GenTF<R>: any n <= 1024, one pencil per 64-thread block, direct DFT from
//     LDS with an exact (n*k mod n) twiddle index.  Used for sizes that have
//     no fast instantiation.
#pragma once

#include "gpe_device.h"

namespace gpe {

enum PassKind : int {
  K_Z_FIRST = 0,   // [scale] -> half(slot_a) -> FFT_z
  K_Z_MID = 1,     // IFFT_z -> half(slot_a) -> half(slot_b) -> FFT_z   (real time)
  K_Z_LAST = 2,    // IFFT_z -> half(slot_a) [-> FP64 norm -> scale]
  K_Y_FWD = 3,
  K_Y_INV = 4,
  K_X_KIN = 5,     // FFT_x -> kinetic -> IFFT_x
  K_Z_DIAG = 6,    // [scale] -> FP64 sums in real space -> FFT_z into work
  K_X_DIAG = 7,    // FFT_x of work -> FP64 kinetic sums and |psi(k)|^2 shells
  K_Z_FWD = 8,     // plain FFT_z (psi -> work)
  K_Y_LZ = 9,      // FFT_y, FP64 sum x * k_y * |F|^2 (no write)
  K_X_LZ = 10,     // FFT_x, FP64 sum y * k_x * |F|^2 (no write)
  // The separable plan: exp(-i k^2 dt) = exp(-i kx^2 dt) exp(-i ky^2 dt)
  // exp(-i kz^2 dt), and a factor of one wavenumber commutes with the
  // transforms along the other axes, so the kinetic step is three 1D
  // propagators, each an FFT -> factor -> inverse FFT inside one pass: one
  // sweep of psi per axis and step.
  K_ZK_FIRST = 11, // [scale] -> half(slot_a) -> FFT_z -> kz factor -> IFFT_z
  K_ZK_MID = 12,   // half(slot_a) -> half(slot_b) -> FFT_z -> kz factor -> IFFT_z   (real time)
  K_YK = 13,       // FFT_y -> ky factor -> IFFT_y
  K_XK = 14,       // FFT_x -> kx factor -> IFFT_x
  K_XK_LAST = 15,  // FFT_x -> kx factor -> IFFT_x -> half(slot_a) [-> FP64 norm partials]
  // Plane passes (gpe_plane.h): one block per x plane runs a Z kind of the
  // separable plan and then YK, so a step sweeps psi twice instead of three
  // times where a plane fits the LDS.
  K_ZYK_FIRST = 16, // ZK_FIRST then YK
  K_ZYK_MID = 17,   // ZK_MID then YK   (real time)
  // Plane passes of the momentum-cutoff step (real time): the field between
  // passes is transformed along y and z, and X_KIN applies the 3D kinetic
  // factor, so a step sweeps psi twice instead of four times.
  K_ZY_FIRST = 18,  // Z_FIRST then Y_FWD
  K_ZY_MID = 19,    // Y_INV then Z_MID then Y_FWD
  K_ZY_LAST = 20,   // Y_INV then Z_LAST
  // Plain steps of the reference-structured (unfused) real-time step: one
  // launch per reference kernel and transform axis, for measurement.
  K_Z_HALF = 21,   // half(slot_a) alone, in real space
  K_Z_INV = 22,    // plain IFFT_z
  K_X_FWD = 23,    // plain FFT_x
  K_X_INV = 24,    // plain IFFT_x
};

// Kinds run by k_plane (one block per x plane).
__host__ __device__ constexpr bool is_plane_kind(int kind) {
  return kind == K_ZYK_FIRST || kind == K_ZYK_MID || kind == K_ZY_FIRST || kind == K_ZY_MID || kind == K_ZY_LAST;
}

enum PassFlags : int {
  F_PRESCALE = 1,  // multiply the loaded field by scal[1] (pending 1/sqrt(norm))
  F_IN_WORK = 2,   // read from the workspace instead of psi
  F_OUT_WORK = 4,  // write to the workspace instead of psi
  F_NORM = 8,      // Z_LAST / XK_LAST: FP64 norm partials; k_norm_finish computes the scale
  F_SPECTRUM = 16, // X_DIAG: also bin |F|^2 into the shells and the view
  F_STAMP = 32,    // record the launch's end on the device wall clock
  F_KIN1_LDS = 64, // separable kinds: the launch's dynamic LDS holds the 1D factor table
  F_KIN1_RT = 128, // separable kinds, real time: stage the 1D factor table in LDS where the launcher
                   // finds that it costs no occupancy (the launcher turns it into F_KIN1_LDS)
  F_NORM_FOLD = 256,  // the first pass of an imaginary step, with F_PRESCALE: finish the previous step's norm
                      // from its partials (norm_parts of them) in every block, instead of k_norm_finish
  F_NORM_DEFER = 512, // the norm kinds with F_NORM: write the partials only; the next step's first
                      // pass finishes the norm (F_NORM_FOLD)
};

// Per-solve device descriptor.  Lives in device memory; kernels index it with
// blockIdx.z so equal-shape solves can share one launch.
template <typename R>
struct SolveDev {
  cpx<R>* psi;
  cpx<R>* work;
  const cpx<double>* tw;   // tw[r] = exp(-2*pi*i*r/n), r < n, FP64
  const cpx<double>* kin_x;  // kinetic factors per axis, FP64 (1/n^3 folded into x)
  const cpx<double>* kin_y;
  const cpx<double>* kin_z;
  const cpx<double>* kin1;   // 1D factor of one axis by FFT index (n entries, 1/n folded in), FP64
  const double* karr;      // k axis, FP64
  const double* xarr;      // x axis, FP64
  const R* pxy;            // potential table over (x, y): [slot][n*n]
  const R* pz;             // potential table over z: [slot][n]
  const R* vre;            // static volume, real part (nullable)
  const R* vim;            // static volume, imaginary part (nullable)
  const R* vdyn;           // time-dependent volume [slot][n^3] (nullable)
  long long pxy_ss, pz_ss, vdyn_ss;   // slot strides (0 = static table)
  double* red;             // per-block partial sums
  double* scal;            // [0] norm, [1] 1/sqrt(norm)
  double* hist;            // norm history per step (nullable)
  unsigned long long* stamps;  // wall-clock slots: block ends and reference times (F_STAMP)
  double* nkocc;           // [block][nbins] |F|^2 partials
  unsigned int* nkcnt;     // [block][nbins] mode-count partials
  double dx3;              // dx*dx*dx
  double kcut2;            // k_cut^2 (masked kinetic)
  R beta;
  R hdt;                   // dt/2
  R bhdt;                  // beta * dt/2, rounded once from FP64
  int n;
  int nbins;               // dk shells of the reference's histogram
  int nview;               // equal-width bins of the view over [0, k_max] (after the shells)
  int vh;                  // largest |integer frequency| on an axis: n/2
  int imag;                // 1: imaginary time
  int masked;              // 1: zero modes with k^2 > kcut^2
};

template <typename R>
struct PassArgs {
  const SolveDev<R>* s;
  int slot_a, slot_b;
  int flags;
  int hist_index;          // step index into s->hist (Z_LAST with F_NORM; the folding pass with F_NORM_FOLD)
  int stamp_index;         // first of this launch's per-block end slots in s->stamps (F_STAMP)
  int norm_parts;          // F_NORM_FOLD: partials the previous step's last pass wrote per solve
  double* red;             // norm partials of Z_LAST / k_norm_finish; null: the solve's own (s->red)
};

// ------------------------------------------------------------ transforms --
template <typename R, int N_, int TP_, int B_, int W_, int RD_ = 1>
struct FastTF {
  static constexpr bool kFast = true;
  static constexpr int N = N_, TP = TP_, B = B_;
  static constexpr int M = N / TP;
  static constexpr int Q = M / TP;
  static constexpr int THREADS = TP * B;
  // Waves per SIMD the register allocation must allow (a tuning parameter:
  // more waves hide more latency but cap the registers per thread).
  static constexpr int MIN_WAVES = W_;
  // Exchange rounds of the Z passes: with 2, half of the block's pencils go
  // through the LDS exchange at a time, which halves its LDS (more blocks
  // per CU) at the cost of a second pair of barriers.  A Z-pass wave's
  // pencils are consecutive, so each wave belongs to one half.
  static constexpr int ROUNDS = RD_;
  static_assert(ROUNDS == 1 || (ROUNDS == 2 && (B / 2) % (64 / TP) == 0), "a round must hold whole waves");
  // LDS exchange layout: a pencil's M rows of TP values, row_stride<AXIS>()
  // apart, pencils pencil_stride<AXIS>() apart.
  // Z passes: a pencil's TP threads are adjacent lanes, and in the transpose
  // read lane t reads row t + TP*q, so rows are padded to TP + 1 to put
  // those lanes on different banks.
  // X/Y passes: adjacent lanes own adjacent pencils at the same t, and an
  // odd pencil stride puts up to 256 bytes of such pencils on distinct
  // banks (an even stride would put neighbours on the same bank).  When a
  // block's pencils span those 256 bytes (B * sizeof(complex) >= 256), the
  // lanes the LDS serves together hold pencils at one t, so the rows need no
  // padding; with fewer pencils they span two t and the padded rows keep
  // them apart.  Measured with the LDS bank-conflict counter, unpadded rows
  // conflict exactly as often as padded ones in these variants (never for
  // complex64 with 32 pencils; 1.3 cycles per LDS instruction for
  // complex128 with 16, either way).  They save a ninth to a fifth of the
  // exchange buffer, which in the Y passes at side 192 fits one more block
  // per CU (about 10% faster per launch).  The X passes keep the padding:
  // the smaller buffer there raises the occupancy the compiler derives from
  // LDS, it then holds the kinetic pass to the variant's register budget,
  // and at side 96 (complex64, 3 waves) that pass spills and runs 30% slower.
  static constexpr bool kPackedY = B * int(sizeof(cpx<R>)) >= 256;
  template <int AXIS>
  __host__ __device__ static constexpr int row_stride() {
    return AXIS == 1 && kPackedY ? TP : TP + 1;
  }
  // Z passes: the lanes served together hold whole pencils of TP lanes each,
  // and their rows are spread over the banks when the pencils start an odd
  // number of TP-element units apart: the smallest such stride that holds
  // M padded rows, TP * (M + Q) elements, rounded up to an odd multiple of TP.
  template <int AXIS>
  __host__ __device__ static constexpr int pencil_stride() {
    if (AXIS == 2) return TP * ((M + Q) % 2 == 0 ? M + Q + 1 : M + Q);
    return M * row_stride<AXIS>() + ((M * row_stride<AXIS>()) % 2 == 0 ? 1 : 0);
  }
  static_assert(TP * (M + Q) >= M * (TP + 1), "a Z pencil stride must hold M padded rows");
  template <int AXIS>
  __host__ __device__ static constexpr int lds_elems() {
    return B * pencil_stride<AXIS>();
  }
  static_assert(N % TP == 0 && M % TP == 0, "N must be TP*TP*Q");
  // X/Y tiles are B consecutive pencils of one row of n
  static_assert(N % B == 0, "B pencils per block must divide N");

  __device__ static __forceinline__ int pos(int t, int m) { return t + TP * m; }
  __device__ static __forceinline__ bool valid(int, int, int) { return true; }

  // Twiddles W_N^{t*k1}, k1 >= 1 (k1 = 0 is the unit factor), staged once
  // per block in LDS as twl[(k1-1)*TP + t], converted from the FP64 table to
  // the field's constant representation (cconst).
  using TwT = cconst<R>;
  static constexpr int TWL_ELEMS = (M - 1) * TP;
  __device__ static __forceinline__ void stage_twiddles(TwT* twl, const cpx<double>* tw) {
    for (int e = threadIdx.x; e < TWL_ELEMS; e += THREADS) {
      const cpx<double> w = gload(tw + (e % TP) * (e / TP + 1));
      twl[e] = make_cconst<R>(w.x, w.y);
    }
    __syncthreads();
  }

  // RNDS exchange rounds; `group` is the round in which this thread's
  // pencil uses the LDS (0 with one round).  Every thread passes every
  // barrier; only the LDS accesses depend on the group.
  //
  // In a Z pass with one round, a pencil's TP lanes sit in one wave and its
  // LDS rows are its own, so the exchange needs only the wave's own ordering
  // (wave_lds_sync), not a workgroup barrier: the block's waves then drift
  // apart freely, one loading while another computes.  With two rounds the
  // halves share the buffer, and in X/Y passes a pencil's threads span waves,
  // so those keep the barriers.
  template <int DIR, int RNDS = 1, int AXIS = 2>
  __device__ static __forceinline__ void fft(cpx<R>* v, cpx<R>* lp, int t, const TwT* twl, int, int group = 0) {
    constexpr int LSTRIDE = row_stride<AXIS>();
    constexpr bool kWaveLocal = AXIS == 2 && RNDS == 1 && 64 % TP == 0;
    auto sync = [] {
      if constexpr (kWaveLocal) wave_lds_sync();
      else __syncthreads();
    };
    dft<M, DIR, R>(v);
    static_for<1, M>([&](auto k1c) {
      constexpr int k1 = decltype(k1c)::value;
      TwT w = twl[(k1 - 1) * TP + t];
      if constexpr (DIR > 0) w = conj_c(w);
      v[k1] = cmul_c(v[k1], w);
      // scheduling fence every 4 twiddles: the twiddles (4 registers each)
      // are read from LDS as they are used, not all up front
      if constexpr (k1 % 4 == 0) __builtin_amdgcn_sched_barrier(0);
    });
    static_for<0, RNDS>([&](auto rc) {
      constexpr int r = decltype(rc)::value;
      const bool mine = RNDS == 1 || group == r;
      sync();
      if (mine) {
        static_for<0, M>([&](auto k1c) {
          constexpr int k1 = decltype(k1c)::value;
          lp[k1 * LSTRIDE + t] = v[k1];
        });
      }
      sync();
      if (mine) {
        static_for<0, Q>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          const int k1 = t + TP * q;
          cpx<R> u[TP];
          static_for<0, TP>([&](auto sc) { u[decltype(sc)::value] = lp[k1 * LSTRIDE + decltype(sc)::value]; });
          dft<TP, DIR, R>(u);
          static_for<0, TP>([&](auto k2c) {
            constexpr int k2 = decltype(k2c)::value;
            v[q + Q * k2] = u[k2];
          });
        });
      }
    });
  }
};

template <typename R>
struct GenTF {
  static constexpr bool kFast = false;
  static constexpr int N = 0, TP = 64, B = 1, M = 16;
  static constexpr int THREADS = 64;
  static constexpr int MIN_WAVES = 1;
  static constexpr int ROUNDS = 1;
  template <int AXIS>
  __host__ __device__ static constexpr int pencil_stride() {
    return 1024;
  }
  template <int AXIS>
  __host__ __device__ static constexpr int lds_elems() {
    return 1024;
  }

  __device__ static __forceinline__ int pos(int t, int m) { return t + 64 * m; }
  __device__ static __forceinline__ bool valid(int t, int m, int n) { return t + 64 * m < n; }
  using TwT = cpx<double>;
  static constexpr int TWL_ELEMS = 1;
  __device__ static __forceinline__ void stage_twiddles(TwT*, const cpx<double>*) {}

  template <int DIR, int RNDS = 1, int AXIS = 2>
  __device__ static __forceinline__ void fft(cpx<R>* v, cpx<R>* lp, int t, const cpx<double>* __restrict__ tw,
                                             int n, int = 0) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (t + 64 * m < n) lp[t + 64 * m] = v[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int k = t + 64 * m;
      if (k < n) {
        cpx<R> acc{R(0), R(0)};
        int idx = 0;
        for (int s = 0; s < n; ++s) {
          cpx<double> w = gload(tw + idx);
          if constexpr (DIR > 0) w.y = -w.y;
          acc = acc + cmul_d(lp[s], w);
          idx += k;
          if (idx >= n) idx -= n;
        }
        v[m] = acc;
      }
    }
  }
};

// ------------------------------------------------------------- pointwise --
// exp of the decay exponents.  complex64: 2^(x log2 e) on the hardware exp2
// (about 1 ulp), with log2 e as a hi + lo pair so that the exponent is formed
// to a fraction of an ulp; three instructions per point against the library
// expf's range handling.  complex128: the library exp.
__device__ __forceinline__ float rexp(float x) {
  constexpr rconst<float> kLog2e = make_rconst<float>(1.44269504088896340736);
  return __builtin_amdgcn_exp2f(scale_r(x, kLog2e));
}
__device__ __forceinline__ double rexp(double x) { return exp(x); }

// One local half-step at a point.  `ph` is the potential's share of the
// phase, (static + dynamic real part) * dt/2, so the exponent is one FMA:
// e = (beta dt/2) |psi|^2 + ph.  `vim` is the static imaginary part.  IMAG and
// VIM (an absorber is present) are compile-time choices: the pass picks the
// instantiation once per tile, so no point evaluates a factor it discards.
// The exponent of one local phase, (beta dt/2)|psi|^2 + ph, and of two merged
// ones without an absorber, 2 (beta dt/2)|psi|^2 + ph.
template <bool MERGED, typename R>
__device__ __forceinline__ R phase_exponent(cpx<R> v, R ph, const SolveDev<R>& s) {
  const R a2 = v.x * v.x + v.y * v.y;
  if constexpr (MERGED) return (s.bhdt + s.bhdt) * a2 + ph;
  else return s.bhdt * a2 + ph;
}
// v * exp(-i e) from (sin e, cos e).
template <typename R>
__device__ __forceinline__ cpx<R> rotate_by(cpx<R> v, R sn, R cs) {
  return cpx<R>{v.x * cs + v.y * sn, v.y * cs - v.x * sn};
}

template <bool IMAG, bool VIM, typename R>
__device__ __forceinline__ cpx<R> half_step(cpx<R> v, R ph, R vim, const SolveDev<R>& s) {
  const R e = phase_exponent<false>(v, ph, s);
  if constexpr (IMAG) {
    const R f = rexp(-e);
    return cpx<R>{v.x * f, v.y * f};
  } else {
    R sn, cs;
    phase_sincos(e, &sn, &cs);
    cpx<R> r = rotate_by(v, sn, cs);
    if constexpr (VIM) {
      const R d = rexp(vim * s.hdt);
      r = cpx<R>{r.x * d, r.y * d};
    }
    return r;
  }
}

// Two consecutive real-time half-steps as one phase, `ph` the potential share
// of both (pa + pb) * dt/2: the first multiplies |psi| by d = exp(vim dt/2),
// so the second sees |psi|^2 d^2, and the product of the two factors is
// d^2 exp(-i(ea + eb)) with ea + eb = (beta dt/2)(|psi|^2 + |psi|^2 d^2) + ph.
template <bool VIM, typename R>
__device__ __forceinline__ cpx<R> two_half_steps(cpx<R> v, R ph, R vim, const SolveDev<R>& s) {
  R d2 = R(1), e;
  if constexpr (VIM) {
    const R a2 = v.x * v.x + v.y * v.y;
    const R d = rexp(vim * s.hdt);
    d2 = d * d;
    e = s.bhdt * (a2 + a2 * d2) + ph;
  } else {
    e = phase_exponent<true>(v, ph, s);
  }
  R sn, cs;
  phase_sincos(e, &sn, &cs);
  cpx<R> r = rotate_by(v, sn, cs);
  if constexpr (VIM) r = cpx<R>{r.x * d2, r.y * d2};
  return r;
}

__device__ __forceinline__ int freq_index(int i, int n) { return i <= (n - 1) / 2 ? i : i - n; }

// Shell index floor(|k|/dk + 1/2) from integer frequencies: with m = isqrt(q),
// sqrt(q) rounds up exactly when q > m^2 + m (sqrt(q) never equals m + 1/2).
__device__ __forceinline__ int shell_index(int p, int q, int r) {
  const int s = p * p + q * q + r * r;
  int m = int(sqrtf(float(s)));
  while (m * m > s) --m;
  while ((m + 1) * (m + 1) <= s) ++m;
  return m + (s > m * m + m ? 1 : 0);
}

// Bin of the equal-width view: floor(nview * |m| / (sqrt(3) * h)) for the
// integer frequency vector m and h = the largest |frequency| on an axis, i.e.
// |k| / k_max in nview bins with k_max the corner wavenumber.  Exact: b is the
// largest integer with 3 h^2 b^2 <= nview^2 |m|^2.  The float estimate
// sqrt(|m|^2) * scale, scale = nview / (sqrt(3) h), is off by at most one (a
// few ulp on a value below nview + 1), and one correction step settles it with
// FP64 products of integers below 2^53, which are exact.  The corner itself
// (b = nview) goes to the last bin.
__device__ __forceinline__ int view_index(int p, int q, int r, int h, int nview, float scale) {
  const int s = p * p + q * q + r * r;
  const double lhs = double(nview * nview) * double(s);
  const double c = 3.0 * double(h * h);
  int b = int(__builtin_amdgcn_sqrtf(float(s)) * scale);
  if (c * double(b * b) > lhs) {
    --b;
  } else if (c * double((b + 1) * (b + 1)) <= lhs) {
    ++b;
  }
  return b < nview ? b : nview - 1;
}

// The norm of the previous imaginary-time step, finished inside the first
// pass of the next step (F_NORM_FOLD) instead of by k_norm_finish.  Every
// block sums the FP64 partials exactly as k_norm_finish does (256 strided
// sums, each wave's shuffle tree, the four wave sums in order from zero), so
// each block gets the finisher's bits; block 0 publishes norm, scale and
// history as the finisher would.  The block's threads play the finisher's
// 256 in turn (blockDim.x a multiple of 64; the launcher folds only then).
// `sh` is 5 doubles of LDS.  Returns 1/sqrt(norm).
template <typename R>
__device__ __forceinline__ double fold_norm(const PassArgs<R>& a, const SolveDev<R>& s, double* sh) {
  const double* red = a.red ? a.red : s.red;
  const int nblk = a.norm_parts;
  const int lane = threadIdx.x & 63;
  for (int base = 0; base < 256; base += blockDim.x) {
    const int v = base + int(threadIdx.x);
    double tot = 0.0;
    if (v < 256)
      for (int q = v; q < nblk; q += 256) tot += red[q];
    tot = wave_sum(tot);
    if (lane == 0 && v < 256) sh[v >> 6] = tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int w = 0; w < 4; ++w) sum += sh[w];
    const double norm = s.dx3 * sum;
    const double inv = 1.0 / sqrt(norm);
    sh[4] = inv;
    if (blockIdx.x == 0) {
      s.scal[0] = norm;
      s.scal[1] = inv;
      if (s.hist) s.hist[a.hist_index] = norm;
    }
  }
  __syncthreads();
  return sh[4];
}

// ------------------------------------------------------------ pass kernel --
// Tile = B pencils; one block per tile.  (A multi-tile variant that prefetched
// the next tile into registers measured slower at every size: the live
// prefetch registers cost more occupancy than the overlap gained.)
// The separable plan's kinds (FFT along one axis, its 1D factor, inverse FFT).
__host__ __device__ constexpr bool is_kin1_kind(int kind) {
  return kind == K_ZK_FIRST || kind == K_ZK_MID || kind == K_YK || kind == K_XK || kind == K_XK_LAST;
}

// Dynamic LDS bytes a pass launch carries: the 1D factor table of an
// imaginary-time separable pass (F_KIN1_LDS), else none.
template <class TF, int KIND, typename R>
__host__ __device__ constexpr unsigned kin1_lds_bytes(int flags) {
  return TF::kFast && is_kin1_kind(KIND) && (flags & F_KIN1_LDS) ? unsigned(TF::N / 2 + 1) * sizeof(cconst<R>) : 0u;
}

// A pencil's first element is base = ubase + lane: ubase, the tile's first
// element, is the same for the whole block; lane is the pencil's offset in
// the tile.
template <class TF, int AXIS>
struct Tile {
  long long base, ubase;
  unsigned lane;
  int pi, pj, pk;  // pencil coordinates (the axis coordinate is unused)
  __device__ __forceinline__ Tile(int tile, int b, int n) : pi(0), pj(0), pk(0) {
    if constexpr (AXIS == 2) {
      const long long row = (long long)tile * TF::B + b;
      pi = int(row / n);
      pj = int(row % n);
      ubase = (long long)tile * TF::B * n;
      lane = unsigned(b) * unsigned(n);
    } else {
      const int tiles_per_row = n / TF::B;
      const int outer = tile / tiles_per_row;
      const int pk0 = (tile - outer * tiles_per_row) * TF::B;
      pk = pk0 + b;
      lane = unsigned(b);
      if constexpr (AXIS == 1) {
        pi = outer;
        ubase = (long long)pi * n * n + pk0;
      } else {
        pj = outer;
        ubase = (long long)pj * n + pk0;
      }
    }
    base = ubase + lane;
  }
};

// Element m of thread t's pencil in a fast-transform tile, addressed as the
// block-uniform pointer tile + TP*m*stride plus the thread's 32-bit byte
// offset, so the compiler forms each address as a scalar base and a vector
// offset (no 64-bit vector add per access).  The offset (lane + t*stride
// elements) stays far below 2^32 bytes for every fast side.
// (The base passes through readfirstlane, which pins it to scalar registers:
// without it the compiler folds the offsets together and forms a 64-bit
// vector address per element.)
template <class TF, typename T>
__device__ __forceinline__ T* tile_elem(T* tile, unsigned off_bytes, int m, long long stride) {
  using Byte = std::conditional_t<std::is_const_v<T>, const char, char>;
  const unsigned long long p = reinterpret_cast<unsigned long long>(tile + (long long)TF::TP * m * stride);
  const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(p));
  const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(p >> 32));
  Byte* base = reinterpret_cast<Byte*>((static_cast<unsigned long long>(hi) << 32) | lo);
  return reinterpret_cast<T*>(base + off_bytes);
}

template <class TF, int AXIS>
__host__ __device__ __forceinline__ int tile_count(int n) {
  return AXIS == 2 ? int((long long)n * n / TF::B) : (n / TF::B) * n;
}

template <class TF, typename R>
__device__ __forceinline__ void load_tile(cpx<R>* v, const cpx<R>* src, long long base, long long stride, int t,
                                          int n) {
#pragma unroll
  for (int m = 0; m < TF::M; ++m)
    if (TF::valid(t, m, n)) v[m] = gload(src + base + (long long)TF::pos(t, m) * stride);
}

template <class TF, int AXIS, int KIND, typename R>
__global__ __launch_bounds__(TF::THREADS, TF::MIN_WAVES) void k_pass(PassArgs<R> a) {
  // Only Z passes exchange in rounds: there a wave's lanes own whole
  // consecutive pencils; in the X/Y passes a pencil's threads span waves.
  constexpr int kRounds = AXIS == 2 ? TF::ROUNDS : 1;
  // Exchange-buffer layout: the pass's own axis, except that the complex64
  // separable X kinds use the Y passes' rows (unpadded where the block's
  // pencils span the banks): their exchange pattern is the Y passes', and
  // the smaller buffer fits more blocks per CU (96^3 batch of 64: xk 218 ->
  // 182 us).  X_KIN keeps the padded rows (with the smaller buffer it spills
  // at 96), and so does complex128 (its XK_LAST spills at 96 packed).
  constexpr int LAX = AXIS == 0 && sizeof(R) == 4 && (KIND == K_XK || KIND == K_XK_LAST) ? 1 : AXIS;
  __shared__ cpx<R> lds[TF::template lds_elems<LAX>() / kRounds];
  __shared__ typename TF::TwT twl_lds[TF::TWL_ELEMS];
  const SolveDev<R>& s = a.s[blockIdx.z];
  const int n = TF::kFast ? TF::N : s.n;
  constexpr int M = TF::M;
  // Launch timing from the device's constant-rate wall clock: when a wave is
  // done, its first lane raises its block's slot to the time (an atomic with
  // no return, so the wave does not wait; no barrier, which would make every
  // wave wait for its stores).  One slot per block: a single slot for the
  // launch would serialise thousands of atomics in L2.  The host takes the
  // latest block end as the launch's end, and a launch's span from the
  // previous timed launch's end (or a reference time) to its own, with no
  // timing events between launches.  The clock is read only at the end: the
  // clock intrinsic counts as a side effect, and one before the descriptor
  // reads would keep the compiler from reading the descriptor through the
  // scalar unit.
  const bool stamp = (a.flags & F_STAMP) != 0;

  int t, b;
  if constexpr (AXIS == 2) {
    t = threadIdx.x % TF::TP;
    b = threadIdx.x / TF::TP;
  } else {
    b = threadIdx.x % TF::B;
    t = threadIdx.x / TF::B;
  }
  const long long stride = AXIS == 2 ? 1LL : (AXIS == 1 ? (long long)n : (long long)n * n);
  const int ntiles = tile_count<TF, AXIS>(n);
  const int group = b / (TF::B / kRounds);
  cpx<R>* lp = lds + (b % (TF::B / kRounds)) * TF::template pencil_stride<LAX>();
  // Block-shared tables staged next to the twiddles (the twiddle barrier
  // covers them): the FP64 x kinetic factors for X_KIN and the z potential
  // line(s) for the Z passes, so they are read from LDS where used instead of
  // being loaded (and held in registers) ahead of use.
  constexpr bool kStageKin = TF::kFast && KIND == K_X_KIN;
  constexpr bool kStagePz =
      TF::kFast && (KIND == K_Z_FIRST || KIND == K_Z_MID || KIND == K_Z_LAST || KIND == K_Z_HALF);
  __shared__ cpx<double> kin_lds[kStageKin ? TF::N : 1];
  __shared__ double kx2_lds[kStageKin ? TF::N : 1];   // kx*kx for the momentum cutoff
  // the z line of the potential (Z_MID: both slots summed), so a point adds
  // it to its pencil's (x, y) share
  __shared__ R pz_lds[kStagePz ? TF::N : 1];
  if constexpr (kStageKin) {
    for (int e = threadIdx.x; e < TF::N; e += TF::THREADS) kin_lds[e] = gload(s.kin_x + e);
    if (s.masked) {
      for (int e = threadIdx.x; e < TF::N; e += TF::THREADS) {
        const double kx = gload(s.karr + e);
        kx2_lds[e] = mul_rn(kx, kx);
      }
    }
  }
  if constexpr (kStagePz) {
    for (int e = threadIdx.x; e < TF::N; e += TF::THREADS) {
      R pz = gload(s.pz + (a.slot_a * s.pz_ss + e));
      if constexpr (KIND == K_Z_MID) pz += gload(s.pz + (a.slot_b * s.pz_ss + e));
      pz_lds[e] = pz;
    }
  }
  // The separable plan's line of the potential along the pass's axis (Z
  // kinds: pz(z), both slots summed for ZK_MID; XK_LAST: the block's pxy(x, y)
  // column, every pencil of an X block sharing its y), staged in LDS where
  // that measured faster: every complex128 kind, and complex64 ZK_MID at the
  // sides below (65.8 -> 57.2 us at 256^3 when it was staged there, though
  // its blocks per CU fell from 4 to 3; the complex64 imaginary-time kinds
  // ran slower staged, at 3 blocks, and read the line from the cache).
  //
  // Complex64 ZK_FIRST / ZK_MID at sides that are multiples of 64 up to 256
  // hold the z line in registers instead: lane L of every wave keeps entries
  // L + 64 j (N / 64 registers), and a point fetches its entry from the lane
  // that holds it (ds_bpermute: the LDS crossbar, no LDS allocation).  A
  // wave's pencils all run along z and share the line, and every lane of the
  // block is active (a block is one whole tile).  The 1 KiB this frees lets
  // the real-time 1D factor table into LDS at 4 blocks per CU.
  constexpr bool kRegLine = TF::kFast && sizeof(R) == 4 && (KIND == K_ZK_FIRST || KIND == K_ZK_MID) &&
                            64 % TF::TP == 0 && TF::N % 64 == 0 && TF::N <= 256;
  constexpr int kRegLineN = kRegLine ? TF::N / 64 : 1;
  R line_reg[kRegLineN];
  if constexpr (kRegLine) {
    const int lane = int(threadIdx.x % 64);
#pragma unroll
    for (int j = 0; j < kRegLineN; ++j) {
      R pz = gload(s.pz + (a.slot_a * s.pz_ss + lane + 64 * j));
      if constexpr (KIND == K_ZK_MID) pz += gload(s.pz + (a.slot_b * s.pz_ss + lane + 64 * j));
      line_reg[j] = pz;
    }
  }
  constexpr bool kStageLine =
      TF::kFast && !kRegLine &&
      (KIND == K_ZK_MID || (sizeof(R) == 8 && (KIND == K_ZK_FIRST || KIND == K_XK_LAST)));
  __shared__ R line_lds[kStageLine ? TF::N : 1];
  if constexpr (kStageLine) {
    if constexpr (KIND == K_XK_LAST) {
      const int pj_block = int(blockIdx.x) / (TF::N / TF::B);   // the X tile's y (Tile<TF, 0>)
      for (int e = threadIdx.x; e < TF::N; e += TF::THREADS)
        line_lds[e] = gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)e * TF::N + pj_block));
    } else {
      for (int e = threadIdx.x; e < TF::N; e += TF::THREADS) {
        R pz = gload(s.pz + (a.slot_a * s.pz_ss + e));
        if constexpr (KIND == K_ZK_MID) pz += gload(s.pz + (a.slot_b * s.pz_ss + e));
        line_lds[e] = pz;
      }
    }
  }
  // The separable plan's 1D factor (s.kin1, by FFT index).  Real time reads
  // the FP64 table at each element's index (one vector offset and immediate
  // offsets for all of a thread's elements).  Imaginary time stages entries
  // 0..n/2, read by |frequency|, in the field's constant representation
  // (float hi + lo for complex64, so the product is one rounding of a factor
  // carried to ~2^-48 in float arithmetic alone), in dynamic LDS that only
  // imaginary-time launches carry (F_KIN1_LDS): without the 2 KiB table a
  // complex64 256^3 separable block fits 4 times per CU instead of 3.
  constexpr bool kKin1 = is_kin1_kind(KIND);
  constexpr bool kStageKin1 = TF::kFast && kKin1;
  // Real time stages the same entries as FP64 (16 bytes each, as the
  // complex64 hi + lo pair), where that keeps the launch's blocks per CU
  // (F_KIN1_LDS set by the launcher on a real-time pass): the factor table is
  // symmetric in the frequency, bit for bit, so the entry at |f| is the entry
  // at the FFT index, and the loads of a thread's 16 factors (four round
  // trips to the cache between the transforms) become LDS reads.
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  cconst<R>* kin1_lds = reinterpret_cast<cconst<R>*>(dyn_lds);
  cpx<double>* kin1_rt_lds = reinterpret_cast<cpx<double>*>(dyn_lds);
  static_assert(sizeof(cconst<R>) == sizeof(cpx<double>), "both staged factor tables take 16 bytes an entry");
  // (imaginary-time launches of a fast transform always carry the table;
  // real-time ones when the launcher set F_KIN1_LDS)
  const bool kin1_staged = kStageKin1 && (a.flags & F_KIN1_LDS) != 0;
  if (kin1_staged) {
    if (s.imag) {
      for (int e = threadIdx.x; e <= TF::N / 2; e += TF::THREADS) {
        const cpx<double> f = gload(s.kin1 + e);
        kin1_lds[e] = make_cconst<R>(f.x, f.y);
      }
    } else {
      for (int e = threadIdx.x; e <= TF::N / 2; e += TF::THREADS) kin1_rt_lds[e] = gload(s.kin1 + e);
    }
  }
  // |frequency| of element m of thread t's pencil.  In a fast transform the
  // first M/2 elements of every thread sit below n/2 (pos = t + TP*m with
  // TP*M/2 = n/2) and the rest at or above it, so |f| is pos or n - pos by m
  // alone.
  auto kin1_index = [&](int t_, int m) -> int {
    const int pos = TF::pos(t_, m);
    if constexpr (TF::kFast) {
      static_assert(TF::M % 2 == 0, "a fast transform's thread holds M even elements");
      return m < M / 2 ? pos : TF::N - pos;
    } else {
      const int f = freq_index(pos, n);
      return f < 0 ? -f : f;
    }
  };
  auto kin1_at = [&](int af) -> cconst<R> {
    if constexpr (kStageKin1) {
      return kin1_lds[af];
    } else {
      const cpx<double> w = gload(s.kin1 + af);
      return make_cconst<R>(w.x, w.y);
    }
  };
  auto kin_at = [&](int i) -> cpx<double> {
    if constexpr (kStageKin) return kin_lds[i];
    else return gload(s.kin_x + i);
  };
  auto kx2_at = [&](int i) -> double {
    if constexpr (kStageKin) {
      return kx2_lds[i];
    } else {
      const double kx = gload(s.karr + i);
      return mul_rn(kx, kx);
    }
  };
  // Elements per scheduling chunk in the kinetic loop: the compiler may not
  // move work across chunk boundaries, which bounds the FP64 factors it keeps
  // live at once.
  constexpr int CHUNK = 4;
  TF::stage_twiddles(twl_lds, s.tw);
  const typename TF::TwT* twp;
  if constexpr (TF::kFast) twp = twl_lds;
  else twp = s.tw;
  const cpx<R>* src = (a.flags & F_IN_WORK) ? s.work : s.psi;
  cpx<R>* dst = (a.flags & F_OUT_WORK) ? s.work : s.psi;
  const bool has_vim = s.vim != nullptr;
  const bool prescale = (a.flags & F_PRESCALE) != 0;
  const bool fold = (a.flags & F_NORM_FOLD) != 0;
  const bool norm = (a.flags & F_NORM) != 0;
  // the pending 1/sqrt(norm) is applied from its FP64 value (float hi + lo
  // for complex64) so the renormalised state has unit norm to rounding, not
  // to the float representation of the scale
  rconst<R> sc = make_rconst<R>(prescale && !fold ? s.scal[1] : 1.0);

  // Reductions accumulated over all tiles of the block.
  double acc[3] = {0.0, 0.0, 0.0};
  __shared__ double hocc[KIND == K_X_DIAG ? 1024 : 1];
  __shared__ unsigned int hcnt[KIND == K_X_DIAG ? 1024 : 1];
  const bool spectrum = KIND == K_X_DIAG && (a.flags & F_SPECTRUM) != 0;
  if (spectrum) {
    for (int q = threadIdx.x; q < s.nbins + s.nview; q += blockDim.x) {
      hocc[q] = 0.0;
      hcnt[q] = 0u;
    }
  }

  const int tile = blockIdx.x;
  if (tile >= ntiles) return;
  cpx<R> v[M];
  {
    const Tile<TF, AXIS> g(tile, b, n);
    if constexpr (TF::kFast) {
      const unsigned off = (g.lane + unsigned(t) * unsigned(stride)) * unsigned(sizeof(cpx<R>));
#pragma unroll
      for (int m = 0; m < M; ++m) v[m] = gload(tile_elem<TF>(src + g.ubase, off, m, stride));
    } else {
      load_tile<TF>(v, src, g.base, stride, t, n);
    }
  }
  if constexpr (KIND == K_Z_FIRST || KIND == K_ZK_FIRST) {
    // (while the tile's loads are in flight)
    if (fold) {
      __shared__ double fold_lds[5];
      sc = make_rconst<R>(fold_norm(a, s, fold_lds));
    }
  }
  {
    const Tile<TF, AXIS> g(tile, b, n);
    const int pi = g.pi, pj = g.pj, pk = g.pk;
    bool write = true;

    if constexpr (KIND == K_Y_FWD || KIND == K_Z_FWD || KIND == K_X_FWD) {
      TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
    } else if constexpr (KIND == K_Y_INV || KIND == K_Z_INV || KIND == K_X_INV) {
      TF::template fft<+1, kRounds, LAX>(v, lp, t, twp, n, group);
    } else if constexpr (KIND == K_X_KIN) {
      TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
      if (s.imag) {
        const double eyz = gload(s.kin_y + pj).x * gload(s.kin_z + pk).x;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (TF::valid(t, m, n)) {
            const double f = kin_at(TF::pos(t, m)).x * eyz;
            v[m] = cscale_d(v[m], f);
          }
          if (m % CHUNK == CHUNK - 1) __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        const cpx<double> eyz = cmul(gload(s.kin_y + pj), gload(s.kin_z + pk));
        double ky2 = 0.0, kz2 = 0.0;
        if (s.masked) {
          ky2 = mul_rn(gload(s.karr + pj), gload(s.karr + pj));
          kz2 = mul_rn(gload(s.karr + pk), gload(s.karr + pk));
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (TF::valid(t, m, n)) {
            const int i = TF::pos(t, m);
            cpx<double> f = cmul(kin_at(i), eyz);
            if (s.masked) {
              // (kx*kx + ky*ky) + kz*kz with every product and sum rounded
              const double k2 = add_rn(add_rn(kx2_at(i), ky2), kz2);
              if (k2 > s.kcut2) f = cpx<double>{0.0, 0.0};
            }
            v[m] = cmul_d(v[m], f);
          }
          if (m % CHUNK == CHUNK - 1) __builtin_amdgcn_sched_barrier(0);
        }
      }
      TF::template fft<+1, kRounds, LAX>(v, lp, t, twp, n, group);
    } else if constexpr (KIND == K_Z_FIRST || KIND == K_Z_MID || KIND == K_Z_LAST || KIND == K_Z_HALF) {
      if constexpr (KIND == K_Z_MID || KIND == K_Z_LAST) {
        TF::template fft<+1, kRounds, LAX>(v, lp, t, twp, n, group);
      } else if (prescale) {
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = scale_r(v[m], sc);
      }
      const long long rowoff = g.base;
      // the pencil's (x, y) share of the potential (Z_MID: both slots).  The
      // potential is summed first and scaled by dt/2 once per point: scaling
      // the shares apart rounds the phase so that the complex64 norm falls
      // steadily (3e-10 per step in the 256^3 stirring run, against a drift
      // without trend when summed first).
      R pxy = gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)pi * n + pj));
      if constexpr (KIND == K_Z_MID) pxy += gload(s.pxy + (a.slot_b * s.pxy_ss + (long long)pi * n + pj));
      const R cxy = pxy;
      // VOL: a static or time-dependent volume adds a per-point term, read
      // from HBM; without one the potential comes from the two staged shares.
      auto points = [&](auto imag_tag, auto vim_tag, auto vol_tag) {
        constexpr bool IMAG = decltype(imag_tag)::value;
        constexpr bool VIM = decltype(vim_tag)::value;
        constexpr bool VOL = decltype(vol_tag)::value;
        // the potential's share of the phase at element m, times dt/2
        auto point_phase = [&](int m) -> R {
          const int k = TF::pos(t, m);
          R pzk;
          if constexpr (kStagePz) {
            pzk = pz_lds[k];
          } else {
            R pz = gload(s.pz + (a.slot_a * s.pz_ss + k));
            if constexpr (KIND == K_Z_MID) pz += gload(s.pz + (a.slot_b * s.pz_ss + k));
            pzk = pz;
          }
          R ph = cxy + pzk;
          if constexpr (VOL) {
            R pv = R(0);
            if (s.vre) pv = gload(s.vre + (rowoff + k));
            if constexpr (KIND == K_Z_MID) pv += pv;
            if (s.vdyn) {
              pv += gload(s.vdyn + (a.slot_a * s.vdyn_ss + rowoff + k));
              if constexpr (KIND == K_Z_MID) pv += gload(s.vdyn + (a.slot_b * s.vdyn_ss + rowoff + k));
            }
            ph += pv;
          }
          return ph * s.hdt;
        };
        auto keep = [&](int m, cpx<R> w) {
          if constexpr (KIND == K_Z_LAST) {
            if (norm) acc[0] += double(w.x) * double(w.x) + double(w.y) * double(w.y);
          }
          v[m] = w;
        };
        if constexpr (!IMAG && !VIM) {
          // every exponent first, then one branch per thread (as in the
          // separable passes' local half-steps)
          R e[M];
          bool small = true;
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (TF::valid(t, m, n)) {
              e[m] = phase_exponent<KIND == K_Z_MID>(v[m], point_phase(m), s);
              small = small && fabs(e[m]) < R(0.78);
            }
          }
          auto finish = [&](auto fast_tag) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
              if (TF::valid(t, m, n)) {
                R sn, cs;
                if constexpr (decltype(fast_tag)::value) sincos_poly(e[m], &sn, &cs);
                else phase_sincos(e[m], &sn, &cs);
                keep(m, rotate_by(v[m], sn, cs));
              }
            }
          };
          if (small) finish(std::true_type{});
          else finish(std::false_type{});
        } else {
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (TF::valid(t, m, n)) {
              const R ph = point_phase(m);
              R vi = R(0);
              if constexpr (VIM) vi = gload(s.vim + (rowoff + TF::pos(t, m)));
              if constexpr (KIND == K_Z_MID) keep(m, two_half_steps<VIM>(v[m], ph, vi, s));
              else keep(m, half_step<IMAG, VIM>(v[m], ph, vi, s));
            }
          }
        }
      };
      // Z_MID is real-time only; an absorber exists only in real time.
      const bool vol = s.vre != nullptr || s.vdyn != nullptr;
      auto with_vol = [&](auto imag_tag, auto vim_tag) {
        if (vol) points(imag_tag, vim_tag, std::true_type{});
        else points(imag_tag, vim_tag, std::false_type{});
      };
      if (KIND != K_Z_MID && s.imag) {
        if constexpr (KIND != K_Z_MID) with_vol(std::true_type{}, std::false_type{});
      } else if (has_vim) {
        with_vol(std::false_type{}, std::true_type{});
      } else {
        with_vol(std::false_type{}, std::false_type{});
      }
      if constexpr (KIND == K_Z_FIRST || KIND == K_Z_MID) TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
    } else if constexpr (kKin1) {
      // The separable plan.  Z kinds open with the local half-step(s) in real
      // space, XK_LAST closes with one after its inverse transform; between,
      // FFT along the pass's axis, that axis's 1D kinetic factor, inverse FFT.
      constexpr bool kLocalZ = KIND == K_ZK_FIRST || KIND == K_ZK_MID;
      constexpr bool kLocalX = KIND == K_XK_LAST;
      constexpr bool kMerged = KIND == K_ZK_MID;
      auto local = [&](auto imag_tag, auto vim_tag, auto vol_tag) {
        constexpr bool IMAG = decltype(imag_tag)::value;
        constexpr bool VIM = decltype(vim_tag)::value;
        constexpr bool VOL = decltype(vol_tag)::value;
        // the pencil's share of the potential (Z: its (x, y) entry; X: its z
        // entry) and, per point, the line along the pass's axis (Z: pz(z);
        // X: pxy(x, y) at the block's y); summed first, then scaled by dt/2
        R pen;
        if constexpr (kLocalZ) {
          pen = gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)pi * n + pj));
          if constexpr (kMerged) pen += gload(s.pxy + (a.slot_b * s.pxy_ss + (long long)pi * n + pj));
        } else {
          pen = gload(s.pz + (a.slot_a * s.pz_ss + pk));
        }
        // the potential's share of the phase at element m, times dt/2
        auto point_phase = [&](int m) -> R {
          const int idx = TF::pos(t, m);
          R line;
          if constexpr (kRegLine) {
            // idx = t + TP m = (t + TP (m % per)) + 64 (m / per)
            constexpr int per = 64 / TF::TP;
            const int src_lane = t + TF::TP * (m % per);
            line = __int_as_float(__builtin_amdgcn_ds_bpermute(src_lane * 4, __float_as_int(line_reg[m / per])));
          } else if constexpr (kStageLine) {
            line = line_lds[idx];
          } else if constexpr (kLocalZ) {
            line = gload(s.pz + (a.slot_a * s.pz_ss + idx));
            if constexpr (kMerged) line += s.pz_ss == 0 ? line : gload(s.pz + (a.slot_b * s.pz_ss + idx));
          } else {
            line = gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)idx * n + pj));
          }
          R ph = pen + line;
          if constexpr (VOL) {
            const long long vidx = g.base + (long long)idx * stride;
            R pv = R(0);
            if (s.vre) pv = gload(s.vre + vidx);
            if constexpr (kMerged) pv += pv;
            if (s.vdyn) {
              pv += gload(s.vdyn + (a.slot_a * s.vdyn_ss + vidx));
              if constexpr (kMerged) pv += gload(s.vdyn + (a.slot_b * s.vdyn_ss + vidx));
            }
            ph += pv;
          }
          return ph * s.hdt;
        };
        auto keep = [&](int m, cpx<R> w) {
          if constexpr (kLocalX) {
            if (norm) acc[0] += double(w.x) * double(w.x) + double(w.y) * double(w.y);
          }
          v[m] = w;
        };
        if constexpr (!IMAG && !VIM) {
          // Real time without an absorber: every exponent first, then one
          // branch per thread for all of its points (the polynomials alone
          // when every phase is below 0.78 rad, as split-step phases are),
          // instead of a branch around each point's sincos.  Same arithmetic
          // as half_step / two_half_steps, so the same bits.
          R e[M];
          bool small = true;
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (TF::valid(t, m, n)) {
              e[m] = phase_exponent<kMerged>(v[m], point_phase(m), s);
              small = small && fabs(e[m]) < R(0.78);
            }
          }
          auto finish = [&](auto fast_tag) {
#pragma unroll
            for (int m = 0; m < M; ++m) {
              if (TF::valid(t, m, n)) {
                R sn, cs;
                if constexpr (decltype(fast_tag)::value) sincos_poly(e[m], &sn, &cs);
                else phase_sincos(e[m], &sn, &cs);
                keep(m, rotate_by(v[m], sn, cs));
              }
            }
          };
          if (small) finish(std::true_type{});
          else finish(std::false_type{});
        } else {
#pragma unroll
          for (int m = 0; m < M; ++m) {
            if (TF::valid(t, m, n)) {
              const R ph = point_phase(m);
              R vi = R(0);
              if constexpr (VIM) vi = gload(s.vim + (g.base + (long long)TF::pos(t, m) * stride));
              if constexpr (kMerged) keep(m, two_half_steps<VIM>(v[m], ph, vi, s));
              else keep(m, half_step<IMAG, VIM>(v[m], ph, vi, s));
            }
          }
        }
      };
      const bool vol = s.vre != nullptr || s.vdyn != nullptr;
      auto with_vol = [&](auto imag_tag, auto vim_tag) {
        if (vol) local(imag_tag, vim_tag, std::true_type{});
        else local(imag_tag, vim_tag, std::false_type{});
      };
      // merged half-steps are real-time only; an absorber exists only in real time
      auto half_steps = [&]() {
        if (!kMerged && s.imag) {
          if constexpr (!kMerged) with_vol(std::true_type{}, std::false_type{});
        } else if (has_vim) {
          with_vol(std::false_type{}, std::true_type{});
        } else {
          with_vol(std::false_type{}, std::false_type{});
        }
      };
      if constexpr (kLocalZ) {
        if constexpr (KIND == K_ZK_FIRST) {
          if (prescale) {
#pragma unroll
            for (int m = 0; m < M; ++m) v[m] = scale_r(v[m], sc);
          }
        }
        half_steps();
      }
      TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
      // The axis's factor.  Imaginary time scales by a real factor (hi + lo:
      // one significant product, so one rounding).  Real time rotates: the
      // product is formed in FP64 from the FP64 table and rounded once
      // (cmul_d).  The hi + lo complex product rounds twice (the cross term,
      // then the sum), and rounding a rotation by a tiny angle twice leans
      // outward: the complex64 norm of the 10000-step 128^3 quench (dt =
      // 5e-5) drifted +2.5e-5 that way, against -4e-7 rounded once, which
      // is what the four-sweep plan's single 3D rotation gives.
      if (s.imag) {
#pragma unroll
        for (int m = 0; m < M; ++m)
          if (TF::valid(t, m, n)) v[m] = scale_re(v[m], kin1_at(kin1_index(t, m)));
      } else if (kin1_staged) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (TF::valid(t, m, n)) v[m] = cmul_d(v[m], kin1_rt_lds[kin1_index(t, m)]);
          if (m % CHUNK == CHUNK - 1) __builtin_amdgcn_sched_barrier(0);
        }
      } else {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          if (TF::valid(t, m, n)) v[m] = cmul_d(v[m], gload(s.kin1 + TF::pos(t, m)));
          if (m % CHUNK == CHUNK - 1) __builtin_amdgcn_sched_barrier(0);
        }
      }
      TF::template fft<+1, kRounds, LAX>(v, lp, t, twp, n, group);
      if constexpr (kLocalX) half_steps();
    } else if constexpr (KIND == K_Z_DIAG) {
      // Real-space FP64 sums at the diagnostic slot: |psi|^2, V|psi|^2, |psi|^4.
      if (prescale) {
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = scale_r(v[m], sc);
      }
      const long long rowoff = g.base;
      const double pxa = double(gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)pi * n + pj)));
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (TF::valid(t, m, n)) {
          const int k = TF::pos(t, m);
          double pot = pxa + double(gload(s.pz + (a.slot_a * s.pz_ss + k)));
          if (s.vre) pot += double(gload(s.vre + (rowoff + k)));
          if (s.vdyn) pot += double(gload(s.vdyn + (a.slot_a * s.vdyn_ss + rowoff + k)));
          const double a2 = double(v[m].x) * double(v[m].x) + double(v[m].y) * double(v[m].y);
          acc[0] += a2;
          acc[1] += pot * a2;
          acc[2] += a2 * a2;
        }
      }
      TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
    } else if constexpr (KIND == K_Y_LZ || KIND == K_X_LZ) {
      // Angular momentum by Parseval along one axis:
      //   sum_j psi^* (-i d/dy) psi = (1/n) sum_ky k_y |FFT_y psi|^2.
      if (prescale) {
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = scale_r(v[m], sc);
      }
      TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
      const double coord = KIND == K_Y_LZ ? gload(s.xarr + pi) : gload(s.xarr + pj);
      double part = 0.0;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (TF::valid(t, m, n)) {
          const double f2 = double(v[m].x) * double(v[m].x) + double(v[m].y) * double(v[m].y);
          part += gload(s.karr + TF::pos(t, m)) * f2;
        }
      }
      acc[0] += coord * part;
      write = false;
    } else if constexpr (KIND == K_X_DIAG) {
      TF::template fft<-1, kRounds, LAX>(v, lp, t, twp, n, group);
      const double ky = gload(s.karr + pj), kz = gload(s.karr + pk);
      const int fq = freq_index(pj, n), fr = freq_index(pk, n);
      const float view_scale = s.nview ? float(s.nview) / (1.7320508f * float(s.vh)) : 0.f;
#pragma unroll
      for (int m = 0; m < M; ++m) {
        if (TF::valid(t, m, n)) {
          const int i = TF::pos(t, m);
          const double kx = gload(s.karr + i);
          const double k2 = kx * kx + ky * ky + kz * kz;
          const double f2 = double(v[m].x) * double(v[m].x) + double(v[m].y) * double(v[m].y);
          acc[0] += f2 * k2;
          acc[1] += f2;
          if (spectrum) {
            const int fp = freq_index(i, n);
            const int sh = shell_index(fp, fq, fr);
            atomicAdd(&hocc[sh], f2);
            atomicAdd(&hcnt[sh], 1u);
            if (s.nview) {
              const int vb = s.nbins + view_index(fp, fq, fr, s.vh, s.nview, view_scale);
              atomicAdd(&hocc[vb], f2);
              atomicAdd(&hcnt[vb], 1u);
            }
          }
        }
      }
      write = false;
    }

    if (write) {
      if constexpr (TF::kFast) {
        const unsigned off = (g.lane + unsigned(t) * unsigned(stride)) * unsigned(sizeof(cpx<R>));
#pragma unroll
        for (int m = 0; m < M; ++m) gstore(tile_elem<TF>(dst + g.ubase, off, m, stride), v[m]);
      } else {
        const long long base = g.base;
#pragma unroll
        for (int m = 0; m < M; ++m)
          if (TF::valid(t, m, n)) gstore(dst + base + (long long)TF::pos(t, m) * stride, v[m]);
      }
    }
  }

  // ---- per-block reductions
  const unsigned nblk = gridDim.x;
  const unsigned bid = blockIdx.x;
  if constexpr (KIND == K_Z_LAST || KIND == K_XK_LAST) {
    if (a.flags & F_NORM) {
      // FP64 block partial only; k_norm_finish (launched right after on the
      // same stream) sums the partials.  An agent-scope fence + ticket here
      // would write back the XCD's L2 once per block, which costs more than
      // the tile itself on small grids.  XK_LAST sums through the exchange
      // buffer, free after its last transform (block_sum's first barrier
      // orders the exchange reads before the sums' writes), which keeps the
      // pass within its LDS budget.
      double* red_scratch;
      if constexpr (KIND == K_XK_LAST) {
        red_scratch = reinterpret_cast<double*>(lds);
      } else {
        __shared__ double red_lds[16];
        red_scratch = red_lds;
      }
      double part[1] = {acc[0]};
      block_sum<1>(part, red_scratch);
      double* red = a.red ? a.red : s.red;
      if (threadIdx.x == 0) red[bid] = part[0];
    }
  } else if constexpr (KIND == K_Z_DIAG) {
    __shared__ double red_scratch[48];
    block_sum<3>(acc, red_scratch);
    if (threadIdx.x == 0) {
      s.red[bid * 3 + 0] = acc[0];
      s.red[bid * 3 + 1] = acc[1];
      s.red[bid * 3 + 2] = acc[2];
    }
  } else if constexpr (KIND == K_Y_LZ || KIND == K_X_LZ) {
    __shared__ double red_scratch[16];
    double part[1] = {acc[0]};
    block_sum<1>(part, red_scratch);
    if (threadIdx.x == 0) s.red[bid] = part[0];
  } else if constexpr (KIND == K_X_DIAG) {
    __shared__ double red_scratch[32];
    double part[2] = {acc[0], acc[1]};
    block_sum<2>(part, red_scratch);
    if (threadIdx.x == 0) {
      s.red[bid * 2 + 0] = part[0];
      s.red[bid * 2 + 1] = part[1];
    }
    const int nb = spectrum ? s.nbins + s.nview : 0;
    if (nb) __syncthreads();
    for (int q = threadIdx.x; q < nb; q += blockDim.x) {
      s.nkocc[(long long)bid * nb + q] = hocc[q];
      s.nkcnt[(long long)bid * nb + q] = hcnt[q];
    }
  }
  if (stamp && (threadIdx.x & 63) == 0)
    atomicMax(s.stamps + a.stamp_index + blockIdx.x, (unsigned long long)wall_clock64());
}

// Sums the Z_LAST norm partials of one solve per block (blockIdx.x = member),
// in a fixed order, and publishes norm, 1/sqrt(norm) and the per-step norm.
template <class R>
__global__ void __launch_bounds__(256) k_norm_finish(PassArgs<R> a, int nblk) {
  const SolveDev<R>& s = a.s[blockIdx.x];
  __shared__ double red_scratch[16];
  double tot[1] = {0.0};
  const double* red = a.red ? a.red : s.red;
  for (int q = threadIdx.x; q < nblk; q += blockDim.x) tot[0] += red[q];
  block_sum<1>(tot, red_scratch);
  if (threadIdx.x == 0) {
    const double norm = s.dx3 * tot[0];
    s.scal[0] = norm;
    s.scal[1] = 1.0 / sqrt(norm);
    if (s.hist) s.hist[a.hist_index] = norm;
  }
}

}  // namespace gpe
