// Plane passes: the separable step's z and y axes in one sweep of psi.
//
// The separable plan (gpe/make_propagator.py:38-101 restated as one FFT ->
// 1D factor -> inverse FFT pass per axis) reads and writes psi once per axis.
// Where one x plane of the field (n^2 values) fits a CU's LDS (complex64 at
// sides 64, 96 and 128, complex128 at 64), one block takes the whole plane
// and runs ZK_FIRST / ZK_MID and then YK on it: the z pencils are loaded,
// given their local half-step(s), transformed along z, scaled by the kz
// factor and transformed back; the plane then turns through LDS so that the
// same threads hold y pencils, which take FFT_y, the ky factor and the
// inverse FFT before the one store.  The sweeps per step fall from three to
// two, and so do the launches, which is what small grids are bound by.
//
// Every value goes through the same operations, in the same order, as in the
// ZK_* and YK passes (same transforms, twiddles, factors and half-step
// arithmetic; the LDS only moves values), so a step is bit-identical to the
// separable plan's.
#pragma once

#include "gpe_passes.h"

namespace gpe {

// Transform shape of a plane pass: the side's fast transform with one block
// per plane (B = N pencils of TP threads).
template <typename R, int N, int TP>
using PlaneTF = FastTF<R, N, TP, N, 1, 1>;

template <typename R, int N, int TP>
__host__ __device__ constexpr int plane_lds_elems() {
  using TF = PlaneTF<R, N, TP>;
  constexpr int z = TF::template lds_elems<2>();
  constexpr int y = TF::template lds_elems<1>();
  constexpr int p = N * (N + TP);
  return z > y ? (z > p ? z : p) : (y > p ? y : p);
}

// LDS bytes of a plane block: exchange / transpose buffer, twiddles, the 1D
// factor table (entries 0..N/2) and the z line of the potential.
template <typename R, int N, int TP>
__host__ __device__ constexpr int plane_lds_bytes() {
  using TF = PlaneTF<R, N, TP>;
  return plane_lds_elems<R, N, TP>() * int(sizeof(cpx<R>)) + TF::TWL_ELEMS * int(sizeof(typename TF::TwT)) +
         (N / 2 + 1) * 16 + N * int(sizeof(R));
}

// Whether a plane pass exists for (precision, side): the plane block must fit
// the 160 KiB of LDS.
__host__ __device__ constexpr bool plane_fits(int bytes) { return bytes <= 160 * 1024; }

// K_ZYK_FIRST: [scale] -> half(slot_a) -> FFT_z -> kz -> IFFT_z -> FFT_y -> ky -> IFFT_y
// K_ZYK_MID:   half(slot_a) -> half(slot_b) -> FFT_z -> kz -> IFFT_z -> FFT_y -> ky -> IFFT_y
// blockIdx.x = x plane, blockIdx.z = solve of the batch.
template <typename R, int N, int TP, int KIND>
__global__ __launch_bounds__(TP * N) void k_plane(PassArgs<R> a) {
  using TF = PlaneTF<R, N, TP>;
  constexpr int M = TF::M;
  constexpr bool kMerged = KIND == K_ZYK_MID;
  static_assert(KIND == K_ZYK_FIRST || KIND == K_ZYK_MID, "plane kinds only");
  static_assert(64 % TP == 0, "a z pencil's threads sit in one wave");

  __shared__ cpx<R> lds[plane_lds_elems<R, N, TP>()];
  __shared__ typename TF::TwT twl[TF::TWL_ELEMS];
  __shared__ cpx<double> kin1_raw[N / 2 + 1];   // FP64 (real time) or cconst<R> (imaginary time)
  __shared__ R line[N];
  cconst<R>* kin1_c = reinterpret_cast<cconst<R>*>(kin1_raw);
  static_assert(sizeof(cconst<R>) == sizeof(cpx<double>), "both factor tables take 16 bytes an entry");

  const SolveDev<R>& s = a.s[blockIdx.z];
  const int x = blockIdx.x;
  const bool stamp = (a.flags & F_STAMP) != 0;

  // block-shared tables; stage_twiddles' barrier covers them all
  for (int e = threadIdx.x; e <= N / 2; e += TF::THREADS) {
    const cpx<double> f = gload(s.kin1 + e);
    if (s.imag) kin1_c[e] = make_cconst<R>(f.x, f.y);
    else kin1_raw[e] = f;
  }
  for (int e = threadIdx.x; e < N; e += TF::THREADS) {
    R pz = gload(s.pz + (a.slot_a * s.pz_ss + e));
    if constexpr (kMerged) pz += gload(s.pz + (a.slot_b * s.pz_ss + e));
    line[e] = pz;
  }
  TF::stage_twiddles(twl, s.tw);

  const cpx<R>* src = (a.flags & F_IN_WORK) ? s.work : s.psi;
  cpx<R>* dst = (a.flags & F_OUT_WORK) ? s.work : s.psi;
  const long long plane = (long long)x * N * N;
  // |frequency| of element m of thread t (entries 0..N/2 of the table)
  auto kin1_index = [](int t_, int m) { return m < M / 2 ? TF::pos(t_, m) : N - TF::pos(t_, m); };
  auto factor = [&](cpx<R>* v, int t_) {
    if (s.imag) {
#pragma unroll
      for (int m = 0; m < M; ++m) v[m] = scale_re(v[m], kin1_c[kin1_index(t_, m)]);
    } else {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        v[m] = cmul_d(v[m], kin1_raw[kin1_index(t_, m)]);
        if (m % 4 == 3) __builtin_amdgcn_sched_barrier(0);
      }
    }
  };

  cpx<R> v[M];
  {
    // ---- z pencils: thread (y, t) holds z = t + TP m
    const int t = threadIdx.x % TP, y = threadIdx.x / TP;
    const long long base = plane + (long long)y * N;
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = gload(src + base + TF::pos(t, m));
    if constexpr (KIND == K_ZYK_FIRST) {
      if (a.flags & F_PRESCALE) {
        // the previous step's norm from its partials (F_NORM_FOLD), or the
        // finisher's scale
        __shared__ double fold_lds[5];
        const rconst<R> sc =
            make_rconst<R>((a.flags & F_NORM_FOLD) ? fold_norm(a, s, fold_lds) : s.scal[1]);
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = scale_r(v[m], sc);
      }
    }
    // local half-step(s): the pencil's (x, y) share plus the z line, summed
    // first and scaled by dt/2, then the phase (as the ZK_* passes form it)
    R pen = gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)x * N + y));
    if constexpr (kMerged) pen += gload(s.pxy + (a.slot_b * s.pxy_ss + (long long)x * N + y));
    auto local = [&](auto imag_tag, auto vim_tag, auto vol_tag) {
      constexpr bool IMAG = decltype(imag_tag)::value;
      constexpr bool VIM = decltype(vim_tag)::value;
      constexpr bool VOL = decltype(vol_tag)::value;
      auto point_phase = [&](int m) -> R {
        const int idx = TF::pos(t, m);
        R ph = pen + line[idx];
        if constexpr (VOL) {
          const long long vidx = base + idx;
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
      if constexpr (!IMAG && !VIM) {
        R e[M];
        bool small = true;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          e[m] = phase_exponent<kMerged>(v[m], point_phase(m), s);
          small = small && fabs(e[m]) < R(0.78);
        }
        auto finish = [&](auto fast_tag) {
#pragma unroll
          for (int m = 0; m < M; ++m) {
            R sn, cs;
            if constexpr (decltype(fast_tag)::value) sincos_poly(e[m], &sn, &cs);
            else phase_sincos(e[m], &sn, &cs);
            v[m] = rotate_by(v[m], sn, cs);
          }
        };
        if (small) finish(std::true_type{});
        else finish(std::false_type{});
      } else {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const R ph = point_phase(m);
          R vi = R(0);
          if constexpr (VIM) vi = gload(s.vim + (base + TF::pos(t, m)));
          if constexpr (kMerged) v[m] = two_half_steps<VIM>(v[m], ph, vi, s);
          else v[m] = half_step<IMAG, VIM>(v[m], ph, vi, s);
        }
      }
    };
    const bool vol = s.vre != nullptr || s.vdyn != nullptr;
    auto with_vol = [&](auto imag_tag, auto vim_tag) {
      if (vol) local(imag_tag, vim_tag, std::true_type{});
      else local(imag_tag, vim_tag, std::false_type{});
    };
    if (!kMerged && s.imag) {
      if constexpr (!kMerged) with_vol(std::true_type{}, std::false_type{});
    } else if (s.vim != nullptr) {
      with_vol(std::false_type{}, std::true_type{});
    } else {
      with_vol(std::false_type{}, std::false_type{});
    }
    cpx<R>* lp = lds + y * TF::template pencil_stride<2>();
    TF::template fft<-1, 1, 2>(v, lp, t, twl, N);
    factor(v, t);
    TF::template fft<+1, 1, 2>(v, lp, t, twl, N);
    // ---- turn the plane: row y of N values along z, rows N + TP apart (a
    // wave's lanes then write 64 distinct elements modulo the 32 that span
    // the banks, and the y pencils below read consecutive z)
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) lds[y * (N + TP) + TF::pos(t, m)] = v[m];
    __syncthreads();
  }
  {
    // ---- y pencils: thread (z, t) holds y = t + TP m (adjacent lanes,
    // adjacent z, as in the Y passes)
    const int z = threadIdx.x % N, t = threadIdx.x / N;
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = lds[TF::pos(t, m) * (N + TP) + z];
    // (the transform's first barrier orders these reads before its writes)
    cpx<R>* lp = lds + z * TF::template pencil_stride<1>();
    TF::template fft<-1, 1, 1>(v, lp, t, twl, N);
    factor(v, t);
    TF::template fft<+1, 1, 1>(v, lp, t, twl, N);
    const long long base = plane + z;
#pragma unroll
    for (int m = 0; m < M; ++m) gstore(dst + base + (long long)TF::pos(t, m) * N, v[m]);
  }
  if (stamp && (threadIdx.x & 63) == 0)
    atomicMax(s.stamps + a.stamp_index + blockIdx.x, (unsigned long long)wall_clock64());
}

}  // namespace gpe
