// Plane passes: a step's z and y axes in one sweep of psi.
//
// The separable plan (gpe/make_propagator.py:38-101 restated as one FFT ->
// 1D factor -> inverse FFT pass per axis) reads and writes psi once per axis.
// Where one x plane of the field (n^2 values) fits a CU's LDS (complex64 at
// sides 32, 48, 64, 96 and 128, complex128 at 32, 48 and 64), one block
// takes the whole plane
// and runs ZK_FIRST / ZK_MID and then YK on it: the z pencils are loaded,
// given their local half-step(s), transformed along z, scaled by the kz
// factor and transformed back; the plane then turns through LDS so that the
// same threads hold y pencils, which take FFT_y, the ky factor and the
// inverse FFT before the one store.  The sweeps per step fall from three to
// two, and so do the launches, which is what small grids are bound by.
//
// With the momentum cutoff the kinetic factor is not separable, and the step
// keeps psi transformed along y and z between its passes (the four-sweep
// plan's Y_FWD ... Y_INV): the ZY kinds run the inverse y and z transforms,
// the half-step(s) and the forward z and y transforms of one plane, and X_KIN
// applies the 3D factor between them.  A step then sweeps psi twice instead
// of four times.  (The cutoff applies to real time only, so the ZY kinds
// have no norm.)
//
// Every value goes through the same operations, in the same order, as in the
// passes a plane pass replaces (same transforms, twiddles, factors and
// half-step arithmetic; the LDS only moves values), so a step matches the
// separable or four-sweep plan's to rounding (the compiler may contract a
// product into an FMA differently in the two kernels).
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
// K_ZY_FIRST:  [scale] -> half(slot_a) -> FFT_z -> FFT_y
// K_ZY_MID:    IFFT_y -> IFFT_z -> half(slot_a) -> half(slot_b) -> FFT_z -> FFT_y
// K_ZY_LAST:   IFFT_y -> IFFT_z -> half(slot_a)
// blockIdx.x = x plane, blockIdx.z = solve of the batch.  A field that is
// transformed along y and z (the ZY kinds' input or output) is held in the
// canonical layout, as the four-sweep plan's Y_FWD leaves it.
template <typename R, int N, int TP, int KIND>
__global__ __launch_bounds__(TP * N) void k_plane(PassArgs<R> a) {
  using TF = PlaneTF<R, N, TP>;
  constexpr int M = TF::M;
  static_assert(is_plane_kind(KIND), "plane kinds only");
  static_assert(64 % TP == 0, "a z pencil's threads sit in one wave");
  constexpr bool kKin = KIND == K_ZYK_FIRST || KIND == K_ZYK_MID;   // 1D factors inside the pass
  constexpr bool kMerged = KIND == K_ZYK_MID || KIND == K_ZY_MID;
  constexpr bool kFirst = KIND == K_ZYK_FIRST || KIND == K_ZY_FIRST;
  constexpr bool kFromK = KIND == K_ZY_MID || KIND == K_ZY_LAST;     // input transformed along y and z
  constexpr bool kToReal = KIND == K_ZY_LAST;                         // output in real space along y, z

  __shared__ cpx<R> lds[plane_lds_elems<R, N, TP>()];
  __shared__ typename TF::TwT twl[TF::TWL_ELEMS];
  __shared__ cpx<double> kin1_raw[kKin ? N / 2 + 1 : 1];   // FP64 (real time) or cconst<R> (imaginary time)
  __shared__ R line[N];
  cconst<R>* kin1_c = reinterpret_cast<cconst<R>*>(kin1_raw);
  static_assert(sizeof(cconst<R>) == sizeof(cpx<double>), "both factor tables take 16 bytes an entry");

  const SolveDev<R>& s = a.s[blockIdx.z];
  const int x = blockIdx.x;
  const bool stamp = (a.flags & F_STAMP) != 0;

  // block-shared tables; stage_twiddles' barrier covers them all
  if constexpr (kKin) {
    for (int e = threadIdx.x; e <= N / 2; e += TF::THREADS) {
      const cpx<double> f = gload(s.kin1 + e);
      if (s.imag) kin1_c[e] = make_cconst<R>(f.x, f.y);
      else kin1_raw[e] = f;
    }
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
  // Field accesses take a scalar base per element row and one 32-bit byte
  // offset per thread (tile_elem), as the pencil passes do: a 64-bit vector
  // address per element would hold two VGPRs each, and the ZY kinds, which
  // load and store y pencils at the same offsets, would keep them live
  // across the whole pass.
  // z pencils: thread (yz, tz) holds z = tz + TP m of row y = yz;
  // y pencils: thread (zy, ty) holds y = ty + TP m of column z = zy (adjacent
  // lanes, adjacent z, as in the Y passes)
  const int tz = threadIdx.x % TP, yz = threadIdx.x / TP;
  const int zy = threadIdx.x % N, ty = threadIdx.x / N;
  const unsigned zoff = unsigned(yz * N + tz) * unsigned(sizeof(cpx<R>));   // element m: + TP m
  const unsigned yoff = unsigned(zy + ty * N) * unsigned(sizeof(cpx<R>));   // element m: + TP m N
  cpx<R>* const lpz = lds + yz * TF::template pencil_stride<2>();
  cpx<R>* const lpy = lds + zy * TF::template pencil_stride<1>();
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
  // Turn the plane through LDS: row y of N values along z, rows N + TP apart
  // (a wave's lanes then touch 64 distinct elements modulo the 32 that span
  // the banks, and the y pencils read consecutive z).  The leading barrier
  // orders the transform's last exchange reads before these writes; the next
  // transform's first barrier orders these reads before its writes.
  auto z_to_y = [&](cpx<R>* v) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) lds[yz * (N + TP) + TF::pos(tz, m)] = v[m];
    __syncthreads();
    // ZY_MID wrote these same elements in its first turn (y_to_z).  Their
    // addresses beyond the 64 KiB reach of a ds offset take one VGPR each, and
    // the compiler kept them from the first turn to here (at 128 complex64,
    // ten of them, spilled: 5 MiB of scratch writes per launch).  The empty
    // asm hides that this column is the same value, so they are formed again.
    int col = zy;
    if constexpr (KIND == K_ZY_MID) asm volatile("" : "+v"(col));
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = lds[TF::pos(ty, m) * (N + TP) + col];
  };
  auto y_to_z = [&](cpx<R>* v) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) lds[TF::pos(ty, m) * (N + TP) + zy] = v[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = lds[yz * (N + TP) + TF::pos(tz, m)];
  };

  // local half-step(s) of a z pencil: the pencil's (x, y) share plus the z
  // line, summed first and scaled by dt/2, then the phase (as the Z passes
  // form it)
  auto local_step = [&](cpx<R>* v) {
    const long long base = plane + (long long)yz * N;
    R pen = gload(s.pxy + (a.slot_a * s.pxy_ss + (long long)x * N + yz));
    if constexpr (kMerged) pen += gload(s.pxy + (a.slot_b * s.pxy_ss + (long long)x * N + yz));
    auto keep = [&](int m, cpx<R> w) { v[m] = w; };
    auto local = [&](auto imag_tag, auto vim_tag, auto vol_tag) {
      constexpr bool IMAG = decltype(imag_tag)::value;
      constexpr bool VIM = decltype(vim_tag)::value;
      constexpr bool VOL = decltype(vol_tag)::value;
      auto point_phase = [&](int m) -> R {
        const int idx = TF::pos(tz, m);
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
            keep(m, rotate_by(v[m], sn, cs));
          }
        };
        if (small) finish(std::true_type{});
        else finish(std::false_type{});
      } else {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const R ph = point_phase(m);
          R vi = R(0);
          if constexpr (VIM) vi = gload(s.vim + (base + TF::pos(tz, m)));
          if constexpr (kMerged) keep(m, two_half_steps<VIM>(v[m], ph, vi, s));
          else keep(m, half_step<IMAG, VIM>(v[m], ph, vi, s));
        }
      }
    };
    const bool vol = s.vre != nullptr || s.vdyn != nullptr;
    auto with_vol = [&](auto imag_tag, auto vim_tag) {
      if (vol) local(imag_tag, vim_tag, std::true_type{});
      else local(imag_tag, vim_tag, std::false_type{});
    };
    // the merged kinds are real-time only; an absorber exists only in real time
    if (!kMerged && s.imag) {
      if constexpr (!kMerged) with_vol(std::true_type{}, std::false_type{});
    } else if (s.vim != nullptr) {
      with_vol(std::false_type{}, std::true_type{});
    } else {
      with_vol(std::false_type{}, std::false_type{});
    }
  };

  cpx<R> v[M];
  if constexpr (kFromK) {
    // ---- y pencils from HBM, back to real space along y, then along z
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = gload(tile_elem<TF>(src + plane, yoff, m, N));
    TF::template fft<+1, 1, 1>(v, lpy, ty, twl, N);
    y_to_z(v);
    TF::template fft<+1, 1, 2>(v, lpz, tz, twl, N);
  } else {
    // ---- z pencils from HBM, in real space
#pragma unroll
    for (int m = 0; m < M; ++m) v[m] = gload(tile_elem<TF>(src + plane, zoff, m, 1));
    if constexpr (kFirst) {
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
  }
  local_step(v);
  if constexpr (kToReal) {
#pragma unroll
    for (int m = 0; m < M; ++m) gstore(tile_elem<TF>(dst + plane, zoff, m, 1), v[m]);
  } else {
    TF::template fft<-1, 1, 2>(v, lpz, tz, twl, N);
    if constexpr (kKin) {
      factor(v, tz);
      TF::template fft<+1, 1, 2>(v, lpz, tz, twl, N);
    }
    z_to_y(v);
    TF::template fft<-1, 1, 1>(v, lpy, ty, twl, N);
    if constexpr (kKin) {
      factor(v, ty);
      TF::template fft<+1, 1, 1>(v, lpy, ty, twl, N);
    }
#pragma unroll
    for (int m = 0; m < M; ++m) gstore(tile_elem<TF>(dst + plane, yoff, m, N), v[m]);
  }
  if (stamp && (threadIdx.x & 63) == 0)
    atomicMax(s.stamps + a.stamp_index + blockIdx.x, (unsigned long long)wall_clock64());
}

}  // namespace gpe
