// One translation unit per (precision, grid size, variant): the build compiles
// this file once for each GPE_PREC (0 = complex64, 1 = complex128), GPE_N
// (0 = generic direct-DFT transform, otherwise a fast side) and GPE_V (tile
// variant from the list in gpe_shapes.h), so the heavy template
// instantiations compile in parallel.
#include <hip/hip_ext.h>

#include "gpe_passes.h"
#include "gpe_plane.h"
#include "gpe_shapes.h"

#ifndef GPE_PREC
#error "GPE_PREC must be defined"
#endif
#ifndef GPE_N
#error "GPE_N must be defined"
#endif
#ifndef GPE_V
#error "GPE_V must be defined"
#endif

namespace gpe {

#if GPE_PREC == 0
using Real = float;
#else
using Real = double;
#endif

#if GPE_N == 0
using TF = GenTF<Real>;
#else
using TF = typename Shape<GPE_PREC, GPE_N, GPE_V>::TF;
#endif

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)
#define GPE_SUFFIX GPE_CAT(GPE_CAT(GPE_CAT(GPE_CAT(GPE_CAT(p, GPE_PREC), _n), GPE_N), _v), GPE_V)
#define GPE_LAUNCHER GPE_CAT(launch_pass_, GPE_SUFFIX)

// Blocks of a launch that fit one CU's LDS (160 KiB, allocated in 512-byte
// units) with `dyn` bytes of dynamic LDS beside the kernel's static LDS.
static int blocks_per_cu_by_lds(unsigned static_bytes, unsigned dyn) {
  const unsigned need = (static_bytes + dyn + 511u) / 512u * 512u;
  return need ? int((160u * 1024u) / need) : 1 << 20;
}

// F_KIN1_RT: a real-time separable pass stages its factor table in LDS when
// the table costs no blocks per CU (the kernel's static LDS is read once).
template <int AXIS, int KIND>
static int resolve_flags(int flags) {
  if (!(flags & F_KIN1_RT)) return flags;
  flags &= ~F_KIN1_RT;
  if constexpr (TF::kFast && is_kin1_kind(KIND)) {
    static const bool fits = [] {
      hipFuncAttributes attr{};
      if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(&k_pass<TF, AXIS, KIND, Real>)) != hipSuccess)
        return false;
      const unsigned s = unsigned(attr.sharedSizeBytes);
      const unsigned dyn = unsigned(TF::N / 2 + 1) * sizeof(cconst<Real>);
      return blocks_per_cu_by_lds(s, dyn) >= blocks_per_cu_by_lds(s, 0);
    }();
    if (fits) flags |= F_KIN1_LDS;
  }
  return flags;
}

template <int AXIS, int KIND>
static void go(int n, int batch, hipStream_t st, const PassArgs<Real>& a_in, hipEvent_t t0, hipEvent_t t1) {
  PassArgs<Real> a = a_in;
  a.flags = resolve_flags<AXIS, KIND>(a.flags);
  const dim3 grid((unsigned)tile_count<TF, AXIS>(n), 1, (unsigned)batch);
  const unsigned dyn = kin1_lds_bytes<TF, KIND, Real>(a.flags);
  if (t0 || t1)
    hipExtLaunchKernelGGL((k_pass<TF, AXIS, KIND, Real>), grid, dim3(TF::THREADS), dyn, st, t0, t1, 0, a);
  else
    hipLaunchKernelGGL((k_pass<TF, AXIS, KIND, Real>), grid, dim3(TF::THREADS), dyn, st, a);
}

static int axis_of(int kind) {
  switch (kind) {
    case K_Y_FWD: case K_Y_INV: case K_Y_LZ: case K_YK: return 1;
    case K_X_KIN: case K_X_DIAG: case K_X_LZ: case K_X_FWD: case K_X_INV: case K_XK: case K_XK_LAST: return 0;
    default: return 2;
  }
}

// Plane passes (one block per x plane) are compiled once per (precision,
// side), in the first variant's unit, where the plane fits the LDS.
template <class T>
static constexpr bool has_plane() {
  if constexpr (T::kFast) {
    return GPE_V == 0 && 64 % T::TP == 0 && T::TP * T::N <= 1024 &&
           plane_fits(plane_lds_bytes<Real, T::N, T::TP>());
  } else {
    return false;
  }
}

template <int KIND, class T = TF>
static void go_plane(int batch, hipStream_t st, const PassArgs<Real>& a, hipEvent_t t0, hipEvent_t t1) {
  if constexpr (has_plane<T>()) {
    constexpr int N = T::N, TP = T::TP;
    const dim3 grid((unsigned)N, 1, (unsigned)batch);
    if (t0 || t1)
      hipExtLaunchKernelGGL((k_plane<Real, N, TP, KIND>), grid, dim3(TP * N), 0, st, t0, t1, 0, a);
    else
      hipLaunchKernelGGL((k_plane<Real, N, TP, KIND>), grid, dim3(TP * N), 0, st, a);
  }
}

// A pass that writes FP64 norm partials, followed by k_norm_finish (one block
// per solve) on the same stream; the finisher carries the end event.
template <int AXIS, int KIND>
static void go_norm(int n, int batch, hipStream_t st, const PassArgs<Real>& a, hipEvent_t t0, hipEvent_t t1) {
  go<AXIS, KIND>(n, batch, st, a, t0, nullptr);
  const dim3 grid((unsigned)batch), block(256);
  const int nblk = (int)tile_count<TF, AXIS>(n);
  if (t1)
    hipExtLaunchKernelGGL((k_norm_finish<Real>), grid, block, 0, st, nullptr, t1, 0, a, nblk);
  else
    hipLaunchKernelGGL((k_norm_finish<Real>), grid, block, 0, st, a, nblk);
}

// Launches one pass over `batch` solves of side n sharing one descriptor
// array; one block per tile of B pencils.  With t0/t1 the dispatch itself
// timestamps its start and end into the events (the norm finisher, when it
// follows Z_LAST, carries the end event).
hipError_t GPE_LAUNCHER(int kind, const void* args, int n, int batch, hipStream_t st, hipEvent_t t0,
                        hipEvent_t t1) {
  const PassArgs<Real>& a = *static_cast<const PassArgs<Real>*>(args);
  switch (kind) {
    case K_Z_FIRST: go<2, K_Z_FIRST>(n, batch, st, a, t0, t1); break;
    case K_Z_MID: go<2, K_Z_MID>(n, batch, st, a, t0, t1); break;
    case K_Z_LAST:
      if ((a.flags & F_NORM) && !(a.flags & F_NORM_DEFER)) go_norm<2, K_Z_LAST>(n, batch, st, a, t0, t1);
      else go<2, K_Z_LAST>(n, batch, st, a, t0, t1);
      break;
    case K_ZK_FIRST: go<2, K_ZK_FIRST>(n, batch, st, a, t0, t1); break;
    case K_ZK_MID: go<2, K_ZK_MID>(n, batch, st, a, t0, t1); break;
    case K_YK: go<1, K_YK>(n, batch, st, a, t0, t1); break;
    case K_XK: go<0, K_XK>(n, batch, st, a, t0, t1); break;
    case K_XK_LAST:
      if ((a.flags & F_NORM) && !(a.flags & F_NORM_DEFER)) go_norm<0, K_XK_LAST>(n, batch, st, a, t0, t1);
      else go<0, K_XK_LAST>(n, batch, st, a, t0, t1);
      break;
    case K_Z_DIAG: go<2, K_Z_DIAG>(n, batch, st, a, t0, t1); break;
    case K_Z_FWD: go<2, K_Z_FWD>(n, batch, st, a, t0, t1); break;
    case K_Y_FWD: go<1, K_Y_FWD>(n, batch, st, a, t0, t1); break;
    case K_Y_INV: go<1, K_Y_INV>(n, batch, st, a, t0, t1); break;
    case K_X_KIN: go<0, K_X_KIN>(n, batch, st, a, t0, t1); break;
    case K_X_DIAG: go<0, K_X_DIAG>(n, batch, st, a, t0, t1); break;
    case K_Y_LZ: go<1, K_Y_LZ>(n, batch, st, a, t0, t1); break;
    case K_X_LZ: go<0, K_X_LZ>(n, batch, st, a, t0, t1); break;
    case K_Z_HALF: go<2, K_Z_HALF>(n, batch, st, a, t0, t1); break;
    case K_Z_INV: go<2, K_Z_INV>(n, batch, st, a, t0, t1); break;
    case K_X_FWD: go<0, K_X_FWD>(n, batch, st, a, t0, t1); break;
    case K_X_INV: go<0, K_X_INV>(n, batch, st, a, t0, t1); break;
    case K_ZYK_FIRST:
      if (!has_plane<TF>()) return hipErrorInvalidValue;
      go_plane<K_ZYK_FIRST>(batch, st, a, t0, t1);
      break;
    case K_ZYK_MID:
      if (!has_plane<TF>()) return hipErrorInvalidValue;
      go_plane<K_ZYK_MID>(batch, st, a, t0, t1);
      break;
    case K_ZY_FIRST:
      if (!has_plane<TF>()) return hipErrorInvalidValue;
      go_plane<K_ZY_FIRST>(batch, st, a, t0, t1);
      break;
    case K_ZY_MID:
      if (!has_plane<TF>()) return hipErrorInvalidValue;
      go_plane<K_ZY_MID>(batch, st, a, t0, t1);
      break;
    case K_ZY_LAST:
      if (!has_plane<TF>()) return hipErrorInvalidValue;
      go_plane<K_ZY_LAST>(batch, st, a, t0, t1);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Blocks per solve that a pass of `kind` launches (sizes partial buffers);
// 0 for a plane kind that this unit does not compile.
long long GPE_CAT(pass_blocks_, GPE_SUFFIX)(int kind, int n) {
  if (is_plane_kind(kind)) return has_plane<TF>() ? n : 0;
  const int axis = axis_of(kind);
  return axis == 2 ? tile_count<TF, 2>(n) : (axis == 1 ? tile_count<TF, 1>(n) : tile_count<TF, 0>(n));
}

}  // namespace gpe
