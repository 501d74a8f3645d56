// One translation unit per (precision, grid size): the build compiles this
// file once for each GPE_PREC (0 = complex64, 1 = complex128) and GPE_N
// (0 = generic direct-DFT transform, otherwise a fast pencil shape), so the
// heavy template instantiations compile in parallel.
#include "gpe_passes.h"
#include "gpe_shapes.h"

#ifndef GPE_PREC
#error "GPE_PREC must be defined"
#endif
#ifndef GPE_N
#error "GPE_N must be defined"
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
using TF = typename Shape<Real, GPE_N>::TF;
#endif

#define GPE_CAT2(a, b) a##b
#define GPE_CAT(a, b) GPE_CAT2(a, b)
#define GPE_LAUNCHER GPE_CAT(GPE_CAT(GPE_CAT(launch_pass_p, GPE_PREC), _n), GPE_N)

template <int AXIS, int KIND>
static void go(dim3 grid, hipStream_t st, const PassArgs<Real>& a) {
  hipLaunchKernelGGL((k_pass<TF, AXIS, KIND, Real>), grid, dim3(TF::THREADS), 0, st, a);
}

// Launches one pass over `batch` solves of side n sharing one descriptor array.
hipError_t GPE_LAUNCHER(int kind, const void* args, int n, int batch, hipStream_t st) {
  const PassArgs<Real>& a = *static_cast<const PassArgs<Real>*>(args);
  const dim3 gz((unsigned)((long long)n * n / TF::B), 1, (unsigned)batch);
  const dim3 gxy((unsigned)(n / TF::B), (unsigned)n, (unsigned)batch);
  switch (kind) {
    case K_Z_FIRST: go<2, K_Z_FIRST>(gz, st, a); break;
    case K_Z_MID: go<2, K_Z_MID>(gz, st, a); break;
    case K_Z_LAST: go<2, K_Z_LAST>(gz, st, a); break;
    case K_Z_DIAG: go<2, K_Z_DIAG>(gz, st, a); break;
    case K_Z_FWD: go<2, K_Z_FWD>(gz, st, a); break;
    case K_Y_FWD: go<1, K_Y_FWD>(gxy, st, a); break;
    case K_Y_INV: go<1, K_Y_INV>(gxy, st, a); break;
    case K_X_KIN: go<0, K_X_KIN>(gxy, st, a); break;
    case K_X_DIAG: go<0, K_X_DIAG>(gxy, st, a); break;
    case K_Y_LZ: go<1, K_Y_LZ>(gxy, st, a); break;
    case K_X_LZ: go<0, K_X_LZ>(gxy, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Number of blocks one pass of `kind` uses per solve (for partial buffers).
long long GPE_CAT(GPE_CAT(GPE_CAT(pass_blocks_p, GPE_PREC), _n), GPE_N)(int kind, int n) {
  if (kind == K_Z_FIRST || kind == K_Z_MID || kind == K_Z_LAST || kind == K_Z_DIAG || kind == K_Z_FWD)
    return (long long)n * n / TF::B;
  return (long long)(n / TF::B) * n;
}

}  // namespace gpe
