// Pencil tile variants per (precision, grid side).
//
// TP threads cooperate on one pencil and hold M = N/TP values each (N must be
// TP*TP*Q); B pencils share a block; W is the waves per SIMD the register
// allocation must allow.  For the strided
// X/Y passes the B pencils are adjacent along z, so every load instruction
// reads B*sizeof(complex) contiguous bytes per row; in Z passes a block's tile
// is B*N contiguous elements.  Each pass kind of a solve can run a different
// variant (all read and write the same canonical field), and the autotuner
// picks per kind; variant 0 is the default before tuning.  All variants of one
// (precision, side) share TP, hence the codelets, twiddles and order of every
// sum: they differ only in tiling, so a field's arithmetic, and its result to
// the last bit, does not depend on which variant the tuner picked.
//
// One line per variant: X(precision 0/1 = complex64/complex128, side, variant,
// TP, B, W, RD), RD the exchange rounds of the Z passes (2: half the block's
// pencils at a time through half the LDS; see FastTF::ROUNDS).  The build
// compiles one translation unit per line.
#pragma once

#include "gpe_passes.h"

#define GPE_VARIANTS(X)                                                                                   \
  X(0, 32, 0, 4, 16, 4, 1) X(0, 32, 1, 4, 32, 4, 1)                                                       \
  X(0, 48, 0, 4, 16, 4, 1) X(0, 48, 1, 4, 48, 4, 1)                                                       \
  X(0, 64, 0, 8, 32, 3, 1) X(0, 64, 1, 8, 16, 4, 1) X(0, 64, 2, 8, 32, 4, 1)                              \
  X(0, 96, 0, 4, 32, 3, 1) X(0, 96, 1, 4, 16, 3, 1) X(0, 96, 2, 4, 32, 4, 1) X(0, 96, 3, 4, 32, 3, 2)     \
  X(0, 128, 0, 8, 32, 3, 1) X(0, 128, 1, 8, 16, 3, 1) X(0, 128, 2, 8, 16, 4, 1)                           \
  X(0, 192, 0, 8, 16, 3, 1) X(0, 192, 1, 8, 32, 3, 1) X(0, 192, 2, 8, 16, 4, 1) X(0, 192, 3, 8, 16, 3, 2) \
  X(0, 256, 0, 16, 16, 3, 1) X(0, 256, 1, 16, 16, 4, 1) X(0, 256, 2, 16, 32, 3, 1)                        \
  X(0, 512, 0, 16, 8, 2, 1) X(0, 512, 1, 16, 16, 2, 1)                                                     \
  X(0, 1024, 0, 32, 4, 1, 1) X(0, 1024, 1, 32, 8, 1, 1)                                                    \
  X(1, 32, 0, 4, 16, 2, 1) X(1, 32, 1, 4, 32, 2, 1)                                                       \
  X(1, 48, 0, 4, 16, 2, 1) X(1, 48, 1, 4, 48, 2, 1)                                                       \
  X(1, 64, 0, 8, 16, 2, 1) X(1, 64, 1, 8, 32, 2, 1)                                                       \
  X(1, 96, 0, 4, 16, 2, 1) X(1, 96, 1, 4, 32, 2, 1) X(1, 96, 2, 4, 8, 2, 1) X(1, 96, 3, 4, 32, 2, 2)      \
  X(1, 128, 0, 8, 16, 2, 1) X(1, 128, 1, 8, 32, 2, 1) X(1, 128, 2, 8, 8, 2, 1)                            \
  X(1, 192, 0, 8, 8, 2, 1) X(1, 192, 1, 8, 16, 2, 1) X(1, 192, 2, 8, 16, 2, 2)                            \
  X(1, 256, 0, 16, 8, 2, 1) X(1, 256, 1, 16, 16, 2, 1) X(1, 256, 2, 16, 4, 2, 1)                          \
  X(1, 512, 0, 16, 4, 1, 1) X(1, 512, 1, 16, 8, 1, 1)                                                      \
  X(1, 1024, 0, 32, 2, 1, 1) X(1, 1024, 1, 32, 4, 1, 1)

namespace gpe {

template <int P>
using prec_t = typename std::conditional<P == 0, float, double>::type;

template <int P, int N, int V>
struct Shape;

#define GPE_SHAPE_SPEC(P, N, V, TP, B, W, RD)        \
  template <>                                         \
  struct Shape<P, N, V> {                             \
    using TF = FastTF<prec_t<P>, N, TP, B, W, RD>;    \
  };
GPE_VARIANTS(GPE_SHAPE_SPEC)
#undef GPE_SHAPE_SPEC

}  // namespace gpe
