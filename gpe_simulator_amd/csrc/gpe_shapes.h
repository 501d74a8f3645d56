// Pencil tile shapes per (precision, grid side).  TP threads cooperate on one
// pencil and hold M = N/TP values each; B pencils share a block.  For the
// strided X/Y passes the B pencils are adjacent along z, so every load
// instruction reads B*sizeof(complex) = 128 contiguous bytes per row.
#pragma once

#include "gpe_passes.h"

namespace gpe {

template <typename R, int N>
struct Shape;

// complex64
template <> struct Shape<float, 64> { using TF = FastTF<float, 64, 8, 32>; };
template <> struct Shape<float, 96> { using TF = FastTF<float, 96, 4, 32>; };
template <> struct Shape<float, 128> { using TF = FastTF<float, 128, 8, 32>; };
template <> struct Shape<float, 192> { using TF = FastTF<float, 192, 8, 16>; };
template <> struct Shape<float, 256> { using TF = FastTF<float, 256, 16, 16>; };
// complex128
template <> struct Shape<double, 64> { using TF = FastTF<double, 64, 8, 16>; };
template <> struct Shape<double, 96> { using TF = FastTF<double, 96, 4, 16>; };
template <> struct Shape<double, 128> { using TF = FastTF<double, 128, 8, 16>; };
template <> struct Shape<double, 192> { using TF = FastTF<double, 192, 8, 8>; };
template <> struct Shape<double, 256> { using TF = FastTF<double, 256, 16, 8>; };

// Sides with a fast instantiation; every other side 2..1024 uses GenTF.
#define GPE_FAST_SIDES(X) X(64) X(96) X(128) X(192) X(256)

}  // namespace gpe
