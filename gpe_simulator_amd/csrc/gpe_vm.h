// Postfix program format for potential expressions.
//
// A potential expression of the schema-2 grammar (numbers, x y z pi [t],
// + - * / **, unary +/-, sin cos tan exp sqrt tanh abs log) is compiled on the
// host into this postfix code and evaluated in FP64 on the device, once per
// table entry.  The propagation passes never run this interpreter: they read
// the tables it fills (per (x,y) column, per z, or per volume point).
#pragma once

namespace gpe {

enum VmOp : int {
  OP_CONST = 1,  // followed by an index into the constant array
  OP_X = 2,
  OP_Y = 3,
  OP_Z = 4,
  OP_T = 5,
  OP_ADD = 10,
  OP_SUB = 11,
  OP_MUL = 12,
  OP_DIV = 13,
  OP_POW = 14,
  OP_NEG = 15,
  OP_MAX = 16,   // used by the built-in box and absorber programs
  OP_SIN = 20,
  OP_COS = 21,
  OP_TAN = 22,
  OP_EXP = 23,
  OP_SQRT = 24,
  OP_TANH = 25,
  OP_FABS = 26,
  OP_LOG = 27,
};

constexpr int kVmStack = 32;
constexpr int kVmMaxCode = 4096;

// Table targets.
enum Target : int {
  TGT_XY = 0,    // n*n entries indexed i*n + j  (x = X[i], y = X[j])
  TGT_Z = 1,     // n entries indexed k          (z = X[k])
  TGT_VOL = 2,   // n^3 entries, real part of the static volume
  TGT_VOL_IM = 3 // n^3 entries, imaginary part of the static volume
};

}  // namespace gpe
