// Native evaluation of time-dependent potential programs.
//
// A postfix program (gpe_vm.h) is turned into HIP C++ source and compiled at
// run time with hiprtc for the current device, the way the reference inlines
// its expressions into kernels it compiles at run time
// (gpe/make_kernel.py:4-84, through CuPy's RawKernel).  The source depends on
// the program's structure only: its constants are read from the program's
// device array, so programs that differ in their numbers (stirrers of other
// speeds, traps of other strengths) share one compiled module.
//
// Every subexpression that does not depend on x, y or z (cos(Omega t), a
// product of constants) is evaluated once per time slot by a small kernel,
// and the per-point kernel reads it.  All arithmetic is FP64, operation by
// operation, without contraction into FMAs, with the same device math
// functions as the interpreter: the tables are bit-identical to the
// interpreter's.
#pragma once

#include <hip/hip_runtime.h>

namespace gpe {

struct JitProgram;

// The compiled program for `code` on the current device, compiling it on
// first use (a few tenths of a second); null when hiprtc is unavailable or
// the compile fails (the caller keeps the interpreter).
JitProgram* jit_get(const int* code, int ncode, int nconst);

// Hoisted subexpressions per time slot (the length of one slot's row of
// the `hv` scratch that jit_eval needs).
int jit_hoisted(const JitProgram* p);

// out[s][i] = R(master[i] + f(x_i, y_i, z_i, t_s)) for s < nslots, i < size,
// R = float (f32) or double, on `st`; `hv` holds nslots * jit_hoisted(p)
// doubles (unused when that is 0), `consts` the program's constants.
hipError_t jit_eval(const JitProgram* p, bool f32, const double* consts, int target, int n, const double* xa,
                    long long size, const double* master, const double* tlist, int nslots, double* hv, void* out,
                    hipStream_t st);

// Process-wide counters: modules compiled, table evaluations launched,
// compiles that failed.
void jit_stats(long long* compiles, long long* launches, long long* failures);

}  // namespace gpe
