// Host-side declarations of the launchers in gpe_misc.hip and the pass
// launchers compiled from gpe_pass_inst.hip.
#pragma once

#include "gpe_shapes.h"

#include "gpe_passes.h"
#include "gpe_vm.h"

namespace gpe {

constexpr int kNormBlocks = 1024;

template <typename R> void launch_twiddle(void* tw, int n, hipStream_t st);
template <typename R>
void launch_kinetic(void* kx, void* ky, void* kz, void* k1, const double* karr, int n, double dt, double coef,
                    int imag, double inv_n3, hipStream_t st);
void launch_vm_static(const int* code, int ncode, const double* consts, int target, int n, const double* xa,
                      long long size, double scale, double* master, hipStream_t st);
template <typename R>
void launch_vm_dynamic(const int* code, int ncode, const double* consts, int target, int n, const double* xa,
                       long long size, const double* master, const double* tlist, int nslots, void* out,
                       hipStream_t st);
template <typename R> void launch_cast(const double* in, void* out, long long size, hipStream_t st);
void launch_fill_f64(double* out, long long size, double v, hipStream_t st);
void launch_add_f64(double* out, const double* in, long long size, hipStream_t st);
template <typename R> void launch_scale(void* psi, long long count, const double* scal, hipStream_t st);
template <typename R> void launch_gaussian(void* psi, const double* xa, int n, double sigma, hipStream_t st);
template <typename R>
void launch_kinetic_points(void* psi, const void* kx, const void* ky, const void* kz, const double* karr, int n,
                           int masked, double kcut2, hipStream_t st);
template <typename R> void launch_from_c128(const void* in, void* out, long long count, hipStream_t st);
template <typename R> void launch_to_c128(const void* in, void* out, long long count, hipStream_t st);
template <typename R>
void launch_norm(const void* psi, long long count, double* red, unsigned int* counter, double* scal, double dx3,
                 int prescale, hipStream_t st);
template <typename R> void launch_abs2_sum(const void* f, int n, int axis, double* out, hipStream_t st);
template <typename R>
void launch_plane(const void* f, int n, int axis, int index, void* out_c128, double* out_abs2, hipStream_t st);
void launch_finish_sums(const double* red, long long nblk, int nq, double* out, hipStream_t st);
void launch_finish_nk(const double* occ, const unsigned int* cnt, long long nblk, int nbins, double* out_occ,
                      long long* out_cnt, hipStream_t st);
void launch_copy(const void* in, void* out, long long bytes, hipStream_t st);
void launch_clear_stamps(unsigned long long* stamps, long long count, hipStream_t st);
void launch_stamp_now(unsigned long long* slot, hipStream_t st);

// Pass launchers: one per (precision, side, variant) instantiation; side 0 is
// the generic transform (one variant).
// t0 / t1 (nullable): events the launch's own dispatch timestamps at its start
// and end (hipExtLaunchKernel), for per-pass timing.
using PassLauncher = hipError_t (*)(int kind, const void* args, int n, int batch, hipStream_t st, hipEvent_t t0,
                                    hipEvent_t t1);
using PassBlocks = long long (*)(int kind, int n);

#define GPE_DECL_PASS(P, N, V, ...)                                                                 \
  hipError_t launch_pass_p##P##_n##N##_v##V(int kind, const void* args, int n, int batch, hipStream_t st,     \
                                            hipEvent_t t0, hipEvent_t t1);                                 \
  long long pass_blocks_p##P##_n##N##_v##V(int kind, int n);
GPE_DECL_PASS(0, 0, 0)
GPE_DECL_PASS(1, 0, 0)
GPE_VARIANTS(GPE_DECL_PASS)
#undef GPE_DECL_PASS

}  // namespace gpe
