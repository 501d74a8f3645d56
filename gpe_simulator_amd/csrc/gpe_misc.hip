// Setup, table, reduction and data-movement kernels of libgpe_hip.so.  None of
// them is on the per-step path except k_scale (deferred renormalisation) and
// k_vm_eval (time-dependent potential tables, a few n^2 entries per step).
#include "gpe_misc.h"

namespace gpe {

// ------------------------------------------------------------ twiddles --
// tw[r] = exp(-2*pi*i*r/n) in FP64: exact octant reduction on integers, then
// sincospi of an argument in [0, 1/4].
template <typename R>
__global__ void k_twiddle(cpx<R>* tw, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const long long u = 8LL * r;
  const long long oct = u / n, f = u % n;
  const double a = double(f) / (4.0 * n);       // in units of pi, <= 1/4
  const double b = double(n - f) / (4.0 * n);
  double sa, ca, sb, cb;
  sincospi(a, &sa, &ca);
  sincospi(b, &sb, &cb);
  double s, c;
  switch (oct) {
    case 0: s = sa; c = ca; break;
    case 1: s = cb; c = sb; break;
    case 2: s = ca; c = -sa; break;
    case 3: s = sb; c = -cb; break;
    case 4: s = -sa; c = -ca; break;
    case 5: s = -cb; c = -sb; break;
    case 6: s = -ca; c = sa; break;
    default: s = -sb; c = cb; break;
  }
  tw[r] = cpx<R>{R(c), R(-s)};
}

// Kinetic factors per axis in FP64: exp(-i*coef*k^2*dt) (real time) or
// exp(-coef*k^2*dt) (imaginary time).  The x factor also carries 1/n^3.  k1
// (n entries, by FFT index like the others) is the separable plan's 1D
// factor, which carries 1/n for its axis's inverse FFT.
template <typename R>
__global__ void k_kinetic(cpx<R>* kx, cpx<R>* ky, cpx<R>* kz, cpx<R>* k1, const double* karr, int n, double dt,
                          double coef, int imag, double inv_n3) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double k = karr[i];
  const double ph = coef * (k * k) * dt;
  cpx<R> e;
  double re, im;
  if (imag) {
    re = exp(-ph);
    im = 0.0;
  } else {
    double s, c;
    sincos(ph, &s, &c);
    re = c;
    im = -s;
  }
  e = cpx<R>{R(re), R(im)};
  ky[i] = e;
  kz[i] = e;
  kx[i] = cpx<R>{R(re * inv_n3), R(im * inv_n3)};
  k1[i] = cpx<R>{R(re / n), R(im / n)};
}

// ------------------------------------------------------ expression VM --
__device__ double vm_run(const int* __restrict__ code, int ncode, const double* __restrict__ consts, double x,
                         double y, double z, double t) {
  double st[kVmStack];
  int sp = 0;
  for (int pc = 0; pc < ncode; ++pc) {
    const int op = code[pc];
    switch (op) {
      case OP_CONST: st[sp++] = consts[code[++pc]]; break;
      case OP_X: st[sp++] = x; break;
      case OP_Y: st[sp++] = y; break;
      case OP_Z: st[sp++] = z; break;
      case OP_T: st[sp++] = t; break;
      case OP_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
      case OP_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
      case OP_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
      case OP_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
      case OP_POW: --sp; st[sp - 1] = pow(st[sp - 1], st[sp]); break;
      case OP_MAX: --sp; st[sp - 1] = fmax(st[sp - 1], st[sp]); break;
      case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case OP_SIN: st[sp - 1] = sin(st[sp - 1]); break;
      case OP_COS: st[sp - 1] = cos(st[sp - 1]); break;
      case OP_TAN: st[sp - 1] = tan(st[sp - 1]); break;
      case OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case OP_TANH: st[sp - 1] = tanh(st[sp - 1]); break;
      case OP_FABS: st[sp - 1] = fabs(st[sp - 1]); break;
      case OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
      default: break;
    }
  }
  return sp > 0 ? st[0] : 0.0;
}

__device__ __forceinline__ void target_coords(int target, long long p, int n, const double* xa, double* x,
                                              double* y, double* z) {
  if (target == TGT_XY) {
    *x = xa[p / n];
    *y = xa[p % n];
    *z = 0.0;
  } else if (target == TGT_Z) {
    *x = 0.0;
    *y = 0.0;
    *z = xa[p];
  } else {
    const long long nn = (long long)n * n;
    *x = xa[p / nn];
    *y = xa[(p / n) % n];
    *z = xa[p % n];
  }
}

// Static programs: master[p] += scale * f(x, y, z) in FP64.
__global__ void k_vm_static(const int* code, int ncode, const double* consts, int target, int n,
                            const double* xa, long long size, double scale, double* master) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= size) return;
  double x, y, z;
  target_coords(target, p, n, xa, &x, &y, &z);
  master[p] += scale * vm_run(code, ncode, consts, x, y, z, 0.0);
}

// Time-dependent programs: out[s][p] = R(master[p] + f(x, y, z, t_s)) with
// the FP64 times t_s taken from tlist; master may be null.
template <typename R>
__global__ void k_vm_dynamic(const int* code, int ncode, const double* consts, int target, int n,
                             const double* xa, long long size, const double* master, const double* tlist,
                             R* out) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = blockIdx.y;
  if (p >= size) return;
  double x, y, z;
  target_coords(target, p, n, xa, &x, &y, &z);
  const double base = master ? master[p] : 0.0;
  out[(long long)slot * size + p] = R(base + vm_run(code, ncode, consts, x, y, z, tlist[slot]));
}

template <typename R>
__global__ void k_cast(const double* in, R* out, long long size) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < size) out[p] = R(in[p]);
}

__global__ void k_fill_f64(double* out, long long size, double v) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < size) out[p] = v;
}

__global__ void k_add_f64(double* out, const double* in, long long size) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < size) out[p] += in[p];
}

// ------------------------------------------------------------ fields --
// The real-time kinetic factor applied point by point over the whole k-space
// field, one thread per point, as the reference's elementwise kinetic kernels
// do (gpe/kernels.py:3-45): the per-axis FP64 factors (1/n^3 folded into x)
// multiplied as in the fused x pass, the same cutoff rule, one FP64 product
// rounded once.  Used by the reference-structured step.
template <typename R>
__global__ void k_kinetic_points(cpx<R>* psi, const cpx<double>* kx, const cpx<double>* ky, const cpx<double>* kz,
                                 const double* karr, int n, int masked, double kcut2) {
  const long long nn = (long long)n * n, count = nn * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = int(e / nn), j = int((e / n) % n), k = int(e % n);
    cpx<double> f = cmul(kx[i], cmul(ky[j], kz[k]));
    if (masked) {
      const double k2 =
          add_rn(add_rn(mul_rn(karr[i], karr[i]), mul_rn(karr[j], karr[j])), mul_rn(karr[k], karr[k]));
      if (k2 > kcut2) f = cpx<double>{0.0, 0.0};
    }
    psi[e] = cmul_d(psi[e], f);
  }
}

// Launch-timing slots start at 0 (raised by atomicMax).
__global__ void k_clear_stamps(unsigned long long* stamps, long long count) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
       e += (long long)gridDim.x * blockDim.x)
    stamps[e] = 0ULL;
}

// A reference time on the stream: the wall clock when this one-thread launch
// runs, i.e. when the work enqueued before it has finished.
__global__ void k_stamp_now(unsigned long long* slot) { *slot = (unsigned long long)wall_clock64(); }

template <typename R>
__global__ void k_scale(cpx<R>* psi, long long count, const double* scal) {
  const double s = scal[1];
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x) {
    const cpx<R> v = psi[p];
    psi[p] = cpx<R>{R(double(v.x) * s), R(double(v.y) * s)};
  }
}

// psi[i,j,k] = (g[i]*g[j])*g[k], g = exp(-x^2/(2 sigma^2)), all in FP64.
template <typename R>
__global__ void k_gaussian(cpx<R>* psi, const double* xa, int n, double sigma) {
  const long long count = (long long)n * n * n;
  const double den = 2.0 * sigma * sigma;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x) {
    const long long i = p / ((long long)n * n), j = (p / n) % n, k = p % n;
    const double gi = exp(-xa[i] * xa[i] / den);
    const double gj = exp(-xa[j] * xa[j] / den);
    const double gk = exp(-xa[k] * xa[k] / den);
    psi[p] = cpx<R>{R((gi * gj) * gk), R(0)};
  }
}

template <typename R>
__global__ void k_from_c128(const cpx<double>* in, cpx<R>* out, long long count) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x) {
    const cpx<double> v = in[p];
    out[p] = cpx<R>{R(v.x), R(v.y)};
  }
}

template <typename R>
__global__ void k_to_c128(const cpx<R>* in, cpx<double>* out, long long count) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x) {
    const cpx<R> v = in[p];
    out[p] = cpx<double>{double(v.x), double(v.y)};
  }
}

// FP64 sum of |psi|^2 (times prescale^2 when pending) with the last block
// publishing norm = dx3 * sum into scal[0] and 1/sqrt(norm) into scal[1].
template <typename R>
__global__ void k_norm(const cpx<R>* psi, long long count, double* red, unsigned int* counter, double* scal,
                       double dx3, int prescale) {
  __shared__ double scratch[16];
  __shared__ int is_last;
  double acc[1] = {0.0};
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x) {
    const cpx<R> v = psi[p];
    acc[0] += double(v.x) * double(v.x) + double(v.y) * double(v.y);
  }
  block_sum<1>(acc, scratch);
  if (threadIdx.x == 0) {
    red[blockIdx.x] = acc[0];
    __threadfence();
    is_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) {
    __threadfence();
    double part[1] = {0.0};
    for (unsigned q = threadIdx.x; q < gridDim.x; q += blockDim.x) part[0] += coherent_load(red + q);
    block_sum<1>(part, scratch);
    if (threadIdx.x == 0) {
      double sum = part[0];
      if (prescale) sum *= scal[1] * scal[1];
      const double norm = dx3 * sum;
      scal[0] = norm;
      scal[1] = 1.0 / sqrt(norm);
      *counter = 0u;
    }
  }
}

// Deterministic finish of per-block partials red[b*nq + q] -> out[q].
__global__ void k_finish_sums(const double* red, long long nblk, int nq, double* out) {
  __shared__ double scratch[16];
  for (int q = 0; q < nq; ++q) {
    double acc[1] = {0.0};
    for (long long b = threadIdx.x; b < nblk; b += blockDim.x) acc[0] += red[b * nq + q];
    block_sum<1>(acc, scratch);
    if (threadIdx.x == 0) out[q] = acc[0];
    __syncthreads();
  }
}

// Per-bin finish of the shell histogram partials: one block per bin, fixed
// per-thread strides and a fixed tree, so the result is deterministic.
__global__ void k_finish_nk(const double* occ, const unsigned int* cnt, long long nblk, int nbins, double* out_occ,
                            long long* out_cnt) {
  __shared__ double scratch[16];
  const int q = blockIdx.x;
  double acc[2] = {0.0, 0.0};
  for (long long b = threadIdx.x; b < nblk; b += blockDim.x) {
    acc[0] += occ[b * nbins + q];
    acc[1] += double(cnt[b * nbins + q]);
  }
  block_sum<2>(acc, scratch);
  if (threadIdx.x == 0) {
    out_occ[q] = acc[0];
    out_cnt[q] = (long long)acc[1];
  }
}

// ------------------------------------------------------ monitor images --
// Projections of |f|^2 for the monitor images, summed in FP64 in a fixed
// order (deterministic):
//   k_abs2_sum_z:  out[i*n + j] = sum_k |f[i][j][k]|^2, one wave per (i, j)
//                  row (contiguous loads, lanes stride the row, then a wave
//                  reduction);
//   k_abs2_sum_x:  out[j*n + k] = sum_i |f[i][j][k]|^2, one thread per (j, k)
//                  column walking i (adjacent threads read adjacent
//                  elements of each plane).
template <typename R>
__global__ void __launch_bounds__(256) k_abs2_sum_z(const cpx<R>* f, int n, double* out) {
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= (long long)n * n) return;
  const cpx<R>* r = f + row * n;
  double acc = 0.0;
  for (int k = lane; k < n; k += 64) {
    const cpx<R> v = gload(r + k);
    acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
  }
  acc = wave_sum(acc);
  if (lane == 0) out[row] = acc;
}

template <typename R>
__global__ void __launch_bounds__(256) k_abs2_sum_x(const cpx<R>* f, int n, double* out) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)n * n;
  if (col >= plane) return;
  double acc = 0.0;
  for (int i = 0; i < n; ++i) {
    const cpx<R> v = gload(f + i * plane + col);
    acc += double(v.x) * double(v.x) + double(v.y) * double(v.y);
  }
  out[col] = acc;
}

// One plane of f as complex128: axis 2 gives out[i*n + j] = f[i][j][index]
// (a strided gather), axis 0 gives out[j*n + k] = f[index][j][k].  With
// `abs2`, |.|^2 in FP64 is written to out_abs2 instead.
template <typename R>
__global__ void __launch_bounds__(256) k_plane(const cpx<R>* f, int n, int axis, int index, cpx<double>* out,
                                               double* out_abs2) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)n * n;
  if (p >= plane) return;
  const long long src = axis == 2 ? p * n + index : (long long)index * plane + p;
  const cpx<R> v = gload(f + src);
  if (out_abs2) {
    out_abs2[p] = double(v.x) * double(v.x) + double(v.y) * double(v.y);
  } else {
    out[p] = cpx<double>{double(v.x), double(v.y)};
  }
}

// Streaming copy used to calibrate bandwidth counters.
__global__ void k_copy(const float4* in, float4* out, long long count) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < count;
       p += (long long)gridDim.x * blockDim.x)
    out[p] = in[p];
}

// ------------------------------------------------------------ launchers --
static inline unsigned nblocks(long long count, int threads) {
  long long b = (count + threads - 1) / threads;
  return (unsigned)(b < 1 ? 1 : b);
}
static inline unsigned nblocks_stride(long long count, int threads) {
  long long b = (count + threads - 1) / threads;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

template <typename R>
void launch_twiddle(void* tw, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_twiddle<R>, dim3(nblocks(n, 256)), dim3(256), 0, st, (cpx<R>*)tw, n);
}
template <typename R>
void launch_kinetic(void* kx, void* ky, void* kz, void* k1, const double* karr, int n, double dt, double coef,
                    int imag, double inv_n3, hipStream_t st) {
  hipLaunchKernelGGL(k_kinetic<R>, dim3(nblocks(n, 256)), dim3(256), 0, st, (cpx<R>*)kx, (cpx<R>*)ky,
                     (cpx<R>*)kz, (cpx<R>*)k1, karr, n, dt, coef, imag, inv_n3);
}
void launch_vm_static(const int* code, int ncode, const double* consts, int target, int n, const double* xa,
                      long long size, double scale, double* master, hipStream_t st) {
  hipLaunchKernelGGL(k_vm_static, dim3(nblocks(size, 256)), dim3(256), 0, st, code, ncode, consts, target, n, xa,
                     size, scale, master);
}
template <typename R>
void launch_vm_dynamic(const int* code, int ncode, const double* consts, int target, int n, const double* xa,
                       long long size, const double* master, const double* tlist, int nslots, void* out,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_vm_dynamic<R>, dim3(nblocks(size, 256), nslots), dim3(256), 0, st, code, ncode, consts,
                     target, n, xa, size, master, tlist, (R*)out);
}
template <typename R>
void launch_cast(const double* in, void* out, long long size, hipStream_t st) {
  hipLaunchKernelGGL(k_cast<R>, dim3(nblocks(size, 256)), dim3(256), 0, st, in, (R*)out, size);
}
void launch_fill_f64(double* out, long long size, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f64, dim3(nblocks(size, 256)), dim3(256), 0, st, out, size, v);
}
void launch_add_f64(double* out, const double* in, long long size, hipStream_t st) {
  hipLaunchKernelGGL(k_add_f64, dim3(nblocks(size, 256)), dim3(256), 0, st, out, in, size);
}
template <typename R>
void launch_scale(void* psi, long long count, const double* scal, hipStream_t st) {
  hipLaunchKernelGGL(k_scale<R>, dim3(nblocks_stride(count, 256)), dim3(256), 0, st, (cpx<R>*)psi, count, scal);
}
void launch_clear_stamps(unsigned long long* stamps, long long count, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_stamps, dim3(nblocks_stride(count, 256)), dim3(256), 0, st, stamps, count);
}
void launch_stamp_now(unsigned long long* slot, hipStream_t st) {
  hipLaunchKernelGGL(k_stamp_now, dim3(1), dim3(1), 0, st, slot);
}
template <typename R>
void launch_kinetic_points(void* psi, const void* kx, const void* ky, const void* kz, const double* karr, int n,
                           int masked, double kcut2, hipStream_t st) {
  hipLaunchKernelGGL(k_kinetic_points<R>, dim3(nblocks_stride((long long)n * n * n, 256)), dim3(256), 0, st,
                     (cpx<R>*)psi, (const cpx<double>*)kx, (const cpx<double>*)ky, (const cpx<double>*)kz, karr, n,
                     masked, kcut2);
}
template <typename R>
void launch_gaussian(void* psi, const double* xa, int n, double sigma, hipStream_t st) {
  hipLaunchKernelGGL(k_gaussian<R>, dim3(nblocks_stride((long long)n * n * n, 256)), dim3(256), 0, st,
                     (cpx<R>*)psi, xa, n, sigma);
}
template <typename R>
void launch_from_c128(const void* in, void* out, long long count, hipStream_t st) {
  hipLaunchKernelGGL(k_from_c128<R>, dim3(nblocks_stride(count, 256)), dim3(256), 0, st, (const cpx<double>*)in,
                     (cpx<R>*)out, count);
}
template <typename R>
void launch_to_c128(const void* in, void* out, long long count, hipStream_t st) {
  hipLaunchKernelGGL(k_to_c128<R>, dim3(nblocks_stride(count, 256)), dim3(256), 0, st, (const cpx<R>*)in,
                     (cpx<double>*)out, count);
}
template <typename R>
void launch_norm(const void* psi, long long count, double* red, unsigned int* counter, double* scal, double dx3,
                 int prescale, hipStream_t st) {
  hipLaunchKernelGGL(k_norm<R>, dim3(kNormBlocks), dim3(256), 0, st, (const cpx<R>*)psi, count, red, counter, scal,
                     dx3, prescale);
}
void launch_finish_sums(const double* red, long long nblk, int nq, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(256), 0, st, red, nblk, nq, out);
}
void launch_finish_nk(const double* occ, const unsigned int* cnt, long long nblk, int nbins, double* out_occ,
                      long long* out_cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_finish_nk, dim3(nbins), dim3(256), 0, st, occ, cnt, nblk, nbins, out_occ, out_cnt);
}
void launch_copy(const void* in, void* out, long long bytes, hipStream_t st) {
  const long long count = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(nblocks_stride(count, 256)), dim3(256), 0, st, (const float4*)in, (float4*)out,
                     count);
}

template <typename R>
void launch_abs2_sum(const void* f, int n, int axis, double* out, hipStream_t st) {
  const long long plane = (long long)n * n;
  if (axis == 2)
    hipLaunchKernelGGL(k_abs2_sum_z<R>, dim3((unsigned)((plane + 3) / 4)), dim3(256), 0, st, (const cpx<R>*)f, n,
                       out);
  else
    hipLaunchKernelGGL(k_abs2_sum_x<R>, dim3(nblocks(plane, 256)), dim3(256), 0, st, (const cpx<R>*)f, n, out);
}
template <typename R>
void launch_plane(const void* f, int n, int axis, int index, void* out_c128, double* out_abs2, hipStream_t st) {
  hipLaunchKernelGGL(k_plane<R>, dim3(nblocks((long long)n * n, 256)), dim3(256), 0, st, (const cpx<R>*)f, n, axis,
                     index, (cpx<double>*)out_c128, out_abs2);
}

#define GPE_INST(R)                                                                                            \
  template void launch_twiddle<R>(void*, int, hipStream_t);                                                  \
  template void launch_kinetic<R>(void*, void*, void*, void*, const double*, int, double, double, int,      \
                                  double, hipStream_t);                                                      \
  template void launch_vm_dynamic<R>(const int*, int, const double*, int, int, const double*, long long,     \
                                     const double*, const double*, int, void*, hipStream_t);                 \
  template void launch_cast<R>(const double*, void*, long long, hipStream_t);                                \
  template void launch_scale<R>(void*, long long, const double*, hipStream_t);                               \
  template void launch_gaussian<R>(void*, const double*, int, double, hipStream_t);                          \
  template void launch_kinetic_points<R>(void*, const void*, const void*, const void*, const double*, int, int,  \
                                         double, hipStream_t);                                               \
  template void launch_from_c128<R>(const void*, void*, long long, hipStream_t);                             \
  template void launch_to_c128<R>(const void*, void*, long long, hipStream_t);                               \
  template void launch_norm<R>(const void*, long long, double*, unsigned int*, double*, double, int,         \
                               hipStream_t);                                                              \
  template void launch_abs2_sum<R>(const void*, int, int, double*, hipStream_t);                             \
  template void launch_plane<R>(const void*, int, int, int, void*, double*, hipStream_t);
GPE_INST(float)
GPE_INST(double)

}  // namespace gpe
