/* libgpe_hip.so — C ABI of the MI355X split-step Gross-Pitaevskii engine.
 *
 * Plain C types only: opaque handles, host pointers and sizes.  Every call
 * returns a status (0 = ok, negative = error; gpe_last_error() describes the
 * most recent failure of the calling thread).  The library owns all device
 * memory; host pointers are borrowed for the duration of a call and never
 * retained.  Each solve owns one HIP stream; calls on one solve are ordered,
 * calls on different solves may overlap on the device.
 *
 * Equation (reference units, PoyntingZhang/GPE_simulator):
 *     i dpsi/dt = [-coef*lap + V + beta*|psi|^2] psi      (coef = 1 in the reference)
 * Strang step: local half-step -> 3D FFT -> kinetic -> inverse FFT ->
 * local half-step; imaginary time renormalises after every step.
 * Reference citations are paths inside the reference repository.
 */
#ifndef GPE_HIP_H
#define GPE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpe_ctx gpe_ctx;
typedef struct gpe_solve gpe_solve;
typedef struct gpe_batch gpe_batch;

enum {
  GPE_OK = 0,
  GPE_ERR_INVALID = -1,   /* bad argument or contract violation (reference: AssertionError) */
  GPE_ERR_NOMEM = -2,     /* device allocation failed */
  GPE_ERR_HIP = -3,       /* HIP runtime error */
  GPE_ERR_STATE = -4      /* call not valid in the solve's current state */
};

enum { GPE_C64 = 0, GPE_C128 = 1 };            /* field precision */
enum { GPE_REAL = 0, GPE_IMAG = 1 };           /* evolution kind */
enum { GPE_TGT_XY = 0, GPE_TGT_Z = 1, GPE_TGT_VOL = 2, GPE_TGT_VOL_IM = 3 };

/* Library version string. */
const char* gpe_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* gpe_last_error(void);

/* Context on one HIP device.  No reference counterpart (CuPy's current
 * device, gpe/main.py:27-38 runs on it implicitly). */
int gpe_ctx_create(int device, gpe_ctx** out);
int gpe_ctx_destroy(gpe_ctx* ctx);
/* Free / total device memory in bytes (serving admission). */
int gpe_mem_info(gpe_ctx* ctx, size_t* free_bytes, size_t* total_bytes);
/* Device name and architecture, e.g. "<name> (gfx950:...)", or the
 * architecture alone when the runtime reports no name (NUL-terminated,
 * truncated to name_len bytes); HIP runtime and driver versions of the
 * context's device: the report metadata that
 * gpe/benchmark.py:204-233 reads from CuPy's runtime. */
int gpe_device_info(gpe_ctx* ctx, char* name, int32_t name_len, int32_t* runtime_version,
                    int32_t* driver_version);

/* Pass tile variants.  A step is a sequence of passes (kinds 0..GPE_PASS_KINDS-1:
 * z_first, z_mid, z_last, y_fwd, y_inv, x_kin, z_diag, x_diag, z_fwd, y_lz,
 * x_lz, zk_first, zk_mid, yk, xk, xk_last, zyk_first, zyk_mid, zy_first,
 * zy_mid, zy_last); each pass kind of a solve runs one of the compiled tile
 * variants of its (side, precision), all producing the same transform (the
 * plane kinds zyk_* and zy_* have one shape per side, whatever the
 * variant).  An autotuner times the variants and sets the
 * fastest per kind.  No reference counterpart: the reference's FFT plans are
 * cuFFT's (gpe/make_propagator.py:50,52). */
enum { GPE_PASS_KINDS = 21 };
/* Bins of the equal-width spectrum view (gpe_spectrum_view). */
enum { GPE_VIEW_BINS = 128 };
/* Number of compiled variants for (side, precision); 1 for sides without a
 * fast transform. */
int gpe_pass_variants(int32_t n, int32_t precision, int32_t* count);
/* Geometry of a variant: threads per pencil, pencils per block, and the waves
 * per SIMD its register budget allows. */
int gpe_pass_variant_info(int32_t n, int32_t precision, int32_t variant, int32_t* threads_per_pencil,
                          int32_t* pencils_per_block, int32_t* waves);
/* Variants a context gives new solves of (side, precision, kind); `variants`
 * has GPE_PASS_KINDS entries. */
int gpe_ctx_set_variants(gpe_ctx* ctx, int32_t n, int32_t precision, int32_t kind, const int32_t* variants,
                         int32_t nkinds);

/* Solve description.  Replaces what make_real_propagator / make_imag_propagator
 * (gpe/make_propagator.py:5, :106) read from `stage` (stage.beta, stage.dt,
 * stage.N_grids, stage.x_array, stage.k_array, stage.k_cut_params;
 * gpe/stage.py:11-51). */
typedef struct gpe_solve_desc {
  int32_t n;                 /* grid side N (N^3 points), 2..1024            */
  int32_t precision;         /* GPE_C64 or GPE_C128                           */
  int32_t kind;              /* GPE_REAL or GPE_IMAG                          */
  int32_t kcut_enable;       /* masked kinetic step (gpe/kernels.py:24-45)    */
  double dt;                 /* time step (reference units)                   */
  double beta;               /* interaction strength (gpe/helper.py:109-113)  */
  double kcut;               /* cutoff wavenumber when kcut_enable            */
  double kinetic_coef;       /* 1.0 reproduces the reference's -lap           */
  const double* x_array;     /* host, n values (gpe/grids.py:3-6)             */
  const double* k_array;     /* host, n values (gpe/grids.py:3-6)             */
} gpe_solve_desc;

/* Allocate psi (N^3 complex of the requested precision) and the step tables
 * on the device.  Reference: stage.prepare building the propagator
 * (gpe/stage.py:53-66, gpe/make_propagator.py:5-152). */
int gpe_solve_create(gpe_ctx* ctx, const gpe_solve_desc* desc, gpe_solve** out);
int gpe_solve_destroy(gpe_solve* solve);
/* Device bytes a solve of this description holds (psi + tables), and the
 * extra bytes the diagnostics workspace adds when first used. */
int gpe_solve_footprint(const gpe_solve_desc* desc, size_t* resident_bytes, size_t* diag_bytes);
/* Step plan of a solve.  Both give the reference's Strang step
 * (gpe/make_propagator.py:38-152) to rounding.  GPE_PLAN_SEPARABLE, the
 * default: the kinetic factor exp(-i k^2 dt) is the product of one factor per
 * axis, and each commutes with the transforms along the other axes, so a step
 * is three passes (z, y, x), each FFT -> 1D factor -> inverse FFT, with the
 * local half-steps fused into the z pass (and the last one into the x pass):
 * 3 sweeps of psi per step.  GPE_PLAN_FOUR_SWEEP: the 3D FFT, the kinetic
 * factor and the inverse 3D FFT in four passes (z, y, x, y; 5 in imaginary
 * time).  GPE_PLAN_PLANE: one block per x plane of the field (complex64 at
 * sides 32, 48, 64, 96 and 128, complex128 at 32, 48 and 64, where a plane
 * fits the LDS) runs a step's z and y passes in one: the separable plan's
 * (2 sweeps of psi per step) or, with a momentum cutoff, the four-sweep
 * plan's y, z, y passes around the 3D kinetic pass (2 sweeps per step
 * instead of 4); either matches the plan it replaces to rounding.  A momentum cutoff (kcut_enable) is not separable, so such
 * solves use the four-sweep or the plane plan.  New solves with the cutoff
 * start on the plane plan where it exists, else on the four-sweep plan; new
 * solves without it on the plane plan where it exists and the side is a
 * multiple of 64, else on the separable plan.  Batch
 * members share a plan (on the plane plan, all with the cutoff or none). */
#define GPE_PLAN_SEPARABLE 0
#define GPE_PLAN_FOUR_SWEEP 1
#define GPE_PLAN_PLANE 2
int gpe_solve_set_plan(gpe_solve* solve, int32_t plan);
int gpe_solve_get_plan(gpe_solve* solve, int32_t* plan);

/* Pass variant per kind of one solve (GPE_PASS_KINDS entries each).  Batch
 * members must use the same variants. */
int gpe_solve_set_variants(gpe_solve* solve, const int32_t* variants, int32_t nkinds);
int gpe_solve_get_variants(gpe_solve* solve, int32_t* variants, int32_t nkinds);

/* Copy elements [offset, offset+count) of psi (flat C-order index) from / to
 * a host complex128 buffer (interleaved re, im).  complex64 solves convert
 * on the device.  Reference: stage.psi assignment / cp.asnumpy and the slab
 * copies of gpe/state_io.py:14-58. */
int gpe_psi_upload(gpe_solve* solve, const double* host_c128, int64_t offset, int64_t count);
int gpe_psi_download(gpe_solve* solve, double* host_c128, int64_t offset, int64_t count);
/* dst.psi = src.psi on the device (same side; precisions may differ).
 * Reference: a `continue` stage reusing the previous stage's array
 * (gpe/stage.py:74-79, gpe/main.py:31-36). */
int gpe_psi_copy(gpe_solve* dst, gpe_solve* src);
/* psi = separable Gaussian exp(-r^2/(2 sigma^2)), then normalised.
 * Reference: gpe/initial_states.py:4-14. */
int gpe_psi_gaussian(gpe_solve* solve, double sigma);

/* Potential programs (postfix code, see gpe_vm.h in the sources).
 * gpe_potential_static adds scale * f(x, y, z) to the FP64 table of `target`
 * (evaluated once, on the device).  gpe_potential_dynamic registers f(x, y, z, t)
 * for `target` (GPE_TGT_XY, GPE_TGT_Z or GPE_TGT_VOL), evaluated in FP64 at
 * every local half-step time t_mid.  Replaces the expression splicing of
 * make_real_spatial_kernel / make_imag_spatial_kernel (gpe/make_kernel.py:4-120). */
int gpe_potential_static(gpe_solve* solve, int32_t target, const int32_t* code, int32_t ncode,
                         const double* consts, int32_t nconst, double scale);
int gpe_potential_dynamic(gpe_solve* solve, int32_t target, const int32_t* code, int32_t ncode,
                          const double* consts, int32_t nconst);
/* Time-dependent programs are evaluated by kernels compiled at run time
 * (hiprtc) from the program's structure, as the reference compiles its
 * spliced expressions at run time (CuPy RawKernel, gpe/make_kernel.py:4-84);
 * programs that differ only in their constants share one compiled module,
 * and the tables are bit-identical to the interpreter's.  gpe_set_jit(0)
 * makes programs registered afterwards use the on-device interpreter (the
 * path taken anyway when a compile fails); gpe_jit_stats reports modules
 * compiled, table evaluations launched and failed compiles, process-wide.
 * No reference counterpart: CuPy's kernel cache plays this part there. */
int gpe_set_jit(int32_t enable);
int gpe_jit_stats(int64_t* compiles, int64_t* launches, int64_t* failures);
/* Static volume V_static = re + i*im (target GPE_TGT_VOL / GPE_TGT_VOL_IM):
 * fill with a constant, or upload a slab of host float64 values that replace
 * (accumulate = 0) or are added to (accumulate = 1) elements [offset,
 * offset+count).  Reference: stage.staticV (gpe/stage.py:56-62,
 * gpe/potentials.py:31-61). */
int gpe_volume_fill(gpe_solve* solve, int32_t target, double value);
int gpe_volume_upload(gpe_solve* solve, int32_t target, const double* host, int64_t offset, int64_t count,
                      int32_t accumulate);
/* Copies elements of the FP64 static volume (GPE_TGT_VOL: real part;
 * GPE_TGT_VOL_IM: imaginary part, the negated absorber of
 * gpe/potentials.py:31-37) to the host; GPE_ERR_STATE when the solve has none. */
int gpe_volume_download(gpe_solve* solve, int32_t target, double* host, int64_t offset, int64_t count);

/* Advance nsteps Strang steps.  t_mid holds the FP64 midpoint time of each
 * step (t_i + dt/2, as gpe/make_propagator.py:94); it may be NULL when no
 * time-dependent program is registered.  Consecutive steps inside one call
 * fuse the second half-step of step i with the first of step i+1.
 * Reference: the propagator closures (gpe/make_propagator.py:12-54, 57-103,
 * 109-151) called once per step by launch_stage (gpe/launch.py:8-12). */
int gpe_step(gpe_solve* solve, int64_t nsteps, const double* t_mid);
/* gpe_step that allocates nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same kernels, same
 * results as gpe_step (bit-identical).  In imaginary time the FP64 norm
 * partials live in the library's file-scope device array (one per GPU, held
 * until the process ends) instead of the solve's buffer, and the per-step
 * norm history goes to the solve's existing history buffer, which is never
 * grown.  Limits (GPE_ERR_INVALID beyond them): at most 65536 norm partials
 * (every fast side up to 1024, the generic transform up to side 256), and
 * nsteps at most the solve's history capacity (1024 steps from creation, more
 * after a gpe_step of more steps).  Real time uses no scratch.  Calls that
 * overlap on one GPU (solves on different streams or threads) share the
 * array and race: serialise them. */
int gpe_step_noalloc(gpe_solve* solve, int64_t nsteps, const double* t_mid);

/* The same real-time steps with the reference's launch structure: one
 * launch per reference kernel and transform axis (half-step, FFT z, y, x,
 * kinetic factor, inverse FFT x, y, z, half-step; gpe/make_propagator.py:38-101
 * with cuFFT's fftn / ifftn as three axis passes each), 9 read + write sweeps
 * of psi per step instead of the fused plan's 4.  Same results to rounding.
 * For measuring what the fusion gains; real-time solves only. */
int gpe_step_unfused(gpe_solve* solve, int64_t nsteps, const double* t_mid);
/* Norm per step of the last imaginary-time gpe_step call (before
 * renormalisation), `count` values ending with the last step. */
int gpe_norm_history(gpe_solve* solve, double* out, int64_t count);

/* FP64 norm dx^3 * sum |psi|^2.  Reference: get_N (gpe/helper.py:81-88). */
int gpe_norm(gpe_solve* solve, double* out);
/* psi /= sqrt(norm).  Reference: normalise (gpe/helper.py:91-93). */
int gpe_normalise(gpe_solve* solve);

/* All diagnostics from one forward FFT of the completed psi, FP64 sums:
 *   scalars[0] norm      dx^3 sum |psi|^2                 (get_N)
 *   scalars[1] E_kin     dx^3/N^3 sum k^2 |F|^2           (get_E, gpe/helper.py:135-140)
 *   scalars[2] E_pot     dx^3 sum V(t)|psi|^2
 *   scalars[3] E_int     beta/2 dx^3 sum |psi|^4
 *   scalars[4] E         E_kin + E_pot + E_int
 *   scalars[5] mu        (E_kin + E_pot + 2 E_int) / norm
 *   scalars[6] parseval  dx^3/N^3 sum |F|^2
 * counts / occupation: per shell floor(|k|/dk + 1/2), nbins values, raw
 * sum |F|^2 (multiply by dx^3/N^3 for Nk).  With both null the modes are not
 * binned at all (scalars only; gpe_spectrum_view keeps the view of the last
 * call that binned them).  Reference: save_outputs'
 * shared fftn (gpe/save.py:31-57) and nk_extraction (gpe/nk_extraction.py:16-45).
 * t is the time at which the time-dependent potential enters E_pot. */
int gpe_diagnostics(gpe_solve* solve, double t, double* scalars, int64_t* counts, double* occupation,
                    int32_t nbins);
/* The same spectrum in GPE_VIEW_BINS equal-width bins over [0, k_max], with
 * k_max = sqrt(3) * (n/2) * dk the grid's corner wavenumber: mode m (integer
 * frequencies) goes to bin floor(GPE_VIEW_BINS * |m| / (sqrt(3) * (n/2))),
 * the corner to the last bin, decided with exact integer arithmetic.
 * counts / occupation as in gpe_diagnostics (nbins must be GPE_VIEW_BINS),
 * from the most recent gpe_diagnostics call on the solve that asked for the
 * spectrum (non-null counts or occupation).  No reference
 * counterpart: the reference bins by dk shells only
 * (gpe/nk_extraction.py:16-45); this is the fixed-size view of that same
 * |F|^2 for comparing runs of different grid sides. */
int gpe_spectrum_view(gpe_solve* solve, int64_t* counts, double* occupation, int32_t nbins);

/* Lz = Re int psi^* (x p_y - y p_x) psi dV with spectral derivatives, FP64.
 * Reference: get_L (gpe/helper.py:143-156). */
int gpe_angular_momentum(gpe_solve* solve, double* out);

/* Batches: equal-shape solves (same side, precision and kind) advanced
 * together, every pass one launch over all members (the member index is the
 * third grid dimension).  Each member keeps its own psi, potentials (static
 * and time-dependent), beta, dt, pass variants and reductions, so its result
 * is the same as when it runs alone.  With use_graph != 0 the nsteps-step
 * launch sequence is captured once in a hipGraph and replayed.  The batch
 * borrows its members: do not step a member alone while its batch is
 * stepping.  No reference counterpart (the reference runs one stage at a
 * time, gpe/main.py:30-38). */
int gpe_batch_create(gpe_ctx* ctx, gpe_solve* const* solves, int32_t count, gpe_batch** out);
int gpe_batch_destroy(gpe_batch* batch);
/* Steps batches whose members carry no time-dependent program. */
int gpe_batch_step(gpe_batch* batch, int64_t nsteps, int32_t use_graph);
/* Steps a real-time batch whose members may carry time-dependent programs
 * (e.g. stirrers of different speeds): member i takes its FP64 midpoint times
 * from t_mid + i*stride (stride 0: all members share t_mid[0..nsteps)), as
 * gpe_step does for one solve.  Every member's tables are evaluated on the
 * batch's stream before each chunk of steps; no graph capture. */
int gpe_batch_step_times(gpe_batch* batch, int64_t nsteps, const double* t_mid, int64_t stride);
/* Asynchronous form: enqueue the steps on the batch's own stream and return;
 * gpe_batch_wait blocks until they are done.  Batches of different shapes
 * launched back to back overlap on the device.  Members must not be used
 * until the wait returns. */
int gpe_batch_launch(gpe_batch* batch, int64_t nsteps, int32_t use_graph);
int gpe_batch_wait(gpe_batch* batch);
/* Non-blocking form of the wait: *done = 1 when everything the batch has
 * enqueued has finished, else 0.  A server polls its batches with it and
 * relaunches each one as soon as its own steps are done. */
int gpe_batch_done(gpe_batch* batch, int32_t* done);
/* Per-step norms of the last `count` steps of every member (imaginary time),
 * member-major: out[i*count + j]; one synchronisation for the whole batch.
 * Reference: the per-step normalise of make_imag_propagator
 * (gpe/make_propagator.py:148-150), kept per member. */
int gpe_batch_norm_history(gpe_batch* batch, double* out, int64_t count);
int gpe_batch_step_timed(gpe_batch* batch, int64_t nsteps, double* ms_per_kind, int64_t* launches_per_kind);

/* Wait for all work of the solve. */
int gpe_sync(gpe_solve* solve);

/* Timing marks on the solve's stream: the CuPy event pair that brackets one
 * timed batch in gpe/benchmark.py:59-69.  gpe_mark(solve, 0) records the
 * start mark and gpe_mark(solve, 1) the stop mark, after whatever the solve
 * has queued by then; gpe_mark_elapsed waits for the stop mark and returns the
 * milliseconds between the two.  Like the CuPy pair, the interval includes
 * device idle time while the host was still submitting. */
int gpe_mark(gpe_solve* solve, int32_t which);
int gpe_mark_elapsed(gpe_solve* solve, double* ms);

/* Measurement hooks (no reference counterpart; gpe/benchmark.py:59-69 times
 * whole batches).  gpe_step_timed runs gpe_step with a HIP event pair around
 * every pass launch on the solve's stream, without synchronising between
 * launches, and reports the summed milliseconds per pass kind
 * (GPE_PASS_KINDS entries, in the order listed above) and launch counts.
 * gpe_batch_step_timed does the same for a batch (eager launches, no graph). */
int gpe_step_timed(gpe_solve* solve, int64_t nsteps, const double* t_mid, double* ms_per_kind,
                   int64_t* launches_per_kind);
/* gpe_step with each pass launch timed by the device's constant-rate wall
 * clock instead of events: every wave of a launch records its end time, the
 * latest is the launch's end, and a launch's span runs from the previous
 * timed launch's end (or from a reference time taken just before it, when
 * other work such as a time-dependent table evaluation came in between) to
 * its own end: completion to completion, with no events between the
 * launches.  Reports summed microseconds per pass kind and launch counts
 * like gpe_step_timed.  Launches beyond 4M block slots per call run untimed
 * (about 1000 launches at 256^3). */
int gpe_step_stamped(gpe_solve* solve, int64_t nsteps, const double* t_mid, double* us_per_kind,
                     int64_t* launches_per_kind);
/* fftn(psi), raw and unnormalised as numpy's and CuPy's fftn (the psi_k
 * of gpe/save.py:31), computed on the device into the solve's work field;
 * psi is unchanged.  gpe_transform_download copies elements of it to the
 * host as complex128 (GPE_ERR_STATE unless a gpe_transform of the current
 * field came before: any step, upload or diagnostics call invalidates it). */
int gpe_transform(gpe_solve* solve);
int gpe_transform_download(gpe_solve* solve, double* host_c128, int64_t offset, int64_t count);

/* The monitor images of gpe/save.py:59-90 as raw FP64 sums on the device
 * (sum_abs2 of gpe/helper.py:3-12), each an n*n C-order plane; the caller
 * applies the reference's scale factors and fftshift:
 *   GPE_MON_TOF_XY        sum_z |F|^2, F = fftn(psi)   -> tof_xy[i*n + j]
 *   GPE_MON_TOF_YZ        sum_x |F|^2                  -> tof_yz[j*n + k]
 *   GPE_MON_TOF_SLICE_XY  |F[:, :, 0]|^2               -> tof_slice_xy
 *   GPE_MON_TOF_SLICE_YZ  |F[0, :, :]|^2               -> tof_slice_yz
 *   GPE_MON_DENSITY_XY    sum_z |psi|^2                -> density_xy
 *   GPE_MON_PLANE_MID     psi[:, :, n/2] as complex128 -> plane_mid (2*n*n doubles)
 * `which` ORs the images wanted; outputs of other images may be NULL.  One
 * transform serves every momentum image (and is reused while the field is
 * unchanged since gpe_transform). */
#define GPE_MON_TOF_XY 1
#define GPE_MON_TOF_YZ 2
#define GPE_MON_TOF_SLICE_XY 4
#define GPE_MON_TOF_SLICE_YZ 8
#define GPE_MON_DENSITY_XY 16
#define GPE_MON_PLANE_MID 32
#define GPE_MON_ALL 63
int gpe_monitor_images(gpe_solve* solve, int32_t which, double* tof_xy, double* tof_yz, double* tof_slice_xy,
                       double* tof_slice_yz, double* density_xy, double* plane_mid);

/* gpe_step with every pass launch carrying its own start and stop event,
 * both timestamped by the dispatch itself (hipExtLaunchKernel), and nothing
 * recorded between launches: a span is the kernel's execution alone, the
 * duration a kernel trace reports for the same launch.  Reports summed
 * microseconds per pass kind and launch counts like gpe_step_timed. */
int gpe_step_kernel_times(gpe_solve* solve, int64_t nsteps, const double* t_mid, double* us_per_kind,
                          int64_t* launches_per_kind);
/* Device-to-device streaming copy of `bytes` (multiple of 16), repeated
 * `reps` times; returns milliseconds per copy.  Bandwidth calibration. */
int gpe_bench_copy(gpe_ctx* ctx, int64_t bytes, int32_t reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* GPE_HIP_H */
