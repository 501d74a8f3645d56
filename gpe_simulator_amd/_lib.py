"""ctypes binding of libgpe_hip.so (the C ABI declared in include/gpe_hip.h).

The library is loaded from this package directory.  There is no fallback: if
libgpe_hip.so is missing or no HIP device is present, every call that needs the
device raises GpeError.
"""

from __future__ import annotations

import ctypes
import threading
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

import numpy as np

LIBRARY = Path(__file__).resolve().parent / "libgpe_hip.so"

GPE_OK, GPE_ERR_INVALID, GPE_ERR_NOMEM, GPE_ERR_HIP, GPE_ERR_STATE = 0, -1, -2, -3, -4
C64, C128 = 0, 1
VIEW_BINS = 128                  # GPE_VIEW_BINS: equal-width bins of the spectrum view
REAL, IMAG = 0, 1
TGT_XY, TGT_Z, TGT_VOL, TGT_VOL_IM = 0, 1, 2, 3

PASS_KINDS = ("z_first", "z_mid", "z_last", "y_fwd", "y_inv", "x_kin", "z_diag", "x_diag", "z_fwd", "y_lz", "x_lz",
              "zk_first", "zk_mid", "yk", "xk", "xk_last", "zyk_first", "zyk_mid", "zy_first", "zy_mid", "zy_last")
# step plans (GPE_PLAN_*): three sweeps per step with the kinetic factor split
# per axis, the four-sweep 3D-transform plan (cutoff allowed), or the plane
# plan, which runs the z and y axes of each x plane in one pass (sides whose
# plane fits the LDS): the separable plan in two sweeps, or with a cutoff the
# four-sweep plan in two (the cutoff is real-time only)
PLAN_SEPARABLE, PLAN_FOUR_SWEEP, PLAN_PLANE = 0, 1, 2


# sides with a plane pass (an x plane of the field fits the LDS) per precision
PLANE_SIDES = {0: (32, 48, 64, 96, 128), 1: (32, 48, 64)}


def default_plan(precision, kind, masked=False, n=None):
    """The plan a new solve steps with (gpe_solve_create's rule).  With a
    momentum cutoff: the plane plan (its ZY passes around the 3D kinetic
    pass) at every plane side, else four-sweep.  Without: the plane plan at
    the plane sides that are multiples of 64 (faster alone and in batches),
    else separable."""
    plane = n is not None and n in PLANE_SIDES[int(precision)]
    if masked:
        return PLAN_PLANE if plane else PLAN_FOUR_SWEEP
    return PLAN_PLANE if plane and n % 64 == 0 else PLAN_SEPARABLE

EXPORTS = (
    "gpe_version", "gpe_last_error", "gpe_ctx_create", "gpe_ctx_destroy", "gpe_mem_info",
    "gpe_solve_create", "gpe_solve_destroy", "gpe_solve_footprint", "gpe_psi_upload",
    "gpe_psi_download", "gpe_psi_copy", "gpe_psi_gaussian", "gpe_potential_static",
    "gpe_potential_dynamic", "gpe_volume_fill", "gpe_volume_upload", "gpe_step",
    "gpe_norm_history", "gpe_norm", "gpe_normalise", "gpe_diagnostics", "gpe_sync",
    "gpe_step_timed", "gpe_bench_copy", "gpe_angular_momentum", "gpe_batch_create", "gpe_batch_destroy",
    "gpe_batch_step", "gpe_batch_step_timed", "gpe_batch_launch", "gpe_batch_wait", "gpe_batch_done",
    "gpe_pass_variants", "gpe_pass_variant_info", "gpe_ctx_set_variants", "gpe_solve_set_variants",
    "gpe_solve_get_variants", "gpe_batch_step_times", "gpe_batch_norm_history", "gpe_device_info",
    "gpe_mark", "gpe_mark_elapsed", "gpe_spectrum_view", "gpe_step_unfused", "gpe_step_stamped",
    "gpe_step_kernel_times", "gpe_transform", "gpe_transform_download", "gpe_monitor_images",
    "gpe_volume_download", "gpe_step_noalloc", "gpe_solve_set_plan", "gpe_solve_get_plan",
    "gpe_set_jit", "gpe_jit_stats",
)

# gpe_monitor_images bits (GPE_MON_* in include/gpe_hip.h), in output order
MONITOR_IMAGES = ("TOF_xy", "TOF_yz", "TOF_slice_xy", "TOF_slice_yz", "density_xy", "plane_mid")


class GpeError(RuntimeError):
    """A libgpe_hip call failed (status code and the library's message)."""

    def __init__(self, code, message):
        self.code = code
        super().__init__(f"libgpe_hip error {code}: {message}")


class GpeInvalid(GpeError, AssertionError):
    """Contract violation; an AssertionError like the reference's checks."""


class SolveDesc(ctypes.Structure):
    _fields_ = [
        ("n", c_int32), ("precision", c_int32), ("kind", c_int32), ("kcut_enable", c_int32),
        ("dt", c_double), ("beta", c_double), ("kcut", c_double), ("kinetic_coef", c_double),
        ("x_array", POINTER(c_double)), ("k_array", POINTER(c_double)),
    ]


_lib = None
_lock = threading.RLock()


def _declare(lib):
    vp, i32, i64, dp = c_void_p, c_int32, c_int64, POINTER(c_double)
    sig = {
        "gpe_version": (c_char_p, []),
        "gpe_set_jit": (c_int, [i32]),
        "gpe_jit_stats": (c_int, [POINTER(i64), POINTER(i64), POINTER(i64)]),
        "gpe_last_error": (c_char_p, []),
        "gpe_ctx_create": (c_int, [c_int, POINTER(vp)]),
        "gpe_ctx_destroy": (c_int, [vp]),
        "gpe_mem_info": (c_int, [vp, POINTER(c_size_t), POINTER(c_size_t)]),
        "gpe_device_info": (c_int, [vp, c_char_p, i32, POINTER(i32), POINTER(i32)]),
        "gpe_solve_create": (c_int, [vp, POINTER(SolveDesc), POINTER(vp)]),
        "gpe_solve_destroy": (c_int, [vp]),
        "gpe_solve_footprint": (c_int, [POINTER(SolveDesc), POINTER(c_size_t), POINTER(c_size_t)]),
        "gpe_psi_upload": (c_int, [vp, dp, i64, i64]),
        "gpe_psi_download": (c_int, [vp, dp, i64, i64]),
        "gpe_psi_copy": (c_int, [vp, vp]),
        "gpe_psi_gaussian": (c_int, [vp, c_double]),
        "gpe_potential_static": (c_int, [vp, i32, POINTER(i32), i32, dp, i32, c_double]),
        "gpe_potential_dynamic": (c_int, [vp, i32, POINTER(i32), i32, dp, i32]),
        "gpe_volume_fill": (c_int, [vp, i32, c_double]),
        "gpe_volume_upload": (c_int, [vp, i32, dp, i64, i64, i32]),
        "gpe_step": (c_int, [vp, i64, dp]),
        "gpe_step_unfused": (c_int, [vp, i64, dp]),
        "gpe_norm_history": (c_int, [vp, dp, i64]),
        "gpe_norm": (c_int, [vp, dp]),
        "gpe_normalise": (c_int, [vp]),
        "gpe_diagnostics": (c_int, [vp, c_double, dp, POINTER(i64), dp, i32]),
        "gpe_spectrum_view": (c_int, [vp, POINTER(i64), dp, i32]),
        "gpe_sync": (c_int, [vp]),
        "gpe_mark": (c_int, [vp, i32]),
        "gpe_mark_elapsed": (c_int, [vp, dp]),
        "gpe_step_timed": (c_int, [vp, i64, dp, dp, POINTER(i64)]),
        "gpe_step_stamped": (c_int, [vp, i64, dp, dp, POINTER(i64)]),
        "gpe_step_kernel_times": (c_int, [vp, i64, dp, dp, POINTER(i64)]),
        "gpe_transform": (c_int, [vp]),
        "gpe_volume_download": (c_int, [vp, i32, dp, i64, i64]),
        "gpe_step_noalloc": (c_int, [vp, i64, dp]),
        "gpe_solve_set_plan": (c_int, [vp, i32]),
        "gpe_solve_get_plan": (c_int, [vp, POINTER(i32)]),
        "gpe_transform_download": (c_int, [vp, dp, i64, i64]),
        "gpe_monitor_images": (c_int, [vp, i32, dp, dp, dp, dp, dp, dp]),
        "gpe_bench_copy": (c_int, [vp, i64, i32, dp]),
        "gpe_angular_momentum": (c_int, [vp, dp]),
        "gpe_batch_create": (c_int, [vp, POINTER(vp), i32, POINTER(vp)]),
        "gpe_batch_destroy": (c_int, [vp]),
        "gpe_batch_step": (c_int, [vp, i64, i32]),
        "gpe_batch_step_timed": (c_int, [vp, i64, dp, POINTER(i64)]),
        "gpe_batch_launch": (c_int, [vp, i64, i32]),
        "gpe_batch_wait": (c_int, [vp]),
        "gpe_batch_done": (c_int, [vp, POINTER(c_int32)]),
        "gpe_pass_variants": (c_int, [i32, i32, POINTER(i32)]),
        "gpe_pass_variant_info": (c_int, [i32, i32, i32, POINTER(i32), POINTER(i32), POINTER(i32)]),
        "gpe_ctx_set_variants": (c_int, [vp, i32, i32, i32, POINTER(i32), i32]),
        "gpe_solve_set_variants": (c_int, [vp, POINTER(i32), i32]),
        "gpe_solve_get_variants": (c_int, [vp, POINTER(i32), i32]),
        "gpe_batch_step_times": (c_int, [vp, i64, dp, i64]),
        "gpe_batch_norm_history": (c_int, [vp, dp, i64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded library (raises GpeError if it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY.exists():
                raise GpeError(GPE_ERR_STATE, f"{LIBRARY} is not built; run gpe_simulator_amd.build")
            handle = ctypes.CDLL(str(LIBRARY))
            _declare(handle)
            _lib = handle
    return _lib


def check(code):
    if code == GPE_OK:
        return
    message = lib().gpe_last_error().decode(errors="replace")
    if code == GPE_ERR_INVALID:
        raise GpeInvalid(code, message)
    raise GpeError(code, message)


def dptr(array):
    return array.ctypes.data_as(POINTER(c_double))


def i32ptr(array):
    return array.ctypes.data_as(POINTER(c_int32))


def i64ptr(array):
    return array.ctypes.data_as(POINTER(c_int64))


class Context:
    """A device context (one HIP device)."""

    def __init__(self, device=0):
        handle = c_void_p()
        check(lib().gpe_ctx_create(int(device), byref(handle)))
        self.handle = handle
        self.device = device
        self.variant_defaults = {}     # (side, precision, kind) -> variants a caller set on this context
        self.tuned_defaults = {}       # (side, precision, kind) -> variants the autotuner set
        self.tunings = {}              # autotuner results cached on this context

    def mem_info(self):
        free, total = c_size_t(), c_size_t()
        check(lib().gpe_mem_info(self.handle, byref(free), byref(total)))
        return free.value, total.value

    def set_variants(self, n, precision, kind, variants, tuned=False):
        """Pass variants (one per kind in PASS_KINDS) for solves created later.
        The autotuner passes tuned=True; a caller's own choice (tuned=False)
        is never replaced by tuning."""
        v = np.ascontiguousarray(variants, dtype=np.int32)
        check(lib().gpe_ctx_set_variants(self.handle, int(n), int(precision), int(kind), i32ptr(v), v.size))
        key = (int(n), int(precision), int(kind))
        if tuned:
            self.tuned_defaults[key] = [int(x) for x in v]
        else:
            self.variant_defaults[key] = [int(x) for x in v]

    def default_variants(self, n, precision, kind):
        """Variants new solves of (side, precision, kind) start from."""
        key = (int(n), int(precision), int(kind))
        return self.variant_defaults.get(key) or self.tuned_defaults.get(key) or [0] * len(PASS_KINDS)

    def device_info(self):
        """{"name", "runtime_version", "driver_version"} of the context's device."""
        name, rt, drv = ctypes.create_string_buffer(256), c_int32(), c_int32()
        check(lib().gpe_device_info(self.handle, name, len(name), byref(rt), byref(drv)))
        return dict(name=name.value.decode(errors="replace"), runtime_version=rt.value,
                    driver_version=drv.value)

    def bench_copy(self, nbytes, reps=20):
        out = c_double()
        check(lib().gpe_bench_copy(self.handle, int(nbytes), int(reps), byref(out)))
        return out.value

    def close(self):
        if self.handle:
            lib().gpe_ctx_destroy(self.handle)
            self.handle = None


_default_ctx = None


def set_jit(enable):
    """Time-dependent programs registered from now on run as kernels compiled
    at run time (True, the default) or through the on-device interpreter."""
    check(lib().gpe_set_jit(1 if enable else 0))


def jit_stats():
    """{"compiles", "launches", "failures"} of the run-time compiled programs
    in this process."""
    c, l, f = c_int64(), c_int64(), c_int64()
    check(lib().gpe_jit_stats(byref(c), byref(l), byref(f)))
    return {"compiles": c.value, "launches": l.value, "failures": f.value}


def default_context():
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
    return _default_ctx


def make_desc(n, precision, kind, dt, beta, x_array, k_array, kcut=None, kinetic_coef=1.0):
    x = np.ascontiguousarray(x_array, dtype=np.float64)
    k = np.ascontiguousarray(k_array, dtype=np.float64)
    desc = SolveDesc(n=int(n), precision=int(precision), kind=int(kind),
                     kcut_enable=1 if kcut is not None else 0, dt=float(dt), beta=float(beta),
                     kcut=float(kcut or 0.0), kinetic_coef=float(kinetic_coef),
                     x_array=dptr(x), k_array=dptr(k))
    desc._keep = (x, k)
    return desc


def footprint(n, precision, kind=REAL):
    x = np.zeros(max(2, n))
    desc = make_desc(n, precision, kind, 1.0, 0.0, x, x)
    resident, diag = c_size_t(), c_size_t()
    check(lib().gpe_solve_footprint(byref(desc), byref(resident), byref(diag)))
    return resident.value, diag.value


def pass_variants(n, precision):
    """[(threads per pencil, pencils per block, waves per SIMD)] of the
    compiled pass variants of (side, precision)."""
    count = c_int32()
    check(lib().gpe_pass_variants(int(n), int(precision), byref(count)))
    out = []
    for v in range(count.value):
        tp, b, w = c_int32(), c_int32(), c_int32()
        check(lib().gpe_pass_variant_info(int(n), int(precision), v, byref(tp), byref(b), byref(w)))
        out.append((tp.value, b.value, w.value))
    return out


def nbins_for(n):
    import math
    return int(math.floor(math.sqrt(3) * (n // 2) + 0.5)) + 1


class Solve:
    """One device-resident split-step solve (psi + tables) behind the C ABI."""

    def __init__(self, n, precision, kind, dt, beta, x_array, k_array, kcut=None, kinetic_coef=1.0,
                 ctx=None):
        self.ctx = ctx or default_context()
        self.n, self.precision, self.kind = int(n), int(precision), int(kind)
        self.dt, self.beta = float(dt), float(beta)
        self.nbins = nbins_for(self.n)
        desc = make_desc(n, precision, kind, dt, beta, x_array, k_array, kcut, kinetic_coef)
        handle = c_void_p()
        check(lib().gpe_solve_create(self.ctx.handle, byref(desc), byref(handle)))
        self.handle = handle
        self.has_dynamic = False

    # -- lifecycle
    def close(self):
        if getattr(self, "handle", None):
            lib().gpe_solve_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self):
        return self.n ** 3

    # -- field transfer
    def upload(self, psi, offset=0):
        data = np.ascontiguousarray(psi, dtype=np.complex128).reshape(-1)
        check(lib().gpe_psi_upload(self.handle, dptr(data.view(np.float64)), int(offset), data.size))

    def download(self, offset=0, count=None):
        count = self.size - offset if count is None else count
        out = np.empty(count, dtype=np.complex128)
        check(lib().gpe_psi_download(self.handle, dptr(out.view(np.float64)), int(offset), int(count)))
        return out

    def to_numpy(self):
        return self.download().reshape((self.n,) * 3)

    def copy_from(self, other):
        check(lib().gpe_psi_copy(self.handle, other.handle))

    def gaussian(self, sigma):
        check(lib().gpe_psi_gaussian(self.handle, float(sigma)))

    # -- potentials
    def potential_static(self, target, program, scale=1.0):
        code, consts = program
        code = np.ascontiguousarray(code, dtype=np.int32)
        consts = np.ascontiguousarray(consts, dtype=np.float64)
        check(lib().gpe_potential_static(self.handle, int(target), i32ptr(code), code.size, dptr(consts),
                                         consts.size, float(scale)))

    def potential_dynamic(self, target, program):
        code, consts = program
        code = np.ascontiguousarray(code, dtype=np.int32)
        consts = np.ascontiguousarray(consts, dtype=np.float64)
        check(lib().gpe_potential_dynamic(self.handle, int(target), i32ptr(code), code.size, dptr(consts),
                                          consts.size))
        self.has_dynamic = True

    def volume_fill(self, target, value):
        check(lib().gpe_volume_fill(self.handle, int(target), float(value)))

    def volume_upload(self, target, values, offset=0, accumulate=False):
        data = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        check(lib().gpe_volume_upload(self.handle, int(target), dptr(data), int(offset), data.size,
                                      1 if accumulate else 0))

    # -- stepping
    def step(self, nsteps=1, t_mid=None):
        tm = None
        if t_mid is not None:
            tm = np.ascontiguousarray(t_mid, dtype=np.float64)
            if tm.size < nsteps:
                raise ValueError("t_mid needs one time per step")
        check(lib().gpe_step(self.handle, int(nsteps), dptr(tm) if tm is not None else None))

    def step_unfused(self, nsteps=1, t_mid=None):
        """Real-time steps with the reference's launch structure (9 sweeps of
        psi per step instead of 4); same results to rounding."""
        tm = None
        if t_mid is not None:
            tm = np.ascontiguousarray(t_mid, dtype=np.float64)
            if tm.size < nsteps:
                raise ValueError("t_mid needs one time per step")
        check(lib().gpe_step_unfused(self.handle, int(nsteps), dptr(tm) if tm is not None else None))

    def step_noalloc(self, nsteps=1, t_mid=None):
        """gpe_step_noalloc: gpe_step without any device allocation (imaginary
        time keeps its norm partials in the library's static device array)."""
        tm = None
        if t_mid is not None:
            tm = np.ascontiguousarray(t_mid, dtype=np.float64)
            if tm.size < nsteps:
                raise ValueError("t_mid needs one time per step")
        check(lib().gpe_step_noalloc(self.handle, int(nsteps), dptr(tm) if tm is not None else None))

    def step_timed(self, nsteps, t_mid=None):
        tm = None if t_mid is None else np.ascontiguousarray(t_mid, dtype=np.float64)
        ms = np.zeros(len(PASS_KINDS))
        cnt = np.zeros(len(PASS_KINDS), dtype=np.int64)
        check(lib().gpe_step_timed(self.handle, int(nsteps), dptr(tm) if tm is not None else None, dptr(ms),
                                   i64ptr(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PASS_KINDS)}

    def step_stamped(self, nsteps, t_mid=None):
        """gpe_step with device wall-clock spans per launch:
        {kind: (summed us, launches)}."""
        tm = None if t_mid is None else np.ascontiguousarray(t_mid, dtype=np.float64)
        us = np.zeros(len(PASS_KINDS))
        cnt = np.zeros(len(PASS_KINDS), dtype=np.int64)
        check(lib().gpe_step_stamped(self.handle, int(nsteps), dptr(tm) if tm is not None else None, dptr(us),
                                     i64ptr(cnt)))
        return {k: (float(us[i]), int(cnt[i])) for i, k in enumerate(PASS_KINDS)}

    def step_kernel_times(self, nsteps, t_mid=None):
        """gpe_step with each launch's own dispatch start and end timestamps
        (the durations a kernel trace reports): {kind: (summed us, launches)}."""
        tm = None if t_mid is None else np.ascontiguousarray(t_mid, dtype=np.float64)
        us = np.zeros(len(PASS_KINDS))
        cnt = np.zeros(len(PASS_KINDS), dtype=np.int64)
        check(lib().gpe_step_kernel_times(self.handle, int(nsteps), dptr(tm) if tm is not None else None,
                                          dptr(us), i64ptr(cnt)))
        return {k: (float(us[i]), int(cnt[i])) for i, k in enumerate(PASS_KINDS)}

    def set_plan(self, plan):
        """Step plan: PLAN_SEPARABLE (default), PLAN_FOUR_SWEEP or PLAN_PLANE."""
        check(lib().gpe_solve_set_plan(self.handle, int(plan)))

    @property
    def plan(self):
        out = c_int32()
        check(lib().gpe_solve_get_plan(self.handle, byref(out)))
        return out.value

    def set_variants(self, variants):
        """Pass variant per kind (sequence in PASS_KINDS order, or a dict)."""
        if isinstance(variants, dict):
            current = self.variants()
            variants = [variants.get(k, current[i]) for i, k in enumerate(PASS_KINDS)]
        v = np.ascontiguousarray(variants, dtype=np.int32)
        check(lib().gpe_solve_set_variants(self.handle, i32ptr(v), v.size))

    def variants(self):
        v = np.zeros(len(PASS_KINDS), dtype=np.int32)
        check(lib().gpe_solve_get_variants(self.handle, i32ptr(v), v.size))
        return [int(x) for x in v]

    def norm_history(self, count):
        out = np.empty(int(count))
        check(lib().gpe_norm_history(self.handle, dptr(out), int(count)))
        return out

    def sync(self):
        check(lib().gpe_sync(self.handle))

    def mark(self, which):
        """Record timing mark 0 (start) or 1 (stop) on the solve's stream."""
        check(lib().gpe_mark(self.handle, int(which)))

    def marked_ms(self):
        """Milliseconds between the marks, after waiting for the stop mark."""
        out = c_double()
        check(lib().gpe_mark_elapsed(self.handle, byref(out)))
        return out.value

    # -- reductions
    def norm(self):
        out = c_double()
        check(lib().gpe_norm(self.handle, byref(out)))
        return out.value

    def angular_momentum(self):
        out = c_double()
        check(lib().gpe_angular_momentum(self.handle, byref(out)))
        return out.value

    def volume(self, target):
        """The FP64 static volume (TGT_VOL real part, TGT_VOL_IM imaginary
        part) as an (n, n, n) host array."""
        out = np.empty(self.n ** 3)
        check(lib().gpe_volume_download(self.handle, int(target), dptr(out), 0, out.size))
        return out.reshape((self.n,) * 3)

    def spectrum(self):
        """fftn(psi), raw and unnormalised, computed on the device and copied
        to the host as a complex128 cube."""
        check(lib().gpe_transform(self.handle))
        out = np.empty(self.n ** 3, dtype=np.complex128)
        check(lib().gpe_transform_download(self.handle, dptr(out.view(np.float64)), 0, out.size))
        return out.reshape((self.n,) * 3)

    def monitor_images(self, names):
        """Raw device sums of the monitor images named in `names` (a subset of
        MONITOR_IMAGES): {name: (n, n) array}; plane_mid is complex128."""
        which, outs = 0, {}
        for bit, name in enumerate(MONITOR_IMAGES):
            if name in names:
                which |= 1 << bit
                outs[name] = np.zeros((self.n, self.n), dtype=np.complex128 if name == "plane_mid" else np.float64)
        unknown = set(names) - set(MONITOR_IMAGES)
        if unknown:
            raise ValueError(f"unknown monitor images {sorted(unknown)}")
        ptrs = [dptr(outs[name].view(np.float64)) if name in outs else None for name in MONITOR_IMAGES]
        check(lib().gpe_monitor_images(self.handle, which, *ptrs))
        return outs

    def normalise(self):
        check(lib().gpe_normalise(self.handle))

    def diagnostics(self, t=0.0, spectrum=True):
        scal = np.zeros(8)
        counts = np.zeros(self.nbins, dtype=np.int64) if spectrum else None
        occ = np.zeros(self.nbins) if spectrum else None
        check(lib().gpe_diagnostics(self.handle, float(t), dptr(scal),
                                    i64ptr(counts) if spectrum else None,
                                    dptr(occ) if spectrum else None, self.nbins))
        names = ("norm", "E_kin", "E_pot", "E_int", "E", "mu", "parseval")
        result = {name: float(scal[i]) for i, name in enumerate(names)}
        if spectrum:
            result["counts"] = counts
            result["occupation"] = occ
            view_counts = np.zeros(VIEW_BINS, dtype=np.int64)
            view_occ = np.zeros(VIEW_BINS)
            check(lib().gpe_spectrum_view(self.handle, i64ptr(view_counts), dptr(view_occ), VIEW_BINS))
            result["view_counts"] = view_counts
            result["view_occupation"] = view_occ
        return result


class Batch:
    """Equal-shape solves advanced together, one launch per pass (C ABI
    gpe_batch_*).  Members keep their own fields and parameters."""

    def __init__(self, solves, ctx=None):
        self.solves = list(solves)
        self.ctx = ctx or self.solves[0].ctx
        arr = (c_void_p * len(self.solves))(*[s.handle for s in self.solves])
        handle = c_void_p()
        check(lib().gpe_batch_create(self.ctx.handle, arr, len(self.solves), byref(handle)))
        self.handle = handle

    def step(self, nsteps, graph=True, t_mid=None):
        """Advance every member nsteps.  Members with time-dependent programs
        need t_mid: shape (nsteps,) shared by all members, or
        (members, nsteps) per member; such batches are not graph-captured."""
        if t_mid is None:
            check(lib().gpe_batch_step(self.handle, int(nsteps), 1 if graph else 0))
            return
        tm = np.ascontiguousarray(t_mid, dtype=np.float64)
        if tm.ndim == 1:
            if tm.size < nsteps:
                raise GpeInvalid(GPE_ERR_INVALID, "t_mid needs one time per step")
            stride = 0
        else:
            if tm.shape[0] != len(self.solves) or tm.shape[1] < nsteps:
                raise GpeInvalid(GPE_ERR_INVALID, "t_mid must be (members, nsteps)")
            stride = tm.shape[1]
        check(lib().gpe_batch_step_times(self.handle, int(nsteps), dptr(tm), int(stride)))

    def launch(self, nsteps, graph=True):
        check(lib().gpe_batch_launch(self.handle, int(nsteps), 1 if graph else 0))

    def wait(self):
        check(lib().gpe_batch_wait(self.handle))

    def done(self):
        """True when everything launched on the batch has finished (no wait)."""
        out = c_int32()
        check(lib().gpe_batch_done(self.handle, byref(out)))
        return bool(out.value)

    def norm_history(self, count):
        """Per-step norms of the last `count` steps, shape (members, count)."""
        out = np.empty((len(self.solves), int(count)))
        check(lib().gpe_batch_norm_history(self.handle, dptr(out), int(count)))
        return out

    def step_timed(self, nsteps):
        ms = np.zeros(len(PASS_KINDS))
        cnt = np.zeros(len(PASS_KINDS), dtype=np.int64)
        check(lib().gpe_batch_step_timed(self.handle, int(nsteps), dptr(ms), i64ptr(cnt)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PASS_KINDS)}

    def close(self):
        if getattr(self, "handle", None):
            lib().gpe_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
