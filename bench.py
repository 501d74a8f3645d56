"""Benchmark: 3D split-step stirring run at 256^3 on one MI355X.

Workload (reference units: beta = 2g, dt = dt'/2 and Omega = 2 Omega' from the
oscillator-unit parameters g = 2000, dt' = 5e-4, Omega' = 0.5):
256^3 grid on [-20, 20)^3, static trap V = x^2 + y^2 + z^2, beta = 4000,
Gaussian stirring beam 10*exp(-((x-4cos t)^2 + (y-4sin t)^2)/2), dt = 2.5e-4,
complex64 field with FP64 norm/energy and FP64 time, |psi(k)|^2 shell
histogram every 500 steps.  A step is one Strang step over the whole grid;
the state is a seeded synthetic Thomas-Fermi profile plus noise, resident in
HBM before timing starts.

Beside it, the same line reports: the complex128 run of the same workload,
the same complex128 run with the reference's launch structure (one launch per
reference kernel and transform axis, 9 sweeps of psi per step instead of 4),
per-kernel launch durations (each dispatch's own start and end timestamps),
the oracle on the host as the CPU baseline (one thread, and a threaded run
whose final state is the FP64 reference the complex64 and complex128 engines
are held to after the same steps), a
short seeded serving stream (p50/p99 time-to-result), the 64-solve 96^3
batched ground-state sweep against sequential complex128 solves, and a
16-speed stirring sweep at 128^3 batched against the same runs in turn.

Prints ONE JSON line.  `value` = grid-point-steps/s of the complex64 run.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N = 256
LENGTH = 40.0
BETA = 4000.0
DT = 2.5e-4
STATIC = "x**2 + y**2 + z**2"
BEAM = "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)"
NK_INTERVAL = 500
HBM_PEAK_GBS = 8000.0
# steady-state psi sweeps per real-time step: the separable plan (the
# default) runs one pass per axis (z, y, x), each FFT -> 1D kinetic factor ->
# inverse FFT, with the local half-steps fused into the z pass; the four-sweep
# plan (3D FFT, kinetic factor, inverse 3D FFT in passes z, y, x, y) needs 4
PASSES_PER_STEP = 3
PASSES_FOUR_SWEEP = 4
TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_stirring256_c64.txt"
STATS_FILE = ROOT / "profiles" / "kernel_stats_stirring256.md"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def synthetic_state(x, seed=0):
    """Thomas-Fermi profile sqrt(max(mu_TF - V, 0)/beta) + 1e-3 complex noise."""
    rng = np.random.default_rng(seed)
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    V = X * X + Y * Y + Z * Z
    mu_tf = (15 * BETA / (8 * np.pi)) ** 0.4        # TF chemical potential for V = r^2
    psi = np.sqrt(np.maximum(mu_tf - V, 0) / BETA).astype(np.complex128)
    psi += 1e-3 * (rng.standard_normal(psi.shape) + 1j * rng.standard_normal(psi.shape))
    dx = x[1] - x[0]
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * dx ** 3)
    return psi


def run_engine(precision, psi0, x, k, steps, warmup, unfused=False, plan=None):
    """Timed steps of the workload; `unfused` runs the reference-structured
    step (one launch per reference kernel and axis) instead of the fused
    plan; `plan` overrides the solve's default step plan."""
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    from gpe_simulator_amd import tune
    solve = make_solve(N, x, k, "real", DT, BETA, static=STATIC, dynamic=BEAM, precision=precision)
    if plan is not None and plan != solve.plan:
        tune.ensure_tuned(N, solve.precision, solve.kind, plan=plan)
        fresh = make_solve(N, x, k, "real", DT, BETA, static=STATIC, dynamic=BEAM, precision=precision)
        solve.close()
        solve = fresh
        solve.set_plan(plan)
    solve.upload(psi0)
    tm = midpoint_times(0, warmup + steps, DT)
    step = solve.step_unfused if unfused else solve.step
    step(warmup, tm[:warmup])
    solve.diagnostics(warmup * DT)
    solve.sync()
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        chunk = min(NK_INTERVAL, steps - done)
        step(chunk, tm[warmup + done: warmup + done + chunk])
        done += chunk
        if done % NK_INTERVAL == 0:
            solve.diagnostics((warmup + done) * DT)      # nk histogram + FP64 scalars (syncs)
    solve.sync()
    elapsed = time.perf_counter() - t0
    diag = solve.diagnostics((warmup + steps) * DT, spectrum=False)
    return solve, elapsed, diag


def kernel_timing(solve, steps, repeats=3):
    """Per-pass launch durations over `steps` steps of the workload, two ways:
    {kind: us per launch} from each launch's own dispatch start and end
    timestamps (a HIP event pair handed to the dispatch, nothing recorded
    between launches: the duration a kernel trace reports; median of
    `repeats` runs), and {kind: us per launch} as spans from one launch's end
    to the next one's (chained stop events; includes the gap between
    launches)."""
    import statistics
    from gpe_simulator_amd.engine import midpoint_times
    tm = midpoint_times(0, steps, DT)
    runs = [solve.step_kernel_times(steps, tm) for _ in range(repeats)]
    own = {kk: statistics.median(r[kk][0] / r[kk][1] for r in runs) for kk in runs[0] if runs[0][kk][1]}
    launches = {kk: runs[0][kk][1] for kk in own}
    events = solve.step_timed(steps, tm)
    spans = {kk: v[0] / v[1] * 1e3 for kk, v in events.items() if v[1]}
    return own, launches, spans


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary (FETCH_SIZE
    + WRITE_SIZE, each scaled by the copy-kernel calibration), else None."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        m = re.match(r"\s*traffic\s+(\w+)\s+([0-9.eE+]+)\s*$", line)
        if m and m.group(1) == kernel:
            return float(m.group(2))
    return None


def traced_avg_us(kind, tp, b, waves):
    """(average duration in us, row label) of the complex64 256^3 `kind` pass
    in the committed rocprofv3 kernel-trace summary of this workload: the row
    of variant (tp, b, waves) if the profiled run tuned the same one, else the
    kind's row with the most calls (the variant that run tuned); (None, None)
    without a summary."""
    if not STATS_FILE.exists():
        return None, None
    want = f"{kind} c64 n{N} (TP={tp},B={b},W={waves})"
    rows = []
    for line in STATS_FILE.read_text().splitlines():
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 4 and cells[1].startswith(f"{kind} c64 n{N} (") and cells[2].isdigit():
            rows.append((cells[1], int(cells[2]), float(cells[3])))
    for lab, _, us in rows:
        if lab == want:
            return us, lab
    if rows:
        lab, _, us = max(rows, key=lambda r: r[1])
        return us, lab
    return None, None


def serving(count, seed=0):
    """Seeded Poisson stream of mixed requests (sides 64..256, 70% ground
    states to |dmu| < 2e-9, 30% 2000-step real-time runs, 80% complex64).
    The pass variants of every shape in the stream are tuned before the first
    arrival (reported as warm_up_s), as a deployed server would do once."""
    from gpe_simulator_amd.serve import Server, request_stream
    reqs = request_stream(count, rate_per_s=4.0, seed=seed)
    server = Server(check_every=50)
    # one-time tuning of every shape the stream can hold, before the first arrival
    shapes = {(r.n, p, r.kind) for r in reqs
              for p in {r.precision} | ({"c128"} if r.kind == "imaginary" else set())}
    warm_s = server.warm(sorted(shapes))
    outcomes, stats = server.run(reqs)
    stats["warm_up_s"] = warm_s
    stats["converged"] = sum(o.converged for o in outcomes)
    stats["stream"] = f"{count} requests, Poisson 4/s, seed {seed}"
    return stats


def sweep(count, every=1, seed=0):
    """64 ground states at 96^3 (seeded g, omega_z, widths) batched, against
    sequential complex128 solves of every `every`-th request (all of them by
    default; with every > 1 the sequential time is extrapolated).  Pass
    variants are tuned for both shapes before either is timed."""
    from gpe_simulator_amd import _lib, tune
    from gpe_simulator_amd.sweep import config_sweep, solve_sequential, sweep_batched
    reqs = config_sweep(count=count, seed=seed)
    for prec in (_lib.C64, _lib.C128):
        tune.ensure_tuned(reqs[0].n, prec, _lib.IMAG, batch=min(count, 64))
    tune.ensure_tuned(reqs[0].n, _lib.C128, _lib.IMAG)
    t0 = time.perf_counter()
    res = sweep_batched(reqs)
    batched = time.perf_counter() - t0
    sample = list(range(0, count, every))
    t0 = time.perf_counter()
    seq = {i: solve_sequential(reqs[i]) for i in sample}
    per_solve = (time.perf_counter() - t0) / len(sample)
    return {"requests": count, "grid": reqs[0].n, "batched_s": batched,
            "sequential": f"{len(sample)} of {count} requests solved one after another"
                          + ("" if every == 1 else " (time extrapolated)"),
            "sequential_s": per_solve * count, "speedup": per_solve * count / batched,
            "converged": sum(r.converged for r in res),
            "steps_batched_mean": float(np.mean([r.steps for r in res])),
            "steps_sequential_mean": float(np.mean([s.steps for s in seq.values()])),
            "mu_rel_diff_vs_sequential": max(abs(res[i].mu - s.mu) / abs(s.mu) for i, s in seq.items())}


def stirring(count, n=128, steps=500):
    """`count` stirring runs at different stirrer speeds (n^3, g = 500,
    beam amplitude 5 on radius 3, dt' = 5e-4, complex64), from one seeded
    state, batched against the same runs one after another."""
    from gpe_simulator_amd.sweep import stir_batched, stir_sequential, stirrer_runs
    from gpe_simulator_amd.grids import get_grids
    runs = stirrer_runs(np.linspace(0.1, 1.5, count), n=n, length=24.0, g=500.0, steps=steps)
    x, _ = get_grids(n, 24.0)
    rng = np.random.default_rng(1)
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    psi0 = np.exp(-(X * X + Y * Y + Z * Z) / 8).astype(np.complex128)
    psi0 += 1e-3 * (rng.standard_normal(psi0.shape) + 1j * rng.standard_normal(psi0.shape))
    stir_batched(runs[:2], psi0)                       # tuning and first-use setup outside the timing
    t0 = time.perf_counter()
    batched = stir_batched(runs, psi0)
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    seq = [stir_sequential(r, psi0) for r in runs]
    ts = time.perf_counter() - t0
    return {"runs": count, "grid": n, "steps": steps, "batched_s": tb, "sequential_s": ts, "speedup": ts / tb,
            "identical_to_sequential": all(a == b for a, b in zip(batched, seq)),
            "Lz_range": [min(r["Lz"] for r in batched), max(r["Lz"] for r in batched)]}


def probe_small_grids():
    """Per shape (64, 96, 128; both precisions; imaginary time, real time,
    and real time with a momentum cutoff at 0.6 k_max): device time per step
    of a solo solve between two marks, the sum of its kernels' own durations
    per step, and host time per step of the same solve as a one-member batch
    replayed from a hipGraph and launched eagerly."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.grids import get_grids
    rows = []
    for n in (64, 96, 128):
        for prec in ("c64", "c128"):
            for kind in ("real", "imaginary", "real_cutoff"):
                x, k = get_grids(n, 20.0)
                kcut = 0.6 * np.pi * n / 20.0 if kind == "real_cutoff" else None
                s = make_solve(n, x, k, kind.split("_")[0], 5e-4, 300.0, static="x**2 + y**2 + z**2", kcut=kcut,
                               precision=prec)
                s.gaussian(1.0)
                s.step(20)
                steps = 400
                s.mark(0)
                s.step(steps)
                s.mark(1)
                solo = s.marked_ms() * 1e3 / steps
                kernels = sum(v[0] for v in s.step_kernel_times(steps).values()) / steps
                b = _lib.Batch([s])
                b.step(steps, graph=True)
                t0 = time.perf_counter()
                b.step(steps, graph=True)
                graph = (time.perf_counter() - t0) * 1e6 / steps
                t0 = time.perf_counter()
                b.step(steps, graph=False)
                eager = (time.perf_counter() - t0) * 1e6 / steps
                b.close()
                rows.append({"n": n, "precision": prec, "kind": kind, "plan": s.plan, "solo_us": solo,
                             "kernels_us": kernels, "batch1_graph_us": graph, "batch1_eager_us": eager})
                s.close()
    return rows


def probe_mixed_sides(steps=200):
    """Solves of different sides advance as one batch per side, each on its
    own stream.  Grid-point-steps/s of three concurrent batches (8 x 64^3,
    8 x 96^3, 2 x 128^3; imaginary time, complex64) against each batch alone
    and against one equal-side batch of about the same points (16 x 96^3)."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.grids import get_grids

    def batch(n, count):
        x, k = get_grids(n, 20.0)
        solves = [make_solve(n, x, k, "imaginary", 5e-4, 300.0, static="x**2 + y**2 + z**2", precision="c64")
                  for _ in range(count)]
        for s in solves:
            s.gaussian(1.0)
        return _lib.Batch(solves), solves, n ** 3 * count

    groups = [batch(64, 8), batch(96, 8), batch(128, 2)]
    single = batch(96, 16)
    out = {"steps": steps}
    for b, _, _ in groups + [single]:
        b.step(steps, graph=True)                      # capture and warm
    for b, _, pts in groups + [single]:
        t0 = time.perf_counter()
        b.step(steps, graph=True)
        out.setdefault("alone", []).append({"points": pts, "gpts_per_s": pts * steps / (time.perf_counter() - t0)})
    t0 = time.perf_counter()
    for b, _, _ in groups:
        b.launch(steps, graph=True)
    for b, _, _ in groups:
        b.wait()
    el = time.perf_counter() - t0
    total = sum(p for _, _, p in groups)
    out["concurrent"] = {"points": total, "gpts_per_s": total * steps / el,
                         "sum_of_alone_s": sum(p * steps / a["gpts_per_s"] for (_, _, p), a in
                                               zip(groups, out["alone"][:3])), "seconds": el}
    out["one_equal_side_batch"] = out["alone"][3]
    for b, solves, _ in groups + [single]:
        b.close()
        for s in solves:
            s.close()
    return out


def probe_large_grids(steps=20):
    """ms per step of the stirring workload's complex64 and complex128 runs at
    512^3 (L = 80, the headline's spacing) against 8x the 256^3 step."""
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    from gpe_simulator_amd.grids import get_grids
    out = {}
    for n, length in ((256, LENGTH), (512, 2 * LENGTH)):
        x, k = get_grids(n, length)
        for prec in ("c64", "c128"):
            s = make_solve(n, x, k, "real", DT, BETA, static=STATIC, dynamic=BEAM, precision=prec)
            s.gaussian(2.0)
            tm = midpoint_times(0, steps + 2, DT)
            s.step(2, tm[:2])
            s.mark(0)
            s.step(steps, tm[2:])
            s.mark(1)
            out[f"{prec}_{n}_ms"] = s.marked_ms() / steps
            s.close()
    for prec in ("c64", "c128"):
        out[f"{prec}_512_over_8x256"] = out[f"{prec}_512_ms"] / (8 * out[f"{prec}_256_ms"])
    return out


def cpu_baseline(psi0, samples, workers=None):
    """The FP64 oracle on a bounded sample of the workload: numpy.fft on one
    thread, or scipy.fft with `workers` threads.  Steps 0..samples from psi0
    (t_i = i*dt); step 0 is the untimed warm-up.  Returns (rate, elapsed,
    final psi, the oracle's FP64 diagnostics of it)."""
    from oracle import gpe_oracle as orc
    orc.use_threads(workers)
    try:
        problem = orc.Problem(N, LENGTH, "real", DT, BETA, static=STATIC, dynamic=BEAM)
        psi = psi0.copy()
        psi = problem.step(psi, 0.0)                     # warm-up step
        t0 = time.perf_counter()
        for i in range(1, samples + 1):
            psi = problem.step(psi, i * DT)
        elapsed = time.perf_counter() - t0
        energies = problem.energies(psi, (samples + 1) * DT)
    finally:
        orc.use_threads(None)
    return N ** 3 * samples / elapsed, elapsed, psi, energies


def accuracy_vs_oracle(psi0, x, k, steps, reference, energies):
    """The complex64 and complex128 engines after `steps` steps of the
    workload from psi0, against the FP64 oracle's state and diagnostics after
    the same steps (computed in the CPU-baseline leg)."""
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    dx3 = float(x[1] - x[0]) ** 3
    n0 = float(np.sum(np.abs(psi0) ** 2) * dx3)
    rho_ref = np.abs(reference) ** 2
    out = {"against": f"FP64 oracle after {steps} steps of the same workload from the same state",
           "steps": steps, "mu_oracle": energies["mu"], "E_oracle": energies["E"],
           "norm_drift_oracle": abs(energies["norm"] - n0) / n0}
    for precision in ("c64", "c128"):
        solve = make_solve(N, x, k, "real", DT, BETA, static=STATIC, dynamic=BEAM, precision=precision)
        solve.upload(psi0)
        solve.step(steps, midpoint_times(0, steps, DT))
        d = solve.diagnostics(steps * DT, spectrum=False)
        psi = solve.to_numpy()
        solve.close()
        out[precision] = {
            "mu_rel": abs(d["mu"] - energies["mu"]) / abs(energies["mu"]),
            "E_rel": abs(d["E"] - energies["E"]) / abs(energies["E"]),
            "norm_drift": abs(d["norm"] - n0) / n0,
            "psi_rel_l2": float(np.linalg.norm(psi - reference) / np.linalg.norm(reference)),
            "density_rel_l2": float(np.linalg.norm(np.abs(psi) ** 2 - rho_ref) / np.linalg.norm(rho_ref))}
    return out


def accuracy(psi0, x, run64, run128):
    """complex64 against the complex128 engine after the whole timed run (the
    complex128 path is held to the FP64 oracle by the parity tests and by
    accuracy_vs_oracle over the first steps)."""
    dx3 = float(x[1] - x[0]) ** 3
    n0 = float(np.sum(np.abs(psi0) ** 2) * dx3)
    (psi64, d64), (psi128, d128) = run64, run128
    rho64, rho128 = np.abs(psi64) ** 2, np.abs(psi128) ** 2
    return {"against": "complex128 engine after the same steps",
            "mu_rel": abs(d64["mu"] - d128["mu"]) / abs(d128["mu"]),
            "E_rel": abs(d64["E"] - d128["E"]) / abs(d128["E"]),
            "density_rel_l2": float(np.linalg.norm(rho64 - rho128) / np.linalg.norm(rho128)),
            "norm_drift_c64": abs(d64["norm"] - n0) / n0,
            "norm_drift_c128": abs(d128["norm"] - n0) / n0}


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=2000)
    parser.add_argument("--warmup", type=int, default=100)
    parser.add_argument("--c128-steps", type=int, default=None, help="default: --steps")
    parser.add_argument("--timed-steps", type=int, default=50)
    parser.add_argument("--cpu-samples", type=int, default=3)
    parser.add_argument("--oracle-steps", type=int, default=25,
                        help="steps of the threaded oracle run that times the multi-core CPU baseline and "
                             "is the FP64 reference of accuracy_c64")
    parser.add_argument("--skip-cpu", action="store_true")
    parser.add_argument("--skip-c128", action="store_true")
    parser.add_argument("--unfused-steps", type=int, default=500,
                        help="steps of the reference-structured c128 run (at most the c128 steps)")
    parser.add_argument("--serve-requests", type=int, default=100, help="0 skips the serving stream")
    parser.add_argument("--sweep-count", type=int, default=64, help="0 skips the batched sweep")
    parser.add_argument("--stir-count", type=int, default=16, help="0 skips the stirring-speed sweep")
    parser.add_argument("--probe", choices=("small", "mixed", "large"), default=None,
                        help="print one JSON line of a side measurement instead of the bench line: small-grid "
                             "step cost, concurrent batches of different sides, or 512^3 against 256^3")
    args = parser.parse_args(argv)
    if args.probe:
        fn = {"small": probe_small_grids, "mixed": probe_mixed_sides, "large": probe_large_grids}[args.probe]
        print(json.dumps({"probe": args.probe, "result": fn()}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); "
              "the engine runs a solve on one GPU and has no multi-GPU path", file=sys.stderr)
        return 2

    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.grids import get_grids
    x, k = get_grids(N, LENGTH)
    log("building synthetic state")
    psi0 = synthetic_state(x)

    log(f"c64 run: {args.warmup} warm-up + {args.steps} timed steps")
    solve64, el64, diag64 = run_engine("c64", psi0, x, k, args.steps, args.warmup)
    final64 = (solve64.to_numpy(), diag64)
    value = N ** 3 * args.steps / el64
    log(f"c64: {el64 * 1e3 / args.steps:.4f} ms/step, {value:.4e} grid-pt-steps/s, mu={diag64['mu']:.9g}")

    log(f"per-kernel timing over {args.timed_steps} steps (dispatch timestamps, then chained spans)")
    own, launches, spans = kernel_timing(solve64, args.timed_steps)
    dominant = max(own, key=lambda kk: own[kk] * launches[kk])
    dom_ms = own[dominant] * 1e-3
    bytes_per_launch = N ** 3 * 2 * 8                 # one read + one write sweep of complex64 psi
    achieved = bytes_per_launch / (dom_ms * 1e-3) / 1e9
    per_kernel = {kk: {"avg_us": round(us, 3), "launches": launches[kk],
                       "span_avg_us": round(spans[kk], 3) if kk in spans else None}
                  for kk, us in own.items()}
    dom_variant = _lib.pass_variants(N, _lib.C64)[solve64.variants()[_lib.PASS_KINDS.index(dominant)]]
    dom_traced_us, dom_traced_row = traced_avg_us(dominant, *dom_variant)
    solve64.close()

    copy_bytes = 256 << 20
    ms_copy = _lib.default_context().bench_copy(copy_bytes, reps=20)
    copy_gbs = 2 * copy_bytes / (ms_copy * 1e-3) / 1e9

    result_c128, acc, four64 = None, None, None
    c128_steps = args.c128_steps or args.steps
    if not args.skip_c128:
        from gpe_simulator_amd import _lib
        log(f"c128 run: {c128_steps} steps, the same (separable) plan as the complex64 run")
        solve128, el128, diag128 = run_engine("c128", psi0, x, k, c128_steps, args.warmup)
        v128 = N ** 3 * c128_steps / el128
        result_c128 = {"value": v128, "ms_per_step": el128 * 1e3 / c128_steps, "plan": "separable",
                       "hbm_frac": v128 * PASSES_PER_STEP * 2 * 16 / (HBM_PEAK_GBS * 1e9),
                       "speedup_c64_over_c128": value / v128, "mu": diag128["mu"]}
        if c128_steps == args.steps:
            acc = accuracy(psi0, x, final64, (solve128.to_numpy(), diag128))
        solve128.close()
        log(f"four-sweep plan: c128 and c64, {c128_steps} steps each")
        solve_f, el_f, _ = run_engine("c128", psi0, x, k, c128_steps, args.warmup, plan=_lib.PLAN_FOUR_SWEEP)
        solve_f.close()
        vf = N ** 3 * c128_steps / el_f
        result_c128["four_sweep"] = {"value": vf, "ms_per_step": el_f * 1e3 / c128_steps, "sweeps_per_step": 4,
                                     "speedup_c64_over_it": value / vf}
        solve_f, el_f, _ = run_engine("c64", psi0, x, k, c128_steps, args.warmup, plan=_lib.PLAN_FOUR_SWEEP)
        solve_f.close()
        four64 = {"value": N ** 3 * c128_steps / el_f, "ms_per_step": el_f * 1e3 / c128_steps,
                  "sweeps_per_step": 4, "hbm_frac_at_64B": N ** 3 * c128_steps / el_f * 64 / (HBM_PEAK_GBS * 1e9)}
        unfused_steps = min(c128_steps, args.unfused_steps)
        log(f"reference-structured c128 run: {unfused_steps} steps")
        solve_u, el_u, diag_u = run_engine("c128", psi0, x, k, unfused_steps, args.warmup, unfused=True)
        solve_u.close()
        vu = N ** 3 * unfused_steps / el_u
        result_c128["unfused"] = {
            "value": vu, "ms_per_step": el_u * 1e3 / unfused_steps, "steps": unfused_steps,
            "launches_per_step": 9, "bytes_per_point": 9 * 2 * 16,
            "hbm_frac": vu * 9 * 2 * 16 / (HBM_PEAK_GBS * 1e9),
            "speedup_fused_c128": v128 / vu, "speedup_c64": value / vu}
    del final64

    cpu, cpu_threads, acc_oracle = None, None, None
    if not args.skip_cpu:
        log(f"CPU oracle baseline: {args.cpu_samples} steps at {N}^3")
        rate, el, _, _ = cpu_baseline(psi0, args.cpu_samples)
        cpu = {"value": rate, "unit": "grid-pt-steps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_samples} Strang steps of the same 256^3 stirring workload "
                         f"(complex128 numpy.fft, 1 thread) after 1 warm-up step, {el:.1f} s"}
        workers = min(16, os.cpu_count() or 1)
        samples = max(1, args.oracle_steps - 1)
        log(f"threaded CPU oracle: {samples + 1} steps at {N}^3 on {workers} threads")
        rate, el, ref_psi, ref_energies = cpu_baseline(psi0, samples, workers=workers)
        cpu_threads = {"value": rate, "unit": "grid-pt-steps/s", "cores": workers, "kind": "port",
                       "sample": f"{samples} Strang steps of the same workload after 1 warm-up step, scipy.fft "
                                 f"on {workers} threads, {el:.1f} s"}
        log(f"engine accuracy against the oracle after {samples + 1} steps")
        acc_oracle = accuracy_vs_oracle(psi0, x, k, samples + 1, ref_psi, ref_energies)
        del ref_psi

    served = None
    if args.serve_requests > 0:
        log(f"serving stream: {args.serve_requests} requests")
        served = serving(args.serve_requests)

    swept = None
    if args.sweep_count > 0:
        log(f"batched sweep: {args.sweep_count} ground states at 96^3")
        swept = sweep(args.sweep_count)

    stirred = None
    if args.stir_count > 0:
        log(f"stirring sweep: {args.stir_count} stirrer speeds at 128^3")
        stirred = stirring(args.stir_count)

    step_bytes = PASSES_PER_STEP * 2 * 8
    line = {
        "metric": "3D SSFM grid-pt-steps/s at 256^3 c64 (% HBM roofline) vs c128; p50/p99 solve ms",
        "value": value,
        "unit": "grid-pt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el64 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64-reductions",
        "data": "synthetic (seeded Thomas-Fermi + noise), resident in HBM",
        "config": {"workload": "stirring run 256^3, L=40, beta=4000, dt=2.5e-4, nk every 500 steps "
                               "(dk shells and the 128-bin view up to k_max)",
                   "grid": N, "precision": "complex64 field, FP64 reductions/time", "step_plan": "separable"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant),
                     "algorithmic_bytes": bytes_per_launch, "avg_us": dom_ms * 1e3,
                     "timing": f"each launch's own dispatch start/end timestamps (HIP event pair on the "
                               f"solve's stream), median over 3 runs of {args.timed_steps} steps",
                     # launch end to launch end, gaps between launches included
                     "span_avg_us": spans[dominant],
                     # the kind's average in the committed kernel trace of this workload
                     # (traced_row names the variant it comes from)
                     "traced_avg_us": dom_traced_us, "traced_row": dom_traced_row,
                     "traced_frac": (bytes_per_launch / (dom_traced_us * 1e-6) / 1e9 / HBM_PEAK_GBS
                                     if dom_traced_us else None)},
        # at the bytes this plan moves (3 sweeps: 48 B per point), and at the
        # 64 B per point of a four-sweep step
        "step_roofline_frac": value * step_bytes / (HBM_PEAK_GBS * 1e9),
        "step_bytes_per_point": step_bytes,
        "step_roofline_frac_at_64B": value * PASSES_FOUR_SWEEP * 2 * 8 / (HBM_PEAK_GBS * 1e9),
        "c64_four_sweep": four64,
        "stream_copy_gbs": copy_gbs,
        "per_kernel": per_kernel,
        "mu_c64": diag64["mu"],
        "c128": result_c128,
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_threads,
        "accuracy_c64": acc_oracle,
        "accuracy_c64_vs_c128": acc,
        "p50_solve_ms": served["p50_ms"] if served else None,
        "p99_solve_ms": served["p99_ms"] if served else None,
        "serving": served,
        "sweep": swept,
        "stirring_sweep": stirred,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
