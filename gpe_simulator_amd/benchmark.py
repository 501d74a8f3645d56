"""Propagator benchmark: python -m gpe_simulator_amd.benchmark --help.

The harness of gpe/benchmark.py:1-261 of the reference, run on libgpe_hip.
It takes the same command line and writes the same report folder, so a report
from either engine reads the same way:

* the same ten cases per grid (VARIANTS, gpe/benchmark.py:18-30), generated as
  a schema-2 config (benchmark_config, :33-55) and parsed back through
  parse_yaml, like the reference;
* the same measurement (:58-109): setup (stage.prepare), one first step,
  warm-up steps, then `repeats` timed batches of `steps` steps through
  launch_stage, each from a freshly built Gaussian, with a host wall clock
  and a device event pair around every batch;
* the same files (:112-179): results.json (every repeat), summary.csv,
  report.md and config_N*.yaml, rewritten after every completed case, plus
  engine_config.txt in place of cupy_config.txt;
* case order shuffled per grid with random.Random(seed), as in :240-246.

Row keys match the reference's, so tools that read its summary.csv read this
one.  `cuda_ms_per_step` keeps its name for that reason; here it holds the HIP
event interval (gpe_mark / gpe_mark_elapsed on the stage's stream).

Beyond the reference, `--precision complex64` runs the same cases on
complex64 fields (FP64 reductions and constants).
"""

from __future__ import annotations

import argparse
import csv
import gc
import hashlib
import json
import math
import platform
import random
from datetime import datetime, timezone
from pathlib import Path
from statistics import median
from time import perf_counter

from .configuration import PRECISIONS, dump_yaml, new_config, new_stage

# name, evolution, static potential, dynamic potential, momentum cutoff
# (the case table of gpe/benchmark.py:18-30)
VARIANTS = [
    ("imaginary_none", "imaginary", False, False, False),
    ("imaginary_static", "imaginary", True, False, False),
    ("real_none", "real", False, False, False),
    ("real_static", "real", True, False, False),
    ("real_dynamic", "real", False, True, False),
    ("real_static_dynamic", "real", True, True, False),
    ("real_none_cutoff", "real", False, False, True),
    ("real_static_cutoff", "real", True, False, True),
    ("real_dynamic_cutoff", "real", False, True, True),
    ("real_static_dynamic_cutoff", "real", True, True, True),
]
VARIANT_ORDER = {name: index for index, (name, *_rest) in enumerate(VARIANTS)}

BOX_LENGTH = 70
SIGMA = 10
STATIC_EXPRESSION = "0.002*(x*x + y*y + z*z)"
DYNAMIC_EXPRESSION = "0.1*cos(0.5*t)*y"
CUTOFF_FRACTION = 0.7          # of the axis Nyquist wavenumber pi*N/L

ROW_KEYS = ("grid", "variant", "steps", "setup_ms", "first_step_ms", "wall_ms_per_step",
            "wall_min_ms_per_step", "wall_max_ms_per_step", "cuda_ms_per_step", "steps_per_second",
            "final_norm")


def benchmark_config(n, steps, dt):
    """Schema-2 config with one stage per case (gpe/benchmark.py:33-55)."""
    config = new_config()
    config["simulation"].update(N_grids=n, simulation_range=BOX_LENGTH)
    config["stages"] = []
    for name, evolution, static, dynamic, cutoff in VARIANTS:
        item = new_stage(imaginary=evolution == "imaginary")
        item["name"] = name
        item["evolution"].update(type=evolution, steps=steps, dt=dt)
        item["initial_state"].update(type="gaussian", sigma=SIGMA)
        item["cloud"].update(a=100, N_atom=150000)
        for group in ("box", "absorber", "noise"):
            item[group]["enabled"] = False
        item["potential"]["static"].update(enabled=static,
                                           expression=STATIC_EXPRESSION if static else "0")
        item["potential"]["dynamic"].update(enabled=dynamic,
                                            expression=DYNAMIC_EXPRESSION if dynamic else "0")
        item["k_cut"].update(enabled=cutoff,
                             value=CUTOFF_FRACTION * math.pi * n / BOX_LENGTH if cutoff else 0)
        item["monitor"] = dict.fromkeys(item["monitor"], 0)
        item["save"].update(directory="", scalar_interval=0, nk_interval=0, psi_interval=0,
                            save_initial=False, save_final=False)
        config["stages"].append(item)
    return config


def timed(solve, function):
    """(wall seconds, device milliseconds) of `function`'s work on `solve`.

    The stream is drained first; the wall clock stops once the stop mark has
    completed, so both cover the same batch (gpe/benchmark.py:58-69)."""
    solve.sync()
    before = perf_counter()
    solve.mark(0)
    function()
    solve.mark(1)
    device_ms = solve.marked_ms()
    return perf_counter() - before, device_ms


def check_state(stage, repeat):
    """Validity of the state after a timed repeat (gpe/benchmark.py:89-94).

    The FP64 norm is a sum of |psi|^2 over every point, so a NaN or an infinity
    anywhere in psi makes it non-finite: one device reduction stands in for the
    reference's element-wise isfinite over the whole field."""
    from .helper import get_N
    norm = float(get_N(stage.psi))
    if not math.isfinite(norm):
        raise RuntimeError(f"{stage.name}: non-finite wavefunction after repeat {repeat}")
    if norm <= 0:
        raise RuntimeError(f"{stage.name}: invalid norm {norm}")
    return norm


def benchmark_stage(stage, steps, warmup, repeats):
    """One case: setup, first step, warm-up, then `repeats` timed batches."""
    from .initial_states import gaussian
    from .launch import launch_stage

    try:
        before = perf_counter()
        stage.prepare()
        stage.solve.sync()
        setup_seconds = perf_counter() - before
        stage.steps = 1
        first_seconds, _ = timed(stage.solve, lambda: launch_stage(stage))
        stage.steps = warmup
        launch_stage(stage)
        stage.solve.sync()

        samples = []
        stage.steps = steps
        for repeat in range(1, repeats + 1):
            stage.psi = gaussian(stage.solve, stage.initial_state["sigma"])
            wall_seconds, device_ms = timed(stage.solve, lambda: launch_stage(stage))
            norm = check_state(stage, repeat)
            samples.append(dict(repeat=repeat, wall_seconds=wall_seconds, cuda_ms=device_ms,
                                final_norm=norm))

        wall = [sample["wall_seconds"] * 1000 / steps for sample in samples]
        return dict(grid=stage.N_grids, variant=stage.name, steps=steps,
                    setup_ms=setup_seconds * 1000, first_step_ms=first_seconds * 1000,
                    wall_ms_per_step=median(wall), wall_min_ms_per_step=min(wall),
                    wall_max_ms_per_step=max(wall),
                    cuda_ms_per_step=median(sample["cuda_ms"] / steps for sample in samples),
                    steps_per_second=1000 / median(wall), final_norm=samples[-1]["final_norm"],
                    samples=samples)
    finally:
        stage.release()


def sorted_results(results):
    return sorted(results, key=lambda row: (row["grid"], VARIANT_ORDER[row["variant"]]))


def report_lines(metadata, results):
    args = metadata["options"]
    field = ("complex128" if metadata["precision"] == "complex128" else
             "complex64 wavefunctions with FP64 reductions and constants; complex64")
    lines = [
        "# GPE propagator benchmark", "",
        f"Completed cases: {len(results)}/{metadata['expected_cases']}. Started: {metadata['started_utc']}.",
        f"GPU: {metadata['gpu']} ({metadata['gpu_memory_gib']:.2f} GiB). "
        f"Python {metadata['python']}; engine {metadata['engine']}; "
        f"HIP runtime {metadata['hip_runtime']}; driver {metadata['hip_driver']}.",
        "", f"Each repeat runs {args['steps']} steps; {args['repeats']} repeats per case. "
        f"One first step plus {args['warmup']} warm-up steps precede measurement. "
        f"dt={args['dt']}; {field} wavefunctions and float64 coordinates.",
        "", "The primary metric is synchronized wall time per step through launch_stage, "
        "including its output-schedule checks; with no output due, launch_stage hands the "
        "whole batch to the engine in one call. Initial-state resets, validation, setup, "
        "warm-up and report writing are outside the repeated timings. All snapshots and "
        "diagnostics are disabled. The Gaussian is rebuilt on the device before each repeat. "
        "A HIP event pair on the stage's stream covers the same batch (column `cuda_ms_per_step` "
        "in summary.csv); its interval can include device idle time while the host submits.",
        "", f"All cases start each repeat from the same normalized Gaussian (sigma={SIGMA}, "
        f"box length={BOX_LENGTH}), with a=100 and reference N_atom=150000. Box, absorber and "
        f"noise are off. Static expression: `{STATIC_EXPRESSION}`. Dynamic expression: "
        f"`{DYNAMIC_EXPRESSION}`. Cutoff: {CUTOFF_FRACTION} times the axis Nyquist wavenumber. "
        "Imaginary time includes per-step normalization; real time does not. "
        "These are different workloads, not interchangeable numerical methods.",
        "", "## Warmed-up stepping", "",
        "Values are per-step times averaged within each batch, then summarized across repeats. "
        "Min/max show the observed spread, not a confidence interval.", "",
        "| Grid | Variant | Wall median ms/step | Wall min–max | Device median ms/step | Steps/s |",
        "| --- | --- | ---: | ---: | ---: | ---: |",
    ]
    for row in results:
        lines.append(f"| {row['grid']}³ | {row['variant']} | {row['wall_ms_per_step']:.4f} | "
                     f"{row['wall_min_ms_per_step']:.4f}–{row['wall_max_ms_per_step']:.4f} | "
                     f"{row['cuda_ms_per_step']:.4f} | {row['steps_per_second']:.1f} |")
    lines += [
        "", "## Observed setup and first step", "",
        "Each value is a single observation. Setup times stage.prepare (device allocation, "
        "coordinate and potential tables, the Gaussian); YAML parsing and axis creation are "
        "excluded. Kernels are compiled ahead of time, so the first step includes no "
        f"compilation. Case order is shuffled with seed {args['seed']}. Every stage's device "
        "memory is released when its case ends.", "",
        "| Grid | Variant | Setup ms | First step ms |", "| --- | --- | ---: | ---: |",
    ]
    for row in results:
        lines.append(f"| {row['grid']}³ | {row['variant']} | {row['setup_ms']:.3f} | "
                     f"{row['first_step_ms']:.3f} |")
    lines += [
        "", "Exact generated configurations are in config_N*.yaml. summary.csv holds the tables; "
        "results.json includes individual repeat timings, norms, execution order and source "
        "hashes. engine_config.txt records the engine, device and pass tile variants.", "",
    ]
    return lines


def write_report(folder, metadata, results):
    """Rewrite results.json, summary.csv and report.md (gpe/benchmark.py:112-179)."""
    folder = Path(folder)
    results = sorted_results(results)
    payload = dict(metadata=metadata, completed_cases=len(results),
                   complete=len(results) == metadata["expected_cases"], results=results)
    (folder / "results.json").write_text(json.dumps(payload, indent=2) + "\n")
    if results:
        with (folder / "summary.csv").open("w", newline="") as file:
            writer = csv.DictWriter(file, fieldnames=[key for key in results[0] if key != "samples"])
            writer.writeheader()
            writer.writerows({key: value for key, value in row.items() if key != "samples"}
                             for row in results)
    (folder / "report.md").write_text("\n".join(report_lines(metadata, results)))


def positive_int(value):
    number = int(value)
    if number < 1:
        raise argparse.ArgumentTypeError("must be a positive integer")
    return number


def make_parser():
    parser = argparse.ArgumentParser(prog="python -m gpe_simulator_amd.benchmark",
                                     description="Compare the propagators on libgpe_hip.")
    parser.add_argument("--grids", nargs="+", type=positive_int, default=[64, 128, 256])
    parser.add_argument("--steps", type=positive_int, default=50, help="steps per timed repeat (default: 50)")
    parser.add_argument("--warmup", type=positive_int, default=5, help="warm-up steps after first step (default: 5)")
    parser.add_argument("--repeats", type=positive_int, default=3)
    parser.add_argument("--dt", type=float, default=0.002)
    parser.add_argument("--seed", type=int, default=42, help="shuffle case order with this seed")
    parser.add_argument("--output", default="benchmark_results", help="parent folder for a new timestamped report")
    parser.add_argument("--precision", choices=PRECISIONS, default="complex128",
                        help="field precision (default: complex128, the reference's)")
    return parser


def parse_args(argv=None):
    parser = make_parser()
    args = parser.parse_args(argv)
    if min(args.grids) < 2 or len(set(args.grids)) != len(args.grids):
        parser.error("grids must be distinct integers of at least 2")
    if not math.isfinite(args.dt) or args.dt <= 0:
        parser.error("dt must be finite and positive")
    return parser, args


def source_hashes():
    root = Path(__file__).resolve().parents[1]
    package = root / "gpe_simulator_amd"
    paths = [*sorted(package.glob("*.py")), *sorted((package / "csrc").glob("*")),
             *sorted((root / "include").glob("*.h"))]
    library = package / "libgpe_hip.so"
    if library.exists():
        paths.append(library)
    return {str(path.relative_to(root)): hashlib.sha256(path.read_bytes()).hexdigest()
            for path in paths if path.is_file()}


def engine_config_text(ctx, device, grids, precision):
    from . import _lib
    code = _lib.C64 if precision == "complex64" else _lib.C128
    lines = [f"engine: {_lib.lib().gpe_version().decode()}",
             f"library: {_lib.LIBRARY}",
             f"device: {device['name']}",
             f"HIP runtime: {device['runtime_version']}; driver: {device['driver_version']}",
             f"device memory (free, total bytes): {ctx.mem_info()}",
             f"precision: {precision}", "", "pass tile variants (TP, B, waves) per grid side:"]
    for n in grids:
        lines.append(f"  {n}: {_lib.pass_variants(n, code)}")
    return "\n".join(lines) + "\n"


def main(argv=None):
    parser, args = parse_args(argv)
    from . import _lib
    try:
        ctx = _lib.default_context()
        device = ctx.device_info()
    except _lib.GpeError as exc:
        parser.error(f"Benchmarking requires libgpe_hip.so and a HIP device ({exc}).")
    from .yaml_parser import parse_yaml

    started = datetime.now(timezone.utc)
    folder = Path(args.output).expanduser().resolve() / started.strftime("%Y%m%dT%H%M%S%fZ")
    folder.mkdir(parents=True)
    (folder / "engine_config.txt").write_text(engine_config_text(ctx, device, args.grids, args.precision))
    metadata = dict(started_utc=started.isoformat(), options=vars(args), gpu=device["name"],
                    gpu_memory_gib=ctx.mem_info()[1] / 2 ** 30, python=platform.python_version(),
                    platform=platform.platform(), engine=_lib.lib().gpe_version().decode(),
                    hip_runtime=device["runtime_version"], hip_driver=device["driver_version"],
                    precision=args.precision, expected_cases=len(args.grids) * len(VARIANTS),
                    execution_order=[], source_sha256=source_hashes())
    results = []
    write_report(folder, metadata, results)
    print(f"GPU: {device['name']}\nReports: {folder}", flush=True)
    rng = random.Random(args.seed)
    for n in args.grids:
        path = folder / f"config_N{n}.yaml"
        path.write_text(dump_yaml(benchmark_config(n, args.steps, args.dt)))
        stages = parse_yaml(path, precision=args.precision)
        rng.shuffle(stages)
        for item in stages:
            print(f"Benchmarking {n}³ {item.name} ...", flush=True)
            metadata["execution_order"].append(dict(grid=n, variant=item.name))
            result = benchmark_stage(item, args.steps, args.warmup, args.repeats)
            results.append(result)
            write_report(folder, metadata, results)
            print(f"  {result['wall_ms_per_step']:.4f} ms/step; "
                  f"{result['steps_per_second']:.1f} steps/s", flush=True)
        del stages
        gc.collect()
    print(f"Finished {len(results)} cases. Report: {folder / 'report.md'}", flush=True)
    return folder


if __name__ == "__main__":
    main()
