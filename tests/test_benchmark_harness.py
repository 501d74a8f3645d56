"""The propagator benchmark harness (gpe_simulator_amd/benchmark.py), the
counterpart of gpe/benchmark.py of the reference.

CPU: the generated config holds the reference's ten cases with its settings
(gpe/benchmark.py:18-55) and passes the schema-2 validator; the report files
have the reference's row keys, order and completeness flag (:112-179); the
command line rejects what the reference rejects (:188-199); and the
reference's own harness tests (tests/test_benchmark.py:36-102) restated with
fakes: the timer's order, repeats that restart from a fresh Gaussian with the
same per-step figures, and a failed setup that still releases the stage.
GPU: a full run at 32^3 and 48^3 writes a complete report whose norms obey
the physics of each case (imaginary time renormalises; real time without
absorber or cutoff conserves the norm; the cutoff only removes weight).
"""

import csv
import json
import math
import random

import pytest

from gpe_simulator_amd import benchmark
from gpe_simulator_amd.configuration import dump_yaml, load_yaml
from gpe_simulator_amd.yaml_parser import parse_yaml

NAMES = [v[0] for v in benchmark.VARIANTS]


def test_case_table_covers_every_combination_once():
    assert len(NAMES) == len(set(NAMES)) == 10
    flags = {(evo, st, dy, cut) for _, evo, st, dy, cut in benchmark.VARIANTS}
    assert len(flags) == 10
    # imaginary time has neither a dynamic potential nor a cutoff case
    assert all(not (evo == "imaginary" and (dy or cut)) for evo, st, dy, cut in flags)
    # every real-time combination of the three switches is present
    assert {(st, dy, cut) for evo, st, dy, cut in flags if evo == "real"} == {
        (a, b, c) for a in (False, True) for b in (False, True) for c in (False, True)}
    for name, evo, st, dy, cut in benchmark.VARIANTS:
        assert name.startswith(evo)
        assert ("static" in name) == st and ("dynamic" in name) == dy and ("cutoff" in name) == cut
        assert ("_none" in name) == (not st and not dy)


@pytest.mark.parametrize("n", [32, 64, 256])
def test_config_round_trips_the_schema(n):
    config = benchmark.benchmark_config(n, steps=7, dt=0.003)
    loaded = load_yaml(dump_yaml(config))
    assert loaded == config
    assert loaded["simulation"]["N_grids"] == n and loaded["simulation"]["simulation_range"] == 70
    for item, (name, evo, st, dy, cut) in zip(loaded["stages"], benchmark.VARIANTS):
        assert item["name"] == name
        assert item["evolution"] == {"type": evo, "steps": 7, "dt": 0.003}
        assert item["initial_state"]["type"] == "gaussian" and item["initial_state"]["sigma"] == 10
        assert item["cloud"] == {"a": 100, "N_atom": 150000}
        assert not any(item[g]["enabled"] for g in ("box", "absorber", "noise"))
        assert item["potential"]["static"] == {
            "enabled": st, "expression": "0.002*(x*x + y*y + z*z)" if st else "0"}
        assert item["potential"]["dynamic"] == {
            "enabled": dy, "expression": "0.1*cos(0.5*t)*y" if dy else "0"}
        assert item["k_cut"]["enabled"] == cut
        # 0.7 of the axis Nyquist wavenumber pi/dx = pi*N/L
        assert item["k_cut"]["value"] == (pytest.approx(0.7 * math.pi / (70 / n), rel=1e-15) if cut else 0)
        assert set(item["monitor"].values()) == {0}
        assert item["save"]["scalar_interval"] == item["save"]["psi_interval"] == 0
        assert not item["save"]["save_initial"] and not item["save"]["save_final"]


def test_parsed_stages_carry_the_cases(tmp_path):
    path = tmp_path / "config_N32.yaml"
    path.write_text(dump_yaml(benchmark.benchmark_config(32, 5, 0.002)))
    stages = parse_yaml(path, precision="complex64")
    assert [s.name for s in stages] == NAMES
    for s, (name, evo, st, dy, cut) in zip(stages, benchmark.VARIANTS):
        assert s.type == evo and s.steps == 5 and s.dt == 0.002 and s.precision == "complex64"
        assert s.potential_params["static"]["activate"] == st
        assert s.potential_params["dynamic"]["activate"] == dy
        assert s.k_cut_params["activate"] == cut
        assert s.solve is None          # nothing is allocated before prepare


def fake_row(grid, name, ms):
    return dict(grid=grid, variant=name, steps=5, setup_ms=1.0, first_step_ms=0.5, wall_ms_per_step=ms,
                wall_min_ms_per_step=ms * 0.9, wall_max_ms_per_step=ms * 1.1, cuda_ms_per_step=ms * 0.95,
                steps_per_second=1000 / ms, final_norm=1.0,
                samples=[dict(repeat=1, wall_seconds=ms * 5e-3, cuda_ms=ms * 4.75, final_norm=1.0)])


def metadata(grids, expected=None):
    _, args = benchmark.parse_args(["--grids", *map(str, grids), "--steps", "5"])
    return dict(started_utc="2026-01-01T00:00:00+00:00", options=vars(args), gpu="test device",
                gpu_memory_gib=288.0, python="3", platform="linux", engine="gpe_hip", hip_runtime=1,
                hip_driver=2, precision="complex128",
                expected_cases=expected if expected is not None else len(grids) * 10,
                execution_order=[], source_sha256={})


def test_report_files(tmp_path):
    rng = random.Random(3)
    rows = [fake_row(n, name, 0.1 * (i + 1)) for n in (64, 32) for i, name in enumerate(NAMES)]
    rng.shuffle(rows)
    meta = metadata([64, 32])
    benchmark.write_report(tmp_path, meta, rows[:7])
    payload = json.loads((tmp_path / "results.json").read_text())
    assert payload["completed_cases"] == 7 and payload["complete"] is False
    benchmark.write_report(tmp_path, meta, rows)
    payload = json.loads((tmp_path / "results.json").read_text())
    assert payload["complete"] is True and payload["completed_cases"] == 20
    order = [(r["grid"], r["variant"]) for r in payload["results"]]
    assert order == [(n, name) for n in (32, 64) for name in NAMES]
    assert all(len(r["samples"]) == 1 for r in payload["results"])
    with (tmp_path / "summary.csv").open() as f:
        reader = csv.DictReader(f)
        assert tuple(reader.fieldnames) == benchmark.ROW_KEYS
        table = list(reader)
    assert [(int(r["grid"]), r["variant"]) for r in table] == order
    report = (tmp_path / "report.md").read_text()
    assert "Completed cases: 20/20." in report
    for n in (32, 64):
        for name in NAMES:
            assert report.count(f"| {n}³ | {name} |") == 2     # stepping and setup tables


def test_empty_report_is_written(tmp_path):
    benchmark.write_report(tmp_path, metadata([64]), [])
    payload = json.loads((tmp_path / "results.json").read_text())
    assert payload["completed_cases"] == 0 and payload["results"] == []
    assert not (tmp_path / "summary.csv").exists()
    assert "Completed cases: 0/10." in (tmp_path / "report.md").read_text()


@pytest.mark.parametrize("argv", [["--grids", "1"], ["--grids", "64", "64"], ["--dt", "0"],
                                  ["--dt", "-1"], ["--dt", "nan"], ["--dt", "inf"], ["--steps", "0"],
                                  ["--repeats", "-2"], ["--precision", "complex32"]])
def test_command_line_rejections(argv):
    with pytest.raises(SystemExit):
        benchmark.parse_args(argv)


def test_command_line_defaults():
    _, args = benchmark.parse_args([])
    assert args.grids == [64, 128, 256] and args.steps == 50 and args.warmup == 5
    assert args.repeats == 3 and args.dt == 0.002 and args.seed == 42
    assert args.output == "benchmark_results" and args.precision == "complex128"


@pytest.mark.gpu
@pytest.mark.parametrize("precision,norm_tol", [("complex128", 1e-11), ("complex64", 2e-6)])
def test_full_run(tmp_path, precision, norm_tol, require_gpu):
    folder = benchmark.main(["--grids", "32", "48", "--steps", "6", "--warmup", "2", "--repeats", "2",
                             "--seed", "7", "--output", str(tmp_path), "--precision", precision])
    payload = json.loads((folder / "results.json").read_text())
    assert payload["complete"] and payload["completed_cases"] == 20
    expected = []
    rng = random.Random(7)
    for n in (32, 48):
        names = list(NAMES)
        rng.shuffle(names)
        expected += [dict(grid=n, variant=name) for name in names]
    assert payload["metadata"]["execution_order"] == expected
    assert payload["metadata"]["precision"] == precision
    for row in payload["results"]:
        assert set(row) == set(benchmark.ROW_KEYS) | {"samples"}
        assert len(row["samples"]) == 2
        for sample in row["samples"]:
            # the event pair sits inside the wall-clock interval
            assert 0 < sample["cuda_ms"] <= sample["wall_seconds"] * 1e3 * 1.01 + 0.05
        norm, name = row["final_norm"], row["variant"]
        if name.startswith("imaginary") or "cutoff" not in name:
            assert abs(norm - 1) < norm_tol, (row["grid"], name, norm)
        else:
            assert 0.5 < norm <= 1 + norm_tol, (row["grid"], name, norm)
    for n in (32, 48):
        assert (folder / f"config_N{n}.yaml").exists()
    assert "pass tile variants" in (folder / "engine_config.txt").read_text()


class FakeSolve:
    """Records the calls the harness makes on a solve."""

    def __init__(self, calls):
        self.calls = calls

    def sync(self):
        self.calls.append("drain stream")

    def mark(self, which):
        self.calls.append("start mark" if which == 0 else "stop mark")

    def marked_ms(self):
        self.calls.append("wait for stop mark")
        return 12.5


def test_timer_brackets_the_batch(monkeypatch):
    """gpe/benchmark.py's measure order (tests/test_benchmark.py:36-54 of the
    reference): drain, wall clock, start, work, stop, wait, wall clock."""
    calls = []
    times = iter((1.0, 1.02))

    def clock():
        calls.append("wall clock")
        return next(times)

    monkeypatch.setattr(benchmark, "perf_counter", clock)
    wall, device = benchmark.timed(FakeSolve(calls), lambda: calls.append("work"))
    assert wall == pytest.approx(0.02) and device == 12.5
    assert calls == ["drain stream", "wall clock", "start mark", "work", "stop mark", "wait for stop mark",
                     "wall clock"]


def test_repeats_reset_state_and_report_per_step_times(monkeypatch):
    """Each timed repeat starts from a freshly built Gaussian; the row holds
    per-step medians and spreads (tests/test_benchmark.py:56-92 of the
    reference, same figures)."""
    import sys
    from types import SimpleNamespace
    from unittest.mock import Mock
    calls = []
    stage = SimpleNamespace(name="imaginary_none", N_grids=2, initial_state={"type": "gaussian", "sigma": 10},
                            psi=None, solve=None, steps=0)
    stage.prepare = lambda: setattr(stage, "solve", FakeSolve([]))
    stage.release = Mock()

    def launch(item):
        calls.append(item.steps)
        if item.steps != 2:                       # the first step and every timed repeat
            assert item.psi == "fresh gaussian", item.psi
        item.psi = "advanced"
        return item.psi

    def gaussian(solve, sigma):
        assert solve is stage.solve and sigma == 10
        return "fresh gaussian"

    stage.psi = "fresh gaussian"                  # prepare builds the initial state
    timings = iter(((0.01, 9), (0.008, 7), (0.012, 11), (0.016, 15)))
    monkeypatch.setattr(benchmark, "timed", lambda solve, fn: (fn(), next(timings))[1])
    clock = iter((0.0, 0.02))
    monkeypatch.setattr(benchmark, "perf_counter", lambda: next(clock))
    monkeypatch.setitem(sys.modules, "gpe_simulator_amd.launch", SimpleNamespace(launch_stage=launch))
    monkeypatch.setitem(sys.modules, "gpe_simulator_amd.initial_states", SimpleNamespace(gaussian=gaussian))
    monkeypatch.setitem(sys.modules, "gpe_simulator_amd.helper", SimpleNamespace(get_N=lambda psi: 1.0))
    row = benchmark.benchmark_stage(stage, steps=4, warmup=2, repeats=3)
    assert calls == [1, 2, 4, 4, 4]
    stage.release.assert_called_once()
    assert row["setup_ms"] == pytest.approx(20)
    assert row["first_step_ms"] == pytest.approx(10)
    assert row["wall_ms_per_step"] == pytest.approx(3)
    assert row["wall_min_ms_per_step"] == pytest.approx(2)
    assert row["wall_max_ms_per_step"] == pytest.approx(4)
    assert row["cuda_ms_per_step"] == pytest.approx(2.75)
    assert row["steps_per_second"] == pytest.approx(1000 / 3)
    assert len(row["samples"]) == 3 and all(s["final_norm"] == 1.0 for s in row["samples"])


def test_failed_setup_releases_stage():
    """tests/test_benchmark.py:94-102 of the reference."""
    from types import SimpleNamespace
    from unittest.mock import Mock
    stage = SimpleNamespace(prepare=Mock(side_effect=RuntimeError("allocation failed")), release=Mock())
    with pytest.raises(RuntimeError, match="allocation failed"):
        benchmark.benchmark_stage(stage, 4, 2, 3)
    stage.release.assert_called_once()


@pytest.mark.parametrize("norm,message", [(float("nan"), "non-finite wavefunction after repeat 2"),
                                          (float("inf"), "non-finite wavefunction after repeat 2"),
                                          (0.0, "invalid norm 0.0")])
def test_invalid_state_after_a_repeat_is_an_error(monkeypatch, norm, message):
    import sys
    from types import SimpleNamespace
    stage = SimpleNamespace(name="real_none", psi="psi")
    monkeypatch.setitem(sys.modules, "gpe_simulator_amd.helper", SimpleNamespace(get_N=lambda psi: norm))
    with pytest.raises(RuntimeError, match=message):
        benchmark.check_state(stage, 2)
