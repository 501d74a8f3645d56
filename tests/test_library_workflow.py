"""C-ABI library and workflow plumbing (CPU, no device calls).

The library test checks that libgpe_hip.so loads and exports every function
include/gpe_hip.h declares.  The workflow tests mirror the reference's CPU
tests of the stage driver (tests/test_workflow.py of the reference): parsing
defers device work and shares axes, the stage clock is t_i = i*dt and resets
per stage, output folders follow the same rules, and --check never loads the
engine.
"""

import json
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "gpe_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gpe_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def library():
    from gpe_simulator_amd import _lib
    if not _lib.LIBRARY.exists():
        from gpe_simulator_amd import build
        build.build()
    return _lib


def test_library_exports_every_declared_function(library):
    names = header_functions()
    assert len(names) >= 20
    handle = library.lib()
    for name in names:
        assert hasattr(handle, name), name
    assert sorted(library.EXPORTS) == names
    assert handle.gpe_version().startswith(b"gpe_hip")


def test_library_reports_errors_without_device(library):
    handle = library.lib()
    assert handle.gpe_last_error() is not None
    # a null context is rejected before any device work
    assert handle.gpe_mem_info(None, None, None) == library.GPE_ERR_INVALID
    assert b"null" in handle.gpe_last_error()


def test_library_is_gfx950_code():
    lib = (ROOT / "gpe_simulator_amd" / "libgpe_hip.so")
    if not lib.exists():
        pytest.skip("library not built")
    assert b"gfx950" in lib.read_bytes()


def test_default_plan_rule():
    """The host's restatement of gpe_solve_create's plan rule (the GPU tests
    check it against the library): with a cutoff, the plane plan at every
    side with a plane pass, else four-sweep; without, the plane plan at the
    plane sides that are multiples of 64, else separable."""
    from gpe_simulator_amd import _lib
    for prec, sides in _lib.PLANE_SIDES.items():
        for n in (16, 32, 48, 64, 96, 128, 192, 256, 512):
            plane = n in sides
            assert _lib.default_plan(prec, _lib.REAL, masked=True, n=n) == \
                (_lib.PLAN_PLANE if plane else _lib.PLAN_FOUR_SWEEP), (prec, n)
            for kind in (_lib.REAL, _lib.IMAG):
                assert _lib.default_plan(prec, kind, n=n) == \
                    (_lib.PLAN_PLANE if plane and n % 64 == 0 else _lib.PLAN_SEPARABLE), (prec, n)
    assert _lib.default_plan(_lib.C128, _lib.REAL, masked=True, n=128) == _lib.PLAN_FOUR_SWEEP
    assert _lib.default_plan(_lib.C64, _lib.REAL, masked=True, n=96) == _lib.PLAN_PLANE
    assert _lib.default_plan(_lib.C64, _lib.REAL, n=96) == _lib.PLAN_SEPARABLE


# ------------------------------------------------------------ workflow --
def write_config(root, stages=1, **first_updates):
    from gpe_simulator_amd.configuration import dump_yaml, new_config
    config = new_config()
    config["simulation"].update(N_grids=4, simulation_range=8, data_root=str(root / "data"),
                                monitor_root=str(root / "monitor"))
    first = config["stages"][0]
    first["evolution"].update(type="real", steps=6, dt=0.125)
    first["box"]["enabled"] = False
    first["monitor"] = dict.fromkeys(first["monitor"], 0)
    first["save"].update(scalar_interval=0, save_final=False)
    for key, value in first_updates.items():
        first[key].update(value)
    for _ in range(stages - 1):
        second = json.loads(json.dumps(first))
        second["initial_state"]["type"] = "continue"
        config["stages"].append(second)
    path = root / "experiment.yaml"
    path.write_text(dump_yaml(config))
    return path, config


def test_parser_shares_axes_and_defers_device_work(tmp_path):
    from gpe_simulator_amd.yaml_parser import parse_yaml
    path, config = write_config(tmp_path, stages=2, potential={"static": {"enabled": True,
                                                                          "expression": "1/2*y**2"}})
    stages = parse_yaml(path)
    assert stages[0].x_array is stages[1].x_array and stages[0].k_array is stages[1].k_array
    for item in stages:
        assert item.psi is None and item.solve is None and item.propagator is None
        assert not item.data_dir.exists()
    assert stages[0].potential_params["static"]["expression"] == "((1.0 / 2.0) * pow(y, 2.0))"
    assert stages[0].data_dir.parent.name == stages[1].monitor_dir.parent.name
    assert stages[0].precision == "complex128"
    assert parse_yaml(path, precision="complex64")[0].precision == "complex64"


def test_duplicate_and_escaping_output_folders(tmp_path):
    from gpe_simulator_amd.configuration import dump_yaml
    from gpe_simulator_amd.yaml_parser import parse_yaml
    path, config = write_config(tmp_path)
    first = config["stages"][0]
    first["save"]["directory"] = "same_folder"
    second = json.loads(json.dumps(first))
    second["save"]["directory"] = "nested/../same_folder"
    config["stages"].append(second)
    path.write_text(dump_yaml(config))
    with pytest.raises(ValueError, match="reuses another stage's output directory"):
        parse_yaml(path)
    config["stages"] = [first]
    first["save"]["directory"] = "../shared"
    path.write_text(dump_yaml(config))
    with pytest.raises(ValueError, match="stay inside the run folder"):
        parse_yaml(path)


def test_stage_folders_names_and_load_paths(tmp_path):
    """Default folders are stage_NN plus the name reduced to [A-Za-z0-9_-]
    runs, under one run stamp in both roots; an absolute save.directory holds
    the data itself while the monitor files keep the default folder; a
    relative load path is taken from the YAML file's folder; noise and
    switch groups arrive with the stage's keyword names."""
    from gpe_simulator_amd.configuration import dump_yaml
    from gpe_simulator_amd.yaml_parser import parse_yaml
    path, config = write_config(tmp_path, stages=3)
    config["stages"][0]["name"] = "  Ground state: g=100!  "
    config["stages"][1]["name"] = "***"
    config["stages"][2]["save"]["directory"] = str(tmp_path / "elsewhere")
    config["stages"][0]["initial_state"].update(type="load", path="states/psi.npy")
    path.write_text(dump_yaml(config))
    first, second, third = parse_yaml(path)
    stamp = first.data_dir.parent.name
    assert first.name == "Ground state: g=100!" and second.name == "***"
    assert first.data_dir == tmp_path / "data" / stamp / "stage_01_Ground_state_g_100"
    assert first.monitor_dir == tmp_path / "monitor" / stamp / "stage_01_Ground_state_g_100"
    assert second.data_dir.name == "stage_02" and second.data_dir.parent.name == stamp
    assert third.data_dir == tmp_path / "elsewhere" / stamp
    assert third.monitor_dir == tmp_path / "monitor" / stamp / "stage_03_Preparation"   # the default name
    assert first.initial_state["path"] == str(tmp_path / "states" / "psi.npy")
    assert "fraction" in first.noise_params and "fraction_of_trap_depth" not in first.noise_params
    assert first.box_params["activate"] is False and "enabled" not in first.box_params
    assert first.k_cut_params == {"activate": config["stages"][0]["k_cut"]["enabled"],
                                  "k_cut": config["stages"][0]["k_cut"]["value"]}


def test_stage_clock_and_per_step_propagator(tmp_path):
    """A propagator without `advance` is called once per step with t = i*dt,
    and the clock restarts in each stage (tests/test_workflow.py:235-272)."""
    from gpe_simulator_amd.launch import launch_stage
    from gpe_simulator_amd.yaml_parser import parse_yaml
    path, _ = write_config(tmp_path, stages=2)
    stages = parse_yaml(path)
    stages[1].potential_params["dynamic"]["activate"] = True
    times = []

    def step(current, t=None):
        times.append(t)
        return current.psi * 0.5

    for item in stages:
        item.steps = 2
        item.psi = np.full((4, 4, 4), 0.1 + 0.2j)
        item.propagator = step
        launch_stage(item)
    assert times == [None, None, 0, 0.125]


def test_launch_stage_batches_steps_between_due_outputs(tmp_path):
    from gpe_simulator_amd import launch
    from gpe_simulator_amd.yaml_parser import parse_yaml
    path, _ = write_config(tmp_path)
    item = parse_yaml(path)[0]
    item.steps = 23
    item.save_params.update(scalar_interval=5, nk_interval=0, psi_interval=0, save_final=False)
    item.monitor_params["show_nk_interval"] = 7
    calls, saved = [], []

    def step(current):
        raise AssertionError("the batched path should be used")

    def advance(current, first, count):
        calls.append((first, count))
        return current.psi

    step.advance = advance
    item.propagator = step
    item.psi = np.zeros((4, 4, 4), complex)
    original = launch.save_outputs
    launch.save_outputs = lambda stage, i, **kw: saved.append(i)
    try:
        launch.launch_stage(item)
    finally:
        launch.save_outputs = original
    assert calls == [(0, 5), (5, 2), (7, 3), (10, 4), (14, 1), (15, 5), (20, 1), (21, 2)]
    assert saved == [0, 5, 7, 10, 14, 15, 20, 21, 23]


def test_cli_check_does_not_load_the_engine(tmp_path):
    path, _ = write_config(tmp_path)
    code = ("import sys, runpy; sys.argv=['main', sys.argv[1], '--check']; "
            "runpy.run_module('gpe_simulator_amd.main', run_name='__main__'); "
            "print('ENGINE' if 'gpe_simulator_amd._lib' in sys.modules else 'NOENGINE')")
    out = subprocess.run([sys.executable, "-c", code, str(path)], capture_output=True, text=True, cwd=ROOT,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Valid configuration: 1 stages" in out.stdout
    assert "NOENGINE" in out.stdout


@pytest.mark.parametrize("i,interval,expected", [(0, 1, False), (1, 0, False), (2, None, False),
                                                 (1, 2, False), (2, 2, True), (6, 3, True)])
def test_output_schedule(i, interval, expected):
    """The output schedule predicate (tests/test_diagnostics.py:186-189 of the
    reference): never at step 0, never for a 0 or missing interval."""
    from gpe_simulator_amd.helper import is_due
    assert is_due(i, interval) == expected


def test_next_due_stops_at_every_due_output_and_the_last_step():
    """launch_stage hands the engine every step up to the next step at which
    save_outputs has work, so outputs land on the same steps as with one call
    per step."""
    from types import SimpleNamespace
    from gpe_simulator_amd.helper import is_due
    from gpe_simulator_amd.launch import MAX_STEPS_PER_CALL, next_due
    stage = SimpleNamespace(steps=50, save_params={"psi_interval": 7, "scalar_interval": 0, "nk_interval": 5},
                            monitor_params={"print_scalar_interval": 0, "show_nk_interval": 12})
    stops, i = [], 0
    while i < stage.steps:
        i = next_due(stage, i)
        stops.append(i)
    due = [s for s in range(1, 51) if any(is_due(s, iv) for iv in (7, 5, 12))] + [50]
    assert stops == sorted(set(due))
    long = SimpleNamespace(steps=3 * MAX_STEPS_PER_CALL, save_params={"psi_interval": 0, "scalar_interval": 0,
                                                                      "nk_interval": 0},
                           monitor_params={"print_scalar_interval": 0})
    assert next_due(long, 0) == MAX_STEPS_PER_CALL


def host_stage(tmp_path, n=4):
    """A parsed stage holding a host psi (no solve: the snapshot code then
    reads and writes host arrays only)."""
    from gpe_simulator_amd.yaml_parser import parse_yaml
    path, _ = write_config(tmp_path)
    item = parse_yaml(path)[0]
    if n != item.N_grids:
        raise ValueError("host_stage builds the 4^3 config")
    rng = np.random.default_rng(2)
    item.psi = rng.normal(size=(n,) * 3) + 1j * rng.normal(size=(n,) * 3)
    return item


def test_snapshot_roundtrip_preserves_values_and_checks_metadata(tmp_path):
    """tests/test_workflow.py:110-130 of the reference."""
    from gpe_simulator_amd.state_io import load_wavefunction, save_wavefunction
    item = host_stage(tmp_path)
    path = item.data_dir / "psi_00000006.npy"
    save_wavefunction(path, item, 6)
    loaded = load_wavefunction(path, item)
    np.testing.assert_array_equal(loaded, item.psi)
    metadata = json.loads(path.with_suffix(".json").read_text())
    assert (metadata["step"], metadata["time"]) == (6, 0.75)
    assert metadata["stage"] == item.config
    item.simulation_range += 1
    with pytest.raises(ValueError, match="simulation range"):
        load_wavefunction(path, item)
    item.simulation_range -= 1
    item.N_atoms += 1
    with pytest.raises(ValueError, match="reference N_atom"):
        load_wavefunction(path, item)
    item.N_atoms -= 1
    bad = tmp_path / "bad.npy"
    np.save(bad, np.ones((4, 4, 4), dtype=np.complex64))
    with pytest.raises(ValueError, match="complex128"):
        load_wavefunction(bad, item)
    np.save(bad, np.ones((5, 5, 5), dtype=np.complex128))
    with pytest.raises(ValueError, match="N_grids"):
        load_wavefunction(bad, item)


def test_snapshot_load_checks_legacy_grid_companion(tmp_path):
    """tests/test_workflow.py:132-141 of the reference."""
    from gpe_simulator_amd.state_io import load_wavefunction
    item = host_stage(tmp_path)
    path = tmp_path / "psi_final.npy"
    np.save(path, item.psi)
    grid = path.with_name("psi_final_grid.npz")
    np.savez(grid, x_array=item.x_array + 0.25, k_array=item.k_array)
    with pytest.raises(ValueError, match="grid does not match"):
        load_wavefunction(path, item)
    np.savez(grid, x_array=item.x_array, k_array=item.k_array)
    np.testing.assert_array_equal(load_wavefunction(path, item), item.psi)
