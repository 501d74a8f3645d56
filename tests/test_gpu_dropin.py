"""The drop-in API on the device: the reference's closed-form propagator tests
(tests/test_propagators.py:32-200 of the reference) restated through this
package's stage / parse_yaml / launch_stage / main."""

import csv
import json
from unittest.mock import patch

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


@pytest.fixture
def env(tmp_path, require_gpu):
    from gpe_simulator_amd.configuration import dump_yaml, new_config
    from gpe_simulator_amd.yaml_parser import parse_yaml

    config = new_config()
    config["simulation"].update(N_grids=8, simulation_range=TWO_PI, data_root=str(tmp_path / "data"),
                                monitor_root=str(tmp_path / "monitor"))
    settings = config["stages"][0]
    settings["box"]["enabled"] = False
    settings["cloud"]["a"] = 0
    settings["evolution"].update(steps=3, dt=0.01)
    settings["save"].update(scalar_interval=0, save_final=False)
    settings["monitor"] = dict.fromkeys(settings["monitor"], 0)

    class Env:
        pass

    e = Env()
    e.config, e.root = config, tmp_path

    def stage(precision="complex128"):
        path = tmp_path / "test.yaml"
        path.write_text(dump_yaml(config))
        item = parse_yaml(path, precision=precision)[0]
        item.prepare()
        return item

    e.stage = stage
    return e


def plane_wave(item, mode=1):
    wave = np.exp(1j * mode * item.x_array)[:, None, None]
    return np.broadcast_to(wave, (item.N_grids,) * 3).copy() / TWO_PI ** 1.5


@pytest.mark.parametrize("precision,rtol", [("complex128", 1e-11), ("complex64", 3e-6)])
def test_real_phase_for_all_potential_and_cutoff_combinations(env, precision, rtol):
    from gpe_simulator_amd.helper import get_N
    from gpe_simulator_amd.launch import launch_stage
    settings = env.config["stages"][0]
    settings["evolution"]["type"] = "real"
    settings["cloud"]["a"] = 10
    for static in (False, True):
        for dynamic in (False, True):
            for cutoff in (False, True):
                settings["potential"]["static"].update(enabled=static, expression="1/2")
                settings["potential"]["dynamic"].update(enabled=dynamic, expression="t/4")
                settings["k_cut"].update(enabled=cutoff, value=1.5)
                item = env.stage(precision)
                initial = plane_wave(item)
                item.psi = initial
                time = item.steps * item.dt
                phase = (1 + item.beta / TWO_PI ** 3) * time
                phase += (0.5 * time if static else 0) + (time ** 2 / 8 if dynamic else 0)
                result = launch_stage(item)
                np.testing.assert_allclose(np.asarray(result), initial * np.exp(-1j * phase), rtol=rtol,
                                           atol=rtol * 1e-1)
                assert get_N(result) == pytest.approx(1, abs=10 * rtol)
                item.release()


def test_per_step_propagator_matches_batched(env):
    """Calling the propagator once per step (reference style) and launch_stage's
    batched calls give the same state."""
    from gpe_simulator_amd.launch import launch_stage
    settings = env.config["stages"][0]
    settings["evolution"].update(type="real", steps=5)
    settings["cloud"]["a"] = 10
    settings["potential"]["static"].update(enabled=True, expression="x**2/3 + cos(y)")
    settings["potential"]["dynamic"].update(enabled=True, expression="sin(t)*z")
    a, b = env.stage(), env.stage()
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(8, 8, 8)) + 1j * rng.normal(size=(8, 8, 8))
    # normalised, so the nonlinear phase per step (beta |psi|^2 dt) stays small
    # and rounding differences are not amplified by the dynamics
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2) * (TWO_PI / 8) ** 3)
    a.psi = psi
    b.psi = psi
    for i in range(5):
        a.psi = a.propagator(a, i * a.dt)
    launch_stage(b)
    np.testing.assert_allclose(np.asarray(a.psi), np.asarray(b.psi), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("static", [False, True])
def test_imaginary_decay_and_normalization(env, static):
    from gpe_simulator_amd.launch import launch_stage
    env.config["stages"][0]["potential"]["static"].update(enabled=static, expression="1/2")
    item = env.stage()
    low, high = plane_wave(item, 1), plane_wave(item, 2)
    item.psi = low + high
    time = item.steps * item.dt
    expected = low * np.exp(-time) + high * np.exp(-4 * time)
    dx = float(item.x_array[1] - item.x_array[0])
    expected /= np.sqrt(np.sum(np.abs(expected) ** 2) * dx ** 3)
    np.testing.assert_allclose(np.asarray(launch_stage(item)), expected, rtol=1e-11, atol=1e-12)


def test_absorber_keeps_analytical_norm_loss(env):
    from gpe_simulator_amd.helper import get_N
    from gpe_simulator_amd.launch import launch_stage
    env.config["stages"][0]["evolution"]["type"] = "real"
    item = env.stage()
    initial = plane_wave(item)
    item.psi = initial
    gamma = 0.4
    item.staticV.fill(-1j * gamma)
    time = item.steps * item.dt
    result = launch_stage(item)
    np.testing.assert_allclose(np.asarray(result), initial * np.exp((-gamma - 1j) * time), rtol=1e-11, atol=1e-12)
    assert get_N(result) == pytest.approx(np.exp(-2 * gamma * time), abs=1e-11)


def test_cutoff_removes_high_mode_without_renormalizing(env):
    from gpe_simulator_amd.helper import get_N
    from gpe_simulator_amd.launch import launch_stage
    settings = env.config["stages"][0]
    settings["evolution"]["type"] = "real"
    settings["k_cut"].update(enabled=True, value=1.5)
    item = env.stage()
    low, high = plane_wave(item, 1), plane_wave(item, 2)
    item.psi = (low + high) / np.sqrt(2)
    result = launch_stage(item)
    expected = low * np.exp(-1j * item.steps * item.dt) / np.sqrt(2)
    np.testing.assert_allclose(np.asarray(result), expected, rtol=1e-11, atol=1e-12)
    assert get_N(result) == pytest.approx(0.5, abs=1e-11)


def test_snapshot_roundtrip_across_slab_boundary(env):
    from gpe_simulator_amd.state_io import load_wavefunction, save_wavefunction
    env.config["simulation"]["N_grids"] = 18
    item = env.stage()
    rng = np.random.default_rng(5)
    original = rng.normal(size=(18,) * 3) + 1j * rng.normal(size=(18,) * 3)
    item.psi = original
    path = env.root / "state.npy"
    save_wavefunction(path, item, 0)
    np.testing.assert_array_equal(np.load(path), original)
    item.psi = np.zeros((18,) * 3)
    field = load_wavefunction(path, item)
    np.testing.assert_array_equal(np.asarray(field), original)


def test_angular_momentum_of_vortex(require_gpu):
    """Lz = +-N for (x +- i y) exp(-r^2/2) (tests/test_diagnostics.py:42-49)."""
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.grids import get_grids
    x, k = get_grids(48, 16)
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    for charge in (-1, 1):
        psi = (X + charge * 1j * Y) * np.exp(-(X * X + Y * Y + Z * Z) / 2)
        solve = make_solve(48, x, k, "real", 1e-3, 0.0)
        solve.upload(psi)
        norm = solve.norm()
        assert solve.angular_momentum() == pytest.approx(charge * norm, rel=1e-9, abs=1e-10)


def test_yaml_run_with_two_stages_and_all_saved_diagnostics(env):
    from gpe_simulator_amd.configuration import dump_yaml
    from gpe_simulator_amd.main import main
    first = env.config["stages"][0]
    first["evolution"]["steps"] = 2
    first["save"].update(save_final=True, scalar_interval=1, nk_interval=1)
    first["monitor"].update({name: 1 for name in first["monitor"] if name.startswith("show_")})
    second = json.loads(json.dumps(first))
    second["evolution"]["type"] = "real"
    second["initial_state"]["type"] = "continue"
    second["box"].update(enabled=True, R=1, L=2, wall_thickness=0.5, trap_depth=0.3)
    second["noise"]["enabled"] = True
    second["absorber"].update(enabled=True, activation_distance=0.5, strength=0.2)
    second["potential"]["static"].update(enabled=True, expression="x**2/20")
    second["potential"]["dynamic"].update(enabled=True, expression="cos(t)*y/10")
    second["k_cut"].update(enabled=True, value=2)
    env.config["stages"].append(second)
    path = env.root / "stages.yaml"
    path.write_text(dump_yaml(env.config))
    with patch("builtins.print"):
        main([str(path)])
    states = sorted((env.root / "data").glob("*/stage_*/psi_*.npy"))
    assert len(states) == 2
    for state in states:
        saved = np.load(state)
        assert saved.dtype == np.complex128 and np.isfinite(saved).all()
    scalar_files = list((env.root / "monitor").glob("*/stage_*/scalars.csv"))
    assert len(scalar_files) == 2
    for scalar in scalar_files:
        with scalar.open() as file:
            assert [row["step"] for row in csv.DictReader(file)] == ["1", "2"]
        assert len(list(scalar.parent.glob("*_00000002.npz"))) == 8


def test_schema3_nk_view_output(env):
    """save.nk_view (schema 3) writes the 128-bin view beside every nk file;
    without it the reference's file set is unchanged."""
    from gpe_simulator_amd.configuration import dump_yaml
    from gpe_simulator_amd.main import main
    env.config["schema_version"] = 3
    env.config["simulation"]["N_grids"] = 16
    settings = env.config["stages"][0]
    settings["evolution"].update(steps=2)
    settings["save"].update(nk_interval=1, nk_view=True)
    second = json.loads(json.dumps(settings))
    second["name"] = "plain"
    second["initial_state"]["type"] = "continue"
    second["save"]["nk_view"] = False
    env.config["stages"].append(second)
    path = env.root / "view.yaml"
    path.write_text(dump_yaml(env.config))
    with patch("builtins.print"):
        main([str(path)])
    folders = sorted((env.root / "monitor").glob("*/stage_*"))
    assert len(folders) == 2
    views = sorted(folders[0].glob("nkview_*.npz"))
    assert [v.name for v in views] == ["nkview_00000001.npz", "nkview_00000002.npz"]
    assert len(list(folders[0].glob("nk_*.npz"))) == 2
    for view in views:
        with np.load(view) as data:
            assert data["counts"].sum() == 16 ** 3 and len(data["counts"]) == 128
            assert data["Nk"].sum() == pytest.approx(1.0, rel=1e-10)      # imaginary time: unit norm
            assert data["edges"][-1] == pytest.approx(float(data["k_max"]), rel=1e-15)
    assert not list(folders[1].glob("nkview_*.npz"))
    assert len(list(folders[1].glob("nk_*.npz"))) == 2


@pytest.mark.parametrize("n", [18, 64])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_device_monitor_images_and_spectrum_match_the_oracle(n, precision, require_gpu):
    """gpe_monitor_images' raw device sums, turned into the reference's arrays
    the way save.py does, against the oracle's restatement of
    gpe/save.py:59-90; the device fftn against numpy's; on the generic (18)
    and a fast (64) transform, with a pending renormalisation in imaginary
    time."""
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.field import DeviceField
    from gpe_simulator_amd.helper import fftn, get_E
    from gpe_simulator_amd.nk_extraction import nk_extraction
    length = 12.0
    x, k = orc.get_grids(n, length)
    solve = make_solve(n, x, k, "imaginary", 1e-3, 50.0, static="x**2 + y**2 + z**2", precision=precision)
    rng = np.random.default_rng(n)
    psi0 = orc.gaussian(x, 1.3) * (1 + 0.05 * rng.standard_normal((n,) * 3)) * np.exp(0.3j * x[None, :, None])
    solve.upload(psi0)
    solve.step(3)                                   # leaves a pending 1/sqrt(norm) on the device
    psi = solve.to_numpy()
    tol = 1e-12 if precision == "c128" else 3e-6
    spec = solve.spectrum()
    np.testing.assert_allclose(spec, np.fft.fftn(psi), rtol=0, atol=tol * np.abs(spec).max())
    names = ("TOF_xy", "TOF_yz", "TOF_slice_xy", "TOF_slice_yz", "density_xy", "plane_mid")
    raw = solve.monitor_images(names)
    dx, dk = x[1] - x[0], abs(k[1] - k[0])
    scale = dx ** 3 / (n ** 3 * dk ** 3)
    shift = np.fft.fftshift
    got = {"TOF_xy": shift(raw["TOF_xy"]) * scale * dk, "TOF_yz": shift(raw["TOF_yz"]) * scale * dk,
           "TOF_slice_xy": shift(raw["TOF_slice_xy"]) * scale, "TOF_slice_yz": shift(raw["TOF_slice_yz"]) * scale,
           "density_xy": raw["density_xy"] * dx, "density_slice_xy": np.abs(raw["plane_mid"]) ** 2}
    ref = orc.monitor_images(psi, x, k)
    for name, values in got.items():
        np.testing.assert_allclose(values, ref[name], rtol=0, atol=10 * tol * np.abs(ref[name]).max(), err_msg=name)
    np.testing.assert_array_equal(raw["plane_mid"], psi[:, :, n // 2])
    # the reference's helper signatures: get_E and nk_extraction take fftn(psi)
    field = DeviceField(solve)
    psi_k = fftn(field)
    assert get_E(psi_k, x, k) == pytest.approx(orc.get_E(np.fft.fftn(psi), x, k), rel=10 * tol)
    shells = nk_extraction(psi_k, x, k)
    np.testing.assert_array_equal(shells["counts"], orc.nk_extraction(np.fft.fftn(psi), x, k)["counts"])
    with pytest.raises(TypeError):
        get_E(np.fft.fftn(psi), x, k)
    solve.close()


@pytest.mark.parametrize("n", [18, 64])
@pytest.mark.parametrize("kind", ["real", "imaginary"])
def test_box_and_absorber_volumes_match_the_oracle(n, kind, require_gpu):
    """The static volumes that add_box / add_absorber evaluate on the device
    (FP64 programs with a max operation) against the oracle's restatement of
    gpe/potentials.py:3-37: the box as the real part, the absorber as the
    negated imaginary part, and no imaginary part in imaginary time
    (gpe/stage.py:57-58)."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.potentials import add_absorber, add_box
    length = 16.0
    x, k = orc.get_grids(n, length)
    box = {"R": 5.0, "L": 6.0, "wall_thickness": 1.0, "trap_depth": 40.0}
    absorber = {"activation_distance": 2.0, "strength": 3.0}
    solve = make_solve(n, x, k, kind, 1e-3, 10.0, precision="c64")
    add_box(solve, box)
    want_re = orc.box_potential(x, box["R"], box["L"], box["wall_thickness"], box["trap_depth"])
    np.testing.assert_allclose(solve.volume(_lib.TGT_VOL), want_re, rtol=1e-14, atol=1e-300)
    if kind == "real":
        add_absorber(solve, absorber, length)
        ramps = orc.absorption_potential(absorber["activation_distance"], absorber["strength"], x, length)
        want = orc.total_static_potential(want_re, ramps)
        np.testing.assert_allclose(solve.volume(_lib.TGT_VOL_IM), want.imag, rtol=1e-14, atol=1e-14)
        assert solve.volume(_lib.TGT_VOL_IM).min() == pytest.approx(-3 * 2.0 * 3.0)   # a corner: three faces
    else:
        with pytest.raises(_lib.GpeError):
            solve.volume(_lib.TGT_VOL_IM)
    solve.close()
