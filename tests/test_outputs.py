"""Stage outputs (gpe_simulator_amd/save.py) on the host.

The reference's save-path tests (tests/test_workflow.py:173-233 of the
reference) restated.  save_outputs reaches the device only through the
DeviceField's solve (diagnostics, angular momentum, downloads), so a host
stand-in for the solve runs them on the CPU: it holds psi as a numpy array
and answers diagnostics with the oracle's formulas.  The stand-in counts its
diagnostics calls where the reference counts its fftn calls.
"""

import csv
import json

import numpy as np
import pytest

from gpe_simulator_amd.field import DeviceField
from oracle import gpe_oracle as orc


class HostSolve:
    """The part of a solve that save_outputs and save_wavefunction use."""

    precision = 1

    def __init__(self, psi, x, k):
        self.psi, self.x, self.k = psi, x, k
        self.n = psi.shape[0]
        self.diagnostic_calls = []

    def to_numpy(self):
        return self.psi.copy()

    def download(self, offset=0, count=None):
        flat = self.psi.ravel()
        return flat[offset:offset + (flat.size - offset if count is None else count)].copy()

    def diagnostics(self, t=0.0, spectrum=True):
        self.diagnostic_calls.append((t, spectrum))
        psi_k = np.fft.fftn(self.psi)
        out = {"norm": orc.get_N(self.psi, self.x), "E_kin": orc.get_E(psi_k, self.x, self.k)}
        if spectrum:
            shells = orc.nk_extraction(psi_k, self.x, self.k)
            view = orc.spectrum_view(psi_k, self.x, self.k)
            dx = self.x[1] - self.x[0]
            # raw per-shell sums of |F|^2, as the engine returns them
            out.update(counts=shells["counts"], occupation=shells["Nk"] * self.psi.size / dx ** 3,
                       view_counts=view["counts"], view_occupation=view["occupation"])
        return out

    def angular_momentum(self):
        return 2.0

    def monitor_images(self, names):
        """The raw sums gpe_monitor_images returns: unshifted, unscaled."""
        power = np.abs(np.fft.fftn(self.psi)) ** 2
        raw = {"TOF_xy": power.sum(axis=2), "TOF_yz": power.sum(axis=0), "TOF_slice_xy": power[:, :, 0],
               "TOF_slice_yz": power[0, :, :], "density_xy": (np.abs(self.psi) ** 2).sum(axis=2),
               "plane_mid": self.psi[:, :, self.n // 2].copy()}
        return {name: raw[name] for name in names}


def workflow_stage(tmp_path, n=4, **save):
    from gpe_simulator_amd.configuration import dump_yaml, new_config
    from gpe_simulator_amd.yaml_parser import parse_yaml
    config = new_config()
    config["simulation"].update(N_grids=n, simulation_range=8, data_root=str(tmp_path / "data"),
                                monitor_root=str(tmp_path / "monitor"))
    first = config["stages"][0]
    first["evolution"].update(type="real", steps=6, dt=0.125)
    first["box"]["enabled"] = False
    first["monitor"] = dict.fromkeys(first["monitor"], 0)
    first["save"].update(scalar_interval=0, save_final=False)
    path = tmp_path / "experiment.yaml"
    path.write_text(dump_yaml(config))
    item = parse_yaml(path)[0]
    item.save_params.update(save)
    rng = np.random.default_rng(4)
    host = rng.normal(size=(n,) * 3) + 1j * rng.normal(size=(n,) * 3)
    solve = HostSolve(host, item.x_array, item.k_array)
    item._psi = DeviceField(solve)
    return item, solve


def test_save_schedules_are_independent_and_do_not_duplicate_final(tmp_path):
    from unittest.mock import patch
    from gpe_simulator_amd import save, state_io
    item, solve = workflow_stage(tmp_path, psi_interval=2, scalar_interval=3, nk_interval=4,
                                 save_initial=True, save_final=True)
    with patch.object(save, "save_wavefunction", wraps=state_io.save_wavefunction) as snapshot:
        save.save_outputs(item, 0, initial=True)
        for step in range(1, 7):
            save.save_outputs(item, step, final=(step == 6))
    assert [call.args[2] for call in snapshot.call_args_list] == [0, 2, 4, 6]
    # one diagnostics evaluation per step with scalar or spectrum output: 3, 4, 6
    assert [t for t, _ in solve.diagnostic_calls] == [0.375, 0.5, 0.75]
    with (item.monitor_dir / "scalars.csv").open() as file:
        rows = list(csv.DictReader(file))
    assert [int(row["step"]) for row in rows] == [3, 6]
    norm = orc.get_N(solve.psi, item.x_array)
    assert float(rows[0]["norm"]) == pytest.approx(norm, rel=1e-12)
    assert float(rows[0]["N_atoms"]) == pytest.approx(item.N_atoms * norm, rel=1e-12)
    with np.load(item.monitor_dir / "nk_00000004.npz") as saved:
        assert saved["time"] == 0.5 and saved["step"] == 4
        assert saved["Nk"].sum() == pytest.approx(norm, rel=1e-12)
        assert saved["counts"].sum() == 4 ** 3
    snap = json.loads((item.data_dir / "psi_00000006.json").read_text())
    assert (snap["step"], snap["time"]) == (6, 0.75)
    np.testing.assert_array_equal(np.load(item.data_dir / "psi_00000006.npy"), solve.psi)


def test_disabled_output_does_no_diagnostics_or_file_io(tmp_path):
    from gpe_simulator_amd import save
    item, solve = workflow_stage(tmp_path)
    save.save_outputs(item, 0, initial=True)
    save.save_outputs(item, 6, final=True)
    assert solve.diagnostic_calls == []
    assert not item.data_dir.exists() and not item.monitor_dir.exists()


def test_monitor_arrays_include_axes_and_correct_projection_norm(tmp_path):
    from gpe_simulator_amd import save
    item, solve = workflow_stage(tmp_path)
    solve.psi[0, 0, 2] = 0
    item.monitor_params.update({key: 1 for key in item.monitor_params
                                if key.startswith("show_") and key != "show_nk_interval"})
    save.save_outputs(item, 1)
    norm = orc.get_N(solve.psi, item.x_array)
    dk = item.k_array[1] - item.k_array[0]
    for name in ("TOF_xy", "TOF_yz"):
        with np.load(item.monitor_dir / f"{name}_00000001.npz") as saved:
            assert saved["values"].sum() * dk ** 2 == pytest.approx(norm, rel=1e-12)
            assert saved["values"].shape == (4, 4)
    with np.load(item.monitor_dir / "density_xy_00000001.npz") as saved:
        dx = item.x_array[1] - item.x_array[0]
        assert saved["values"].sum() * dx ** 2 == pytest.approx(norm, rel=1e-12)
        np.testing.assert_array_equal(saved["x"], item.x_array)
    with np.load(item.monitor_dir / "phase_xy_00000001.npz") as saved:
        assert np.isnan(saved["values"][0, 0])
        assert saved["z"] == 0


def test_two_point_grid_has_positive_momentum_density(tmp_path):
    from gpe_simulator_amd import save
    item, solve = workflow_stage(tmp_path, n=2)
    solve.psi[:] = 1.0
    item.monitor_params["show_TOF_slice_interval"] = 1
    save.save_outputs(item, 1)
    with np.load(item.monitor_dir / "TOF_slice_xy_00000001.npz") as saved:
        assert saved["values"].sum() > 0


def test_every_monitor_image_matches_the_oracle(tmp_path):
    """Each saved monitor array equals the oracle's restatement of
    gpe/save.py:59-90 (scale factors, fftshift, the mid plane and the phase
    mask) for the same field."""
    from gpe_simulator_amd import save
    item, solve = workflow_stage(tmp_path, n=6)
    solve.psi[1, 2, 3] = 0
    item.monitor_params.update({key: 1 for key in item.monitor_params
                                if key.startswith("show_") and key != "show_nk_interval"})
    save.save_outputs(item, 1)
    expected = orc.monitor_images(solve.psi, np.asarray(item.x_array), np.asarray(item.k_array))
    for name, ref in expected.items():
        with np.load(item.monitor_dir / f"{name}_00000001.npz") as saved:
            np.testing.assert_allclose(saved["values"], ref, rtol=1e-12, atol=0)
