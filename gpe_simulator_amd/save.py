"""Snapshots and diagnostics after a completed step (gpe/save.py:13-90).

Same schedule, files and columns as the reference.  Norm, kinetic energy, the
|psi(k)|^2 shell histogram and Lz come from the engine's FP64 device
reductions.  The monitor images come from device sums as well: one fftn of
the field on the device serves every momentum image (the reference shares one
fftn, gpe/save.py:31), and the projections and slices are FP64 sums over the
device fields (sum_abs2, gpe/helper.py:3-12), so only n x n planes cross to the
host, where the reference's scale factors, fftshift and the phase mask apply.
"""

from __future__ import annotations

import csv

import numpy as np

from .helper import is_due
from .nk_extraction import shell_summary, view_summary
from .state_io import save_wavefunction

# monitor schedule key -> device image (gpe_monitor_images)
_IMAGE_OF = {"show_TOF_image_interval": "TOF_xy", "show_verticle_TOF_interval": "TOF_yz",
             "show_TOF_slice_interval": "TOF_slice_xy", "show_verticle_TOF_slice_interval": "TOF_slice_yz",
             "show_magnifier_interval": "density_xy"}


def save_outputs(stage, step, *, initial=False, final=False):
    save, monitor = stage.save_params, stage.monitor_params
    if (is_due(step, save["psi_interval"]) or (initial and save["save_initial"])
            or (final and save["save_final"])):
        save_wavefunction(stage.data_dir / f"psi_{step:08d}.npy", stage, step)
    if initial:
        return
    scalar_save = is_due(step, save["scalar_interval"])
    scalar_print = is_due(step, monitor["print_scalar_interval"])
    spectrum_due = is_due(step, save["nk_interval"]) or is_due(step, monitor["show_nk_interval"])
    due = {name: is_due(step, interval) for name, interval in monitor.items()}
    time = step * stage.dt
    psi = stage.psi

    def save_npz(name, **arrays):
        stage.monitor_dir.mkdir(parents=True, exist_ok=True)
        np.savez(stage.monitor_dir / f"{name}_{step:08d}.npz", step=step, time=time,
                 **{key: np.asarray(value) for key, value in arrays.items()})

    diag = None
    if scalar_save or scalar_print or spectrum_due:
        diag = psi.solve.diagnostics(time, spectrum=spectrum_due)
    if scalar_save or scalar_print:
        norm, energy = diag["norm"], diag["E_kin"]
        angular_momentum = psi.solve.angular_momentum()
        if scalar_print:
            print(f"{stage.name}: step={step} t={time:g} N={norm:.8g} "
                  f"E_kin={energy:.8g} Lz={angular_momentum:.8g}", flush=True)
        if scalar_save:
            stage.monitor_dir.mkdir(parents=True, exist_ok=True)
            path = stage.monitor_dir / "scalars.csv"
            new_file = not path.exists()
            with path.open("a", newline="") as file:
                writer = csv.writer(file)
                if new_file:
                    writer.writerow(["step", "time", "norm", "N_atoms", "E_kin", "Lz"])
                writer.writerow([step, time, norm, stage.N_atoms * norm, energy, angular_momentum])
    if spectrum_due:
        save_npz("nk", **shell_summary(diag["counts"], diag["occupation"], stage.x_array, stage.k_array,
                                       stage.N_grids))
        if save.get("nk_view"):
            # the equal-width 128-bin view of the same spectrum (schema 3
            # only; the name does not match the reference's nk_* files)
            save_npz("nkview", **view_summary(diag["view_counts"], diag["view_occupation"], stage.x_array,
                                              stage.k_array, stage.N_grids))

    wanted = [name for field, name in _IMAGE_OF.items() if due[field]]
    if due["show_magnifier_slice_interval"] or due["show_phase_interval"]:
        wanted.append("plane_mid")
    if not wanted:
        return
    images = psi.solve.monitor_images(wanted)
    x, k = np.asarray(stage.x_array), np.asarray(stage.k_array)
    dx = float(x[1] - x[0])
    dk = abs(float(k[1] - k[0]))
    scale = dx ** 3 / (stage.N_grids ** 3 * dk ** 3)
    ks = np.fft.fftshift(k)
    if "TOF_xy" in images:
        save_npz("TOF_xy", values=np.fft.fftshift(images["TOF_xy"]) * (scale * dk), kx=ks, ky=ks)
    if "TOF_yz" in images:
        save_npz("TOF_yz", values=np.fft.fftshift(images["TOF_yz"]) * (scale * dk), ky=ks, kz=ks)
    if "TOF_slice_xy" in images:
        save_npz("TOF_slice_xy", values=np.fft.fftshift(images["TOF_slice_xy"]) * scale, kx=ks, ky=ks)
    if "TOF_slice_yz" in images:
        save_npz("TOF_slice_yz", values=np.fft.fftshift(images["TOF_slice_yz"]) * scale, ky=ks, kz=ks)
    if "density_xy" in images:
        save_npz("density_xy", values=images["density_xy"] * dx, x=x, y=x)
    if "plane_mid" in images:
        middle = stage.N_grids // 2
        plane = images["plane_mid"]
        density = plane.real ** 2 + plane.imag ** 2
        if due["show_magnifier_slice_interval"]:
            save_npz("density_slice_xy", values=density, x=x, y=x, z=x[middle])
        if due["show_phase_interval"]:
            phase = np.where(density > density.max() * 1e-12, np.angle(plane), np.nan)
            save_npz("phase_xy", values=phase, x=x, y=x, z=x[middle])
