"""Snapshots and diagnostics after a completed step (gpe/save.py:13-90).

Same schedule, files and columns as the reference.  Norm, kinetic energy, the
|psi(k)|^2 shell histogram and Lz come from the engine's FP64 device
reductions (one forward transform per due step, as the reference shares one
fftn).  The monitor images (TOF and density projections, slices, phase) are
small host-side post-processing of a downloaded copy on due steps only.
"""

from __future__ import annotations

import csv

import numpy as np

from .field import asnumpy
from .helper import is_due
from .nk_extraction import shell_summary, view_summary
from .state_io import save_wavefunction

TOF_FIELDS = ("show_TOF_image_interval", "show_verticle_TOF_interval", "show_TOF_slice_interval",
              "show_verticle_TOF_slice_interval")


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

    images = any(due[name] for name in TOF_FIELDS) or due["show_magnifier_interval"] \
        or due["show_magnifier_slice_interval"] or due["show_phase_interval"]
    if not images:
        return
    host = asnumpy(psi)
    x, k = np.asarray(stage.x_array), np.asarray(stage.k_array)
    if any(due[name] for name in TOF_FIELDS):
        psi_k = np.fft.fftn(host)
        dx = float(x[1] - x[0])
        dk = abs(float(k[1] - k[0]))
        scale = dx ** 3 / (host.size * dk ** 3)
        ks = np.fft.fftshift(k)
        power = np.abs(psi_k) ** 2
        if due["show_TOF_image_interval"]:
            save_npz("TOF_xy", values=np.fft.fftshift(power.sum(axis=2)) * (scale * dk), kx=ks, ky=ks)
        if due["show_verticle_TOF_interval"]:
            save_npz("TOF_yz", values=np.fft.fftshift(power.sum(axis=0)) * (scale * dk), ky=ks, kz=ks)
        if due["show_TOF_slice_interval"]:
            save_npz("TOF_slice_xy", values=np.fft.fftshift(power[:, :, 0]) * scale, kx=ks, ky=ks)
        if due["show_verticle_TOF_slice_interval"]:
            save_npz("TOF_slice_yz", values=np.fft.fftshift(power[0, :, :]) * scale, ky=ks, kz=ks)
    if due["show_magnifier_interval"]:
        save_npz("density_xy", values=(np.abs(host) ** 2).sum(axis=2) * float(x[1] - x[0]), x=x, y=x)
    if due["show_magnifier_slice_interval"] or due["show_phase_interval"]:
        middle = stage.N_grids // 2
        plane = host[:, :, middle]
        density = np.abs(plane) ** 2
        if due["show_magnifier_slice_interval"]:
            save_npz("density_slice_xy", values=density, x=x, y=x, z=x[middle])
        if due["show_phase_interval"]:
            phase = np.where(density > density.max() * 1e-12, np.angle(plane), np.nan)
            save_npz("phase_xy", values=phase, x=x, y=x, z=x[middle])
