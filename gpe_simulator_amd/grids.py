"""Coordinate and wavenumber axes (gpe/grids.py:3-6 of the reference).

Axes are small float64 host arrays; the engine copies them to the device
once per solve."""

import numpy as np


def get_grids(N_grids, simulation_range):
    """x = linspace(-L/2, L/2, N, endpoint=False); k = 2*pi*fftfreq(N, L/N)."""
    x_grid = np.linspace(-simulation_range / 2, simulation_range / 2, N_grids, endpoint=False)
    k_grid = 2 * np.pi * np.fft.fftfreq(N_grids, simulation_range / N_grids)
    return x_grid, k_grid
