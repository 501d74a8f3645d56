"""Initial states (gpe/initial_states.py:4-14): built on the device."""

from .field import DeviceField


def gaussian(solve, sigma):
    """Separable exp(-r^2/(2 sigma^2)), normalised, into `solve`'s field."""
    solve.gaussian(sigma)
    return DeviceField(solve)
