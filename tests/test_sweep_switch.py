"""Host logic of the complex64 -> complex128 switch (no device needed)."""

import numpy as np
import pytest

from gpe_simulator_amd.sweep import FieldFloor, Switch, mu_from_norms


def feed(switch, steps, dmus):
    for s, d in zip(steps, dmus):
        if switch.update(s, d):
            return s
    return None


def decaying(rate, amplitude, floor, steps, seed=0):
    """Exponential |dmu| decay meeting a noisy rounding floor."""
    rng = np.random.default_rng(seed)
    clean = amplitude * np.exp(-rate * steps)
    noise = floor * rng.uniform(0.2, 2.0, steps.size)
    return np.maximum(clean, noise), clean


def test_target_is_where_the_clean_decay_reaches_tol():
    steps = np.arange(50, 20001, 50)
    rate, amp, floor, tol = np.log(4) / 400, 1e-1, 1e-6, 2e-9
    dmus, _ = decaying(rate, amp, floor, steps)
    switch = Switch(tol)
    at = feed(switch, steps, dmus)
    exact = np.log(amp / tol) / rate
    assert switch.target == pytest.approx(exact, rel=0.02)
    assert at == steps[np.searchsorted(steps, switch.target)]


def test_no_switch_while_dmu_still_falls():
    steps = np.arange(50, 3001, 50)
    switch = Switch(1e-12)
    assert feed(switch, steps, 1e-2 * np.exp(-steps / 300.0)) is None
    assert switch.target is None


def test_dmu_below_tol_switches_at_once():
    switch = Switch(1e-9)
    assert not switch.update(50, 1e-3)
    assert switch.update(100, 5e-10)


def test_flat_history_switches_when_the_plateau_shows():
    # no clean decay to fit: switch as soon as the floor is recognised
    steps = np.arange(50, 5001, 50)
    rng = np.random.default_rng(1)
    dmus = 1e-6 * rng.uniform(0.5, 1.5, steps.size)
    switch = Switch(1e-9, plateau=10)
    at = feed(switch, steps, dmus)
    assert at is not None and switch.target is None
    # the first check at which no new minimum appeared in the last 10
    first = next(j for j in range(11, steps.size) if min(dmus[j - 9:j + 1]) >= min(dmus[:j - 9]))
    assert at == steps[first]


def test_mu_from_norms_inverts_norm_decay():
    dt, mu = 5e-4, 7.25
    norms = np.full(10, np.exp(-2 * mu * dt))
    np.testing.assert_allclose(mu_from_norms(norms, dt), mu, rtol=1e-13)


def test_measured_field_floor_moves_the_switch_earlier():
    steps = np.arange(50, 20001, 50)
    rate, amp, floor, tol = np.log(4) / 400, 1e-1, 1e-6, 2e-9
    dmus, _ = decaying(rate, amp, floor, steps)
    field = FieldFloor(samples=2)
    late = Switch(tol, field_floor=field)
    feed(late, steps, dmus)
    assert late.late
    field.add(1e-8)
    assert field.value is None
    field.add(3e-8)
    field.add(1e-3)                   # later readings do not feed back
    assert field.value == pytest.approx(2e-8)
    early = Switch(tol, field_floor=field)
    at = feed(early, steps, dmus)
    assert not early.late
    # the fitted line reaches the measured floor after the plateau is seen
    assert at == steps[np.searchsorted(steps, np.log(amp / 2e-8) / rate)]
    assert at < late.target


def test_mu_below_tol_marks_the_switch_late():
    switch = Switch(1e-9, field_floor=FieldFloor())
    assert switch.update(50, 1e-10) and switch.late


def test_shallow_fit_is_bounded_by_the_horizon():
    # a slow member: the clean part of its history is a transient that
    # extrapolates far beyond where the floor was recognised
    steps = np.arange(50, 40001, 50)
    dmus = np.maximum(1e-3 * np.exp(-steps / 3000.0), 1e-6 * np.random.default_rng(2).uniform(0.2, 2.0, steps.size))
    switch = Switch(2e-9, horizon=1.5)
    at = feed(switch, steps, dmus)
    assert switch.target > 1.5 * switch.detected
    assert at == steps[np.searchsorted(steps, 1.5 * switch.detected)]
