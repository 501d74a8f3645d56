"""Static volume potentials built on the device.

Counterparts of gpe/potentials.py:3-61 of the reference.  The box trap and the
absorber are expressed as FP64 programs (with an internal max operation) and
evaluated by the engine into the static volume; imaginary-time stages get a
real volume (gpe/stage.py:57-58), real-time stages box - i*absorber
(gpe/potentials.py:31-37).
"""

from __future__ import annotations

import ast

import numpy as np

from . import _lib
from .expressions import compile_program, vm_max


def _c(value):
    return ast.Constant(value=float(value))


def _name(name):
    return ast.Name(id=name, ctx=ast.Load())


def _bin(left, op, right):
    return ast.BinOp(left=left, op=op, right=right)


def _call(name, arg):
    return ast.Call(func=_name(name), args=[arg], keywords=[])


def box_tree(box_radius, box_length, wall_thickness, trap_depth):
    """depth * max(tube, sheet), tube = exp(-(sqrt(x^2+y^2)-R)^2/(2w^2)),
    sheet = exp(-(|z|-L/2)^2/(2w^2))  (gpe/potentials.py:3-11)."""
    two_w2 = _c(2 * wall_thickness ** 2)
    rho = _call("sqrt", _bin(_bin(_name("x"), ast.Pow(), _c(2)), ast.Add(), _bin(_name("y"), ast.Pow(), _c(2))))
    tube = _call("exp", ast.UnaryOp(op=ast.USub(), operand=_bin(
        _bin(_bin(rho, ast.Sub(), _c(box_radius)), ast.Pow(), _c(2)), ast.Div(), two_w2)))
    sheet = _call("exp", ast.UnaryOp(op=ast.USub(), operand=_bin(
        _bin(_bin(_call("abs", _name("z")), ast.Sub(), _c(box_length / 2)), ast.Pow(), _c(2)), ast.Div(), two_w2)))
    return _bin(vm_max(tube, sheet), ast.Mult(), _c(trap_depth))


def absorber_tree(activation_distance, amplitude, simulation_range):
    """amplitude * sum over faces of max(d - distance to face, 0)
    (gpe/potentials.py:13-28)."""
    half = simulation_range / 2
    terms = []
    for axis in ("x", "y", "z"):
        low = vm_max(_bin(_c(activation_distance), ast.Sub(), _bin(_name(axis), ast.Add(), _c(half))), _c(0))
        high = vm_max(_bin(_c(activation_distance), ast.Sub(), _bin(_c(half), ast.Sub(), _name(axis))), _c(0))
        terms += [low, high]
    total = terms[0]
    for term in terms[1:]:
        total = _bin(total, ast.Add(), term)
    return _bin(total, ast.Mult(), _c(amplitude))


def add_box(solve, box_params):
    tree = box_tree(box_params["R"], box_params["L"], box_params["wall_thickness"], box_params["trap_depth"])
    solve.potential_static(_lib.TGT_VOL, compile_program(tree))


def add_absorber(solve, absorber_params, simulation_range):
    """Adds -absorber to the imaginary part of the static volume."""
    tree = absorber_tree(absorber_params["activation_distance"], absorber_params["strength"], simulation_range)
    solve.potential_static(_lib.TGT_VOL_IM, compile_program(tree), scale=-1.0)


def add_white_noise(solve, N_grid, trap_depth, fraction, seed=42):
    """Add Gaussian disorder of width fraction*trap_depth to the real part.

    The reference draws from CuPy's generator (gpe/potentials.py:39-61), which
    is not available here; this draws numpy's PCG64 stream with the same seed,
    16 x-planes at a time, and accumulates it into the device volume, so the
    disorder values differ from the reference's (parity of the disorder field
    is not pinned; its statistics are)."""
    rng = np.random.default_rng(seed)
    sigma = float(fraction * trap_depth)
    planes = min(16, N_grid)
    for start in range(0, N_grid, planes):
        stop = min(start + planes, N_grid)
        noise = rng.standard_normal((stop - start, N_grid, N_grid)) * sigma
        solve.volume_upload(_lib.TGT_VOL, noise.ravel(), offset=start * N_grid * N_grid, accumulate=True)
