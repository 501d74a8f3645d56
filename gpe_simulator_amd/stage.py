"""Stage lifecycle: the reference's `stage` class, backed by a device solve.

Same constructor and attributes as gpe/stage.py:11-51; `prepare` (:53-66)
creates the libgpe_hip solve, loads the potential programs and the static
volume (box, absorber, disorder), sets the initial state and builds the
propagator; `release` (:68-72) drops the stage's references.  The returned
wavefunction (a DeviceField) keeps its solve alive, so a `continue` stage can
copy it on the device after the previous stage is released.
"""

from __future__ import annotations

import os

import numpy as np

from . import _lib
from .engine import make_solve
from .expressions import tree_from_c
from .field import DeviceField
from .helper import get_beta
from .make_propagator import make_imag_propagator, make_real_propagator
from .potentials import add_absorber, add_box, add_white_noise
from .state_io import load_wavefunction

PRECISION_ENV = "GPE_HIP_PRECISION"


def default_precision():
    return os.environ.get(PRECISION_ENV, "complex128")


class StaticVolume:
    """The static potential volume of a prepared stage (stage.staticV)."""

    def __init__(self, solve, kind):
        self.solve, self.kind = solve, kind

    @property
    def shape(self):
        return (self.solve.n,) * 3

    @property
    def dtype(self):
        return np.dtype(np.float64 if self.kind == "imaginary" else np.complex128)

    def fill(self, value):
        value = complex(value)
        self.solve.volume_fill(_lib.TGT_VOL, value.real)
        if self.kind == "real":
            self.solve.volume_fill(_lib.TGT_VOL_IM, value.imag)
        elif value.imag != 0:
            raise TypeError("imaginary time evolution requires a real static potential")


class stage:
    def __init__(self, simulation_range: float, x_array, k_array, N_grids: int, type: str, steps: int,
                 dt: float, initial_state: dict, box_params: dict, absorber_params: dict, noise_params: dict,
                 cloud_params: dict, monitor_params: dict, save_params: dict, potential_params: dict,
                 k_cut_params: dict, precision=None):
        if not isinstance(x_array, np.ndarray):
            raise TypeError("x_array should be a numpy array")
        if not isinstance(k_array, np.ndarray):
            raise TypeError("k_array should be a numpy array")
        if not x_array.dtype == np.float64:
            raise TypeError("x_array should be a real array")
        if not k_array.dtype == np.float64:
            raise TypeError("k_array should be a real array")
        if not isinstance(type, str):
            raise TypeError("type of the stage should be a string")
        if type not in {"imaginary", "real"}:
            raise KeyError("type of the stage should either be 'real' or 'imaginary")
        self.absorber_params = absorber_params
        self.noise_params = noise_params
        self.simulation_range = simulation_range
        self.x_array = x_array
        self.k_array = k_array
        self.N_grids = N_grids
        self.type = type
        self.steps = steps
        self.dt = dt
        self.initial_state = initial_state
        self.box_params = box_params
        self.a = cloud_params["a"]
        self.N_atoms = cloud_params["N_atom"]
        self.save_params = save_params
        self.monitor_params = monitor_params
        self.potential_params = potential_params
        self.beta = get_beta(self.N_atoms, self.a)
        self.k_cut_params = k_cut_params
        self.precision = precision or default_precision()
        self.solve = None
        self._psi = None
        self.staticV = None
        self.propagator = None
        self.name = ""
        self.config = {}

    # -- wavefunction
    @property
    def psi(self):
        return self._psi

    @psi.setter
    def psi(self, value):
        if value is None or self.solve is None:
            self._psi = value
            return
        if isinstance(value, DeviceField):
            if value.solve is not self.solve:
                self.solve.copy_from(value.solve)
        else:
            self.solve.upload(np.asarray(value))
        self._psi = DeviceField(self.solve)

    # -- lifecycle
    def prepare(self, previous_psi=None):
        """Create the device solve and everything its propagator reads."""
        self.solve = None
        kind = self.type
        statics = self.potential_params["static"]
        dynamics = self.potential_params["dynamic"]
        static_tree = tree_from_c(statics["expression"]) if statics["activate"] else None
        dynamic_tree = (tree_from_c(dynamics["expression"], time_dependent=True)
                        if dynamics["activate"] and kind == "real" else None)
        kcut = self.k_cut_params["k_cut"] if self.k_cut_params["activate"] else None
        solve = make_solve(self.N_grids, self.x_array, self.k_array, kind, self.dt, self.beta,
                           static=static_tree, dynamic=dynamic_tree, kcut=kcut, precision=self.precision)
        if self.box_params["activate"]:
            add_box(solve, self.box_params)
        if kind == "real" and self.absorber_params["activate"]:
            add_absorber(solve, self.absorber_params, self.simulation_range)
        if self.noise_params["activate"]:
            add_white_noise(solve, self.N_grids, self.box_params["trap_depth"], self.noise_params["fraction"],
                            self.noise_params["seed"])
        self.solve = solve
        self._psi = None
        self.psi = self._get_initial_state(previous_psi)
        self.staticV = StaticVolume(solve, kind)
        self.propagator = make_real_propagator(self) if kind == "real" else make_imag_propagator(self)

    def release(self):
        """Drop the stage's arrays and propagator; a returned psi stays valid."""
        self._psi = None
        self.staticV = None
        self.propagator = None
        self.solve = None

    def _get_initial_state(self, previous_psi):
        kind = self.initial_state["type"]
        if kind == "continue":
            if previous_psi is None:
                raise ValueError("A continue stage needs the preceding stage's wavefunction.")
            return previous_psi
        if kind == "load":
            return load_wavefunction(self.initial_state["path"], self)
        if kind == "gaussian":
            self.solve.gaussian(self.initial_state["sigma"])
            return DeviceField(self.solve)
        raise ValueError(f"Unknown initial state: {kind}")
