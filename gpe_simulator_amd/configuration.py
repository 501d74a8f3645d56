"""Experiment configuration: the reference's schema-2 YAML, unchanged.

Defaults, validation rules and error messages follow
config_editor/configuration.py:41-266 of the reference so that a schema-2 file
that the reference accepts is accepted here with the same meaning, and one it
rejects is rejected with the same message.  Defaults come from the reference's
gpe/config.py:1-15 (N_grids 256, range 70, R 25, L 50, wall 2, U_D 49,
r_abs 5, A_abs 100, dt 0.01, N_atom 150000, a 100).

Schema 2 rejects unknown keys, so engine options (precision) are not YAML keys
here: they are chosen with the CLI flag / environment variable documented in
main.py, or with `schema_version: 3`, which is schema 2 plus an optional
`simulation.precision` ("complex128" default, or "complex64") and an optional
per-stage `save.nk_view` (false by default): when true, each nk output is
accompanied by the 128-bin equal-width view of the same spectrum
(nkview_<step>.npz).
"""

from __future__ import annotations

import copy
import math
from pathlib import Path

import yaml

from .expressions import is_zero, parse_expression

MONITOR_FIELDS = (
    "print_scalar_interval", "show_nk_interval", "show_TOF_image_interval",
    "show_verticle_TOF_interval", "show_TOF_slice_interval",
    "show_verticle_TOF_slice_interval", "show_magnifier_interval",
    "show_magnifier_slice_interval", "show_phase_interval",
)
SAVE_INTERVALS = ("scalar_interval", "nk_interval", "psi_interval")
PRECISIONS = ("complex128", "complex64")

PROJECT_DEFAULTS = dict(N_grids=256, simulation_range=70, R=25, L=50, U_D=49, wallthickness=2,
                        r_abs=5, A_abs=100, a=100, N_atom=150000, dt=0.01)


def project_defaults():
    return dict(PROJECT_DEFAULTS)


def new_stage(imaginary=False):
    cfg = project_defaults()
    monitor = {name: 0 for name in MONITOR_FIELDS}
    monitor["print_scalar_interval"] = 1000
    return {
        "name": "",
        "evolution": {"type": "imaginary" if imaginary else "real", "steps": 2000, "dt": cfg["dt"]},
        "initial_state": {"type": "gaussian" if imaginary else "continue", "sigma": 3.0, "path": ""},
        "potential": {"static": {"enabled": False, "expression": "0"},
                      "dynamic": {"enabled": False, "expression": "0"}},
        "k_cut": {"enabled": False, "value": 0.0},
        "box": {"enabled": True, "R": cfg["R"], "L": cfg["L"], "trap_depth": cfg["U_D"],
                "wall_thickness": cfg["wallthickness"]},
        "absorber": {"enabled": not imaginary, "activation_distance": cfg["r_abs"], "strength": cfg["A_abs"]},
        "noise": {"enabled": False, "fraction_of_trap_depth": 0.01, "seed": 42, "correlation_length": 0.0},
        "cloud": {"a": cfg["a"], "N_atom": cfg["N_atom"]},
        "monitor": monitor,
        "save": {"directory": "", "scalar_interval": 1000, "nk_interval": 0, "psi_interval": 0,
                 "save_initial": False, "save_final": True},
    }


def new_config():
    cfg = project_defaults()
    first = new_stage(imaginary=True)
    first["name"] = "Preparation"
    return {"schema_version": 2,
            "simulation": {"simulation_range": cfg["simulation_range"], "N_grids": cfg["N_grids"],
                           "data_root": "data", "monitor_root": "monitor"},
            "stages": [first]}


class ConfigError(ValueError):
    """Validation failure; `errors` lists {path, message} items."""

    def __init__(self, errors):
        self.errors = errors
        super().__init__("; ".join(item["message"] for item in errors))


class _Validator:
    def __init__(self):
        self.errors = []

    def error(self, path, message):
        self.errors.append({"path": path, "message": message})

    def merge(self, value, default, path):
        """Overlay `value` on `default`, reporting keys the default lacks."""
        if not isinstance(value, dict):
            self.error(path, f"{path or 'Configuration'} must be a mapping.")
            return copy.deepcopy(default)
        merged = copy.deepcopy(default)
        for key, item in value.items():
            child = f"{path}.{key}" if path else str(key)
            if key not in default:
                self.error(child, f"Unknown setting: {child}.")
            elif isinstance(default[key], dict):
                merged[key] = self.merge(item, default[key], child)
            else:
                merged[key] = copy.deepcopy(item)
        return merged

    def number(self, box, key, path, minimum=None, positive=False, integer=False):
        value = box[key]
        ok = type(value) in (int, float) and math.isfinite(value)
        if integer:
            ok = ok and type(value) is int and abs(value) <= 2 ** 53 - 1
        if not ok or (positive and value <= 0) or (minimum is not None and value < minimum):
            qualifier = "positive " if positive else "non-negative " if minimum == 0 else ""
            kind = "integer" if integer else "finite number"
            tail = f" ≥ {minimum}." if minimum not in (None, 0) else "."
            self.error(f"{path}.{key}", f"{path}.{key} must be a {qualifier}{kind}{tail}")

    def text(self, box, key, path, required=False):
        value = box[key]
        if not isinstance(value, str) or (required and not value.strip()):
            self.error(f"{path}.{key}", f"{path}.{key} must be {'non-empty ' if required else ''}text.")

    def choice(self, box, key, path, options):
        if box[key] not in options:
            self.error(f"{path}.{key}", f"{path}.{key} must be one of: {', '.join(options)}.")

    def boolean(self, box, key, path):
        if type(box[key]) is not bool:
            self.error(f"{path}.{key}", f"{path}.{key} must be true or false.")


def normalize_config(raw):
    """Fill optional settings, reject unknown fields, and validate physical inputs.

    Follows config_editor/configuration.py:83-228."""
    if not isinstance(raw, dict):
        raise ConfigError([{"path": "", "message": "Configuration must be a YAML mapping."}])
    check = _Validator()
    raw = copy.deepcopy(raw)
    simulation = raw.get("simulation", {})
    if isinstance(simulation, dict) and "output_root" in simulation:
        root = simulation.pop("output_root")
        if not isinstance(root, str) or not root.strip():
            raise ConfigError([{"path": "simulation.output_root", "message": "Output root must be non-empty text."}])
        simulation.setdefault("data_root", str(Path(root) / "data"))
        simulation.setdefault("monitor_root", str(Path(root) / "monitor"))
    schema3 = raw.get("schema_version") == 3 and type(raw.get("schema_version")) is int
    base = new_config()
    if schema3:
        base["schema_version"] = 3
        base["simulation"]["precision"] = "complex128"
    config = check.merge(raw, base, "")
    version = config["schema_version"]
    if type(version) is not int or version not in ((2, 3) if schema3 else (2,)):
        check.error("schema_version", "Use schema_version: 2 with separate static and dynamic potential expressions.")

    sim = config["simulation"]
    check.number(sim, "N_grids", "simulation", minimum=2, integer=True)
    check.number(sim, "simulation_range", "simulation", positive=True)
    check.text(sim, "data_root", "simulation", required=True)
    check.text(sim, "monitor_root", "simulation", required=True)
    if schema3:
        check.choice(sim, "precision", "simulation", PRECISIONS)
    stages = raw.get("stages")
    if not isinstance(stages, list) or not stages:
        check.error("stages", "Add at least one stage.")
        stages = []
    config["stages"] = []
    for index, raw_stage in enumerate(stages):
        path = f"stages.{index}"
        if isinstance(raw_stage, dict) and isinstance(raw_stage.get("save"), dict):
            old = raw_stage["save"]
            legacy = old.pop("checkpoint_interval", 0)
            if legacy:
                if old.get("psi_interval", 0) not in (0, legacy):
                    check.error(f"{path}.save.psi_interval",
                                "Choose one wavefunction snapshot interval; restart checkpoints are not supported.")
                else:
                    old["psi_interval"] = legacy
        default = new_stage(imaginary=index == 0)
        if schema3:
            default["save"]["nk_view"] = False
        stage = check.merge(raw_stage, default, path)
        config["stages"].append(stage)
        _validate_stage(check, config, stage, index, path)
    if check.errors:
        raise ConfigError(check.errors)
    return config


def _validate_stage(check, config, stage, index, path):
    check.text(stage, "name", path)
    evo = stage["evolution"]
    check.choice(evo, "type", f"{path}.evolution", ("imaginary", "real"))
    check.number(evo, "steps", f"{path}.evolution", positive=True, integer=True)
    check.number(evo, "dt", f"{path}.evolution", positive=True)
    initial = stage["initial_state"]
    check.choice(initial, "type", f"{path}.initial_state", ("gaussian", "load", "continue"))
    check.text(initial, "path", f"{path}.initial_state", required=initial["type"] == "load")
    check.number(initial, "sigma", f"{path}.initial_state", positive=True)
    if index == 0 and initial["type"] == "continue":
        check.error(f"{path}.initial_state.type",
                    "Stage 1 needs a Gaussian or a loaded state; there is no previous stage.")
    for kind in ("static", "dynamic"):
        part = stage["potential"][kind]
        part_path = f"{path}.potential.{kind}"
        check.boolean(part, "enabled", part_path)
        if part["enabled"] is False:
            part["expression"] = "0"
        try:
            tree = parse_expression(part["expression"], time_dependent=kind == "dynamic")
            if is_zero(tree):
                part["enabled"] = False
                part["expression"] = "0"
        except ValueError as exc:
            check.error(f"{part_path}.expression", str(exc))
    imaginary = evo["type"] == "imaginary"
    if imaginary and stage["potential"]["dynamic"]["enabled"]:
        check.error(f"{path}.potential.dynamic.enabled",
                    "Imaginary time evolution does not accept a time dependent potential.")
    cutoff = stage["k_cut"]
    check.boolean(cutoff, "enabled", f"{path}.k_cut")
    check.number(cutoff, "value", f"{path}.k_cut", minimum=0, positive=cutoff["enabled"] is True)
    if imaginary and cutoff["enabled"]:
        check.error(f"{path}.k_cut.enabled", "The momentum cutoff is available only for real time evolution.")
    for group in ("box", "absorber", "noise"):
        check.boolean(stage[group], "enabled", f"{path}.{group}")
    if imaginary and stage["absorber"]["enabled"]:
        check.error(f"{path}.absorber.enabled", "Imaginary time evolution does not accept an absorber.")
    for key in ("R", "L", "wall_thickness"):
        check.number(stage["box"], key, f"{path}.box", positive=True)
    check.number(stage["box"], "trap_depth", f"{path}.box", minimum=0)
    for key in ("activation_distance", "strength"):
        check.number(stage["absorber"], key, f"{path}.absorber", minimum=0)
    for key in ("fraction_of_trap_depth", "correlation_length"):
        check.number(stage["noise"], key, f"{path}.noise", minimum=0)
    check.number(stage["noise"], "seed", f"{path}.noise", minimum=0, integer=True)
    if stage["noise"]["enabled"] and not stage["box"]["enabled"]:
        check.error(f"{path}.noise.enabled", "Trap-dependent noise needs an enabled box to define its depth.")
    if stage["noise"]["enabled"] and stage["noise"]["correlation_length"] != 0:
        check.error(f"{path}.noise.correlation_length",
                    "Correlated disorder is not implemented; use correlation_length: 0.")
    check.number(stage["cloud"], "a", f"{path}.cloud")
    check.number(stage["cloud"], "N_atom", f"{path}.cloud", positive=True, integer=True)
    if (index > 0 and initial["type"] == "continue"
            and stage["cloud"]["N_atom"] != config["stages"][index - 1]["cloud"]["N_atom"]):
        check.error(f"{path}.cloud.N_atom",
                    "A continued state must keep the previous stage's reference N_atom. "
                    "Change a to adjust the interaction.")
    for key in MONITOR_FIELDS:
        check.number(stage["monitor"], key, f"{path}.monitor", minimum=0, integer=True)
    for key in SAVE_INTERVALS:
        check.number(stage["save"], key, f"{path}.save", minimum=0, integer=True)
    for key in ("save_initial", "save_final", *(("nk_view",) if "nk_view" in stage["save"] else ())):
        check.boolean(stage["save"], key, f"{path}.save")
    check.text(stage["save"], "directory", f"{path}.save")


class _UniqueKeyLoader(yaml.SafeLoader):
    """Safe YAML loader that rejects duplicate and non-scalar mapping keys."""


def _mapping(loader, node, deep=False):
    loader.flatten_mapping(node)
    result = {}
    for key_node, value_node in node.value:
        key = loader.construct_object(key_node, deep=deep)
        if not isinstance(key, (str, int, float, bool, type(None))):
            raise ValueError("YAML mapping keys must be scalar values.")
        if key in result:
            raise ValueError(f"Duplicate YAML key: {key}")
        result[key] = loader.construct_object(value_node, deep=deep)
    return result


_UniqueKeyLoader.add_constructor(yaml.resolver.BaseResolver.DEFAULT_MAPPING_TAG, _mapping)


def load_yaml(source):
    """Parse and validate YAML text (config_editor/configuration.py:251-257)."""
    try:
        return normalize_config(yaml.load(source, Loader=_UniqueKeyLoader))
    except ConfigError:
        raise
    except (yaml.YAMLError, ValueError, RecursionError) as exc:
        raise ConfigError([{"path": "", "message": f"Cannot read YAML: {exc}"}]) from exc


def dump_yaml(config):
    """Validated YAML text with the schema header (config_editor/configuration.py:260-266)."""
    config = normalize_config(config)
    header = (f"# GPE experiment configuration — schema {config['schema_version']}\n"
              "# Double precision. Lengths and times are dimensionless; a is in Bohr radii.\n"
              "# Every stage starts at t = 0. Intervals count completed steps; 0 disables.\n"
              "# Gaussian disorder covers the whole grid. Potential expressions use Python syntax.\n")
    return header + yaml.safe_dump(config, sort_keys=False, allow_unicode=True, width=100)
