"""Experiment YAML -> ordered stage objects (gpe/yaml_parser.py:17-71).

Same folder layout and rules as the reference: one UTC run stamp per parse,
relative output roots resolved against the project root, duplicate or escaping
stage folders rejected, axes shared by all stages and no device allocation
until a stage is prepared.  The field precision comes from `precision`
(argument), `simulation.precision` in a schema-3 file, or GPE_HIP_PRECISION.
"""

from __future__ import annotations

import copy
import re
from datetime import datetime, timezone
from pathlib import Path

from .configuration import load_yaml
from .expressions import cuda_expression
from .grids import get_grids
from .stage import default_precision, stage

PROJECT_ROOT = Path(__file__).resolve().parents[1]


def parse_yaml(path, precision=None):
    path = Path(path).expanduser().resolve()
    config = load_yaml(path.read_text())
    simulation = config["simulation"]
    precision = precision or simulation.get("precision") or default_precision()
    run_id = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%S%fZ")

    def output_root(value):
        folder = Path(value).expanduser()
        return folder if folder.is_absolute() else PROJECT_ROOT / folder

    data_root = output_root(simulation["data_root"])
    monitor_root = output_root(simulation["monitor_root"])
    x, k = get_grids(simulation["N_grids"], simulation["simulation_range"])
    stages, data_dirs, monitor_dirs = [], set(), set()
    for index, settings in enumerate(config["stages"], 1):
        initial = copy.deepcopy(settings["initial_state"])
        if initial["type"] == "load":
            source = Path(initial["path"]).expanduser()
            initial["path"] = str(source if source.is_absolute() else path.parent / source)

        def active(group):
            values = dict(settings[group])
            values["activate"] = values.pop("enabled")
            return values

        potential = {kind: dict(activate=values["enabled"],
                                expression=cuda_expression(values["expression"] if values["enabled"] else "0",
                                                           time_dependent=(kind == "dynamic")))
                     for kind, values in settings["potential"].items()}
        noise = active("noise")
        noise["fraction"] = noise.pop("fraction_of_trap_depth")
        item = stage(simulation_range=simulation["simulation_range"], x_array=x, k_array=k,
                     N_grids=simulation["N_grids"], **settings["evolution"], initial_state=initial,
                     box_params=active("box"), absorber_params=active("absorber"), noise_params=noise,
                     cloud_params=settings["cloud"], monitor_params=settings["monitor"],
                     save_params=settings["save"], potential_params=potential,
                     k_cut_params=dict(activate=settings["k_cut"]["enabled"], k_cut=settings["k_cut"]["value"]),
                     precision=precision)
        item.name = settings["name"].strip() or f"Stage {index}"
        item.config = copy.deepcopy(settings)
        slug = re.sub(r"[^A-Za-z0-9_-]+", "_", settings["name"].strip()).strip("_")
        default_folder = f"stage_{index:02d}" + (f"_{slug}" if slug else "")
        folder = Path(settings["save"]["directory"] or default_folder).expanduser()
        item.data_dir = folder / run_id if folder.is_absolute() else data_root / run_id / folder
        item.monitor_dir = monitor_root / run_id / (default_folder if folder.is_absolute() else folder)
        if not folder.is_absolute():
            if not (item.data_dir.resolve().is_relative_to((data_root / run_id).resolve()) and
                    item.monitor_dir.resolve().is_relative_to((monitor_root / run_id).resolve())):
                raise ValueError("Relative save.directory must stay inside the run folder; "
                                 "use an absolute path for another location.")
        if item.data_dir.resolve() in data_dirs or item.monitor_dir.resolve() in monitor_dirs:
            raise ValueError(f"Stage {index} reuses another stage's output directory; "
                             "choose a distinct save.directory.")
        data_dirs.add(item.data_dir.resolve())
        monitor_dirs.add(item.monitor_dir.resolve())
        stages.append(item)
    return stages
