"""Experiment YAML -> ordered stage objects (drop-in for gpe/yaml_parser.py:17-71).

Behaviour the reference's callers rely on:

* every parse gets one UTC run stamp (``%Y%m%dT%H%M%S%fZ``); each stage
  writes its data under ``<data_root>/<stamp>/<folder>`` and its monitor
  files under ``<monitor_root>/<stamp>/<folder>``, with relative roots taken
  from the project root;
* a stage's folder is its ``save.directory``, else ``stage_NN`` plus the
  stage name reduced to ``[A-Za-z0-9_-]`` runs; an absolute directory holds
  the data itself (``<dir>/<stamp>``) while the monitor files keep the default
  folder under the monitor root;
* a relative directory may not leave the run folder, and no two stages may
  share a data or monitor folder (ValueError, the reference's messages);
* ``load`` initial states resolve relative paths against the YAML file;
* expressions become the reference's C text (disabled ones read ``0``);
  the YAML groups map onto the stage keywords (``enabled`` -> ``activate``,
  ``fraction_of_trap_depth`` -> ``fraction``, ``k_cut.value`` -> ``k_cut``).

Stages share one pair of axes and allocate nothing until prepared.  The field
precision comes from `precision`, a schema-3 ``simulation.precision``, or
GPE_HIP_PRECISION.
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

_ESCAPE_ERROR = ("Relative save.directory must stay inside the run folder; "
                 "use an absolute path for another location.")


def _project_path(value):
    folder = Path(value).expanduser()
    return folder if folder.is_absolute() else PROJECT_ROOT / folder


def _default_folder(index, name):
    slug = re.sub(r"[^A-Za-z0-9_-]+", "_", name.strip()).strip("_")
    return f"stage_{index:02d}_{slug}" if slug else f"stage_{index:02d}"


class _RunFolders:
    """Output folders of one parse: the run stamp, both roots and the folders
    already given to earlier stages."""

    def __init__(self, simulation):
        self.stamp = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%S%fZ")
        self.data_run = _project_path(simulation["data_root"]) / self.stamp
        self.monitor_run = _project_path(simulation["monitor_root"]) / self.stamp
        self.taken_data, self.taken_monitor = set(), set()

    def assign(self, index, name, directory):
        """(data_dir, monitor_dir) of stage `index`, checked and recorded."""
        default = _default_folder(index, name)
        folder = Path(directory or default).expanduser()
        if folder.is_absolute():
            data_dir, monitor_dir = folder / self.stamp, self.monitor_run / default
        else:
            data_dir, monitor_dir = self.data_run / folder, self.monitor_run / folder
            inside = (data_dir.resolve().is_relative_to(self.data_run.resolve())
                      and monitor_dir.resolve().is_relative_to(self.monitor_run.resolve()))
            if not inside:
                raise ValueError(_ESCAPE_ERROR)
        data_key, monitor_key = data_dir.resolve(), monitor_dir.resolve()
        if data_key in self.taken_data or monitor_key in self.taken_monitor:
            raise ValueError(f"Stage {index} reuses another stage's output directory; "
                             "choose a distinct save.directory.")
        self.taken_data.add(data_key)
        self.taken_monitor.add(monitor_key)
        return data_dir, monitor_dir


def _switch_group(group, **renames):
    """A YAML on/off group as the stage expects it: `enabled` -> `activate`,
    plus any other key renames."""
    out = {("activate" if key == "enabled" else renames.get(key, key)): value for key, value in group.items()}
    return out


def _potentials(groups):
    out = {}
    for kind, group in groups.items():
        source = group["expression"] if group["enabled"] else "0"
        out[kind] = {"activate": group["enabled"],
                     "expression": cuda_expression(source, time_dependent=kind == "dynamic")}
    return out


def _initial_state(settings, yaml_dir):
    initial = copy.deepcopy(settings)
    if initial["type"] == "load":
        source = Path(initial["path"]).expanduser()
        if not source.is_absolute():
            source = yaml_dir / source
        initial["path"] = str(source)
    return initial


def parse_yaml(path, precision=None):
    """The stages of the experiment file at `path`, in order."""
    path = Path(path).expanduser().resolve()
    config = load_yaml(path.read_text())
    simulation = config["simulation"]
    precision = precision or simulation.get("precision") or default_precision()
    folders = _RunFolders(simulation)
    n, length = simulation["N_grids"], simulation["simulation_range"]
    x, k = get_grids(n, length)
    stages = []
    for index, settings in enumerate(config["stages"], 1):
        item = stage(simulation_range=length, x_array=x, k_array=k, N_grids=n, **settings["evolution"],
                     initial_state=_initial_state(settings["initial_state"], path.parent),
                     box_params=_switch_group(settings["box"]),
                     absorber_params=_switch_group(settings["absorber"]),
                     noise_params=_switch_group(settings["noise"], fraction_of_trap_depth="fraction"),
                     cloud_params=settings["cloud"], monitor_params=settings["monitor"],
                     save_params=settings["save"], potential_params=_potentials(settings["potential"]),
                     k_cut_params={"activate": settings["k_cut"]["enabled"], "k_cut": settings["k_cut"]["value"]},
                     precision=precision)
        item.name = settings["name"].strip() or f"Stage {index}"
        item.config = copy.deepcopy(settings)
        item.data_dir, item.monitor_dir = folders.assign(index, settings["name"], settings["save"]["directory"])
        stages.append(item)
    return stages
