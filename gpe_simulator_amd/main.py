"""Run an experiment: python -m gpe_simulator_amd.main experiment.yaml

Command-line drop-in for ``python -m gpe.main`` (gpe/main.py:7-38): the same
positional YAML path and ``--check`` (validate only, print the stage count),
the same progress lines, and the same rule that only a ``continue`` stage
starts from the previous stage's final field; ``--precision`` adds the choice
of field precision (complex128, the reference's, or complex64 with FP64
reductions and time)."""

import argparse
from pathlib import Path


def _parser():
    parser = argparse.ArgumentParser(description="Run the stages in a GPE experiment YAML file.")
    parser.add_argument("config", help="Path to the experiment YAML file")
    parser.add_argument("--check", action="store_true",
                        help="Validate YAML settings without loading the engine or running a simulation")
    parser.add_argument("--precision", choices=("complex128", "complex64"), default=None,
                        help="field precision (default: schema-3 simulation.precision, "
                             "$GPE_HIP_PRECISION, else complex128)")
    return parser


def run_stages(stages):
    """Prepare, propagate and release each stage in turn, handing the final
    field to the next stage only when that one continues from it."""
    from .launch import launch_stage
    carried = None
    for current in stages:
        print(f"Starting {current.name} ({current.type}, {current.steps} steps)", flush=True)
        start = carried if current.initial_state["type"] == "continue" else None
        carried = None              # the stage now owns the field it starts from
        current.prepare(start)
        del start
        carried = launch_stage(current)
        current.release()
        print(f"Finished {current.name}", flush=True)
    return carried


def main(argv=None):
    parser = _parser()
    args = parser.parse_args(argv)
    from .configuration import ConfigError, load_yaml
    try:
        config = load_yaml(Path(args.config).expanduser().read_text())
    except (OSError, ConfigError) as exc:
        parser.error(str(exc))
    if args.check:
        print(f"Valid configuration: {len(config['stages'])} stages. No simulation was run.")
        return
    from .yaml_parser import parse_yaml
    run_stages(parse_yaml(args.config, precision=args.precision))


if __name__ == "__main__":
    main()
