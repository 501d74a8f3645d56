"""Run an experiment: python -m gpe_simulator_amd.main experiment.yaml

Drop-in for `python -m gpe.main` (gpe/main.py:7-38): same arguments and
stage loop, plus --precision to pick the field precision (complex128, the
reference's, or complex64 with FP64 reductions and time)."""

import argparse
from pathlib import Path


def main(argv=None):
    parser = argparse.ArgumentParser(description="Run the stages in a GPE experiment YAML file.")
    parser.add_argument("config", help="Path to the experiment YAML file")
    parser.add_argument("--check", action="store_true",
                        help="Validate YAML settings without loading the engine or running a simulation")
    parser.add_argument("--precision", choices=("complex128", "complex64"), default=None,
                        help="field precision (default: schema-3 simulation.precision, "
                             "$GPE_HIP_PRECISION, else complex128)")
    args = parser.parse_args(argv)
    from .configuration import ConfigError, load_yaml
    try:
        config = load_yaml(Path(args.config).expanduser().read_text())
    except (OSError, ConfigError) as exc:
        parser.error(str(exc))
    if args.check:
        print(f"Valid configuration: {len(config['stages'])} stages. No simulation was run.")
        return
    from .launch import launch_stage
    from .yaml_parser import parse_yaml
    stages = parse_yaml(args.config, precision=args.precision)
    psi = None
    for current in stages:
        print(f"Starting {current.name} ({current.type}, {current.steps} steps)", flush=True)
        if current.initial_state["type"] != "continue":
            psi = None
        current.prepare(psi)
        psi = None
        psi = launch_stage(current)
        current.release()
        print(f"Finished {current.name}", flush=True)


if __name__ == "__main__":
    main()
