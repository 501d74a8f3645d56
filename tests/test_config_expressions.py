"""Schema-2 configuration and expression front end (CPU).

Mirrors the behaviour the reference's config_editor tests pin
(config_editor/tests/test_configuration.py, test_expressions.py) and the C
rendering golden strings of tests/test_workflow.py:25-32.
"""

import math

import numpy as np
import pytest

from gpe_simulator_amd.configuration import (ConfigError, dump_yaml, load_yaml, new_config, new_stage,
                                             normalize_config)
from gpe_simulator_amd.expressions import (TGT_VOL, TGT_XY, TGT_Z, additive_terms, compile_program,
                                           cuda_expression, evaluate, is_zero, join_terms, parse_expression,
                                           plan_potential, tree_from_c)


# ------------------------------------------------------------ expressions --
def test_cuda_expression_golden_strings():
    assert cuda_expression("1/2*y**2") == "((1.0 / 2.0) * pow(y, 2.0))"
    assert cuda_expression("abs(-x) + cos(t)", True) == "(fabs((-x)) + cos(t))"
    assert cuda_expression("pi") == repr(np.pi)
    with pytest.raises(ValueError):
        cuda_expression("t*y")
    with pytest.raises(ValueError):
        cuda_expression("x; return 0")


@pytest.mark.parametrize("formula,message", [
    ("A*y", "Unrecognised variable A"), ("sinn(x)", "Unrecognised function sinn"),
    ("sin(x,y)", "one positional argument"), ("cos", "Use cos(...)"), ("x^2", "Use ** for powers"),
    ("np.sin(x)", "without np. or cp.")])
def test_grammar_messages(formula, message):
    with pytest.raises(ValueError) as info:
        parse_expression(formula, True)
    assert message in str(info.value)


@pytest.mark.parametrize("formula", ["__import__('os').system('echo unsafe')", "x.__class__", "[x for x in y]",
                                     "(lambda: 1)()", "x[0]", "True", "1j", "1e999", "sin(x=1)",
                                     "x if y else z", "", " ", "sin(", "x y", "x=2"])
def test_grammar_rejects_non_arithmetic(formula):
    with pytest.raises(ValueError):
        parse_expression(formula, True)


def test_zero_detection():
    for formula in ("0", "0.0", "(-0.0)", "+0", "0e10"):
        assert is_zero(parse_expression(formula))
    for formula in ("0.03*x", "x-x", "sin(0)"):
        assert not is_zero(parse_expression(formula))


@pytest.mark.parametrize("source", ["x/2 + y**2/10 + abs(z)/5", "sin(t)*x + cos(pi)*z",
                                    "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)",
                                    "0.002*(x*x + y*y + z*z)", "-(x - 2*y)/3 + z*(1+x)", "(x+y)*(2+cos(pi))",
                                    "sqrt(x**2+y**2+1)*z/10 - tanh(z) + log(2+x*x)"])
def test_c_text_round_trip_and_additive_split(source):
    tree = parse_expression(source, time_dependent=True)
    back = tree_from_c(cuda_expression(source, True), time_dependent=True)
    x = np.linspace(-3, 3, 7)
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    full = evaluate(tree, X, Y, Z, 0.7)
    np.testing.assert_allclose(evaluate(back, X, Y, Z, 0.7), full, rtol=1e-14, atol=1e-14)
    split = evaluate(join_terms(additive_terms(tree)), X, Y, Z, 0.7)
    np.testing.assert_allclose(split, full, rtol=1e-13, atol=1e-13)


def run_program(program, x, y, z, t=0.0):
    """Host interpreter of the device postfix format (test-only)."""
    code, consts = program
    funcs = {20: math.sin, 21: math.cos, 22: math.tan, 23: math.exp, 24: math.sqrt, 25: math.tanh,
             26: abs, 27: math.log}
    stack, pc = [], 0
    while pc < len(code):
        op = code[pc]
        if op == 1:
            pc += 1
            stack.append(consts[code[pc]])
        elif op in (2, 3, 4, 5):
            stack.append({2: x, 3: y, 4: z, 5: t}[op])
        elif op in (10, 11, 12, 13, 14, 16):
            b, a = stack.pop(), stack.pop()
            stack.append({10: a + b, 11: a - b, 12: a * b, 13: a / b if op == 13 else 0,
                          14: a ** b if op == 14 else 0, 16: max(a, b)}[op])
        elif op == 15:
            stack.append(-stack.pop())
        else:
            stack.append(funcs[op](stack.pop()))
        pc += 1
    assert len(stack) == 1
    return stack[0]


def test_plan_groups_terms_by_coordinates():
    plan = plan_potential("x**2 + y**2 + 16*z**2", "10*exp(-((x-4*cos(t))**2+(y-4*sin(t))**2)/2)")
    assert set(plan.static) == {TGT_XY, TGT_Z} and set(plan.dynamic) == {TGT_XY}
    plan = plan_potential("sqrt(x**2 + y**2 + 1) * z / 10", "cos(t)*x*z/20 + t/4 + sin(t)*z")
    assert set(plan.static) == {TGT_VOL} and set(plan.dynamic) == {TGT_XY, TGT_Z, TGT_VOL}
    plan = plan_potential(None, "cos(pi)*z + 0*t")
    assert TGT_Z in plan.static


@pytest.mark.parametrize("static,dynamic", [
    ("x/2 + y**2/10 + abs(z)/5", "sin(t)*x + cos(pi)*z"),
    ("0.002*(x*x + y*y + z*z)", "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)"),
    ("sqrt(x**2 + y**2 + 1) * z / 10", "cos(t)*x*z/20"),
    ("x**2 + y**2 + 16*z**2", None)])
def test_planned_programs_sum_to_the_expression(static, dynamic):
    """Static + per-step programs evaluated by a host interpreter of the device
    format reproduce the full potential at every point."""
    plan = plan_potential(static, dynamic)
    t = 0.41
    full_static = parse_expression(static)
    full_dynamic = parse_expression(dynamic, True) if dynamic else None
    for x in (-1.7, 0.0, 2.3):
        for y in (-0.4, 1.9):
            for z in (-2.2, 0.6):
                want = evaluate(full_static, x, y, z)
                if full_dynamic is not None:
                    want += evaluate(full_dynamic, x, y, z, t)
                got = sum(run_program(p, x, y, z) for p in plan.static.values())
                got += sum(run_program(p, x, y, z, t) for p in plan.dynamic.values())
                assert got == pytest.approx(want, rel=1e-13, abs=1e-13)


def test_deep_expression_is_rejected_for_device():
    deep = "+".join(["(x" for _ in range(40)]) + ")" * 40
    with pytest.raises(ValueError):
        tree = parse_expression(deep.replace("+", "*"))
        compile_program(tree)


# ------------------------------------------------------------ configuration --
def test_roundtrip_preserves_custom_settings():
    config = new_config()
    real = new_stage()
    real["name"] = "No"
    real["potential"]["static"].update(enabled=True, expression="0.03*x")
    real["potential"]["dynamic"].update(enabled=True, expression="0.125*cos(0.065*t)*y**2")
    real["noise"].update(enabled=True, seed=12345, fraction_of_trap_depth=0.01)
    real["save"].update(save_initial=True, psi_interval=10000, directory="results with spaces")
    config["stages"].append(real)
    text = dump_yaml(config)
    assert load_yaml(text) == config
    assert "precision:" not in text


def test_first_stage_cannot_continue():
    config = new_config()
    config["stages"][0]["initial_state"]["type"] = "continue"
    with pytest.raises(ConfigError) as info:
        normalize_config(config)
    assert info.value.errors[0]["path"] == "stages.0.initial_state.type"


@pytest.mark.parametrize("mutate,message", [
    (lambda c: c["stages"][0]["potential"]["dynamic"].update(enabled=True, expression="cos(t)*y"),
     "Imaginary time evolution does not accept a time dependent potential"),
    (lambda c: (c["stages"][0]["box"].update(enabled=False), c["stages"][0]["noise"].update(enabled=True)),
     "Trap-dependent noise needs"),
    (lambda c: c["stages"][0]["noise"].update(enabled=True, correlation_length=1),
     "Correlated disorder is not implemented"),
    (lambda c: c["stages"][0]["k_cut"].update(enabled=True, value=2), "only for real time"),
])
def test_physics_rules(mutate, message):
    config = new_config()
    mutate(config)
    with pytest.raises(ConfigError, match=message):
        normalize_config(config)


def test_continued_state_keeps_reference_atom_number():
    config = new_config()
    second = new_stage()
    second["cloud"]["N_atom"] += 1
    config["stages"].append(second)
    with pytest.raises(ConfigError, match="reference N_atom"):
        normalize_config(config)
    second["initial_state"]["type"] = "gaussian"
    normalize_config(config)


def test_invalid_numbers_report_every_path():
    config = new_config()
    config["stages"][0]["initial_state"]["type"] = "load"
    config["stages"][0]["evolution"]["steps"] = 1.5
    config["simulation"]["N_grids"] = True
    config["stages"][0]["evolution"]["dt"] = float("nan")
    with pytest.raises(ConfigError) as info:
        normalize_config(config)
    paths = {item["path"] for item in info.value.errors}
    assert {"stages.0.initial_state.path", "stages.0.evolution.steps", "simulation.N_grids",
            "stages.0.evolution.dt"} <= paths


def test_zero_disables_and_disabled_clears():
    config = new_config()
    config["stages"][0]["potential"]["static"].update(enabled=True, expression="(-0.0)")
    assert not normalize_config(config)["stages"][0]["potential"]["static"]["enabled"]
    config["stages"][0]["potential"]["static"].update(enabled=False, expression="0.03*y")
    assert load_yaml(dump_yaml(config))["stages"][0]["potential"]["static"] == {"enabled": False,
                                                                              "expression": "0"}


def test_legacy_output_root_and_checkpoint():
    result = load_yaml("schema_version: 2\nsimulation: {output_root: results}\n"
                       "stages: [{save: {checkpoint_interval: 50}}]")
    assert result["simulation"]["data_root"] == "results/data"
    assert result["stages"][0]["save"]["psi_interval"] == 50
    with pytest.raises(ConfigError, match="Choose one wavefunction snapshot interval"):
        load_yaml("schema_version: 2\nstages: [{save: {checkpoint_interval: 50, psi_interval: 20}}]")


@pytest.mark.parametrize("source", ["!!python/object/apply:os.system ['echo should-not-run']",
                                    "schema_version: 1\nschema_version: 2\n",
                                    "schema_version: 2\nstages: [{}]\nsimulation: {precision: single}\n",
                                    "schema_version: 1\nstages: [{}]",
                                    "schema_version: 2\nstages: [{noise: {region: inside_box}}]"])
def test_unsafe_duplicate_unknown_rejected(source):
    with pytest.raises(ConfigError):
        load_yaml(source)


def test_schema3_adds_precision_and_nk_view_only():
    result = load_yaml("schema_version: 3\nsimulation: {precision: complex64}\nstages: [{}]")
    assert result["simulation"]["precision"] == "complex64"
    with pytest.raises(ConfigError):
        load_yaml("schema_version: 3\nsimulation: {precision: half}\nstages: [{}]")
    default = load_yaml("schema_version: 3\nstages: [{}]")
    assert default["simulation"]["precision"] == "complex128"
    assert default["stages"][0]["save"]["nk_view"] is False
    assert load_yaml(dump_yaml(result)) == result
    view = load_yaml("schema_version: 3\nstages: [{save: {nk_view: true, nk_interval: 5}}]")
    assert view["stages"][0]["save"]["nk_view"] is True
    assert load_yaml(dump_yaml(view)) == view
    with pytest.raises(ConfigError, match="nk_view must be true or false"):
        load_yaml("schema_version: 3\nstages: [{save: {nk_view: 1}}]")
    # schema 2 keeps the reference's keys only
    assert "nk_view" not in load_yaml("schema_version: 2\nstages: [{}]")["stages"][0]["save"]
    with pytest.raises(ConfigError, match="Unknown setting: stages.0.save.nk_view"):
        load_yaml("schema_version: 2\nstages: [{save: {nk_view: true}}]")
