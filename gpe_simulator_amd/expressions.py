"""Potential expressions: the schema-2 grammar, C rendering and device programs.

The grammar is the one the reference's YAML accepts
(config_editor/expressions.py:8-58): numbers, x y z pi (and t for dynamic
potentials), + - * / **, unary +/-, and sin cos tan exp sqrt tanh abs log with
one argument.  `cuda_expression` renders the same C arithmetic text as the
reference (gpe/expressions.py:9-30); the engine itself does not splice source
text into kernels.  Instead an expression is split into additive terms, the
terms are grouped by the coordinates they read, and each group becomes a
postfix program that the device evaluates in FP64 into a table:

    XY  -- terms reading at most x, y (and t): one value per (x, y) column
    Z   -- terms reading z only (and t):        one value per z plane
    VOL -- terms mixing z with x or y:          one value per grid point

The split distributes factors that do not read x, y or z over sums
(c*(a+b) -> c*a + c*b, (a+b)/c -> a/c + b/c), so a harmonic trap becomes three
one-dimensional terms and a beam moving in the xy-plane becomes a per-step
(x, y) table.  Values agree with a direct evaluation to rounding.
"""

from __future__ import annotations

import ast
import math
from dataclasses import dataclass, field

FUNCTIONS = {"sin", "cos", "tan", "exp", "sqrt", "tanh", "abs", "log"}
BINARY = (ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow)

# Opcodes of the device interpreter (gpe_vm.h).
OP_CONST, OP_X, OP_Y, OP_Z, OP_T = 1, 2, 3, 4, 5
OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_NEG, OP_MAX = 10, 11, 12, 13, 14, 15, 16
OP_FUNC = {"sin": 20, "cos": 21, "tan": 22, "exp": 23, "sqrt": 24, "tanh": 25, "abs": 26, "log": 27}
OP_BIN = {ast.Add: OP_ADD, ast.Sub: OP_SUB, ast.Mult: OP_MUL, ast.Div: OP_DIV, ast.Pow: OP_POW}
VM_STACK = 32

TGT_XY, TGT_Z, TGT_VOL = 0, 1, 2


def parse_expression(source, time_dependent=False):
    """Validate `source` against the schema-2 grammar and return its AST body.

    Follows config_editor/expressions.py:12-58 (same rules and messages)."""
    if not isinstance(source, str) or not source.strip():
        raise ValueError("Enter a formula, or 0 to disable this potential.")
    if len(source) > 4096:
        raise ValueError("Keep the formula within 4096 characters.")
    try:
        tree = ast.parse(source.strip(), mode="eval").body
    except (SyntaxError, RecursionError) as exc:
        message = exc.msg if isinstance(exc, SyntaxError) else "expression is too deeply nested"
        raise ValueError(f"Syntax error: {message}.") from exc
    if sum(1 for _ in ast.walk(tree)) > 256:
        raise ValueError("The formula is too complex; use a shorter expression.")
    allowed = {"x", "y", "z", "pi"} | ({"t"} if time_dependent else set())
    _check(tree, allowed)
    return tree


def _check(node, allowed):
    if isinstance(node, ast.Constant):
        try:
            ok = type(node.value) in (int, float) and math.isfinite(node.value)
        except OverflowError:
            ok = False
        if not ok:
            raise ValueError("Use finite real numbers in the formula.")
        return
    if isinstance(node, ast.Name):
        if node.id in allowed:
            return
        if node.id in FUNCTIONS:
            raise ValueError(f"Use {node.id}(...) to call this function.")
        raise ValueError(f"Unrecognised variable {node.id}.")
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.UAdd, ast.USub)):
        _check(node.operand, allowed)
        return
    if isinstance(node, ast.BinOp) and isinstance(node.op, BINARY):
        _check(node.left, allowed)
        _check(node.right, allowed)
        return
    if isinstance(node, ast.Call):
        if not isinstance(node.func, ast.Name):
            raise ValueError("Use function names directly, such as sin(x), without np. or cp.")
        if node.func.id not in FUNCTIONS:
            raise ValueError(f"Unrecognised function {node.func.id}.")
        if len(node.args) != 1 or node.keywords:
            raise ValueError(f"{node.func.id} expects one positional argument.")
        _check(node.args[0], allowed)
        return
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.BitXor):
        raise ValueError("Use ** for powers, for example x**2, rather than ^.")
    raise ValueError("Use numbers, coordinates, + - * / **, parentheses and the listed functions.")


def is_zero(node):
    """True for a literal zero, possibly signed (config_editor/expressions.py:61-64)."""
    while isinstance(node, ast.UnaryOp):
        node = node.operand
    return isinstance(node, ast.Constant) and node.value == 0


def cuda_expression(source, time_dependent=False):
    """C arithmetic text of an expression, as gpe/expressions.py:9-30 renders it."""
    return _render_c(parse_expression(source, time_dependent=time_dependent))


def _render_c(node):
    if isinstance(node, ast.Constant):
        return repr(float(node.value))
    if isinstance(node, ast.Name):
        return repr(math.pi) if node.id == "pi" else node.id
    if isinstance(node, ast.UnaryOp):
        return f"({'-' if isinstance(node.op, ast.USub) else '+'}{_render_c(node.operand)})"
    if isinstance(node, ast.Call):
        name = "fabs" if node.func.id == "abs" else node.func.id
        return f"{name}({_render_c(node.args[0])})"
    left, right = _render_c(node.left), _render_c(node.right)
    if isinstance(node.op, ast.Pow):
        return f"pow({left}, {right})"
    symbol = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}[type(node.op)]
    return f"({left} {symbol} {right})"


def tree_from_c(text, time_dependent=False):
    """Parse the C text `cuda_expression` renders (the form the reference keeps
    in stage.potential_params, gpe/yaml_parser.py:44-46) back into a grammar
    tree: pow(a, b) -> a**b, fabs(a) -> abs(a)."""

    class Undo(ast.NodeTransformer):
        def visit_Call(self, node):
            self.generic_visit(node)
            if isinstance(node.func, ast.Name) and node.func.id == "pow" and len(node.args) == 2:
                return ast.BinOp(left=node.args[0], op=ast.Pow(), right=node.args[1])
            if isinstance(node.func, ast.Name) and node.func.id == "fabs":
                node.func = ast.Name(id="abs", ctx=ast.Load())
            return node

    try:
        tree = ast.parse(str(text).strip(), mode="eval").body
    except SyntaxError as exc:
        raise ValueError(f"Syntax error: {exc.msg}.") from exc
    tree = Undo().visit(tree)
    _check(tree, {"x", "y", "z", "pi"} | ({"t"} if time_dependent else set()))
    return tree


# ------------------------------------------------------------ analysis --
def variables(node):
    """Coordinates and time read by an expression tree (pi is a constant)."""
    names = set()
    for item in ast.walk(node):
        if isinstance(item, ast.Name) and item.id != "pi":
            names.add(item.id)
    return names


def _spatial(node):
    return variables(node) & {"x", "y", "z"}


def additive_terms(node, sign=1):
    """Split a tree into signed additive terms [(sign, term)], distributing
    factors that read no coordinate over sums."""
    if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub)):
        right_sign = sign if isinstance(node.op, ast.Add) else -sign
        return additive_terms(node.left, sign) + additive_terms(node.right, right_sign)
    if isinstance(node, ast.UnaryOp):
        return additive_terms(node.operand, sign if isinstance(node.op, ast.UAdd) else -sign)
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Mult):
        if not _spatial(node.left):
            inner = additive_terms(node.right)
            if len(inner) > 1:
                return [(sign * s, ast.BinOp(left=node.left, op=ast.Mult(), right=term)) for s, term in inner]
        if not _spatial(node.right):
            inner = additive_terms(node.left)
            if len(inner) > 1:
                return [(sign * s, ast.BinOp(left=term, op=ast.Mult(), right=node.right)) for s, term in inner]
    if isinstance(node, ast.BinOp) and isinstance(node.op, ast.Div) and not _spatial(node.right):
        inner = additive_terms(node.left)
        if len(inner) > 1:
            return [(sign * s, ast.BinOp(left=term, op=ast.Div(), right=node.right)) for s, term in inner]
    return [(sign, node)]


def term_target(term):
    space = _spatial(term)
    if "z" not in space:
        return TGT_XY
    if space == {"z"}:
        return TGT_Z
    return TGT_VOL


def join_terms(terms):
    """Rebuild one tree from signed terms."""
    tree = None
    for sign, term in terms:
        if tree is None:
            tree = term if sign > 0 else ast.UnaryOp(op=ast.USub(), operand=term)
        else:
            tree = ast.BinOp(left=tree, op=ast.Add() if sign > 0 else ast.Sub(), right=term)
    return tree if tree is not None else ast.Constant(value=0.0)


# ------------------------------------------------------------ programs --
def compile_program(node):
    """Postfix program (code, consts) for the device interpreter."""
    code, consts = [], []

    def const(value):
        consts.append(float(value))
        code.extend((OP_CONST, len(consts) - 1))

    def emit(item):
        if isinstance(item, ast.Constant):
            const(item.value)
        elif isinstance(item, ast.Name):
            if item.id == "pi":
                const(math.pi)
            else:
                code.append({"x": OP_X, "y": OP_Y, "z": OP_Z, "t": OP_T}[item.id])
        elif isinstance(item, ast.UnaryOp):
            emit(item.operand)
            if isinstance(item.op, ast.USub):
                code.append(OP_NEG)
        elif isinstance(item, ast.Call):
            emit(item.args[0])
            code.append(OP_FUNC[item.func.id])
        elif isinstance(item, ast.BinOp):
            emit(item.left)
            emit(item.right)
            code.append(OP_BIN[type(item.op)])
        elif isinstance(item, _Max):
            emit(item.left)
            emit(item.right)
            code.append(OP_MAX)
        else:  # pragma: no cover - guarded by parse_expression
            raise ValueError(f"cannot compile {ast.dump(item)}")

    emit(node)
    if _stack_depth(code) > VM_STACK:
        raise ValueError("The formula is too deeply nested for the device evaluator.")
    return code, consts


def _stack_depth(code):
    depth = peak = 0
    pc = 0
    while pc < len(code):
        op = code[pc]
        if op == OP_CONST:
            depth += 1
            pc += 1
        elif op in (OP_X, OP_Y, OP_Z, OP_T):
            depth += 1
        elif op in (OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_MAX):
            depth -= 1
        peak = max(peak, depth)
        pc += 1
    return peak


class _Max(ast.AST):
    """max(left, right); only built internally (box and absorber potentials)."""
    _fields = ("left", "right")


def vm_max(left, right):
    return _Max(left=left, right=right)


@dataclass
class PotentialPlan:
    """Per-target programs of a stage's potential."""
    static: dict = field(default_factory=dict)   # target -> (code, consts)
    dynamic: dict = field(default_factory=dict)  # target -> (code, consts)


def plan_potential(static_source=None, dynamic_source=None):
    """Group the static and time-dependent expressions into device programs.

    Terms that ignore t are static (evaluated once into an FP64 table); the
    t-dependent terms of a target form its per-step program, whose values the
    device adds to that target's static table at every half-step time."""
    groups = {TGT_XY: ([], []), TGT_Z: ([], []), TGT_VOL: ([], [])}
    for source, dynamic in ((static_source, False), (dynamic_source, True)):
        if source is None:
            continue
        tree = source if isinstance(source, ast.AST) else parse_expression(source, time_dependent=dynamic)
        if is_zero(tree):
            continue
        for sign, term in additive_terms(tree):
            timed = "t" in variables(term)
            groups[term_target(term)][1 if timed else 0].append((sign, term))
    plan = PotentialPlan()
    for target, (fixed, timed) in groups.items():
        if fixed:
            plan.static[target] = compile_program(join_terms(fixed))
        if timed:
            plan.dynamic[target] = compile_program(join_terms(timed))
    return plan


def evaluate(node, x, y, z, t=0.0, np=None):
    """Evaluate a tree with numpy broadcasting (host-side, used by tests)."""
    import numpy
    np = np or numpy
    funcs = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "sqrt": np.sqrt,
             "tanh": np.tanh, "abs": np.abs, "log": np.log}

    def ev(item):
        if isinstance(item, ast.Constant):
            return float(item.value)
        if isinstance(item, ast.Name):
            return {"x": x, "y": y, "z": z, "t": t, "pi": math.pi}[item.id]
        if isinstance(item, ast.UnaryOp):
            value = ev(item.operand)
            return -value if isinstance(item.op, ast.USub) else value
        if isinstance(item, ast.Call):
            return funcs[item.func.id](ev(item.args[0]))
        if isinstance(item, _Max):
            return np.maximum(ev(item.left), ev(item.right))
        a, b = ev(item.left), ev(item.right)
        return {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b,
                ast.Div: lambda: a / b, ast.Pow: lambda: np.power(a, b)}[type(item.op)]()

    return ev(node)
