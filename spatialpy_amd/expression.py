"""Propensity expressions: from the text a SpatialPy model holds to the
bytecode the engine's expression VM runs.

The reference turns a propensity string into C++ in two stages and lets g++
give it meaning:

  A. ``Reaction._create_custom_propensity`` (core/reaction.py:166-253,
     325-330): ``^`` -> ``**``, Python parse, ``**``/``^`` -> ``pow(a,b)``,
     ``e`` -> 2.718281828459045, then a printer that parenthesises every
     binary operation, prints any unary operation as ``(-x)`` and prints
     nothing for operators it has no visitor for;
  B. ``BuildExpression.getexpr_cpp`` (solvers/build_expression.py:241-262,
     381-753) with the model namespace (core/model.py:270-278): names are
     checked and replaced (species -> ``x[i]``, parameters -> ``P<i>``,
     data functions -> ``data_fn[k]``, ``math`` names and ``vol``, ``t``
     kept), then printed as C++ -- binary operations parenthesised,
     comparisons and ``&&``/``||`` NOT parenthesised.

The result is the body of ``double f(const int *x, double t, const double
vol, const double *data_fn, int sd)`` (solver.py:344-373), compiled as
C++14 with <math.h>; the deterministic right-hand side uses the same text
with ``const double *x`` (solver.py:160-190).  So the C++ text, read with
C precedence and C types, is what a propensity means: ``x[i]`` is an
``int`` in a propensity and a ``double`` in a right-hand side, integer
literals are ``int`` (``1/2`` is 0), parameters are ``double``.

This module restates both stages, parses the C++ text with C precedence and
the usual arithmetic conversions, and compiles it to the engine's VM
bytecode (include/ssasdpd.h, ``ssasdpd_program``).  ``run_program`` is the
CPU mirror of the device interpreter, used to check the compiler.
"""
import ast
import math
import struct

import numpy as np

# ---------------------------------------------------------------------------
# bytecode (mirrors spatialpy_amd/csrc/vm.hpp)
OP_RET, OP_LD = 0, 1
OP_ADD_D, OP_SUB_D, OP_MUL_D, OP_DIV_D = 2, 3, 4, 5
OP_ADD_I, OP_SUB_I, OP_MUL_I, OP_DIV_I = 6, 7, 8, 9
OP_NEG_D, OP_NEG_I = 10, 11
OP_LT, OP_LE, OP_GT, OP_GE, OP_EQ, OP_NE = 12, 13, 14, 15, 16, 17
OP_AND, OP_OR = 18, 19
OP_CALL1, OP_CALL2 = 20, 21
N_OPS = 22

SRC_STACK, SRC_CONST, SRC_X, SRC_VOL, SRC_T, SRC_DF = 0, 1, 2, 3, 4, 5

STACK_MAX = 8          # device stack depth (registers)
IDX_MAX = (1 << 22) - 1

FUNC1 = ["exp", "log", "log10", "log2", "log1p", "expm1", "sqrt", "sin", "cos", "tan", "asin", "acos",
         "atan", "sinh", "cosh", "tanh", "asinh", "acosh", "atanh", "erf", "erfc", "lgamma", "fabs",
         "floor", "ceil", "trunc"]
FUNC2 = ["pow", "atan2", "hypot", "fmod", "copysign", "remainder"]
FUNC_ID = {f: i for i, f in enumerate(FUNC1)}
FUNC_ID.update({f: 32 + i for i, f in enumerate(FUNC2)})
# results the device gives bit-identical to the host libm: exact operations,
# and pow, which the device rounds once from a double-double power for
# integral exponents |n| <= 64 (the correctly rounded value; other
# exponents use the device libm)
EXACT_FUNCS = {"sqrt", "fabs", "floor", "ceil", "trunc", "fmod", "copysign", "remainder", "pow"}

I32_MIN, I32_MAX = -(1 << 31), (1 << 31) - 1


class ExpressionError(ValueError):
    pass


def encode(op, src=SRC_STACK, rev=0, idx=0):
    assert 0 <= op < 64 and 0 <= src < 8 and 0 <= idx <= IDX_MAX
    return op | (src << 6) | (rev << 9) | (idx << 10)


def decode(w):
    return w & 63, (w >> 6) & 7, (w >> 9) & 1, w >> 10


# ---------------------------------------------------------------------------
# stage A: core/reaction.py:166-253, 325-330
class _PowAndE(ast.NodeTransformer):
    """reaction.py:166-189: ``a**b`` / ``a^b`` -> ``pow(a, b)``; ``e`` ->
    its value."""

    def visit_BinOp(self, node):
        node.left = self.visit(node.left)
        node.right = self.visit(node.right)
        if isinstance(node.op, (ast.BitXor, ast.Pow)):
            return ast.copy_location(ast.Call(func=ast.Name(id="pow", ctx=ast.Load()),
                                              args=[node.left, node.right], keywords=[]), node)
        return node

    def visit_Name(self, node):
        if node.id == "e":
            return ast.copy_location(ast.Constant(value=float(np.e)), node)
        return node


class _ReactionText(ast.NodeVisitor):
    """reaction.py:191-253: the printer.  Operators other than + - * / print
    nothing, any unary operation prints as ``(-x)``, only positional
    arguments of a call are printed, and nodes without a visitor print their
    children back to back."""
    _subst = {"ln": "log"}
    _ops = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}

    def __init__(self):
        self.s = ""

    def visit_BinOp(self, node):
        self.s += "("
        self.visit(node.left)
        self.s += self._ops.get(type(node.op), "")
        self.visit(node.right)
        self.s += ")"

    def visit_Name(self, node):
        self.s += node.id

    def visit_Constant(self, node):
        # NodeVisitor routes int/float/complex constants to visit_Num, which
        # prints str(value); other constants have no visitor
        if isinstance(node.value, (int, float, complex)) and not isinstance(node.value, bool):
            self.s += str(node.value)

    def visit_Call(self, node):
        if not isinstance(node.func, ast.Name):
            raise ExpressionError("only plain function names can be called in a propensity")
        self.s += self._subst.get(node.func.id, node.func.id) + "("
        for k, arg in enumerate(node.args):
            if k:
                self.s += ","
            self.visit(arg)
        self.s += ")"

    def visit_UnaryOp(self, node):
        self.s += "(-"
        self.visit(node.operand)
        self.s += ")"


def reaction_text(expr):
    """The string Reaction keeps for a propensity (stage A)."""
    if isinstance(expr, (int, float)) and not isinstance(expr, bool):
        expr = str(expr)
    tree = ast.parse(str(expr).replace("^", "**"), mode="eval")
    tree = _PowAndE().visit(tree)
    p = _ReactionText()
    p.visit(tree)
    return p.s


def mass_action_texts(rate, reactants):
    """Reaction._create_mass_action (reaction.py:260-313): the propensity
    and deterministic-rate strings before stage A.  `rate` is a parameter
    name or a number's text; `reactants` maps species names to
    stoichiometry, in insertion order."""
    total = sum(reactants.values())
    if total > 2:
        raise ExpressionError("A mass-action reaction cannot involve more than two of one species "
                              "or one of two species")
    prop = str(rate)
    ode = str(rate)
    for name, st in reactants.items():
        if st == 2:
            prop = f"{prop} * {name} * ({name} - 1) / vol"
            ode += f" * {name} * {name}"
        else:
            prop += f" * {name}"
            ode += f" * {name}"
    order = len(reactants)
    if order == 2:
        prop += " / vol"
    elif order == 0:
        prop += " * vol"
    return prop, ode


# ---------------------------------------------------------------------------
# stage B: solvers/build_expression.py
class _Validate(ast.NodeTransformer):
    """ValidationVisitor (build_expression.py:53-195) with sanitize=True and
    the blacklist ["="]."""

    def __init__(self, namespace):
        self.ns = namespace
        self.bad_names = []
        self.bad_ops = []

    def visit_Name(self, node):
        if node.id not in self.ns:
            self.bad_names.append(node.id)
        else:
            node.id = self.ns[node.id]
        self.generic_visit(node)
        return node

    def visit_Call(self, node):
        if not isinstance(node.func, ast.Name):
            raise ExpressionError("only plain function names can be called")
        if node.func.id not in self.ns:
            self.bad_names.append(node.func.id)
        else:
            node.func.id = self.ns[node.func.id]
        self.generic_visit(node)
        return node

    def visit_Assign(self, node):
        self.bad_ops.append("=")
        self.generic_visit(node)
        return node


class _PowToCall(ast.NodeTransformer):
    """CppConverter.CppExpressionTransformer (build_expression.py:718-736)."""

    def visit_BinOp(self, node):
        self.generic_visit(node)
        if isinstance(node.op, ast.Pow):
            node = ast.copy_location(ast.Call(func=ast.Name(id="pow", ctx=ast.Load()),
                                              args=[node.left, node.right], keywords=[]), node)
        return node


class _CppText(ast.NodeVisitor):
    """ExpressionConverter + CppConverter (build_expression.py:381-753): a
    stack of strings; the right operand is visited first so that the left
    one is popped first.  Nodes without a visitor push their children's
    strings without combining them."""
    _arith = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.Pow: "**"}
    _cmp = {ast.Eq: "==", ast.NotEq: "!=", ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">="}

    def __init__(self):
        self.e = []

    def visit_Name(self, node):
        self.e.append(node.id)

    def visit_Constant(self, node):
        self.e.append(str(node.value))

    def visit_BinOp(self, node):
        self.visit(node.right)
        self.visit(node.left)
        op = self._arith.get(type(node.op))
        if op is not None:
            self.e.append(f"({self.e.pop()}{op}{self.e.pop()})")

    def visit_Call(self, node):
        args = []
        for a in node.args:
            self.visit(a)
            args.append(self.e.pop())
        self.e.append(f"{node.func.id}({','.join(args)})")

    def visit_BoolOp(self, node):
        for v in reversed(node.values):
            self.visit(v)
        op = "&&" if isinstance(node.op, ast.And) else "||"
        for _ in range(len(node.values) - 1):
            self.e.append(f"{self.e.pop()} {op} {self.e.pop()}")

    def visit_Compare(self, node):
        for c in node.comparators:
            self.visit(c)
        self.visit(node.left)
        for op in node.ops:
            sym = self._cmp.get(type(op))
            if sym is not None:
                self.e.append(f"{self.e.pop()} {sym} {self.e.pop()}")

    def visit_UnaryOp(self, node):
        self.visit(node.operand)
        if isinstance(node.op, ast.USub):
            self.e.append(f"-{self.e.pop()}")


def cpp_text(statement, namespace):
    """BuildExpression(namespace, blacklist=["="], sanitize=True)
    .getexpr_cpp(statement): the C++ text, or ExpressionError where the
    reference would return None (and write uncompilable code)."""
    tree = ast.parse(str(statement).replace("^", "**"))
    v = _Validate(namespace)
    v.visit(tree)
    if v.bad_ops or v.bad_names:
        raise ExpressionError(f"invalid names {v.bad_names} / operators {v.bad_ops} in '{statement}'")
    tree = _PowToCall().visit(tree)
    c = _CppText()
    c.visit(tree)
    return "".join(c.e)


def model_namespace(species, parameters, data_functions):
    """Model.__get_expression_utility (core/model.py:270-278) with the
    sanitised names of model.py:557-568, 703-714, 932-943."""
    ns = {name: name for name in math.__dict__}
    ns.update({s: f"x[{i}]" for i, s in enumerate(species)})
    ns.update({p: f"P{i}" for i, p in enumerate(parameters)})
    ns.update({d: f"data_fn[{i}]" for i, d in enumerate(data_functions)})
    ns.update({"vol": "vol", "t": "t"})
    return ns


# ---------------------------------------------------------------------------
# the C++ text, read with C precedence and types
I, D = "int", "double"


class _Node:
    __slots__ = ("k", "a", "b", "ty", "v")

    def __init__(self, k, ty, a=None, b=None, v=None):
        self.k, self.ty, self.a, self.b, self.v = k, ty, a, b, v


def _tokens(s):
    out, i, n = [], 0, len(s)
    two = ("&&", "||", "==", "!=", "<=", ">=")
    while i < n:
        c = s[i]
        if c.isspace():
            i += 1
            continue
        if c.isdigit() or (c == "." and i + 1 < n and s[i + 1].isdigit()):
            j = i
            while j < n and s[j].isdigit():
                j += 1
            isf = False
            if j < n and s[j] == ".":
                isf = True
                j += 1
                while j < n and s[j].isdigit():
                    j += 1
            if j < n and s[j] in "eE":
                k = j + 1
                if k < n and s[k] in "+-":
                    k += 1
                if k < n and s[k].isdigit():
                    isf = True
                    j = k
                    while j < n and s[j].isdigit():
                        j += 1
            if j < n and (s[j].isalpha() or s[j] == "_"):
                raise ExpressionError(f"bad number near '{s[i:j + 1]}'")
            out.append(("f" if isf else "i", s[i:j]))
            i = j
            continue
        if c.isalpha() or c == "_":
            j = i
            while j < n and (s[j].isalnum() or s[j] == "_"):
                j += 1
            out.append(("n", s[i:j]))
            i = j
            continue
        if s[i:i + 2] in two:
            out.append(("o", s[i:i + 2]))
            i += 2
            continue
        if c in "+-*/<>!()[],":
            out.append(("o", c))
            i += 1
            continue
        raise ExpressionError(f"unexpected character '{c}'")
    return out


class _Parser:
    """Precedence climbing over the C grammar subset the converter emits."""
    _levels = [("||",), ("&&",), ("==", "!="), ("<", "<=", ">", ">="), ("+", "-"), ("*", "/")]

    def __init__(self, text, sym):
        self.t = _tokens(text)
        self.p = 0
        self.sym = sym

    def peek(self):
        return self.t[self.p] if self.p < len(self.t) else (None, None)

    def take(self, val=None):
        tok = self.peek()
        if tok[0] is None or (val is not None and tok[1] != val):
            raise ExpressionError(f"expected '{val}' in C expression")
        self.p += 1
        return tok

    def parse(self):
        if not self.t:
            raise ExpressionError("empty expression")
        e = self.binary(0)
        if self.p != len(self.t):
            raise ExpressionError(f"unexpected '{self.t[self.p][1]}' in C expression")
        return e

    def binary(self, lvl):
        if lvl == len(self._levels):
            return self.unary()
        left = self.binary(lvl + 1)
        while True:
            kind, val = self.peek()
            if kind == "o" and val in self._levels[lvl]:
                self.p += 1
                right = self.binary(lvl + 1)
                left = _binop(val, left, right)
            else:
                return left

    def unary(self):
        kind, val = self.peek()
        if kind == "o" and val == "-":
            self.p += 1
            x = self.unary()
            return _Node("neg", x.ty, x)
        if kind == "o" and val == "+":
            self.p += 1
            return self.unary()
        if kind == "o" and val == "!":
            self.p += 1
            return _Node("not", I, self.unary())
        return self.primary()

    def primary(self):
        kind, val = self.take()
        if kind == "i":
            v = int(val)
            if v > I32_MAX:
                raise ExpressionError(f"integer literal {val} does not fit an int")
            return _Node("const", I, v=v)
        if kind == "f":
            return _Node("const", D, v=float(val))
        if kind == "o" and val == "(":
            e = self.binary(0)
            self.take(")")
            return e
        if kind == "n":
            nxt = self.peek()
            if nxt == ("o", "("):
                self.p += 1
                args = []
                if self.peek() != ("o", ")"):
                    while True:
                        args.append(self.binary(0))
                        if self.peek() == ("o", ","):
                            self.p += 1
                            continue
                        break
                self.take(")")
                return self.call(val, args)
            if nxt == ("o", "["):
                self.p += 1
                ik, iv = self.take()
                if ik != "i":
                    raise ExpressionError("array index must be an integer literal")
                self.take("]")
                return self.element(val, int(iv))
            return self.name(val)
        raise ExpressionError(f"unexpected '{val}' in C expression")

    def call(self, fn, args):
        if fn in FUNC1:
            if len(args) != 1:
                raise ExpressionError(f"{fn} takes one argument in C")
        elif fn in FUNC2:
            if len(args) != 2:
                raise ExpressionError(f"{fn} takes two arguments in C")
        else:
            raise ExpressionError(f"function '{fn}' is not available to propensities (not a C math function "
                                  "with a double result)")
        return _Node("call", D, a=args, v=fn)

    def element(self, name, k):
        if name == "x":
            if k >= self.sym["n_x"]:
                raise ExpressionError(f"species index {k} out of range")
            return _Node("x", self.sym["x_type"], v=k)
        if name == "data_fn":
            if k >= self.sym["n_df"]:
                raise ExpressionError(f"data function index {k} out of range")
            return _Node("df", D, v=k)
        raise ExpressionError(f"'{name}' is not an array")

    def name(self, nm):
        if nm == "vol":
            return _Node("vol", D)
        if nm == "t":
            return _Node("t", D)
        if nm in self.sym["params"]:
            return _Node("const", D, v=float(self.sym["params"][nm]))
        raise ExpressionError(f"'{nm}' is not defined in the generated C++")


def _binop(op, l, r):
    if op in ("+", "-", "*", "/"):
        ty = I if (l.ty == I and r.ty == I) else D
        return _Node("bin", ty, l, r, v=op)
    if op in ("&&", "||"):
        return _Node("logic", I, l, r, v=op)
    return _Node("cmp", I, l, r, v=op)


def parse_c(text, params, n_x, n_df=0, x_type=I):
    """Typed tree of a C++ propensity text.  params: {"P<i>": value}."""
    return _Parser(text, {"params": params, "n_x": n_x, "n_df": n_df, "x_type": x_type}).parse()


# ---------------------------------------------------------------------------
# compiler
_BIN = {("+", I): OP_ADD_I, ("-", I): OP_SUB_I, ("*", I): OP_MUL_I, ("/", I): OP_DIV_I,
        ("+", D): OP_ADD_D, ("-", D): OP_SUB_D, ("*", D): OP_MUL_D, ("/", D): OP_DIV_D}
_CMP = {"<": OP_LT, "<=": OP_LE, ">": OP_GT, ">=": OP_GE, "==": OP_EQ, "!=": OP_NE}
_LOGIC = {"&&": OP_AND, "||": OP_OR}


class Program:
    """Flat code and constants of several expressions."""

    def __init__(self):
        self.code = []
        self.consts = []
        self._kidx = {}
        self.depth = 0

    def const(self, v):
        key = struct.pack("<d", float(v))
        if key not in self._kidx:
            self._kidx[key] = len(self.consts)
            self.consts.append(float(v))
        return self._kidx[key]

    def leaf(self, n):
        """(src, idx) of a leaf operand, or None."""
        if n.k == "const":
            return SRC_CONST, self.const(n.v)
        if n.k == "x":
            return SRC_X, n.v
        if n.k == "vol":
            return SRC_VOL, 0
        if n.k == "t":
            return SRC_T, 0
        if n.k == "df":
            return SRC_DF, n.v
        return None

    def _need(self, n):
        """Stack slots needed to evaluate n."""
        lf = self.leaf(n)
        if lf is not None:
            return 1
        if n.k in ("neg", "not"):
            return self._need(n.a)
        if n.k == "call":
            if len(n.a) == 1:
                return self._need(n.a[0])
            l, r = n.a
            return self._pair_need(l, r)
        return self._pair_need(n.a, n.b)

    def _pair_need(self, l, r):
        if self.leaf(r) is not None:
            return self._need(l)
        if self.leaf(l) is not None:
            return self._need(r)
        return max(self._need(l), 1 + self._need(r))

    def _emit(self, n):
        lf = self.leaf(n)
        if lf is not None:
            self.code.append(encode(OP_LD, lf[0], 0, lf[1]))
            return
        if n.k == "neg":
            self._emit(n.a)
            self.code.append(encode(OP_NEG_I if n.ty == I else OP_NEG_D))
            return
        if n.k == "not":
            # !a == (a == 0)
            self._emit(n.a)
            self.code.append(encode(OP_EQ, SRC_CONST, 0, self.const(0.0)))
            return
        if n.k == "call":
            f = FUNC_ID[n.v]
            if len(n.a) == 1:
                self._emit(n.a[0])
                self.code.append(encode(OP_CALL1, SRC_STACK, 0, f))
                return
            self._pair(OP_CALL2, n.a[0], n.a[1], f)
            return
        if n.k == "bin":
            self._pair(_BIN[(n.v, n.ty)], n.a, n.b, 0)
        elif n.k == "cmp":
            self._pair(_CMP[n.v], n.a, n.b, 0)
        else:
            self._pair(_LOGIC[n.v], n.a, n.b, 0)

    def _pair(self, op, l, r, fn):
        """Binary operation; the right (or, reversed, the left) operand can
        be a leaf read in place.  CALL2 keeps its function id in idx, so its
        operands always come from the stack."""
        rl, ll = self.leaf(r), self.leaf(l)
        if op != OP_CALL2 and rl is not None:
            self._emit(l)
            self.code.append(encode(op, rl[0], 0, rl[1]))
        elif op != OP_CALL2 and ll is not None:
            self._emit(r)
            self.code.append(encode(op, ll[0], 1, ll[1]))
        else:
            self._emit(l)
            self._emit(r)
            self.code.append(encode(op, SRC_STACK, 0, fn))

    def add(self, tree):
        """Append one expression; returns its start index."""
        need = self._need(tree)
        if need > STACK_MAX:
            raise ExpressionError(f"expression needs a stack of {need} (the engine's VM holds {STACK_MAX})")
        start = len(self.code)
        self._emit(tree)
        self.code.append(encode(OP_RET))
        self.depth = max(self.depth, need)
        return start

    def arrays(self):
        return (np.asarray(self.code, dtype=np.uint32), np.asarray(self.consts, dtype=np.float64))


def check_program(code, consts, start, n_x, n_df):
    """Host-side validation the engine repeats before any upload: opcodes,
    operand indices, stack depth within STACK_MAX, a RET with one value."""
    d, pc = 0, start
    while True:
        if pc >= len(code):
            raise ExpressionError("program runs past the end of the code")
        op, src, rev, idx = decode(int(code[pc]))
        pc += 1
        if op >= N_OPS:
            raise ExpressionError("bad opcode")
        if op == OP_RET:
            if d != 1:
                raise ExpressionError("program leaves a bad stack")
            return pc
        if src != SRC_STACK:
            if src > SRC_DF or (src == SRC_CONST and idx >= len(consts)) or (src == SRC_X and idx >= n_x) or \
                    (src == SRC_DF and idx >= n_df):
                raise ExpressionError("operand out of range")
        if op == OP_LD:
            if src == SRC_STACK:
                raise ExpressionError("load from the stack")
            d += 1
        elif op in (OP_NEG_D, OP_NEG_I, OP_CALL1):
            if d < 1 or src != SRC_STACK:
                raise ExpressionError("unary operation on an empty stack")
            if op == OP_CALL1 and idx not in range(len(FUNC1)):
                raise ExpressionError("bad function")
        else:
            if op == OP_CALL2 and (idx < 32 or idx >= 32 + len(FUNC2) or src != SRC_STACK):
                raise ExpressionError("bad function")
            if src == SRC_STACK:
                if d < 2:
                    raise ExpressionError("binary operation on a short stack")
                d -= 1
            elif d < 1:
                raise ExpressionError("binary operation on an empty stack")
        if d > STACK_MAX:
            raise ExpressionError("stack overflow")


# ---------------------------------------------------------------------------
# CPU mirror of the device interpreter (C semantics; this host's libm)
def _wrap(v):
    return ((int(v) - I32_MIN) & 0xFFFFFFFF) + I32_MIN


def _idiv(a, b):
    a, b = int(a), int(b)
    if b == 0:
        return 0            # the reference's program dies with SIGFPE here
    if a == I32_MIN and b == -1:
        return I32_MIN
    q = abs(a) // abs(b)
    return q if (a < 0) == (b < 0) else -q


_NP = {"asin": "arcsin", "acos": "arccos", "atan": "arctan", "asinh": "arcsinh", "acosh": "arccosh",
       "atanh": "arctanh", "pow": "power", "atan2": "arctan2"}


def _libm(fn, *a):
    """math.<fn> is the C library's function; where Python raises instead
    of returning the IEEE result (overflow, pole, domain), numpy's
    float64 result (inf / -inf / nan) is used."""
    try:
        return float(getattr(math, fn)(*a))
    except (OverflowError, ValueError):
        if fn == "lgamma":
            return math.inf
        if fn == "remainder":
            return math.nan
        with np.errstate(all="ignore"):
            return float(getattr(np, _NP.get(fn, fn))(*[np.float64(v) for v in a]))


def _ddiv(a, b):
    with np.errstate(all="ignore"):
        return float(np.float64(a) / np.float64(b))


def run_program(code, consts, start, x, vol, t, df=()):
    st = []
    pc = start
    while True:
        op, src, rev, idx = decode(int(code[pc]))
        pc += 1
        if op == OP_RET:
            return float(st[-1])

        def operand():
            if src == SRC_CONST:
                return float(consts[idx])
            if src == SRC_X:
                return x[idx]
            if src == SRC_VOL:
                return float(vol)
            if src == SRC_T:
                return float(t)
            return float(df[idx])
        if op == OP_LD:
            st.append(operand())
            continue
        if op == OP_NEG_D:
            st[-1] = -float(st[-1])
            continue
        if op == OP_NEG_I:
            st[-1] = _wrap(-int(st[-1]))
            continue
        if op == OP_CALL1:
            st[-1] = _libm(FUNC1[idx], float(st[-1]))
            continue
        if src == SRC_STACK:
            b = st.pop()
        else:
            b = operand()
        a = st[-1]
        if rev:
            a, b = b, a
        if op == OP_CALL2:
            st[-1] = _libm(FUNC2[idx - 32], float(a), float(b))
        elif op in (OP_ADD_I, OP_SUB_I, OP_MUL_I):
            a, b = int(a), int(b)
            st[-1] = _wrap(a + b if op == OP_ADD_I else a - b if op == OP_SUB_I else a * b)
        elif op == OP_DIV_I:
            st[-1] = _idiv(a, b)
        elif op in (OP_ADD_D, OP_SUB_D, OP_MUL_D, OP_DIV_D):
            a, b = float(a), float(b)
            if op == OP_ADD_D:
                st[-1] = a + b
            elif op == OP_SUB_D:
                st[-1] = a - b
            elif op == OP_MUL_D:
                st[-1] = a * b
            else:
                st[-1] = _ddiv(a, b)
        elif op in (OP_LT, OP_LE, OP_GT, OP_GE, OP_EQ, OP_NE):
            a, b = float(a), float(b)
            st[-1] = int({OP_LT: a < b, OP_LE: a <= b, OP_GT: a > b, OP_GE: a >= b, OP_EQ: a == b,
                          OP_NE: a != b}[op])
        elif op == OP_AND:
            st[-1] = int(float(a) != 0.0 and float(b) != 0.0)
        elif op == OP_OR:
            st[-1] = int(float(a) != 0.0 or float(b) != 0.0)
        else:
            raise ExpressionError("bad opcode")


# ---------------------------------------------------------------------------
# mass-action shapes the engine evaluates from its table
K_ZERO, K_FIRST, K_AB, K_2X, K_PROGRAM = 0, 1, 2, 3, 4


def _is_k(n):
    return n.k == "const" and n.ty == D


def _is_bin(n, op):
    return n.k == "bin" and n.v == op and n.ty == D


def table_shape(prop, ode):
    """(kind, a, b, k) when the propensity tree and the deterministic-rate
    tree are exactly the four mass-action shapes the engine's table
    evaluates (same operations, same order), else None:
      0: (k*vol) / k          1: (k*X) / (k*X)
      2: (((k*X)*Y)/vol) / ((k*X)*Y)      3: (((k*X)*(X-1))/vol) / ((k*X)*X)"""
    p, o = prop, ode
    if _is_bin(p, "*") and _is_k(p.a) and p.b.k == "vol":
        if _is_k(o) and o.v == p.a.v:
            return K_ZERO, 0, 0, p.a.v
        return None
    if _is_bin(p, "*") and _is_k(p.a) and p.b.k == "x":
        k, a = p.a.v, p.b.v
        if _is_bin(o, "*") and _is_k(o.a) and o.a.v == k and o.b.k == "x" and o.b.v == a:
            return K_FIRST, a, 0, k
        return None
    if not (_is_bin(p, "/") and p.b.k == "vol" and _is_bin(p.a, "*")):
        return None
    m = p.a
    if not (_is_bin(m.a, "*") and _is_k(m.a.a) and m.a.b.k == "x"):
        return None
    k, a = m.a.a.v, m.a.b.v
    if not (_is_bin(o, "*") and _is_bin(o.a, "*") and _is_k(o.a.a) and o.a.a.v == k and o.a.b.k == "x"
            and o.a.b.v == a and o.b.k == "x"):
        return None
    if m.b.k == "x":
        if o.b.v == m.b.v:
            return K_AB, a, m.b.v, k
        return None
    if m.b.k == "bin" and m.b.v == "-" and m.b.ty == I and m.b.a.k == "x" and m.b.a.v == a and \
            m.b.b.k == "const" and m.b.b.ty == I and m.b.b.v == 1 and o.b.v == a:
        return K_2X, a, 0, k
    return None
