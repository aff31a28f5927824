"""Write tests/golden/expr_cases.json: propensity expressions, the C++ text
the reference's two-stage conversion gives them (restated in
spatialpy_amd/expression.py), and what that text computes when compiled by
g++ -std=c++14 with <math.h> -- the way the reference builds its generated
model file (solver.py:443-507) and the way its own expression test
evaluates converted expressions (test/integration_tests/test_solver.py:
202-255 compiles them with gcc).

Cases:
  * ref_*: the expression and comparison cases of test_solver.py:65-117,
    stage B only (BuildExpression.getexpr_cpp) over double arguments, with
    the Python value that test compares against;
  * prop_* / ode_*: custom propensity strings (stage A + B) in a model
    namespace -- species as `int` (propensities) or `double` (deterministic
    rates), parameters, a data function, vol and t.

Run where g++ is available:  python tests/golden/make_expr_fixtures.py
"""
import itertools
import json
import math
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from spatialpy_amd import expression as E  # noqa: E402

REF_EXPR = [
    ("x*2", ["x"], [[0.0], [1.0], [-1.0], [9.999], [-9.999]]),
    ("x*2 + x/2 - (x*3)^2 + x/3^2", ["x"], [[0.0], [1.0], [-1.0], [3.333], [-3.333], [9.8765], [-9.8765]]),
    ("(x-1)*y^2+x", ["x", "y"], [[1.0, 2.4], [5.1, 0.0], [5.1, 1.0], [5.1, -1.0], [9.8765, -1.0], [-1.0, 9.8765]]),
    ("(x^2/y^2/z^2)/x^2/y^2/z^2**1/x**1/y**1/z", ["x", "y", "z"],
     [[5.1, 0.1, 2.0], [0.1, 5.1, 2.0], [2.0, 0.1, 5.1], [2.0, 5.1, 0.1]]),
]
REF_BOOL = [
    ("x > 0", ["x"], [[100], [0], [0.001], [-1]]),
    ("x > y", ["x", "y"], [[100, 99], [99, 100], [-10, 10], [10, -10], [0.001, 0.0], [0.0, 0.001], [-99.999, -99.998]]),
    ("x > 0 and y < x", ["x", "y"], [[100, 99], [99, 100], [0, -100], [-0.001, -99.0], [0, 0.001], [-0.001, 0]]),
    ("x > 0 and y < 10 and x > y", ["x", "y"],
     [[100, 9], [0.01, 0.00], [100, 200], [0.01, 0.02], [0, 0], [-0.01, -0.02], [-0.01, 0]]),
    ("x > 0 and y < 10 or y > 100", ["x", "y"], [[10, 9], [0.01, 9.99], [0, 10], [-1.0, -1.0]]),
    ("x^2>x and y<y^2 or z^2!=z^3 and y!=z", ["x", "y", "z"],
     [[1.0, 1.0, 1.0], [99.9, 99.9, 100.0], [0.0, -1.0, 99.9], [-1.0, -1.0, 0.00]]),
]
# C-precedence probes of stage B (comparisons and && / || are printed
# without parentheses)
REF_QUIRK = [
    ("(x > 1) * y", ["x", "y"], [[2.0, 3.0], [0.5, 3.0], [2.0, 0.5]]),
    ("x > 0 and (y > 0 or x > 5)", ["x", "y"], [[1.0, -1.0], [6.0, -1.0], [-1.0, 1.0], [1.0, 1.0]]),
]

SPECIES = ["A", "B"]
PARAMS = {"k1": 2.5, "K": 3.0, "kd": 0.05}
DFS = ["g"]
PROP = [
    "A/2", "A*(A-1)/2", "k1*A/2", "A/2*k1", "(A+B)/(1+A)", "2^B", "-A+3", "k1*A/(1+pow(B/K,4))",
    "e*A", "ln(A+1)", "sqrt(A)*k1", "exp(-t)*A", "vol*k1", "g*A", "A*A*A*A*A", "B/A", "floor(k1*A)/2",
    "fmod(A, 3)", "hypot(A,B)", "atan2(B, A+1)", "1/2*A", "1.0/2*A", "A/B*B", "-(A-B)/2", "K*K-A*A/3",
    "k1*A*B/vol", "kd*A*(A-1)/vol", "tanh(A/K)*log10(B+1)", "1e-3*A+B*2.5e2", "+A", "A-B-1",
    "copysign(k1, B-A)", "expm1(-kd*t)*A", "lgamma(A+1)", "erf(kd*A)", "cosh(kd*B)",
]
ODE = ["A/2", "A*(A-1)/2", "kd*A*A", "B/3+kd*B", "pow(A,2)/(K*K+pow(A,2))", "sqrt(B)*k1*g", "(A+B)/(1+A)"]
ERRORS = ["A%B", "k1*(A>3)", "abs(A)", "pi*A", "math.exp(A)", "log(A, 2)", "q*A"]

A_VALS = [0, 1, 2, 3, 7, 100, 46341]
B_VALS = [0, 1, 5, 13]
OTHER = [(1.0 / 144, 0.0, 1.0), (0.3, 1.7, 2.5)]


def _c_number(v):
    # shortest round-trip decimal (hex literals are not C++14)
    return repr(float(v))


def gcc_eval(cpp, x_type, n_x, n_df, inputs, bool_result=False):
    """Compile the C++ text inside a generated-style function and evaluate
    it for each input row [x..., vol, t, df...]; hex results."""
    xt = "int" if x_type == "int" else "double"
    rows = ",".join("{" + ",".join(_c_number(v) for v in r) + "}" for r in inputs)
    width = n_x + 2 + n_df
    src = f"""#include <math.h>
#include <stdio.h>
#include <stddef.h>
{''.join(f'static const double P{i} = {float(v)!r};' for i, v in enumerate(PARAMS.values()))}
static double f(const {xt} *x, double t, const double vol, const double *data_fn, int sd) {{
    (void)x; (void)t; (void)vol; (void)data_fn; (void)sd;
    return {cpp};
}}
static const double rows[][{max(width, 1)}] = {{{rows}}};
int main() {{
    for (size_t r = 0; r < sizeof rows / sizeof rows[0]; r++) {{
        {xt} x[{max(n_x, 1)}];
        for (int k = 0; k < {n_x}; k++) x[k] = ({xt})rows[r][k];
        double df[{max(n_df, 1)}];
        for (int k = 0; k < {n_df}; k++) df[k] = rows[r][{n_x} + 2 + k];
        double y = f(x, rows[r][{n_x} + 1], rows[r][{n_x}], df, 1);
        printf("%a\\n", y);
    }}
    return 0;
}}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "e.cpp")
        exe = os.path.join(d, "e")
        with open(c, "w") as fh:
            fh.write(src)
        subprocess.run(["g++", "-std=c++14", "-w", c, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    return out


def _no_int_div_by_zero(prog, start, x, vol, t, df):
    orig = E._idiv

    def strict(a, b):
        if int(b) == 0:
            raise ZeroDivisionError
        return orig(a, b)
    E._idiv = strict
    try:
        E.run_program(prog[0], prog[1], start, x, vol, t, df)
        return True
    except ZeroDivisionError:
        return False
    finally:
        E._idiv = orig


def _exact(cpp):
    """True when every function the text calls is one the device gives
    bit-identical to the host libm (all pow calls here have integral
    exponents)."""
    return not any(f"{fn}(" in cpp for fn in E.FUNC1 + E.FUNC2 if fn not in E.EXACT_FUNCS)


def main():
    cases = []
    # reference cases: doubles through stage B, x/y/z -> x[0..2]
    for group, items in (("ref_expr", REF_EXPR), ("ref_bool", REF_BOOL), ("ref_quirk", REF_QUIRK)):
        for k, (text, args, values) in enumerate(items):
            ns = dict({name: name for name in math.__dict__})
            ns.update({a: f"x[{i}]" for i, a in enumerate(args)})
            cpp = E.cpp_text(text, ns)
            inputs = [[float(v) for v in vals] + [1.0, 0.0] for vals in values]
            out = gcc_eval(cpp, "double", len(args), 0, inputs)
            py = [float(eval(text.replace("^", "**"), {}, dict(zip(args, vals)))) for vals in values]
            cases.append(dict(name=f"{group}_{k}", text=text, stage="cpp", args=args, x_type="double",
                              cpp=cpp, inputs=inputs, expect_hex=out, python=py,
                              boolean=group == "ref_bool", exact=_exact(cpp)))
    ns = E.model_namespace(SPECIES, list(PARAMS), DFS)
    cparams = {f"P{i}": v for i, v in enumerate(PARAMS.values())}
    for group, items, x_type in (("prop", PROP, E.I), ("ode", ODE, E.D)):
        for k, text in enumerate(items):
            stage_a = E.reaction_text(text)
            cpp = E.cpp_text(stage_a, ns)
            tree = E.parse_c(cpp, cparams, len(SPECIES), len(DFS), x_type)
            p = E.Program()
            start = p.add(tree)
            prog = p.arrays()
            inputs = []
            for a, b, (vol, t, g) in itertools.product(A_VALS, B_VALS, OTHER):
                if _no_int_div_by_zero(prog, start, [a, b] if x_type == E.I else [float(a), float(b)], vol, t, [g]):
                    inputs.append([float(a), float(b), vol, t, g])
            out = gcc_eval(cpp, "int" if x_type == E.I else "double", len(SPECIES), len(DFS), inputs)
            exact = _exact(cpp)
            cases.append(dict(name=f"{group}_{k}", text=text, stage="reaction", stage_a=stage_a, x_type=x_type,
                              cpp=cpp, inputs=inputs, expect_hex=out, exact=exact))
    errors = []
    for text in ERRORS:
        try:
            E.parse_c(E.cpp_text(E.reaction_text(text), ns), cparams, len(SPECIES), len(DFS), E.I)
            raise SystemExit(f"expected an error for {text}")
        except (E.ExpressionError, SyntaxError) as e:
            errors.append(dict(text=text, error=type(e).__name__))
    fixture = dict(species=SPECIES, params=PARAMS, data_functions=DFS, cases=cases, errors=errors)
    with open(os.path.join(HERE, "expr_cases.json"), "w") as fh:
        json.dump(fixture, fh, indent=1)
    print(len(cases), "cases,", sum(len(c["inputs"]) for c in cases), "evaluations,", len(errors), "error cases")


if __name__ == "__main__":
    main()
