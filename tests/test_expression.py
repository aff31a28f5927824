"""CPU tests of the custom-propensity pipeline: the restated string stages
against the reference's own unit-test expectations, the C++ text's values
against g++ (tests/golden/expr_cases.json, written by
make_expr_fixtures.py), the bytecode compiler through its CPU mirror, the
host-side program check, the lowering of custom reactions, and the oracle
against the reference engine running g++-compiled custom propensities
(tests/golden/c6_*.npz)."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

import spatialpy_amd as sp
from spatialpy_amd import expression as E, lowering, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _fixture():
    with open(os.path.join(GOLDEN, "expr_cases.json")) as f:
        return json.load(f)


def _hex(v):
    return float.fromhex(v)


def _ulps(a, b):
    if math.isnan(a) and math.isnan(b):
        return 0
    if a == b:
        return 0
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    return abs(int(ia) - int(ib))


# ---------------------------------------------------------------- stage A
@pytest.mark.parametrize("text,expect", [
    # test/unit_tests/test_reaction.py:47-85, 221-241, 738-789
    ("k1 * A", "(k1*A)"), ("k1 * A * B", "((k1*A)*B)"), ("k1 * A * B / vol", "(((k1*A)*B)/vol)"),
    ("k1*A/2", "((k1*A)/2)"), ("k1*A/3", "((k1*A)/3)"), (20, "20"), (0.5, "0.5"), (5, "5"),
    # reaction.py:166-253: pow, e, ln, unary operators, dropped operators
    ("A^2", "pow(A,2)"), ("A**2", "pow(A,2)"), ("e*A", "(2.718281828459045*A)"), ("ln(A)", "log(A)"),
    ("-A", "(-A)"), ("+A", "(-A)"), ("not A", "(-A)"), ("A%B", "(AB)"), ("A > 3", "A3"),
])
def test_stage_a_matches_reference_unit_tests(text, expect):
    assert E.reaction_text(text) == expect


def test_mass_action_strings_match_reference_unit_tests():
    # test_reaction.py:540-585
    X, A = sp.Species("X", 1.0), sp.Species("A", 1.0)
    r = sp.Reaction("r", reactants={A: 1, X: 1}, products={}, rate="k1")
    assert (r.propensity_function, r.ode_propensity_function) == ("(((k1*A)*X)/vol)", "((k1*A)*X)")
    r = sp.Reaction("r", reactants={X: 2}, products={}, rate="k1")
    assert (r.propensity_function, r.ode_propensity_function) == ("(((k1*X)*(X-1))/vol)", "((k1*X)*X)")
    r = sp.Reaction("r", reactants={X: 1}, products={}, rate="k1")
    assert (r.propensity_function, r.ode_propensity_function) == ("(k1*X)", "(k1*X)")
    r = sp.Reaction("r", reactants={}, products={X: 1}, rate="k1")
    assert (r.propensity_function, r.ode_propensity_function) == ("(k1*vol)", "k1")
    # custom: each function defaults to the other (reaction.py:126-135)
    r = sp.Reaction("r", reactants={A: 1}, products={}, ode_propensity_function="k1 * A * B")
    assert r.propensity_function == r.ode_propensity_function == "((k1*A)*B)"


# ------------------------------------------------------- stage B and C types
def test_reference_expression_cases_behave_as_the_reference_test_requires():
    """test_solver.py:202-255: the converted C of each case, compiled by g++,
    agrees with Python's value to 3 places (expressions) or exactly
    (comparisons and boolean operators)."""
    fx = _fixture()
    for c in fx["cases"]:
        if not c["name"].startswith(("ref_expr", "ref_bool")):
            continue
        for hexv, py in zip(c["expect_hex"], c["python"]):
            got = _hex(hexv)
            if c["boolean"]:
                assert bool(got) == bool(py), c
            else:
                assert abs(got - py) < 0.5e-3 * max(1.0, abs(py)) or round(got - py, 3) == 0, c


def test_c_text_and_compiled_programs_match_gxx_bit_for_bit():
    """The restated conversion gives the fixture's C++ text, and the
    compiled bytecode run by the CPU mirror of the device interpreter gives
    g++'s result bits (same libm here) for every input -- int arithmetic,
    precedence quirks and all."""
    fx = _fixture()
    ns = E.model_namespace(fx["species"], list(fx["params"]), fx["data_functions"])
    cparams = {f"P{i}": v for i, v in enumerate(fx["params"].values())}
    checked = 0
    for c in fx["cases"]:
        if c["stage"] == "cpp":
            ns_c = dict({name: name for name in math.__dict__})
            ns_c.update({a: f"x[{i}]" for i, a in enumerate(c["args"])})
            cpp = E.cpp_text(c["text"], ns_c)
            n_x, n_df, params = len(c["args"]), 0, {}
        else:
            assert E.reaction_text(c["text"]) == c["stage_a"]
            cpp = E.cpp_text(c["stage_a"], ns)
            n_x, n_df, params = len(fx["species"]), len(fx["data_functions"]), cparams
        assert cpp == c["cpp"], c["name"]
        tree = E.parse_c(cpp, params, n_x, n_df, c["x_type"])
        p = E.Program()
        start = p.add(tree)
        code, consts = p.arrays()
        E.check_program(code, consts, start, n_x, n_df)
        for row, hexv in zip(c["inputs"], c["expect_hex"]):
            x = row[:n_x]
            if c["x_type"] == E.I:
                x = [int(v) for v in x]
            got = E.run_program(code, consts, start, x, row[n_x], row[n_x + 1], row[n_x + 2:])
            want = _hex(hexv)
            # CPython's own lgamma/erf differ from glibc's by an ulp or so
            tol = 0 if not any(f in cpp for f in ("lgamma(", "erf(", "erfc(")) else 8
            assert _ulps(got, want) <= tol, (c["name"], row, got, want)
            checked += 1
    assert checked > 2000


def test_c_semantics_the_generated_code_has():
    ns = E.model_namespace(["A", "B"], ["k"], [])
    run = lambda s, x, prm=(2.5,): E.run_program(*_prog(s, ns, prm, 2), x, 0.1, 0.0)

    def _prog(s, ns, prm, nx):
        t = E.parse_c(E.cpp_text(E.reaction_text(s), ns), {f"P{i}": v for i, v in enumerate(prm)}, nx, 0, E.I)
        p = E.Program()
        st = p.add(t)
        code, consts = p.arrays()
        return code, consts, st
    assert run("A/2", [7, 0]) == 3                    # int division truncates
    assert run("-A/2", [7, 0]) == -3                   # toward zero
    assert run("1/2*A", [7, 0]) == 0                   # int literal division
    assert run("k*A/2", [7, 0]) == 2.5 * 7 / 2
    assert run("A*A*A*A", [50000, 0]) == ((50000 ** 4 + 2 ** 31) % 2 ** 32) - 2 ** 31   # 32-bit wrap
    assert run("A^2", [3, 0]) == 9.0


@pytest.mark.parametrize("text", ["A%B", "k1*(A>3)", "abs(A)", "pi*A", "math.exp(A)", "log(A, 2)", "q*A"])
def test_expressions_the_reference_cannot_compile_are_rejected(text):
    ns = E.model_namespace(["A", "B"], ["k1"], [])
    with pytest.raises((E.ExpressionError, SyntaxError)):
        E.parse_c(E.cpp_text(E.reaction_text(text), ns), {"P0": 1.0}, 2, 0, E.I)


def test_stack_limit_and_program_check():
    ns = E.model_namespace(["A", "B"], [], [])
    deep = "A*(B+1)"
    for _ in range(6):
        deep = f"A*(B+({deep}))"
    t = E.parse_c(E.cpp_text(E.reaction_text(deep), ns), {}, 2, 0, E.I)
    p = E.Program()
    p.add(t)                      # leaves fold into instructions: shallow stack
    assert p.depth <= 2
    wide = "(A+B)"
    for _ in range(E.STACK_MAX):
        wide = f"(A+B)*({wide}+(A-B))"
    with pytest.raises(E.ExpressionError, match="stack"):
        E.Program().add(E.parse_c(E.cpp_text(E.reaction_text(wide), ns), {}, 2, 0, E.I))
    code, consts = p.arrays()
    bad = code.copy()
    bad[-1] = E.encode(E.OP_ADD_D)             # no RET
    with pytest.raises(E.ExpressionError):
        E.check_program(bad, consts, 0, 2, 0)
    with pytest.raises(E.ExpressionError):
        E.check_program(code, consts, 0, 1, 0)   # species index out of range


def test_table_shapes_for_mass_action_and_equivalent_custom_strings():
    ns = E.model_namespace(["A", "B"], ["k"], [])
    prm = {"P0": 0.7}

    def shape(prop, ode):
        tp = E.parse_c(E.cpp_text(E.reaction_text(prop), ns), prm, 2, 0, E.I)
        to = E.parse_c(E.cpp_text(E.reaction_text(ode), ns), prm, 2, 0, E.D)
        return E.table_shape(tp, to)
    for reac, kind, a, b in (({}, 0, 0, 0), ({"A": 1}, 1, 0, 0), ({"A": 1, "B": 1}, 2, 0, 1), ({"B": 2}, 3, 1, 0)):
        assert shape(*E.mass_action_texts("k", reac)) == (kind, a, b, 0.7)
    assert shape("k*A", "k*A") == (1, 0, 0, 0.7)
    assert shape("A/2", "A/2") is None                 # int arithmetic: a program
    assert shape(*E.mass_action_texts("2", {"A": 1})) is None   # int rate: int product in C


# --------------------------------------------------------------- lowering
def test_lowering_of_custom_reactions():
    sa = workloads.custom_chemistry_2d()
    kinds = [r[0] for r in sa.reactions]
    assert kinds == [4, 4, 4, 4, 1, 4]
    # any custom propensity makes the dependency graph dense
    # (core/model.py:187-196)
    S, R = sa.n_species, sa.n_reactions
    assert np.array_equal(sa.jcG, np.arange(0, (S + R + 1) * R, R))
    assert np.array_equal(sa.irG, np.tile(np.arange(R), S + R))
    assert sa.data_fn.shape == (sa.n, 1)
    assert np.array_equal(sa.data_fn[:, 0], 1.0 + 2.0 * sa.x[:, 0])
    for r in sa.reactions:
        if r[0] == 4:
            E.check_program(sa.vm_code, sa.vm_consts, r[1], S, 1)
            E.check_program(sa.vm_code, sa.vm_consts, r[2], S, 1)
    # the reaction restricted to type 1 carries the generated `sd == 1`
    assert sa.reactions[2][4] == 1 << 1


def test_bad_programs_are_rejected_before_the_gpu():
    from spatialpy_amd import _abi
    lib = _abi.load_library()
    sa = workloads.custom_chemistry_2d()
    for patch in ("no_ret", "bad_species", "bad_const"):
        s2 = workloads.custom_chemistry_2d()
        code = s2.vm_code.copy()
        if patch == "no_ret":
            code[:] = E.encode(E.OP_LD, E.SRC_VOL)
        elif patch == "bad_species":
            code[0] = E.encode(E.OP_LD, E.SRC_X, 0, 7)
        else:
            code[0] = E.encode(E.OP_LD, E.SRC_CONST, 0, 999)
        s2.vm_code = code
        dh = lowering.DescHandle(s2)
        h = ctypes.c_void_p()
        rc = lib.ssasdpd_create(dh.byref(), ctypes.byref(h))
        assert rc == 1 and b"program" in lib.ssasdpd_last_error(), patch
    assert sa.vm_code is not None


# ------------------------------------------- oracle vs the reference engine
@pytest.mark.parametrize("name", ["c6_custom_chemistry", "c6_custom_chemistry_libm"])
def test_oracle_custom_propensities_bit_exact_with_reference(oracle, name):
    """The reference engine ran the g++-compiled C++ text of each reaction
    (oracle/ref_harness.cpp loads it); the oracle runs the compiled VM
    programs: same events, counts and deterministic concentrations."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    sa = getattr(workloads, meta["gen"])(**meta["params"])
    assert np.array_equal(sa.x, z["x"]) and np.array_equal(sa.u0, z["u0"])
    o = oracle.Oracle(sa)
    res = o.run(seed=meta["seed"])
    st = o.stats()
    assert [st.reactions, st.diffusions] == z["counters"].tolist()
    assert np.array_equal(res["xx"].ravel(), z["xx"])
    assert np.array_equal(res["C"].ravel().view(np.uint64), z["snap_C"].view(np.uint64))
