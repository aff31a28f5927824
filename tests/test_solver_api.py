"""The Solver drop-in, tested like the reference's own integration tests
(test/integration_tests/test_solver.py:120-200, test/system_tests/
test_compiler.py:31-36 with test/models/birth_death.py)."""
import numpy as np
import pytest

import spatialpy_amd as sp
from spatialpy_amd import workloads


def birth_death_model(seed=1):
    """test/models/birth_death.py:19-50 written against this package."""
    model = sp.Model(name="Spatial Birth-Death")
    domain = sp.Domain.create_2D_domain(xlim=(0, 1), ylim=(0, 1), numx=10, numy=10, type_id="Habitat", fixed=True)
    model.add_domain(domain)
    model.add_species(sp.Species(name="Rabbits", diffusion_coefficient=0.1))
    model.add_initial_condition(sp.ScatterInitialCondition(species="Rabbits", count=100, seed=seed))
    model.add_parameter([sp.Parameter(name="k_birth", expression=10), sp.Parameter(name="k_death", expression=0.1)])
    model.add_reaction([sp.Reaction(name="birth", reactants={}, products={"Rabbits": 1}, rate="k_birth"),
                        sp.Reaction(name="death", reactants={"Rabbits": 1}, products={}, rate="k_death")])
    model.timespan(sp.TimeSpan.linspace(t=10, num_points=11, timestep_size=1))
    return model


def diffusion_debug_model():
    """test/integration_tests/test_solver.py:28-46 (50x50 at D=0.01) with 5000
    molecules at the origin and t=0..1 instead of 1e5 and t=0..10, to keep
    the event count small."""
    model = sp.Model("diffusion_debug_test")
    model.add_species(sp.Species(name="A", diffusion_coefficient=0.01))
    domain = sp.Domain.create_2D_domain(xlim=[-1, 1], ylim=[-1, 1], numx=50, numy=50, type_id=1, mass=1.0, nu=1.0,
                                        fixed=True, rho0=1.0, c0=1.0, P0=1.0)
    model.add_domain(domain)
    model.add_initial_condition(sp.PlaceInitialCondition("A", 5000, [0, 0, 0]))
    model.timespan(sp.TimeSpan.linspace(t=1, num_points=11, timestep_size=0.1))
    return model


class _NoEngineSolver(sp.Solver):
    """Lowering only (the engine needs a GPU)."""

    def compile(self, debug=False, profile=False):
        try:
            super().compile(debug, profile)
        except sp.SimulationError as e:
            if e.code != 7:      # only "no HIP device" is expected here
                raise


def test_lowering_of_the_reference_birth_death_model():
    s = _NoEngineSolver(birth_death_model())
    s.compile()
    ref = workloads.birth_death_2d()
    assert np.array_equal(s.sa.x, ref.x) and s.sa.h == ref.h and np.array_equal(s.sa.rho, ref.rho)
    assert s.sa.dimension == 2 and s.sa.nt == 10 and s.sa.output_steps.tolist() == list(range(11))
    assert [r[:5] for r in s.sa.reactions] == [(0, 0, 0, 10.0, 0), (1, 0, 0, 0.1, 0)]
    assert s.sa.u0.sum() == 100


def test_propensity_expressions_match_reference_strings():
    """The golden strings of test/unit_tests/test_reaction.py:47-85, 401-462."""
    A, B, X = sp.Species("A", 1.0), sp.Species("B", 1.0), sp.Species("X", 1.0)
    k1 = sp.Parameter("k1", expression=1.0)
    assert sp.Reaction("r", reactants={A: 1}, products={}, rate=k1).propensity_expression() == "(k1*A)"
    assert sp.Reaction("r", reactants={X: 1, sp.Species("Y", 1.0): 1}, products={}, rate=k1) \
        .propensity_expression() == "(((k1*X)*Y)/vol)"
    assert sp.Reaction("r", reactants={X: 2}, products={}, rate=k1).propensity_expression() == \
        "(((k1*X)*(X-1))/vol)"
    assert sp.Reaction("r", reactants={}, products={A: 1}, rate=k1).propensity_expression() == "(k1*vol)"


def test_custom_propensities_lower_to_table_rows_or_programs():
    m = birth_death_model()
    m.add_reaction(sp.Reaction("custom", reactants={"Rabbits": 1}, products={}, propensity_function="k_death*Rabbits"))
    m.add_reaction(sp.Reaction("halve", reactants={"Rabbits": 1}, products={}, propensity_function="Rabbits/2"))
    sa = sp.Solver(m).lower()
    # the same C text as a mass-action first-order row runs from the table;
    # int division needs a program
    assert [r[0] for r in sa.reactions] == [0, 1, 1, 4]
    assert sa.custom_propensities and sa.irG.size == sa.n_reactions * (sa.n_reactions + sa.n_species)
    m.add_reaction(sp.Reaction("bad", reactants={"Rabbits": 1}, products={}, propensity_function="k_death*q"))
    with pytest.raises(sp.SimulationError, match="invalid names"):
        sp.Solver(m).lower()


@pytest.mark.gpu
def test_solver_io_t0_equals_u0():
    model = diffusion_debug_model()
    result = model.run(seed=3)
    A = result.get_species("A", 0)
    assert not (A - model.u0[0]).any()


@pytest.mark.gpu
def test_same_seed_same_result_and_different_seeds_differ():
    model = diffusion_debug_model()
    solver = sp.Solver(model)
    r1 = solver.run(seed=1)
    r2 = solver.run(seed=1)
    r3 = solver.run(seed=100)
    assert r1 == r2
    assert r1 != r3
    assert model.run(seed=1) == model.run(seed=1)


@pytest.mark.gpu
def test_exact_replay_solver_matches_oracle_birth_death(oracle):
    model = birth_death_model()
    s = sp.Solver(model, exact_replay=True)
    res = s.run(seed=7)
    o = oracle.Oracle(s.sa)
    ref = o.run(seed=7)
    assert np.array_equal(res.data["xx"], ref["xx"])
    counts = res.get_species("Rabbits")
    assert counts.shape == (11, 121) and counts[0].sum() == 100


@pytest.mark.gpu
def test_ensemble_length():
    model = birth_death_model()
    res = model.run(number_of_trajectories=3, seed=2)
    assert len(res) == 3


@pytest.mark.gpu
def test_result_export_vtk_matches_oracle_files(tmp_path, oracle):
    """Result.export_vtk writes the reference format; on the exact-replay
    path its files equal those written from the FP64 oracle's run."""
    from spatialpy_amd import vtk
    from spatialpy_amd.lowering import output_schedule
    model = birth_death_model()
    s = sp.Solver(model, exact_replay=True)
    res = s.run(seed=7)
    res.export_vtk(str(tmp_path / "gpu"))
    o = oracle.Oracle(s.sa)
    ref = o.run(seed=7)
    steps = output_schedule(s.sa.output_steps, s.sa.nt)
    vtk.write_result(str(tmp_path / "cpu"), ref, s.sa, ["Rabbits"], steps)
    import os
    files = sorted(os.listdir(tmp_path / "cpu"))
    assert len(files) == 12
    for f in files:
        assert (tmp_path / "gpu" / f).read_text() == (tmp_path / "cpu" / f).read_text(), f


def test_place_initial_condition_follows_closest_vertex():
    """Domain.closest_vertex (core/domain.py:460-477) skips a vertex at
    distance exactly 0: a PlaceInitialCondition at a vertex's coordinates
    lands on its nearest other vertex (ties: the lower index)."""
    from spatialpy_amd.core.model import closest_vertex
    x = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.4, 0.0, 0.0]])
    assert closest_vertex(x, [0.0, 0.0, 0.0]) == 3
    assert closest_vertex(x, [0.1, 0.0, 0.0]) == 0
    assert closest_vertex(x, [0.4, 0.0, 0.0]) == 0       # itself skipped; 0 at 0.4 beats 1 at 0.6
    assert closest_vertex(x[:3], [0.0, 0.0, 0.0]) == 1   # 1 and 2 tie at distance 1
    model = birth_death_model()
    model.listOfInitialConditions = [sp.PlaceInitialCondition("Rabbits", 5, model.domain.vertices[0])]
    model.compile_prep()
    assert model.u0[0].sum() == 5 and model.u0[0, 0] == 0
