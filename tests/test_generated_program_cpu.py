"""The oracle and the lowering against the program SpatialPy itself
generates (fixtures from tests/golden/make_generated_fixtures.py: the model
built with spatialpy's API, written by Solver._Solver__create_propensity_file
(spatialpy/solvers/solver.py:100-158), compiled against the reference's
engine sources and run as `exe -t 1 -s SEED`).

* Lowering: the same model built with this package's Model API lowers to
  exactly the arrays the generator baked into the program (particles,
  u0, stoichiometry and dependency CSC, diffusion matrix, scalars, output
  steps, boundary conditions, data functions).
* The oracle on those arrays, seeded as the program, prints the same event
  sequence (time, voxel, reaction or destination: the program's debug-level-1
  lines, simulate_rdme.cpp:282-283, 405-413) and writes the same VTK files
  byte for byte (output.cpp), including positions at %.10e, velocities and
  densities at %lf and every count.  Static and moving domains, a boundary
  condition, two particle types, custom propensities and a data function.
  The only differences allowed are the program's undefined reads, listed in
  genfix.compare_vtk: leftover memory in the uninitialised F / Fbp of step 0
  (denormals on exactly-zero coordinates) and the chem-RHS read past
  input_N_dense (the C[...] fields of the custom-chemistry models).  Every
  count, event, velocity and density matches exactly."""
import pytest

import genfix

CASES = genfix.CASES


@pytest.mark.parametrize("name", CASES)
def test_lowering_equals_the_generated_program(name):
    sa, _ = genfix.lower(name)
    genfix.GM.check_lowering(name, sa, genfix.load(name))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_the_generated_program(name, oracle, tmp_path):
    fx = genfix.load(name)
    sa, model = genfix.lower(name)
    o = oracle.Oracle(sa)
    o.set_trace(int(fx["n_events"]) + 10)
    res = o.run(seed=genfix.GM.RUN_SEED[name])
    ev, n = o.trace(int(fx["n_events"]) + 10)
    genfix.check_events(name, fx, ev, n)
    st = o.stats()
    assert (st.reactions, st.diffusions) == (int(fx["n_reactions"]), int(fx["n_diffusions"]))
    genfix.write_vtk(str(tmp_path), name, res, sa, model)
    assert genfix.compare_vtk(str(tmp_path), name, sa) == []


@pytest.mark.parametrize("name", ["dam_break", "cylinder"])
def test_jacobi_order_oracle_stays_within_the_gpu_tolerance(name, oracle, tmp_path):
    """The GPU reads the neighbours' densities inside the boundary volume
    fraction after the whole substep (Jacobi order), the one-thread
    reference in index order (model.cpp:285-293).  The oracle in that order
    keeps the program's events and counts and stays inside the tolerance the
    GPU test allows (genfix.compare_vtk, printed_unit)."""
    fx = genfix.load(name)
    sa, model = genfix.lower(name)
    o = oracle.Oracle(sa, gauss_seidel=False)
    o.set_trace(int(fx["n_events"]) + 10)
    res = o.run(seed=genfix.GM.RUN_SEED[name])
    ev, n = o.trace(int(fx["n_events"]) + 10)
    genfix.check_events(name, fx, ev, n)
    genfix.write_vtk(str(tmp_path), name, res, sa, model)
    assert genfix.compare_vtk(str(tmp_path), name, sa, printed_unit=True) == []
