"""The HIP engine against the program SpatialPy itself generates (the
fixtures of tests/golden/make_generated_fixtures.py, see
tests/test_generated_program_cpu.py for what they hold).

The model is built with this package's Model API and lowered by its Solver;
the engine runs in replay mode on the std::mt19937_64 stream of the
program's seed (exe -s SEED) and must print the program's event sequence
(every event: time at %e, voxel, reaction or destination; checked by count,
field by field for the first 4,000 and by SHA-256 over all), its reaction
and diffusion totals, and its VTK files -- byte for byte on static domains,
except where the program reads undefined memory; on moving domains every
count exactly and positions, velocities and densities to one unit in their
last printed place (genfix.compare_vtk)."""
import pytest

import genfix
from spatialpy_amd import _abi
from spatialpy_amd.engine import Engine, mt19937_64_stream

pytestmark = pytest.mark.gpu

FIELDS = ("xx", "C", "x", "v", "rho", "bvf_phi", "nu")


@pytest.mark.parametrize("name", genfix.CASES)
def test_engine_reproduces_the_generated_program(name, tmp_path):
    fx = genfix.load(name)
    sa, model = genfix.lower(name)
    if not sa.static_domain:
        # lattices hold exact r^2 ties at step 0: the reference's list order
        # among them is the kd-tree's (static domains always use it)
        sa.flags |= _abi.FLAG_ANN_ORDER_MOVING
    cap = int(fx["n_events"]) + 10
    print(f"case {name}: n={sa.n} events={cap - 10}", flush=True)
    # at most 4 words per event (2 uniforms, 2 exponentials) plus the
    # per-step queue builds of moving domains
    words, expo = mt19937_64_stream(genfix.GM.RUN_SEED[name], 5 * cap + 50 * sa.n * int(sa.nt) + 1000)
    eng = Engine(sa, trace_capacity=cap)
    res = eng.run(words=words, expo=expo, fields=FIELDS)
    ev, n = eng.trace(cap)
    genfix.check_events(name, fx, ev, n)
    st = eng.stats()
    assert (st.reactions, st.diffusions) == (int(fx["n_reactions"]), int(fx["n_diffusions"]))
    genfix.write_vtk(str(tmp_path), name, res, sa, model)
    bad = genfix.compare_vtk(str(tmp_path), name, sa, printed_unit=not sa.static_domain)
    assert bad == [], repr(bad)
    eng.close()
