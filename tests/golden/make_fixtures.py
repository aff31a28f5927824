"""Write the golden fixtures under tests/golden/ by running SpatialPy's own C++
engine (oracle/_ref/ref_harness_O0, built from the reference sources by
`make -C oracle ref`) on small seeded models.

Run in the container that holds the reference:  python tests/golden/make_fixtures.py
Each fixture holds the model's inputs (generator name, parameters, seed and
the generated positions / initial counts) and the reference engine's
outputs: neighbour lists after the force substep (index, r, dW/dr, D_ij),
forces, snapshots of counts / positions / velocities / densities /
concentrations, and the reaction / diffusion event counters.
"""
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from spatialpy_amd import workloads, lowering  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness_O0")
ANN_TEST = "/root/reference/spatialpy/solvers/c_base/ssa_sdpd-c-simulation-engine/external/ANN/test"

CASES = {
    "c1_birth_death": dict(gen="birth_death_2d", params=dict(seed=1), seed=7, kw={}),
    "c2_static_rd_small": dict(gen="static_rd_3d",
                               params=dict(n_side=8, nA=800, nB=800, t_end=3, num_points=4, seed=2),
                               seed=11, kw={}),
    "c3_dam_break_small": dict(gen="dam_break_2d",
                               params=dict(nx=20, ny=20, spacing=0.025, n_mol=100, nt=45, out_every=15, dt=2e-4),
                               seed=5, kw=dict(force_step=25, nbr_step=25)),
    "c5_cylinder_small": dict(gen="cylinder_3d",
                              params=dict(n_target=1200, nt=10, out_every=5, seed=4, dt=2.5e-4),
                              seed=9, kw=dict(force_step=6, nbr_step=6)),
    # custom propensity functions (the reference compiles their generated
    # C++; the harness loads it from props.so)
    "c6_custom_chemistry": dict(gen="custom_chemistry_2d", params=dict(seed=3), seed=13, kw={}),
    "c6_custom_chemistry_libm": dict(gen="custom_chemistry_2d", params=dict(seed=3, libm=True), seed=13, kw={}),
}


def add_propensity_library(sa, d):
    """Compile the model's generated reaction functions the way the
    reference builds its model file (g++ -std=c++14, no optimisation) into
    d/props.so and point the harness at it."""
    if sa.vm_code is None or sa.prop_cpp is None:
        return
    src = os.path.join(d, "props.cpp")
    with open(src, "w") as f:
        f.write(lowering.harness_propensity_source(sa))
    subprocess.run(["g++", "-std=c++14", "-shared", "-fPIC", "-w", src, "-o", os.path.join(d, "props.so")],
                   check=True)
    with open(os.path.join(d, "meta.txt"), "a") as f:
        f.write("prop_lib props.so\n")


def make_case(name, spec):
    sa = getattr(workloads, spec["gen"])(**spec["params"])
    with tempfile.TemporaryDirectory() as d:
        lowering.write_harness_dir(sa, d, seed=spec["seed"], **spec["kw"])
        add_propensity_library(sa, d)
        r = subprocess.run([HARNESS, d], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: harness failed: {r.stdout[-500:]} {r.stderr[-500:]}")
        ld = lambda f, dt: np.fromfile(os.path.join(d, f), dtype=dt)
        out = dict(
            meta=json.dumps(dict(gen=spec["gen"], params=spec["params"], seed=spec["seed"], kw=spec["kw"])),
            x=sa.x, u0=sa.u0,
            nbr_rowptr=ld("out_nbr_rowptr.i64", np.int64), nbr_idx=ld("out_nbr_idx.i32", np.int32),
            nbr_dist=ld("out_nbr_dist.f64", np.float64), nbr_dwdr=ld("out_nbr_dwdr.f64", np.float64),
            nbr_dij=ld("out_nbr_dij.f64", np.float64),
            F=ld("out_F.f64", np.float64), Fbp=ld("out_Fbp.f64", np.float64),
            Frho=ld("out_Frho.f64", np.float64), Q=ld("out_Q.f64", np.float64),
            xx=ld("out_xx.u32", np.uint32), snap_x=ld("out_x.f64", np.float64),
            snap_v=ld("out_v.f64", np.float64), snap_rho=ld("out_rho.f64", np.float64),
            snap_C=ld("out_C.f64", np.float64), snap_steps=ld("out_steps.u32", np.uint32),
            counters=ld("out_counters.i64", np.int64),
        )
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(name, "n =", sa.n, "events =", out["counters"].tolist())


VTK_CASES = {
    "vtk_c1_birth_death": dict(gen="birth_death_2d", params=dict(seed=1, t_end=2, num_points=3), seed=7),
    "vtk_c3_dam_break_tiny": dict(gen="dam_break_2d",
                                  params=dict(nx=6, ny=6, spacing=0.08, n_mol=20, nt=4, out_every=2, dt=2e-4),
                                  seed=5),
}


def make_vtk_case(name, spec):
    """The reference engine's own VTK files (output_vtk__sync_step +
    output_vtk__async_step, run in the harness's working directory)."""
    sa = getattr(workloads, spec["gen"])(**spec["params"])
    dst = os.path.join(HERE, name)
    os.makedirs(dst, exist_ok=True)
    with tempfile.TemporaryDirectory() as d:
        lowering.write_harness_dir(sa, d, seed=spec["seed"], write_vtk=1)
        r = subprocess.run([HARNESS, d], capture_output=True, text=True, cwd=d)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: harness failed: {r.stdout[-500:]} {r.stderr[-500:]}")
        files = sorted(f for f in os.listdir(d) if f.endswith(".vtk"))
        for f in files:
            with open(os.path.join(d, f)) as a, open(os.path.join(dst, f), "w") as b:
                b.write(a.read())
    with open(os.path.join(dst, "case.json"), "w") as f:
        json.dump(dict(gen=spec["gen"], params=spec["params"], seed=spec["seed"], files=files), f)
    print(name, "n =", sa.n, "files", files)


def make_ann_test1():
    """ANN's own test 1: 20 2-D points, bucket 1, sliding midpoint ('suggest');
    test1.save records the tree's node counts and depth."""
    pts = np.loadtxt(os.path.join(ANN_TEST, "test1-data.pts"))
    q = np.loadtxt(os.path.join(ANN_TEST, "test1-query.pts"))
    save = open(os.path.join(ANN_TEST, "test1.save")).read()
    get = lambda key: int(re.search(key + r"\s*=\s*(\d+)", save).group(1))
    np.savez_compressed(os.path.join(HERE, "ann_test1.npz"), data=pts, query=q,
                        n_leaves=get("n_leaves"), n_splits=get("n_splits"), depth=get("depth"))
    print("ann_test1", pts.shape, get("n_leaves"), get("n_splits"), get("depth"))


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    only = sys.argv[1:]
    for name, spec in CASES.items():
        if not only or name in only:
            make_case(name, spec)
    for name, spec in VTK_CASES.items():
        if not only or name in only:
            make_vtk_case(name, spec)
    if not only:
        make_ann_test1()
