"""The oracle against the golden vectors written by the reference's own compiled objects
(tests/golden/make_golden.py): element data, assembled CSR, CG solution and 3-step
trajectories must agree bit for bit."""
import glob
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN
from ffea_amd.mesh import Mesh
from ffea_amd.units import SimParams

CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
               if not os.path.basename(p).startswith("rng_"))


def load_case(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    p = g["params"]
    prm = SimParams(dt=p[0], kT=p[1], epsilon2=p[2], max_iterations_cg=int(p[3]),
                    calc_noise=int(p[4]), calc_stokes=int(p[5]), stokes_visc=p[6],
                    rng_seed=int(p[7]))
    m = Mesh(pos=g["rest"], topo=g["topo"], material=g["material"],
             stokes_radius=g["stokes_radius"], num_linear=int(np.unique(g["topo"][:, :4]).size),
             pinned=g["pinned"])
    return g, m, prm, float(g["scale"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden(oracle, name):
    g, m, prm, sc = load_case(name)
    oracle.set_package_seed(g["seed"])
    s = oracle.stream()
    b = oracle.blob(m, prm, sc, pos=g["x0"])
    for k in ("J_inv_0", "vol_0", "stokes_drag"):
        assert np.array_equal(b.get(k), g[k]), k
    assert np.array_equal(b.get("material"), g["material_meso"])
    inv, rhs = b.forces_only(s)
    assert inv == 0
    for k in ("dpsi", "vol", "visc", "node_force"):
        assert np.array_equal(b.get(k), g[k]), k
    assert np.array_equal(rhs, g["rhs"])
    assert np.array_equal(s.state(), g["rng_after_forces"])
    st, v = b.solve(rhs)
    assert st == 0
    key, col, val = b.csr()
    assert np.array_equal(key, g["csr_key"]) and np.array_equal(col, g["csr_col"])
    assert np.array_equal(val, g["csr_val"])
    assert np.array_equal(v, g["solve_v"])
    b.close()
    oracle.set_package_seed(g["seed"])
    s = oracle.stream()
    b = oracle.blob(m, prm, sc, pos=g["x0"])
    for k in range(g["traj"].shape[0]):
        assert pyoracle.world_step(oracle, [b], s) == 0
        assert np.array_equal(b.get("pos"), g["traj"][k]), k
    assert np.array_equal(b.get("vel"), g["vel"])
    assert np.array_equal(s.state(), g["rng_after_steps"])
