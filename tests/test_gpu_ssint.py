"""GPU parity of the Lennard-Jones surface-surface interactions (ffh_blob_set_ssint) against
the oracle (pinned in tests/test_oracle_ssint.py by the reference's cubes_lj energy test).

On the reference's two-cube fixture (second-order meshes: the surface triangles have edge
midpoints as corners) with MassLumpedSolver (no reductions in the solve), dt = 1 fs and noise
on: force vectors, positions and interaction energies bit-exact step by step, for the cubes_lj
set-up (top faces active, other blob only), every face active with self-interaction, lookup
grids of one and two cells per axis (neighbour offsets wrapping onto the same cell: pairs met
several times), es_update 3 (geometry and grid held between refreshes), ljsteric with the
cubes 0.6 nm apart (its soft core below Rmin) and with every face of both cubes active and
self-interaction on (its in-front and shared-node tests), and gensoft (Rmin the cutoff, k0).  The massless
solver: force vector bit-exact, positions within 1e-9.  A centroid outside the grid fails the
step, as in the oracle.  Graph replay equals ordinary launches.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CUBES = os.path.join(GOLDEN, "physics", "cubes_lj", "cube_veryFine.ffea")


def cubes(oracle, solver="masslumped", seed=11, all_faces=False, shift=0.0):
    from ffea_amd import formats as F, units as U
    from ffea_amd.blob import Batch
    from ffea_amd.world import linearised, place_blobs
    s = F.read_script(CUBES)
    prm = U.SimParams.from_si(dt=1e-15, epsilon=1e-6, calc_noise=1)
    meshes, scales = [], []
    for b in s["blobs"]:
        meshes.append(F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin")))
        scales.append(U.blob_scale(float(b["scale"])))
    pos = place_blobs(s["params"], s["blobs"], [linearised(m.pos * sc, m.topo) for m, sc in zip(meshes, scales)])
    # midpoints as averages of the placed vertices, as the device evaluates them (the reference
    # keeps the shifted loaded midpoints until the first step's end)
    pos[1][:, 1] += shift * 1e-9 / U.MESO_LENGTH
    pos = [linearised(p, m.topo) for p, m in zip(pos, meshes)]
    sd = oracle.package_seed(seed)
    oracle.set_package_seed(sd)
    st = oracle.stream()
    bo, faces, types, opps = [], [], [], []
    for b, m, sc, p in zip(s["blobs"], meshes, scales, pos):
        o = oracle.blob(m, prm, sc)
        o.set_solver(solver)
        o.set_positions(p.ravel())
        surf = F.read_surface(b["surface"])
        fc = surf[:, 1:4]
        op = F.surface_opposite(surf, m.topo)
        ty = F.read_ssint_types(b["vdw"], fc.shape[0])
        if all_faces:
            ty = np.ones_like(ty)
        o.set_surface_forces(fc, [], np.zeros((0, 3)))
        o.set_ssint_types(ty, op)
        bo.append(o)
        faces.append(fc)
        types.append(ty)
        opps.append(op)
    g = Batch(meshes, prm, scales, rng_state=sd)
    g.set_positions(np.concatenate(pos))
    g.set_solver(solver)
    off = g.node_off
    gf = np.concatenate([fc + off[k] for k, fc in enumerate(faces)])
    gop = np.concatenate([op + off[k] for k, op in enumerate(opps)])
    lj = os.path.join(os.path.dirname(CUBES), s["params"]["vdw_forcefield_params"])
    E, R, _ = F.read_ssint_params(lj, "lennard-jones")
    return g, bo, st, gf, np.concatenate(types), E, R, 3e-9 / U.MESO_LENGTH, gop


@pytest.fixture
def ssint_off(oracle):
    yield
    oracle.set_ssint()


def run_both(oracle, g, bo, st, nsteps, exact=True):
    import pyoracle
    xs = np.concatenate([o.get("pos") for o in bo])
    for k in range(nsteps):
        assert pyoracle.world_step(oracle, bo, st) == 0
        # the split step, as the World driver runs a check step: the energy of this step's
        # solve is read between the element and the position phase
        g.update_internal_forces()
        e = g.ssint_energy()
        g.update_positions()
        print("step", k, flush=True)
        rhs = np.concatenate([x.ravel() for x in g.split(g.rhs())])
        want = np.concatenate([o.get("rhs") for o in bo])
        if exact or k == 0:
            assert np.array_equal(rhs, want), k
            assert e == oracle.ssint_energy(), k
        else:
            # the massless solver's positions after the first step agree to rounding only (CG
            # dot products summed in tree order), so the energy of those positions does too
            assert e == pytest.approx(oracle.ssint_energy(), rel=1e-9, abs=0), k
        g.step(1)  # and an ordinary step
        assert pyoracle.world_step(oracle, bo, st) == 0
        if exact:
            assert np.array_equal(np.concatenate([x.ravel() for x in g.split(g.rhs())]),
                                  np.concatenate([o.get("rhs") for o in bo])), k
        got = np.concatenate([x.ravel() for x in g.split(g.positions())])
        po = np.concatenate([o.get("pos") for o in bo])
        if exact:
            assert np.array_equal(got, po), k
        else:
            assert np.linalg.norm(got - po) <= 1e-9 * np.linalg.norm(po - xs), k
    assert np.array_equal(g.rng_state(), st.state())


@pytest.mark.parametrize("case", ["cubes_lj", "all_self", "grid1", "grid2", "es3", "ljsteric_close",
                                  "ljsteric_self", "gensoft"])
def test_trajectory(oracle, ssint_off, case):
    g, bo, st, faces, types, E, R, h, opp = cubes(oracle, all_faces=case in ("all_self", "ljsteric_self"),
                                                  shift=-0.4 if case == "ljsteric_close" else 0.0)
    grid, cell, inc, es = (10, 10, 10), h, 0, 1
    kind, K = "lennard-jones", np.zeros_like(E)
    if case in ("all_self", "ljsteric_self"):
        inc = 1
    if case.startswith("ljsteric"):
        kind = "ljsteric"
    if case == "gensoft":  # Rmin is the cutoff; a spring constant as well
        kind, R, K = "gensoft", np.full_like(R, h), np.full_like(E, 0.02)
    elif case == "grid1":
        grid, cell = (1, 1, 1), 1e3
    elif case == "grid2":
        grid, cell = (2, 2, 2), 44.0  # faces in cells 1 and 2 (wrapped to 0) per axis
    elif case == "es3":
        es = 3
    print(case, int((types >= 0).sum()), "active faces", flush=True)
    oracle.set_ssint(E, R, grid, cell, inc_self=inc, es_update=es, kind=kind, k0=K)
    g.set_ssint(faces, types, E, R, grid, cell, inc, es, kind=kind, opposite=opp, k0=K)
    run_both(oracle, g, bo, st, 4 if case != "es3" else 5)


def test_massless(oracle, ssint_off):
    g, bo, st, faces, types, E, R, h, _ = cubes(oracle, solver="CG_nomass")
    oracle.set_ssint(E, R, (10, 10, 10), h, inc_self=0, es_update=1)
    g.set_ssint(faces, types, E, R, (10, 10, 10), h, 0, 1)
    run_both(oracle, g, bo, st, 3, exact=False)


def test_grid_overflow_fails(oracle, ssint_off):
    import pyoracle
    from ffea_amd import _lib
    g, bo, st, faces, types, E, R, h, _ = cubes(oracle)
    oracle.set_ssint(E, R, (10, 10, 10), 1.0, inc_self=0, es_update=1)
    assert pyoracle.world_step(oracle, bo, st) == -2
    g.set_ssint(faces, types, E, R, (10, 10, 10), 1.0, 0, 1)
    with pytest.raises(_lib.FFHError):
        g.step(1)


def test_graph_replay(oracle):
    g, _, _, faces, types, E, R, h, _ = cubes(oracle, seed=3)
    a = cubes(oracle, seed=3)[0]
    for x in (a, g):
        x.set_ssint(faces, types, E, R, (10, 10, 10), h, 0, 2)
    a.set_graphs(False)
    for k in range(3):
        a.step(2)
        g.step(2)
        assert np.array_equal(a.positions(), g.positions()), k
    for x in (a, g):
        x.set_ssint()
    a.step(1)
    g.step(1)
    assert np.array_equal(a.positions(), g.positions())
