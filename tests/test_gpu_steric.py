"""GPU parity of the steric volume exclusion (ffh_steric.h: FFH_SSINT_STERIC, and ljsteric /
gensoft where face elements overlap) against the oracle, whose geometry is pinned in
tests/test_oracle_steric.py by the reference's compiled tet_a_tetII / volumeIntersectionII, its
known answer and closed forms.

- The two sphere_63_120 blobs of the reference's tests/physics/squidgy_steric pushed 3 meso
  into each other (the face elements at the rim of the contact overlap), MassLumpedSolver (no
  reductions in the solve), noise on, steric_factor 2e3 as the script sets it: the force
  vectors, positions and the interaction energy bit-exact against the oracle step by step, for
  steric, ljsteric (Lennard-Jones elsewhere) and gensoft; apart (no overlap) steric adds
  nothing.  Graph replay equals ordinary launches.
- The reference's own squidgy_steric run through the World driver (mass-ful CG, the blobs at
  7000 m/s towards each other, steric only, 25,000 steps), with its checker's criteria
  (tests/physics/squidgy_steric/checkSquidgySpheres.cpp): the centres' distance turns from
  falling to rising (the blobs bounce), at most 0.5 % of the element pairs over all frames
  intersect (the oracle's volume, as the checker uses the reference's), and neither blob turns
  by more than 3 degrees.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN
from test_oracle_steric import lj_tables, spheres

pytestmark = pytest.mark.gpu


@pytest.fixture
def ssint_off(oracle):
    yield
    oracle.set_ssint()


def batch(oracle, c, kind, E, R, K, seed_vec, es_update=1):
    from ffea_amd.blob import Batch
    g = Batch(c["meshes"], c["prm"], c["scales"], rng_state=seed_vec)
    g.set_positions(np.concatenate(c["pos"]))
    g.set_solver("masslumped")
    off = g.node_off
    faces = np.concatenate([fc + off[k] for k, fc in enumerate(c["faces"])])
    opp = np.concatenate([op + off[k] for k, op in enumerate(c["opp"])])
    g.set_steric(c["factor"], 5e-3)
    g.set_ssint(faces, np.concatenate(c["types"]), E, R, c["grid"], c["h"], 0, es_update, kind=kind, opposite=opp,
                k0=K)
    return g


def tables(kind, h):
    if kind == "steric":
        return np.zeros((1, 1)), np.zeros((1, 1)), None
    E, R = lj_tables()
    if kind == "gensoft":  # Rmin is the cutoff
        return E, np.full_like(R, h), np.full_like(E, 0.02)
    return E, R, None


@pytest.mark.parametrize("kind,shift", [("steric", 3.0), ("steric", 0.0), ("ljsteric", 3.0), ("gensoft", 3.0)])
def test_spheres_against_oracle(oracle, ssint_off, kind, shift):
    """At shift 3.0 the first step has steric pairs, at 0.0 none
    (tests/test_oracle_steric.py::test_spheres_in_contact, the same positions)."""
    from test_gpu_ssint import run_both
    c = spheres(oracle, shift)
    E, R, K = tables(kind, c["h"])
    sd = oracle.package_seed(c["prm"].rng_seed)
    oracle.set_package_seed(sd)
    st = oracle.stream()
    oracle.set_ssint(E, R, c["grid"], c["h"], inc_self=0, es_update=1, kind=kind, k0=K, steric_factor=c["factor"])
    g = batch(oracle, c, kind, E, R, K, sd)
    print(kind, shift, flush=True)
    run_both(oracle, g, c["blobs"], st, 3)


def test_spheres_graph_replay(oracle):
    from ffea_amd.blob import package_seed
    c = spheres(oracle, 3.0)
    E, R, K = tables("ljsteric", c["h"])
    sd = package_seed(c["prm"].rng_seed)
    a, g = (batch(oracle, c, "ljsteric", E, R, K, sd, es_update=2) for _ in range(2))
    a.set_graphs(False)
    for k in range(3):
        a.step(2)
        g.step(2)
        assert np.array_equal(a.positions(), g.positions()), k


def kabsch_angle(p, q):
    """The rotation angle (degrees) of the best-fit rotation taking point set p to q."""
    p, q = p - p.mean(axis=0), q - q.mean(axis=0)
    u, _, vt = np.linalg.svd(p.T @ q)
    d = np.sign(np.linalg.det(u @ vt))
    r = (u @ np.diag([1, 1, d]) @ vt).T
    return np.degrees(np.arccos(np.clip((np.trace(r) - 1) / 2, -1, 1)))


def test_squidgy_spheres_bounce(oracle, tmp_path):
    from ffea_amd import formats as F, units as U
    from ffea_amd.world import World, read_trajectory
    shutil.copytree(os.path.join(GOLDEN, "physics", "sphere_63_120_structure"), tmp_path / "sphere_63_120_structure")
    run = tmp_path / "squidgy_steric"
    shutil.copytree(os.path.join(GOLDEN, "physics", "squidgy_steric"), run)
    w = World(str(run / "sphere_63_120.ffea"))
    w.run()
    frames = read_trajectory(str(run / "sphere_63_120_trajectory.ftj"))
    print("frames", len(frames), flush=True)
    assert len(frames) >= 300
    topo = [m.topo[:, :4] for m in w.meshes]
    d, hits, checks = [], 0, 0
    for step, blobs, _ in frames:
        p = [b[:, :3] / U.MESO_LENGTH for b in blobs]
        d.append(np.linalg.norm(p[1].mean(axis=0) - p[0].mean(axis=0)))
        hits += oracle.count_tet_intersections(p[0][topo[0]], p[1][topo[1]])
        checks += topo[0].shape[0] * topo[1].shape[0]
    d = np.array(d)
    bounce = np.flatnonzero(np.diff(d) >= 0)
    print(f"closest {d.min():.3f} meso at frame {d.argmin()}, first rise at frame {bounce[:1]}, "
          f"intersecting pairs {hits} of {checks}", flush=True)
    assert bounce.size > 0 and d[-1] > d.min()
    assert hits / checks <= 0.005
    p0 = [b[:, :3] for b in frames[0][1]]
    p1 = [b[:, :3] for b in frames[-2][1]]
    for k in range(2):
        assert kabsch_angle(p0[k], p1[k]) < 3.0, k
