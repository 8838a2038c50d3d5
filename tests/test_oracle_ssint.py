"""Lennard-Jones surface-surface interactions (VdW_solver with ssint_type lennard-jones,
src/VdW_solver.cpp:158-384, 663-731) in the oracle.  VdW_solver, Face and the lookup grid need
Blob and World, which do not compile here (Eigen), so they are pinned by:

- the reference's own physics test tests/physics/cubes_lj (checkE.py): two unit cubes whose top
  faces (200 active triangles each) face each other 1 nm apart, Emin = 2.537e17 J/m^4,
  Rmin = 1 nm; the double surface integral of E(1/r^12 - 2/r^6) gives -1.7870e-19 J (CUBA), and
  the measured SurfSurfEnergy must be within 1 %;
- Newton's third law: every pair adds equal and opposite totals to its two faces, so the
  interaction forces sum to zero over all nodes;
- the lookup grid only decides which pairs are considered: with every face in range of every
  other (a grid of one cell per axis, whose 27 neighbours are all that cell) the energy is that
  of all pairs, each counted 27 times;
- inc_self_ssint 0 leaves a blob alone; faces whose normals point the same way do not interact.
"""
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN

CUBES = os.path.join(GOLDEN, "physics", "cubes_lj", "cube_veryFine.ffea")


def cubes_world(oracle):
    from ffea_amd import formats as F, units as U
    from ffea_amd.world import linearised, place_blobs
    s = F.read_script(CUBES)
    prm = F.params_from_script(s["params"])
    meshes, scales = [], []
    for b in s["blobs"]:
        meshes.append(F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin")))
        scales.append(U.blob_scale(float(b["scale"])))
    pos = place_blobs(s["params"], s["blobs"], [linearised(m.pos * sc, m.topo) for m, sc in zip(meshes, scales)])
    blobs = []
    for b, m, sc, p in zip(s["blobs"], meshes, scales, pos):
        o = oracle.blob(m, prm, sc)
        o.set_positions(p.ravel())
        surf = F.read_surface(b["surface"])
        faces = surf[:, 1:4]
        o.set_surface_forces(faces, [], np.zeros((0, 3)))
        o.set_ssint_types(F.read_ssint_types(b["vdw"], faces.shape[0]), F.surface_opposite(surf, m.topo))
        blobs.append(o)
    lj = os.path.join(os.path.dirname(CUBES), s["params"]["vdw_forcefield_params"])
    E, R, _ = F.read_ssint_params(lj, "lennard-jones")
    cutoff = 3e-9 / U.MESO_LENGTH
    return blobs, E, R, cutoff, s


@pytest.fixture
def ssint_off(oracle):
    yield
    oracle.set_ssint()


def test_cubes_lj_energy(oracle, ssint_off):
    from ffea_amd import units as U
    blobs, E, R, h, s = cubes_world(oracle)
    oracle.set_ssint(E, R, (10, 10, 10), h, inc_self=0, es_update=1)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    assert pyoracle.world_step(oracle, blobs, oracle.stream()) == 0
    e = oracle.ssint_energy() * U.MESO_ENERGY
    assert abs(e - -1.7870e-19) / 1.7870e-19 < 1e-2, e


def interaction_rhs(oracle, blobs, E, R, grid, h, inc_self):
    """The force vectors handed to the solvers with and without the interaction (noise off)."""
    out = []
    pos = [b.get("pos").copy() for b in blobs]
    for on in (False, True):
        for b, p in zip(blobs, pos):
            b.set_positions(p)
        if on:
            oracle.set_ssint(E, R, grid, h, inc_self=inc_self, es_update=1)
        else:
            oracle.set_ssint()
        oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
        assert pyoracle.world_step(oracle, blobs, oracle.stream()) == 0
        out.append([b.get("rhs").reshape(-1, 3) for b in blobs])
    for b, p in zip(blobs, pos):
        b.set_positions(p)
    return out


def test_momentum_conserved(oracle, ssint_off):
    blobs, E, R, h, _ = cubes_world(oracle)
    for b in blobs:
        b.set_solver("masslumped")  # the force vector is the solver's input, pins zeroed
    (a0, a1), (b0, b1) = interaction_rhs(oracle, blobs, E, R, (10, 10, 10), h, 0)
    d = [b0 - a0, b1 - a1]
    tot = d[0].sum(axis=0) + d[1].sum(axis=0)
    scale = np.abs(d[0]).sum()
    assert scale > 0
    assert np.abs(tot).max() < 1e-9 * scale
    # the cubes attract: blob 0's interaction force points to blob 1 (+y), and back
    assert d[0].sum(axis=0)[1] > 0 and d[1].sum(axis=0)[1] < 0


def test_grid_only_selects_pairs(oracle, ssint_off):
    """One cell per axis: all 27 neighbour offsets wrap onto the same cell, so every pair is met
    27 times; the energy is 27 times the 10-cell grid's, whose cells of 3 nm already hold both
    top faces (1 nm apart) as neighbours."""
    blobs, E, R, h, _ = cubes_world(oracle)
    oracle.set_ssint(E, R, (10, 10, 10), h, inc_self=0, es_update=1)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    pos = [b.get("pos").copy() for b in blobs]
    assert pyoracle.world_step(oracle, blobs, oracle.stream()) == 0
    e10 = oracle.ssint_energy()
    for b, p in zip(blobs, pos):
        b.set_positions(p)
    oracle.set_ssint(E, R, (1, 1, 1), 1e3, inc_self=0, es_update=1)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    assert pyoracle.world_step(oracle, blobs, oracle.stream()) == 0
    assert oracle.ssint_energy() == pytest.approx(27 * e10, rel=1e-12)


def test_self_interaction_switch_and_normals(oracle, ssint_off):
    """With inc_self_ssint 0 a lone cube feels nothing; with 1 its active top face (coplanar
    triangles, normals all the same way) still feels nothing."""
    blobs, E, R, h, _ = cubes_world(oracle)
    one = blobs[:1]
    for inc in (0, 1):
        (a,), (b,) = interaction_rhs(oracle, one, E, R, (10, 10, 10), h, inc)
        assert np.array_equal(a, b)
        assert oracle.ssint_energy() == 0.0


# ---- the other types' pair potentials and checks (ljsteric's soft core, gensoft) ------------

W2 = 0.333333333333333 * 0.333333333333333


@pytest.mark.parametrize("kind", ["lennard-jones", "ljsteric", "gensoft"])
def test_pair_potentials(oracle, kind):
    """Closed forms of one Gauss-point pair's energy and force magnitude
    (src/VdW_solver.cpp:713-784): Lennard-Jones has its minimum -Emin (times both weights) at
    Rmin with zero force; ljsteric's soft core below Rmin joins it there with value and force,
    and is -Emin x^2 (3 - 2x) in x = r / Rmin; gensoft (Rmin = the cutoff) rises from -Emin at
    r = 0 to 0 at the cutoff, force zero at both.  The force magnitude is minus the energy's
    derivative for the first two and plus it for gensoft, as the reference writes them."""
    E, R, k0 = 0.7, 1.3, 0.4
    f, e = oracle.ssint_factors(kind, R, E, R, k0)
    assert f == pytest.approx(0.0, abs=1e-12)
    assert e == pytest.approx(0.0 if kind == "gensoft" else -W2 * E, abs=1e-12)
    for r in (0.3 * R, 0.8 * R, 1.2 * R, 1.7 * R):
        f, e = oracle.ssint_factors(kind, r, E, R, k0)
        h = 1e-6 * r
        de = (oracle.ssint_factors(kind, r + h, E, R, k0)[1] - oracle.ssint_factors(kind, r - h, E, R, k0)[1]) / (2 * h)
        sign = 1 if kind == "gensoft" else -1
        assert f == pytest.approx(sign * de / W2, rel=1e-6, abs=1e-9)
        if kind == "ljsteric" and r < R:
            x = r / R
            assert e == pytest.approx(-W2 * E * x * x * (3 - 2 * x), rel=1e-13)
            assert e < oracle.ssint_factors("lennard-jones", r, E, R)[1]
    if kind == "gensoft":
        assert oracle.ssint_factors(kind, 0.0, E, R, k0)[1] == pytest.approx(-W2 * E, rel=1e-13)


def test_ljsteric_is_lennard_jones_beyond_rmin(oracle, ssint_off):
    """With every Gauss-point distance above Rmin (the cubes 1.1 nm apart) ljsteric takes the
    Lennard-Jones branch everywhere and its extra tests reject nothing on facing planes: the
    same energy to the last bit; closer (0.6 nm) its soft core is below Lennard-Jones."""
    from ffea_amd import units as U
    out = {}
    for gap, kinds in ((0.1, ("lennard-jones", "ljsteric")), (-0.4, ("lennard-jones", "ljsteric"))):
        for kind in kinds:
            blobs, E, R, h, _ = cubes_world(oracle)
            p = blobs[1].get("pos").reshape(-1, 3)
            p[:, 1] += gap * 1e-9 / U.MESO_LENGTH
            blobs[1].set_positions(p.ravel())
            oracle.set_ssint(E, R, (10, 10, 10), h, inc_self=0, es_update=1, kind=kind)
            oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
            assert pyoracle.world_step(oracle, blobs, oracle.stream()) == 0
            out[gap, kind] = oracle.ssint_energy()
    assert out[0.1, "ljsteric"] == out[0.1, "lennard-jones"]
    assert out[-0.4, "ljsteric"] < out[-0.4, "lennard-jones"]


def test_steric_checks_exclude_neighbours(oracle, ssint_off):
    """With self-interaction on and every face of one cube active, ljsteric's shared-node test
    drops adjacent faces (which Lennard-Jones keeps, at tiny distances): a finite energy far
    below the Lennard-Jones one."""
    from ffea_amd import formats as F
    blobs, E, R, h, s = cubes_world(oracle)
    b = s["blobs"][0]
    m = F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"])
    surf = F.read_surface(b["surface"])
    opp = F.surface_opposite(surf, m.topo)
    one = blobs[:1]
    one[0].set_ssint_types(np.ones(surf.shape[0], np.int32), opp)
    res = {}
    for kind in ("lennard-jones", "ljsteric"):
        pos = one[0].get("pos").copy()
        oracle.set_ssint(E, R, (10, 10, 10), h, inc_self=1, es_update=1, kind=kind)
        oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
        assert pyoracle.world_step(oracle, one, oracle.stream()) == 0
        res[kind] = oracle.ssint_energy()
        one[0].set_positions(pos)
    assert np.isfinite(res["ljsteric"]) and abs(res["ljsteric"]) < abs(res["lennard-jones"])


@pytest.mark.parametrize("kind", ["lennard-jones", "ljsteric", "gensoft"])
def test_face_pair_force_spread_moments(oracle, ssint_off, kind):
    """One facing pair of faces, one active face per cube: the forces the pair leaves on each
    face's three corners (src/VdW_solver.cpp:288-320) against the Gauss-point forces
    F_kl = f(|p_k - q_l|) (p_k - q_l) / |p_k - q_l| (k <= l, F_lk = F_kl: :663-697), with
    p, q the 3-point rule's points on the two faces (eta = 2/3, 1/6, 1/6) and W = 1/3:
      sum_j f1_j = +Ap Aq W^2 sum_kl F_kl,   sum_j x1_j (x) f1_j = +Ap Aq W^2 sum_kl p_l (x) F_kl,
      sum_j f2_j = -Ap Aq W^2 sum_kl F_kl,   sum_j x2_j (x) f2_j = -Ap Aq W^2 sum_kl q_k (x) F_kl.
    Linear interpolation carries the Gauss points' moments to the corners, so a corner or
    Gauss index out of place, the faces' roles swapped or a wrong area moves these moments."""
    from ffea_amd import formats as F
    blobs, E, R, h, s = cubes_world(oracle)
    tops = [F.read_ssint_types(b["vdw"], F.read_surface(b["surface"]).shape[0]) for b in s["blobs"]]
    faces = [F.read_surface(b["surface"])[:, 1:4] for b in s["blobs"]]
    pos = [b.get("pos").reshape(-1, 3).copy() for b in blobs]
    cen = [p[fc].mean(axis=1) for p, fc in zip(pos, faces)]
    a = np.flatnonzero(tops[0] >= 0)[17]
    cand = np.flatnonzero(tops[1] >= 0)
    b = cand[np.argmin(np.linalg.norm(cen[1][cand] - cen[0][a], axis=1))]
    opps = [F.surface_opposite(F.read_surface(x["surface"]), F.load_blob(x["nodes"], x["topology"], x["material"],
                                                                      x["stokes"]).topo) for x in s["blobs"]]
    for o, t, keep, op in ((blobs[0], tops[0], a, opps[0]), (blobs[1], tops[1], b, opps[1])):
        one = np.full_like(t, -1)
        one[keep] = t[keep]
        o.set_ssint_types(one, op)
    k0 = None
    if kind == "gensoft":  # Rmin is the cutoff
        R, k0 = np.full_like(R, h), np.full_like(E, 0.02)
    oracle.set_ssint(E, R, (10, 10, 10), h, inc_self=0, es_update=1, kind=kind, k0=k0)
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    assert pyoracle.world_step(oracle, blobs, oracle.stream()) == 0
    f1, f2 = (o.get("face_force").reshape(-1, 4, 3)[k] for o, k in ((blobs[0], a), (blobs[1], b)))
    assert not f1[3].any() and not f2[3].any()  # the opposite vertices: steric forces only
    f1, f2 = f1[:3], f2[:3]
    x1, x2 = pos[0][faces[0][a]], pos[1][faces[1][b]]
    eta = np.full((3, 3), 1 / 6) + np.eye(3) / 2
    p, q = eta @ x1, eta @ x2
    t0, t1 = int(tops[0][a]), int(tops[1][b])
    Fkl = np.zeros((3, 3, 3))
    for k in range(3):
        for l in range(k, 3):
            d = p[k] - q[l]
            r = np.linalg.norm(d)
            fm = oracle.ssint_factors(kind, r, E[t0, t1], R[t0, t1], 0.0 if k0 is None else k0[t0, t1])[0]
            Fkl[k, l] = Fkl[l, k] = fm * d / r
    area = [0.5 * np.linalg.norm(np.cross(x[1] - x[0], x[2] - x[0])) for x in (x1, x2)]
    c = area[0] * area[1] / 9
    tot = c * Fkl.sum(axis=(0, 1))
    m1 = c * np.einsum("li,klj->ij", p, Fkl)
    m2 = -c * np.einsum("ki,klj->ij", q, Fkl)
    scale = np.abs(tot).max()
    assert scale > 0
    assert np.abs(f1.sum(axis=0) - tot).max() <= 1e-12 * scale
    assert np.abs(f2.sum(axis=0) + tot).max() <= 1e-12 * scale
    mscale = scale * np.abs(x1).max()
    assert np.abs(np.einsum("ji,jk->ik", x1, f1) - m1).max() <= 1e-11 * mscale
    assert np.abs(np.einsum("ji,jk->ik", x2, f2) - m2).max() <= 1e-11 * mscale
    # the moments do not follow from the totals: the corners differ from the centroid's share
    assert np.abs(f1 - f1.mean(axis=0)).max() > 1e-6 * scale
