"""The steric volume exclusion in the oracle (ffea_oracle.c): the overlap test tet_a_tetII
(src/CheckTetrahedraOverlap.cpp:208-337), the intersection volume volumeIntersectionII
(src/VolumeIntersection.cpp:951-1024), the local coordinates (src/mat_vec_fns_II.cpp:725-733)
and the pair forces of Face::getTetraIntersectionVolumeTotalGradientAndShapeFunctions with
VdW_solver::do_steric_interaction (src/Face.cpp:473-514, src/VdW_solver.cpp:458-487).

Pinned by:
- the reference's own known answer, tests/consistency/volume/testVolume.cpp: the corner
  tetrahedron shifted by (0.05, 0.02, 0) against it shifted by (0, 0, 0.5) overlaps, with a
  common volume that prints as 0.020833333333333 at 14 digits; its second pair (tetC, tetD)
  has a nonzero common volume;
- closed forms: two translates T + a, T + b of the corner tetrahedron T (x, y, z >= 0,
  x + y + z <= 1) meet in the corner tetrahedron of side s = 1 + min(sum a, sum b) -
  sum max(a, b), volume s^3 / 6, under any linear map M times |det M|; that also gives the
  one-sided differences of the gradient in closed form; a tetrahedron inside another has its
  own volume as the common one (or the reference's bound 8 d^3 where that is smaller, d the
  distance from its vertices' mean to its last vertex); far apart ones neither overlap nor
  have a common volume;
- the local coordinates sum to 1 and reproduce the point; the pair forces sum to zero
  (Newton's third law) and push the tetrahedra apart;
- where oracle/_ref is built, the reference's own compiled tet_a_tetII, volumeIntersectionII
  and getLocalCoordinatesForLinTet: every result bit for bit on random pairs (shared vertices,
  touching faces, one inside the other, degenerate shifts included).
"""
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN

T0 = np.array([[0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, 1]], np.float64)  # testVolume.cpp's vertex order


def oriented(T):
    """The vertex order the overlap test expects, as a surface face and its opposite vertex
    give it: (T1 - T0) x (T2 - T0) points away from T3."""
    T = np.array(T, np.float64)
    if np.linalg.det(T[1:] - T[0]) > 0:
        T[[1, 2]] = T[[2, 1]]
    return T


def corner_volume(a, b):
    s = 1 + min(a.sum(), b.sum()) - np.maximum(a, b).sum()
    return max(s, 0.0) ** 3 / 6


def test_reference_known_answer(oracle):
    A = T0 + [0.05, 0.02, 0.0]
    B = T0 + [0.0, 0.0, 0.5]
    assert oracle.tets_overlap(A, B)
    v, cm = oracle.tet_intersection(A, B, want_cm=True)
    assert f"{v:.14g}" == "0.020833333333333"
    # the intersection is the corner tetrahedron of side 1/2 at (0.05, 0.02, 0.5): cm its
    # vertices' mean
    assert np.allclose(cm, [0.05 + 0.125, 0.02 + 0.125, 0.5 + 0.125], rtol=0, atol=1e-15)
    C = np.array([[5.16991, 5.11394, -2.59001], [2.68658, 5.11124, -5.03959], [5.12737, 5.10932, -5.0372],
                  [2.98162, 2.85562, -2.78657]])
    D = np.array([[4.8437, 2.47363, -5.0158], [4.84606, 5.11758, -2.44064], [7.20236, 5.11031, -5.00642],
                  [4.88715, 5.11134, -5.00755]])
    assert oracle.tet_intersection(C, D)[0] > 0


REGULAR = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64)


def random_map(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q @ np.diag(rng.uniform(0.5, 2.0, 3))


def test_translates_closed_form(oracle):
    rng = np.random.default_rng(3)
    n = 0
    for t in range(400):
        a, b = rng.uniform(-0.3, 0.3, 3), rng.uniform(-0.3, 0.3, 3)
        want = corner_volume(a, b)
        if want < 1e-4:
            continue
        M = random_map(rng) if t % 2 else np.eye(3)
        A, B = oriented((T0 + a) @ M.T), oriented((T0 + b) @ M.T)
        assert oracle.tets_overlap(A, B)
        v, _ = oracle.tet_intersection(A, B)
        assert v == pytest.approx(abs(np.linalg.det(M)) * want, rel=1e-11), t
        n += 1
    assert n > 200


def test_inside_and_apart(oracle):
    rng = np.random.default_rng(5)
    for t in range(100):
        B = oriented(REGULAR @ random_map(rng).T + rng.normal(size=(4, 3)) * 0.1 + rng.normal(size=3))
        c = B.mean(axis=0)
        A = oriented(c + 0.3 * (B - c) + rng.normal(size=3) * 0.01)
        bary = np.linalg.solve(np.vstack([B.T, np.ones(4)]), np.vstack([A.T, np.ones(4)]))
        assert (bary > 1e-3).all()  # every vertex of A well inside B
        vA = abs(np.linalg.det(A[1:] - A[0])) / 6
        # the reference's bound 8 d^3, d the distance from the points' mean to the last point
        # listed (A's vertices in order: maxVolume's loop leaves the last distance)
        vA = min(vA, 8 * np.linalg.norm(A.mean(axis=0) - A[3]) ** 3)
        assert oracle.tets_overlap(A, B) and oracle.tets_overlap(B, A)
        assert oracle.tet_intersection(A, B)[0] == pytest.approx(vA, rel=1e-10)
        assert oracle.tet_intersection(B, A)[0] == pytest.approx(vA, rel=1e-10)
        far = B + [0, 0, 10 * np.ptp(B)]
        assert not oracle.tets_overlap(B, far)
        assert oracle.tet_intersection(B, far)[0] == 0.0


def test_local_coordinates(oracle):
    rng = np.random.default_rng(7)
    for _ in range(50):
        T = rng.normal(size=(4, 3))
        p = rng.normal(size=3)
        phi = oracle.tet_local_coords(T, p)
        assert phi.sum() == pytest.approx(1.0, abs=1e-12)
        assert np.allclose(phi @ T, p, atol=1e-12)


def test_pair_forces_closed_form(oracle):
    """Translates of the corner tetrahedron: V, the gradient's one-sided differences from the
    closed form, the forces phi_j g on A and -phi_j g on B with phi the local coordinates of
    the intersection's vertex mean, energy factor V."""
    rng = np.random.default_rng(11)
    dr, fac = 5e-3, 2e3
    n = 0
    for t in range(200):
        a, b = rng.uniform(-0.25, 0.25, 3), rng.uniform(-0.25, 0.25, 3)
        V = corner_volume(a, b)
        if V < 1e-3:
            continue
        A, B = T0 + a, T0 + b
        rc, e, fA, fB = oracle.steric_forces(A, B, dr, fac)
        assert rc == 1
        assert e == pytest.approx(fac * V, rel=1e-11)
        g = np.zeros(3)
        for d in range(3):
            dx = np.eye(3)[d] * dr
            g[d] = (corner_volume(a, b + dx) - V) / dr - (corner_volume(b, a + dx) - V) / dr
        g *= fac
        m = np.maximum(a, b)
        s = 1 + min(a.sum(), b.sum()) - m.sum()
        cm = m + s / 4  # the intersection's vertices m, m + s e_i
        phi1, phi2 = oracle.tet_local_coords(A, cm), oracle.tet_local_coords(B, cm)
        scale = np.abs(g).max()
        assert np.allclose(fA, np.outer(phi1, g), rtol=0, atol=1e-7 * scale)
        assert np.allclose(fB, -np.outer(phi2, g), rtol=0, atol=1e-7 * scale)
        assert np.abs(fA.sum(axis=0) + fB.sum(axis=0)).max() <= 1e-12 * scale
        # repulsion: the force on A points from B's centroid towards A's
        assert fA.sum(axis=0) @ (A.mean(axis=0) - B.mean(axis=0)) > 0
        n += 1
    assert n > 50


def ref_or_skip():
    if not pyoracle.have_reference():
        pytest.skip("oracle/_ref is not built here")
    return pyoracle.Reference()


def random_pairs(n, seed):
    rng = np.random.default_rng(seed)
    for t in range(n):
        A = rng.normal(size=(4, 3))
        B = A.mean(axis=0) + rng.normal(size=(4, 3)) * rng.uniform(0.3, 1.5)
        k = t % 6
        if k == 1:
            B[0] = A[0]  # a shared vertex
        elif k == 2:
            B[:3] = A[:3]  # a shared face, opposite vertices on either side or not
            B[3] = A[:3].mean(axis=0) - (A[3] - A[:3].mean(axis=0)) * rng.uniform(-1, 1)
        elif k == 3:
            B = A + rng.normal(size=3) * 1e-3  # a small shift
        elif k == 4:
            B = A.copy()  # identical
        yield A, B


def test_against_reference_objects(oracle):
    ref = ref_or_skip()
    n = ov = 0
    for A, B in random_pairs(3000, 1):
        assert oracle.tets_overlap(A, B) == ref.tets_overlap(A, B)
        v1, c1 = oracle.tet_intersection(A, B, True)
        v2, c2 = ref.tet_intersection(A, B, True)
        assert np.array_equal([v1], [v2], equal_nan=True) and np.array_equal(c1, c2, equal_nan=True)
        p = A.mean(axis=0) + 0.1
        assert np.array_equal(oracle.tet_local_coords(A, p), ref.tet_local_coords(A, p))
        r1, r2 = oracle.steric_forces(A, B, 5e-3, 1.7), ref.steric_forces(A, B, 5e-3, 1.7)
        assert r1[0] == r2[0]
        if r1[0] == 1:
            assert np.array_equal([r1[1]], [r2[1]], equal_nan=True)
            assert np.array_equal(r1[2], r2[2], equal_nan=True) and np.array_equal(r1[3], r2[3], equal_nan=True)
        n += 1
        ov += r1[0] == 1 and np.isfinite(r1[1])
    assert ov > 500


# ---- world level: the reference's squidgy_steric spheres pushed into each other -------------

SQUIDGY = os.path.join(GOLDEN, "physics", "squidgy_steric", "sphere_63_120.ffea")


def lj_tables():
    """A one-type Lennard-Jones table from the cubes_lj fixture's parameters (meso units)."""
    from ffea_amd import formats as F
    E, R, _ = F.read_ssint_params(os.path.join(GOLDEN, "physics", "cubes_lj", "veryFine.lj"), "lennard-jones")
    return E[:1, :1].copy(), R[:1, :1].copy()


def spheres(oracle, shift, solver="masslumped", noise=1, seed=7):
    """The two sphere_63_120 blobs of tests/physics/squidgy_steric (placed as its script places
    them, 0.5 nm apart), each moved `shift` meso towards the other; midpoints as the averages
    of their edge ends.  Returns the oracle blobs, their faces, types and opposite vertices, the
    positions, the meshes, scales and parameters, and the script's lookup grid."""
    from ffea_amd import formats as F, units as U
    from ffea_amd.world import linearised, place_blobs
    s = F.read_script(SQUIDGY)
    prm = U.SimParams.from_si(dt=1e-15, epsilon=1e-6, calc_noise=noise, rng_seed=seed)
    meshes = [F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin")) for b in s["blobs"]]
    scales = [U.blob_scale(float(b["scale"])) for b in s["blobs"]]
    pos = place_blobs(s["params"], s["blobs"], [linearised(m.pos * sc, m.topo) for m, sc in zip(meshes, scales)])
    out = dict(blobs=[], faces=[], types=[], opp=[], pos=[], meshes=meshes, scales=scales, prm=prm)
    for k, (b, m, sc, p) in enumerate(zip(s["blobs"], meshes, scales, pos)):
        q = p.copy()
        q[:, 0] += shift if k == 0 else -shift
        q = linearised(q, m.topo)
        o = oracle.blob(m, prm, sc)
        o.set_solver(solver)
        o.set_positions(q.ravel())
        surf = F.read_surface(b["surface"])
        fc, op = surf[:, 1:4], F.surface_opposite(surf, m.topo)
        ty = F.read_ssint_types(b["vdw"], fc.shape[0])
        o.set_surface_forces(fc, [], np.zeros((0, 3)))
        o.set_ssint_types(ty, op)
        for key, v in (("blobs", o), ("faces", fc), ("types", ty), ("opp", op), ("pos", q)):
            out[key].append(v)
    out["grid"] = [int(s["params"][f"es_N_{a}"]) for a in "xyz"]
    out["h"] = float(s["params"]["vdw_cutoff"]) / U.MESO_LENGTH
    out["factor"] = float(s["params"]["vdw_steric_factor"])
    return out


@pytest.fixture
def ssint_off(oracle):
    yield
    oracle.set_ssint()


@pytest.mark.parametrize("shift", [2.0, 3.0])
def test_spheres_in_contact(oracle, ssint_off, shift):
    """The spheres start 3 meso apart.  Moved 2.0 each they overlap by 1 meso, but the faces
    that cross are no longer in front of each other (consider_interaction) and no pair
    applies; moved 3.0 each the face elements at the rim of the contact overlap: a positive
    energy, forces on both spheres that sum to zero (each pair's local coordinates sum to 1 on
    both sides) and push them apart, the opposite vertices loaded too."""
    import pyoracle
    c = spheres(oracle, shift, noise=0)
    oracle.set_ssint(np.zeros((1, 1)), np.zeros((1, 1)), c["grid"], c["h"], inc_self=0, es_update=1, kind="steric",
                     steric_factor=c["factor"])
    oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
    assert pyoracle.world_step(oracle, c["blobs"], oracle.stream()) == 0
    ff = [o.get("face_force").reshape(-1, 4, 3) for o in c["blobs"]]
    e = oracle.ssint_energy()
    if shift == 2.0:
        assert e == 0.0 and not ff[0].any() and not ff[1].any()
        return
    assert e > 0
    t0, t1 = ff[0].sum(axis=(0, 1)), ff[1].sum(axis=(0, 1))
    assert np.abs(t0 + t1).max() <= 1e-12 * np.abs(t0).max()
    assert t0[0] < 0 < t1[0]  # blob 0 sits at lower x
    assert (np.abs(ff[0][:, 3]).sum(axis=1) > 0).any()  # opposite vertices take forces too
