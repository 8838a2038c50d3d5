"""Forces from outside the element loop in the oracle: springs (World::apply_springs,
src/World.cpp:3808-3846), linear constant forces (CTF 1) and linear surface forces (CTF 3) of
Blob::apply_ctforces (src/Blob.cpp:2094-2162).  No compiled reference object holds these (they
live in World.cpp / Blob.cpp), so they are pinned by what the mathematics fixes, on the
force vector handed to the solver (noise off, so a run with and without the forces differs only
by them):

- a spring adds k (d - l) along the separation to its first node and the opposite to its
  second, nothing elsewhere; at its rest length it adds exactly nothing; its energy is
  k (d - l)^2 / 2 and minus the energy's gradient is the force (central differences);
- a constant force on a vertex adds itself there; on an edge midpoint it reaches the two edge
  ends half each (linearise_force, src/Blob.cpp:1066-1085);
- a surface set spreads its total force over its faces in proportion to their areas, a third
  to each face vertex, so the additions sum to the set's force.
"""
import numpy as np
import pytest

import pyoracle
from conftest import box_case


@pytest.fixture
def clean_springs(oracle):
    oracle.set_springs([], [], [], [], [], [])
    yield
    oracle.set_springs([], [], [], [], [], [])


def run_pair(oracle, m, prm, sc, x0, setup, nsteps=1):
    """One world step each for a plain blob and one with `setup(blob)` applied; returns the
    two force vectors handed to the solver ([N_t, 3])."""
    out = []
    for with_forces in (False, True):
        oracle.set_package_seed(np.arange(1, 7, dtype=np.uint64))
        s = oracle.stream()
        b = oracle.blob(m, prm, sc, pos=x0)
        if with_forces:
            setup(b)
        else:
            oracle.set_springs([], [], [], [], [], [])
        for _ in range(nsteps):
            assert pyoracle.world_step(oracle, [b], s) == 0
        out.append(b.get("rhs").reshape(-1, 3))
    return out


def case():
    m, prm, sc, x0 = box_case((3, 3, 2), noise=0, sigma=0.05, seed=5, jitter=0.1)
    m.pinned = np.zeros(0, np.int32)
    return m, prm, sc, x0


def test_spring_force_and_rest_length(oracle, clean_springs):
    m, prm, sc, x0 = case()
    lin = np.unique(m.topo[:, :4])
    n0, n1 = int(lin[3]), int(lin[17])
    p = x0 * sc
    d = np.sqrt(((p[n1] - p[n0]) ** 2).sum())
    k, l = 0.7, 0.6 * d
    a, b = run_pair(oracle, m, prm, sc, x0, lambda bl: oracle.set_springs([0], [n0], [0], [n1], [k], [l]))
    diff = b - a
    sep = p[n1] - p[n0]
    want = k * (d - l) * sep / d
    scale = np.abs(a).max()
    assert np.allclose(diff[n0], want, rtol=0, atol=1e-12 * scale)
    assert np.allclose(diff[n1], -want, rtol=0, atol=1e-12 * scale)
    others = np.setdiff1d(np.arange(m.num_nodes), [n0, n1])
    assert np.array_equal(b[others], a[others])
    # at rest length (the oracle's own distance formula) nothing changes at all
    l0 = np.sqrt((sep[0] * sep[0] + sep[1] * sep[1]) + sep[2] * sep[2])
    a, b = run_pair(oracle, m, prm, sc, x0, lambda bl: oracle.set_springs([0], [n0], [0], [n1], [k], [l0]))
    assert np.array_equal(a, b)


def test_spring_energy_gradient(oracle, clean_springs):
    m, prm, sc, x0 = case()
    lin = np.unique(m.topo[:, :4])
    n0, n1 = int(lin[2]), int(lin[11])
    bl = oracle.blob(m, prm, sc, pos=x0)
    p = (x0 * sc).copy()
    k, l = 1.3, 0.5
    oracle.set_springs([0], [n0], [0], [n1], [k], [l])
    d = np.linalg.norm(p[n1] - p[n0])
    assert oracle.spring_energy([bl]) == pytest.approx(0.5 * k * (d - l) ** 2, rel=1e-13)
    h = 1e-6
    grad = np.zeros(3)
    for c in range(3):
        for sgn in (1, -1):
            q = p.copy()
            q[n0, c] += sgn * h
            bl.set_positions(q.ravel())
            grad[c] += sgn * oracle.spring_energy([bl]) / (2 * h)
    sep = p[n1] - p[n0]
    force0 = k * (d - l) * sep / d
    assert np.allclose(-grad, force0, rtol=1e-6, atol=1e-9)


def test_constant_force_on_vertex_and_midpoint(oracle, clean_springs):
    m, prm, sc, x0 = case()
    lin = np.unique(m.topo[:, :4])
    v = int(lin[5])
    e = 4
    mid, ea, eb = int(m.topo[e, 4]), int(m.topo[e, 0]), int(m.topo[e, 1])  # midpoint of edge (0, 1)
    F = np.array([[0.3, -0.2, 0.5], [-0.4, 0.1, 0.25]])
    a, b = run_pair(oracle, m, prm, sc, x0, lambda bl: bl.set_constant_forces([v, mid], F))
    diff = b - a
    scale = np.abs(a).max()
    assert np.allclose(diff[v], F[0], rtol=0, atol=1e-12 * scale)
    assert np.allclose(diff[ea], 0.5 * F[1], rtol=0, atol=1e-12 * scale)
    assert np.allclose(diff[eb], 0.5 * F[1], rtol=0, atol=1e-12 * scale)
    assert np.all(b[mid] == 0)
    assert np.allclose(diff.sum(axis=0), F.sum(axis=0), rtol=0, atol=1e-11 * scale)


def test_surface_forces_spread_by_area(oracle, clean_springs):
    m, prm, sc, x0 = case()
    faces = np.array([m.topo[e, [0, 1, 2]] for e in range(0, 12)])  # vertex triangles
    sets = [[0, 3, 5, 7], [2, 3, 11]]  # face 3 is in both sets
    SF = np.array([[1.0, 0.0, 0.5], [-0.2, 0.7, 0.0]])
    a, b = run_pair(oracle, m, prm, sc, x0, lambda bl: bl.set_surface_forces(faces, sets, SF))
    diff = b - a
    p = x0 * sc
    want = np.zeros_like(diff)
    for s, fl in enumerate(sets):
        ar = np.array([0.5 * np.linalg.norm(np.cross(p[faces[f][1]] - p[faces[f][0]], p[faces[f][2]] - p[faces[f][0]]))
                       for f in fl])
        for f, A in zip(fl, ar):
            for nd in faces[f]:
                want[nd] += A / (3 * ar.sum()) * SF[s]
    scale = np.abs(a).max()
    assert np.allclose(diff, want, rtol=0, atol=1e-12 * scale)
    assert np.allclose(diff.sum(axis=0), SF.sum(axis=0), rtol=0, atol=1e-11 * scale)


def test_surface_forces_on_midpoint_corners_are_conserved(oracle, clean_springs):
    """Second-order surface triangles have edge midpoints as corners; their tractions reach the
    vertices through linearise_force, half to each edge end, so the total is still the set's."""
    m, prm, sc, x0 = case()
    faces = np.array([m.topo[e, [0, 4, 5]] for e in range(6)] + [m.topo[e, [4, 7, 5]] for e in range(6, 9)])
    sets = [list(range(9))]
    SF = np.array([[0.3, -1.1, 0.4]])
    a, b = run_pair(oracle, m, prm, sc, x0, lambda bl: bl.set_surface_forces(faces, sets, SF))
    diff = b - a
    scale = np.abs(a).max()
    assert np.allclose(diff.sum(axis=0), SF[0], rtol=0, atol=1e-11 * scale)
    mids = np.unique(m.topo[:, 4:])
    assert np.all(b[mids] == 0)


# ---- rotational constant forces (CTF 2, src/Blob.cpp:2104-2136) -----------------------------

def rot_cases(seed=3, n=200):
    rng = np.random.default_rng(seed)
    for k in range(n):
        p = rng.normal(size=3) * 10 ** rng.uniform(-2, 2)
        d = rng.normal(size=3)
        d /= np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        a = np.concatenate([rng.normal(size=3), d])
        yield p, a, rng.uniform(0.1, 3.0), k & 1


def test_rot_force_matches_reference_helpers(oracle, reference):
    """The restatement against the reference's own intersectingPointToLine, cross product,
    mag and resize (src/mat_vec_fns_II.cpp), bit for bit, including a point on the axis."""
    for p, a, F, t in rot_cases():
        assert np.array_equal(oracle.rot_force(p, a, F, t), reference.rot_force(p, a, F, t))
    a = np.array([0.5, -1.0, 2.0, 0.0, 0.6, 0.8])
    p = a[:3] + 2.5 * a[3:]  # on the axis: its own foot
    for t in (0, 1):
        assert np.array_equal(oracle.rot_force(p, a, 1.2, t), reference.rot_force(p, a, 1.2, t), equal_nan=True)


def test_rot_force_geometry(oracle):
    """r runs from the node to its foot on the axis, perpendicular to the axis, as long as the
    node's distance to it; f = F (r x d): tangential, |f| = F |r| for 'f' and F for 't'."""
    for p, a, F, t in rot_cases(seed=11, n=100):
        x0, d = a[:3], a[3:]
        r = x0 + np.dot(p - x0, d) * d - p
        dist = np.linalg.norm(np.cross(p - x0, d))
        f = oracle.rot_force(p, a, F, t)
        want = (F / dist if t else F) * np.cross(r, d)
        assert np.allclose(f, want, rtol=1e-9, atol=1e-12 * np.linalg.norm(want))
        assert np.linalg.norm(f) == pytest.approx(F if t else F * dist, rel=1e-9)
        assert abs(np.dot(f, d)) <= 1e-9 * np.linalg.norm(f)
    p, a = np.array([1.0, 2.0, 3.0]), np.array([1.0, 2.0, 0.0, 0.0, 0.0, 1.0])
    assert np.array_equal(oracle.rot_force(p, a, 2.0, 0), np.zeros(3))


def test_rotational_force_on_vertex_and_midpoint(oracle, clean_springs):
    """Added to the node's force after the linear constant forces; on an edge midpoint it is
    evaluated at the midpoint and passed on half to each edge end."""
    m, prm, sc, x0 = case()
    lin = np.unique(m.topo[:, :4])
    v = int(lin[7])
    e = 6
    mid, ea, eb = int(m.topo[e, 5]), int(m.topo[e, 0]), int(m.topo[e, 2])  # midpoint of edge (0, 2)
    ax = np.array([[0.1, 0.2, 0.0, 0.0, 0.6, 0.8], [0.0, 0.0, 0.3, 1.0, 0.0, 0.0]])
    F, T = [0.7, 0.4], [0, 1]
    a, b = run_pair(oracle, m, prm, sc, x0, lambda bl: bl.set_rotational_forces([v, mid], F, ax, T))
    diff = b - a
    p = x0 * sc
    pm = 0.5 * (p[ea] + p[eb])
    fv = oracle.rot_force(p[v], ax[0], F[0], T[0])
    fm = oracle.rot_force(pm, ax[1], F[1], T[1])
    scale = np.abs(a).max()
    assert np.allclose(diff[v], fv, rtol=0, atol=1e-12 * scale)
    assert np.allclose(diff[ea], 0.5 * fm, rtol=0, atol=1e-12 * scale)
    assert np.allclose(diff[eb], 0.5 * fm, rtol=0, atol=1e-12 * scale)
    others = np.setdiff1d(np.arange(m.num_nodes), [v, ea, eb])
    assert np.array_equal(b[others], a[others])
