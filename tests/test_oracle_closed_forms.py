"""Closed-form pins of the oracle's element and solver math (independent of the reference
objects): unit tetrahedron gradients and volume, stress-free rest state, rigid motions,
viscosity-matrix symmetry / null space / positivity, and CG against a dense solve."""
import numpy as np
import pytest

from conftest import box_case
from ffea_amd import mesh as M, units as U
from ffea_amd.mesh import Mesh


def one_tet(verts):
    verts = np.asarray(verts, float)
    mids = [0.5 * (verts[a] + verts[b]) for a, b in M.EDGES]
    pos = np.vstack([verts, mids])
    return Mesh(pos=pos, topo=np.arange(10, dtype=np.int32)[None, :],
                material=np.array([M.DEFAULT_MATERIAL]), stokes_radius=np.r_[np.full(4, 0.5), np.zeros(6)],
                num_linear=4)


def noiseless(eps=1e-6):
    return U.SimParams.from_si(epsilon=eps, calc_noise=0)


def test_unit_tet_gradients_and_volume(oracle):
    m = one_tet([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    b = oracle.blob(m, noiseless(), 1.0)
    dpsi = b.get("dpsi").reshape(3, 4)  # [dim][node]
    assert np.array_equal(dpsi, np.array([[-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]], float))
    assert b.get("vol")[0] == 1.0 / 6.0
    assert b.get("vol_0")[0] == 1.0 / 6.0
    assert np.array_equal(b.get("J_inv_0").reshape(3, 3), np.eye(3))


def test_scaled_tet_volume_and_gradient_sum(oracle):
    rng = np.random.default_rng(3)
    for _ in range(20):
        v = rng.normal(size=(4, 3))
        if np.linalg.det(v[1:] - v[0]) == 0:
            continue
        b = oracle.blob(one_tet(v), noiseless(), 1.0)
        dpsi = b.get("dpsi").reshape(3, 4)
        vol = abs(np.linalg.det(v[1:] - v[0])) / 6.0
        assert b.get("vol")[0] == pytest.approx(vol, rel=1e-13)
        # gradients of the barycentric coordinates: sum to zero, grad(psi_a).(x_b - x_0) = delta
        assert np.allclose(dpsi.sum(axis=1), 0.0, atol=1e-12 * np.abs(dpsi).max())
        E = (v[1:] - v[0]) @ dpsi[:, 1:]
        assert np.allclose(E, np.eye(3), atol=1e-12)


def rhs_of(oracle, m, prm, sc, pos):
    oracle.set_package_seed(np.array([1, 2, 3, 4, 5, 6], np.uint64))
    s = oracle.stream()
    b = oracle.blob(m, prm, sc, pos=pos)
    inv, rhs = b.forces_only(s)
    assert inv == 0
    return b, rhs


def test_rest_state_and_rigid_motions_are_force_free(oracle):
    m, prm, sc, x0 = box_case((3, 3, 3), noise=0)
    scale_f = None
    # deformed reference force magnitude
    _, f_def = rhs_of(oracle, m, prm, sc, x0)
    scale_f = np.abs(f_def).max()
    _, f_rest = rhs_of(oracle, m, prm, sc, m.pos)
    assert np.abs(f_rest).max() < 1e-9 * scale_f
    _, f_tr = rhs_of(oracle, m, prm, sc, m.pos + np.array([3.0, -1.0, 0.5]))
    assert np.abs(f_tr).max() < 1e-9 * scale_f
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    _, f_rot = rhs_of(oracle, m, prm, sc, m.pos @ R.T)
    assert np.abs(f_rot).max() < 1e-9 * scale_f
    # a rigidly moved deformed state has the rotated forces
    _, f_def_rot = rhs_of(oracle, m, prm, sc, x0 @ R.T)
    assert np.allclose(f_def_rot.reshape(-1, 3), f_def.reshape(-1, 3) @ R.T, atol=1e-9 * scale_f)


def test_viscosity_matrix_symmetric_psd_translation_null(oracle):
    m, prm, sc, x0 = box_case((2, 2, 2), noise=0, jitter=0.15, seed=5)
    b, _ = rhs_of(oracle, m, prm, sc, x0)
    V = b.get("visc").reshape(-1, 12, 12)
    for Ve in V:
        assert np.array_equal(Ve, Ve.T)
        w = np.linalg.eigvalsh(Ve)
        assert w.min() > -1e-12 * w.max()
        # rows are (direction i, node a): a rigid translation t gives V . t = 0
        for d in range(3):
            t = np.zeros(12)
            t[4 * d: 4 * d + 4] = 1.0
            assert np.abs(Ve @ t).max() < 1e-12 * np.abs(Ve).max()


def test_cg_solution_matches_dense_solve(oracle):
    m, prm, sc, x0 = box_case((2, 3, 2), noise=1, eps=1e-10)
    b, rhs = rhs_of(oracle, m, prm, sc, x0)
    st, v = b.solve(rhs)
    assert st == 0
    key, col, val = b.csr()
    n = key.size - 1
    K = np.zeros((n, n))
    for i in range(n):
        K[i, col[key[i]:key[i + 1]]] = val[key[i]:key[i + 1]]
    assert np.allclose(K, K.T, rtol=0, atol=1e-13 * np.abs(K).max())
    vd = np.linalg.solve(K, rhs)
    assert np.linalg.norm(v - vd) <= 1e-8 * np.linalg.norm(vd)
    # stopping invariant of NoMassCGSolver: ||f - K v||^2 / ||f||^2 < eps^2
    r = rhs - K @ v
    assert (r @ r) / (rhs @ rhs) < 1e-20 * 1.01
    # midpoint rows are identity with zero force -> zero velocity
    mid = np.setdiff1d(np.arange(m.num_nodes), np.unique(m.topo[:, :4]))
    assert np.all(v.reshape(-1, 3)[mid] == 0)


def test_strain_energy_is_the_potential_of_the_elastic_forces(oracle):
    """The measured strain energy (src/Blob.cpp:1424-1432) and the elastic nodal forces
    (src/tetra_element_linear.cpp:378-408, :441-490) are computed by separate code.  With noise
    off the force vector is purely elastic and must be minus the energy's gradient; a dropped
    term, a wrong factor or a wrong sign in either breaks this.  Also: zero at rest, invariant
    under rigid motion."""
    m, prm, sc, x0 = box_case((2, 2, 1), noise=0, sigma=0.05, seed=5, jitter=0.1)

    def energy_and_rhs(pos):
        b, rhs = rhs_of(oracle, m, prm, sc, pos)
        return oracle.strain_energy(b), rhs

    U0, f = energy_and_rhs(x0)
    assert U0 > 0
    U_rest, _ = energy_and_rhs(m.pos)
    assert abs(U_rest) < 1e-12 * U0
    th = 0.4
    R = np.array([[1, 0, 0], [0, np.cos(th), -np.sin(th)], [0, np.sin(th), np.cos(th)]])
    U_rot, _ = energy_and_rhs(x0 @ R.T + np.array([1.0, 2.0, -3.0]))
    assert U_rot == pytest.approx(U0, rel=1e-11)
    h = 1e-5  # input units; central differences, O(h^2) truncation
    fmax = np.abs(f).max()
    for node in range(m.num_linear):
        for c in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[node, c] += h
            xm[node, c] -= h
            dU = (energy_and_rhs(xp)[0] - energy_and_rhs(xm)[0]) / (2 * h * sc)
            assert -dU == pytest.approx(f[3 * node + c], abs=1e-6 * fmax), (node, c)
