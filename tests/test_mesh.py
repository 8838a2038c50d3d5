"""Box meshes (ffea_amd.mesh): the direct midpoint numbering of a Kuhn box equals the generic
construction -- every element edge, sorted by (low vertex, high vertex), numbered in that order
after the linear nodes -- and midpoints are edge averages."""
import numpy as np
import pytest

from ffea_amd import mesh as M


def generic_midpoints(m):
    tet4 = m.topo[:, :4].astype(np.int64)
    n = m.num_linear
    ends = np.stack([np.sort(tet4[:, list(ab)], axis=1) for ab in M.EDGES], axis=1)
    key = ends[..., 0] * n + ends[..., 1]
    uniq, inv = np.unique(key.ravel(), return_inverse=True)
    lin = m.pos[:n]
    return n + inv.reshape(-1, 6), np.concatenate([lin, 0.5 * (lin[uniq // n] + lin[uniq % n])])


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 4), (5, 1, 2), (7, 6, 5)])
def test_box_mesh_midpoints_match_generic_numbering(dims):
    m = M.box_mesh(*dims, jitter=0.15, seed=3)
    mids, pos = generic_midpoints(m)
    assert np.array_equal(m.topo[:, 4:], mids)
    assert np.array_equal(m.pos, pos)
    assert m.num_nodes == m.num_linear + len(np.unique(mids))
    # linearise through the edge list equals the element-by-element restatement
    x = M.perturb(m, 0.02, seed=1)
    plain = M.Mesh(pos=m.pos, topo=m.topo, material=m.material, stokes_radius=m.stokes_radius,
                   num_linear=m.num_linear)
    assert np.array_equal(x, M.perturb(plain, 0.02, seed=1))
