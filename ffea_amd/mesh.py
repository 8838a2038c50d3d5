"""Box meshes of second-order (10-node) tetrahedra in FFEA's layout.

A box of nx*ny*nz cubic cells is split 6 tetrahedra per cell (Kuhn split along
the cell's main diagonal).  Every tetrahedron then gets 6 edge midpoints in
FFEA's local order n[4..9] = edges (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)
(src/tetra_element_linear.cpp:567-591, ffeatools/modules/FFEA_topology.py:418-467),
with midpoint nodes appended after the linear nodes and zero Stokes radius on
midpoints (ffeatools/modules/FFEA_topology.py:456-458).

The result is plain numpy arrays in the N_t layout the reference loads: node
positions for all nodes, a [E, 10] topology, per-element SI material rows and
per-node Stokes radii, all in the blob's input length unit before `scale`.
"""
from __future__ import annotations

import dataclasses
from itertools import permutations

import numpy as np

EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))

# Default material of the workloads (SI): density, shear_visc, bulk_visc,
# shear_mod, bulk_mod, dielectric -- the column order of a .mat row
# (src/Blob.cpp:2855-2866).
DEFAULT_MATERIAL = (1500.0, 1e-3, 1e-3, 3.7e8, 1.1e9, 1.0)


@dataclasses.dataclass
class Mesh:
    pos: np.ndarray           # [N_t, 3] float64, input units (before scale)
    topo: np.ndarray          # [E, 10] int32
    material: np.ndarray      # [E, 6] float64 SI
    stokes_radius: np.ndarray  # [N_t] float64, input units
    num_linear: int           # linear nodes are 0..num_linear-1
    pinned: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    # [N_t - num_linear, 2] edge ends of each midpoint node, when known (box meshes)
    mid_ends: np.ndarray | None = None

    @property
    def num_nodes(self) -> int:
        return int(self.pos.shape[0])

    @property
    def num_elements(self) -> int:
        return int(self.topo.shape[0])

    def linearise(self, pos: np.ndarray | None = None) -> np.ndarray:
        """Midpoints re-averaged from their edge ends (linearise_element)."""
        p = (self.pos if pos is None else pos).copy()
        if self.mid_ends is not None:  # one write per midpoint (the same value every element writes)
            e = self.mid_ends
            p[self.num_linear:] = 0.5 * (p[e[:, 0]] + p[e[:, 1]])
            return p
        for m, (a, b) in enumerate(EDGES):
            p[self.topo[:, 4 + m]] = 0.5 * (p[self.topo[:, a]] + p[self.topo[:, b]])
        return p


def _kuhn_cell_tets():
    """6 tets of the unit cube sharing the (0,0,0)-(1,1,1) diagonal, positively oriented."""
    corner = lambda v: v[0] + 2 * v[1] + 4 * v[2]
    tets = []
    for perm in permutations(range(3)):
        v = [0, 0, 0]
        path = [corner(v)]
        for ax in perm:
            v[ax] = 1
            path.append(corner(v))
        pts = np.array([[(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1] for c in path], float)
        det = np.linalg.det(pts[1:] - pts[0])
        if det < 0:
            path[1], path[2] = path[2], path[1]
        tets.append(path)
    return np.array(tets, np.int64)


def _element_volumes6(pos, tet4):
    p0 = pos[tet4[:, 0]]
    d1 = pos[tet4[:, 1]] - p0
    d2 = pos[tet4[:, 2]] - p0
    d3 = pos[tet4[:, 3]] - p0
    return np.einsum("ij,ij->i", d1, np.cross(d2, d3))


def box_mesh(nx: int, ny: int, nz: int, h: float = 1.0, material=DEFAULT_MATERIAL,
             stokes_radius: float = 0.5, jitter: float = 0.0, seed: int = 0) -> Mesh:
    """Kuhn-split box mesh with quadratic midpoints.

    jitter: linear nodes are moved by U(-jitter*h, +jitter*h) per coordinate, redrawing
    the nodes of any tetrahedron whose volume is not positive.
    """
    nx, ny, nz = int(nx), int(ny), int(nz)
    sx, sy = nx + 1, (nx + 1) * (ny + 1)
    n_lin = (nx + 1) * (ny + 1) * (nz + 1)
    g = np.indices((nx + 1, ny + 1, nz + 1)).reshape(3, -1).T  # (i, j, k) with k fastest
    # node id = i + sx*j + sy*k
    ids = g[:, 0] + sx * g[:, 1] + sy * g[:, 2]
    lin = np.empty((n_lin, 3))
    lin[ids] = g * h
    cell = np.indices((nx, ny, nz)).reshape(3, -1).T
    cell = cell[np.lexsort((cell[:, 0], cell[:, 1], cell[:, 2]))]  # x fastest
    base = cell[:, 0] + sx * cell[:, 1] + sy * cell[:, 2]
    off = np.array([(c & 1) + sx * ((c >> 1) & 1) + sy * ((c >> 2) & 1) for c in range(8)])
    kt = _kuhn_cell_tets()
    tet4 = (base[:, None, None] + off[kt][None, :, :]).reshape(-1, 4)

    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        rest = lin.copy()
        move = np.arange(n_lin)
        for _ in range(1000):
            lin[move] = rest[move] + rng.uniform(-jitter * h, jitter * h, size=(move.size, 3))
            bad = _element_volumes6(lin, tet4) <= 0.0
            if not bad.any():
                break
            move = np.unique(tet4[bad])
        else:
            raise RuntimeError("could not jitter mesh without inverting elements")

    # Midpoint ids appended after the linear nodes, one per edge in ascending (low, high)
    # vertex order.  Every edge of a Kuhn box runs from its lower vertex a to a + o with o one
    # of 7 offsets, ascending: +x, +y, +x+y, +z, +x+z, +y+z, +x+y+z, present wherever the cells
    # it borders exist; so the sorted edge list is counted per vertex instead of sorted.
    offs = np.array([1, sx, 1 + sx, sy, 1 + sy, sx + sy, 1 + sx + sy], np.int64)
    gi, gj, gk = (np.arange(n_lin) % sx, (np.arange(n_lin) // sx) % (ny + 1), np.arange(n_lin) // sy)
    ex, ey, ez = gi < nx, gj < ny, gk < nz
    has = np.stack([ex, ey, ex & ey, ez, ex & ez, ey & ez, ex & ey & ez], axis=1)  # [n_lin, 7]
    rank = np.cumsum(has, axis=1) - has
    start = np.concatenate([[0], np.cumsum(has.sum(axis=1))[:-1]])
    mid_of = n_lin + start[:, None] + rank  # valid where has
    mids = np.empty((tet4.shape[0], 6), np.int64)
    for m, (ea, eb) in enumerate(EDGES):
        lo, hi = np.minimum(tet4[:, ea], tet4[:, eb]), np.maximum(tet4[:, ea], tet4[:, eb])
        mids[:, m] = mid_of[lo, np.searchsorted(offs, hi - lo)]
    a = np.repeat(np.arange(n_lin), has.sum(axis=1))
    b = a + offs[np.nonzero(has)[1]]
    pos = np.concatenate([lin, 0.5 * (lin[a] + lin[b])], axis=0)
    topo = np.concatenate([tet4, mids], axis=1).astype(np.int32)
    mat = np.tile(np.asarray(material, float), (topo.shape[0], 1))
    rad = np.zeros(pos.shape[0])
    rad[:n_lin] = stokes_radius
    return Mesh(pos=pos, topo=topo, material=mat, stokes_radius=rad, num_linear=n_lin,
                mid_ends=np.stack([a, b], axis=1).astype(np.int32))


def perturb(mesh: Mesh, sigma: float, seed: int) -> np.ndarray:
    """Rest positions plus N(0, sigma^2) on every linear coordinate; midpoints re-averaged."""
    rng = np.random.default_rng(seed)
    p = mesh.pos.copy()
    p[: mesh.num_linear] += rng.normal(0.0, sigma, size=(mesh.num_linear, 3))
    return mesh.linearise(p)


def scale_material(mesh: Mesh, lo: float, hi: float, seed: int) -> None:
    """Scale each material parameter by one log-uniform factor in [lo, hi] per blob."""
    rng = np.random.default_rng(seed)
    f = np.exp(rng.uniform(np.log(lo), np.log(hi), size=5))
    mesh.material[:, :5] *= f[None, :]


def stiff_core(mesh: Mesh, axis: int = 0, factor: float = 10.0) -> None:
    """Multiply shear and bulk moduli by `factor` in the central third along `axis`."""
    p = mesh.pos[: mesh.num_linear, axis]
    lo, hi = p.min(), p.max()
    c = mesh.pos[mesh.topo[:, :4], axis].mean(axis=1)
    core = (c > lo + (hi - lo) / 3.0) & (c < lo + 2.0 * (hi - lo) / 3.0)
    mesh.material[core, 3] *= factor
    mesh.material[core, 4] *= factor
