"""Host-side mirror of the reference's Blob / NoMassCGSolver interface over the C-ABI.

`Blob` follows include/Blob.h (init :124, update_internal_forces :130,
update_positions :140) and `GpuNoMassCGSolver` follows include/Solver.h:30-55
(init / solve / apply_matrix).  Arrays use FFEA's full node numbering (N_t nodes
including quadratic midpoints) and meso units, as the reference holds them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .units import SimParams

PRECISIONS = {"fp64": _lib.FFH_PRECISION_FP64, "mixed": _lib.FFH_PRECISION_MIXED}


def package_seed(rng_seed: int) -> np.ndarray:
    """World::init RNG.2 (src/World.cpp:299-305): six RngStream seeds drawn from the host
    libc's srand/rand, (rand() + rand()) % twoMax[i/3].

    The int sum can exceed INT_MAX, which is undefined behaviour in C++. The reference's
    optimised build widens both operands to the 64-bit type of twoMax before adding, so the
    sum does not wrap: its own checkpoint tests/physics/cyl_160_fine/cyl_160_fine.fcp
    (rng_seed 753, no draws made) holds exactly the streams of the non-wrapping sum.
    With one thread the first RngStream starts from exactly these seeds."""
    libc = C.CDLL(None)
    libc.srand(C.c_uint(rng_seed & 0xffffffff))
    two_max = (4294967087, 4294944443)
    out = np.zeros(6, np.uint64)
    for i in range(6):
        out[i] = (int(libc.rand()) + int(libc.rand())) % two_max[i // 3]
    return out


class Blob:
    """One dynamic blob on the GPU (Blob + NoMassCGSolver + its RngStream).

    mesh: an object with pos [N_t,3], topo [E,10], material [E,6] (SI), stokes_radius [N_t],
    pinned (node ids), in the blob's input units; scale: the blob `scale` already divided by
    the meso length (units.blob_scale).
    """

    def __init__(self, mesh, params: SimParams, scale: float, rng_state=None, precision="fp64",
                 pos=None):
        self._L = _lib.load()
        self.num_nodes = int(mesh.pos.shape[0])
        self.num_elements = int(mesh.topo.shape[0])
        self.params = params
        rest = np.ascontiguousarray(np.asarray(mesh.pos, np.float64) * scale).ravel()
        topo = np.ascontiguousarray(mesh.topo, np.int32).ravel()
        mat = np.ascontiguousarray(mesh.material, np.float64).ravel()
        rad = np.ascontiguousarray(np.asarray(mesh.stokes_radius, np.float64) * scale)
        pinned = np.ascontiguousarray(getattr(mesh, "pinned", np.zeros(0)), np.int32)
        if rng_state is None:
            rng_state = package_seed(params.rng_seed)
        st = np.ascontiguousarray(rng_state, np.uint64)
        self._h = self._L.ffh_blob_create(
            self.num_nodes, self.num_elements, rest, topo, mat, rad, int(pinned.size),
            pinned if pinned.size else np.zeros(1, np.int32), params.dt, params.kT,
            params.epsilon2, params.max_iterations_cg, params.calc_noise, params.calc_stokes,
            params.stokes_visc, st, PRECISIONS[precision])
        if not self._h:
            raise _lib.FFHError(-1, _lib.last_error())
        self.precision = precision
        self.solver = "CG_nomass"
        if pos is not None:
            self.set_positions(np.asarray(pos, np.float64) * scale)

    # -- lifetime -----------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._L.ffh_blob_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- the step -------------------------------------------------------------------
    def update_internal_forces(self) -> int:
        n = C.c_int(0)
        _lib.check(self._L.ffh_blob_update_internal_forces(self._h, C.byref(n)))
        return n.value

    def update_positions(self) -> None:
        _lib.check(self._L.ffh_blob_update_positions(self._h))

    def step(self, nsteps: int = 1) -> None:
        _lib.check(self._L.ffh_blob_step(self._h, int(nsteps)))

    TUNING_KEYS = ("element_block", "assembly_block", "assembly_tiles", "spmv_waves_per_slice",
                   "spmv_nontemporal")

    def tuning(self) -> dict:
        v = [C.c_int() for _ in range(5)]
        _lib.check(self._L.ffh_blob_tuning(self._h, *[C.byref(x) for x in v]))
        return dict(zip(self.TUNING_KEYS, (x.value for x in v)))

    def set_tuning(self, element_block: int, assembly_block: int, assembly_tiles: int,
                   spmv_waves_per_slice: int, spmv_nontemporal: int = 0) -> None:
        _lib.check(self._L.ffh_blob_set_tuning(self._h, int(element_block), int(assembly_block),
                                               int(assembly_tiles), int(spmv_waves_per_slice),
                                               int(spmv_nontemporal)))

    def set_external_forces(self, force_meso=None) -> None:
        """Forces added to every following position phase ([N_t, 3] meso units; None clears)."""
        if force_meso is None:
            _lib.check(self._L.ffh_blob_set_external_forces(self._h, None))
            return
        f = np.ascontiguousarray(force_meso, np.float64).ravel()
        if f.size != 3 * self.num_nodes:
            raise ValueError("external forces need 3 values per node")
        _lib.check(self._L.ffh_blob_set_external_forces(self._h, f.ctypes.data_as(C.c_void_p)))

    def set_springs(self, node0, node1, k, l) -> None:
        """Springs between nodes (N_t ids of this handle), k and l in meso units; empty clears."""
        n0, n1 = (np.ascontiguousarray(x, np.int32).ravel() for x in (node0, node1))
        kk, ll = (np.ascontiguousarray(x, np.float64).ravel() for x in (k, l))
        if not (n0.size == n1.size == kk.size == ll.size):
            raise ValueError("springs need node0, node1, k and l of one length")
        z = np.zeros(1, np.int32)
        _lib.check(self._L.ffh_blob_set_springs(self._h, n0.size, n0 if n0.size else z, n1 if n1.size else z,
                                                kk if kk.size else np.zeros(1), ll if ll.size else np.zeros(1)))

    def spring_energy(self) -> float:
        e = C.c_double(0)
        _lib.check(self._L.ffh_blob_spring_energy(self._h, C.byref(e)))
        return e.value

    def set_constant_forces(self, nodes, forces) -> None:
        """Constant forces [n, 3] (meso units) on nodes [n] (N_t ids), in the reference's order."""
        nd = np.ascontiguousarray(nodes, np.int32).ravel()
        fr = np.ascontiguousarray(forces, np.float64).ravel()
        if fr.size != 3 * nd.size:
            raise ValueError("one 3-vector per node")
        _lib.check(self._L.ffh_blob_set_constant_forces(self._h, nd.size, nd if nd.size else np.zeros(1, np.int32),
                                                        fr if fr.size else np.zeros(3)))

    def set_rotational_forces(self, nodes, F, axis, torque) -> None:
        """Rotational constant forces on nodes [n] (N_t ids): F [n] (meso force), axis [n, 6]
        (point in meso length, unit direction), torque [n] (1: 't', 0: 'f'); empty clears."""
        nd = np.ascontiguousarray(nodes, np.int32).ravel()
        fr = np.ascontiguousarray(F, np.float64).ravel()
        ax = np.ascontiguousarray(axis, np.float64).ravel()
        tq = np.ascontiguousarray(torque, np.int32).ravel()
        if fr.size != nd.size or ax.size != 6 * nd.size or tq.size != nd.size:
            raise ValueError("one F, axis and type per node")
        if nd.size == 0:
            nd, fr, ax, tq = np.zeros(1, np.int32), np.zeros(1), np.zeros(6), np.zeros(1, np.int32)
            n = 0
        else:
            n = nd.size
        _lib.check(self._L.ffh_blob_set_rotational_forces(self._h, n, nd, fr, ax, tq))

    def set_ssint(self, faces=None, types=None, emin=None, rmin=None, grid=(1, 1, 1), cell_size=1.0,
                  inc_self: int = 1, es_update: int = 1, kind: str = "lennard-jones", opposite=None,
                  k0=None) -> None:
        """Surface-surface interactions: faces [F, 3] (N_t ids), types [F] (-1 inactive),
        emin / rmin (/ k0) [ntypes, ntypes] meso, the lookup grid (es_N per axis, cell size =
        ssint_cutoff in meso), inc_self_ssint, es_update, the type ("lennard-jones", "ljsteric",
        "gensoft", "steric") and each face's opposite element vertex (all types but
        lennard-jones; steric needs no emin / rmin).  faces None removes them."""
        if faces is None:
            z = np.zeros(1)
            _lib.check(self._L.ffh_blob_set_ssint(self._h, 1, 0, np.zeros(3, np.int32), np.zeros(1, np.int32),
                                                  np.zeros(1, np.int32), 0, z, z, z, np.ones(3, np.int32), 1.0, 0, 1))
            return
        fc = np.ascontiguousarray(faces, np.int32).ravel()
        ty = np.ascontiguousarray(types, np.int32).ravel()
        if fc.size != 3 * ty.size:
            raise ValueError("one type per face")
        op = np.full(ty.size, -1, np.int32) if opposite is None else np.ascontiguousarray(opposite, np.int32).ravel()
        if kind == "steric" and emin is None:
            emin = rmin = np.zeros((1, 1))
        e = np.ascontiguousarray(emin, np.float64)
        r = np.ascontiguousarray(rmin, np.float64)
        k = np.zeros_like(e) if k0 is None else np.ascontiguousarray(k0, np.float64)
        if e.ndim != 2 or e.shape[0] != e.shape[1] or r.shape != e.shape or k.shape != e.shape:
            raise ValueError("emin, rmin and k0 are square tables of one size")
        _lib.check(self._L.ffh_blob_set_ssint(self._h, _lib.FFH_SSINT_KINDS[kind], ty.size,
                                              fc if fc.size else np.zeros(3, np.int32),
                                              op if op.size else np.zeros(1, np.int32),
                                              ty if ty.size else np.zeros(1, np.int32), e.shape[0], e.ravel(),
                                              r.ravel(), k.ravel(), np.ascontiguousarray(grid, np.int32),
                                              float(cell_size), int(inc_self), int(es_update)))

    def set_steric(self, steric_factor: float = 1.0, steric_dr: float = 5e-3) -> None:
        """steric_factor and steric_dr (meso, as SimulationParams holds them) of the steric
        forces of the ljsteric, gensoft and steric interactions."""
        _lib.check(self._L.ffh_blob_set_steric(self._h, float(steric_factor), float(steric_dr)))

    def set_stress_output(self, enable: bool = True) -> None:
        """Store each element's internal_stress_mag in every element phase (Blob's stress file)."""
        _lib.check(self._L.ffh_blob_set_stress_output(self._h, int(bool(enable))))

    def stress_magnitudes(self) -> np.ndarray:
        """internal_stress_mag of the last element phase, one per element (meso pressure)."""
        out = np.zeros(self.num_elements, np.float64)
        _lib.check(self._L.ffh_blob_get_stress_magnitudes(self._h, out))
        return out

    def ssint_energy(self) -> float:
        """The surface-surface interaction energy of the last step's solve (meso units)."""
        v = C.c_double()
        _lib.check(self._L.ffh_blob_ssint_energy(self._h, C.byref(v)))
        return v.value

    def set_box(self, box=None, periodic_mask: int = 0b111111) -> None:
        """World::run's periodic box (meso units): each step first moves a blob whose centroid
        has left the box through a periodic wall back in.  periodic_mask bits 0..5:
        wall_x_1, wall_x_2, wall_y_1, wall_y_2, wall_z_1, wall_z_2.  box None removes it."""
        if box is None:
            _lib.check(self._L.ffh_blob_set_box(self._h, np.zeros(3), 0))
            return
        bx = np.ascontiguousarray(box, np.float64).ravel()
        if bx.size != 3:
            raise ValueError("box is three lengths")
        _lib.check(self._L.ffh_blob_set_box(self._h, bx, int(periodic_mask)))

    def set_surface_forces(self, faces, sets, set_forces) -> None:
        """Linear surface forces: faces [F, 3] (vertex ids), sets: lists of face indices,
        set_forces [S, 3] total force of each set (meso units); empty sets clear."""
        fc = np.ascontiguousarray(faces, np.int32).ravel()
        off = np.concatenate([[0], np.cumsum([len(x) for x in sets])]).astype(np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in sets]) if sets else np.zeros(1),
                                    np.int32)
        sf = np.ascontiguousarray(set_forces, np.float64).ravel()
        if sf.size != 3 * len(sets):
            raise ValueError("one force per set")
        _lib.check(self._L.ffh_blob_set_surface_forces(self._h, fc.size // 3, fc if fc.size else np.zeros(3, np.int32),
                                                       len(sets), off, flat, sf if sf.size else np.zeros(3)))

    def set_graphs(self, enable: bool) -> None:
        _lib.check(self._L.ffh_blob_set_graphs(self._h, int(bool(enable))))

    def steps_done(self) -> int:
        n = C.c_longlong(0)
        _lib.check(self._L.ffh_blob_steps_done(self._h, C.byref(n)))
        return n.value

    # -- solver interface -------------------------------------------------------------
    def set_solver(self, name: str) -> None:
        """The Blob's linear solver by its .ffea name (src/World.cpp:2545-2560): "CG_nomass"
        (default), "CG" (mass matrix, ConjugateGradientSolver) or "masslumped"."""
        if name not in _lib.FFH_SOLVERS:
            raise ValueError(f"solver {name!r}: one of {sorted(_lib.FFH_SOLVERS)}")
        _lib.check(self._L.ffh_blob_set_solver(self._h, _lib.FFH_SOLVERS[name]))
        self.solver = name

    def solve(self, force) -> np.ndarray:
        x = np.ascontiguousarray(force, np.float64).copy().ravel()
        _lib.check(self._L.ffh_blob_solve(self._h, x))
        return x

    def apply_matrix(self, vec) -> np.ndarray:
        out = np.zeros(3 * self.num_nodes, np.float64)
        _lib.check(self._L.ffh_blob_apply_matrix(
            self._h, np.ascontiguousarray(vec, np.float64).ravel(), out))
        return out

    # -- state ----------------------------------------------------------------------------
    def set_positions(self, pos_meso) -> None:
        _lib.check(self._L.ffh_blob_set_positions(
            self._h, np.ascontiguousarray(pos_meso, np.float64).ravel()))

    def set_velocities(self, vel_meso) -> None:
        _lib.check(self._L.ffh_blob_set_velocities(
            self._h, np.ascontiguousarray(vel_meso, np.float64).ravel()))

    def _get(self, fn) -> np.ndarray:
        out = np.zeros(3 * self.num_nodes, np.float64)
        _lib.check(fn(self._h, out))
        return out.reshape(-1, 3)

    def positions(self) -> np.ndarray:
        return self._get(self._L.ffh_blob_get_positions)

    def velocities(self) -> np.ndarray:
        return self._get(self._L.ffh_blob_get_velocities)

    def rhs(self) -> np.ndarray:
        return self._get(self._L.ffh_blob_get_rhs)

    def rng_state(self) -> np.ndarray:
        out = np.zeros(6, np.uint64)
        _lib.check(self._L.ffh_blob_get_rng_state(self._h, out))
        return out

    def set_rng_state(self, st) -> None:
        _lib.check(self._L.ffh_blob_set_rng_state(self._h, np.ascontiguousarray(st, np.uint64)))

    def last_cg(self) -> tuple[int, float]:
        it, res = C.c_int(0), C.c_double(0)
        _lib.check(self._L.ffh_blob_last_cg(self._h, C.byref(it), C.byref(res)))
        return it.value, res.value

    def sizes(self) -> dict:
        n, e, b, s = C.c_int(), C.c_int(), C.c_longlong(), C.c_longlong()
        _lib.check(self._L.ffh_blob_sizes(self._h, C.byref(n), C.byref(e), C.byref(b), C.byref(s)))
        return {"num_linear": n.value, "num_elements": e.value, "num_blocks": b.value,
                "num_slots": s.value}

    def stream(self) -> int:
        return int(self._L.ffh_blob_stream(self._h) or 0)

    def synchronize(self) -> None:
        _lib.check(self._L.ffh_blob_synchronize(self._h))

    def set_profiling(self, enable: bool) -> None:
        _lib.check(self._L.ffh_blob_set_profiling(self._h, int(bool(enable))))

    def kernel_times(self) -> dict:
        ms = np.zeros(_lib.FFH_NUM_KERNELS, np.float64)
        cnt = np.zeros(_lib.FFH_NUM_KERNELS, np.int64)
        _lib.check(self._L.ffh_blob_kernel_times(self._h, ms, cnt))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(_lib.KERNEL_NAMES)}

    def strain_energy(self) -> float:
        e = C.c_double(0)
        _lib.check(self._L.ffh_blob_strain_energy(self._h, C.byref(e)))
        return e.value

    def kinetic_energy(self) -> float:
        e = C.c_double(0)
        _lib.check(self._L.ffh_blob_kinetic_energy(self._h, C.byref(e)))
        return e.value


    # -- learned preconditioner (no counterpart in the reference) -------------------------
    def mpnn_init(self, seed: int = 0) -> None:
        _lib.check(self._L.ffh_mpnn_init(self._h, int(seed)))

    def mpnn_weights(self) -> np.ndarray:
        w = np.zeros(self._L.ffh_mpnn_num_params(), np.float32)
        _lib.check(self._L.ffh_mpnn_get_weights(self._h, w))
        return w

    def mpnn_set_weights(self, w) -> None:
        _lib.check(self._L.ffh_mpnn_set_weights(self._h, np.ascontiguousarray(w, np.float32)))

    def set_preconditioner(self, kind: str = "jacobi", fallback_iterations: int = 0) -> None:
        k = {"jacobi": _lib.FFH_PRECOND_JACOBI, "learned": _lib.FFH_PRECOND_LEARNED}[kind]
        _lib.check(self._L.ffh_blob_set_preconditioner(self._h, k, int(fallback_iterations)))

    def mpnn_apply(self, r) -> np.ndarray:
        out = np.zeros(3 * self.num_nodes, np.float64)
        _lib.check(self._L.ffh_mpnn_apply(self._h, np.ascontiguousarray(r, np.float64).ravel(), out))
        return out.reshape(-1, 3)

    def mpnn_harvest(self, max_snapshots: int) -> None:
        _lib.check(self._L.ffh_mpnn_harvest(self._h, int(max_snapshots)))

    def mpnn_harvest_count(self) -> tuple[int, int]:
        n, p = C.c_int(0), C.c_longlong(0)
        _lib.check(self._L.ffh_mpnn_harvest_count(self._h, C.byref(n), C.byref(p)))
        return n.value, p.value

    def mpnn_train(self, epochs: int, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        loss = np.zeros(int(epochs), np.float32)
        _lib.check(self._L.ffh_mpnn_train(self._h, int(epochs), float(lr), int(seed), loss))
        return loss

    def mpnn_loss_grad(self, r, e) -> tuple[float, np.ndarray]:
        g = np.zeros(self._L.ffh_mpnn_num_params(), np.float32)
        loss = C.c_double(0)
        _lib.check(self._L.ffh_mpnn_loss_grad(self._h, np.ascontiguousarray(r, np.float64).ravel(),
                                               np.ascontiguousarray(e, np.float64).ravel(), C.byref(loss), g))
        return loss.value, g


class Batch(Blob):
    """Several blobs of one World stepped together on the GPU (ffh_batch_create): one launch
    per phase for all of them, one RngStream in World::run's draw order (the element phases
    of every blob, then their position phases, src/World.cpp:2020-2041).

    meshes: one mesh per blob (as for Blob); scales: one blob scale each (or one for all).
    Node arrays (positions, velocities, rhs, solve/apply_matrix vectors) are the blobs'
    arrays concatenated in blob order; `split` cuts such an array back into blobs.
    """

    def __init__(self, meshes, params: SimParams, scales, rng_state=None, precision="fp64"):
        self._L = _lib.load()
        meshes = list(meshes)
        nb = len(meshes)
        if np.ndim(scales) == 0:
            scales = [float(scales)] * nb
        if len(scales) != nb or nb == 0:
            raise ValueError("need one scale per blob and at least one blob")
        nn = np.array([m.pos.shape[0] for m in meshes], np.int32)
        ne = np.array([m.topo.shape[0] for m in meshes], np.int32)
        pins = [np.asarray(getattr(m, "pinned", np.zeros(0)), np.int32).ravel() for m in meshes]
        npin = np.array([p.size for p in pins], np.int32)
        pos = np.ascontiguousarray(np.concatenate(
            [np.asarray(m.pos, np.float64) * sc for m, sc in zip(meshes, scales)])).ravel()
        topo = np.ascontiguousarray(np.concatenate([m.topo for m in meshes]), np.int32).ravel()
        mat = np.ascontiguousarray(np.concatenate([m.material for m in meshes]), np.float64).ravel()
        rad = np.ascontiguousarray(np.concatenate(
            [np.asarray(m.stokes_radius, np.float64) * sc for m, sc in zip(meshes, scales)]))
        pin = np.ascontiguousarray(np.concatenate(pins) if npin.sum() else np.zeros(1), np.int32)
        if rng_state is None:
            rng_state = package_seed(params.rng_seed)
        st = np.ascontiguousarray(rng_state, np.uint64)
        self.num_blobs = nb
        self.num_nodes = int(nn.sum())
        self.num_elements = int(ne.sum())
        self.node_off = np.concatenate([[0], np.cumsum(nn)]).astype(np.int64)
        self.params = params
        self._h = self._L.ffh_batch_create(
            nb, nn, ne, pos, topo, mat, rad, npin, pin, params.dt, params.kT, params.epsilon2,
            params.max_iterations_cg, params.calc_noise, params.calc_stokes, params.stokes_visc,
            st, PRECISIONS[precision])
        if not self._h:
            raise _lib.FFHError(-1, _lib.last_error())
        self.precision = precision
        self.solver = "CG_nomass"

    def split(self, arr) -> list:
        """Cut a concatenated per-node array ([N_t,...] or [3 N_t]) into per-blob pieces."""
        a = np.asarray(arr)
        if a.ndim == 1:
            a = a.reshape(-1, 3)
        return [a[self.node_off[i]:self.node_off[i + 1]] for i in range(self.num_blobs)]

    def last_cg_per_blob(self):
        """(iterations, residual2, status) arrays of the last solve, one entry per blob."""
        it = np.zeros(self.num_blobs, np.int32)
        res = np.zeros(self.num_blobs, np.float64)
        stat = np.zeros(self.num_blobs, np.int32)
        _lib.check(self._L.ffh_batch_last_cg(self._h, it, res, stat))
        return it, res, stat


class GpuNoMassCGSolver:
    """The Solver interface (include/Solver.h:30-55) over the GPU blob.

    init() mirrors NoMassCGSolver::init's role (pattern + matrix at current positions);
    solve(x) takes the force [N_t,3] and returns the velocity; apply_matrix(in) returns K*in
    (a real product, where the reference's NoMassCGSolver stub does nothing,
    include/NoMassCGSolver.h:103-104).
    """

    def __init__(self):
        self.blob: Blob | None = None

    def init(self, mesh, params: SimParams, scale: float, rng_state=None, precision="fp64"):
        self.blob = Blob(mesh, params, scale, rng_state=rng_state, precision=precision)
        return _lib.FFH_OK

    def solve(self, x) -> np.ndarray:
        return self.blob.solve(x).reshape(-1, 3)

    def apply_matrix(self, vec) -> np.ndarray:
        return self.blob.apply_matrix(vec)
