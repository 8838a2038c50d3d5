"""Benchmark of the massless FFEA time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1|c0|c2|c4] [--no-cpu]

Workload (default "c1"): one dynamic blob, a cube of 48^3 unit cells (h = 1 nm) split 6 tets
per cell (Kuhn), 117,649 linear nodes / 912,673 nodes with midpoints / 663,552 tets; material
1500 kg/m^3, shear/bulk viscosity 1e-3 Pa s, shear modulus 3.7e8 Pa, bulk modulus 1.1e9 Pa,
Stokes radius 0.5 nm on linear nodes, solvent viscosity 1e-3 Pa s; T = 300 K, dt = 1e-14 s,
Jacobi-PCG to relative eps = 1e-6, thermal noise on (RngStream seeded by rng_seed 44),
initial positions rest + N(0, (0.02 h)^2).  A step = one World::run iteration: element phase,
assembly, force gather + noise, CG solve, integration -- all on the GPU, inputs resident in HBM.

Other workloads: "c0" (8^3 cells), "c4" (192x96x48 jittered box with a 10x stiffer core) and
"c2", a batch of 256 independent blobs of uneven size (box meshes with nx, ny, nz uniform in
[6, 32], nodes jittered +-0.15 h, every material parameter scaled by a per-blob log-uniform
factor in [0.5, 2]) stepped together: one launch per phase for all blobs.

`value` is FP64-path steps/s (the reference's arithmetic); the mixed path (FP32 element,
assembly and SpMV storage) is reported beside it.  Kernel durations come from HIP events the
library records on its own stream around every launch in the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    # name: (cells, jitter (fraction of h), stiff-core factor, description); cells None: batch
    "c0": ((8, 8, 8), 0.0, 1.0, "cube 8^3 cells, Kuhn 6-tet split: 729 linear nodes, 3,072 tets"),
    "c3": (None, 0.15, 1.0,
           "learned preconditioner: trained on 20,000 (residual, correction) pairs harvested from plain CG "
           "solves of a 64-blob batch drawn like c2 (n in [6, 32]); evaluated on 64 held-out jittered box "
           "meshes with n in [24, 64] per axis, flexible CG with the network against Jacobi-PCG"),
    "c2": (None, 0.15, 1.0,
           "batch of 256 blobs, box meshes with nx, ny, nz uniform in [6, 32] (seeded), Kuhn 6-tet split, "
           "nodes jittered +-0.15h (no inverted tets), each material parameter scaled by a per-blob "
           "log-uniform factor in [0.5, 2]: 2,111,146 linear nodes, 10,886,874 tets"),
    "c1": ((48, 48, 48), 0.0, 1.0, "cube 48^3 cells, Kuhn 6-tet split: 117,649 linear nodes, 663,552 tets"),
    "c4": ((192, 96, 48), 0.15, 10.0,
           "box 192x96x48 cells, Kuhn 6-tet split, nodes jittered +-0.15h (seeded, no inverted tets), "
           "central third along x 10x stiffer (shear and bulk moduli): 917,329 linear nodes, 5,308,416 tets"),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def box_batch(count, lo, hi, seed, jitter=0.15):
    """`count` box meshes, nx, ny, nz uniform in [lo, hi], jittered, per-blob material scaling."""
    from ffea_amd import mesh as M
    dims = np.random.default_rng(seed).integers(lo, hi + 1, size=(count, 3))
    ms, xs = [], []
    for k, d in enumerate(dims):
        m = M.box_mesh(*map(int, d), jitter=jitter, seed=seed * 1000 + k)
        M.scale_material(m, 0.5, 2.0, seed=seed * 1000 + 500 + k)
        ms.append(m)
        xs.append(M.perturb(m, 0.02, seed=seed * 1000 + 900 + k))
    return ms, xs


def make_case(cfg):
    """Meshes (one per blob), parameters, scale, initial positions and description."""
    from ffea_amd import mesh as M, units as U
    cells, jitter, core, desc = CONFIGS[cfg]
    prm = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=1, calc_stokes=1, rng_seed=44)
    if cfg == "c3":
        ms, xs = box_batch(64, 24, 64, seed=31)
        return ms, prm, U.blob_scale(1e-9), xs, desc
    if cells is None:
        dims = np.random.default_rng(2).integers(6, 33, size=(256, 3))
        ms, xs = [], []
        for k, d in enumerate(dims):
            m = M.box_mesh(*map(int, d), jitter=jitter, seed=1000 + k)
            M.scale_material(m, 0.5, 2.0, seed=2000 + k)
            ms.append(m)
            xs.append(M.perturb(m, 0.02, seed=3000 + k))
        return ms, prm, U.blob_scale(1e-9), xs, desc
    m = M.box_mesh(*cells, jitter=jitter, seed=4)
    if core != 1.0:
        M.stiff_core(m, axis=0, factor=core)
    x0 = M.perturb(m, 0.02, seed=1)
    return [m], prm, U.blob_scale(1e-9), [x0], desc


RR_PERIOD = int(os.environ.get("FFH_RR_PERIOD", "16"))  # the library's true-residual period


def algorithmic_bytes(kernel, sizes, bm, bv=8):
    """Bytes each launch must move, each operand counted once (bm = bytes per matrix/element
    value: 8 on the FP64 path, 4 on the mixed path; bv = bytes per CG work-vector entry r, p, q,
    z: 8 on the FP64 path, 4 on the mixed path, whose iterate x, right-hand side f and D^-1 stay
    FP64).

    N linear nodes, E elements, B 3x3 blocks, S SELL slots (blocks + padding)."""
    N, E, B = sizes["num_linear"], sizes["num_elements"], sizes["num_blocks"]
    S = sizes.get("num_slots", B)
    v = 24 * N  # one FP64 3-vector over the linear nodes
    w = 3 * bv * N  # one work vector
    return {
        # enodes 16, J0^-1, vol0, A B G E (14 constants: FP64, or their FP32 copies on the mixed
        # path), the determinant's sign read and written (2); writes the 12 nodal forces
        "element": E * (16 + 14 * bm + 2 + 12 * bm) + v,
        # per slice its element list (4 B per slice-element pair, L pairs), enodes + A, B once
        # (32E), contribution codes (2 B per contribution, C of them: the diagonal and upper
        # blocks' element terms), per summed block (N diagonal + (B - N) / 2 upper) its count,
        # SELL position and mirror slot (6 B), the diagonal slots' flags (N), positions (32N),
        # drag (8N); writes every block once (9 bm B) and D^-1 (24N)
        "assemble": 4 * sizes.get("slice_elements", E) + 32 * E + 2 * sizes.get("contributions", 10 * E)
                    + 6 * (N + (B - N) // 2) + 65 * N + 9 * bm * B,
        # element forces + codes + per-node data; writes f, x (FP64) and r, p, q, z
        "rhs": 12 * bm * E + 16 * E + 39 * N + 2 * v + 4 * w,
        # matrix + column indices, z gathered once, p q read and written
        "cg_spmv": 9 * bm * B + 4 * B + 5 * w,
        # x read and written, D^-1; p, q, z and r read and written
        "cg_update": 3 * v + 5 * w,
        "integrate": 4 * v,
    }.get(kernel)


def assembly_lists(meshes):
    """Sizes of the assembly's lists for these meshes (blobs numbered on from slice boundaries):
    L, the distinct (element, 64-row slice) pairs, and C, the element terms of the summed blocks
    (each element's unpinned vertices on the diagonal, its unpinned vertex pairs above it)."""
    L = C = 0
    for m in meshes:
        t = np.asarray(m.topo)[:, :4].astype(np.int64)
        E = t.shape[0]
        sl = t // 64
        pairs = np.unique(np.concatenate([np.arange(E, dtype=np.int64) * (1 << 32) + sl[:, k] for k in range(4)]))
        L += int(pairs.size)
        free = np.ones(m.num_linear, bool)
        free[np.asarray(m.pinned, np.int64)[np.asarray(m.pinned) < m.num_linear]] = False
        f = free[t]
        nf = f.sum(axis=1)
        C += int(nf.sum() + (nf * (nf - 1) // 2).sum())
    return {"slice_elements": L, "contributions": C}


def true_residual_bytes(sizes):
    """Extra bytes of a mixed-path SpMV that also forms f - K x: x gathered and f read (FP64),
    r written (FP32)."""
    N = sizes["num_linear"]
    return 24 * N + 24 * N + 12 * N


def checks_per_solve(iters, period=RR_PERIOD):
    """True-residual passes of a mixed-path solve of `iters` iterations: every `period`
    iterations and at the final stop."""
    n = int(round(iters))
    return 1 + ((n - 1) // period if period > 0 else 0)


def run_gpu(ms, prm, sc, xs, precision, steps, warmup):
    """Warm-up, then two timed regions of `steps` steps: the first without per-kernel events
    (the throughput), the second with start/stop events carried in every dispatch packet on
    the blob's stream (the per-kernel durations)."""
    from ffea_amd.blob import Batch, Blob, package_seed
    t0 = time.time()
    if len(ms) == 1:
        b = Blob(ms[0], prm, sc, rng_state=package_seed(prm.rng_seed), precision=precision, pos=xs[0])
    else:
        b = Batch(ms, prm, sc, rng_state=package_seed(prm.rng_seed), precision=precision)
        b.set_positions(np.concatenate(xs) * sc)
    log(f"[{precision}] {len(ms)} blob(s) created in {time.time() - t0:.1f}s: {b.sizes()}, tuned {b.tuning()}")
    b.step(warmup)
    b.synchronize()
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step(1)
        iters.append(b.last_cg()[0])
    b.synchronize()
    dt = time.perf_counter() - t0
    b.set_profiling(True)
    t1 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    dt_prof = time.perf_counter() - t1
    kt = b.kernel_times()
    b.set_profiling(False)
    sizes = b.sizes()
    sizes.update(assembly_lists(ms))
    sizes["tuning"] = b.tuning()
    log(f"[{precision}] {steps} steps in {dt:.3f}s = {steps / dt:.1f} steps/s (with per-kernel events "
        f"{steps / dt_prof:.1f} steps/s), CG iterations mean {np.mean(iters):.2f}")
    b.close()
    return steps / dt, dt, kt, sizes, float(np.mean(iters)), dt_prof / steps


def algorithmic_bytes_mass(kernel, sizes, bm, solver):
    """The same count for the mass-ful solvers: the element kernel also gathers the node
    velocities (K_e u_e), the force kernel reads them for the drag (MassLumpedSolver: it writes
    f and x = f / m only), the mass matrix holds one FP64 value per block, the integration reads
    the velocity it updates.  No assembly."""
    N, B = sizes["num_linear"], sizes["num_blocks"]
    v = 24 * N
    base = algorithmic_bytes(kernel, sizes, bm)
    E = sizes["num_elements"]
    return {
        "element": base + v if base is not None else None,
        "rhs": (12 * bm * E + 16 * E + 39 * N + 3 * v + 8 * N) if solver == "masslumped" else base + v,
        "cg_spmv": 12 * B + 5 * v,
        "cg_update": 8 * v,
        "integrate": 5 * v,
    }.get(kernel)


def kernel_table(kt, sizes, bm, solver="CG_nomass", bv=8, solves=0, iters=0.0):
    """Per-kernel launches, average duration and achieved algorithmic GB/s.  On the mixed path
    (bv = 4) the SpMV's bytes include its true-residual passes, checks_per_solve of them per
    solve, spread over its launches."""
    out = {}
    for k, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        avg_s = ms / cnt / 1e3
        byts = algorithmic_bytes(k, sizes, bm, bv) if solver == "CG_nomass" else algorithmic_bytes_mass(k, sizes, bm, solver)
        if k == "cg_spmv" and bv == 4 and solver == "CG_nomass" and solves:
            byts += true_residual_bytes(sizes) * checks_per_solve(iters) * solves / cnt
        out[k] = {"launches": cnt, "avg_us": avg_s * 1e6, "total_ms": ms}
        if byts is not None:
            out[k].update(alg_bytes=int(byts), achieved_GBs=byts / avg_s / 1e9)
    return out


def massful_cpu(ms, p, sc, xs, max_seconds=20.0):
    """MassLumpedSolver steps of the reference's own objects (oracle/_ref fast build, OpenMP
    within the blob) on this host, a bounded sample; single blobs up to 2 M tets.  The
    reference's ConjugateGradientSolver is not timed: its init builds a dense N_t x N_t matrix
    (src/ConjugateGradientSolver.cpp:78-84)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if len(ms) != 1 or ms[0].num_elements > 2_000_000 or not os.path.exists(pyoracle.REF_FAST_SO):
        return None
    lib = pyoracle.Reference(pyoracle.REF_FAST_SO)
    lib.set_package_seed(lib.package_seed(p.rng_seed))
    s = lib.stream(per_thread=True)
    t0 = time.time()
    b = lib.blob(ms[0], p, sc, pos=xs[0])
    b.set_solver("masslumped")
    setup = time.time() - t0
    assert pyoracle.world_step(lib, [b], s) == 0
    n, t0 = 0, time.perf_counter()
    while True:
        assert pyoracle.world_step(lib, [b], s) == 0
        n += 1
        el = time.perf_counter() - t0
        if el > max_seconds or n >= 200:
            break
    b.close()
    log(f"[cpu reference masslumped] {n} steps in {el:.2f}s")
    return {"value": n / el, "unit": "steps/s", "cores": lib.num_threads(), "kind": "reference",
            "sample": f"{n} noise-on MassLumpedSolver steps of the same mesh after 1 warm-up step (reference "
                      f"objects -O3 -ffast-math, OpenMP within blob), setup {setup:.0f}s excluded"}


def massful(ms, prm, sc, xs, steps, warmup, cpu_seconds=0.0):
    """The mass-ful solvers on the same workload (FP64): ConjugateGradientSolver on the
    consistent mass matrix and MassLumpedSolver, steps/s and per-kernel times.  Time step 1 fs:
    the explicit mass-ful update of these nanometre elements is unstable at 10 fs (in the
    reference as here); the work per step does not depend on it."""
    import dataclasses
    from ffea_amd.blob import Batch, Blob, package_seed
    p = dataclasses.replace(prm, dt=prm.dt / 10)
    out = {"dt_s": 1e-15}
    for solver in ("CG", "masslumped"):
        if len(ms) == 1:
            b = Blob(ms[0], p, sc, rng_state=package_seed(p.rng_seed), pos=xs[0])
        else:
            b = Batch(ms, p, sc, rng_state=package_seed(p.rng_seed))
            b.set_positions(np.concatenate(xs) * sc)
        b.set_solver(solver)
        b.step(warmup)
        b.synchronize()
        iters = []
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step(1)
            iters.append(b.last_cg()[0])
        b.synchronize()
        dt = time.perf_counter() - t0
        b.set_profiling(True)
        b.step(steps)
        b.synchronize()
        kt = b.kernel_times()
        b.set_profiling(False)
        sizes = b.sizes()
        ke = b.kinetic_energy()
        b.close()
        log(f"[{solver}] {steps} steps in {dt:.3f}s = {steps / dt:.1f} steps/s, CG iterations mean "
            f"{np.mean(iters):.2f}, kinetic energy {ke:.4g} (3/2 N kT = {1.5 * sizes['num_linear'] * p.kT:.4g})")
        out[solver] = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps,
                       "cg_iters_mean": float(np.mean(iters)), "kinetic_energy_meso": ke,
                       "kernels": kernel_table(kt, sizes, 8, solver)}
    if cpu_seconds > 0:
        out["masslumped"]["cpu_baseline"] = massful_cpu(ms, p, sc, xs, cpu_seconds)
    return out


def boundary_faces(t4, p):
    """Outward-oriented boundary triangles of a linear tetrahedral mesh (faces of one element)."""
    f = np.concatenate([t4[:, [1, 2, 3]], t4[:, [0, 3, 2]], t4[:, [0, 1, 3]], t4[:, [0, 2, 1]]])
    opp = np.concatenate([t4[:, 0], t4[:, 1], t4[:, 2], t4[:, 3]])
    _, idx, cnt = np.unique(np.sort(f, axis=1), axis=0, return_index=True, return_counts=True)
    keep = np.sort(idx[cnt == 1])
    f, opp = f[keep].copy(), opp[keep]
    n = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    flip = (n * (p[opp] - p[f[:, 0]])).sum(axis=1) > 0
    f[flip, 1], f[flip, 2] = f[flip, 2].copy(), f[flip, 1].copy()
    return f


def lj_pairs(ps, fs, cut, es):
    """Face pairs the Lennard-Jones solve visits (other blob, neighbour cells with one period,
    normals not facing the same way), counted with the lookup grid of the first step."""
    cen = np.concatenate([p[f].mean(axis=1) for p, f in zip(ps, fs)])
    nrm = np.concatenate([np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]]) for p, f in zip(ps, fs)])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    blob = np.concatenate([np.full(f.shape[0], k) for k, f in enumerate(fs)])
    es = np.array(es)
    cell = np.floor(cen / cut).astype(np.int64) % es
    cid = (cell[:, 0] * es[1] + cell[:, 1]) * es[2] + cell[:, 2]
    members = {}
    for i, c in enumerate(cid):
        members.setdefault(c, []).append(i)
    pairs = 0
    for i in range(cen.shape[0]):
        for d in np.indices((3, 3, 3)).reshape(3, -1).T - 1:
            c = (cell[i] + d) % es
            js = np.array(members.get((c[0] * es[1] + c[1]) * es[2] + c[2], []), np.int64)
            if js.size:
                js = js[(js > i) & (blob[js] != blob[i])]
                pairs += int(((nrm[js] @ nrm[i]) <= 0).sum())
    return pairs


def lj_case(n=(48, 8, 48), gap=1.0):
    """Two Kuhn box blocks, one above the other, their facing 48 x 48 nm sides 1 nm apart, every
    boundary face active with the cubes_lj forcefield (Emin 2.537e17 J/m^4, Rmin 1 nm), centred
    in the lookup grid of 3 nm cells World::init would build."""
    from ffea_amd import mesh as M, units as U
    sc = U.blob_scale(1e-9)
    ms = [M.box_mesh(*n), M.box_mesh(*n)]
    ps = [ms[0].pos * sc, ms[1].pos * sc + np.array([0.0, (n[1] + gap) * sc, 0.0])]
    cut = 3e-9 / U.MESO_LENGTH
    allp = np.concatenate(ps)
    es = [2 * int(np.ceil(d / cut)) for d in allp.max(0) - allp.min(0)]
    shift = np.array(es) * cut / 2 - allp.mean(0)
    ps = [p + shift for p in ps]
    fs = [boundary_faces(m.topo[:, :4], p) for m, p in zip(ms, ps)]
    E = np.array([[2.537e17 * U.MESO_AREA ** 2 / U.MESO_ENERGY]])
    R = np.array([[1e-9 / U.MESO_LENGTH]])
    return ms, sc, ps, fs, E, R, cut, es


def ssint_leg(steps, warmup, cpu_seconds=0.0):
    """Lennard-Jones surface-surface interactions (FP64): steps/s of the two-block workload with
    and without the interaction, the interaction kernels' times, pair evaluations per second,
    and the oracle's C restatement timed on a bounded sample of the same steps (port)."""
    from ffea_amd import units as U
    from ffea_amd.blob import Batch, package_seed
    ms, sc, ps, fs, E, R, cut, es = lj_case()
    prm = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=1, calc_stokes=1, rng_seed=44)
    nf = [f.shape[0] for f in fs]
    faces = np.concatenate([fs[0], fs[1] + ms[0].pos.shape[0]])
    pairs = lj_pairs(ps, fs, cut, es)
    out = {"workload": f"two {ms[0].num_elements}-element Kuhn blocks 1 nm apart, {sum(nf)} active faces, "
                       f"grid {es} of {cut * U.MESO_LENGTH * 1e9:.0f} nm cells, es_update 1, noise on, dt 1e-14 s",
           "faces": sum(nf), "pairs_per_step": pairs}
    for on in (False, True):
        g = Batch(ms, prm, [sc, sc], rng_state=package_seed(44))
        g.set_positions(np.concatenate(ps))
        if on:
            g.set_ssint(faces, np.zeros(faces.shape[0], np.int32), E, R, es, cut, 0, 1)
        g.step(warmup)
        g.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        g.synchronize()
        dt = time.perf_counter() - t0
        key = "with" if on else "without"
        out[key] = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps}
        if on:
            g.set_profiling(True)
            g.step(steps)
            g.synchronize()
            kt = g.kernel_times()
            g.set_profiling(False)
            out["energy_J"] = g.ssint_energy() * U.MESO_ENERGY
            kf, kg = kt["ssint_forces"], kt["ssint_grid"]
            avg = kf[0] / kf[1] / 1e3
            out["kernels"] = {"ssint_forces": {"launches": kf[1], "avg_us": 1e6 * avg},
                              "ssint_grid": {"launches": kg[1], "avg_us": 1e3 * kg[0] / max(kg[1], 1)}}
            # each pair is evaluated by both of its faces
            out["pair_evaluations_per_s"] = 2 * pairs / avg
        g.close()
        log(f"[ssint {key}] {steps} steps in {dt:.3f}s = {steps / dt:.1f} steps/s")
    out["interaction_ms_per_step"] = out["with"]["ms_per_step"] - out["without"]["ms_per_step"]
    if cpu_seconds > 0:
        out["cpu_baseline"] = ssint_cpu(ms, prm, sc, ps, fs, E, R, cut, es, cpu_seconds)
    return out


def ssint_cpu(ms, prm, sc, ps, fs, E, R, cut, es, max_seconds):
    """The oracle's world step with and without the interaction on this host (one thread), a
    bounded sample; the difference is the interaction's time per step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle()
    per = {}
    for on in (False, True):
        o.set_package_seed(o.package_seed(44))
        s = o.stream()
        bo = []
        for m, p, f in zip(ms, ps, fs):
            b = o.blob(m, prm, sc)
            b.set_positions(p.ravel())
            b.set_surface_forces(f, [], np.zeros((0, 3)))
            b.set_ssint_types(np.zeros(f.shape[0], np.int32))
            bo.append(b)
        if on:
            o.set_ssint(E, R, es, cut, inc_self=0, es_update=1)
        n, t0 = 0, time.perf_counter()
        while True:
            assert pyoracle.world_step(o, bo, s) == 0
            n += 1
            el = time.perf_counter() - t0
            if el > max_seconds / 2 or n >= 20:
                break
        o.set_ssint()
        per[on] = el / n
        for b in bo:
            b.close()
    log(f"[cpu oracle ssint] {per[True]:.3f} s/step with, {per[False]:.3f} without")
    return {"value": 1.0 / per[True], "unit": "steps/s", "cores": 1, "kind": "port",
            "interaction_s_per_step": per[True] - per[False],
            "sample": "oracle (FP64 C restatement) world steps of the same two blocks with and without the "
                      "interaction, after set-up"}


def cpu_baseline(cfg, ms, prm, sc, xs, max_seconds=20.0):
    """Runs _cpu_baseline in a thread with a 2 GiB stack: the reference's SpMV keeps two
    3*N_t-double arrays on the stack (src/SparseMatrixFixedPattern.cpp:153-154), 44 MB at c1."""
    import threading
    res = {}
    old = threading.stack_size(2 << 30)
    try:
        t = threading.Thread(target=lambda: res.update(out=_cpu_baseline(cfg, ms, prm, sc, xs, max_seconds)))
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    return res.get("out")


def _cpu_baseline(cfg, ms, prm, sc, xs, max_seconds=20.0):
    """The reference's own compiled objects (oracle/_ref, -O3 -ffast-math, OpenMP within the
    blob) on this host, falling back to the FP64 restatement; a bounded sample of steps.  For a
    batch, the sample is the first blobs up to 1/8 of the elements, and the rate is scaled by
    the element ratio (the reference's step time is linear in the element count)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind, lib = None, None
    if os.path.exists(pyoracle.REF_FAST_SO):
        kind, lib = "reference", pyoracle.Reference(pyoracle.REF_FAST_SO)
    elif os.path.exists(pyoracle.ORACLE_SO):
        kind, lib = "port", pyoracle.Oracle()
    else:
        return None
    cores = os.cpu_count() or 1
    if kind == "reference":
        cores = lib.num_threads()
    seed = lib.package_seed(prm.rng_seed)
    lib.set_package_seed(seed)
    s = lib.stream(per_thread=True) if kind == "reference" else lib.stream()
    etot = sum(m.num_elements for m in ms)
    note = ""
    if len(ms) == 1 and etot > 2_000_000:
        # one large blob: the reference's setup alone would take minutes, so its step is timed on
        # a 48^3-cell box of the same construction and scaled by the element count (its step time
        # is linear in the elements)
        from ffea_amd import mesh as M
        cells, jitter, core, _ = CONFIGS[cfg]
        m1 = M.box_mesh(48, 48, 48, jitter=jitter, seed=4)
        if core != 1.0:
            M.stiff_core(m1, axis=0, factor=core)
        ms, xs = [m1], [M.perturb(m1, 0.02, seed=1)]
        note = f" on a 48^3-cell box of the same construction ({m1.num_elements} tets)"
    nb, esub = 0, 0
    while nb < len(ms) and (nb == 0 or esub < etot / 8):
        esub += ms[nb].num_elements
        nb += 1
    t0 = time.time()
    blobs = [lib.blob(m, prm, sc, pos=x) for m, x in zip(ms[:nb], xs[:nb])]
    setup = time.time() - t0
    log(f"[cpu {kind}] setup {setup:.1f}s")
    assert pyoracle.world_step(lib, blobs, s) == 0  # warm-up step
    if kind == "reference":
        lib.phase_times(reset=True)
    n, t0 = 0, time.perf_counter()
    while True:
        assert pyoracle.world_step(lib, blobs, s) == 0
        n += 1
        el = time.perf_counter() - t0
        if el > max_seconds or n >= 50:
            break
    phases = None
    if kind == "reference":
        ph = lib.phase_times(reset=True) / n
        build = sum(lib.time_build(b) for b in blobs)
        phases = {"element_loop_ms": 1e3 * ph[0], "force_gather_noise_ms": 1e3 * ph[1],
                  "matrix_build_ms": 1e3 * build, "cg_ms": 1e3 * (ph[2] - build),
                  "integration_ms": 1e3 * ph[3]}
        log(f"[cpu {kind}] per step: {phases}")
    for b in blobs:
        b.close()
    del s
    log(f"[cpu {kind}] {n} steps of {nb} blob(s) in {el:.2f}s")
    part = "" if nb == len(ms) else (f" restricted to its first {nb} of {len(ms)} blobs ({esub} of {etot} tets), "
                                     f"rate scaled by {esub}/{etot}")
    if note:
        part = note + f", rate scaled by {esub}/{etot}"
    return {"value": n / el * esub / etot, "unit": "steps/s", "cores": cores, "kind": kind,
            "host": host_description(), "phases_per_step": phases,
            "sample": f"{n} noise-on steps of the same {cfg} workload{part} after 1 warm-up step "
                      f"({'reference objects -O3 -ffast-math, OpenMP within blob' if kind == 'reference' else 'FP64 restatement, 1 thread'}); "
                      f"host {platform.processor() or platform.machine()}, setup {setup:.0f}s excluded"}


def host_description() -> dict:
    """CPU model and logical CPU count of this host (the CPU baseline's machine)."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": usable, "logical_cpus": os.cpu_count() or 1}


def hbm_spmv_leg(steps=10, warmup=3):
    """The FP32 block SpMV where its matrix streams from HBM: the c4 workload (192x96x48 box,
    5.3 M tets; its FP32 matrix and columns, 540 MB, exceed the 256 MB Infinity Cache) on the
    mixed path, per-launch SpMV times from the in-packet events; true residual only at each
    solve's stop (FFH_RR_PERIOD=0), whose extra bytes are counted."""
    from ffea_amd.blob import Blob, package_seed
    ms, prm, sc, xs, desc = make_case("c4")
    old = os.environ.get("FFH_RR_PERIOD")
    os.environ["FFH_RR_PERIOD"] = "0"
    try:
        b = Blob(ms[0], prm, sc, rng_state=package_seed(prm.rng_seed), precision="mixed", pos=xs[0])
    finally:
        if old is None:
            os.environ.pop("FFH_RR_PERIOD")
        else:
            os.environ["FFH_RR_PERIOD"] = old
    b.step(warmup)
    b.set_profiling(True)
    its = []
    for _ in range(steps):
        b.step(1)
        its.append(b.last_cg()[0])
    b.synchronize()
    kt = b.kernel_times()
    b.set_profiling(False)
    sizes = b.sizes()
    tuning = b.tuning()
    b.close()
    ms_tot, cnt = kt["cg_spmv"]
    avg_s = ms_tot / cnt / 1e3
    alg = algorithmic_bytes("cg_spmv", sizes, 4, 4) + true_residual_bytes(sizes) * steps / cnt
    out = {"workload": f"c4: {desc}", "precision": "mixed (FP32 matrix and work vectors)",
           "launches": cnt, "avg_launch_us": avg_s * 1e6, "alg_bytes_per_launch": int(alg),
           "bound": "hbm", "achieved": alg / avg_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": alg / avg_s / 1e9 / HBM_PEAK_GBS, "cg_iters_mean": float(np.mean(its)),
           "traffic": measured_traffic("c4", "cg_spmv", "mixed", tuning), "launch_shapes": tuning}
    log(f"[hbm spmv c4 mixed] {out['avg_launch_us']:.1f} us per launch, {out['achieved']:.0f} GB/s = "
        f"{out['frac']:.3f} of peak")
    return out


def equipartition(ms, prm, sc, xs, precision, steps=4000, start=2000, every=100):
    """Noise-on run from the bench's initial state: mean strain energy over steps start..steps
    (sampled every `every` steps, Blob::make_measurements' observable) against the equipartition
    value (3 N - 6) kT / 2 of a free blob -- the check of tests/physics/sphere_63_120_nomass."""
    from ffea_amd.blob import Batch, Blob, package_seed
    if len(ms) == 1:
        b = Blob(ms[0], prm, sc, rng_state=package_seed(prm.rng_seed), precision=precision, pos=xs[0])
        pinned = ms[0].pinned.size
    else:
        b = Batch(ms, prm, sc, rng_state=package_seed(prm.rng_seed), precision=precision)
        b.set_positions(np.concatenate(xs) * sc)
        pinned = sum(m.pinned.size for m in ms)
    nlin = b.sizes()["num_linear"]
    modes = 3 * (nlin - pinned) - 6 * len(ms)
    e = []
    for k in range(0, steps + 1, every):
        if k >= start:
            e.append(b.strain_energy())
        if k < steps:
            b.step(every)
    b.close()
    ratio = float(np.mean(e)) / (modes * prm.kT / 2)
    log(f"[equipartition {precision}] <U> / ((3N-6) kT/2) = {ratio:.4f} over steps {start}..{steps}")
    return {"steps": steps, "average_from_step": start, "sample_every": every, "modes": modes,
            "mean_energy_over_theory": ratio}


def long_trajectory(cfg, ms, prm, sc, xs, steps):
    """A long noise-off trajectory of the GPU paths against the reference's own objects in its
    fast build (oracle/_ref/libffea_ref_fast.so: -O3 -ffast-math, OpenMP on every host core;
    the strict build would take the better part of an hour for 1,000 c1 steps): positions and
    velocities every `steps // 10` steps, relative to the displacement."""
    import threading
    from ffea_amd import units as U
    from ffea_amd.blob import Blob
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if len(ms) != 1 or not os.path.exists(pyoracle.REF_FAST_SO):
        return None
    m, x0 = ms[0], xs[0]
    p0 = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=0, calc_stokes=1, rng_seed=prm.rng_seed)
    every = max(1, steps // 10)
    ref = []

    def cpu():
        lib = pyoracle.Reference(pyoracle.REF_FAST_SO)
        lib.set_package_seed(lib.package_seed(p0.rng_seed))
        s = lib.stream(per_thread=True)
        b = lib.blob(m, p0, sc, pos=x0)
        for k in range(1, steps + 1):
            assert pyoracle.world_step(lib, [b], s) == 0
            if k % every == 0:
                ref.append((k, b.get("pos"), b.get("vel")))
                log(f"[trajectory {cfg}] reference step {k}")
        b.close()

    old = threading.stack_size(2 << 30)
    try:
        t = threading.Thread(target=cpu)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    xs0 = x0.ravel() * sc
    out = {"reference": "reference objects, fast build (OpenMP)", "steps": steps, "noise": "off"}
    for prec in ("fp64", "mixed"):
        b = Blob(m, p0, sc, rng_state=np.arange(1, 7, dtype=np.uint64), precision=prec, pos=x0)
        worst_p = worst_v = 0.0
        done = 0
        for k, pr, vr in ref:
            b.step(k - done)
            done = k
            p, v = b.positions().ravel(), b.velocities().ravel()
            worst_p = max(worst_p, float(np.linalg.norm(p - pr) / np.linalg.norm(pr - xs0)))
            worst_v = max(worst_v, float(np.linalg.norm(v - vr) / np.linalg.norm(vr)))
        b.close()
        out[prec] = {"max_position_rel_to_displacement": worst_p, "max_velocity_rel": worst_v}
    log(f"[trajectory {cfg}] {out}")
    return out


def mixed_vs_fp64(ms, prm, sc, xs, steps=20):
    """The mixed path against the FP64 path (whose parity with the reference is established at
    c0/c1 and in the tests): noise-off positions and velocities after `steps` steps from the
    bench's initial state, and one CG solution of the same force vector."""
    from ffea_amd import units as U
    from ffea_amd.blob import Batch, Blob
    p0 = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=0, calc_stokes=1, rng_seed=prm.rng_seed)
    out = {}
    for prec in ("fp64", "mixed"):
        if len(ms) == 1:
            b = Blob(ms[0], p0, sc, rng_state=np.arange(1, 7, dtype=np.uint64), precision=prec, pos=xs[0])
        else:
            b = Batch(ms, p0, sc, rng_state=np.arange(1, 7, dtype=np.uint64), precision=prec)
            b.set_positions(np.concatenate(xs) * sc)
        b.update_internal_forces()
        if prec == "fp64":
            b.update_positions()
            f = b.rhs()
        v = b.solve(f).ravel()
        b.set_positions(np.concatenate(xs) * sc)
        b.step(steps)
        out[prec] = (v, b.positions().ravel(), b.velocities().ravel())
        b.close()
    x0 = np.concatenate(xs).ravel() * sc
    (v64, p64, w64), (vm, pm, wm) = out["fp64"], out["mixed"]
    rel = lambda a, r: float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300))
    res = {"steps": steps, "noise": "off", "cg_solution_rel": rel(vm, v64), "velocity_rel": rel(wm, w64),
           "position_rel_to_displacement": float(np.linalg.norm(pm - p64) / np.linalg.norm(p64 - x0))}
    log(f"[mixed vs fp64] {res}")
    return res


def parity_vs_reference(cfg, ms, prm, sc, xs, steps):
    """Noise-off parity of the GPU paths against the reference's own objects built strict
    (oracle/_ref/libffea_ref.so, one thread; the FP64 restatement if absent): the CG solution
    of the first solve (same force vector) and positions / velocities after `steps` steps."""
    import threading
    from ffea_amd import units as U
    from ffea_amd.blob import Blob
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if len(ms) != 1:
        return None
    m, x0 = ms[0], xs[0]
    p0 = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=0, calc_stokes=1, rng_seed=prm.rng_seed)
    ref = {}

    def cpu():
        kind = "reference" if os.path.exists(pyoracle.REF_SO) else "port"
        lib = pyoracle.Reference() if kind == "reference" else pyoracle.Oracle()
        lib.set_package_seed(lib.package_seed(p0.rng_seed))
        s = lib.stream()
        log(f"[parity {cfg}] {kind}: setting up")
        b = lib.blob(m, p0, sc, pos=x0)
        inv, rhs = b.forces_only(s)
        st, v = b.solve(rhs)
        b.close()
        b = lib.blob(m, p0, sc, pos=x0)
        log(f"[parity {cfg}] {kind}: {steps} steps")
        t0 = time.perf_counter()
        for k in range(steps):
            assert pyoracle.world_step(lib, [b], s) == 0
            log(f"[parity {cfg}] step {k + 1} done")
        ref.update(kind=kind, rhs=rhs, v=v, pos=b.get("pos"), vel=b.get("vel"), secs=time.perf_counter() - t0)
        b.close()

    old = threading.stack_size(2 << 30)
    try:
        t = threading.Thread(target=cpu)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "pos" not in ref:
        return None
    xs0 = x0.ravel() * sc
    out = {"reference": ref["kind"], "steps": steps, "noise": "off", "cpu_seconds": ref["secs"]}
    for prec in ("fp64", "mixed"):
        b = Blob(m, p0, sc, rng_state=np.arange(1, 7, dtype=np.uint64), precision=prec, pos=x0)
        b.update_internal_forces()
        v = b.solve(ref["rhs"]).ravel()
        b.set_positions(x0 * sc)
        b.step(steps)
        p, vel = b.positions().ravel(), b.velocities().ravel()
        b.close()
        rel = lambda a, r: float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300))
        out[prec] = {"cg_solution_rel": rel(v, ref["v"]), "velocity_rel": rel(vel, ref["vel"]),
                     "position_rel_to_displacement": float(np.linalg.norm(p - ref["pos"]) /
                                                           np.linalg.norm(ref["pos"] - xs0)),
                     "position_rel": rel(p, ref["pos"])}
    log(f"[parity {cfg}] vs {ref['kind']} over {steps} noise-off steps: {out}")
    return out


def kernel_symbol(kernel, precision, tuning):
    """The kernel's name in rocprofv3 records (template arguments, no namespace or parameters)
    for the launch shapes the tuner picked; massless solver."""
    t = "double" if precision == "fp64" else "float"
    u = tuning.get("assembly_tiles", 2)
    waves = tuning.get("assembly_block", 512) // 64
    nt = "true" if tuning.get("spmv_nontemporal") else "false"
    spmv = "k_cg_spmv1" if tuning.get("spmv_waves_per_slice") == 1 else "k_cg_spmv"
    mixed = precision != "fp64"
    return {"element": f"k_element<{t}, {t}, false>",
            "assemble": f"k_assemble_slice<{t}, {waves}, {u}, true>",
            "rhs": f"k_rhs<{t}, {t}>", "cg_spmv": f"{spmv}<{t}, {nt}, false>",
            "cg_update": "k_cg_update<float, true>" if mixed else "k_cg_update<double, false>",
            "integrate": "k_integrate"}.get(kernel)


def measured_traffic(cfg, kernel, precision, tuning):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary of this workload (profiles/traffic_<cfg>.json, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{cfg}.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))["kernels"].get(kernel_symbol(kernel, precision, tuning))
    return None if rec is None else rec["traffic_bytes"]


def copy_calibration(mbytes: int) -> dict:
    """One device-to-device copy of `mbytes` MiB through the HIP runtime (a plain copy kernel of
    known size): the reference point that scales rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(mbytes) << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        return {"error": "hipMalloc failed"}
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    return {"bytes": n, "seconds": dt}


def cg_phase(b, steps, kind):
    """`steps` time steps with preconditioner `kind`; CG time per step from the kernel events
    (force gather + CG init, SpMV, update, and the network's kernels), per-blob iterations."""
    b.set_preconditioner(kind)
    b.step(2)
    b.set_profiling(True)
    its = []
    b.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step(1)
        its.append(b.last_cg_per_blob()[0].copy() if hasattr(b, "last_cg_per_blob") else np.array([b.last_cg()[0]]))
    b.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kt = b.kernel_times()
    b.set_profiling(False)
    cg_ms = sum(v[0] for k, v in kt.items() if k not in ("element", "assemble", "integrate")) / steps
    its = np.array(its)
    return {"cg_ms_per_solve": cg_ms, "ms_per_step_with_events": 1e3 * wall,
            "iterations_mean_per_blob": float(its.mean()), "iterations_max": int(its.max()),
            "launched_iterations_mean": float(its.max(axis=1).mean()),
            "kernels_us": {k: v[0] / v[1] * 1e3 for k, v in kt.items() if v[1]},
            "launches_per_step": {k: v[1] / steps for k, v in kt.items() if v[1]}}


def kappa_jacobi(mesh, x0, prm, sc, dims):
    """Extreme eigenvalues of the Jacobi-preconditioned viscosity matrix D^-1 K (the operator
    Jacobi-PCG iterates on) of one box mesh, after one step from x0: the scalar diagonal of K by
    probing (nodes (i, j, k) of one residue class mod 3 share no element, so 27 x 3 applies of K
    give it exactly), then Lanczos (scipy eigsh) on D^-1/2 K D^-1/2, whose eigenvalues are D^-1 K's,
    with K applied on the GPU through Solver::apply_matrix.  Also the CG iteration bound for the
    relative tolerance eps at that condition number, ln(2 / eps) / ln((sqrt k + 1) / (sqrt k - 1))."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    from ffea_amd.blob import Blob, package_seed
    b = Blob(mesh, prm, sc, rng_state=package_seed(prm.rng_seed), precision="fp64", pos=x0)
    b.step(1)
    nl, nt = mesh.num_linear, mesh.num_nodes
    nx, ny, nz = dims
    ids = np.arange(nl)
    i, j, k = ids % (nx + 1), (ids // (nx + 1)) % (ny + 1), ids // ((nx + 1) * (ny + 1))
    color = (i % 3) + 3 * (j % 3) + 9 * (k % 3)

    def apply(v_lin):
        full = np.zeros(3 * nt)
        full[:3 * nl] = v_lin
        return b.apply_matrix(full)[:3 * nl]

    diag = np.zeros(3 * nl)
    for col in range(27):
        sel = np.nonzero(color == col)[0]
        for c in range(3):
            v = np.zeros(3 * nl)
            v[3 * sel + c] = 1.0
            diag[3 * sel + c] = apply(v)[3 * sel + c]
    dm = 1.0 / np.sqrt(diag)
    op = LinearOperator((3 * nl, 3 * nl), matvec=lambda v: dm * apply(dm * np.asarray(v).ravel()), dtype=np.float64)
    lmax = float(eigsh(op, k=1, which="LA", tol=1e-4, return_eigenvectors=False)[0])
    lmin = float(eigsh(op, k=1, which="SA", tol=1e-4, return_eigenvectors=False)[0])
    b.close()
    kappa = lmax / lmin
    rho = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
    bound = int(np.ceil(np.log(2.0 / np.sqrt(prm.epsilon2)) / -np.log(rho)))
    return {"cells": [int(nx), int(ny), int(nz)], "linear_nodes": int(nl), "lambda_min": lmin, "lambda_max": lmax,
            "kappa": kappa, "cg_iteration_bound": bound}


def run_c3(a):
    """Learned preconditioner: harvest, train, evaluate on held-out meshes."""
    from ffea_amd import units as U
    from ffea_amd.blob import Batch, package_seed
    prm = U.SimParams.from_si(dt=1e-14, epsilon=1e-6, calc_noise=1, calc_stokes=1, rng_seed=44)
    sc = U.blob_scale(1e-9)
    t0 = time.time()
    tm, tx = box_batch(64, 6, 32, seed=13)
    tb = Batch(tm, prm, sc, rng_state=package_seed(prm.rng_seed), precision="mixed")
    tb.set_positions(np.concatenate(tx) * sc)
    tb.mpnn_init(seed=1)
    tb.mpnn_harvest(1200)
    steps = 0
    while tb.mpnn_harvest_count()[1] < 20000 and steps < 200:
        tb.step(1)
        steps += 1
    snaps, pairs = tb.mpnn_harvest_count()
    log(f"[c3] harvested {pairs} pairs in {snaps} snapshots over {steps} steps ({time.time() - t0:.1f}s)")
    t1 = time.perf_counter()
    loss = tb.mpnn_train(epochs=a.epochs, lr=2e-3, seed=1)
    train_s = time.perf_counter() - t1
    log(f"[c3] trained {a.epochs} epochs in {train_s:.1f}s, loss {loss[0]:.4f} -> {loss[-1]:.4f}")
    w = tb.mpnn_weights()
    train_sizes = tb.sizes()
    tb.close()
    ms, prm, sc, xs, desc = make_case("c3")
    res = {}
    for prec in ("fp64", "mixed"):
        t0 = time.time()
        b = Batch(ms, prm, sc, rng_state=package_seed(prm.rng_seed), precision=prec)
        b.set_positions(np.concatenate(xs) * sc)
        log(f"[c3 {prec}] {len(ms)} held-out blobs created in {time.time() - t0:.1f}s: {b.sizes()}")
        b.mpnn_set_weights(w)
        b.step(a.warmup)
        res[f"{prec}_jacobi"] = cg_phase(b, a.steps, "jacobi")
        res[f"{prec}_learned"] = cg_phase(b, a.steps, "learned")
        for k in ("jacobi", "learned"):
            r = res[f"{prec}_{k}"]
            log(f"[c3 {prec} {k}] CG {r['cg_ms_per_solve']:.2f} ms/solve, iterations per blob "
                f"{r['iterations_mean_per_blob']:.2f} (max {r['iterations_max']})")
        sizes = b.sizes()
        b.close()
    base = res["fp64_jacobi"]["cg_ms_per_solve"]
    # condition numbers of D^-1 K on the smallest held-out meshes (the bound on what any
    # preconditioner can save over Jacobi-PCG)
    dims = np.random.default_rng(31).integers(24, 65, size=(64, 3))
    order = np.argsort(np.prod(dims + 1, axis=1))[:a.kappa_meshes]
    kappas = []
    for m_i in order:
        t0 = time.time()
        kk = kappa_jacobi(ms[m_i], xs[m_i], prm, sc, dims[m_i])
        kk["blob"] = int(m_i)
        kappas.append(kk)
        log(f"[c3 kappa] blob {m_i} {kk['cells']}: lambda {kk['lambda_min']:.4f} .. {kk['lambda_max']:.4f}, "
            f"kappa {kk['kappa']:.3f}, CG bound {kk['cg_iteration_bound']} iterations ({time.time() - t0:.1f}s)")
    out = {
        "metric": "c3: CG time-to-solution at eps=1e-6, learned-preconditioned mixed flexible CG vs plain FP64 "
                  "Jacobi-PCG (ratio of CG ms per solve)",
        "value": base / res["mixed_learned"]["cg_ms_per_solve"],
        "unit": "x (higher is better)",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": res["mixed_learned"]["ms_per_step_with_events"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16 network, f32/f64 CG",
        "data": "synthetic: seeded jittered box meshes, RngStream noise",
        "config": {"workload": f"c3: {desc}", "held_out_linear_nodes": sizes["num_linear"],
                   "held_out_elements": sizes["num_elements"], "train_linear_nodes": train_sizes["num_linear"],
                   "parallelism": "single GPU"},
        "training": {"pairs": pairs, "snapshots": snaps, "epochs": a.epochs, "seconds": train_s,
                     "loss_first_epoch": float(loss[0]), "loss_last_epoch": float(loss[-1])},
        "solvers": res,
        "speedup_mixed_jacobi_vs_fp64_jacobi": base / res["mixed_jacobi"]["cg_ms_per_solve"],
        "jacobi_condition_numbers": kappas,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c1", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--epochs", type=int, default=20, help="c3: training epochs")
    ap.add_argument("--kappa-meshes", type=int, default=6,
                    help="c3: held-out meshes (smallest first) whose D^-1 K condition number is measured")
    ap.add_argument("--parity-steps", type=int, default=None,
                    help="noise-off steps compared with the reference's strict objects (0: skip; default 10 "
                         "for c0 and c1, 0 above: their setup alone takes minutes on the host)")
    ap.add_argument("--trajectory-steps", type=int, default=0,
                    help="noise-off steps compared with the reference's fast build (0: skip; c1 asks for 1000)")
    ap.add_argument("--equipartition-steps", type=int, default=4000,
                    help="noise-on steps of the equipartition check, averaged over the second half (0: skip)")
    ap.add_argument("--massful-steps", type=int, default=100,
                    help="timed steps of each mass-ful solver on the same workload (0: skip)")
    ap.add_argument("--ssint-steps", type=int, default=50,
                    help="timed steps of the Lennard-Jones surface-interaction workload (0: skip)")
    ap.add_argument("--hbm-steps", type=int, default=10,
                    help="c1: also time the FP32 SpMV of the c4 workload, whose matrix streams from HBM (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="also run one device-to-device copy of this size (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported "
              "(independent blobs shard by blob with no communication; one large blob by node slabs with a "
              "boundary exchange and two scalar all-reduces per CG iteration)",
              file=sys.stderr)
        sys.exit(2)
    if a.config == "c3":
        return run_c3(a)
    ms, prm, sc, xs, desc = make_case(a.config)
    log(f"workload {a.config}: {desc}")
    sps, wall, kt, sizes, iters, prof_step = run_gpu(ms, prm, sc, xs, "fp64", a.steps, a.warmup)
    ktab = kernel_table(kt, sizes, 8)
    mixed = None
    if not a.no_mixed:
        sps_m, wall_m, kt_m, sizes_m, iters_m, prof_m = run_gpu(ms, prm, sc, xs, "mixed", a.steps, a.warmup)
        mixed = {"steps_per_s": sps_m, "ns_per_day": sps_m * 1e-5 * 86400, "ms_per_step": 1e3 * wall_m / a.steps,
                 "cg_iters_mean": iters_m, "kernels": kernel_table(kt_m, sizes_m, 4, bv=4, solves=a.steps, iters=iters_m)}
    dom = max(ktab, key=lambda k: ktab[k]["total_ms"])
    d = ktab[dom]
    roof = {"bound": "hbm", "kernel": dom, "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": d["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": measured_traffic(a.config, dom, "fp64", sizes.get("tuning", {})),
            "alg_bytes_per_launch": d["alg_bytes"], "avg_launch_us": d["avg_us"]}
    cpu = None if a.no_cpu else cpu_baseline(a.config, ms, prm, sc, xs, a.cpu_seconds)
    parity = None
    psteps = a.parity_steps if a.parity_steps is not None else (10 if a.config in ("c0", "c1") else 0)
    if not a.no_cpu and psteps > 0:
        parity = parity_vs_reference(a.config, ms, prm, sc, xs, psteps)
    mvf = None if a.no_mixed else mixed_vs_fp64(ms, prm, sc, xs)
    traj = long_trajectory(a.config, ms, prm, sc, xs, a.trajectory_steps) if a.trajectory_steps > 0 else None
    equi = None
    if a.equipartition_steps > 0:
        equi = {p: equipartition(ms, prm, sc, xs, p, a.equipartition_steps, a.equipartition_steps // 2)
                for p in (("fp64",) if a.no_mixed else ("fp64", "mixed"))}
    calib = copy_calibration(a.copy_calibration) if a.copy_calibration else None
    hbm = hbm_spmv_leg(a.hbm_steps, a.warmup) if (a.config == "c1" and a.hbm_steps > 0) else None
    mass = None
    if a.massful_steps > 0:
        mass = massful(ms, prm, sc, xs, a.massful_steps, a.warmup, 0.0 if a.no_cpu else min(a.cpu_seconds, 15.0))
    ssint = None
    if a.ssint_steps > 0:
        ssint = ssint_leg(a.ssint_steps, a.warmup, 0.0 if a.no_cpu else min(a.cpu_seconds, 20.0))
    out = {
        "metric": f"steps/s at CG eps=1e-6 (FP64 path), {a.config}",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Kuhn box mesh, seeded Gaussian initial displacement, RngStream noise",
        "config": {"workload": f"{a.config}: {desc}; eps 1e-6, noise on, dt 1e-14 s", "blobs": len(ms),
                   "linear_nodes": sizes["num_linear"], "elements": sizes["num_elements"],
                   "blocks": sizes["num_blocks"], "parallelism": "single GPU",
                   "launch_shapes": sizes["tuning"]},
        "ns_per_day": sps * 1e-5 * 86400,
        "cg_iters_mean": iters,
        "ms_per_step_with_kernel_events": 1e3 * prof_step,
        "kernels": ktab,
        "mixed": mixed,
        "roofline": roof,
        "spmv_hbm": hbm,
        "cpu_baseline": cpu,
        "parity_vs_reference": parity,
        "mixed_vs_fp64": mvf,
        "trajectory_vs_reference": traj,
        "equipartition": equi,
        "massful": mass,
        "ssint": ssint,
    }
    if calib:
        out["copy_calibration"] = calib
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
