"""World-level driver: a .ffea script's dynamic blobs stepped together on one GPU
(ffea_amd.blob.Batch), with FFEA's output and checkpoint files.  Blobs use the massless
CG_nomass solver, or one of the mass-ful solvers (CG, masslumped); all blobs of a script use the
same one.

Follows World::run's order per step (src/World.cpp:1811-2043):
  1. at check steps, the RngStream state is written to the checkpoint file (before the element
     phase consumes draws; World::print_checkpoints, src/World.cpp:4234-4280);
  2. element phase of every blob (Blob::update_internal_forces);
  3. at check steps, a trajectory frame and a measurement line
     (World::print_trajectory_and_measurement_files, src/World.cpp:4125-4231): node positions
     at the start of the step, strain energy of this element phase, centroid and RMSD;
  4. position phase of every blob (Blob::update_positions).
Steps run from step_initial to num_steps inclusive.

File formats (written as the reference writes them):
  .ftj  header (src/World.cpp:532-548), frames of "Blob b, Conformation 0, step s" + DYNAMIC +
        one "%e" x 10 line per node (Blob::write_nodes_to_file src/Blob.cpp:1143-1180: positions in
        m, then 7 zeros on the massless path; with a mass-ful solver velocities in m/s, the
        potential (0 without electrostatics) and the node force, which World::run has set to zero
        before the frame is written, src/World.cpp:1907), "*", the conformation-change block, "*";
  .fm   "Measurements:" then Time [KineticEnergy when a blob has mass] StrainEnergy
        Centroid.x/y/z RMSD in "%-14.6e" columns (World::write_measurements_to_file,
        src/World.cpp:4342-4366);
  .fcp  "RNGStreams dedicated to the thermal stress: K" then K lines of six integers.
Restart (restart = 1, checkpoint_in; src/World.cpp:645-800): positions (and, with a mass-ful
solver, velocities: Blob::read_nodes_from_file, src/Blob.cpp:1268-1306) from the last complete
trajectory frame (which is removed and written again), the last measurement line removed,
RngStream state from checkpoint_in, stepping resumes at that frame's step.

Placement (World::init, src/World.cpp:2660-2705, :455-506): a blob's <centroid> / <rotation>,
then move_into_box (default 1), after the rest state is taken from the positions as loaded.

Deliberate differences: the checkpoint file is rewritten whole at each check (the reference
rewinds without truncating); the .fm header lists this implementation's parameters; default
output names are <script>.ftj / <script>.fm / <script>.fcp.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from . import _lib
from . import formats as F
from . import units as U
from .blob import Batch, package_seed


# ---------------------------------------------------------------------------------------------
# file formats
# ---------------------------------------------------------------------------------------------
def traj_header(nodes_per_blob) -> str:
    s = "FFEA_trajectory_file\n\nInitialisation:\nNumber of Blobs %d\nNumber of Conformations" % len(nodes_per_blob)
    s += "".join(" %d" % 1 for _ in nodes_per_blob) + "\n"
    for b, n in enumerate(nodes_per_blob):
        s += "Blob %d:\t" % b + "Conformation %d Nodes %d\t" % (0, n) + "\n"
    return s + "\n*\n"


def traj_frame(step: int, positions_m, velocities=None) -> str:
    """One frame; positions_m: per blob [N_t, 3] in metres; velocities (mass-ful solvers): per
    blob [N_t, 3] in m/s, or None for the massless layout."""
    out = []
    for b, p in enumerate(positions_m):
        out.append("Blob %d, Conformation %d, step %d\nDYNAMIC\n" % (b, 0, step))
        if velocities is None:
            out.append("".join("%e %e %e %e %e %e %e %e %e %e\n" % (x, y, z, 0., 0., 0., 0., 0., 0., 0.)
                               for x, y, z in np.asarray(p, np.float64)))
        else:
            pv = np.concatenate([np.asarray(p, np.float64), np.asarray(velocities[b], np.float64)], axis=1)
            out.append("".join("%e %e %e %e %e %e %e %e %e %e\n" % (x, y, z, u, v, w, 0., 0., 0., 0.)
                               for x, y, z, u, v, w in pv))
    out.append("*\nConformation Changes:\n")
    out.append("".join("Blob %d: Conformation %d -> Conformation %d\n" % (b, 0, 0) for b in range(len(positions_m))))
    out.append("*\n")
    return "".join(out)


MEAS_COLUMNS = ("Time", "StrainEnergy", "Centroid.x", "Centroid.y", "Centroid.z", "RMSD")
MEAS_COLUMNS_MASS = ("Time", "KineticEnergy") + MEAS_COLUMNS[1:]


def meas_columns(mass_in_system: bool, springs: bool, ssint: bool = False):
    """World::init's measurement header (src/World.cpp:565-583) for the columns of this path."""
    return ((MEAS_COLUMNS_MASS if mass_in_system else MEAS_COLUMNS) + (("SpringEnergy",) if springs else ())
            + (("SurfSurfEnergy",) if ssint else ()))


def meas_header_line(cols) -> str:
    return "".join(("%-15s" if c == "SurfSurfEnergy" else "%-14s") % c for c in cols) + "\n"


def meas_line(values, ssint: bool = False) -> str:
    """World::write_measurements_to_file (src/World.cpp:4342-4365): %-14.6e, the surface-surface
    energy (the last column when present) %-15.6e."""
    fm = ["%-14.6e"] * len(values)
    if ssint:
        fm[-1] = "%-15.6e"
    return "".join(f % v for f, v in zip(fm, values)) + "\n"


def checkpoint_text(states) -> str:
    s = "RNGStreams dedicated to the thermal stress: %d\n" % len(states)
    return s + "".join("%d %d %d %d %d %d\n" % tuple(int(v) for v in st) for st in states)


def read_checkpoint(path):
    """States of the thermal-stress streams of a .fcp file ([K, 6] uint64)."""
    lines = open(path).read().split("\n")
    head = lines[0].split()
    if not lines[0].startswith("RNGStreams dedicated to the thermal stress:"):
        raise F.FormatError(f"{path}: not an FFEA checkpoint file")
    k = int(head[-1])
    st = np.array([[int(v) for v in lines[1 + i].split()[:6]] for i in range(k)], np.uint64)
    return st


def read_trajectory(path):
    """Complete frames of a .ftj file: list of (step, [per-blob [N_t, 10] float64], start offset).
    A frame is complete when its conformation-change block is closed by '*'."""
    data = open(path, "rb").read()
    text = data.decode()
    head_end = text.index("\n*\n") + 3
    frames, pos = [], head_end
    lines = text[head_end:].split("\n")
    off = head_end
    i = 0

    def take():
        nonlocal i, off
        line = lines[i]
        i += 1
        off += len(line) + 1
        return line

    try:
        while i < len(lines) and lines[i].startswith("Blob "):
            start = off
            blobs, step = [], None
            while lines[i].startswith("Blob ") and ", step " in lines[i]:
                h = take()
                step = int(h.rsplit("step", 1)[1])
                state = take()
                rows = []
                if state != "STATIC":
                    while lines[i] and lines[i][0] not in "B*":
                        rows.append([float(v) for v in take().split()])
                blobs.append(np.array(rows))
            if take() != "*" or take() != "Conformation Changes:":
                break
            while lines[i].startswith("Blob "):
                take()
            if take() != "*":
                break
            frames.append((step, blobs, start))
    except (IndexError, ValueError):
        pass  # a frame cut off part-way: not complete
    return frames


# ---------------------------------------------------------------------------------------------
# placing the blobs (World::init, src/World.cpp:2660-2705 and :455-506)
# ---------------------------------------------------------------------------------------------
def _seq_sum(x):
    """Sum in index order, as the reference's loops add (numpy's sum is pairwise)."""
    return float(np.add.accumulate(np.asarray(x, np.float64))[-1]) if len(x) else 0.0


def _centroid(p):
    """Blob::get_centroid (src/Blob.cpp:971-983): node sums divided by the node count."""
    n = p.shape[0]
    return np.array([_seq_sum(p[:, k]) / n for k in range(3)])


def _vec3(text):
    return [float(v) for v in text.strip().strip("()").split(",")]


def _rotation_matrix(vals):
    """The float matrix Blob::rotate uses (src/Blob.cpp:804-828): three angles in degrees,
    converted and combined in single precision (libm's cosf / sinf), or nine entries."""
    f = np.float32
    if len(vals) == 9:
        return [f(v) for v in vals]
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.cosf.restype = libm.sinf.restype = ctypes.c_float
    libm.cosf.argtypes = libm.sinf.argtypes = [ctypes.c_float]
    deg = np.pi / 180.0
    xa, ya, za = (f(float(f(v)) * deg) for v in vals)
    c = lambda a: f(libm.cosf(float(a)))  # noqa: E731
    s_ = lambda a: f(libm.sinf(float(a)))  # noqa: E731
    cx, sx, cy, sy, cz, sz = c(xa), s_(xa), c(ya), s_(ya), c(za), s_(za)
    return [cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz,
            cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz,
            f(-1) * sy, sx * cy, cx * cy]


def linearised(pos, topo):
    """Midpoints set to their edge averages, as Blob::init's linearise_elements leaves them
    (src/tetra_element_linear.cpp:567-591)."""
    p = np.array(pos, np.float64)
    for k, (a, b) in enumerate(((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))):
        p[topo[:, 4 + k]] = .5 * (p[topo[:, a]] + p[topo[:, b]])
    return p


def place_blobs(params: dict, blobs, positions, return_box: bool = False):
    """Positions (meso units, [N_t, 3] per blob) after World::init places the blobs:
    <centroid> (Blob::position, src/Blob.cpp:886-918: the centroid times 1/N, moved to the
    given point times the blob's scale), <rotation> (Blob::rotate about the centroid), then
    move_into_box (default 1): every blob shifted so that the system centroid sits at the
    centre of a box of es_N * ssint_cutoff per axis, es_N = 2 ceil(extent / ssint_cutoff)
    unless given (and an interaction or electrostatics flag is on).  The box (box_dim,
    src/World.cpp:455-471) is set up whether or not the blobs move into it; return_box adds
    it and es_N (the lookup grid's cells per axis) to the result."""
    out = []
    for b, p in zip(blobs, positions):
        p = np.array(p, np.float64)
        sc = U.blob_scale(float(b.get("scale", 1.0)))
        key = "centroid" if "centroid" in b else ("centroid_pos" if "centroid_pos" in b else None)
        if key:
            target = [v * sc for v in _vec3(b[key])]
            n = p.shape[0]
            cen = [_seq_sum(p[:, k]) * (1.0 / n) for k in range(3)]
            v = [target[k] - cen[k] for k in range(3)]
            for k in range(3):
                p[:, k] += v[k]
        if "rotation" in b:
            r = [float(x) for x in _rotation_matrix(_vec3(b["rotation"]))]
            com = _centroid(p)
            q = p - com
            x, y, z = q[:, 0], q[:, 1], q[:, 2]
            p = np.stack([x * r[0] + y * r[1] + z * r[2] + com[0], x * r[3] + y * r[4] + z * r[5] + com[1],
                          x * r[6] + y * r[7] + z * r[8] + com[2]], axis=1)
        out.append(p)
    # World::get_system_dimensions (src/World.cpp:3413-3466) with Blob::get_min_max
    # (src/Blob.cpp:2383-2411), whose per-axis update is "max, else min"
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for p in out:
        bmin, bmax = np.full(3, np.inf), np.full(3, -np.inf)
        for row in p:
            for k in range(3):
                if row[k] > bmax[k]:
                    bmax[k] = row[k]
                elif row[k] < bmin[k]:
                    bmin[k] = row[k]
        lo, hi = np.minimum(lo, bmin), np.maximum(hi, bmax)
    dim = hi - lo
    cutoff = float(params.get("ssint_cutoff", params.get("vdw_cutoff", 3e-9))) / U.MESO_LENGTH
    flags = [int(float(params.get(k, d))) for k, d in (("calc_steric", 1), ("calc_es", 0), ("calc_preComp", 0))]
    flags.append(int(float(params.get("calc_ssint", params.get("calc_vdw", 1)))))
    es = [int(float(params.get(f"es_N_{a}", -1))) if any(flags) else 0 for a in "xyz"]
    if min(es) < 1:
        es = [2 * int(np.ceil(d / cutoff)) for d in dim]
    box = np.array([cutoff * n for n in es])
    # World::get_system_centroid (:3381-3411): blob centroids weighted by their node counts
    tot = sum(p.shape[0] for p in out)
    wc = np.zeros(3)
    for p in out:
        wc += _centroid(p) * p.shape[0]
    wc /= tot
    shift = box / 2.0 - wc
    if int(float(params.get("move_into_box", 1))) == 1:
        out = [p + shift for p in out]
    return (out, box, es) if return_box else out


# SimulationParams wall keys (src/SimulationParams.cpp:420-480; PBC by default), in the bit
# order of ffh_blob_set_box's periodic mask
WALL_KEYS = ("wall_x_1", "wall_x_2", "wall_y_1", "wall_y_2", "wall_z_1", "wall_z_2")


def periodic_mask(params: dict) -> int:
    """Which walls are periodic (PBC, the default); an unknown value keeps the default, as the
    reference's parser does."""
    m = 0
    for k, key in enumerate(WALL_KEYS):
        if params.get(key, "PBC") not in ("HARD", "STOP"):
            m |= 1 << k
    return m


# ---------------------------------------------------------------------------------------------
# the world
# ---------------------------------------------------------------------------------------------
class World:
    """The dynamic blobs of a .ffea script on one GPU (World::init / World::run)."""

    def __init__(self, script: str, precision: str = "fp64", quiet: bool = True):
        self.script = os.path.abspath(script)
        s = F.read_script(script)
        self.raw = s["params"]
        self.params = F.params_from_script(self.raw)
        base = os.path.dirname(self.script)
        stem = os.path.splitext(os.path.basename(self.script))[0]

        def path(key, default):
            v = self.raw.get(key)
            return os.path.normpath(os.path.join(base, v)) if v else os.path.join(base, default)

        self.traj_path = path("trajectory_out_fname", stem + ".ftj")
        self.meas_path = path("measurement_out_fname", stem + ".fm")
        self.ckpt_out = path("checkpoint_out", stem + ".fcp")
        self.ckpt_in = path("checkpoint_in", "") if self.raw.get("checkpoint_in") else None
        self.restart = int(float(self.raw.get("restart", 0)))
        self.num_steps = int(float(self.raw.get("num_steps", 0)))
        self.check = int(float(self.raw.get("check", 10000)))
        self.dt_si = float(self.raw.get("dt", 1e-14))
        if self.restart and not self.ckpt_in:
            raise F.FormatError("Checkpoint input file required if restart is required.")
        if self.ckpt_in and os.path.abspath(self.ckpt_in) == os.path.abspath(self.ckpt_out):
            raise F.FormatError("checkpoint_in and checkpoint_out must differ")
        self.meshes, self.scales = [], []
        solvers = set()
        for b in s["blobs"]:
            solver = b.get("solver", "CG_nomass")
            if solver not in _lib.FFH_SOLVERS:
                raise F.FormatError(f"blob solver {solver}: one of {sorted(_lib.FFH_SOLVERS)} on this path")
            solvers.add(solver)
            if b.get("motion_state", "DYNAMIC") != "DYNAMIC":
                raise F.FormatError("only DYNAMIC blobs are on this path")
            self.meshes.append(F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin")))
            self.scales.append(U.blob_scale(float(b.get("scale", 1.0))))
        if len(solvers) > 1:
            raise F.FormatError("all blobs of a script must use the same solver on this path")
        self.solver = solvers.pop() if solvers else "CG_nomass"
        self.mass_in_system = self.solver != "CG_nomass"  # src/World.cpp:2545-2556
        self.nodes = [m.num_nodes for m in self.meshes]
        # the rest state comes from the positions as loaded; World::init then places the blobs,
        # and pos_0 (the RMSD reference) is taken after that (src/World.cpp:502-506)
        self.pos0, self.box, self.es_N = place_blobs(self.raw, s["blobs"],
                                          [linearised(m.pos * sc, m.topo) for m, sc in zip(self.meshes, self.scales)],
                                          return_box=True)
        if self.restart:
            st = read_checkpoint(self.ckpt_in)[0]
        else:
            st = package_seed(self.params.rng_seed)
        self.batch = Batch(self.meshes, self.params, self.scales, rng_state=st, precision=precision)
        self.batch.set_positions(np.concatenate(self.pos0))
        self.batch.set_solver(self.solver)
        # <velocity = (vx, vy, vz)>: every node of the blob (m/s -> meso; Blob::velocity_all,
        # src/World.cpp:2577-2584, 2707-2709); the massless solve overwrites it in its first step
        if self.mass_in_system and any("velocity" in b for b in s["blobs"]):
            off = self.batch.node_off
            vel = np.zeros((int(off[-1]), 3))
            for k, b in enumerate(s["blobs"]):
                if "velocity" in b:
                    vel[off[k]:off[k + 1]] = np.array(_vec3(b["velocity"])) / U.MESO_VELOCITY
            self.batch.set_velocities(vel)
        # walls: periodic ones wrap a blob whose centroid leaves the box at the start of every
        # step (src/World.cpp:1850-1901).  A hard wall zeroes outward velocities of surface nodes
        # (Blob::enforce_box_boundaries, src/Blob.cpp:2272-2330), which the massless solve then
        # replaces; with a mass-ful solver that clamp is not on this path.
        if self.mass_in_system and any(self.raw.get(k) == "HARD" for k in WALL_KEYS):
            raise F.FormatError("hard walls with a mass-ful solver are not on this path")
        self.walls = periodic_mask(self.raw)
        if self.walls:
            self.batch.set_box(self.box, self.walls)
        self.calc_springs = int(float(self.raw.get("calc_springs", 0))) != 0
        self.calc_ctforces = int(float(self.raw.get("calc_ctforces", 0))) == 1
        self._interactions(s)
        self._surface_interactions(s)
        self.quiet = quiet
        self.step_initial = 0
        if self.restart:
            self._restart_files()
        else:
            with open(self.traj_path, "w") as f:
                f.write(traj_header(self.nodes))
            with open(self.meas_path, "w") as f:
                f.write(self._meas_header())

    def _interactions(self, s):
        """Springs (World::load_springs) and constant forces (Blob::load_ctforces) of the
        script onto the batch, node and face ids shifted to the batch's numbering."""
        off = self.batch.node_off
        if self.calc_springs:
            if not s.get("springs"):
                raise F.FormatError("calc_springs set but no <springs> file")
            sp = F.read_springs(s["springs"])
            nb = len(self.meshes)
            if (sp["blob0"] >= nb).any() or (sp["blob1"] >= nb).any() or (sp["blob0"] < 0).any() or (sp["blob1"] < 0).any():
                raise F.FormatError("spring blob index out of range")
            if (sp["conf0"] != 0).any() or (sp["conf1"] != 0).any():
                raise F.FormatError("springs on conformations other than 0 are not on this path")
            for e in ("0", "1"):
                bad = (sp["node" + e] < 0) | (sp["node" + e] >= np.array(self.nodes)[sp["blob" + e]])
                if bad.any():
                    raise F.FormatError("spring node index out of range")
            self.batch.set_springs(off[sp["blob0"]] + sp["node0"], off[sp["blob1"]] + sp["node1"], sp["k"], sp["l"])
        if self.calc_ctforces:
            if not s.get("ctforces"):
                raise F.FormatError("ctforces_fname not provided in <ctforces> block, while calc_ctforces set to 1")
            nodes, forces, faces, sets, set_forces = [], [], [], [], []
            rot = [[], [], [], []]
            face_off = 0
            for b, (blob, m) in enumerate(zip(s["blobs"], self.meshes)):
                ct = F.read_ctforces(s["ctforces"], b, m.num_nodes)
                (ln, lf), (bs, bf) = ct.linear, ct.surface
                nodes.append(ln + off[b])
                forces.append(lf)
                rot[0].append(ct.rotational[0] + off[b])
                for k in (1, 2, 3):
                    rot[k].append(ct.rotational[k])
                if bs:
                    if not blob.get("surface"):
                        raise F.FormatError(f"blob {b}: surface forces need its surface file")
                    fc = F.read_surface(blob["surface"])[:, 1:4]
                    faces.append(fc + off[b])
                    sets += [[f + face_off for f in x] for x in bs]
                    set_forces.append(bf)
                    face_off += fc.shape[0]
            self.batch.set_constant_forces(np.concatenate(nodes), np.concatenate(forces))
            rot = [np.concatenate(x) for x in rot]
            if rot[0].size:
                self.batch.set_rotational_forces(*rot)
            if sets:
                self.batch.set_surface_forces(np.concatenate(faces), sets, np.concatenate(set_forces))

    def _surface_interactions(self, s):
        """Surface-surface interactions as World::init sets them up (src/World.cpp:273-280,
        896-970; SimulationParams checks, src/SimulationParams.cpp:626-660): the faces of the
        blobs' .vdw / .ssint files that are active (type != -1).  The SurfSurfEnergy column is
        written whenever calc_ssint or calc_steric is on (the defaults: ljsteric with steric
        on).  lennard-jones, ljsteric, gensoft and the steric volume exclusion alone (steric,
        or calc_ssint 0 with calc_steric 1) run on the device, with steric_factor and steric_dr
        (src/SimulationParams.cpp:412-418).  Periodic blob corrections (force_pbc) and sticky
        walls are not on this path."""
        p = self.raw
        calc_ssint = int(float(p.get("calc_ssint", p.get("calc_vdw", 1))))
        calc_steric = int(float(p.get("calc_steric", 1)))
        kind = p.get("ssint_type", p.get("vdw_type", "ljsteric"))
        self.ssint_column = calc_ssint == 1 or calc_steric == 1
        if int(float(p.get("sticky_wall_xz", 0))) == 1:
            raise F.FormatError("sticky_wall_xz is not on this path")
        if not self.ssint_column:
            return
        if calc_ssint == 1:
            if kind == "steric":
                if calc_steric == 0:
                    raise F.FormatError("ssint_type steric needs calc_steric = 1")
                calc_ssint = 0
            elif kind not in ("lennard-jones", "ljsteric", "gensoft"):
                raise F.FormatError("ssint_type must be lennard-jones, ljsteric or gensoft")
            elif kind != "lennard-jones" and calc_steric == 0:
                raise F.FormatError(f"ssint_type {kind} needs calc_steric = 1")
        fname = None
        for k in ("vdw_in_fname", "ssint_in_fname", "vdw_forcefield_params", "ssint_forcefield_params", "ljparams",
                  "lj_params"):
            if k in p:
                fname = os.path.normpath(os.path.join(os.path.dirname(self.script), p[k]))
        if calc_ssint == 1 and fname is None:
            raise F.FormatError("surface-surface forcefield params file name required (ssint_in_fname)")
        E = R = K = None
        ntypes = 1
        if calc_ssint == 1:
            cutoff = float(p.get("ssint_cutoff", p.get("vdw_cutoff", 3e-9)))
            E, R, K = F.read_ssint_params(fname, kind, cutoff / U.MESO_LENGTH)
            ntypes = E.shape[0]
        off = self.batch.node_off
        faces, types, opposite = [], [], []
        for b, blob in enumerate(s["blobs"]):
            path = blob.get("vdw", blob.get("ssint"))
            if not path or not blob.get("surface"):
                raise F.FormatError(f"blob {b}: surface-surface interactions need its surface and vdw files")
            surf = F.read_surface(blob["surface"])
            fc = surf[:, 1:4]
            if not (calc_ssint == 1 and kind == "lennard-jones"):  # Face::n[3]: pair tests, steric
                opposite.append(F.surface_opposite(surf, self.meshes[b].topo) + off[b])
            ty = F.read_ssint_types(path, fc.shape[0])
            if calc_ssint == 1:
                if (ty > ntypes - 1).any():
                    raise F.FormatError(f"blob {b}: a face type is higher than the forcefield file allows")
            else:
                ty = np.where(ty > ntypes - 1, 0, ty)  # steric only: every type is 0 (src/Blob.cpp:3541-3552)
            faces.append(fc + off[b])
            types.append(ty)
        types = np.concatenate(types)
        if not (types != -1).any():
            return
        if calc_ssint != 1:
            kind = "steric"  # the volume exclusion alone (Steric_solver, or calc_ssint 0)
        if int(float(p.get("force_pbc", 0))) == 1:
            raise F.FormatError("force_pbc is not on this path")
        cutoff = float(p.get("ssint_cutoff", p.get("vdw_cutoff", 3e-9))) / U.MESO_LENGTH
        inc_self = int(float(p.get("inc_self_ssint", p.get("inc_self_vdw", 1))))
        es_update = int(float(p.get("es_update", 10)))
        if kind != "lennard-jones":
            self.batch.set_steric(float(p.get("steric_factor", p.get("vdw_steric_factor", 1))),
                                  float(p.get("steric_dr", p.get("vdw_steric_dr", 5e-3))))
        self.batch.set_ssint(np.concatenate(faces), types, E, R, self.es_N, cutoff, inc_self, es_update, kind=kind,
                             opposite=np.concatenate(opposite) if opposite else None, k0=K)

    def _meas_header(self) -> str:
        p = self.raw
        s = "FFEA Global Measurement File\n\nSimulation Details:\n"
        s += "\tImplementation: ffea_amd (MI355X)\n"
        s += f"\tScript Filename = {os.path.basename(self.script)}\n\tSimulation Type = Full\n\n"
        s += "Parameters:\n\tSystem parameters:\n"
        for k in ("restart", "dt", "num_steps", "check", "rng_seed", "kT", "max_iterations_cg", "epsilon",
                  "stokes_visc", "num_blobs"):
            if k in p:
                s += f"\t{k} = {p[k]}\n"
        s += "\n\tCalculations enabled:\n"
        s += f"\tcalc_noise = {self.params.calc_noise}\n\tcalc_stokes = {self.params.calc_stokes}\n\n\n"
        cols = meas_columns(self.mass_in_system, self.calc_springs, self.ssint_column)
        s += "Measurements:\n" + meas_header_line(cols)
        return s

    def _restart_files(self):
        frames = read_trajectory(self.traj_path)
        if not frames:
            raise F.FormatError(f"{self.traj_path}: no complete frame to restart from")
        step, blobs, start = frames[-1]
        if len(blobs) != len(self.meshes):
            raise F.FormatError(f"{self.traj_path}: frame has {len(blobs)} blobs, script has {len(self.meshes)}")
        pos = np.concatenate([b[:, :3] / U.MESO_LENGTH for b in blobs])  # read_nodes_from_file
        self.batch.set_positions(pos)
        if self.mass_in_system:
            self.batch.set_velocities(np.concatenate([b[:, 3:6] / U.MESO_VELOCITY for b in blobs]))
        with open(self.traj_path, "r+b") as f:
            f.truncate(start)
        with open(self.meas_path, "rb") as f:
            text = f.read()
        cut = text.rstrip(b"\n").rfind(b"\n") + 1  # drop the last measurement line
        with open(self.meas_path, "r+b") as f:
            f.truncate(cut)
        self.step_initial = step

    def measurements(self, step: int):
        """Time, strain energy, centroid, RMSD (World::make_measurements, src/World.cpp:4282-4340;
        Blob::make_measurements, src/Blob.cpp:1363-1528) at the positions of the current step."""
        pos = self.batch.split(self.batch.positions())
        tot = sum(self.nodes)
        cog = np.zeros(3)
        msd = 0.0
        for p, p0, n in zip(pos, self.pos0, self.nodes):
            c = np.zeros(3)
            for k in range(3):  # sequential node sums, as the reference adds them
                c[k] = float(np.add.accumulate(p[:, k])[-1]) / n
            d = p - p0
            r = np.sqrt(float(np.add.accumulate((d * d).sum(axis=1))[-1]) / n)
            cog += c * n
            msd += r * r * n
        cog /= tot
        ke = (self.batch.kinetic_energy() * U.MESO_ENERGY,) if self.mass_in_system else ()
        se = (self.batch.spring_energy() * U.MESO_ENERGY,) if self.calc_springs else ()
        ss = (self.batch.ssint_energy() * U.MESO_ENERGY,) if self.ssint_column else ()
        return (step * self.dt_si, *ke, self.batch.strain_energy() * U.MESO_ENERGY, *(cog * U.MESO_LENGTH),
                np.sqrt(msd / tot) * U.MESO_LENGTH, *se, *ss), pos

    def run(self):
        """World::run from step_initial to num_steps inclusive."""
        b = self.batch
        t0 = time.time()
        step = self.step_initial
        while step <= self.num_steps:
            if step % self.check == 0:
                with open(self.ckpt_out, "w") as f:
                    f.write(checkpoint_text([b.rng_state()]))
                b.update_internal_forces()  # an inversion raises: the run stops, as the reference's does
                vals, pos = self.measurements(step)
                vel = None
                if self.mass_in_system:
                    vel = [v * U.MESO_VELOCITY for v in b.split(b.velocities())]
                with open(self.traj_path, "a") as f:
                    f.write(traj_frame(step, [p * U.MESO_LENGTH for p in pos], vel))
                with open(self.meas_path, "a") as f:
                    f.write(meas_line(vals, self.ssint_column))
                self._positions_phase()
                step += 1
                if not self.quiet:
                    print(f"step = {step - 1} ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
            else:
                n = min(self.check - step % self.check, self.num_steps + 1 - step)
                before = b.steps_done()
                try:
                    b.step(n)
                except _lib.CGError as e:
                    # the failing step completed without moving that blob; the reference carries
                    # on after a failed solve (src/World.cpp:2042), so does this loop
                    print(f"warning: {e}; the blob is not moved that step", file=sys.stderr)
                step += b.steps_done() - before
        b.synchronize()
        return step - 1

    def _positions_phase(self):
        try:
            self.batch.update_positions()
        except _lib.CGError as e:
            print(f"warning: {e}; the blob is not moved this step", file=sys.stderr)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Run a .ffea script's dynamic blobs on one GPU")
    ap.add_argument("script")
    ap.add_argument("--precision", default="fp64", choices=("fp64", "mixed"))
    a = ap.parse_args(argv)
    w = World(a.script, precision=a.precision, quiet=False)
    w.run()
    return 0


if __name__ == "__main__":
    sys.exit(main())
