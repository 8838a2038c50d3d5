"""FFEA text formats on the massless path: .node .top .surf .mat .stokes .pin and the
.ffea script, read with the reference loaders' rules and written as ffeatools writes them.

Readers follow Blob::load_nodes (src/Blob.cpp:2415-2513), load_topology (:2517-2640),
load_surface (:2644-2739), load_material_params (:2821-2876), load_stokes_params
(:2880-2947), load_pinned_nodes (:3691-3755); the script parser follows the
<param>/<system><blob><conformation> layout of FFEA_input_reader
(include/FFEA_input_reader.h:38-76) for the keys this path reads.
"""
from __future__ import annotations

import os
import re
import time
from typing import NamedTuple

import numpy as np

from .mesh import Mesh


class FormatError(ValueError):
    pass


def _lines(path):
    with open(path) as f:
        return f.read().split("\n")


def _header(lines, path, kind):
    if lines[0] not in (f"ffea {kind}", f"walrus {kind}"):
        raise FormatError(f"{path}: this is not a 'ffea {kind}' (read '{lines[0]}')")


def _count(line, key, path):
    m = re.match(rf"{key}\s+(-?\d+)\s*$", line)
    if not m:
        raise FormatError(f"{path}: expected '{key} <n>', read '{line}'")
    return int(m.group(1))


def read_node(path):
    """Returns (pos [N,3] float64, num_surface_nodes)."""
    L = _lines(path)
    _header(L, path, "node file")
    n = _count(L[1], "num_nodes", path)
    ns = _count(L[2], "num_surface_nodes", path)
    _count(L[3], "num_interior_nodes", path)
    if L[4] != "surface nodes:":
        raise FormatError(f"{path}: could not find 'surface nodes:' line")
    pos = np.zeros((n, 3))
    k = 5
    for i in range(n):
        if i == ns:
            if L[k] != "interior nodes:":
                raise FormatError(f"{path}: could not find 'interior nodes:' line")
            k += 1
        v = L[k].split()
        if len(v) < 3:
            raise FormatError(f"{path}: error reading node {i}")
        pos[i] = [float(v[0]), float(v[1]), float(v[2])]
        k += 1
    return pos, ns


def read_topology(path, num_nodes=None):
    """Returns (topo [E,10] int32, num_surface_elements)."""
    L = _lines(path)
    _header(L, path, "topology file")
    e = _count(L[1], "num_elements", path)
    es = _count(L[2], "num_surface_elements", path)
    _count(L[3], "num_interior_elements", path)
    if L[4] != "surface elements:":
        raise FormatError(f"{path}: could not find 'surface elements:' line")
    topo = np.zeros((e, 10), np.int32)
    k = 5
    for i in range(e):
        if i == es:
            if L[k] != "interior elements:":
                raise FormatError(f"{path}: could not find 'interior elements:' line")
            k += 1
        v = L[k].split()
        if len(v) < 10:
            raise FormatError(f"{path}: error reading element {i}")
        topo[i] = [int(x) for x in v[:10]]
        k += 1
    if num_nodes is not None and (topo.min() < 0 or topo.max() >= num_nodes):
        raise FormatError(f"{path}: element references an out of bounds node index")
    return topo, es


def read_material(path, num_elements=None):
    L = _lines(path)
    _header(L, path, "material params file")
    e = _count(L[1], "num_elements", path)
    if num_elements is not None and e != num_elements:
        raise FormatError(f"{path}: {e} elements, topology has {num_elements}")
    mat = np.array([[float(x) for x in L[2 + i].split()[:6]] for i in range(e)], np.float64)
    if mat.shape != (e, 6):
        raise FormatError(f"{path}: there should be 6 values per element")
    return mat


def read_stokes(path, num_nodes=None):
    L = _lines(path)
    _header(L, path, "stokes radii file")
    n = _count(L[1], "num_nodes", path)
    if num_nodes is not None and n != num_nodes:
        raise FormatError(f"{path}: {n} nodes, node file has {num_nodes}")
    return np.array([float(L[2 + i].split()[0]) for i in range(n)], np.float64)


def read_pin(path, num_nodes=None):
    L = _lines(path)
    _header(L, path, "pinned nodes file")
    n = _count(L[1], "num_pinned_nodes", path)
    if L[2] != "pinned nodes:":
        raise FormatError(f"{path}: could not find 'pinned nodes:' line")
    pins = np.array([int(L[3 + i].split()[0]) for i in range(n)], np.int32)
    if num_nodes is not None and n and (pins.min() < 0 or pins.max() >= num_nodes):
        raise FormatError(f"{path}: pinned node references an out of bounds node index")
    return pins


def read_surface(path):
    L = _lines(path)
    _header(L, path, "surface file")
    n = _count(L[1], "num_surface_faces", path)
    if L[2] != "faces:":
        raise FormatError(f"{path}: could not find 'faces:' line")
    return np.array([[int(x) for x in L[3 + i].split()[:4]] for i in range(n)], np.int32)


def read_ssint_types(path, num_faces):
    """Blob::load_ssint (src/Blob.cpp:3494-3578): one interaction type per surface face, -1
    for an inactive face ("ffea vdw file" / "ffea ssint file")."""
    L = _lines(path)
    if L[0] not in ("walrus vdw file", "ffea vdw file", "ffea ssint file"):
        raise FormatError(f"{path}: this is not a 'ffea ssint file' (read '{L[0]}')")
    n = _count(L[1], "num_faces", path)
    if n != num_faces:
        raise FormatError(f"{path}: {n} faces, the surface file has {num_faces}")
    if L[2] not in ("vdw params:", "ssint params:"):
        raise FormatError(f"{path}: could not find 'ssint params:' line")
    if len(L) < 3 + n:
        raise FormatError(f"{path}: expected {n} face types")
    return np.array([int(L[3 + i].split()[0]) for i in range(n)], np.int32)


# tetra_element_linear::get_opposite_node (src/tetra_element_linear.cpp:645-663): the element
# vertex opposite a second-order surface sub-triangle, keyed by the sum of the squares of the
# triangle's local node numbers (e.g. [0, 4, 5] -> 41 -> vertex 3).
_OPPOSITE = {114: 0, 134: 0, 154: 0, 194: 0, 110: 1, 142: 1, 126: 1, 61: 1,
             116: 2, 81: 2, 109: 2, 52: 2, 66: 3, 41: 3, 78: 3, 90: 3}


def surface_opposite(surf, topo):
    """Face::n[3] of every .surf face (element, three nodes): the element's vertex opposite the
    face (Blob::load_surface, src/Blob.cpp:2700-2719); FormatError where the face is not one of
    the split second-order sub-triangles."""
    surf = np.asarray(surf)
    out = np.empty(surf.shape[0], np.int32)
    for i, (e, *nodes) in enumerate(surf[:, :4]):
        local = [int(np.where(topo[e] == n)[0][0]) if (topo[e] == n).any() else -1 for n in nodes]
        op = _OPPOSITE.get(sum(k * k for k in local)) if min(local) >= 0 else None
        if op is None:
            raise FormatError(f"surface face {i}: could not find the opposite node")
        out[i] = topo[e][op]
    return out


def read_ssint_params(path, ssint_type="lennard-jones", ssint_cutoff=None):
    """SSINT_matrix::init_ssint (src/LJ_matrix.cpp:122-199): the ntypes x ntypes table of
    (Emin, Rmin[, k0]) converted to meso units -- Emin * area^2 / Energy, Rmin / length,
    k0 * area / Energy; with "gensoft" Rmin is the cutoff.  Returns (Emin, Rmin, k0) as
    [ntypes, ntypes] arrays, entry [i, j] for face types i and j."""
    from .units import MESO_AREA, MESO_ENERGY, MESO_LENGTH
    with open(path) as f:
        lines = f.read().split("\n")
    if lines[0] not in ("ffea vdw forcefield params file", "ffea ssint forcefield params file"):
        raise FormatError(f"{path}: this is not a 'ffea ssint forcefield params file' (read '{lines[0]}')")
    nt = int(lines[1].split(" ")[1])
    E, R, K = (np.zeros((nt, nt)) for _ in range(3))
    for i in range(nt):
        sets = lines[2 + i].split(")")
        for j in range(nt):
            vals = sets[j].strip().replace("(", "", 1).split(",")
            if len(vals) > 3 or len(vals) < 2:
                raise FormatError(f"{path}: parameter set ({i}, {j}) needs 2 or 3 values")
            v = [float(x) for x in vals] + [0.0] * (3 - len(vals))
            if v[0] <= 0 or v[1] <= 0:
                raise FormatError(f"{path}: Emin and Rmin must be greater than 0")
            if ssint_type == "gensoft":
                v[1] = ssint_cutoff * MESO_LENGTH
            E[i, j] = v[0] * MESO_AREA * MESO_AREA / MESO_ENERGY
            R[i, j] = v[1] / MESO_LENGTH
            K[i, j] = v[2] * MESO_AREA / MESO_ENERGY
    return E, R, K


def _script_lines(path):
    """FFEA_input_reader::file_to_lines (src/FFEA_input_reader.cpp:36-113): <!-- --> comments
    removed (also across lines), each line trimmed, empty lines dropped."""
    text = re.sub(r"<!--.*?(-->|$)", "", open(path).read(), flags=re.S)
    return [l.strip() for l in text.split("\n") if l.strip()]


def read_springs(path):
    """World::load_springs (src/World.cpp:3473-3560): per spring k and l converted to meso
    units (k *= area / Energy, l /= length), blob, conformation and node of each end.
    Returns a dict of arrays k, l, blob0, blob1, conf0, conf1, node0, node1."""
    from .units import MESO_AREA, MESO_ENERGY, MESO_LENGTH
    L = _lines(path)
    if L[0] != "ffea springs file":
        raise FormatError(f"{path}: this is not a 'ffea spring file' (read '{L[0]}')")
    n = _count(L[1], "num_springs", path)
    i = 2 + (L[2].strip() == "springs:")
    rows = []
    for j in range(n):
        v = L[i + j].split()
        if len(v) < 8:
            raise FormatError(f"{path}: problem reading spring data (spring {j})")
        rows.append(v[:8])
    out = {k: np.array([int(r[c]) for r in rows], np.int32)
           for c, k in ((2, "blob0"), (3, "blob1"), (4, "conf0"), (5, "conf1"), (6, "node0"), (7, "node1"))}
    k = np.array([float(r[0]) for r in rows])
    out["k"] = k * (MESO_AREA / MESO_ENERGY)
    out["l"] = np.array([float(r[1]) for r in rows]) / MESO_LENGTH
    if (out["k"] < 0).any() or (out["l"] < 0).any():
        raise FormatError(f"{path}: spring constants and lengths must be >= 0")
    same = (out["blob0"] == out["blob1"]) & (out["conf0"] == out["conf1"]) & (out["node0"] == out["node1"])
    if same.any():
        raise FormatError(f"{path}: spring {int(np.flatnonzero(same)[0])} is connected to the same node")
    return out


def _normalised(d):
    """arr3Normalise (src/mat_vec_fns_II.cpp:149-160)."""
    norm = 0.0
    for x in d:
        norm += x * x
    norm = float(np.sqrt(norm))
    return [x / norm for x in d]


class CtForces(NamedTuple):
    """One blob's constant forces from a .ctforces file (meso units)."""
    linear: tuple        # (nodes [n], forces [n, 3])
    rotational: tuple    # (nodes [n], F [n], axis [n, 6]: point, unit direction, torque [n] 0/1)
    surface: tuple       # (list of face-index lists, forces [S, 3])


def read_ctforces(path, blob_index, num_nodes, conformation_index=0) -> CtForces:
    """Blob::load_ctforces (src/Blob.cpp:3084-3430) for one blob: the linear forces (CTF 1,
    "all" expanded to every node), the rotational forces (CTF 2: F / force unit, the axis point
    scaled by the inverse length unit, the direction normalised, type 'pf' or 'pt'; "all"
    expanded) and the linear surface forces (CTF 3).  A rotational axis given by two nodes
    ('n' types) raises FormatError: the reference leaves its axis undefined (src/Blob.cpp:2114-2124)."""
    from .units import MESO_FORCE, MESO_LENGTH
    L = _script_lines(path)
    if len(L) < 4:
        raise FormatError(f"{path}: too short")
    counts = {"num_ctforces": 0, "num_linear_forces": 0, "num_rot_forces": 0, "num_linear_surface_forces": 0}
    r = 1
    while r < 5 and r < len(L) and ":" not in L[r]:
        w = L[r].split()
        if w[0] in counts:
            counts[w[0]] = int(w[1])
        r += 1
    nl, nr, ns = counts["num_linear_forces"], counts["num_rot_forces"], counts["num_linear_surface_forces"]
    want = counts["num_ctforces"] + r + (nl > 0) + (nr > 0) + (ns > 0)
    if counts["num_ctforces"] != nl + nr + ns or counts["num_ctforces"] == 0 or want != len(L):
        raise FormatError(f"{path}: inconsistent ctforces header")
    nodes, forces = [], []
    if L[r].startswith("linear forces:"):
        r += 1
    elif nl > 0:
        raise FormatError(f"{path}: expected 'linear forces:'")
    for line in L[r:r + nl]:
        w = line.split()
        if len(w) != 7:
            raise FormatError(f"{path}: invalid line '{line}'")
        if int(w[4]) != blob_index or int(w[5]) != conformation_index:
            continue
        d = _normalised([float(w[1]), float(w[2]), float(w[3])])
        F = float(w[0])
        f = [F * d[0] / MESO_FORCE, F * d[1] / MESO_FORCE, F * d[2] / MESO_FORCE]
        for nd in (range(num_nodes) if w[6] == "all" else [int(w[6])]):
            nodes.append(nd)
            forces.append(f)
    r += nl
    if r < len(L) and L[r].startswith("rotational forces:"):
        r += 1
    elif nr > 0:
        raise FormatError(f"{path}: expected 'rotational forces:'")
    mdfm1, mdlm1 = 1.0 / MESO_FORCE, 1.0 / MESO_LENGTH
    rnodes, rF, raxis, rtorque = [], [], [], []
    for line in L[r:r + nr]:
        w = line.split()
        if len(w) != 11:
            raise FormatError(f"{path}: invalid line '{line}'")
        if int(w[8]) != blob_index or int(w[9]) != conformation_index:
            continue
        typ = w[1]
        if typ[:1] == "n":
            raise FormatError(f"{path}: rotational forces about an axis between two nodes are undefined in the reference")
        if typ[:1] != "p":
            raise FormatError(f"{path}: invalid rotational force: {typ}")
        if typ[1:2] not in ("f", "t"):
            raise FormatError(f"{path}: invalid rotational force type: {typ[1:]}")
        F = float(w[0]) * mdfm1
        pt = [mdlm1 * float(x) for x in w[2:5]]
        d = _normalised([float(x) for x in w[5:8]])
        for nd in (range(num_nodes) if w[10] == "all" else [int(w[10])]):
            rnodes.append(nd)
            rF.append(F)
            raxis.append(pt + d)
            rtorque.append(1 if typ[1] == "t" else 0)
    r += nr
    sets, set_forces = [], []
    if r < len(L) and L[r].startswith("linear surface forces:"):
        r += 1
    elif ns > 0:
        raise FormatError(f"{path}: expected 'linear surface forces:'")
    for line in L[r:r + ns]:
        w = line.split()
        if len(w) < 7:
            raise FormatError(f"{path}: invalid line '{line}'")
        if int(w[4]) != blob_index or int(w[5]) != conformation_index:
            continue
        if w[6] == "all":
            raise FormatError(f"{path}: 'all' surface faces are refused by the reference too")
        d = _normalised([float(w[1]), float(w[2]), float(w[3])])
        F = float(w[0])
        set_forces.append([F * d[0] / MESO_FORCE, F * d[1] / MESO_FORCE, F * d[2] / MESO_FORCE])
        sets.append([int(x) for x in w[6:]])
    return CtForces((np.array(nodes, np.int32), np.array(forces, np.float64).reshape(-1, 3)),
                    (np.array(rnodes, np.int32), np.array(rF, np.float64), np.array(raxis, np.float64).reshape(-1, 6),
                     np.array(rtorque, np.int32)),
                    (sets, np.array(set_forces, np.float64).reshape(-1, 3)))


def load_blob(node, top, mat, stokes, pin=None) -> Mesh:
    pos, _ = read_node(node)
    topo, _ = read_topology(top, pos.shape[0])
    material = read_material(mat, topo.shape[0])
    rad = read_stokes(stokes, pos.shape[0])
    pins = read_pin(pin, pos.shape[0]) if pin else np.zeros(0, np.int32)
    lin = np.unique(topo[:, :4])
    return Mesh(pos=pos, topo=topo, material=material, stokes_radius=rad, num_linear=int(lin.size),
                pinned=pins)


# ---------------------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------------------
def write_node(path, pos, num_surface=None):
    pos = np.asarray(pos)
    n = pos.shape[0]
    ns = n if num_surface is None else int(num_surface)
    with open(path, "w") as f:
        f.write(f"ffea node file\nnum_nodes {n}\nnum_surface_nodes {ns}\nnum_interior_nodes {n - ns}\n")
        f.write("surface nodes:\n")
        for i in range(n):
            if i == ns:
                f.write("interior nodes:\n")
            f.write("%.17g %.17g %.17g\n" % tuple(pos[i]))
        if ns == n:
            f.write("interior nodes:\n")


def write_topology(path, topo):
    topo = np.asarray(topo)
    e = topo.shape[0]
    with open(path, "w") as f:
        f.write(f"ffea topology file\nnum_elements {e}\nnum_surface_elements {e}\n"
                f"num_interior_elements 0\nsurface elements:\n")
        for row in topo:
            f.write(" ".join(str(int(v)) for v in row) + "\n")
        f.write("interior elements:\n")


def write_material(path, material):
    material = np.asarray(material)
    with open(path, "w") as f:
        f.write(f"ffea material params file\nnum_elements {material.shape[0]}\n")
        for row in material:
            f.write(" ".join("%.17g" % v for v in row) + "\n")


def write_stokes(path, radius):
    radius = np.asarray(radius)
    with open(path, "w") as f:
        f.write(f"ffea stokes radii file\nnum_nodes {radius.shape[0]}\n")
        for r in radius:
            f.write("%.17g\n" % r)


def write_pin(path, pins):
    pins = np.asarray(pins, np.int64)
    with open(path, "w") as f:
        f.write(f"ffea pinned nodes file\nnum_pinned_nodes {pins.size}\npinned nodes:\n")
        for p in pins:
            f.write(f"{int(p)}\n")


def write_surface(path, faces):
    """faces [F, 4]: element, then the face's three nodes."""
    faces = np.asarray(faces)
    with open(path, "w") as f:
        f.write(f"ffea surface file\nnum_surface_faces {faces.shape[0]}\nfaces:\n")
        for row in faces:
            f.write(" ".join(str(int(v)) for v in row) + "\n")


def write_ssint_types(path, types):
    """A .vdw file (the format Blob::load_ssint reads): one type per surface face, -1 off."""
    types = np.asarray(types)
    with open(path, "w") as f:
        f.write(f"ffea vdw file\nnum_faces {types.size}\nvdw params:\n")
        f.write("".join(f"{int(t)}\n" for t in types))


def write_blob(prefix, mesh: Mesh):
    write_node(prefix + ".node", mesh.pos)
    write_topology(prefix + ".top", mesh.topo)
    write_material(prefix + ".mat", mesh.material)
    write_stokes(prefix + ".stokes", mesh.stokes_radius)
    write_pin(prefix + ".pin", mesh.pinned)


# ---------------------------------------------------------------------------------------
# .ffea script (the keys of the CG_nomass path)
# ---------------------------------------------------------------------------------------
def read_script(path) -> dict:
    """Parses <param> key/values and the first conformation of each <blob>.

    Returns {"params": {key: str}, "blobs": [{"nodes":..., "topology":..., "material":...,
    "stokes":..., "pin":..., "solver":..., "scale":...}], "springs": path or None,
    "ctforces": path or None}; file paths resolved relative to the script.  The <springs> and
    <ctforces> blocks of <interactions> name one file each (World::read_and_build_system,
    src/World.cpp:2262-2297).
    """
    text = open(path).read()
    base = os.path.dirname(os.path.abspath(path))
    params = {}
    m = re.search(r"<param>(.*?)</param>", text, re.S)
    if not m:
        raise FormatError(f"{path}: no <param> block")
    for k, v in re.findall(r"<\s*([A-Za-z_0-9]+)\s*=\s*([^>]*)>", m.group(1)):
        params[k.strip()] = v.strip()
    blobs = []
    for bm in re.finditer(r"<blob>(.*?)</blob>", text, re.S):
        body = bm.group(1)
        b = {}
        for k, v in re.findall(r"<\s*([A-Za-z_0-9]+)\s*=\s*([^>]*)>", body):
            k, v = k.strip(), v.strip()
            if k in b:
                continue  # first conformation only
            b[k] = v
        for k in ("nodes", "topology", "surface", "material", "stokes", "pin", "vdw", "ssint"):
            if k in b:
                b[k] = os.path.normpath(os.path.join(base, b[k]))
        blobs.append(b)
    files = {}
    for block in ("springs", "ctforces"):
        bm = re.search(rf"<{block}>(.*?)</{block}>", text, re.S)
        tags = re.findall(r"<\s*([A-Za-z_0-9]+)\s*=\s*([^>]*)>", bm.group(1)) if bm else []
        files[block] = os.path.normpath(os.path.join(base, tags[0][1].strip())) if tags else None
    return {"params": params, "blobs": blobs, **files}


def params_from_script(params: dict):
    """SimParams from a <param> block with SimulationParams::assign's conversions and
    defaults (src/SimulationParams.cpp:26-60, :211-540)."""
    from .units import SimParams, KB

    def g(k, d):
        return float(params[k]) if k in params else d

    eps = g("epsilon", 0.1)  # default epsilon2 = 0.01 (src/SimulationParams.cpp:36)
    stokes = int(float(params.get("calc_stokes", params.get("do_stokes", 1))))
    # rng_seed "time": the clock's seconds (src/SimulationParams.cpp:328-333)
    seed = int(time.time()) if params.get("rng_seed") == "time" else int(g("rng_seed", 0))
    return SimParams.from_si(dt=g("dt", 1e-14), kT=g("kT", 4.11e-21), epsilon=eps,
                             max_iterations_cg=int(g("max_iterations_cg", 1000)),
                             calc_noise=int(g("calc_noise", 1)), calc_stokes=stokes,
                             stokes_visc=g("stokes_visc", 1e-3), rng_seed=seed)
