"""Output and checkpoint formats of ffea_amd.world against files the reference's own runs wrote
(tests/golden/formats: tet_ascii.1.ftj / .fm from tests/rods/unit/connection_orientation,
cyl_160_fine.fcp from tests/physics/cyl_160_fine): the writers reproduce them byte for byte
from the values the readers parse, and the restart reader keeps only complete frames."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

FMT = os.path.join(GOLDEN, "formats")


def test_trajectory_header_and_frames_byte_identical():
    from ffea_amd import world as W
    text = open(os.path.join(FMT, "tet_ascii.1.ftj")).read()
    frames = W.read_trajectory(os.path.join(FMT, "tet_ascii.1.ftj"))
    assert [f[0] for f in frames] == [1, 2]
    assert [b.shape for b in frames[0][1]] == [(10, 10), (10, 10)]
    rebuilt = W.traj_header([10, 10]) + "".join(W.traj_frame(s, [b[:, :3] for b in blobs]) for s, blobs, _ in frames)
    assert rebuilt == text
    # frame start offsets point at the frame headers
    for s, _, off in frames:
        assert text[off:].startswith("Blob 0, Conformation 0, step %d\n" % s)


def test_partial_frame_is_not_complete(tmp_path):
    from ffea_amd import world as W
    text = open(os.path.join(FMT, "tet_ascii.1.ftj")).read()
    cut = text[: text.rindex("Conformation Changes:") - 40]  # second frame cut inside its rows
    p = tmp_path / "t.ftj"
    p.write_text(cut)
    frames = W.read_trajectory(str(p))
    assert [f[0] for f in frames] == [1]


def test_measurement_lines_match_reference_columns():
    from ffea_amd import world as W
    lines = open(os.path.join(FMT, "tet_ascii.1.fm")).read().split("\n")
    k = lines.index("Measurements:")
    cols = W.meas_columns(False, False, ssint=True)
    assert lines[k + 1] == W.meas_header_line(cols).rstrip("\n")
    for line in lines[k + 2:]:
        if not line.strip():
            continue
        vals = [float(v) for v in line.split()]
        assert W.meas_line(vals[:6]).rstrip("\n") == line[:84]
        assert W.meas_line(vals, ssint=True).rstrip("\n") == line


def test_checkpoint_round_trip():
    from ffea_amd import world as W
    path = os.path.join(FMT, "cyl_160_fine.fcp")
    st = W.read_checkpoint(path)
    assert st.shape == (8, 6) and st.dtype == np.uint64
    assert W.checkpoint_text(st) == open(path).read()


def test_mass_frame_layout_and_round_trip(tmp_path):
    """A mass-ful frame (Blob::write_nodes_to_file, src/Blob.cpp:1154-1161): positions,
    velocities, potential and force per node; read back by the restart reader."""
    from ffea_amd import world as W
    rng = np.random.default_rng(1)
    p = [rng.normal(size=(4, 3)) * 1e-9, rng.normal(size=(2, 3)) * 1e-9]
    v = [rng.normal(size=(4, 3)), rng.normal(size=(2, 3))]
    text = W.traj_header([4, 2]) + W.traj_frame(3, p, v)
    row = text.split("DYNAMIC\n")[1].split("\n")[0].split()
    assert len(row) == 10 and row[6:] == ["0.000000e+00"] * 4
    assert float(row[3]) == float("%e" % v[0][0, 0])
    path = tmp_path / "m.ftj"
    path.write_text(text)
    (step, blobs, _), = W.read_trajectory(str(path))
    assert step == 3
    for b, pp, vv in zip(blobs, p, v):
        assert np.allclose(b[:, :3], pp, rtol=1e-6) and np.allclose(b[:, 3:6], vv, rtol=1e-6)
    assert W.MEAS_COLUMNS_MASS[:3] == ("Time", "KineticEnergy", "StrainEnergy")


def test_springs_file():
    """World::load_springs (src/World.cpp:3473-3560) on the reference's cube_xx.springs
    (tests/physics/cube_springs_structure): k in meso units via area / Energy, l via length."""
    from ffea_amd import formats as F, units as U
    sp = F.read_springs(os.path.join(FMT, "cube_xx.springs"))
    assert sp["k"].size == 8
    assert sp["k"][0] == 1.0e-1 * (U.MESO_AREA / U.MESO_ENERGY)
    assert sp["l"][0] == 4.75e-9 / U.MESO_LENGTH
    assert list(sp["node0"][:2]) == [58, 58] and list(sp["node1"][:2]) == [36, 79]
    assert (sp["blob0"] == 0).all() and (sp["blob1"] == 1).all()


def test_ctforces_file():
    """Blob::load_ctforces (src/Blob.cpp:3084-3430) on the Young's-modulus test's file
    (tests/physics/E_cyl_160_fine): one linear surface set of 220 faces, 1e-11 N along x."""
    from ffea_amd import formats as F, units as U
    path = os.path.join(GOLDEN, "physics", "cyl_160_fine_E", "cyl_160_fine_surf.ctforces")
    ct = F.read_ctforces(path, 0, 49819)
    (nodes, forces), (sets, sf) = ct.linear, ct.surface
    assert ct.rotational[0].size == 0
    assert nodes.size == 0 and len(sets) == 1 and sets[0][:3] == [8580, 8581, 8582] and len(sets[0]) == 220
    assert sf[0, 0] == 1e-11 * 1.0 / U.MESO_FORCE and sf[0, 1] == 0 and sf[0, 2] == 0
    # a blob the file does not name gets nothing
    ct = F.read_ctforces(path, 1, 10)
    (nodes, _), (sets, _) = ct.linear, ct.surface
    assert nodes.size == 0 and sets == []


def test_ctforces_rotational(tmp_path):
    """The example .ctforces file of the reference's user manual (docs/2-UserManual/
    7-FFEAFiles.md:257-270), with all three kinds: per blob the lines that name it, "all"
    expanded; the rotational axis point scaled by the inverse length unit, F by the inverse
    force unit (src/Blob.cpp:3296-3312).  The manual prints num_ctforces 5 for its six forces,
    which the reader's count check refuses (src/Blob.cpp:3152-3170); the fixture says 6."""
    from ffea_amd import formats as F, units as U
    path = os.path.join(GOLDEN, "formats", "manual_example.ctforces")
    ct = F.read_ctforces(path, 0, 12)
    nodes, Fr, axis, torque = ct.rotational
    assert list(nodes) == list(range(12)) and (torque == 1).all()
    assert (Fr == 1e-10 * (1.0 / U.MESO_FORCE)).all()
    want = [(1.0 / U.MESO_LENGTH) * x for x in (-1.9e-9, -1.4313e-9, 8.911e-10)] + [0.0, 1.0, 0.0]
    assert (axis == np.array(want)).all()
    assert ct.linear[0].size == 0 and ct.surface[0] == [[280, 281, 282, 283, 284, 285]]
    ct = F.read_ctforces(path, 1, 7)
    assert list(ct.linear[0]) == list(range(7)) and ct.rotational[0].size == 0 and ct.surface[0] == []
    assert (ct.linear[1] == np.array([0.0, 0.0, 1e-9 / U.MESO_FORCE])).all()
    ct = F.read_ctforces(path, 2, 200)
    assert list(ct.linear[0]) == [104, 134, 152]
    bad = tmp_path / "as_printed.ctforces"
    bad.write_text(open(path).read().replace("num_ctforces 6", "num_ctforces 5"))
    with pytest.raises(F.FormatError):
        F.read_ctforces(str(bad), 0, 12)


def test_ctforces_rotational_types(tmp_path):
    """'pf' is a circular force (torque 0), an unnormalised direction is normalised, and an
    axis through two nodes ('n' types), which the reference leaves undefined, is refused."""
    from ffea_amd import formats as F
    head = "ffea ctforces file\nnum_ctforces 1\nnum_rot_forces 1\nrotational forces:\n"
    p = tmp_path / "a.ctforces"
    p.write_text(head + "2e-11 pf 0 0 0 0 0 2 0 0 5\n")
    nodes, Fr, axis, torque = F.read_ctforces(str(p), 0, 10).rotational
    assert list(nodes) == [5] and list(torque) == [0] and list(axis[0, 3:]) == [0.0, 0.0, 1.0]
    for bad in ("2e-11 nf 0 0 0 0 0 2 0 0 5", "2e-11 px 0 0 0 0 0 1 0 0 5", "2e-11 qt 0 0 0 0 0 1 0 0 5",
                "2e-11 pf 0 0 0 0 0 1 0 0"):
        p.write_text(head + bad + "\n")
        with pytest.raises(F.FormatError):
            F.read_ctforces(str(p), 0, 10)


def test_script_interaction_files():
    from ffea_amd import formats as F
    s = F.read_script(os.path.join(GOLDEN, "physics", "cyl_160_fine_E", "cyl_160_fine-E.ffea"))
    assert s["ctforces"].endswith("cyl_160_fine_E/cyl_160_fine_surf.ctforces") and s["springs"] is None
    assert s["params"]["calc_ctforces"] == "1"


def _sphere():
    from ffea_amd import formats as F, units as U
    from conftest import SPHERE_FFEA
    s = F.read_script(SPHERE_FFEA)
    b = {k: v for k, v in s["blobs"][0].items() if k not in ("centroid", "centroid_pos")}
    m = F.load_blob(b["nodes"], b["topology"], b["material"], b["stokes"], b.get("pin"))
    return s, b, m, U.blob_scale(float(b["scale"]))


def test_move_into_box_centres_the_system():
    """World::init (src/World.cpp:455-489): the box is es_N * ssint_cutoff per axis with
    es_N = 2 ceil(extent / cutoff) (3 nm default cutoff), and the system centroid ends at its
    centre; move_into_box = 0 leaves the positions alone."""
    from ffea_amd import units as U, world as W
    s, b, m, sc = _sphere()
    p = W.linearised(m.pos * sc, m.topo)
    (q,) = W.place_blobs({}, [b], [p])
    cut = 3e-9 / U.MESO_LENGTH
    box = np.array([2 * np.ceil((p[:, k].max() - p[:, k].min()) / cut) * cut for k in range(3)])
    assert np.allclose(q.mean(axis=0), box / 2, rtol=0, atol=1e-12 * box.max())
    assert np.allclose(q - q[0], p - p[0], rtol=0, atol=1e-12 * box.max())  # a pure translation
    (r,) = W.place_blobs({"move_into_box": "0"}, [b], [p])
    assert np.array_equal(r, p)
    # es_N given (and an interaction flag on, as calc_steric is by default): that box
    (q2,) = W.place_blobs({"es_N_x": "10", "es_N_y": "10", "es_N_z": "10"}, [b], [p])
    assert np.allclose(q2.mean(axis=0), np.full(3, 5 * cut), rtol=0, atol=1e-12 * cut)


def test_blob_centroid_and_rotation():
    """<centroid> puts the blob's centroid at the point times its scale; <rotation> (degrees
    about x, y, z, applied about the centroid) of 180 about z maps (x, y) to (-x, -y)."""
    from ffea_amd import world as W
    s, b, m, sc = _sphere()
    p = W.linearised(m.pos * sc, m.topo)
    bb = dict(b, centroid="(1.0,2.0,-3.0)")
    (q,) = W.place_blobs({"move_into_box": "0"}, [bb], [p])
    assert np.allclose(q.mean(axis=0), np.array([1.0, 2.0, -3.0]) * sc, rtol=0, atol=1e-9 * sc)
    br = dict(b, rotation="(0.0,0.0,180)")
    (r,) = W.place_blobs({"move_into_box": "0"}, [br], [p])
    c = p.mean(axis=0)
    want = np.stack([2 * c[0] - p[:, 0], 2 * c[1] - p[:, 1], p[:, 2]], axis=1)
    assert np.allclose(r, want, rtol=0, atol=1e-6 * np.abs(p).max())
    d0 = np.linalg.norm(p[1:] - p[0], axis=1)
    d1 = np.linalg.norm(r[1:] - r[0], axis=1)
    assert np.allclose(d1, d0, rtol=1e-6)


def test_ssint_files():
    """Blob::load_ssint and SSINT_matrix::init_ssint on the cubes_lj fixture (the reference's
    tests/physics/fine_cube_structure): 1256 faces, the 200 of the top face type 0; 7 types,
    entry (0, 0) = (2.537e17 J/m^4, 1 nm) in meso units, the rest (1e-15, 1e-10); no k0 given.
    Face::n[3] of every sub-triangle is a vertex of its element (get_opposite_node)."""
    from ffea_amd import formats as F, units as U
    d = os.path.join(GOLDEN, "physics", "cubes_lj")
    t = F.read_ssint_types(os.path.join(d, "veryFine_topface0.vdw"), 1256)
    assert (t == 0).sum() == 200 and (t == -1).sum() == 1056
    E, R, K = F.read_ssint_params(os.path.join(d, "veryFine.lj"))
    assert E.shape == (7, 7) and E[0, 0] == 2.537e17 * U.MESO_AREA * U.MESO_AREA / U.MESO_ENERGY
    assert R[0, 0] == 1e-9 / U.MESO_LENGTH and R[1, 2] == 1e-10 / U.MESO_LENGTH and (K == 0).all()
    Eg, Rg, _ = F.read_ssint_params(os.path.join(d, "veryFine.lj"), "gensoft", 3e-9 / U.MESO_LENGTH)
    assert np.allclose(Rg, 3e-9 / U.MESO_LENGTH, rtol=1e-15) and (Eg == E).all()
    m = F.load_blob(*(os.path.join(d, "veryFine." + x) for x in ("node", "top", "mat", "stokes")))
    surf = F.read_surface(os.path.join(d, "veryFine.surf"))
    opp = F.surface_opposite(surf, m.topo)
    for (e, *nodes), o in zip(surf, opp):
        assert o in m.topo[e, :4] and o not in nodes
    with pytest.raises(F.FormatError):
        F.read_ssint_types(os.path.join(d, "veryFine_topface0.vdw"), 1000)


def test_analysis_on_reference_files():
    """The checker functions (ffea_amd.analysis) on the reference's own ffeatools fixtures:
    the measurement file's columns and the trajectory's frame centroids; and the closed forms
    of the checkers' formulas."""
    from ffea_amd import analysis as A
    H = A.read_measurements(os.path.join(FMT, "tet_ascii.1.fm"))
    assert list(H)[:2] == ["Time", "StrainEnergy"] and H["Time"][0] == 1e-14
    assert A.surface_energy(H) == H["SurfSurfEnergy"][-1]
    c = A.frame_centroids(os.path.join(FMT, "tet_ascii.1.ftj"))
    assert c.shape[1] == 3 and np.isfinite(c).all()
    # a straight walk of equal steps: <dr^2> is the step length squared
    cs = np.arange(5)[:, None] * np.array([[3e-9, 4e-9, 0.0]])
    r2, th, _ = A.stokes_diffusion(cs, 4.11e-21, 1e-3, 2.0, 1e-9, 1e-14, 100)
    assert r2 == pytest.approx(25e-18) and th == pytest.approx(6 * 4.11e-21 / (6 * np.pi * 1e-3 * 2e-9) * 1e-12)
    assert A.youngs_modulus(1e-9, 1e-11, 160e-9, 10e-9) == pytest.approx(1e-11 * 160e-9 / (np.pi * 1e-16 * 1e-9))
    assert A.flexural_rigidity(2e-9, 1e-12, 160e-9) == pytest.approx(1e-12 * (160e-9) ** 3 / 6e-9)
    e = A.equipartition({"StrainEnergy": np.full(50, 4.11e-21 * 183 / 2)}, 63, 4.11e-21)
    assert e["StrainEnergy"] == pytest.approx(0.0, abs=1e-15)
