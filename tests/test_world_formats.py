"""Output and checkpoint formats of ffea_amd.world against files the reference's own runs wrote
(tests/golden/formats: tet_ascii.1.ftj / .fm from tests/rods/unit/connection_orientation,
cyl_160_fine.fcp from tests/physics/cyl_160_fine): the writers reproduce them byte for byte
from the values the readers parse, and the restart reader keeps only complete frames."""
import os

import numpy as np

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
    assert lines[k + 1].startswith("".join("%-14s" % c for c in W.MEAS_COLUMNS))
    for line in lines[k + 2:]:
        if not line.strip():
            continue
        vals = [float(v) for v in line.split()[:6]]
        assert W.meas_line(vals).rstrip("\n") == line[:84]


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
