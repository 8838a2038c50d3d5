"""Writes tests/golden/physics/cyl_160_fine.npz from the reference's own cyl_160_fine fixture
(tests/physics/cyl_160_fine: node, topology, material, Stokes radii, pins, surface faces) and
copies the Young's-modulus and bending scripts with their ctforces files
(tests/physics/E_cyl_160_fine, tests/physics/EI_cyl_160_fine).  Run in a checkout that has the
reference at /root/reference; the tests rebuild the text files from the arrays."""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from ffea_amd import formats as F  # noqa: E402

REF = "/root/reference/tests/physics"


def main():
    d = os.path.join(REF, "cyl_160_fine")
    m = F.load_blob(*(os.path.join(d, "cyl_160_fine." + x) for x in ("node", "top", "mat", "stokes", "pin")))
    surf = F.read_surface(os.path.join(d, "cyl_160_fine.surf"))
    np.savez_compressed(os.path.join(HERE, "physics", "cyl_160_fine.npz"), pos=m.pos, topo=m.topo,
                        material=m.material, stokes=m.stokes_radius, pinned=m.pinned, surface=surf)
    for t, sub in (("E", "cyl_160_fine_E"), ("EI", "cyl_160_fine_EI")):
        for f in (f"cyl_160_fine-{t}.ffea", "cyl_160_fine_surf.ctforces"):
            shutil.copy(os.path.join(REF, f"{t}_cyl_160_fine", f), os.path.join(HERE, "physics", sub, f))


if __name__ == "__main__":
    main()
