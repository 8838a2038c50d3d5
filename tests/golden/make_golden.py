"""Writes the golden vectors of tests/golden/*.npz from the reference's OWN compiled
objects (oracle/_ref/libffea_ref.so, built by `make -C oracle ref` from /root/reference).

    python tests/golden/make_golden.py

Each scenario stores its inputs and the reference's outputs; tests/test_golden.py checks
the oracle against them bit for bit, and the GPU tests use them as a second anchor.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import pyoracle  # noqa: E402
from conftest import box_case, sphere_case  # noqa: E402


def params_array(prm):
    return np.array([prm.dt, prm.kT, prm.epsilon2, prm.max_iterations_cg, prm.calc_noise,
                     prm.calc_stokes, prm.stokes_visc, prm.rng_seed], np.float64)


def blob_scenario(ref, name, mesh, prm, scale, x0, nsteps=3):
    seed = ref.package_seed(prm.rng_seed)
    ref.set_package_seed(seed)
    s1 = ref.stream()
    b = ref.blob(mesh, prm, scale, pos=x0)
    out = dict(rest=mesh.pos, x0=x0 if x0 is not None else mesh.pos, topo=mesh.topo,
               material=mesh.material, stokes_radius=mesh.stokes_radius, pinned=mesh.pinned,
               params=params_array(prm), scale=np.float64(scale), seed=seed)
    inv, rhs = b.forces_only(s1)
    out.update(rhs=rhs, dpsi=b.get("dpsi"), vol=b.get("vol"), visc=b.get("visc"),
               node_force=b.get("node_force"), stokes_drag=b.get("stokes_drag"),
               material_meso=b.get("material"), J_inv_0=b.get("J_inv_0"), vol_0=b.get("vol_0"))
    st, v = b.solve(rhs)
    key, col, val = b.csr()
    out.update(solve_v=v, csr_key=key, csr_col=col, csr_val=val, rng_after_forces=s1.state())
    b.close()
    # a fresh run of nsteps World::run iterations
    ref.set_package_seed(seed)
    s2 = ref.stream()
    b = ref.blob(mesh, prm, scale, pos=x0)
    traj = []
    for _ in range(nsteps):
        rc = pyoracle.world_step(ref, [b], s2)
        assert rc == 0, rc
        traj.append(b.get("pos"))
    out.update(traj=np.array(traj), vel=b.get("vel"), rng_after_steps=s2.state())
    b.close()
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, {k: np.asarray(v).shape for k, v in out.items()})


def rng_scenario(ref):
    seed = ref.package_seed(44)
    ref.set_package_seed(seed)
    g = ref.stream()
    draws = np.array([g.randu01() for _ in range(64)])
    st = [g.state()]
    for e, c in ((5, 3), (0, 7 * 3072), (10, 12345), (-127, 0), (0, 1)):
        g.advance(e, c)
        st.append(g.state())
    draws2 = np.array([g.randu01() for _ in range(8)])
    np.savez_compressed(os.path.join(HERE, "rng_seed44.npz"), seed=seed, draws=draws,
                        states=np.array(st), draws_after=draws2,
                        advances=np.array([(5, 3), (0, 7 * 3072), (10, 12345), (-127, 0), (0, 1)]))
    print("rng_seed44", draws[:3])


def main():
    pyoracle.build(with_ref=True)
    ref = pyoracle.Reference()
    rng_scenario(ref)
    m, prm, sc, x0 = box_case((3, 3, 2), noise=1)
    blob_scenario(ref, "box332_noise", m, prm, sc, x0)
    m, prm, sc, x0 = box_case((3, 3, 2), noise=0)
    blob_scenario(ref, "box332_nonoise", m, prm, sc, x0)
    m, prm, sc, x0 = box_case((3, 2, 2), noise=1, jitter=0.15, seed=7)
    m.pinned = np.array([0, 1, 5], np.int32)
    blob_scenario(ref, "box322_pinned_jitter", m, prm, sc, x0)
    m, prm, sc = sphere_case()
    blob_scenario(ref, "sphere_63_120", m, prm, sc, None)


if __name__ == "__main__":
    main()
