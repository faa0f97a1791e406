"""Writes tests/golden/fixtures.npz from the CPU oracle: small known outputs
of the implicit-modeling path that the tests check both the oracle (against
drift) and the GPU library (without running the oracle) against.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def sample(e, n, seed):
    """n active voxels of an export (fixed seed): (idx (n,3) int32, values (n,) f32)."""
    li, off = np.nonzero(e.active())
    pick = np.sort(np.random.default_rng(seed).choice(li.size, min(n, li.size), replace=False))
    idx = e.origins[li[pick]] + np.stack([off[pick] >> 6, (off[pick] >> 3) & 7, off[pick] & 7], 1)
    return idx.astype(np.int32), e.values[li[pick], off[pick]]


def build():
    from oracle import oracle as O
    import baby_shark_amd as bs
    from baby_shark_amd import meshgen as mg
    out = {}
    # volume/mod.rs:134-152 up to the mesher: box2.stl at voxel 0.2 (band 0), offset 0.5
    box2 = bs.read_stl(os.path.join(HERE, "box2.stl"))
    v = O.mesh_to_volume(box2, 0.2, 0)
    e = v.export()
    out.update(box2_m2v_origins=e.origins, box2_m2v_masks=e.masks, box2_m2v_values=e.values)
    v.offset(0.5)
    e = v.export()
    out.update(box2_off_origins=e.origins, box2_off_masks=e.masks, box2_off_values=e.values,
               box2_off_dc=v.dual_contour(0.2))
    # the torus of the small workload: torus 128x64, h 0.02, band 3
    e = O.mesh_to_volume(mg.torus(1.0, 0.35, 128, 64, 0.01, seed=1), 0.02, 3).export()
    i, val = sample(e, 5000, 11)
    out.update(torus_origins=e.origins, torus_masks=e.masks, torus_idx=i, torus_values=val)
    # its sphere: r 0.5 at (1, 0, 0), h 0.02
    e = O.sphere(0.5, (1.0, 0.0, 0.0), 0.02).export()
    i, val = sample(e, 5000, 12)
    out.update(sphere_origins=e.origins, sphere_masks=e.masks, sphere_idx=i, sphere_values=val)
    # a small sphere's dual contouring
    out.update(small_sphere_dc=O.sphere(0.12, (0.0, 0.0, 0.0), 0.05).dual_contour(0.05))
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **build())
    print("wrote", os.path.join(HERE, "fixtures.npz"))
