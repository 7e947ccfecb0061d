"""Write the golden fixtures of tests/test_golden.py from the float64 / float32 CPU oracle.

    python tests/golden/make_golden.py

The reference ships no data fixtures (its tests are analytic), so these come from the oracle. It is pinned by
those analytic checks (tests/test_oracle_*.py), and the fixtures freeze its outputs on three seeded cases:
- config1_trajectory.npz: the single-rope config-1 run (33 vertices from (0, 0, 1), root at 0.1 m/s, 60 frames
  of 50 substeps), float64 positions after every frame;
- rope_rope_lists.npz: rope-rope candidate lists and step limits of a 9-rope tangle at given positions and
  paddings (float32 oracle, the arithmetic of the device);
- mesh_broadphase.npz: per-vertex mesh candidate lists of scattered probes among the meshes of a small config-3
  scene (float32 oracle).
Each file also records the SHA-256 of the inputs it was made from, so a drifting generator is caught.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "cosserat-rod-solver-cuda_amd"), os.path.dirname(HERE)]


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


def tangle(rng, n_ropes=9, V=40):
    """Ropes crossing each other inside one small region (many close segment pairs)."""
    ropes = []
    for _ in range(n_ropes):
        c = rng.uniform(-0.02, 0.02, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        w = rng.normal(size=3) * 0.02
        t = np.arange(V)[:, None] * 0.015
        ropes.append(c + (t - t.mean()) * d + w * np.sin(t * 40))
    return ropes


def config1_case(oracle):
    from ropesim import scenes
    s = scenes.config1(frames=60)
    o = oracle.OracleBatch(s.positions, s.density, s.k_stretch, s.k_bend, s.max_h, s.num_itr, prec=64,
                           offsets=s.offsets)
    s.apply(o)
    fis = s.frame_inputs(60)
    xs = []
    for fi in fis:
        s.step(o, fi)
        xs.append(o.positions())
    drive = np.stack([np.concatenate([fi["p0"].ravel(), fi["v0"].ravel()]) for fi in fis])
    return dict(positions=np.stack(xs), input_sha=sha(s.positions, drive))


def rope_rope_case(oracle):
    rng = np.random.default_rng(7)
    ropes = tangle(rng)
    o = oracle.OracleBatch(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, prec=32)
    o.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 3)
    x = np.concatenate(ropes).astype(np.float32)
    x += rng.normal(0, 0.002, x.shape).astype(np.float32)
    pads = rng.uniform(0.004, 0.012, len(x)).astype(np.float32)
    counts, lists, lim = o.detect(x.astype(np.float64), pads.astype(np.float64))
    flat = np.concatenate([lists[i, :min(c, 128)] for i, c in enumerate(counts)]).astype(np.int32)
    offs = np.concatenate([[0], np.cumsum(np.minimum(counts, 128))]).astype(np.int64)
    return dict(rope_offsets=np.cumsum([0] + [len(r) for r in ropes]).astype(np.int32), x=x, pads=pads,
                counts=counts.astype(np.int32), lists=flat, list_offsets=offs, step_limit=lim.astype(np.float32),
                input_sha=sha(x, pads))


def mesh_case(oracle):
    from ropesim import scenes
    s = scenes.config3(n_ropes=8, frames=1, vmax=64, grid_n=16)
    o = oracle.OracleBatch(s.positions, s.density, s.k_stretch, s.k_bend, s.max_h, s.num_itr, prec=32,
                           offsets=s.offsets)
    s.apply(o)
    rng = np.random.default_rng(3)
    nv = s.num_vertices
    x = rng.uniform([-1.4, -1.4, 0.1], [1.4, 1.4, 1.7], (nv, 3)).astype(np.float32)
    dx = rng.normal(0, 0.003, (nv, 3)).astype(np.float32)
    counts, cand = o.broadphase(x.astype(np.float64), dx.astype(np.float64))
    return dict(x=x, dx=dx, counts=counts.astype(np.int32), cand=cand.astype(np.int32),
                input_sha=sha(s.positions, s.mesh_pos0, s.mesh_quat0, *[m.values for m in s.meshes], x, dx))


CASES = {"config1_trajectory": config1_case, "rope_rope_lists": rope_rope_case, "mesh_broadphase": mesh_case}


def main():
    from oracle import pyoracle
    pyoracle.lib()
    for name, fn in CASES.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **fn(pyoracle))
        print("wrote", name)


if __name__ == "__main__":
    main()
