"""Write tests/golden/config4_gradients.npz: float64 references of the differentiation workload's gradients.

    python tests/golden/make_config4_grads.py          (about two minutes on 8 cores)

The workload is config 4 at its full rope geometry: 8 of its ropes (65 vertices, config-1 material, 0.02 m
segments), 50 substeps of 1/3000 s per 1/60 s frame, 8 iterations, plane z = 0 and one sphere SDF (r 0.1 m on a
48^3 lattice) per rope, root program; 6 frames, in which ropes land on their spheres and one reaches the plane.
The target trajectory is the rope tips of the float64 oracle run with hidden per-vertex EA and EI scaled by
U(0.5, 2); the loss is the sum over frames of the squared tip distances to it, taken at the nominal material.

Stored:
- the hidden material, the nominal material (EA, EI, node masses), the drag and the root program;
- float64 torch autograd gradients through oracle/torch_rope.py (the transcription pinned by
  tests/test_torch_transcription.py) for EA, EI, mass, drag, p0, v0, p1;
- float32 autograd of the same transcription at EA (1 + k 1e-7), k = 0..3: how far float32 arithmetic alone moves
  each gradient family;
- central differences of the float64 C++ oracle per rope (the ropes never interact): for each per-vertex or
  per-rope family along the float64 gradient restricted to that rope, and for the drag, which every rope shares,
  from the per-rope loss terms of one pair of runs; step sizes of tests/test_torch_transcription.py (relative
  1e-6 for EA, drag and v0, 1e-5 for EI, 1e-7 for the masses, 1e-9 m for root positions).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "cosserat-rod-solver-cuda_amd"), os.path.dirname(HERE)]

FRAMES, ROPES = 6, 8
STEPS = {"ea": 1e-6, "ei": 1e-5, "m": 1e-7, "drag": 1e-6, "p0": 1e-9, "v0": 1e-6, "p1": 1e-9}


def workload():
    from ropesim import scenes
    return scenes.config4(n_ropes=ROPES, frames=FRAMES, grid_n=48)


def main():
    import torch

    from adjoint_cases import make_oracle, oracle_loss, program, torch_grads
    from oracle import pyoracle
    torch.set_num_threads(os.cpu_count())
    s = workload()
    prog = program(s)
    tips = s.offsets[1:] - 1
    ea_h, ei_h = s.hidden_material()
    o = make_oracle(pyoracle, s)
    o.set_stretch_stiffnesses(ea_h)
    o.set_bend_stiffnesses(ei_h)
    tgt = np.zeros((FRAMES, ROPES, 3))
    for f in range(FRAMES):
        o.step(prog[0][f], prog[1][f], s.dt, s.substeps, root_pos_end=prog[2][f])
        tgt[f] = o.positions()[tips]
    w = make_oracle(pyoracle, s).inv_masses()
    m = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 1.0).astype(np.float32).astype(np.float64)
    nv = s.num_vertices
    mat = (np.full(nv, np.float32(s.k_stretch), np.float64), np.full(nv, np.float32(s.k_bend), np.float64), m)
    t0 = time.time()
    g64, loss = torch_grads(pyoracle, s, mat, s.drag, prog, tips, tgt)
    print(f"float64 autograd: loss {loss:.6e} ({time.time() - t0:.0f} s)", flush=True)
    g32 = []
    for p in (0.0, 1e-7, 2e-7, 3e-7):
        g, _ = torch_grads(pyoracle, s, (mat[0] * (1 + p), mat[1], mat[2]), s.drag, prog, tips, tgt,
                           dtype=torch.float32)
        g32.append(g)
    print(f"float32 autograd x4 ({time.time() - t0:.0f} s)", flush=True)

    V = s.num_vertices // ROPES

    def rope_losses(name, h, v):
        """Per-rope tip losses of the float64 oracle with family `name` moved by h v."""
        mm, pp, dd = list(mat), list(prog), s.drag
        if name in ("ea", "ei", "m"):
            i = ("ea", "ei", "m").index(name)
            mm[i] = mat[i] + h * v
        elif name == "drag":
            dd = s.drag + h * v
        else:
            i = ("p0", "v0", "p1").index(name)
            pp[i] = prog[i] + h * v
        o = make_oracle(pyoracle, s)
        o.set_stretch_stiffnesses(mm[0])
        o.set_bend_stiffnesses(mm[1])
        o.set_masses(mm[2])
        o.drag = dd
        per = np.zeros(ROPES)
        for f in range(FRAMES):
            o.step(pp[0][f], pp[1][f], s.dt, s.substeps, root_pos_end=pp[2][f])
            per += ((o.positions()[tips] - tgt[f]) ** 2).sum(1)
        return per

    out = dict(ea_hidden=ea_h, ei_hidden=ei_h, ea=mat[0], ei=mat[1], m=mat[2], drag=np.float64(s.drag),
               p0=prog[0], v0=prog[1], p1=prog[2], tgt=tgt, tips=tips, loss=np.float64(loss))
    for k in g64:
        out["g64_" + k] = np.asarray(g64[k])
        for j, g in enumerate(g32):
            out[f"g32_{j}_" + k] = np.asarray(g[k])
    # central differences per rope (ropes never interact): for a per-vertex or per-rope family, along the float64
    # gradient restricted to the rope (material scaled by its value: relative steps); for the drag, shared by
    # every rope, the per-rope loss terms of one pair of runs
    for k in ("ea", "ei", "m", "p0", "v0", "p1"):
        dirs, fds = [], []
        for r in range(ROPES):
            g = np.asarray(g64[k])
            v = np.zeros_like(g)
            if k in ("ea", "ei", "m"):
                scale = mat[("ea", "ei", "m").index(k)]
                sl = slice(r * V, (r + 1) * V)
                v[sl] = scale[sl] * (g[sl] * scale[sl])
                v /= np.linalg.norm(g[sl] * scale[sl])
            else:
                v[:, r] = g[:, r] / np.linalg.norm(g[:, r])
            eps = STEPS[k]
            fds.append((rope_losses(k, eps, v)[r] - rope_losses(k, -eps, v)[r]) / (2 * eps))
            dirs.append(v)
        out["dir_" + k] = np.asarray(dirs)
        out["fd_" + k] = np.asarray(fds)
        ad = np.array([float(np.sum(np.asarray(g64[k]) * dirs[r])) for r in range(ROPES)])
        print(f"  {k}: float64 autograd vs central difference per rope, worst {np.max(np.abs(ad - fds) / np.abs(fds)):.1e}",
              flush=True)
    eps = STEPS["drag"]
    out["fd_drag"] = (rope_losses("drag", eps, 1.0) - rope_losses("drag", -eps, 1.0)) / (2 * eps)
    print(f"  drag: per-rope central differences sum {out['fd_drag'].sum():.6e}, float64 autograd {float(g64['drag']):.6e}",
          flush=True)
    np.savez_compressed(os.path.join(HERE, "config4_gradients.npz"), **out)
    print(f"wrote config4_gradients.npz ({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
