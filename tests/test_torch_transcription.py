"""The float64 torch transcription (oracle/torch_rope.py) that serves as the autograd reference of the adjoint.

Pinned two ways, on a scene whose ropes hit a sphere SDF and the ground plane, and on a scene of crossing
ropes in rope-rope contact (barrier, friction and the step limit active):
- its forward follows the C++ float64 oracle (rope_oracle.hpp, itself pinned by the analytic tests) to
  rounding;
- its autograd gradients equal central finite differences of the C++ oracle's loss.
So a dropped term or a wrong sign in the transcription fails here, not only in the GPU comparison.
"""
import numpy as np
import pytest

from adjoint_cases import (contact_scene, make_oracle, oracle_loss, program, rel, rope_contact_scene, targets,
                           torch_grads)


@pytest.fixture(scope="module")
def small(oracle):
    s = contact_scene(n_ropes=4, V=33, frames=6, substeps=20)
    o = make_oracle(oracle, s)
    w = o.inv_masses()
    m = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 1.0)
    rng = np.random.default_rng(3)
    mat = (s.k_stretch * rng.uniform(0.7, 1.3, s.num_vertices), s.k_bend * rng.uniform(0.7, 1.3, s.num_vertices), m)
    prog = program(s)
    tips, tgt = targets(s)
    g, loss = torch_grads(oracle, s, mat, s.drag, prog, tips, tgt)
    return dict(s=s, mat=mat, prog=prog, tips=tips, tgt=tgt, g=g, loss=loss)


def test_scene_has_sphere_and_plane_contact(oracle, small):
    s = small["s"]
    _, touched = oracle_loss(oracle, s, small["mat"], s.drag, small["prog"], small["tips"], small["tgt"], forces=True)
    assert touched == 3, f"contact coverage bits {touched:b} (1: sphere, 2: plane)"


def test_forward_follows_float64_oracle(oracle, small):
    import torch
    from oracle.torch_rope import TorchRope
    s, mat, (p0, v0, p1) = small["s"], small["mat"], small["prog"]
    o = make_oracle(oracle, s)
    o.set_stretch_stiffnesses(mat[0])
    o.set_bend_stiffnesses(mat[1])
    o.set_masses(mat[2])
    tr = TorchRope(make_oracle(oracle, s), s.gravity, planes=s.planes,
                   contact=[s.contact[k] for k in ("radius", "mu_s", "mu_k", "v_slip", "kappa")],
                   mesh_specs=[dict(values=m.values, origin=m.origin, spacing=m.spacing, mu_s=m.mu_s, mu_k=m.mu_k)
                               for m in s.meshes])
    with torch.no_grad():
        xs = tr.rollout(*mat, s.drag, *(torch.tensor(a) for a in (p0, v0, p1)), s.dt, s.substeps, s.num_itr).numpy()
    moved = 0.0
    for f in range(s.frames):
        o.step(p0[f], v0[f], s.dt, s.substeps, root_pos_end=p1[f])
        x = o.positions()
        moved = max(moved, float(np.abs(x - s.positions).max()))
        assert np.abs(xs[f] - x).max() < 1e-8, f
    assert moved > 0.05


@pytest.mark.parametrize("which,eps", [("ea", 1e-6), ("ei", 1e-5), ("m", 1e-7), ("drag", 1e-6), ("p1", 1e-9),
                                       ("p0", 1e-9), ("v0", 1e-6)])
def test_autograd_equals_central_differences(oracle, small, which, eps):
    """Directional derivatives; steps relative to the parameter scale. The loss is strongly curved in the
    root positions and in the mass. The penalty stiffness m (kappa / h)^2 is large, and the friction law has
    kinks. So those steps are small (the central difference converges to the autograd value as 1e-5, 1e-7,
    1e-9 steps give 4e-2, 7e-5, 2e-8 relative error for p1). EI takes 1e-5 because its absolute scale
    (1e-4 N m^2) puts a 1e-6 step into rounding noise."""
    s, mat, prog, tips, tgt, g = small["s"], small["mat"], small["prog"], small["tips"], small["tgt"], small["g"]
    rng = np.random.default_rng({"ea": 1, "ei": 2, "m": 3, "drag": 4, "p1": 5, "p0": 6, "v0": 7}[which])
    if which in ("ea", "ei", "m"):
        i = ("ea", "ei", "m").index(which)
        v = rng.normal(size=mat[i].shape) * mat[i]

        def loss_at(h):
            mm = list(mat)
            mm[i] = mat[i] + h * v
            return oracle_loss(oracle, s, mm, s.drag, prog, tips, tgt)[0]
    elif which == "drag":
        v = 1.0

        def loss_at(h):
            return oracle_loss(oracle, s, mat, s.drag + h, prog, tips, tgt)[0]
    else:
        i = ("p0", "v0", "p1").index(which)
        v = rng.normal(size=prog[i].shape)

        def loss_at(h):
            p = list(prog)
            p[i] = prog[i] + h * v
            return oracle_loss(oracle, s, mat, s.drag, p, tips, tgt)[0]
    fd = (loss_at(eps) - loss_at(-eps)) / (2 * eps)
    ad = float(np.sum(g[which] * v))
    assert rel(ad, fd) < 1e-4, (ad, fd)


@pytest.fixture(scope="module")
def crossing(oracle):
    s = rope_contact_scene()
    o = make_oracle(oracle, s)
    w = o.inv_masses()
    m = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 1.0)
    rng = np.random.default_rng(3)
    mat = (s.k_stretch * rng.uniform(0.7, 1.3, s.num_vertices), s.k_bend * rng.uniform(0.7, 1.3, s.num_vertices), m)
    prog = program(s)
    tips, tgt = targets(s)
    g, loss = torch_grads(oracle, s, mat, s.drag, prog, tips, tgt)
    return dict(s=s, mat=mat, prog=prog, tips=tips, tgt=tgt, g=g, loss=loss)


def test_rope_contact_forward_follows_float64_oracle(oracle, crossing):
    """The transcription's rope-rope detection, barrier, friction and step clamp against the C++ oracle; the
    same rollout without rope contact lands centimetres away, and every mechanism acts."""
    import torch
    from oracle.torch_rope import TorchRope
    s, mat, (p0, v0, p1) = crossing["s"], crossing["mat"], crossing["prog"]
    o, off = make_oracle(oracle, s), make_oracle(oracle, s)
    off.set_rope_contact(False)
    for b in (o, off):
        b.set_stretch_stiffnesses(mat[0])
        b.set_bend_stiffnesses(mat[1])
        b.set_masses(mat[2])
    tr = TorchRope(make_oracle(oracle, s), s.gravity, rope_contact=s.rope_contact)
    with torch.no_grad():
        xs = tr.rollout(*mat, s.drag, *(torch.tensor(a) for a in (p0, v0, p1)), s.dt, s.substeps, s.num_itr).numpy()
    effect = 0.0
    for f in range(s.frames):
        o.step(p0[f], v0[f], s.dt, s.substeps, root_pos_end=p1[f])
        off.step(p0[f], v0[f], s.dt, s.substeps, root_pos_end=p1[f])
        assert np.abs(xs[f] - o.positions()).max() < 1e-8, f
        effect = max(effect, float(np.abs(o.positions() - off.positions()).max()))
    assert effect > 0.01
    assert min(tr.stats.get(k, 0) for k in ("barrier", "friction", "clamped")) > 100, tr.stats


@pytest.mark.parametrize("which,eps", [("ea", 1e-6), ("ei", 1e-5), ("m", 1e-6), ("drag", 1e-6), ("p1", 1e-8),
                                       ("p0", 1e-8), ("v0", 1e-6)])
def test_rope_contact_autograd_equals_central_differences(oracle, crossing, which, eps):
    """Directional derivatives as above, through rope-rope contact. The central differences at these steps sit
    within 1e-5 of autograd (the root positions need 1e-8: a 1e-6 step changes contact pairs' barrier
    curvature enough to move the difference by 3e-3)."""
    s, mat, prog, tips, tgt, g = (crossing[k] for k in ("s", "mat", "prog", "tips", "tgt", "g"))
    rng = np.random.default_rng({"ea": 1, "ei": 2, "m": 3, "drag": 4, "p1": 5, "p0": 6, "v0": 7}[which])
    if which in ("ea", "ei", "m"):
        i = ("ea", "ei", "m").index(which)
        v = rng.normal(size=mat[i].shape) * mat[i]

        def loss_at(h):
            mm = list(mat)
            mm[i] = mat[i] + h * v
            return oracle_loss(oracle, s, mm, s.drag, prog, tips, tgt)[0]
    elif which == "drag":
        v = 1.0

        def loss_at(h):
            return oracle_loss(oracle, s, mat, s.drag + h, prog, tips, tgt)[0]
    else:
        i = ("p0", "v0", "p1").index(which)
        v = rng.normal(size=prog[i].shape)

        def loss_at(h):
            p = list(prog)
            p[i] = prog[i] + h * v
            return oracle_loss(oracle, s, mat, s.drag, p, tips, tgt)[0]
    fd = (loss_at(eps) - loss_at(-eps)) / (2 * eps)
    ad = float(np.sum(g[which] * v))
    assert rel(ad, fd) < 1e-4, (ad, fd)
