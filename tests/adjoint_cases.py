"""Shared cases of the differentiation-form checks: a contact-rich scene, its root program, tip targets,
and the float64 references (torch autograd through oracle/torch_rope.py, and the C++ oracle's loss for
finite differences)."""
import numpy as np


def contact_scene(n_ropes=8, V=33, frames=6, substeps=20, seed=5, with_contact=True):
    """Ropes leaning 17 degrees below horizontal from roots at z = 0.19 m, a sphere (r 0.1) 1 cm below each
    rope's midpoint, plane z = 0: they hit the sphere within ~3 frames and the floor soon after."""
    from ropesim import scenes, sdf
    rng = np.random.default_rng(seed)
    seg, R, r = 0.02, 0.005, 0.1
    a = rng.uniform(0, 2 * np.pi, n_ropes)
    d = np.column_stack([np.cos(a), np.sin(a), np.full(n_ropes, -0.3)])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    roots = np.column_stack([rng.uniform(-0.3, 0.3, (n_ropes, 2)), np.full(n_ropes, 0.19)])
    pos = roots[:, None, :] + seg * np.arange(V)[None, :, None] * d[:, None, :]
    mid = pos[:, V // 2]
    centre = mid - np.array([0.0, 0.0, r + R + 0.01])
    grid_n, margin = 40, R + 0.02
    lo = -(r + margin)
    sp = 2 * (r + margin) / (grid_n - 1)
    g = sdf.from_function(sdf.sdf_sphere(r), (lo, lo, lo), (lo + sp * (grid_n - 1),) * 3, sp)
    s = scenes.Scene("contact", pos.reshape(-1, 3).astype(np.float32), (np.arange(n_ropes + 1) * V).astype(np.int32),
                     0.1, 500.0, 1e-4, 1.0 / 60.0 / substeps, 8, 1.0 / 60.0, substeps, frames, drag=0.2,
                     planes=np.array([[0, 0, 1, 0]], np.float32),
                     contact=dict(radius=R, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0),
                     meshes=[scenes.Mesh(g.values, g.origin, g.spacing, 0.5, 0.4)],
                     mesh_pos0=centre[:, None, :].astype(np.float32),
                     mesh_quat0=np.tile(np.array([0, 0, 0, 1], np.float32), (n_ropes, 1, 1)), seed=seed)
    if not with_contact:
        s.planes, s.contact, s.meshes, s.mesh_pos0, s.mesh_quat0 = None, None, [], None, None
    return s


def rope_contact_scene(n_groups=2, V=25, frames=6, substeps=20, seed=9):
    """Rope-rope contact in groups of two: in each environment rope B starts 1.2 cm above rope A and crosses
    it at right angles, both leaning down from roots that follow the root program, so the ropes press on and
    slide along each other within the first frame. No surfaces (the barrier, friction and step limit carry
    the contact)."""
    from ropesim import scenes
    rng = np.random.default_rng(seed)
    seg = 0.02
    ropes = []
    for g in range(n_groups):
        c = np.array([0.5 * g, 0.0, 0.3])
        for k in range(2):
            ang = rng.uniform(-0.2, 0.2) + (0.0 if k == 0 else np.pi / 2)
            d = np.array([np.cos(ang), np.sin(ang), -0.15 - 0.1 * k])
            d /= np.linalg.norm(d)
            L = seg * (V - 1)
            root = c - 0.5 * L * np.array([np.cos(ang), np.sin(ang), 0.0]) + np.array([0.0, 0.0, 0.012 * k])
            ropes.append(root + seg * np.arange(V)[:, None] * d)
    pos = np.concatenate(ropes).astype(np.float32)
    off = (np.arange(len(ropes) + 1) * V).astype(np.int32)
    rc = dict(radius=0.004, barrier=0.001, k_collision=0.025, mu=0.4, k_friction=1e7, detection_scaler=1.2,
              group_ropes=2)
    return scenes.Scene("rope_contact", pos, off, 0.1, 500.0, 1e-4, 1.0 / 60.0 / substeps, 8, 1.0 / 60.0, substeps,
                        frames, drag=0.2, rope_contact=rc, seed=seed)


def program(s):
    fis = s.frame_inputs(s.frames)
    return [np.stack([f[k] for f in fis]).astype(np.float64) for k in ("p0", "v0", "p1")]


def targets(s):
    rng = np.random.default_rng(11)
    tips = s.offsets[1:] - 1
    return tips, s.positions[tips][None].astype(np.float64) + rng.normal(0.0, 0.05, (s.frames, len(tips), 3))


def nominal_material(s, b):
    nv = s.num_vertices
    w = np.asarray(b.inv_masses(), np.float64).reshape(-1)
    m = np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 1.0).astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(3)
    ea = (s.k_stretch * rng.uniform(0.7, 1.3, nv)).astype(np.float32).astype(np.float64)
    ei = (s.k_bend * rng.uniform(0.7, 1.3, nv)).astype(np.float32).astype(np.float64)
    return ea, ei, m


def make_oracle(oracle, s):
    o = oracle.OracleBatch(s.positions, s.density, s.k_stretch, s.k_bend, s.max_h, s.num_itr, prec=64,
                           offsets=s.offsets)
    s.apply(o)
    return o


def torch_grads(oracle, s, mat, drag, prog, tips, tgt, dtype=None):
    import torch
    dtype = torch.float64 if dtype is None else dtype
    from oracle.torch_rope import TorchRope
    o = make_oracle(oracle, s)
    tr = TorchRope(o, s.gravity, planes=() if s.planes is None else s.planes,
                   contact=[s.contact[k] for k in ("radius", "mu_s", "mu_k", "v_slip", "kappa")] if s.contact else
                   (0.0, 0.5, 0.5, 1e-3, 2.0),
                   mesh_specs=[dict(values=m.values, origin=m.origin, spacing=m.spacing, mu_s=m.mu_s, mu_k=m.mu_k)
                               for m in s.meshes], dtype=dtype, rope_contact=s.rope_contact)
    t = lambda a: torch.tensor(np.asarray(a, np.float64), dtype=dtype, requires_grad=True)
    EA, EI, M = (t(a) for a in mat)
    P0, V0, P1 = (t(a) for a in prog)
    D = torch.tensor(float(drag), dtype=dtype, requires_grad=True)
    x = tr.rollout(EA, EI, M, D, P0, V0, P1, s.dt, s.substeps, s.num_itr)
    loss = ((x[:, tips].double() - torch.tensor(tgt)) ** 2).sum()
    loss.backward()
    g = dict(ea=EA, ei=EI, m=M, drag=D, p0=P0, v0=V0, p1=P1)
    return {k: v.grad.double().numpy() for k, v in g.items()}, float(loss.detach())


def oracle_loss(oracle, s, mat, drag, prog, tips, tgt, forces=False):
    o = make_oracle(oracle, s)
    o.set_stretch_stiffnesses(mat[0])
    o.set_bend_stiffnesses(mat[1])
    o.set_masses(mat[2])
    o.drag = drag
    if forces:
        o.enable_contact_forces(True)
    p0, v0, p1 = prog
    loss, touched = 0.0, 0
    for f in range(s.frames):
        o.step(p0[f], v0[f], s.dt, s.substeps, root_pos_end=p1[f])
        loss += float(((o.positions()[tips] - tgt[f]) ** 2).sum())
        if forces:
            touched |= int(np.abs(o.mesh_wrenches()).sum() > 0)
            touched |= int((o.positions()[:, 2] < s.contact["radius"]).any()) << 1
    return loss, touched


def rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


