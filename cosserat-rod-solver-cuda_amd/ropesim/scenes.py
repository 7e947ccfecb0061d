"""Synthetic workloads of the named rope configurations, with seeded inputs.

A Scene holds everything needed to build the same batch on any object with the RopeBatch surface
(the HIP RopeBatch, or a CPU checker with the same method names) and to drive it frame by frame:
geometry (flat positions + offsets), material, knobs, planes, meshes and their per-frame poses, rope
contact, and the per-frame root program. All draws come from numpy.random.default_rng(seed) in a fixed
order, so both sides see identical float32 inputs.

Configurations (sizes scale down through the arguments for quick checks):
  config1  single rope, 33 vertices hanging from (0, 0, 1), root moving at 0.1 m/s along +x
  config2  N ropes x 65 vertices, lower-hemisphere directions (tips above the floor), root program, ground plane
  config3  N ropes with V ~ U{16..256}, per-segment EA / EI, contact groups of 4 with rope-rope
           contact, 8 kinematic SDF meshes per environment on SE(3) random walks, ground plane
  config4  N ropes x 65 (config-1 material), plane + one static sphere SDF per rope
  config5  16384 ropes x 96 (config-1 material, per-vertex jitter), 64 meshes on 96^3 grids, groups of 4
  resting  N free ropes x 65 lying on one slab mesh (every vertex in contact): the force-readout workload
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import sdf as _sdf


def _quat_exp(w):
    """Unit quaternions (x, y, z, w) of rotation vectors w [..., 3]."""
    th = np.linalg.norm(w, axis=-1, keepdims=True)
    s = np.where(th > 0, np.sin(th / 2) / np.maximum(th, 1e-30), 0.5)
    return np.concatenate([w * s, np.cos(th / 2)], -1)


def _quat_mul(a, b):
    av, aw = a[..., :3], a[..., 3:]
    bv, bw = b[..., :3], b[..., 3:]
    v = aw * bv + bw * av + np.cross(av, bv)
    w = aw * bw - (av * bv).sum(-1, keepdims=True)
    return np.concatenate([v, w], -1)


@dataclass
class Mesh:
    values: np.ndarray
    origin: np.ndarray
    spacing: float
    mu_s: float = 0.5
    mu_k: float = 0.4


@dataclass
class Scene:
    name: str
    positions: np.ndarray          # float32 [nv, 3]
    offsets: np.ndarray            # int32 [N + 1]
    density: float
    k_stretch: float
    k_bend: float
    max_h: float
    num_itr: int
    dt: float
    substeps: int
    frames: int
    gravity: tuple = (0.0, 0.0, -9.81)
    drag: float = 0.0
    damping: float = 0.0
    drag_quad: float = 0.0         # quadratic normal drag [1/m]
    drag_quad_t: float = 0.0       # quadratic tangential drag [1/m] (acts only with drag_quad > 0)
    pinned: np.ndarray | None = None
    fix_root_orientation: bool = False
    ea: np.ndarray | None = None   # float32 [nv] per segment at its first vertex
    ei: np.ndarray | None = None
    masses: np.ndarray | None = None  # float32 [nv] node masses (pinned entries ignored)
    planes: np.ndarray | None = None
    contact: dict | None = None    # radius, mu_s, mu_k, v_slip, kappa
    meshes: list = field(default_factory=list)
    mesh_pos0: np.ndarray | None = None   # [N, M, 3] initial poses (teleport)
    mesh_quat0: np.ndarray | None = None
    rope_contact: dict | None = None
    drive: str = "program"         # "program" | "const" | "free"
    seed: int = 0
    _rng_state: dict | None = None

    @property
    def num_ropes(self):
        return len(self.offsets) - 1

    @property
    def num_vertices(self):
        return int(self.offsets[-1])

    # ---- building ------------------------------------------------------------------------
    def apply(self, b):
        """Configure a freshly constructed batch (HIP RopeBatch or the CPU checker)."""
        b.set_gravity(np.asarray(self.gravity, np.float32))
        b.drag = self.drag
        b.damping = self.damping
        if self.drag_quad > 0.0 or self.drag_quad_t > 0.0:
            b.drag_quad = self.drag_quad
            b.drag_quad_t = self.drag_quad_t
        if self.ea is not None:
            b.set_stretch_stiffnesses(self.ea)
        if self.ei is not None:
            b.set_bend_stiffnesses(self.ei)
        if self.masses is not None:
            b.set_masses(self.masses)
        if self.contact is not None:
            b.set_contact(**self.contact)
        if self.planes is not None:
            b.set_planes(self.planes, **self.contact)
        for m, mesh in enumerate(self.meshes):
            b.add_mesh(mesh.values, mesh.origin, mesh.spacing, pos=self.mesh_pos0[:, m], quat=self.mesh_quat0[:, m],
                       mu_s=mesh.mu_s, mu_k=mesh.mu_k)
        if self.rope_contact is not None:
            b.set_rope_contact(True, **self.rope_contact)
        return b

    def build(self, ropesim_module, **kw):
        b = ropesim_module.RopeBatch.from_positions(self.positions, self.density, self.k_stretch, self.k_bend,
                                                   self.max_h, self.num_itr, pinned=self.pinned,
                                                   fix_root_orientation=self.fix_root_orientation,
                                                   offsets=self.offsets, **kw)
        return self.apply(b)

    def hidden_material(self, seed=7, lo=0.5, hi=2.0):
        """Per-vertex EA and EI (float32 [nv]) scaled by U(lo, hi) from the scene's own: the material a target
        trajectory is generated with when the scene's material is to be identified from it."""
        rng = np.random.default_rng(seed)
        nv = self.num_vertices
        ea = np.full(nv, self.k_stretch, np.float64) if self.ea is None else self.ea.astype(np.float64)
        ei = np.full(nv, self.k_bend, np.float64) if self.ei is None else self.ei.astype(np.float64)
        return (ea * rng.uniform(lo, hi, nv)).astype(np.float32), (ei * rng.uniform(lo, hi, nv)).astype(np.float32)

    # ---- driving -------------------------------------------------------------------------
    def frame_inputs(self, nframes=None):
        """Per-frame float32 arrays: dict of p0 [N,3], v0 [N,3], p1 [N,3] and, with meshes, mesh_pos /
        mesh_quat [N, M, 3/4] (targets at the end of the frame). Deterministic in the scene seed."""
        nframes = self.frames if nframes is None else nframes
        rng = np.random.default_rng(self.seed + 1000)
        N = self.num_ropes
        roots = self.positions[self.offsets[:-1]].astype(np.float64)
        mp = None if self.mesh_pos0 is None else self.mesh_pos0.astype(np.float64).copy()
        mq = None if self.mesh_quat0 is None else self.mesh_quat0.astype(np.float64).copy()
        out = []
        t = 0.0
        for f in range(nframes):
            d = {}
            if self.drive == "const":      # constant velocity 0.1 m/s along +x (config 1)
                v = np.tile([0.1, 0.0, 0.0], (N, 1))
                p0 = roots + v * t
                d.update(p0=p0, v0=v, p1=None)
            elif self.drive == "program":
                v0 = rng.normal(0.0, 0.2, (N, 3))
                p1 = roots + rng.uniform(-0.05, 0.05, (N, 3))
                d.update(p0=roots.copy(), v0=v0, p1=p1)
                roots = p1
            if mp is not None and self.name.startswith("config3"):
                G = 4
                E = (N + G - 1) // G
                M = mp.shape[1]
                dp = rng.normal(0.0, 0.01, (E, M, 3))
                dw = rng.normal(0.0, 0.02, (E, M, 3))
                env = np.arange(N) // G
                mp = mp + dp[env]
                mq = _quat_mul(_quat_exp(dw[env]), mq)
                mq /= np.linalg.norm(mq, axis=-1, keepdims=True)
                d.update(mesh_pos=mp.astype(np.float32), mesh_quat=mq.astype(np.float32))
            out.append({k: (None if v is None else np.ascontiguousarray(v, np.float32)) for k, v in d.items()})
            t += self.dt
        return out

    def step(self, b, fi, to_dev=None):
        """Advance one frame with inputs `fi`; to_dev converts float32 arrays for the HIP batch."""
        cv = (lambda a: a) if to_dev is None else to_dev
        if "mesh_pos" in fi:
            b.set_mesh_poses_all(cv(fi["mesh_pos"]), cv(fi["mesh_quat"]))
        if self.drive == "free":
            b.step_free(self.dt, self.substeps)
        elif self.drive == "const":
            b.step(cv(fi["p0"]), cv(fi["v0"]), self.dt, self.substeps)
        else:
            b.step(cv(fi["p0"]), cv(fi["v0"]), self.dt, self.substeps, root_pos_end=cv(fi["p1"]))


# -------------------------------------------------------------------------------------------- configs

def _lower_hemisphere(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 2] = -np.abs(d[:, 2])
    return d


def _hanging_clear_of_floor(rng, roots, length, clearance=0.05):
    """Lower-hemisphere directions, redrawn per rope until the straight rope's tip is at least `clearance`
    above the plane z = 0: a rope that starts below the floor is thrown out of it by the penalty force in
    its first substeps, which no configuration asks for."""
    d = _lower_hemisphere(rng, len(roots))
    for r in range(len(roots)):
        for _ in range(1000):
            if roots[r, 2] + length * d[r, 2] > clearance:
                break
            d[r] = _lower_hemisphere(rng, 1)[0]
        else:
            d[r] = np.array([1.0, 0.0, 0.0])
    return d


def config1(frames=60):
    """Single rope: 33 vertices straight down from (0, 0, 1), 0.03 m segments, rho 0.1, EA 500,
    EI 1e-4, h 1/3000 (50 substeps per 1/60 s frame), 8 iterations, linear drag 0.2, root at 0.1 m/s."""
    V, seg = 33, 0.03
    pos = np.zeros((V, 3), np.float32)
    pos[:, 2] = 1.0 - seg * np.arange(V)
    return Scene("config1", pos, np.array([0, V], np.int32), 0.1, 500.0, 1e-4, 1.0 / 3000.0, 8, 1.0 / 60.0, 50,
                 frames, drag=0.2, drive="const")


def config2(n_ropes=1024, V=65, frames=120, seed=0):
    """N ropes x V vertices, 0.02 m segments along lower-hemisphere directions, rho 0.117, EA 11473,
    EI 4.48e-3, h 1/6000 (100 substeps), 8 iterations, vertex 0 pinned, root program, plane z = 0."""
    rng = np.random.default_rng(seed)
    p0 = np.column_stack([rng.uniform(-0.5, 0.5, (n_ropes, 2)), rng.uniform(0.8, 1.2, n_ropes)])
    seg = 0.02
    d = _hanging_clear_of_floor(rng, p0, seg * (V - 1))
    pos = (p0[:, None, :] + seg * np.arange(V)[None, :, None] * d[:, None, :]).astype(np.float32)
    contact = dict(radius=0.00635, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0)
    return Scene("config2", pos.reshape(-1, 3), (np.arange(n_ropes + 1) * V).astype(np.int32), 0.117, 11473.0,
                 4.48e-3, 1.0 / 6000.0, 8, 1.0 / 60.0, 100, frames, planes=np.array([[0, 0, 1, 0]], np.float32),
                 contact=contact, seed=seed)


def _shape_library(rng, M, grid_n, R):
    """M closed-form shapes (spheres, boxes, capsules, cycling) sampled on grid_n^3 lattices."""
    meshes, radii = [], []
    for m in range(M):
        kind = m % 3
        if kind == 0:
            r = rng.uniform(0.05, 0.2)
            phi, ext = _sdf.sdf_sphere(r), np.array([r, r, r])
        elif kind == 1:
            half = rng.uniform(0.03, 0.15, 3)
            phi, ext = _sdf.sdf_box(half), half
        else:
            r, hl = rng.uniform(0.03, 0.08), rng.uniform(0.05, 0.15)
            phi, ext = _sdf.sdf_capsule(r, hl), np.array([r, r, r + hl])
        margin = R + 0.02
        lo, hi = -(ext + margin), ext + margin
        spacing = float((hi - lo).max() / (grid_n - 1))
        g = _sdf.from_function(phi, lo, lo + spacing * (grid_n - 1), spacing)
        meshes.append(Mesh(g.values, g.origin, g.spacing, 0.5, 0.4))
        radii.append(float(np.linalg.norm(ext)))
    return meshes, radii


def config3(n_ropes=4096, frames=2, seed=1, n_meshes=8, grid_n=64, vmin=16, vmax=256, group=4):
    """Batching form: n_ropes with V ~ U{vmin..vmax}, 0.015 m segments, per-segment EA ~ U(2e3, 2e4),
    EI ~ U(1e-4, 1e-2), rho 0.117, h 1/6000 (100 substeps), 8 iterations; contact groups of `group`
    ropes with rope-rope contact (radius 0.004 m); n_meshes SDF meshes per environment (grid_n^3 float32
    grids of spheres, boxes and capsules) posed by SE(3) random walks (0.01 m, 0.02 rad per frame); plane
    z = 0; roots pinned and driven by the root program. All environments share one region of space (they
    never interact), which keeps float32 coordinates near the origin."""
    rng = np.random.default_rng(seed)
    R = 0.004
    Vs = rng.integers(vmin, vmax + 1, n_ropes)
    seg = 0.015
    roots = np.column_stack([rng.uniform(-0.5, 0.5, (n_ropes, 2)), rng.uniform(0.8, 1.2, n_ropes)])
    ropes = []
    for r in range(n_ropes):
        L = seg * (Vs[r] - 1)
        for _ in range(1000):
            d = _lower_hemisphere(rng, 1)[0]
            if roots[r, 2] + L * d[2] > 0.05:
                break
        else:
            d = np.array([1.0, 0.0, 0.0])
        ropes.append(roots[r] + seg * np.arange(Vs[r])[:, None] * d[None, :])
    off = np.zeros(n_ropes + 1, np.int32)
    off[1:] = np.cumsum(Vs)
    pos = np.concatenate(ropes).astype(np.float32)
    nv = int(off[-1])
    ea = rng.uniform(2e3, 2e4, nv).astype(np.float32)
    ei = rng.uniform(1e-4, 1e-2, nv).astype(np.float32)
    meshes, radii = _shape_library(rng, n_meshes, grid_n, R)
    E = (n_ropes + group - 1) // group
    # mesh poses per environment: clear of the environment's ropes at the start
    mpos = np.zeros((E, n_meshes, 3))
    mq = np.zeros((E, n_meshes, 4))
    for e in range(E):
        verts = pos[off[e * group]:off[min((e + 1) * group, n_ropes)]].astype(np.float64)
        for m in range(n_meshes):
            for _ in range(200):
                c = np.array([rng.uniform(-1.2, 1.2), rng.uniform(-1.2, 1.2), rng.uniform(0.3, 1.5)])
                if np.linalg.norm(verts - c, axis=1).min() > radii[m] + 0.05:
                    break
            mpos[e, m] = c
            mq[e, m] = _quat_exp(rng.normal(0, 1.0, 3))
    env = np.arange(n_ropes) // group
    contact = dict(radius=R, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0)
    rc = dict(radius=0.9 * R, barrier=0.2 * R, k_collision=0.025, mu=0.4, k_friction=1e7, detection_scaler=1.2,
              group_ropes=group)
    return Scene("config3", pos, off, 0.117, 11473.0, 4.48e-3, 1.0 / 6000.0, 8, 1.0 / 60.0, 100, frames, ea=ea, ei=ei,
                 planes=np.array([[0, 0, 1, 0]], np.float32), contact=contact, meshes=meshes,
                 mesh_pos0=mpos[env].astype(np.float32), mesh_quat0=mq[env].astype(np.float32), rope_contact=rc,
                 seed=seed)


def config4(n_ropes=256, V=65, frames=120, seed=4, grid_n=48):
    """Differentiation form: n_ropes x V with config-1 material (rho 0.1 kg/m, EA 500 N, EI 1e-4 N m^2, linear
    drag 0.2 1/s) on 0.02 m segments along lower-hemisphere directions, h 1/3000 (50 substeps per 1/60 s frame),
    8 iterations, roots driven by the root program; plane z = 0 and one static sphere SDF (r 0.1 m on a
    grid_n^3 lattice) per rope whose top lies 2 cm below the rope's midpoint, so the falling rope lands on it."""
    rng = np.random.default_rng(seed)
    p0 = np.column_stack([rng.uniform(-0.5, 0.5, (n_ropes, 2)), rng.uniform(0.8, 1.2, n_ropes)])
    seg = 0.02
    d = _hanging_clear_of_floor(rng, p0, seg * (V - 1))
    pos = (p0[:, None, :] + seg * np.arange(V)[None, :, None] * d[:, None, :]).astype(np.float32)
    R = 0.005
    r = 0.1
    margin = R + 0.02
    lo = -(r + margin)
    spacing = 2 * (r + margin) / (grid_n - 1)
    g = _sdf.from_function(_sdf.sdf_sphere(r), (lo, lo, lo), (lo + spacing * (grid_n - 1),) * 3, spacing)
    side = np.cross(d, np.array([0.0, 0.0, 1.0]))
    ns = np.linalg.norm(side, axis=1, keepdims=True)
    side = np.where(ns > 1e-3, side / np.maximum(ns, 1e-30), np.array([1.0, 0.0, 0.0]))
    # the sphere's top 2 cm below the rope's midpoint: the middle of the falling rope lands on it within a few
    # frames while the rest swings on
    c = pos[:, V // 2].astype(np.float64) - np.array([0.0, 0.0, r + R + 0.02])
    c[:, 2] = np.maximum(c[:, 2], r + 0.01)
    for i in range(n_ropes):  # keep the start clear of the sphere
        while np.linalg.norm(pos[i].astype(np.float64) - c[i], axis=1).min() < r + 0.02:
            c[i] += 0.02 * side[i]
    mpos = c[:, None, :].astype(np.float32)
    mq = np.tile(np.array([0, 0, 0, 1], np.float32), (n_ropes, 1, 1))
    contact = dict(radius=R, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0)
    return Scene("config4", pos.reshape(-1, 3), (np.arange(n_ropes + 1) * V).astype(np.int32), 0.1, 500.0, 1e-4,
                 1.0 / 3000.0, 8, 1.0 / 60.0, 50, frames, drag=0.2, planes=np.array([[0, 0, 1, 0]], np.float32),
                 contact=contact, meshes=[Mesh(g.values, g.origin, g.spacing, 0.5, 0.4)], mesh_pos0=mpos,
                 mesh_quat0=mq, seed=seed)


def config5(n_ropes=16384, V=96, frames=3, seed=5, n_meshes=64, grid_n=96, group=4, n_shapes=8, rope_contact=True):
    """Scale: n_ropes x V on 0.01 m segments with config-1 material (rho 0.1, EA 500, EI 1e-4) jittered by
    U(-20 %, +20 %) per vertex, h 1/6000 (100 substeps), 8 iterations; contact groups of `group` ropes with
    rope-rope contact; n_meshes SDF meshes per environment on grid_n^3 float32 grids (spheres, boxes, capsules
    and tori from closed forms, n_shapes distinct lattices reused across the slots), placed once per
    environment; plane z = 0; roots driven by the root program. Environments share one region of space."""
    rng = np.random.default_rng(seed)
    seg, R = 0.01, 0.004
    roots = np.column_stack([rng.uniform(-0.5, 0.5, (n_ropes, 2)), rng.uniform(1.0, 1.3, n_ropes)])
    d = _lower_hemisphere(rng, n_ropes)
    d[:, 2] = np.minimum(d[:, 2], -0.5)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (roots[:, None, :] + seg * np.arange(V)[None, :, None] * d[:, None, :]).astype(np.float32)
    nv = n_ropes * V
    jitter = lambda: rng.uniform(0.8, 1.2, nv).astype(np.float32)
    ea, ei = 500.0 * jitter(), 1e-4 * jitter()
    lumped = np.full((n_ropes, V), 0.1 * seg, np.float32)  # rho l; the tip carries half a segment
    lumped[:, -1] *= 0.5
    masses = lumped.reshape(-1) * jitter()
    shapes, radii = [], []
    for k in range(n_shapes):
        kind = k % 4
        if kind == 0:
            r = rng.uniform(0.05, 0.12)
            phi, ext = _sdf.sdf_sphere(r), np.array([r, r, r])
        elif kind == 1:
            half = rng.uniform(0.03, 0.1, 3)
            phi, ext = _sdf.sdf_box(half), half
        elif kind == 2:
            r, hl = rng.uniform(0.02, 0.05), rng.uniform(0.04, 0.1)
            phi, ext = _sdf.sdf_capsule(r, hl), np.array([r, r, r + hl])
        else:
            a, b = rng.uniform(0.05, 0.09), rng.uniform(0.015, 0.03)
            phi, ext = _sdf.sdf_torus(a, b), np.array([a + b, a + b, b])
        margin = R + 0.02
        lo, hi = -(ext + margin), ext + margin
        spacing = float((hi - lo).max() / (grid_n - 1))
        g = _sdf.from_function(phi, lo, lo + spacing * (grid_n - 1), spacing)
        shapes.append(Mesh(g.values, g.origin, g.spacing, 0.5, 0.4))
        radii.append(float(np.linalg.norm(ext)))
    meshes = [shapes[m % n_shapes] for m in range(n_meshes)]
    E = (n_ropes + group - 1) // group
    # one layout of the mesh slots per environment, kept clear of the environment's ropes at the start
    mpos = np.zeros((E, n_meshes, 3))
    mq = np.zeros((E, n_meshes, 4))
    for e in range(E):
        verts = pos[e * group:min((e + 1) * group, n_ropes)].reshape(-1, 3).astype(np.float64)
        for m in range(n_meshes):
            for _ in range(100):
                c = np.array([rng.uniform(-1.5, 1.5), rng.uniform(-1.5, 1.5), rng.uniform(0.15, 1.6)])
                if np.linalg.norm(verts - c, axis=1).min() > radii[m % n_shapes] + 0.03:
                    break
            mpos[e, m] = c
            mq[e, m] = _quat_exp(rng.normal(0, 1.0, 3))
    env = np.arange(n_ropes) // group
    contact = dict(radius=R, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0)
    rc = dict(radius=0.9 * R, barrier=0.2 * R, k_collision=0.025, mu=0.4, k_friction=1e7, detection_scaler=1.2,
              group_ropes=group) if rope_contact else None
    s = Scene("config5", pos.reshape(-1, 3), (np.arange(n_ropes + 1) * V).astype(np.int32), 0.1, 500.0, 1e-4,
              1.0 / 6000.0, 8, 1.0 / 60.0, 100, frames, ea=ea, ei=ei, planes=np.array([[0, 0, 1, 0]], np.float32),
              contact=contact, meshes=meshes, mesh_pos0=mpos[env].astype(np.float32),
              mesh_quat0=mq[env].astype(np.float32), rope_contact=rc, masses=masses, seed=seed)
    return s


def resting(n_ropes=4096, V=65, frames=8, grid_spacing=1 / 64):
    """Force-readout workload: n_ropes free ropes x V vertices (1 cm segments, config-1 material) lying side by
    side on one slab mesh (a box SDF whose top face is z = 0, on a lattice of `grid_spacing`), released 0.5 mm
    above it; h 1/6000 (100 substeps per 1/60 s frame), 8 iterations, no drag. Once settled every vertex
    presses on the slab in every substep, so each wave's wrench contributions all go to one (rope, mesh)
    sum or two."""
    R = 0.004
    pos = np.zeros((n_ropes, V, 3), np.float32)
    pos[:, :, 0] = np.arange(V) * 0.01 - 0.005 * (V - 1)
    pos[:, :, 1] = np.linspace(-0.5, 0.5, n_ropes)[:, None]
    pos[:, :, 2] = R + 0.0005
    half = 0.005 * (V - 1) + 0.1
    g = _sdf.from_function(_sdf.sdf_box((half, 0.6, 0.05), (0, 0, -0.05)), (-half - 0.05, -0.65, -0.12),
                           (half + 0.05, 0.65, 0.03), grid_spacing)
    contact = dict(radius=R, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0)
    return Scene("resting", pos.reshape(-1, 3), (np.arange(n_ropes + 1) * V).astype(np.int32), 0.1, 500.0, 1e-4,
                 1.0 / 6000.0, 8, 1.0 / 60.0, 100, frames, pinned=np.zeros(n_ropes * V, bool), contact=contact,
                 meshes=[Mesh(g.values, g.origin, g.spacing, 0.5, 0.4)],
                 mesh_pos0=np.zeros((n_ropes, 1, 3), np.float32),
                 mesh_quat0=np.tile(np.array([0, 0, 0, 1], np.float32), (n_ropes, 1, 1)), drive="free")
