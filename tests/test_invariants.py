"""Invariants of the reference's own test suite, restated on the oracle and (on a GPU) on the HIP path in both
arithmetic modes:
- bitwise inertness of contact objects that are never touched: reference tests/test_contact.py:108-131,
  tests/test_mesh_broadphase.py:51-68, tests/test_mesh_motion.py:227-238;
- bitwise independence of environments (an environment runs identically alone or in a batch):
  tests/test_contact.py:134-151,203-222, tests/test_mesh_contact.py:258-273;
- deterministic replay in fresh objects: tests/test_deterministic_sweeps.py:13-53;
- contact physics that pins the plane law beyond statics: no pass-through of rope-rope contact, measured with an
  exact segment distance (tests/test_contact.py:25-36,153-171); candidate overflow falls back to scanning every
  mesh (tests/test_mesh_broadphase.py:71-83); impact penetration and bounce, bounce independent of the substep,
  convergence of the contact solve over the iteration count (tests/test_contact.py:225-293).
On the device, environment independence also checks the persistent kernel's packing of contact groups into
cells: a batch of one and a batch of four pack the same rope into different cells and slots.
"""
import numpy as np
import pytest

from conftest import fp32

NV, SEG, RHO, EA, EI = 21, 0.05, 0.117, 11473.0, 4.48e-3
R = 0.00635
H, ITR, G = 1.0 / 6000.0, 16, 9.81
KAPPA, VSLIP = 2.0, 1e-3
PEN0 = G * H * H / KAPPA ** 2
RR = dict(radius=0.9 * R, barrier=0.2 * R, k_collision=0.025, k_friction=1e7)
DT, SUB = 1 / 60, 100
Q0 = np.array([0, 0, 0, 1], np.float32)


def seg_dist(p0, p1, q0, q1):
    """Exact closest distance between segments p0p1 and q0q1 (broadcasting): the textbook clamped solution
    (Ericson, Real-Time Collision Detection, 5.1.9), independent of the solver's closest-point routine."""
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    a, e, f = (d1 * d1).sum(-1), (d2 * d2).sum(-1), (d2 * r).sum(-1)
    c, bb = (d1 * r).sum(-1), (d1 * d2).sum(-1)
    den = a * e - bb * bb
    s = np.where(den > 1e-12, np.clip((bb * f - c * e) / np.maximum(den, 1e-12), 0, 1), 0.0)
    t = (bb * s + f) / np.maximum(e, 1e-12)
    s = np.where(t < 0, np.clip(-c / np.maximum(a, 1e-12), 0, 1),
                 np.where(t > 1, np.clip((bb - c) / np.maximum(a, 1e-12), 0, 1), s))
    t = np.clip(t, 0, 1)
    return np.linalg.norm(p0 + d1 * s[..., None] - (q0 + d2 * t[..., None]), axis=-1)


def line(n, direction, origin):
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    return (np.asarray(origin, np.float64) + SEG * np.arange(n)[:, None] * d).astype(np.float32)


def hanging(solver, roots, direction=(0.0, 0.0, -1.0), itr=ITR):
    """Ropes hanging from pinned roots (the reference's batch() of tests/test_contact.py)."""
    pos = np.stack([line(NV, direction, r) for r in roots])
    o = solver.OracleBatch(pos, RHO, EA, EI, H, itr, prec=64)
    o.damping = 0.0
    o.drag = 0.0
    o.set_gravity([0.0, 0.0, -G])
    return o


def free(solver, pos, drag=0.0, mu_s=0.5, mu_k=None):
    """Unpinned ropes with the contact material set (tests/test_mesh_broadphase.py batch())."""
    pos = np.asarray(pos, np.float32)
    o = solver.OracleBatch(pos, RHO, EA, EI, H, ITR, pinned=np.zeros(pos.shape[:2], bool), prec=64)
    o.damping = 0.0
    o.drag = drag
    o.set_gravity([0.0, 0.0, -G])
    o.set_contact(radius=R, mu_s=mu_s, mu_k=mu_k, v_slip=VSLIP, kappa=KAPPA)
    return o


def rows(n, z=0.02, y=0.0):
    pos = np.zeros((n, NV, 3), np.float32)
    pos[:, :, 0] = np.arange(NV) * SEG
    pos[:, :, 1] = y
    pos[:, :, 2] = z
    return pos


@pytest.fixture(scope="module")
def grids():
    from ropesim import sdf
    slab = sdf.from_function(sdf.sdf_box((3.0, 3.0, 0.125), (0, 0, -0.125)), (-0.3125, -0.6, -0.1875),
                             (1.3125, 0.6, 0.1875), 1 / 128)
    box = sdf.from_function(sdf.sdf_box((3.0, 1.0, 0.125), (0, 0, -0.125)), (-0.3125, -0.3125, -0.1875),
                            (1.3125, 0.3125, 0.1875), 1 / 128)
    return dict(slab=slab, box=box)


def per_rope(o, n):
    return o.positions().reshape(n, NV, 3)


# ------------------------------------------------------------------ bitwise inertness


def test_planes_without_contact_are_bitwise_inert(solver):
    """A plane far below the rope changes nothing (the contact-off path is unchanged)."""
    def run(with_plane):
        o = solver.OracleBatch(line(NV, (1, 0, 0), (0, 0, 0))[None], RHO, EA, EI, H, 8, prec=64)
        o.damping = 0.0
        if with_plane:
            o.set_planes(np.array([[0.0, 1.0, 0.0, -50.0]]), radius=R)
        for f in range(40):
            t = f / 60
            o.step(np.array([[0.1 * np.sin(4 * t), 0.0, 0.0]]), np.array([[0.4 * np.cos(4 * t), 0.0, 0.0]]), DT, SUB)
        return o.positions()
    assert np.array_equal(run(False), run(True))


def test_far_meshes_are_inert(solver, grids):
    """24 meshes, the rope's floor is mesh 20 and the others sit 5 m away (never candidates): the rest sink is
    the one-mesh value and the run is bitwise the one-mesh run."""
    g = grids["slab"]

    def run(many):
        o = free(solver, rows(1), drag=0.1)
        if many:
            for m in range(24):
                p = np.array([0, 0, -0.02]) if m == 20 else np.array([5.0 + m, 5.0, -0.02])
                o.add_mesh(*g.as_args(), pos=p)
        else:
            o.add_mesh(*g.as_args(), pos=np.array([0, 0, -0.02]))
        for _ in range(60):
            o.step_free(DT, SUB)
        return o.positions()
    one, many = run(False), run(True)
    assert one[:, 2].min() == pytest.approx(-0.02 + R - PEN0, abs=1e-7 if fp32(solver) else 3e-8)
    assert np.array_equal(one, many)


def test_static_target_is_inert(solver, grids):
    """set_mesh_poses with the pose the mesh already has leaves the steps bitwise unchanged."""
    g = grids["slab"]

    def run(call):
        o = free(solver, rows(1), drag=0.1)
        o.add_mesh(*g.as_args(), pos=np.array([0, 0, -0.02]))
        for _ in range(30):
            if call:
                o.set_mesh_poses(0, np.array([0, 0, -0.02]), Q0)
            o.step_free(DT, SUB)
        return o.positions()
    assert np.array_equal(run(False), run(True))


# ------------------------------------------------------------------ environment independence


def test_batch_plane_contact_matches_single_env(solver):
    """Plane contact is vertex-local: env 0 is bitwise identical alone or in a batch of four (its root lowered
    0.3 m onto the floor)."""
    def run(n):
        roots = np.array([[3.0 * e, 0.0, 1.1] for e in range(n)])
        tilt = np.radians(2.0)
        o = hanging(solver, roots, direction=(np.sin(tilt), 0.0, -np.cos(tilt)))
        o.set_planes(np.array([[0.0, 0.0, 1.0, 0.0]]), radius=R, mu_s=0.5, mu_k=0.4)
        for f in range(60):
            p0 = roots + np.array([0.0, 0.0, -0.3 * f / 60])
            p1 = roots + np.array([0.0, 0.0, -0.3 * (f + 1) / 60])
            o.step(p0, np.tile([0.0, 0.0, -0.3], (n, 1)), DT, SUB, root_pos_end=p1)
        return per_rope(o, n)[0]
    one = run(1)
    assert (one[:, 2] < 0.05).any() and (one[:, 2] > R - 1e-4).all()  # on the floor, not through it
    assert np.array_equal(one, run(4))


def test_rope_contact_groups_isolate_envs(solver):
    """Two envs at the same spot do not see each other (group = 1 rope): env 0 is bitwise identical while
    env 1's rope swings straight through it."""
    def run(n):
        o = hanging(solver, np.array([[0.0, 0.0, 1.1]] * n))
        o.set_rope_contact(True, RR["radius"], RR["barrier"], RR["k_collision"], 0.4, RR["k_friction"], 1.2, 1)
        for f in range(30):
            p = np.array([[0.0, 0.0, 1.1]] * n)
            v = np.zeros((n, 3))
            p1 = p.copy()
            if n == 2:
                t0, t1 = f / 60, (f + 1) / 60
                p[1, 0] = 0.3 * np.sin(6 * t0)
                p1[1, 0] = 0.3 * np.sin(6 * t1)
                v[1, 0] = (p1[1, 0] - p[1, 0]) * 60
            o.step(p, v, DT, SUB, root_pos_end=p1)
        return per_rope(o, n)[0]
    out1, out2 = run(1), run(2)
    assert np.isfinite(out2).all()
    assert np.array_equal(out1, out2)


def test_batch_mesh_env_independent(solver, grids):
    """Per-env mesh poses: env 0 is bitwise identical alone or beside envs whose meshes sit elsewhere."""
    g = grids["box"]

    def run(n):
        pos = rows(n, z=0.2)
        pos[:, :, 1] = 3.0 * np.arange(n)[:, None]
        o = free(solver, pos, drag=0.1)
        p = np.zeros((n, 3))
        p[:, 1] = 3.0 * np.arange(n)
        p[:, 2] = -0.01 * np.arange(n)
        o.add_mesh(*g.as_args(), pos=p, quat=np.tile(Q0, (n, 1)))
        for _ in range(30):
            o.step_free(DT, SUB)
        return per_rope(o, n)[0]
    one = run(1)
    assert (one[:, 2] < 0.05).all()  # it reached the floor
    assert np.array_equal(one, run(4))


# ------------------------------------------------------------------ deterministic replay


@pytest.mark.parametrize("strings,vertices,iterations", [(1, 3, 3), (1, 63, 3), (1, 64, 3), (1, 127, 3),
                                                         (1, 128, 3), (6, 124, 3), (6, 124, 24)])
def test_deterministic_state_replays(solver, strings, vertices, iterations):
    """Fresh objects driven by the same commands reproduce every frame's state bit for bit (endpoint-pinned
    strings whose first end moves)."""
    dt, substeps = 1 / 480, 32
    u = np.linspace(0, 1, vertices, dtype=np.float32)
    initial = np.zeros((strings, vertices, 3), np.float32)
    initial[:, :, 0] = 0.64 * u
    initial[:, :, 1] = np.arange(strings)[:, None] * 0.009
    initial[:, :, 2] = (0.002 + 0.001 * np.arange(strings)[:, None]) * np.sin(u * np.pi)
    pins = np.zeros(initial.shape[:2], bool)
    pins[:, [0, -1]] = True
    roots = initial[:, 0].astype(np.float64)
    program = []
    for step in range(16):
        start, end = roots.copy(), roots.copy()
        start[:, 0] -= 0.00001 * step
        end[:, 0] -= 0.00001 * (step + 1)
        program.append((start, (end - start) / dt, end))

    def run():
        o = solver.OracleBatch(initial, 0.005, 7500.0, 1e-7, dt / substeps, iterations, pinned=pins,
                               straight_rest=True, prec=64)
        o.set_gravity([0.0, 0.0, -9.81])
        o.damping = 0.0
        o.drag = 20.0
        states = []
        for start, velocity, end in program:
            o.step(start, velocity, dt, substeps, root_pos_end=end)
            states.append((o.positions().copy(), o.velocities().copy(), o.quaternions().copy()))
            assert all(np.isfinite(a).all() for a in states[-1])
        return states
    first, second = run(), run()
    assert not np.array_equal(first[-1][0].reshape(strings, vertices, 3)[:, 1:-1], initial[:, 1:-1])
    for a, b in zip(first, second):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


# ------------------------------------------------------------------ contact physics


def test_rope_rope_no_passthrough(solver):
    """Rope B's root circles rope A: the ropes wind round each other but never come closer than the barrier
    wall (exact segment distances of the end-of-frame positions)."""
    o = hanging(solver, np.array([[0.0, 0.0, 1.1], [0.08, 0.0, 1.1]]))
    o.set_rope_contact(True, RR["radius"], RR["barrier"], RR["k_collision"], 0.4, RR["k_friction"], 1.2, 2)
    w = 2 * np.pi * 0.5
    closest = np.inf
    for f in range(90):
        t0, t1 = f / 60, (f + 1) / 60
        pa = np.array([0.0, 0.0, 1.1])
        p0 = np.stack([pa, [0.08 * np.cos(w * t0), 0.08 * np.sin(w * t0), 1.1]])
        p1 = np.stack([pa, [0.08 * np.cos(w * t1), 0.08 * np.sin(w * t1), 1.1]])
        o.step(p0, (p1 - p0) * 60, DT, SUB, root_pos_end=p1)
        P = per_rope(o, 2)
        closest = min(closest, seg_dist(P[0, :-1, None], P[0, 1:, None], P[1, None, :-1], P[1, None, 1:]).min())
    assert closest > 2 * RR["radius"]  # the barrier wall is at 11.43 mm, contact at 12.7 mm


def test_candidate_overflow_scans_every_mesh(solver, grids):
    """More coincident floors than the 8 candidate slots (the vertex falls back to scanning every mesh): the
    penalties share the weight, so the sink is PEN0 / n, as two coincident floors give PEN0 / 2."""
    g = grids["slab"]

    def sink(n_floors):
        o = free(solver, rows(1), drag=0.1)
        for _ in range(n_floors):
            o.add_mesh(*g.as_args(), pos=np.array([0, 0, -0.02]))
        for _ in range(60):
            o.step_free(DT, SUB)
        return -0.02 + R - o.positions()[:, 2].min()
    assert sink(2) == pytest.approx(PEN0 / 2, rel=0.05)
    assert sink(10) == pytest.approx(PEN0 / 10, rel=0.1)


def drop_flat(solver, v0, substeps=100, itr=ITR):
    """A free flat rope 1e-4 m above a plane at z = -(R + 1e-4), falling at v0."""
    h = 1 / (60 * substeps)
    o = solver.OracleBatch(line(NV, (1, 0, 0), (0, 0, 0))[None], RHO, EA, EI, h, itr,
                           pinned=np.zeros((1, NV), bool), prec=64)
    o.damping = 0.0
    o.drag = 0.0
    o.set_gravity([0.0, 0.0, -G])
    o.set_planes(np.array([[0.0, 0.0, 1.0, -(R + 1e-4)]]), radius=R, kappa=KAPPA)
    v = np.zeros((NV, 3))
    v[:, 2] = -v0
    o.set_state("velocities", v)
    return o


def test_impact_penetration_and_bounce(solver):
    """Flat 10 m/s impact: the penetration stays well under the rope radius; a small numerical bounce."""
    o = drop_flat(solver, 10.0)
    m = 1.0 / o.inv_masses()
    pen, up = 0.0, 0.0
    for _ in range(300):  # 50 ms, one substep per call
        o.step_free(H, 1)
        pen = max(pen, -(o.positions()[:, 2] + 1e-4).min())
        up = max(up, (m * o.velocities()[:, 2]).sum() / m.sum())
    assert pen < 0.6e-3  # the reference measures 0.31 mm
    assert 0.03 < up / 10.0 < 0.2  # numerical bounce at kappa = 2 (~0.10)


def test_bounce_independent_of_substep_size(solver):
    """k = m (kappa / h)^2 makes the (numerical) restitution the same at any substep count."""
    def bounce(substeps):
        h = 1 / (60 * substeps)
        o = drop_flat(solver, 5.0, substeps=substeps)
        m = 1.0 / o.inv_masses()
        up = 0.0
        for _ in range(int(0.06 / h)):
            o.step_free(h, 1)
            up = max(up, (m * o.velocities()[:, 2]).sum() / m.sum())
        return up / 5.0
    assert bounce(50) == pytest.approx(bounce(200), abs=0.04)


def test_contact_solve_converges_with_iterations(solver):
    """No limit cycle: a vertex touching down mid-substep ends at the same gap for 15, 16 and 17 iterations."""
    gaps = []
    for itr in (15, 16, 17):
        o = solver.OracleBatch(line(NV, (1, 0, 0), (0, 0, 0))[None], RHO, EA, EI, H, itr,
                               pinned=np.zeros((1, NV), bool), prec=64)
        o.damping = 0.0
        o.drag = 0.0
        o.set_gravity([0.0, 0.0, -G])
        o.set_planes(np.array([[0.0, 0.0, 1.0, -(R + 0.9 * 5.0 * H)]]), radius=R, kappa=KAPPA)
        v = np.zeros((NV, 3))
        v[:, 2] = -5.0
        o.set_state("velocities", v)
        o.step_free(H, 1)
        gaps.append(o.positions()[NV // 2, 2])
    assert max(gaps) - min(gaps) < (3e-7 if fp32(solver) else 1e-7)
