"""Pin the CPU oracle against what the mathematics fixes; the same analytic checks run on the HIP path.

The reference ships no data fixtures; its tests are analytic (tests/test_statics.py, test_dynamics.py,
test_drag_quad.py, test_contact.py, test_batch.py of the reference). Each check here restates one of
those known answers on the oracle, chosen so that a dropped term, a wrong sign or a transposed index in
the restatement fails it. Tolerances are the reference tests' own. The `solver` fixture runs each analytic
check on the oracle (CPU) and on the HIP RopeBatch (GPU, through tests/hip_adapter.py); checks of exact
identities use float32 bounds on the device. The detection and broad-phase enumerations stay oracle-only
(the device's lists are compared with the oracle's bit for bit in tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

from conftest import fp32

G = 9.81


def rope_line(n, seg, direction, origin=(0, 0, 0)):
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    return (np.asarray(origin, np.float64) + seg * np.arange(n)[:, None] * d).astype(np.float32)


def make(oracle, pos, density=0.1, EA=500.0, EI=1e-6, max_h=1 / 3000, itr=8, prec=64, **kw):
    o = oracle.OracleBatch([pos], density, EA, EI, max_h, itr, prec=prec, **kw)
    o.drag = 0.0
    o.damping = 0.0
    return o


def advance(o, dt, max_h, frames):
    """RopeSim.advance(dt) without explicit substeps: ceil(dt / max_h) in float32."""
    S = max(1, int(np.ceil(np.float32(dt) / np.float32(max_h))))
    for _ in range(frames):
        o.step_free(dt, S)
    return S


def test_advance_substep_count_matches_reference_rounding():
    # 1/60 over 1/3000 in float32 is 50.000004 -> 51 substeps (docs/SOLVER.md of the reference)
    assert int(np.ceil(np.float32(1 / 60) / np.float32(1 / 3000))) == 51


@pytest.mark.parametrize("prec", [64, 32])
def test_hanging_elongation_matches_EA_theory(solver, prec):
    """examples/hanging_rope.py: 33 x 3 cm, rho 0.1, EA 500, EI 1e-4; elongation m g L / 2EA = 9.041e-4 m."""
    if getattr(solver, "FP32", False) and prec == 64:
        pytest.skip("the device computes in float32 (covered by prec=32)")
    n, seg = 33, 0.03
    o = make(solver, rope_line(n, seg, (0, 0, -1)), EA=500.0, EI=1e-4, max_h=(1 / 60) / 50, prec=prec)
    o.set_gravity([0, 0, -G])
    o.drag = 2.0
    for _ in range(240):
        o.step_free(1 / 60, 50)
    X = o.positions()
    L0 = (n - 1) * seg
    dL = np.linalg.norm(np.diff(X, axis=0), axis=1).sum() - L0
    theory = 0.1 * L0 * G * L0 / (2 * 500.0)
    assert dL == pytest.approx(theory, rel=2e-3)
    # The tip stays on the axis. Correctly rounded arithmetic keeps the symmetric cancellations exact
    # (1e-6 m); fused multiply-add rounds one product of a cancelling pair only (fma(a, b, -c d) is not 0
    # when a b == c d), so the device's fast arithmetic drifts by float32 rounding: 1e-5 of rope length.
    lat = 1e-5 * L0 if getattr(solver, "FAST", False) else 1e-6
    assert abs(X[-1, 0]) < lat and abs(X[-1, 1]) < lat


def test_discrete_hang_elongation_per_segment(solver):
    """tests/test_statics.py:31: elongation_i = g * mass_below_i * l_i / EA, summed."""
    n, seg, rho, EA = 51, 0.02, 0.1, 500.0
    o = make(solver, rope_line(n, seg, (0, -1, 0)), density=rho, EA=EA)
    o.set_gravity([0, -9.8, 0])
    o.drag = 3.0
    advance(o, 1 / 60, 1 / 3000, 240)
    inv = o.inv_masses()
    m = np.where(inv > 0, 1 / np.where(inv > 0, inv, 1), 0)
    rest = o.rest_lengths()[: n - 1]
    tension = 9.8 * np.array([m[i + 1:].sum() for i in range(n - 1)])
    theory = (tension * rest / EA).sum()
    X = o.positions()
    meas = np.linalg.norm(np.diff(X, axis=0), axis=1).sum() - rest.sum()
    assert meas == pytest.approx(theory, rel=0.02)


def test_lumped_masses(solver):
    """tests/test_api.py inv_masses: interior 1/(rho l), tip 2/(rho l), pinned root 0."""
    n, seg, rho = 9, 0.02, 0.1
    inv = make(solver, rope_line(n, seg, (1, 0, 0)), density=rho).inv_masses()
    assert inv[0] == 0.0
    np.testing.assert_allclose(inv[1:-1], 1.0 / (rho * seg), rtol=1e-6)
    np.testing.assert_allclose(inv[-1], 1.0 / (rho * seg / 2), rtol=1e-6)


def test_initial_frames_rotate_ex_onto_segments(solver):
    """Frames of a straight rope along +x are the identity (test_api.py:110); in general R(q_i) e_x is the
    segment direction and the frames are twist-free (x-twist taken over from the previous segment)."""
    o = make(solver, rope_line(6, 0.02, (1, 0, 0)))
    np.testing.assert_allclose(o.quaternions()[:-1], np.tile([0, 0, 0, 1.0], (5, 1)), atol=1e-12)
    tol = 2e-6 if fp32(solver) else 1e-7
    rng = np.random.default_rng(0)
    pts = np.cumsum(rng.normal(size=(12, 3)) * 0.02, 0).astype(np.float32)
    o = make(solver, pts)
    q = o.quaternions()
    from scipy.spatial.transform import Rotation
    for i in range(11):
        d = (pts[i + 1] - pts[i]).astype(np.float64)
        ex = Rotation.from_quat(q[i]).apply([1.0, 0, 0])
        np.testing.assert_allclose(ex, d / np.linalg.norm(d), atol=tol)


def test_pendulum_period_of_hanging_chain(solver):
    """tests/test_dynamics.py:22: T = (4 pi / z1) sqrt(L / g), z1 the first zero of J0; rel 5%."""
    n, seg = 33, 0.02
    L = (n - 1) * seg
    th = np.deg2rad(2.0)
    o = make(solver, rope_line(n, seg, (np.sin(th), -np.cos(th), 0)))
    o.set_gravity([0, -9.8, 0])
    ts, xs = [], []
    S = max(1, int(np.ceil(np.float32(1 / 60) / np.float32(1 / 3000))))
    for k in range(360):
        o.step_free(1 / 60, S)
        ts.append((k + 1) / 60)
        xs.append(o.positions()[-1, 0])
    ts, xs = np.array(ts), np.array(xs)
    idx = np.where(np.diff(np.sign(xs)) != 0)[0]
    tc = ts[idx] - xs[idx] * (ts[idx + 1] - ts[idx]) / (xs[idx + 1] - xs[idx])
    T = 2.0 * np.mean(np.diff(tc))
    z1 = 2.404825557695773
    assert T == pytest.approx((4 * np.pi / z1) * np.sqrt(L / 9.8), rel=0.05)


def test_root_program_closed_form(solver):
    """tests/test_batch.py:195: the root lands on p1, the frame on q1; the root velocity is the last
    substep's midpoint sample v0 + a (dt - h / 2)."""
    N, V, H = 3, 10, 1 / 60
    ropes = [rope_line(V, 0.02, (0, -1, 0), (1.5 * e, 0, 0)) for e in range(N)]
    o = solver.OracleBatch(ropes, 0.1, 500.0, 1e-6, 1 / 3000, 8, fix_root_orientation=True, prec=64)
    o.damping = 0.0
    roots = np.array([r[0] for r in ropes], np.float64)
    p1 = roots + np.array([0.04, -0.02, 0.03])
    v0 = np.tile([0.5, 0.2, -0.1], (N, 1))
    th = np.deg2rad([10.0, 20.0, 30.0])
    q0 = np.tile([0, 0, 0, 1.0], (N, 1))
    q1 = np.zeros((N, 4))
    q1[:, 2], q1[:, 3] = np.sin(th / 2), np.cos(th / 2)
    o.step(roots, v0, H, 50, root_q=q0, root_pos_end=p1, root_q_end=q1)
    tol, rtol = (1e-6, 2e-5) if fp32(solver) else (1e-12, 1e-10)
    np.testing.assert_allclose(o.positions()[[0, V, 2 * V]], p1, atol=tol)
    np.testing.assert_allclose(o.quaternions()[[0, V, 2 * V]], q1, atol=tol)
    a = 2.0 * ((p1 - roots) - v0 * H) / H ** 2
    np.testing.assert_allclose(o.velocities()[[0, V, 2 * V]], v0 + a * (H - (H / 50) / 2), rtol=rtol, atol=tol)


def test_constant_velocity_root_advects_exactly(solver):
    """tests/test_api.py:86: advance(H, substeps=50) moves the pinned root by exactly v H."""
    o = make(solver, rope_line(8, 0.02, (0, -1, 0)))
    o.step(np.zeros((1, 3)), np.array([[0.12, 0, 0]]), 1 / 60, 50)
    np.testing.assert_allclose(o.positions()[0], [0.12 / 60, 0, 0], atol=1e-8 if fp32(solver) else 1e-12)


def test_quadratic_drag_closed_form(solver):
    """tests/test_drag_quad.py:56: a translating straight rope in zero gravity, v_n(t) = v0 / (1 + k v0 t),
    measured on the far half before the root's constraint wave arrives."""
    n, KN, V0 = 33, 0.5, 10.0
    o = make(solver, rope_line(n, 0.02, (1, 0, 0)), EA=1.0)
    o.set_gravity([0, 0, 0])
    o.drag_quad = KN
    o.set_state("velocities", np.tile([0.0, V0, 0.0], (n, 1)))
    o.set_state("velocities", np.tile([0.0, V0, 0.0], (n, 1)))
    S = max(1, int(np.ceil(np.float32(1 / 60) / np.float32(1 / 3000))))
    for _ in range(3):
        o.step_free(1 / 60, S)
    v = o.velocities()[16:]
    np.testing.assert_allclose(v[:, 1], V0 / (1 + KN * V0 * 3 / 60), rtol=1e-2)
    assert np.abs(v[:, [0, 2]]).max() < 0.05


def free_rope(solver, gravity, plane_w, mu_s, mu_k, R=0.00635, H=1 / 6000, ITR=16):
    """tests/test_contact.py:43: 21 x 5 cm rope (rho 0.117, EA 11473, EI 4.48e-3) free on a plane."""
    pos = rope_line(21, 0.05, (1, 0, 0))
    o = solver.OracleBatch([pos], 0.117, 11473.0, 4.48e-3, H, ITR, pinned=np.zeros(21, bool), prec=64)
    o.damping = 0.0
    o.drag = 0.0
    o.set_gravity(gravity)
    o.set_planes(np.array([[0.0, 0.0, 1.0, plane_w]]), radius=R, mu_s=mu_s, mu_k=mu_k, v_slip=1e-3, kappa=2.0)
    return o


def test_plane_support_equals_weight_and_sink(solver):
    """tests/test_contact.py:55: sum of penalty forces = weight (abs 5e-3), sink g h^2 / kappa^2 (rel 5%)."""
    R, H = 0.00635, 1 / 6000
    w = -(R + 5e-4)
    o = free_rope(solver, (0, 0, -G), w, 0.5, 0.5)
    m = 1.0 / o.inv_masses()
    for _ in range(60):
        o.step_free(1 / 60, 100)
    z = o.positions()[:, 2]
    pen = np.maximum(-(z - w - R), 0)
    k = m * (2.0 / H) ** 2
    assert (k * pen).sum() / (m.sum() * G) == pytest.approx(1.0, abs=5e-3)
    assert pen.max() == pytest.approx(G * H * H / 4.0, rel=0.05)
    assert np.abs(o.velocities()).max() < 1e-3


@pytest.mark.parametrize("mu,slides", [(0.5, False), (0.2, True)])
def test_plane_incline_coulomb(solver, mu, slides):
    """tests/test_contact.py:68: on a 20 degree incline the rope sticks for mu > tan 20 and slides with
    a = g (sin - mu cos) otherwise (rel 1%)."""
    R, H = 0.00635, 1 / 6000
    th = np.radians(20.0)
    pen0 = G * H * H / 4.0
    o = free_rope(solver, (G * np.sin(th), 0, -G * np.cos(th)), -(R - pen0 * np.cos(th)), mu, mu)
    m = 1.0 / o.inv_masses()
    ts, vx = [], []
    for f in range(72):
        o.step_free(1 / 60, 100)
        ts.append((f + 1) / 60)
        vx.append((m[:, None] * o.velocities()).sum(0)[0] / m.sum())
    ts, vx = np.array(ts), np.array(vx)
    a = np.polyfit(ts[ts > 0.3], vx[ts > 0.3], 1)[0]
    if slides:
        assert a == pytest.approx(G * (np.sin(th) - mu * np.cos(th)), rel=0.01)
    else:
        assert abs(a) < 1e-3 and abs(vx[-1]) < 2e-3


def test_box_sdf_floor_equals_plane(solver):
    """tests/test_mesh_contact.py:70: a box SDF whose top face is z = 0 acts exactly like the plane z = 0."""
    from ropesim import sdf
    R = 0.00635
    g = sdf.from_function(sdf.sdf_box((3.0, 1.0, 0.125), (0, 0, -0.125)), (-0.3125, -0.3125, -0.1875),
                          (1.3125, 0.3125, 0.1875), 1 / 128)
    pos = rope_line(21, 0.05, (1, 0, 0), (0, 0, R + 0.002))

    def run(kind):
        o = solver.OracleBatch([pos], 0.117, 11473.0, 4.48e-3, 1 / 6000, 16, pinned=np.zeros(21, bool), prec=64)
        o.damping = 0.0
        o.drag = 0.0
        o.set_gravity([0, 0, -G])
        if kind == "plane":
            o.set_planes(np.array([[0, 0, 1.0, 0.0]]), radius=R, mu_s=0.5, mu_k=0.5)
        else:
            o.set_contact(R, 0.5, 0.5)
            o.add_mesh(g.values, g.origin, g.spacing, mu_s=0.5, mu_k=0.5)
        for _ in range(20):
            o.step_free(1 / 60, 100)
        return o.positions()
    assert np.abs(run("plane") - run("box")).max() < (1e-5 if fp32(solver) else 1e-6)


def test_trilinear_sample_matches_closed_form_on_planar_regions():
    """ropesim.sdf sample of a half-space (linear field) is exact anywhere inside the lattice."""
    from ropesim import sdf
    n = np.array([0.3, -0.5, 0.8])
    n /= np.linalg.norm(n)
    g = sdf.from_function(lambda p: p @ n - 0.1, (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 1 / 32)
    rng = np.random.default_rng(1)
    pts = rng.uniform(-0.45, 0.45, (500, 3))
    phi, inside = g.sample(pts)
    assert inside.all()
    np.testing.assert_allclose(phi, pts @ n - 0.1, atol=2e-6)


def _seg_dist_exact(p0, p1, q0, q1):
    """Closest distance between segments by dense sampling refined with a fine local search (float64)."""
    t = np.linspace(0, 1, 401)
    A = p0 + (p1 - p0) * t[:, None]
    B = q0 + (q1 - q0) * t[:, None]
    d = np.linalg.norm(A[:, None] - B[None], axis=2)
    i, j = np.unravel_index(d.argmin(), d.shape)
    best = d[i, j]
    for _ in range(3):
        ta = np.clip(t[i] + np.linspace(-0.005, 0.005, 101), 0, 1)
        tb = np.clip(t[j] + np.linspace(-0.005, 0.005, 101), 0, 1)
        A = p0 + (p1 - p0) * ta[:, None]
        B = q0 + (q1 - q0) * tb[:, None]
        d = np.linalg.norm(A[:, None] - B[None], axis=2)
        best = min(best, d.min())
    return best


def test_rope_rope_detection_matches_brute_force(oracle):
    """Candidate lists (csrc/batch.cu:19-49) against an independent float64 segment distance: every pair
    clearly inside the padded radii is listed, every pair clearly outside is not, lists ascend, same-rope
    neighbours |i - j| <= 2 and other groups are excluded."""
    rng = np.random.default_rng(5)
    ropes = []
    for r in range(4):
        c = rng.uniform(-0.004, 0.004, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        ropes.append((c + (np.arange(20)[:, None] * 0.015 - 0.15) * d).astype(np.float32))
    o = oracle.OracleBatch(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, prec=64)
    o.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 2)
    x = np.concatenate(ropes).astype(np.float64)
    pads = rng.uniform(0.004, 0.01, len(x))
    counts, lists, lim = o.detect(x, pads)
    off = o.offsets
    rope_of = np.repeat(np.arange(4), np.diff(off))
    seen = 0
    for i in range(len(x)):
        if i + 1 >= off[rope_of[i] + 1]:
            assert counts[i] == 0
            continue
        got = list(lists[i, : counts[i]])
        assert got == sorted(got)
        ri = max(pads[i], pads[i + 1])
        for j in range(len(x)):
            if j + 1 >= off[rope_of[j] + 1] or rope_of[j] // 2 != rope_of[i] // 2:
                assert j not in got
                continue
            if rope_of[j] == rope_of[i] and abs(j - i) <= 2:
                assert j not in got
                continue
            r = ri + max(pads[j], pads[j + 1])
            dist = _seg_dist_exact(x[i], x[i + 1], x[j], x[j + 1])
            if dist < r * (1 - 1e-4):
                assert j in got, (i, j, dist, r)
                seen += 1
            elif dist > r * (1 + 1e-4):
                assert j not in got, (i, j, dist, r)
        # step limit: min(pad - detR, 0.45 * distance to each listed segment)
        detR = 0.0036 + 0.5 * 0.0008
        ds = [_seg_dist_exact(x[i], x[i + 1], x[j], x[j + 1]) for j in got]
        assert lim[i] == pytest.approx(min([pads[i] - detR] + [0.45 * d for d in ds]), rel=1e-4)
    assert seen > 20


def test_mesh_broadphase_matches_enumeration(oracle):
    """sdfBuildCandidates (YB/sim/contact.cuh:121-136): ascending mesh ids whose posed lattice AABB padded
    by 2|dx| + 2 cm holds the point; more than 8 -> -1; pinned -> 0."""
    from ropesim import sdf
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(2)
    pos = rope_line(40, 0.05, (1, 0, 0), (-1, 0, 0.5))
    o = oracle.OracleBatch([pos], 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, prec=64)
    o.set_contact(0.005, 0.5, 0.5)
    boxes = []
    for m in range(11):
        g = sdf.from_function(sdf.sdf_sphere(0.1), (-0.15,) * 3, (0.15,) * 3, 0.05)
        p = rng.uniform(-1.0, 1.0, 3)
        q = Rotation.random(random_state=m).as_quat()
        o.add_mesh(g.values, g.origin, g.spacing, pos=p, quat=q)
        hi_o = g.origin.astype(np.float64) + (np.array(g.values.shape) - 1) * np.float64(np.float32(g.spacing))
        corners = np.array([[hi_o[0] if c & 1 else g.origin[0], hi_o[1] if c & 2 else g.origin[1],
                             hi_o[2] if c & 4 else g.origin[2]] for c in range(8)])
        w = Rotation.from_quat(q / np.linalg.norm(q)).apply(corners) + p
        boxes.append((w.min(0), w.max(0)))
    x = rng.uniform(-1.2, 1.2, (40, 3))
    dx = rng.normal(0, 0.01, (40, 3))
    counts, cand = o.broadphase(x, dx)
    assert counts[0] == 0  # pinned root
    for i in range(1, 40):
        pad = 2 * np.linalg.norm(dx[i]) + 0.02
        want = [m for m, (lo, hi) in enumerate(boxes) if np.all(x[i] >= lo - pad - 1e-9) and np.all(x[i] <= hi + pad + 1e-9)]
        sure = [m for m, (lo, hi) in enumerate(boxes) if np.all(x[i] >= lo - pad + 1e-6) and np.all(x[i] <= hi + pad - 1e-6)]
        if len(want) > 8:
            assert counts[i] == -1
        else:
            got = list(cand[i, : counts[i]])
            assert set(sure) <= set(got) <= set(want) and got == sorted(got)
