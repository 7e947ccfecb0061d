"""The host-driven RopeSim surface and the RopeBatch bookkeeping calls on the device, against the oracle.

Reference behaviour covered (csrc/bindings.cpp of the reference):
- RopeSim.advance(dt, substeps=0) runs ceil(dt / max_h) substeps in float32 (51 for 1/60 over 1/3000);
  advance(dt, substeps=k) runs exactly k of dt / k (:270-295; tests/test_api.py:84-99);
- RopeSim.set_root(pos, vel) places the pinned root and the solver sweeps it by h v per substep
  (:318-350; tests/test_driving.py:15-46);
- RopeSim.reset restores the construction state, zero velocities, time 0 and the first frame's previous
  substep length (:303-316);
- RopeBatch.reset_envs(ids) restores only those environments (csrc/batch.cu:178-189, 400-413;
  tests/test_batch.py:119);
- a rope-rope candidate list longer than 128 raises RuntimeError and leaves last_error set, and the device
  error word is cleared for the next step (csrc/batch.cu:47; YB/YarnBall.cu:311-339).
A RopeSim drives its root once per frame, which is the oracle's constant-velocity drive (csrc/batch.cu:90-104)
of the same root position and velocity.
"""
import numpy as np
import pytest

from conftest import host, to_dev

pytestmark = pytest.mark.gpu

H = 1 / 60
ARITH = [pytest.param(True, id="fast"), pytest.param(False, id="ieee")]


def line(n, seg, direction):
    d = np.asarray(direction, np.float64)
    d /= np.linalg.norm(d)
    return (seg * np.arange(n)[:, None] * d).astype(np.float32)


def make_sim(rs, fast, n=25, seg=0.02, **kw):
    r = rs.RopeSim(num_verts=n, seg_len=seg, dir=np.array([0.0, -1.0, 0.0], np.float32), density=0.1,
                   k_stretch=500.0, k_bend=1e-6, max_h=1 / 3000, **kw)
    r.fast_math = fast
    r.set_gravity(np.array([0.0, -9.8, 0.0], np.float32))
    r.drag = 0.0
    r.damping = 0.0
    return r


def make_oracle(oracle, n=25, seg=0.02, prec=64):
    o = oracle.OracleBatch([line(n, seg, (0, -1, 0))], 0.1, 500.0, 1e-6, 1 / 3000, 8, prec=prec)
    o.set_gravity([0.0, -9.8, 0.0])
    o.drag = 0.0
    o.damping = 0.0
    return o


def traj(t):
    return np.array([0.1 * np.sin(2 * np.pi * t), 0.0, 0.05 * np.sin(2 * np.pi * 1.7 * t)])


@pytest.mark.parametrize("fast", ARITH)
def test_ropesim_advance_implicit_and_explicit_substeps_vs_oracle(rs, oracle, fast):
    """A swept root over 12 frames: advance(H) (51 substeps) for six frames, then advance(H, substeps=50); every
    frame within 1e-5 of rope length of the float64 oracle driven by the same root positions and velocities."""
    n, seg = 25, 0.02
    L = (n - 1) * seg
    r = make_sim(rs, fast, n, seg)
    o = make_oracle(oracle, n, seg)
    assert int(np.ceil(np.float32(H) / np.float32(r.max_h))) == 51
    worst = 0.0
    for k in range(12):
        t = k * H
        p, v = traj(t), (traj(t + H) - traj(t)) / H
        r.set_root(p.astype(np.float32), v.astype(np.float32))
        explicit = k >= 6
        if explicit:
            r.advance(H, substeps=50)
        else:
            r.advance(H)
        o.step(p[None], v[None], H, 50 if explicit else 51)
        x = np.asarray(r.positions(), np.float64)
        e = np.abs(x - o.positions()).max() / L
        worst = max(worst, e)
        assert e < 1e-5, f"frame {k}: {e:.2e} of rope length"
        # the root is swept exactly: p + H v
        np.testing.assert_allclose(x[0], p + H * v, atol=2e-6)
    assert r.time == pytest.approx(12 * H, rel=1e-6)
    assert abs(r.max_h - 1 / 3000) < 1e-9  # explicit substeps leave max_h as it was
    print(f"\n  RopeSim vs float64 oracle, fast {fast}: worst per-frame error {worst:.2e} of rope length")


def test_ropesim_reset_equals_fresh_object(rs, oracle):
    """After frames with 51 substeps (previous substep 1/3060, not max_h), reset() gives back the fresh object's
    trajectory bit for bit: positions, frames, zero velocities, time 0 and the first frame's previous substep."""
    def run(sim, frames, t0=0):
        for k in range(frames):
            t = (t0 + k) * H
            sim.set_root(traj(t).astype(np.float32), ((traj(t + H) - traj(t)) / H).astype(np.float32))
            sim.advance(H)
        return [np.array(sim.positions()), np.array(sim.velocities()), np.array(sim.quaternions())]

    a = make_sim(rs, True)
    init = [np.array(a.positions()), np.array(a.quaternions())]
    run(a, 7, t0=3)
    a.reset()
    assert a.time == 0.0 and a.last_error == 0
    assert np.array_equal(np.array(a.positions()), init[0]) and np.array_equal(np.array(a.quaternions()), init[1])
    assert not np.array(a.velocities()).any()
    got = run(a, 5)
    want = run(make_sim(rs, True), 5)
    for x, y in zip(got, want):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # and the reset object still follows the oracle
    o = make_oracle(oracle)
    for k in range(5):
        t = k * H
        o.step(traj(t)[None], ((traj(t + H) - traj(t)) / H)[None], H, 51)
    assert np.abs(got[0] - o.positions()).max() / (24 * 0.02) < 1e-5


def test_ropesim_set_root_frame_and_errors(rs):
    """set_root(q) drives the root frame of a fixed-orientation rope; q on a free-pivot root, a zero q and
    negative substeps are ValueErrors (tests/test_driving.py, test_api.py of the reference)."""
    r = make_sim(rs, True, n=8, fix_root_orientation=True)
    th = np.deg2rad(30.0)
    q = np.array([0.0, 0.0, np.sin(th / 2), np.cos(th / 2)], np.float32)
    r.set_root(np.zeros(3, np.float32), np.zeros(3, np.float32), q=q)
    r.advance(H, substeps=50)
    np.testing.assert_allclose(np.array(r.quaternions())[0], q, atol=1e-6)
    with pytest.raises(ValueError):
        r.set_root(np.zeros(3, np.float32), np.zeros(3, np.float32), q=np.zeros(4, np.float32))
    free = make_sim(rs, True, n=8)
    with pytest.raises(ValueError):
        free.set_root(np.zeros(3, np.float32), np.zeros(3, np.float32), q=np.array([0, 0, 0, 1], np.float32))
    with pytest.raises(ValueError):
        free.advance(H, substeps=-1)


def _batch(rs, N=4, V=25, mode=1, fast=False):
    roots = np.zeros((N, 3), np.float32)
    roots[:, 0] = np.arange(N) * 1.5
    b = rs.RopeBatch(num_envs=N, verts_per_rope=V, root_pos=roots, dir=np.array([0, -1, 0], np.float32),
                     fix_root_orientation=False)
    b.set_gravity(np.array([0, -9.8, 0], np.float32))
    b.drag = 0.0
    b.damping = 0.0
    b.set_dispatch(mode)
    b.fast_math = fast
    return b, roots


@pytest.mark.parametrize("mode", [0, 1])
def test_reset_envs_subset(rs, oracle, mode):
    """reset_envs([1, 3]) after 30 driven frames: envs 1 and 3 return to the construction state with zero
    velocities bit for bit, envs 0 and 2 are untouched bit for bit, and the following frames match the float32
    oracle that ran the same sequence (IEEE arithmetic: the oracle's operations)."""
    N, V = 4, 25
    b, roots = _batch(rs, N, V, mode)
    init = (host(b.positions()).copy(), host(b.quaternions()).copy())
    o = oracle.OracleBatch(init[0].reshape(N, V, 3), 0.1, 500.0, 1e-6, 1 / 3000, 8, prec=32)
    o.set_gravity([0, -9.8, 0])
    o.drag = 0.0
    o.damping = 0.0
    vel = np.tile(np.array([0.2, 0.0, 0.0], np.float32), (N, 1))
    for k in range(30):
        pos = roots + np.array([0.2 * k * H, 0.0, 0.0], np.float32)
        b.step(to_dev(pos), to_dev(vel), H, 50)
        o.step(pos, vel, H, 50)
    before = [host(g()).copy() for g in (b.positions, b.velocities, b.quaternions)]
    b.reset_envs(np.array([1, 3], np.int32))
    o.reset_envs([1, 3])
    after = [host(g()) for g in (b.positions, b.velocities, b.quaternions)]
    bits = lambda a: a.view(np.uint32)
    assert np.array_equal(bits(after[0][[1, 3]]), bits(init[0][[1, 3]]))
    assert np.array_equal(bits(after[2][[1, 3]]), bits(init[1][[1, 3]]))
    assert not after[1][[1, 3]].any()
    for x, y in zip(before, after):
        assert np.array_equal(bits(x[[0, 2]]), bits(y[[0, 2]]))
    for k in range(30, 35):
        pos = roots + np.array([0.2 * k * H, 0.0, 0.0], np.float32)
        b.step(to_dev(pos), to_dev(vel), H, 50)
        o.step(pos, vel, H, 50)
    d = np.abs(host(b.positions()).reshape(-1, 3) - o.positions()).max()
    assert d < 1e-6, d


def _parallel_pile(n_side=4, V=40, seg=0.005, gap=0.002):
    """n_side x n_side parallel straight ropes along x, `gap` apart: every segment lies within a few
    centimetres of hundreds of others, and no two are closer than `gap`."""
    ropes = []
    for a in range(n_side):
        for c in range(n_side):
            p = np.zeros((V, 3), np.float32)
            p[:, 0] = seg * np.arange(V)
            p[:, 1] = gap * a
            p[:, 2] = 0.5 + gap * c
            ropes.append(p)
    return ropes


@pytest.mark.parametrize("mode", [0, 1])
def test_collision_list_overflow_raises_and_clears(rs, mode):
    """A detection radius of 2 cm (scaler 40 on the yarn defaults) over a pile of 16 parallel ropes 2 mm apart
    gives segments more than 128 candidates: the step raises RuntimeError, last_error is 1
    (ERROR_MAX_COLLISIONS_PER_SEGMENT_EXCEEDED), and the error word is cleared, so the next step without the
    overflow runs. The barrier acts below 2 r + b = 1 mm only, so the pile sits still under zero gravity."""
    ropes = _parallel_pile()
    nv = sum(len(r) for r in ropes)
    b = rs.RopeBatch.from_positions(ropes, 0.1, 500.0, 1e-4, 1 / 6000, 8, pinned=np.zeros(nv, bool))
    b.set_gravity(np.zeros(3, np.float32))
    b.set_dispatch(mode)
    b.set_rope_contact(True, 1e-4, 8e-4, 1e-5, 0.1, 5.0, 40.0, len(ropes))
    assert b.last_error == 0
    with pytest.raises(RuntimeError, match="MAX_COLLISIONS"):
        b.step_free(H, 10)
    assert b.last_error == 1
    b.set_rope_contact(True, 1e-4, 8e-4, 1e-5, 0.1, 5.0, 1.2, len(ropes))
    b.step_free(H, 10)  # no overflow at the default scaler: the word was cleared
    assert b.last_error == 1  # the code itself stays, as the reference's lastErrorCode
    x = host(b.positions()).reshape(-1, 3)
    assert np.isfinite(x).all() and np.abs(x - np.concatenate(ropes)).max() < 1e-4
