"""Quadratic drag and stretch damping pinned against closed forms, on the CPU oracle and the HIP path.

The reference's endItr (YB/sim/iteration.cu:85-112) splits each free vertex's velocity by its segment
tangent t (the tip vertex uses its one real segment, q_{i-1}) and gives each component the exact
frozen-direction solution of dv/dt = -k |v| v over one substep:
    v_n /= 1 + drag_quad   |v_n| h,    v_t /= 1 + drag_quad_t |v_t| h,
the whole block gated on drag_quad > 0. The checks restate the reference's own (tests/test_drag_quad.py:56-177
of the reference): a straight low-EA rope translating in zero gravity, measured on its far half before the
constraint wave from the pinned root arrives, follows v(t) = v0 / (1 + k v0 t) in a damped component and
keeps v0 in an undamped one.

The stretch damping of cosseratItr (YB/sim/cosserat.cu:66-88) adds -(damping / h) (EA / l) dx_i to each
end of every segment and (EA / l)(damping / h) to its Hessian. A free rope translating rigidly keeps every
strain at zero, so each vertex solves (dx - y) / (h^2 w) = -n (damping / h)(EA / l) dx with n the number of
its segments; with lumped masses (w = 1 / (rho l) inside, 2 / (rho l) at the ends) every vertex then decays
by the same factor per substep,
    v <- v / (1 + 2 damping EA h / (rho l^2)),
and a dropped term, a wrong sign or a missing segment end breaks either the rate or the uniformity.
"""
import numpy as np
import pytest

from conftest import fp32

N = 33
FAR = 16
KN = 0.5
V0 = 10.0
H = 1 / 60
NFRAMES = 3
T = NFRAMES * H


def line(n, seg, direction):
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    return (seg * np.arange(n)[:, None] * d).astype(np.float32)


def substeps(max_h):
    """RopeSim.advance(H) without explicit substeps: ceil(H / max_h) in float32 (csrc/bindings.cpp:276)."""
    return max(1, int(np.ceil(np.float32(H) / np.float32(max_h))))


def decay_rope(solver, max_h=1 / 3000, EA=1.0, pinned=None):
    """Straight +x rope of 32 x 2 cm, rho 0.1, low EA (slow constraint wave), zero gravity, no linear drag."""
    o = solver.OracleBatch([line(N, 0.02, (1, 0, 0))], 0.1, EA, 1e-6, max_h, 8, pinned=pinned, prec=64)
    o.drag = 0.0
    o.damping = 0.0
    o.set_gravity([0.0, 0.0, 0.0])
    return o


def kick_and_run(o, v0, max_h=1 / 3000, frames=NFRAMES):
    o.set_state("velocities", np.tile(np.asarray(v0, np.float64), (o.nv, 1)))
    S = substeps(max_h)
    for _ in range(frames):
        o.step_free(H, S)
    return o.velocities(), S


def closed(v0, k, t):
    return v0 / (1.0 + k * v0 * t) if k > 0 else v0


def test_axial_motion_undamped_when_kt_zero(solver):
    """v along the rope with drag_quad_t = 0: the normal term sees |v_n| = 0 and nothing decays."""
    o = decay_rope(solver)
    o.drag_quad = KN
    v, _ = kick_and_run(o, (V0, 0, 0))
    np.testing.assert_allclose(v[FAR:, 0], V0, rtol=2e-3)


def test_components_decay_separately_at_45_degrees(solver):
    """kt = kn, v at 45 degrees: each component follows its own closed form (not the isotropic one)."""
    o = decay_rope(solver)
    o.drag_quad = KN
    o.drag_quad_t = KN
    c = V0 / np.sqrt(2.0)
    v, _ = kick_and_run(o, (c, c, 0))
    want = closed(c, KN, T)
    np.testing.assert_allclose(v[FAR:, 0], want, rtol=1e-2)
    np.testing.assert_allclose(v[FAR:, 1], want, rtol=1e-2)
    # the isotropic law |v| / (1 + k |v| t) would give a different component speed
    assert abs(closed(V0, KN, T) / np.sqrt(2.0) - want) / want > 0.05


def test_tangential_coefficient_acts_on_axial_motion(solver):
    """kn > 0 and kt > 0 with v along the rope: the tangential closed form, not the normal one."""
    o = decay_rope(solver)
    o.drag_quad = KN
    o.drag_quad_t = 0.2
    v, _ = kick_and_run(o, (V0, 0, 0))
    np.testing.assert_allclose(v[FAR:, 0], closed(V0, 0.2, T), rtol=1e-2)
    assert np.abs(v[FAR:, 1:]).max() < 0.05


def test_decay_is_substep_invariant(solver):
    """The per-substep law is exact for a frozen direction: 25 and 100 substeps per frame agree."""
    vy = []
    for max_h in (1 / 1500, 1 / 6000):
        o = decay_rope(solver, max_h=max_h)
        o.drag_quad = KN
        v, _ = kick_and_run(o, (0, V0, 0), max_h=max_h)
        vy.append(float(np.mean(v[FAR:, 1])))
    assert abs(vy[0] - vy[1]) / vy[1] < 5e-3, vy


def test_tangential_coefficient_inert_without_normal(solver):
    """drag_quad_t alone is inert: the quadratic block is gated on drag_quad > 0 (iteration.cu:90)."""
    o = decay_rope(solver)
    o.drag_quad_t = KN
    v, _ = kick_and_run(o, (0, V0, 0))
    np.testing.assert_allclose(v[FAR:, 1], V0, rtol=2e-3)


def test_tip_uses_previous_segment_tangent(solver):
    """The tip vertex has no segment of its own; it takes the tangent of segment n-2 (q_{i-1}). Built along +x,
    the rope is first let hang along -y (its tip frame rotates by ~90 degrees; the padding frame at the last
    vertex keeps its construction value +x), then kicked along the real tip tangent with a large kn: the
    correct tangent leaves the tip undamped, the padding frame would read the motion as normal and damp it
    by about half in one frame."""
    o = solver.OracleBatch([line(N, 0.02, (1, 0, 0))], 0.1, 500.0, 1e-6, 1 / 3000, 8, prec=64)
    o.damping = 0.0
    o.set_gravity([0.0, -9.8, 0.0])
    o.drag = 3.0
    S = substeps(1 / 3000)
    for _ in range(240):
        o.step_free(H, S)
    p = o.positions()
    t_tip = p[-1] - p[-2]
    t_tip /= np.linalg.norm(t_tip)
    assert t_tip[1] < -0.9, t_tip
    o.drag = 0.0
    o.set_gravity([0.0, 0.0, 0.0])
    o.drag_quad = 5.0
    o.set_state("velocities", np.tile(V0 * t_tip, (o.nv, 1)))
    o.step_free(H, S)
    tip = np.linalg.norm(o.velocities()[-1])
    assert abs(tip - V0) / V0 < 0.05, tip
    # the same kick across the rope decays to about V0 / (1 + kn V0 H) = 0.55 V0: the tangent the tip takes is
    # what keeps the axial kick undamped
    n_tip = np.cross(t_tip, [0.0, 0.0, 1.0])
    o.set_state("velocities", np.tile(V0 * n_tip / np.linalg.norm(n_tip), (o.nv, 1)))
    o.step_free(H, S)
    assert np.linalg.norm(o.velocities()[-1]) < 0.6 * V0


@pytest.mark.parametrize("prec", [64, 32])
def test_stretch_damping_of_rigid_translation(solver, prec):
    """Free rope (no pinned vertex) translating rigidly with damping 1e-4: every vertex decays as
    v0 / (1 + 2 damping EA h / (rho l^2))^n after n substeps."""
    if getattr(solver, "FP32", False) and prec == 64:
        pytest.skip("the device computes in float32 (covered by prec=32)")
    beta, EA, rho, l = 1e-4, 1.0, 0.1, 0.02
    o = solver.OracleBatch([line(N, l, (1, 0, 0))], rho, EA, 1e-6, 1 / 3000, 8, pinned=np.zeros(N, bool), prec=prec)
    o.drag = 0.0
    o.damping = beta
    o.set_gravity([0.0, 0.0, 0.0])
    v0 = np.array([0.3, 0.4, -0.2])
    v, S = kick_and_run(o, v0)
    h = H / S
    a = 2.0 * beta * EA * h / (rho * l * l)
    want = v0 / (1.0 + a) ** (S * NFRAMES)
    assert want[0] < 0.8 * v0[0]  # the decay is large enough to see a wrong rate
    # float32: 153 substeps of rounding; float64: the float32 initial positions leave strains of ~1e-8. A wrong
    # rate is off by more than 1e-2.
    rtol = 2e-4 if fp32(solver, prec) else 2e-6
    np.testing.assert_allclose(v, np.tile(want, (N, 1)), rtol=rtol, atol=rtol * np.abs(v0).max())
