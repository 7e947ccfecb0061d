"""Contact-force readout on the HIP path (reference YB/sim/contact.cuh:172-204, csrc/sdf.cu:89-92, 239-270).

Scene: uneven free ropes, one per environment, dropped onto a slab mesh and draped over a sphere mesh that
pokes up through it; the slab translates and turns a little every frame (moving-mesh branch, the omega x r
term of the friction), and the sphere is static. Every wave of the device holds vertices of several ropes
touching both meshes, so the wave-level wrench sums see mixed (rope, mesh) keys.

Checks:
- per-vertex contact forces and per-(rope, mesh) wrenches after every frame, through the impact: IEEE
  arithmetic against the float32 oracle to summation-order rounding; fast arithmetic against the float64
  oracle to twice the float32 oracle's own gap;
- fixed-point summation (enable_contact_forces(deterministic=True)) gives the same wrenches bit for bit in
  the per-substep kernels, in every persistent-kernel variant, with and without the captured graph, and
  in a fresh batch; float summation agrees with it to rounding;
- the resting total: the wrenches' vertical force equals the ropes' weight once they settle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R, RHO, EA, EI = 0.004, 0.1, 500.0, 1e-4
H, ITR, DT, SUB = 1 / 3000, 8, 1 / 60, 40
G = 9.81


def scene(n=10, seed=3):
    from ropesim import sdf
    rng = np.random.default_rng(seed)
    Vs = rng.integers(20, 71, n)
    ropes = []
    for r in range(n):
        x0 = rng.uniform(-0.45, -0.3)
        p = np.zeros((Vs[r], 3))
        p[:, 0] = x0 + 0.012 * np.arange(Vs[r])
        p[:, 1] = rng.uniform(-0.05, 0.05)
        p[:, 2] = 0.075
        ropes.append(p)
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum(Vs)
    pos = np.concatenate(ropes).astype(np.float32)
    slab = sdf.from_function(sdf.sdf_box((1.0, 1.0, 0.05), (0, 0, -0.05)), (-0.6, -0.6, -0.12), (0.9, 0.6, 0.03),
                             1 / 64)
    ball = sdf.from_function(sdf.sdf_sphere(0.06), (-0.08, -0.08, -0.08), (0.08, 0.08, 0.08), 1 / 200)
    ball_pos = np.column_stack([rng.uniform(-0.1, 0.1, n), pos[off[:-1], 1], np.full(n, 0.005)]).astype(np.float32)
    return pos, off, slab, ball, ball_pos, rng.uniform(-1, 1, n)


def slab_pose(frame, spin):
    """Slab target at the end of `frame`: drifts along y and turns about z, per environment."""
    t = (frame + 1) * DT
    n = len(spin)
    pos = np.zeros((n, 3), np.float32)
    pos[:, 1] = 0.02 * t * spin
    th = 0.3 * t * spin
    q = np.zeros((n, 4), np.float32)
    q[:, 2] = np.sin(th / 2)
    q[:, 3] = np.cos(th / 2)
    return pos, q


def make(kind, rs=None, oracle=None, fast=True, mode=1, variant=-1, graph=False, deterministic=False):
    pos, off, slab, ball, ball_pos, spin = scene()
    n = len(off) - 1
    pin = np.zeros(len(pos), bool)
    if kind == "oracle":
        b = oracle.OracleBatch(pos, RHO, EA, EI, H, ITR, pinned=pin, prec=32, offsets=off)
    else:
        b = rs.RopeBatch.from_positions(pos, RHO, EA, EI, H, ITR, pinned=pin, offsets=off)
        b.fast_math = fast
        b.set_dispatch(mode, variant)
        b.set_graph(graph)
    b.drag = 0.0
    b.damping = 0.0
    b.set_gravity(np.array([0, 0, -G], np.float32))
    b.set_contact(R, mu_s=0.5, mu_k=0.4, v_slip=1e-3, kappa=2.0)
    zero = np.zeros((n, 3), np.float32)
    ident = np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1))
    b.add_mesh(slab.values, slab.origin, slab.spacing, pos=zero, quat=ident, mu_s=0.5, mu_k=0.4)
    b.add_mesh(ball.values, ball.origin, ball.spacing, pos=ball_pos, quat=ident, mu_s=0.3, mu_k=0.2)
    if kind == "oracle":
        b.enable_contact_forces(True)
    else:
        b.enable_contact_forces(True, deterministic=deterministic)
    return b, spin


def frame(b, f, spin):
    p, q = slab_pose(f, spin)
    b.set_mesh_poses(0, p, q)
    b.step_free(DT, SUB)


def readout(b, n_meshes=2):
    F = np.asarray(b.contact_forces(), np.float64).reshape(-1, 3)
    W = np.asarray(b.mesh_wrenches(), np.float64).reshape(-1, n_meshes, 6)
    return F, W


def weights():
    pos, off, *_ = scene()
    seg = np.linalg.norm(np.diff(pos.astype(np.float64), axis=0), axis=1)
    M = np.array([RHO * seg[off[r]:off[r + 1] - 1].sum() for r in range(len(off) - 1)])
    return M * G


class _Prec64:
    def __init__(self, oracle):
        self.OracleBatch = lambda *a, **k: oracle.OracleBatch(*a, **{**k, "prec": 64})


@pytest.mark.parametrize("deterministic", [False, True], ids=["float", "fixed"])
@pytest.mark.parametrize("fast", [pytest.param(True, id="fast"), pytest.param(False, id="ieee")])
def test_readout_matches_oracle(rs, oracle, fast, deterministic):
    """IEEE: within 1e-5 of the largest rope weight of the float32 oracle (the same arithmetic; only the wrench
    summation order differs). Fast: no further from the float64 oracle than twice the float32 oracle's own gap
    (at least 1e-4 of the weight): a stiff penalty turns rounding-level position differences into percent-level
    force differences, 1e-2 of the weight in float32 arithmetic itself on this scene."""
    o32, spin = make("oracle", oracle=oracle)
    o64, _ = make("oracle", oracle=_Prec64(oracle))
    b, _ = make("hip", rs=rs, fast=fast, deterministic=deterministic)
    assert b.deterministic_forces == deterministic
    s = weights().max()
    touched = np.zeros(2, bool)
    for f in range(8):
        for x in (o32, o64, b):
            frame(x, f, spin)
        F32, W32 = readout(o32)
        F64, W64 = readout(o64)
        Fb, Wb = readout(b)
        touched |= np.abs(W32[:, :, 2]).max(0) > 0.05 * s
        if not fast:
            assert np.abs(Wb - W32).max() < 1e-5 * s, (f, np.abs(Wb - W32).max() / s)
            assert np.abs(Fb - F32).max() < 1e-5 * s, (f, np.abs(Fb - F32).max() / s)
        else:
            bar_w = 2 * max(np.abs(W32 - W64).max(), 1e-4 * s)
            bar_f = 2 * max(np.abs(F32 - F64).max(), 1e-4 * s)
            assert np.abs(Wb - W64).max() <= bar_w, (f, np.abs(Wb - W64).max() / s, bar_w / s)
            assert np.abs(Fb - F64).max() <= bar_f, (f, np.abs(Fb - F64).max() / s, bar_f / s)
    assert touched.all(), "both meshes carry load"


def _wrenches(rs, frames=6, **kw):
    b, spin = make("hip", rs=rs, **kw)
    out = []
    for f in range(frames):
        frame(b, f, spin)
        out.append(readout(b)[1])
    return np.stack(out)


@pytest.mark.parametrize("fast", [pytest.param(True, id="fast"), pytest.param(False, id="ieee")])
def test_fixed_point_wrenches_reproduce_bitwise(rs, fast):
    ref = _wrenches(rs, fast=fast, mode=1, deterministic=True)
    assert np.abs(ref).max() > 0
    runs = {"per-substep kernels": dict(mode=0), "graph": dict(mode=1, graph=True), "fresh batch": dict(mode=1)}
    for v in range(rs.num_cell_variants()):
        blk, vpt, ok = rs.cell_variant_info(v)
        if ok and blk * vpt >= 70:
            runs[f"variant {blk}x{vpt}"] = dict(mode=1, variant=v)
    for name, kw in runs.items():
        w = _wrenches(rs, fast=fast, deterministic=True, **kw)
        assert np.array_equal(w, ref), (name, np.abs(w - ref).max())
    flt = _wrenches(rs, fast=fast, mode=1, deterministic=False)
    assert np.abs(flt - ref).max() <= 1e-5 * np.abs(ref).max()


def test_settled_wrenches_carry_the_weight(rs):
    b, spin = make("hip", rs=rs, deterministic=True)
    for f in range(60):
        frame(b, f, 0 * spin)  # slab at rest
    _, W = readout(b)
    Fz = W[:, :, 2].sum(1)
    assert np.allclose(Fz, -weights(), rtol=5e-3), (Fz, -weights())
