"""HIP path vs the CPU oracle on identical seeded inputs (parity tests proper), in both arithmetic modes
of the device ("fast": fused multiply-add and the hardware reciprocal / square root / log, the default;
"ieee": every operation correctly rounded and unfused, the float32 oracle's arithmetic).

Bars: positions within 1e-5 of rope length of the float64 oracle (per frame for the single-rope config,
after one frame for the batches). On ropes where float32 arithmetic itself (the float32 oracle, with and
without contracted multiply-adds, and under rounding-level perturbations of its input) strays past half of
that, the rope is held to twice that float32 floor. IEEE mode also within 1e-5 of the float32 oracle after one
frame. Rope-rope candidate lists, step limits and mesh broad-phase lists bit-exact given identical positions;
the per-substep kernels and the persistent frame kernel bit-identical within each mode.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import host, to_dev

pytestmark = pytest.mark.gpu

TOL_L = 1e-5  # of rope length
ARITH = [pytest.param(True, id="fast"), pytest.param(False, id="ieee")]


def _lib(rs):
    return C.CDLL(os.path.join(os.path.dirname(rs.__file__), "librope_hip.so"))


def build_pair(rs, oracle, scene, prec=64, mode=1, fast=True):
    b = scene.build(rs)
    b.set_dispatch(mode)
    b.fast_math = fast
    o = oracle.OracleBatch(scene.positions, scene.density, scene.k_stretch, scene.k_bend, scene.max_h, scene.num_itr,
                           pinned=scene.pinned, fix_root_orientation=scene.fix_root_orientation, prec=prec,
                           offsets=scene.offsets)
    scene.apply(o)
    return b, o


def rope_lengths(scene):
    p = scene.positions.astype(np.float64)
    L = np.zeros(scene.num_ropes)
    for r in range(scene.num_ropes):
        a, e = scene.offsets[r], scene.offsets[r + 1]
        L[r] = np.linalg.norm(np.diff(p[a:e], axis=0), axis=1).sum()
    return L


def rel_err(scene, x_gpu, x_ref):
    """Max over ropes of (max vertex position error / rope length)."""
    L = rope_lengths(scene)
    err = np.linalg.norm(x_gpu.reshape(-1, 3) - x_ref.reshape(-1, 3), axis=1)
    rope = np.repeat(np.arange(scene.num_ropes), np.diff(scene.offsets))
    per = np.zeros(scene.num_ropes)
    np.maximum.at(per, rope, err)
    return float((per / L).max())


def test_views_are_device_arrays_sharing_memory(rs):
    import torch
    from ropesim import scenes
    s = scenes.config2(n_ropes=2, V=10, frames=1)
    b = s.build(rs)
    v = b.positions()
    t = torch.as_tensor(v, device="cuda")
    assert t.is_cuda and t.dtype == torch.float32 and tuple(t.shape) == (2, 10, 3)
    assert t.data_ptr() == v.ptr
    t0 = t.clone()
    fi = s.frame_inputs(1)[0]
    s.step(b, fi, to_dev)
    torch.cuda.synchronize()
    assert not torch.equal(t, t0)
    q = torch.as_tensor(b.quaternions(), device="cuda")
    n = torch.linalg.vector_norm(q[:, :-1, :], dim=-1)
    assert torch.allclose(n, torch.ones_like(n), atol=1e-5)


@pytest.mark.parametrize("fast", ARITH)
@pytest.mark.parametrize("mode", [0, 1])
def test_config1_every_frame_within_1e5_of_length(rs, oracle, mode, fast):
    from ropesim import scenes
    s = scenes.config1(frames=60)
    b, o = build_pair(rs, oracle, s, 64, mode, fast)
    L = rope_lengths(s)[0]
    worst = 0.0
    for f, fi in enumerate(s.frame_inputs()):
        s.step(b, fi, to_dev)
        s.step(o, fi)
        e = np.abs(host(b.positions()).reshape(-1, 3) - o.positions()).max() / L
        worst = max(worst, e)
        assert e < TOL_L, f"frame {f}: {e:.3e} of rope length"
    print(f"\n  config1 mode {mode} fast {fast}: worst per-frame error {worst:.2e} of rope length")


def test_config1_fp32_oracle_tracks_device(rs, oracle):
    """IEEE mode runs the float32 oracle's operations in its order: the device follows it to rounding."""
    from ropesim import scenes
    s = scenes.config1(frames=10)
    b, o = build_pair(rs, oracle, s, 32, 1, fast=False)
    for fi in s.frame_inputs():
        s.step(b, fi, to_dev)
        s.step(o, fi)
    d = np.abs(host(b.positions()).reshape(-1, 3) - o.positions()).max()
    print(f"\n  config1 device vs float32 oracle after 10 frames: {d:.3e} m")
    assert d < 1e-6


def _run(rs, scene, mode, nframes, variant=-1, fast=True):
    b = scene.build(rs)
    try:
        b.set_dispatch(mode, variant)
    except ValueError:  # a variant that does not take this batch (rope length, contact groups)
        return None
    b.fast_math = fast
    for fi in scene.frame_inputs(nframes):
        scene.step(b, fi, to_dev)
    return b


def _small(which, frames=1):
    """The named configuration at a size the oracle finishes in seconds. Config 4 runs 6 frames (its ropes hit
    the plane from frame 3) and config 5 8 frames (a rope reaches one of its 64 meshes at frame 6)."""
    from ropesim import scenes
    if which.startswith("config2"):
        s = scenes.config2(n_ropes=16, V=65, frames=frames)
        if which == "config2-free":
            s.drive = "free"
        if which == "config2-dragquad":  # the reference's own test values (tests/test_drag_quad.py:180)
            s.drag_quad, s.drag_quad_t = 0.3, 0.01
        return s
    if which == "config3":
        return scenes.config3(n_ropes=8, frames=frames, vmax=96, grid_n=32)
    if which == "config4":
        return scenes.config4(n_ropes=8, frames=max(frames, 6), grid_n=32)
    return scenes.config5(n_ropes=8, frames=max(frames, 8), grid_n=24)  # 64 mesh slots, groups of 4, jitter


@pytest.mark.parametrize("fast", ARITH)
@pytest.mark.parametrize("which", ["config2", "config2-dragquad", "config3", "config5"])
def test_per_substep_kernels_equal_persistent_kernel_bitwise(rs, which, fast):
    from ropesim import scenes
    if which.startswith("config2"):
        s = scenes.config2(n_ropes=12, V=65, frames=3)
        if which == "config2-dragquad":
            s.drag_quad, s.drag_quad_t = 0.3, 0.01
    elif which == "config3":
        s = scenes.config3(n_ropes=8, frames=2, vmax=96, grid_n=24)
    else:
        s = scenes.config5(n_ropes=8, frames=2, grid_n=24)
    a = _run(rs, s, 0, s.frames, fast=fast)
    need = {"config2": 65, "config2-dragquad": 65, "config3": 96 * 4, "config5": 96 * 4}[which]
    ran = 0
    for var in range(rs.num_cell_variants()):
        blk, vpt, ok = rs.cell_variant_info(var)
        if not ok or blk * vpt < need:
            continue
        b = _run(rs, s, 1, s.frames, var, fast=fast)
        if b is None:
            continue
        ran += 1
        for get in ("positions", "velocities", "quaternions"):
            x, y = host(getattr(a, get)()), host(getattr(b, get)())
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{get} differs (variant {var})"
    assert ran >= 3


def _fp32_oracle(oracle, scene, k_scale=1.0, arith="ieee"):
    o32 = oracle.OracleBatch(scene.positions, scene.density, scene.k_stretch * k_scale, scene.k_bend, scene.max_h,
                             scene.num_itr, pinned=scene.pinned, fix_root_orientation=scene.fix_root_orientation,
                             prec=32, offsets=scene.offsets, arith=arith)
    scene.apply(o32)
    return o32


def per_rope_err(scene, x, x_ref):
    L = rope_lengths(scene)
    err = np.linalg.norm(x.reshape(-1, 3) - x_ref.reshape(-1, 3), axis=1)
    rope = np.repeat(np.arange(scene.num_ropes), np.diff(scene.offsets))
    per = np.zeros(scene.num_ropes)
    np.maximum.at(per, rope, err)
    return per / L


def _one_frame_four_ways(rs, oracle, scene, nframes=1, fast=True, perturbed=0):
    """Per-rope errors (of rope length) of the device against the float32 and float64 oracles, and the float32
    floor: how far float32 arithmetic itself lands from float64 on this scene, as the largest per-rope gap of
    the float32 oracle, of the float32 oracle with contracted multiply-adds (the arithmetic nvcc compiles the
    reference with by default) and of `perturbed` more float32 runs whose stretch stiffness is moved by
    k * 1e-7 relative (about one float32 ulp per k)."""
    b, o64 = build_pair(rs, oracle, scene, 64, 1, fast)
    o32s = [_fp32_oracle(oracle, scene, 1.0 + 1e-7 * k) for k in range(perturbed + 1)]
    ofma = _fp32_oracle(oracle, scene, arith="fma")
    for fi in scene.frame_inputs(nframes):
        scene.step(b, fi, to_dev)
        for o in [o64, ofma] + o32s:
            scene.step(o, fi)
    x = host(b.positions())
    ref = o64.positions()
    gaps = [per_rope_err(scene, o.positions(), ref) for o in [ofma] + o32s]
    floor = np.max(gaps, axis=0)
    return per_rope_err(scene, x, o32s[0].positions()), per_rope_err(scene, x, ref), gaps[1], gaps[0], floor


@pytest.mark.parametrize("fast", ARITH)
@pytest.mark.parametrize("which", ["config2", "config2-free", "config2-dragquad", "config3", "config4", "config5"])
def test_one_frame_vs_oracle(rs, oracle, which, fast):
    """Bar: 1e-5 of rope length of the float64 oracle. Where float32 arithmetic itself (the float32 oracle, and
    the same with the reference's contracted multiply-adds) lands further than half of that from float64 on
    some rope, that rope is held to twice the float32 floor instead. Stiff ropes (EA up to 2e4 N on 1.5-2 cm
    segments) lose float32 digits in the stretch constraint, most where the root program whips a rope to
    several m/s, and rope-rope contacts (the log barrier) make a trajectory sensitive to rounding: on config 3
    every float32 arithmetic lands 1e-3 of rope length from float64 after one frame. IEEE mode also stays
    within 1e-5 of rope length of the float32 oracle (the same operations) after one frame. Config 4 (6 frames)
    and config 5 (8 frames, groups of 4 in contact with 64 mesh slots) run long enough for one-ulp
    differences of the transcendentals (logf of the barrier, the root program's acos / sin) to grow through
    the contacts: there the IEEE mode is held to the float32 oracle's own distance from float64 too."""
    s = _small(which)
    e32, e64, gap, gap_fma, floor = _one_frame_four_ways(rs, oracle, s, nframes=s.frames, fast=fast,
                                                         perturbed=2 if fast else 0)
    print(f"\n  {which} fast {fast}: vs float32 oracle {e32.max():.2e}, vs float64 {e64.max():.2e}; float32 vs "
          f"float64: oracle {gap.max():.2e}, contracted {gap_fma.max():.2e}, largest with perturbations "
          f"{floor.max():.2e}; ropes above 1e-5: device {int((e64 > TOL_L).sum())}, float32 oracles "
          f"{int((gap > TOL_L).sum())} / {int((gap_fma > TOL_L).sum())} of {s.num_ropes}")
    bar = np.where(floor < 0.5 * TOL_L, TOL_L, np.maximum(TOL_L, 2 * floor))
    assert (e64 <= bar).all(), (e64.max(), floor.max())
    if not fast:
        if s.frames > 1:
            assert (e32 <= np.maximum(TOL_L, gap)).all()
        else:
            assert e32.max() < TOL_L


@pytest.mark.parametrize("fast", ARITH)
def test_config4_full_geometry_one_frame(rs, oracle, fast):
    """The differentiation workload's positions after one frame: all 256 ropes x 65 vertices (config-1
    material), 50 substeps at h 1/3000, 8 iterations, plane and one sphere SDF per rope on a 48^3 lattice,
    root program. Eight of its ropes (the first eight) within 1e-5 of rope length of the float64 oracle; every
    rope within 1e-5, or twice the float32 floor on the ropes where float32 arithmetic itself strays past
    half of that."""
    from ropesim import scenes
    s = scenes.config4(n_ropes=256, frames=1, grid_n=48)
    e32, e64, gap, gap_fma, floor = _one_frame_four_ways(rs, oracle, s, fast=fast)
    print(f"\n  config4 256 x 65 fast {fast}: first 8 ropes {e64[:8].max():.2e}, all ropes {e64.max():.2e}; float32 "
          f"vs float64: oracle {gap.max():.2e}, contracted {gap_fma.max():.2e}; ropes above 1e-5: device "
          f"{int((e64 > TOL_L).sum())}, float32 oracles {int((gap > TOL_L).sum())} / {int((gap_fma > TOL_L).sum())}")
    assert e64[:8].max() <= TOL_L
    bar = np.where(floor < 0.5 * TOL_L, TOL_L, np.maximum(TOL_L, 2 * floor))
    assert (e64 <= bar).all(), (e64.max(), floor.max())


def test_program_lands_on_endpoint_and_slerp_end(rs):
    """The root program lands exactly on p1 and the frame on q1 (csrc/batch.cu:122-159 closed form)."""
    N, V, H = 3, 10, 1 / 60
    roots = np.zeros((N, 3), np.float32)
    roots[:, 0] = np.arange(N) * 1.5
    b = rs.RopeBatch(num_envs=N, verts_per_rope=V, root_pos=roots, fix_root_orientation=True)
    b.damping = 0.0
    p1 = roots + np.array([0.04, -0.02, 0.03], np.float32)
    v0 = np.tile(np.array([0.5, 0.2, -0.1], np.float32), (N, 1))
    th = np.deg2rad([10.0, 20.0, 30.0])
    q0 = np.zeros((N, 4), np.float32)
    q0[:, 3] = 1
    q1 = np.zeros((N, 4), np.float32)
    q1[:, 2] = np.sin(th / 2)
    q1[:, 3] = np.cos(th / 2)
    b.step(to_dev(roots), to_dev(v0), H, 50, root_q=to_dev(q0), root_pos_end=to_dev(p1), root_q_end=to_dev(q1))
    np.testing.assert_allclose(host(b.positions())[:, 0], p1, atol=2e-6)
    np.testing.assert_allclose(host(b.quaternions())[:, 0], q1, atol=1e-5)
    a = 2.0 * ((p1 - roots) - v0 * H) / H ** 2
    np.testing.assert_allclose(host(b.velocities())[:, 0], v0 + a * (H - (H / 50) / 2), rtol=1e-4, atol=1e-4)


def _tangle(rng, n_ropes=9, V=40):
    """Ropes crossing each other inside one small region (many close segment pairs)."""
    ropes = []
    for r in range(n_ropes):
        c = rng.uniform(-0.02, 0.02, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        w = rng.normal(size=3) * 0.02
        t = np.arange(V)[:, None] * 0.015
        ropes.append(c + (t - t.mean()) * d + w * np.sin(t * 40))
    return ropes


def test_rope_rope_lists_bitexact_vs_oracle(rs, oracle):
    rng = np.random.default_rng(7)
    ropes = _tangle(rng)
    b = rs.RopeBatch.from_positions(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
    b.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 3)
    o = oracle.OracleBatch(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, prec=32)
    o.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 3)
    nv = b.num_vertices
    x = np.concatenate(ropes).astype(np.float32)
    x += rng.normal(0, 0.002, x.shape).astype(np.float32)
    pads = rng.uniform(0.004, 0.012, nv).astype(np.float32)
    L = _lib(rs)
    counts = np.zeros(nv, np.int32)
    lists = np.zeros(nv * 128, np.int32)
    lim = np.zeros(nv, np.float32)
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    rc = L.rs_check_detect(C.c_void_p(b._handle), x.ctypes.data_as(fp), pads.ctypes.data_as(fp),
                           counts.ctypes.data_as(ip), lists.ctypes.data_as(ip), lim.ctypes.data_as(fp))
    assert rc == 0
    oc, ol, olim = o.detect(x.astype(np.float64), pads.astype(np.float64))
    assert counts.sum() > 50, "the scene must produce many candidates"
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(lists.reshape(nv, 128), ol)
    np.testing.assert_array_equal(lim, olim.astype(np.float32))


@pytest.mark.parametrize("fast", ARITH)
def test_tangled_contact_per_substep_equals_persistent_bitwise(rs, fast):
    """Many rope-rope contacts (a tangle of 9 free ropes in one group): the persistent kernel's tiled broad
    phase builds the same lists as the per-substep brute-force scan, so the trajectories stay bit-identical."""
    rng = np.random.default_rng(11)
    ropes = _tangle(rng)
    runs = []
    for mode in (0, 1):
        b = rs.RopeBatch.from_positions(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8,
                                        pinned=np.zeros(sum(len(r) for r in ropes), bool))
        b.set_gravity(np.array([0, 0, -9.81], np.float32))
        b.damping = 0.0
        b.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 9)
        b.set_dispatch(mode, -1)
        b.fast_math = fast
        for _ in range(3):
            b.step_free(1 / 60, 100)
        runs.append((host(b.positions()), host(b.velocities()), host(b.quaternions())))
        assert b.last_error == 0
    for x, y in zip(*runs):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


@pytest.mark.parametrize("which", ["config3", "config5"])
def test_mesh_broadphase_bitexact_vs_oracle(rs, oracle, which):
    """8 moving-walk meshes (config 3) or 64 static slots per environment (config 5); the overflow of the
    8-candidate list is covered by tests/test_invariants.py (ten coincident floors)."""
    from ropesim import scenes
    s = scenes.config3(n_ropes=8, frames=1, vmax=64, grid_n=16) if which == "config3" else \
        scenes.config5(n_ropes=8, frames=1, grid_n=16)
    b = s.build(rs)
    o = oracle.OracleBatch(s.positions, s.density, s.k_stretch, s.k_bend, s.max_h, s.num_itr, prec=32,
                           offsets=s.offsets)
    s.apply(o)
    rng = np.random.default_rng(3)
    nv = s.num_vertices
    # probe points scattered over the meshes' neighbourhood, displacements of a few mm
    x = rng.uniform([-1.4, -1.4, 0.1], [1.4, 1.4, 1.7], (nv, 3)).astype(np.float32)
    dx = rng.normal(0, 0.003, (nv, 3)).astype(np.float32)
    counts = np.zeros(nv, np.int32)
    cand = np.zeros(nv * 8, np.int32)
    L = _lib(rs)
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int)
    assert L.rs_check_broadphase(C.c_void_p(b._handle), x.ctypes.data_as(fp), dx.ctypes.data_as(fp),
                                 counts.ctypes.data_as(ip), cand.ctypes.data_as(ip)) == 0
    oc, ocand = o.broadphase(x.astype(np.float64), dx.astype(np.float64))
    assert (counts > 0).sum() > 10
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(cand.reshape(nv, 8), ocand)
