"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from the CPU oracle).

CPU: the oracle still reproduces them, and the seeded inputs still hash to what they were made from.
GPU: the HIP path against them. Bars: the config-1 trajectory within 1e-5 of rope length every frame;
rope-rope candidate lists, step limits and mesh broad-phase lists bit-exact.
"""
import ctypes as C
import os

import numpy as np
import pytest

from golden.make_golden import config1_case, mesh_case, rope_rope_case, sha, tangle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def unpack_lists(g, nv):
    lists = np.full((nv, 128), -1, np.int32)
    for i in range(nv):
        a, b = g["list_offsets"][i], g["list_offsets"][i + 1]
        lists[i, : b - a] = g["lists"][a:b]
    return lists


# ------------------------------------------------------------------ CPU: the oracle reproduces the fixtures

def test_oracle_reproduces_config1_trajectory(oracle):
    g, now = load("config1_trajectory"), config1_case(oracle)
    np.testing.assert_array_equal(now["input_sha"], g["input_sha"])
    np.testing.assert_allclose(now["positions"], g["positions"], rtol=0, atol=1e-12)


def test_oracle_reproduces_rope_rope_lists(oracle):
    g, now = load("rope_rope_lists"), rope_rope_case(oracle)
    for k in ("input_sha", "counts", "lists", "list_offsets", "step_limit"):
        np.testing.assert_array_equal(now[k], g[k])


def test_oracle_reproduces_mesh_broadphase(oracle):
    g, now = load("mesh_broadphase"), mesh_case(oracle)
    for k in ("input_sha", "counts", "cand"):
        np.testing.assert_array_equal(now[k], g[k])


# ------------------------------------------------------------------ GPU: the HIP path against the fixtures

def _lib(rs):
    return C.CDLL(os.path.join(os.path.dirname(rs.__file__), "librope_hip.so"))


@pytest.mark.gpu
@pytest.mark.parametrize("fast", [True, False], ids=["fast", "ieee"])
@pytest.mark.parametrize("mode", [0, 1])
def test_device_config1_trajectory_matches_golden(rs, mode, fast):
    import torch
    from ropesim import scenes
    g = load("config1_trajectory")
    s = scenes.config1(frames=60)
    b = s.build(rs)
    b.set_dispatch(mode)
    b.fast_math = fast
    L = 0.03 * 32
    worst = 0.0
    for f, fi in enumerate(s.frame_inputs(60)):
        s.step(b, fi, lambda a: torch.as_tensor(a, device="cuda"))
        x = torch.as_tensor(b.positions(), device="cuda").cpu().numpy().reshape(-1, 3).astype(np.float64)
        worst = max(worst, float(np.abs(x - g["positions"][f]).max()) / L)
    assert worst < 1e-5, worst


@pytest.mark.gpu
@pytest.mark.parametrize("broad_phase", ["scan", "bvh"])
def test_device_rope_rope_lists_match_golden(rs, broad_phase):
    """Both rope-rope detections of the per-substep kernels (the full scan and the per-group hierarchy)."""
    g = load("rope_rope_lists")
    ropes = tangle(np.random.default_rng(7))
    b = rs.RopeBatch.from_positions(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
    b.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 3)
    b.rope_broad_phase = broad_phase
    assert b.rope_bvh_active == (broad_phase == "bvh")
    nv = b.num_vertices
    x, pads = g["x"], g["pads"]
    counts = np.zeros(nv, np.int32)
    lists = np.zeros(nv * 128, np.int32)
    lim = np.zeros(nv, np.float32)
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int)
    assert _lib(rs).rs_check_detect(C.c_void_p(b._handle), x.ctypes.data_as(fp), pads.ctypes.data_as(fp),
                                    counts.ctypes.data_as(ip), lists.ctypes.data_as(ip), lim.ctypes.data_as(fp)) == 0
    np.testing.assert_array_equal(counts, g["counts"])
    np.testing.assert_array_equal(lists.reshape(nv, 128), unpack_lists(g, nv))
    np.testing.assert_array_equal(lim, g["step_limit"])


@pytest.mark.gpu
def test_device_mesh_broadphase_matches_golden(rs):
    from ropesim import scenes
    g = load("mesh_broadphase")
    s = scenes.config3(n_ropes=8, frames=1, vmax=64, grid_n=16)
    b = s.build(rs)
    nv = s.num_vertices
    counts = np.zeros(nv, np.int32)
    cand = np.zeros(nv * 8, np.int32)
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int)
    assert _lib(rs).rs_check_broadphase(C.c_void_p(b._handle), g["x"].ctypes.data_as(fp), g["dx"].ctypes.data_as(fp),
                                        counts.ctypes.data_as(ip), cand.ctypes.data_as(ip)) == 0
    np.testing.assert_array_equal(counts, g["counts"])
    np.testing.assert_array_equal(cand.reshape(nv, 8), g["cand"])


def test_config4_gradient_references_agree_with_central_differences():
    """The float64 autograd gradients of the differentiation workload (tests/golden/config4_gradients.npz)
    project onto the float64 oracle's per-rope central differences within 1e-3, and the per-rope drag
    differences sum to the autograd drag gradient."""
    import os
    ref = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config4_gradients.npz"))
    for k in ("ea", "ei", "m", "p0", "v0", "p1"):
        for r in range(8):
            ad = float(np.sum(ref["g64_" + k] * ref["dir_" + k][r]))
            fd = float(ref["fd_" + k][r])
            assert abs(ad - fd) <= 1e-3 * abs(fd), (k, r, ad, fd)
    assert abs(float(ref["fd_drag"].sum()) - float(ref["g64_drag"])) <= 1e-3 * abs(float(ref["g64_drag"]))
    assert ref["tgt"].shape == (6, 8, 3) and np.isfinite(ref["g64_ea"]).all()
