"""Rope-rope detection through the per-group bounding-volume hierarchy (the reference's LBVH path, reached by
RopeSim(collisions=True): third_party/yarnball/KittenEngine/YarnBall/sim/collision.cu:10-101). The hierarchy
must give the full scan's candidate lists (ascending partner index, lowest 128 kept) and step limits bit for
bit, so trajectories with either detection are bit-identical; the scan itself is pinned against the oracle
in test_gpu_parity.py and test_golden.py, and the tangle case here checks the hierarchy against the oracle
directly."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import host

pytestmark = pytest.mark.gpu


def _lib(rs):
    return C.CDLL(os.path.join(os.path.dirname(rs.__file__), "librope_hip.so"))


def _ropes(rng, lengths, box=0.02, seg=0.015):
    """Ropes of the given vertex counts crossing inside one small region."""
    out = []
    for V in lengths:
        c = rng.uniform(-box, box, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        w = rng.normal(size=3) * 0.02
        t = np.arange(V)[:, None] * seg
        out.append(c + (t - t.mean()) * d + w * np.sin(t * 40))
    return out


def _detect(rs, b, x, pads, mode):
    b.rope_broad_phase = mode
    nv = b.num_vertices
    counts = np.zeros(nv, np.int32)
    lists = np.zeros(nv * 128, np.int32)
    lim = np.zeros(nv, np.float32)
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int)
    x = np.ascontiguousarray(x, np.float32)
    pads = np.ascontiguousarray(pads, np.float32)
    rc = _lib(rs).rs_check_detect(C.c_void_p(b._handle), x.ctypes.data_as(fp), pads.ctypes.data_as(fp),
                                  counts.ctypes.data_as(ip), lists.ctypes.data_as(ip), lim.ctypes.data_as(fp))
    assert rc == 0
    return counts, lists.reshape(nv, 128), lim


def test_bvh_lists_match_oracle_on_tangle(rs, oracle):
    rng = np.random.default_rng(7)
    ropes = _ropes(rng, [40] * 9)
    b = rs.RopeBatch.from_positions(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
    b.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 3)
    o = oracle.OracleBatch(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, prec=32)
    o.set_rope_contact(True, 0.0036, 0.0008, 0.025, 0.4, 1e7, 1.2, 3)
    x = np.concatenate(ropes).astype(np.float32)
    x += rng.normal(0, 0.002, x.shape).astype(np.float32)
    pads = rng.uniform(0.004, 0.012, len(x)).astype(np.float32)
    counts, lists, lim = _detect(rs, b, x, pads, "bvh")
    assert b.rope_bvh_active
    oc, ol, olim = o.detect(x.astype(np.float64), pads.astype(np.float64))
    assert counts.sum() > 50
    np.testing.assert_array_equal(counts, oc)
    np.testing.assert_array_equal(lists, ol)
    np.testing.assert_array_equal(lim, olim.astype(np.float32))


CASES = {
    "ragged, groups of 2": dict(seed=2, lengths=[16, 33, 7, 50, 3, 64, 4, 3, 90, 40], group=2, box=0.02),
    "one long rope": dict(seed=3, lengths=[2500], group=1, box=0.0),
    "uneven groups of 3 with 3-vertex ropes": dict(seed=5, lengths=[300, 3, 200, 4, 5, 700, 3, 3, 64], group=3,
                                                   box=0.03),
    "many groups": dict(seed=6, lengths=[24] * 300, group=4, box=0.01),
    "dense pile (lists overflow)": dict(seed=8, lengths=[40] * 60, group=60, box=0.004),
}


@pytest.mark.parametrize("case", list(CASES))
def test_bvh_lists_equal_full_scan(rs, case):
    c = CASES[case]
    rng = np.random.default_rng(c["seed"])
    ropes = _ropes(rng, c["lengths"], box=c["box"], seg=0.004 if c["box"] == 0.0 else 0.015)
    if c["box"] == 0.0:  # a long rope folded into a tight zigzag: most contacts far apart in index
        V = c["lengths"][0]
        k = np.arange(V)
        ropes = [np.column_stack([0.004 * (k % 50), 0.006 * ((k // 50) % 10), 0.005 * (k // 500)])
                 + rng.normal(0, 0.0005, (V, 3))]
    b = rs.RopeBatch.from_positions(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
    b.set_rope_contact(True, 0.0005, 0.0004, 0.025, 0.4, 1e7, 1.2, c["group"])
    x = np.concatenate(ropes).astype(np.float32)
    pads = rng.uniform(0.002, 0.01, len(x)).astype(np.float32)
    scan = _detect(rs, b, x, pads, "scan")
    tree = _detect(rs, b, x, pads, "bvh")
    assert scan[0].sum() > 20
    if case.startswith("dense"):
        assert scan[0].max() == 128, "the pile must overflow some lists"
    for s, t in zip(scan, tree):
        assert np.array_equal(s.view(np.uint32), t.view(np.uint32))


def _helix(V=3000, seg=0.005):
    """One rope wound as a tight helix (turns 3.2 mm apart, just outside the 2.8 mm contact distance) standing
    on its pinned first vertex: gravity settles the coil onto itself."""
    R = 0.03
    dphi = seg / R
    k = np.arange(V)
    phi = k * dphi
    return [np.column_stack([R * np.cos(phi), R * np.sin(phi), 0.0032 * phi / (2 * np.pi)])]


def _run_helix(rs, mode, dispatch=0, graph=False, contact=True, frames=3):
    b = rs.RopeBatch.from_positions(_helix(), 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
    b.set_gravity(np.array([0, 0, -9.81], np.float32))
    if contact:
        b.set_rope_contact(True, 0.001, 0.0008, 0.025, 0.4, 1e7, 1.2, 1)
        b.rope_broad_phase = mode
    b.set_dispatch(dispatch)
    if graph:
        b.set_graph(True)
    out = []
    for _ in range(frames):
        b.step_free(1 / 60, 20)
        out.append(host(b.positions()).copy())
    assert b.last_error == 0
    return b, out


def test_long_rope_self_contact_trajectory_bvh_equals_scan(rs):
    """A 3000-vertex helix settling under gravity with self-contact: bit-identical frames with either
    detection, and contact changes the motion."""
    ts, scan = _run_helix(rs, "scan")
    tb, tree = _run_helix(rs, "bvh")
    ta, auto = _run_helix(rs, "auto")
    assert not ts.rope_bvh_active and tb.rope_bvh_active and ta.rope_bvh_active
    _, free = _run_helix(rs, "scan", contact=False)
    for f, (s, t, a) in enumerate(zip(scan, tree, auto)):
        assert np.array_equal(s.view(np.uint32), t.view(np.uint32)), f"frame {f}"
        assert np.array_equal(s.view(np.uint32), a.view(np.uint32)), f"frame {f}"
    assert np.abs(scan[-1] - free[-1]).max() > 1e-6, "the helix must be in self-contact"


def test_group_larger_than_any_cell_runs_per_substep_kernels(rs):
    """Dispatch mode 1 with a contact group no cell holds runs the per-substep kernels (bit-identical to
    mode 0) instead of failing."""
    _, m0 = _run_helix(rs, "auto", dispatch=0, frames=2)
    _, m1 = _run_helix(rs, "auto", dispatch=1, frames=2)
    for s, t in zip(m0, m1):
        assert np.array_equal(s.view(np.uint32), t.view(np.uint32))


def test_bvh_graph_replays_match_eager(rs):
    """The hierarchy's kernels (sort, build, refit, query) captured into the frame graph and replayed."""
    import torch
    _, eager = _run_helix(rs, "bvh", frames=4)
    b, graph = _run_helix(rs, "bvh", graph=True, frames=4)
    scratch = torch.randn(1 << 16, device="cuda")
    float((scratch * 2).sum())
    for f, (s, t) in enumerate(zip(eager, graph)):
        assert np.array_equal(s.view(np.uint32), t.view(np.uint32)), f"frame {f}"
    assert b.graph_captures >= 1


def test_ropesim_collisions_long_rope(rs):
    """RopeSim(collisions=True) on a rope longer than any cell: runs through the hierarchy (the reference's
    test_api.py:197-203 on a 2000-vertex rope)."""
    r = rs.RopeSim(num_verts=2100, seg_len=0.005, collisions=True)
    assert r.collisions_enabled and r.rope_bvh_active
    r.drag = 0.5
    for _ in range(10):
        r.advance(1 / 60)
    assert np.isfinite(host(r.positions())).all()
    assert r.last_error == 0


def test_broad_phase_setting_and_errors(rs):
    """The setting reads back, an unknown mode raises ValueError, and without rope contact no hierarchy is
    in use."""
    ropes = _ropes(np.random.default_rng(1), [200, 200])
    b = rs.RopeBatch.from_positions(ropes, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
    assert b.rope_broad_phase == "auto" and not b.rope_bvh_active
    b.set_rope_contact(True, 0.0005, 0.0004, 0.025, 0.4, 1e7, 1.2, 2)
    assert b.rope_bvh_active  # a 400-vertex group
    b.rope_broad_phase = "scan"
    assert b.rope_broad_phase == "scan" and not b.rope_bvh_active
    with pytest.raises(ValueError):
        b.rope_broad_phase = "lbvh"
    assert b.rope_broad_phase == "scan"
    b.rope_broad_phase = "bvh"
    b.set_rope_contact(False, 0.0005, 0.0004, 0.025, 0.4, 1e7, 1.2, 2)
    assert b.rope_broad_phase == "bvh" and not b.rope_bvh_active
