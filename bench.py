#!/usr/bin/env python3
"""Throughput of the batched rope solver on one MI355X, in vertex-substeps/s.

Workload (the headline configuration): 4096 ropes with vertex counts U{16..256} (seed 1), 0.015 m
segments, per-segment EA / EI, contact groups of 4 with rope-rope contact, 8 kinematic SDF meshes
(64^3 float32 grids) per environment on SE(3) random walks, ground plane, root program, 100 substeps
of 1/6000 s per 1/60 s frame, 8 iterations. A step is one frame. Inputs (geometry, per-frame root
programs and mesh targets) are resident on the GPU before the timed region.

One JSON line: metric, value, ms_per_step over frames 5..24 by default (the per-frame times in
"frame_times"), roofline of the persistent frame kernel (algorithmic bytes per launch over its
HIP-event-timed duration on the solver stream), a VALU figure, the same window in the other arithmetic mode
("other_arithmetic"), and the CPU baseline (the oracle on the host cores in float32 and float64, a bounded
sample of the same workload, with the device's one-frame position error against float64 on samples of every
configuration). "plain" times the plain method on configs 1 and 2. The "differentiation" object is the second
half of the metric, on the config-4 workload:
- forward time of the per-substep kernels and of the persistent frame kernel;
- a differentiable rollout's forward (with per-frame checkpoints) and adjoint time, and their ratio, with the
  tips tracking a trajectory made with hidden per-vertex EA / EI.
The "readout" object continues the headline batch with the contact-force readout on (float wrench sums, then
the reproducible fixed-point sums) and reports the time per frame of each against the readout off, and does the
same on ropes resting on a slab mesh. The "bake_mesh" object times ropesim.sdf.bake_mesh on a 32768-triangle
torus. The "rope_bvh" object times one long rope in self-contact, and the headline batch, through the
per-substep kernels with the full-scan rope-rope detection and with the per-group bounding-volume hierarchy.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "cosserat-rod-solver-cuda_amd")]

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
B_ALG = 136.0  # bytes per vertex-substep: the float32 state round trip (x, v, v_last, q, w, l, k, rest q, flags)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_cell_4096_uneven_ropes_hbm_traffic.md")


def profiled_traffic():
    """HBM bytes of one frame-kernel launch on the default workload, as measured with two rocprofv3 --pmc runs
    (FETCH_SIZE and WRITE_SIZE, scaled by a plain device copy of known size) and recorded in the traffic
    profile; None when the profile is absent."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                m = re.match(r"traffic_bytes_per_launch = (\d+)\s*$", line)
                if m:
                    return int(m.group(1))
    except OSError:
        pass
    return None


# Flops of the method per unit, counted by running the oracle on an operation-counting scalar
# (oracle/opcount.cpp; add, sub, mul, div, sqrt and transcendentals one flop each, unfused). Checked against a
# fresh count by tests/test_opcount.py.
FLOP = {
    "init_per_vertex_substep": 41.615,
    "position_sweep_per_vertex_iteration": 160.508,
    "orientation_sweep_per_vertex_iteration": 243.969,
    "end_per_vertex_substep": 11.877,
    "plane_per_free_vertex_iteration": 16.000,
    "mesh_broad_phase_per_vertex_substep_per_mesh": 14.769,
    "mesh_candidate_per_iteration": 106.000,
    "pair_test": 108.016,
}


def flops_per_vertex_substep(iters, planes=0, meshes=0):
    """Counted flops per vertex-substep of the solver (init, iterations, end), the plane tests and the mesh
    broad phase. Mesh-candidate terms and rope-rope pair terms depend on the contact state and are not
    included."""
    f = FLOP["init_per_vertex_substep"] + FLOP["end_per_vertex_substep"]
    f += iters * (FLOP["position_sweep_per_vertex_iteration"] + FLOP["orientation_sweep_per_vertex_iteration"])
    f += iters * planes * FLOP["plane_per_free_vertex_iteration"]
    f += meshes * FLOP["mesh_broad_phase_per_vertex_substep_per_mesh"]
    return f


def per_rope_error(scene, x, ref):
    """Largest vertex distance of each rope over the rope's length."""
    import numpy as np
    p = scene.positions.astype(np.float64)
    off = scene.offsets
    d = np.linalg.norm(np.asarray(x, np.float64).reshape(-1, 3) - np.asarray(ref, np.float64).reshape(-1, 3), axis=1)
    out = np.zeros(scene.num_ropes)
    for r in range(scene.num_ropes):
        L = np.linalg.norm(np.diff(p[off[r]:off[r + 1]], axis=0), axis=1).sum()
        out[r] = d[off[r]:off[r + 1]].max() / L
    return out


def device_frames(scene, fast, frames=1):
    """Positions of the HIP batch after `frames` frames of the scene, in the given arithmetic mode."""
    import torch

    import ropesim
    b = scene.build(ropesim)
    b.fast_math = fast
    for fi in scene.frame_inputs(frames):
        scene.step(b, fi, lambda a: torch.as_tensor(a, device="cuda"))
    return torch.as_tensor(b.positions(), device="cuda").cpu().numpy().reshape(-1, 3)


def one_frame_errors(scene, o64, o32):
    """Largest per-rope position error (of rope length) after one frame against the float64 oracle: the device
    in fast and IEEE arithmetic, and the float32 oracle itself (the floor float32 arithmetic sets)."""
    import numpy as np
    ref = o64.positions()
    out = {"ropes": scene.num_ropes}
    for name, fast in (("device_fast", True), ("device_ieee", False)):
        e = per_rope_error(scene, device_frames(scene, fast), ref)
        out[name] = float(e.max())
        out[name + "_ropes_above_1e-5"] = int((e > 1e-5).sum())
    e32 = per_rope_error(scene, o32.positions(), ref)
    out["float32_oracle"] = float(e32.max())
    out["float32_oracle_ropes_above_1e-5"] = int((e32 > 1e-5).sum())
    return out


def cpu_baseline(sample_ropes, threads):
    """The oracle (OpenMP on the host cores) on a bounded sample of the headline workload, float32 (the
    reference's precision) and float64, one frame each; and, from the same runs, the one-frame position error of
    the device against float64 there and on samples of the other configurations."""
    import numpy as np

    from oracle.pyoracle import OracleBatch
    from ropesim import scenes

    def oracle_frame(scene, prec):
        o = OracleBatch(scene.positions, scene.density, scene.k_stretch, scene.k_bend, scene.max_h, scene.num_itr,
                        pinned=scene.pinned, prec=prec, offsets=scene.offsets)
        scene.apply(o)
        fi = scene.frame_inputs(1)[0]
        t0 = time.perf_counter()
        scene.step(o, fi)
        dt = time.perf_counter() - t0
        assert np.isfinite(o.positions()).all()
        return o, dt

    s = scenes.config3(n_ropes=sample_ropes, frames=1)
    o32, t32 = oracle_frame(s, 32)
    o64, t64 = oracle_frame(s, 64)
    units = s.num_vertices * s.substeps
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    errors = {"config3_sample": one_frame_errors(s, o64, o32)}
    others = {"config1": scenes.config1(frames=1), "config2_sample": scenes.config2(n_ropes=64, frames=1),
              "config4": scenes.config4(n_ropes=256, frames=1), "config5_sample": scenes.config5(n_ropes=64, frames=1)}
    for name, sc in others.items():
        errors[name] = one_frame_errors(sc, oracle_frame(sc, 64)[0], oracle_frame(sc, 32)[0])
    return {"value": units / t32, "unit": "vertex-substeps/s", "cores": threads, "kind": "port",
            "sample": f"the oracle (OpenMP, {threads} threads on {model}) on {sample_ropes} ropes ({s.num_vertices} "
                      f"vertices, {sample_ropes // 4} environments) of the headline generator, 1 frame of "
                      f"{s.substeps} substeps: float32 {t32:.1f} s (value), float64 {t64:.1f} s",
            "float64": {"value": units / t64, "unit": "vertex-substeps/s", "seconds": t64},
            "one_frame_error_vs_float64": errors,
            "one_frame_error_note": "largest per-rope vertex error over rope length after one frame; samples: "
                                    "config 1 whole, config 2 64 ropes, config 3 the CPU sample, config 4 all 256 "
                                    "ropes, config 5 64 ropes with its 64 mesh slots"}


def plain_legs(frames1=60, frames2=120, warm=5):
    """The plain method on configs 1 and 2 (the reference's throughput harness shape, examples/throughput.py:
    a warm-up frame that captures the graph, then frames timed back to back): config 1 through the drop-in
    RopeSim (set_root + advance(1/60, substeps=50) per frame) and config 2 through RopeBatch with the root
    program, each with the per-substep kernels (graph-replayed) and with the persistent frame kernel."""
    import numpy as np
    import torch

    import ropesim
    from ropesim import scenes
    out = {}
    c1 = scenes.config1(frames=frames1 + warm)
    row = {"workload": "config1: one rope of 33 vertices (seg 0.03 m, rho 0.1, EA 500, EI 1e-4), drag 0.2, root at "
                       "0.1 m/s, RopeSim.set_root + advance(1/60, substeps=50) per frame", "frames": frames1}
    for mode in (0, 1):
        r = ropesim.RopeSim(num_verts=33, seg_len=0.03, dir=np.array([0, 0, -1], np.float32), density=0.1,
                            k_stretch=500.0, k_bend=1e-4, max_h=1 / 3000)
        r.set_gravity(np.array([0, 0, -9.81], np.float32))
        r.drag = 0.2
        r.damping = 0.0
        r.set_dispatch(mode)
        t0 = None
        for f in range(frames1 + warm):
            if f == warm:
                t0 = time.perf_counter()
            r.set_root(np.array([0.1 * f / 60, 0, 0], np.float32), np.array([0.1, 0, 0], np.float32))
            r.advance(1 / 60, substeps=50)
        wall = time.perf_counter() - t0
        assert np.isfinite(np.asarray(r.positions())).all()
        key = "per_substep_kernels" if mode == 0 else "persistent_kernel"
        row[key + "_frames_per_s"] = frames1 / wall
        row[key + "_vertex_substeps_per_s"] = 33 * 50 * frames1 / wall
    out["config1"] = row
    s = scenes.config2(frames=frames2 + warm)
    fis = s.frame_inputs(frames2 + warm)
    dev = [{k: torch.as_tensor(v, device="cuda") for k, v in fi.items() if v is not None} for fi in fis]
    row = {"workload": f"config2: {s.num_ropes} ropes x 65 vertices, root program, plane, 100 substeps x 8 iterations "
                       "per 1/60 s frame, RopeBatch.step", "frames": frames2}
    for mode in (0, 1):
        b = s.build(ropesim)
        b.set_dispatch(mode, -1)
        if mode == 1:
            b.autotune(s.dt, s.substeps, True)
        b.set_graph(True)
        for f in range(warm):
            s.step(b, dev[f])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(warm, warm + frames2):
            s.step(b, dev[f])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        assert bool(torch.isfinite(torch.as_tensor(b.positions(), device="cuda")).all())
        key = "per_substep_kernels" if mode == 0 else "persistent_kernel"
        row[key + "_ms_per_frame"] = 1e3 * wall / frames2
        row[key + "_vertex_substeps_per_s"] = s.num_vertices * s.substeps * frames2 / wall
        del b
    row["persistent_speedup"] = row["per_substep_kernels_ms_per_frame"] / row["persistent_kernel_ms_per_frame"]
    out["config2"] = row
    return out


def resting_readout_bench(n_ropes, frames=5):
    """Readout on ropes resting on one slab mesh (every vertex in contact in every substep): time per frame
    with the readout off, with float wrench sums and with fixed-point sums, after two warm-up frames each."""
    import numpy as np
    import torch

    import ropesim
    from ropesim import scenes
    s = scenes.resting(n_ropes=n_ropes)
    b = s.build(ropesim)
    b.set_dispatch(1, -1)
    b.autotune(s.dt, s.substeps, False)
    b.set_graph(True)
    out = {"workload": f"{n_ropes} free ropes x 65 vertices resting on one slab SDF mesh, 100 substeps x 8 "
                       "iterations per 1/60 s frame"}
    for name, on, det in (("off", False, False), ("float", True, False), ("fixed_point", True, True)):
        b.enable_contact_forces(on, deterministic=det)
        for _ in range(2):
            b.step_free(s.dt, s.substeps)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(frames):
            b.step_free(s.dt, s.substeps)
        torch.cuda.synchronize()
        out[name + "_ms_per_frame"] = 1e3 * (time.perf_counter() - t) / frames
    w = np.asarray(b.mesh_wrenches(), np.float64).reshape(n_ropes, 6)
    m = 1.0 / np.asarray(b.inv_masses(), np.float64)
    out["wrench_z_over_weight"] = float(-w[:, 2].sum() / (m.sum() * 9.81))
    return out


def torus_mesh(major, minor, nu, nv):
    """Closed triangulated torus around the z axis: nu x nv quads, two triangles each."""
    import numpy as np
    u = 2 * np.pi * np.arange(nu) / nu
    v = 2 * np.pi * np.arange(nv) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    pts = np.stack([(major + minor * np.cos(V)) * np.cos(U), (major + minor * np.cos(V)) * np.sin(U),
                    minor * np.sin(V)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    tri = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return pts, tri


def bake_bench():
    """Mesh distance grid on the GPU: a 32768-triangle torus on a ~96^3 lattice, signed by the winding number;
    node-triangle pairs per second (each pair: closest point on the triangle and its solid angle)."""
    import numpy as np

    from ropesim import sdf
    v, t = torus_mesh(0.2, 0.06, 256, 64)
    spacing, margin = 0.005, 0.03
    sdf.bake_mesh(v, t, spacing, margin)  # warm-up
    t0 = time.perf_counter()
    g = sdf.bake_mesh(v, t, spacing, margin)
    dt = time.perf_counter() - t0
    nodes = int(np.prod(g.values.shape))
    ring = np.hypot(*np.meshgrid(g.origin[0] + spacing * np.arange(g.values.shape[0]),
                                 g.origin[1] + spacing * np.arange(g.values.shape[1]), indexing="ij")) - 0.2
    exact = np.sqrt(ring[:, :, None] ** 2 + (g.origin[2] + spacing * np.arange(g.values.shape[2]))[None, None] ** 2) - 0.06
    return {"workload": f"closed torus mesh (R 0.2 m, r 0.06 m, {len(t)} triangles) on a "
                        f"{'x'.join(map(str, g.values.shape))} lattice, signed by the winding number",
            "ms": 1e3 * dt, "node_triangle_pairs_per_s": nodes * len(t) / dt,
            "max_abs_dev_from_torus_sdf_m": float(np.abs(g.values - exact).max()),
            "timing": "host to host: upload, kernel, download"}


def rope_bvh_bench(sizes, frames=3):
    """Rope-rope detection in one long rope with self-contact, through the per-substep kernels (what
    RopeSim(collisions=True) runs for ropes no cell holds): a helix of n vertices (5 mm segments, turns 3.2 mm
    apart) settling on its pinned first vertex, 20 substeps x 8 iterations per 1/60 s frame, frames captured
    as graphs. ms per frame with the full scan and with the hierarchy, and whether their positions agree bit
    for bit; then the same comparison on many small groups (the headline batch through the per-substep
    kernels)."""
    import numpy as np
    import torch

    import ropesim as rs
    out = {"workload": "one rope coiled as a helix (5 mm segments, turns 3.2 mm apart) in self-contact, per-substep "
                       "kernels, 20 substeps x 8 iterations per 1/60 s frame",
           "timing": f"wall clock of {frames} frames after one warm-up frame", "sizes": []}
    for n in sizes:
        phi = np.arange(n) * 0.005 / 0.03
        helix = [np.column_stack([0.03 * np.cos(phi), 0.03 * np.sin(phi), 0.0032 * phi / (2 * np.pi)])]
        row, pos = {"vertices": n}, {}
        for mode in ("scan", "bvh"):
            b = rs.RopeBatch.from_positions(helix, 0.117, 11473.0, 4.48e-3, 1 / 6000, 8, pinned=None)
            b.set_gravity(np.array([0, 0, -9.81], np.float32))
            b.set_rope_contact(True, 0.001, 0.0008, 0.025, 0.4, 1e7, 1.2, 1)
            b.rope_broad_phase = mode
            b.set_dispatch(0)
            b.set_graph(True)
            b.step_free(1 / 60, 20)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(frames):
                b.step_free(1 / 60, 20)
            torch.cuda.synchronize()
            row[f"{mode}_ms_per_frame"] = 1e3 * (time.perf_counter() - t0) / frames
            pos[mode] = torch.as_tensor(b.positions(), device="cuda").cpu().numpy()
        row["bvh_speedup"] = row["scan_ms_per_frame"] / row["bvh_ms_per_frame"]
        row["bitwise_equal"] = bool(np.array_equal(pos["scan"].view(np.uint32), pos["bvh"].view(np.uint32)))
        out["sizes"].append(row)
        print(f"# rope bvh leg: {row}", file=sys.stderr, flush=True)
    # many groups: the headline batch through the per-substep kernels
    from ropesim import scenes
    s = scenes.config3(n_ropes=4096, frames=frames + 1)
    fis = s.frame_inputs(frames + 1)
    row, pos = {"workload": "config3 batch (4096 ropes, contact groups of 4) through the per-substep kernels"}, {}
    for mode in ("scan", "bvh"):
        b = s.build(rs)
        b.set_dispatch(0)
        b.rope_broad_phase = mode
        b.set_graph(True)
        s.step(b, fis[0], lambda a: torch.as_tensor(a, device="cuda"))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(1, frames + 1):
            s.step(b, fis[f], lambda a: torch.as_tensor(a, device="cuda"))
        torch.cuda.synchronize()
        row[f"{mode}_ms_per_frame"] = 1e3 * (time.perf_counter() - t0) / frames
        pos[mode] = torch.as_tensor(b.positions(), device="cuda").cpu().numpy()
    row["bvh_speedup"] = row["scan_ms_per_frame"] / row["bvh_ms_per_frame"]
    row["bitwise_equal"] = bool(np.array_equal(pos["scan"].view(np.uint32), pos["bvh"].view(np.uint32)))
    out["groups"] = row
    return out


def differentiation_bench(n_ropes, frames):
    """Config-4 workload: forward throughput of the plain per-substep kernels against the persistent frame
    kernel, and the differentiable rollout (forward with per-frame checkpoints, then the adjoint over every
    frame) with its backward/forward time ratio."""
    import numpy as np
    import torch

    import ropesim
    from ropesim import scenes
    s = scenes.config4(n_ropes=n_ropes, frames=frames)
    b = s.build(ropesim)
    fis = s.frame_inputs(frames)
    dev = lambda k: torch.tensor(np.stack([f[k] for f in fis]), device="cuda")
    P0, V0, P1 = dev("p0"), dev("v0"), dev("p1")
    units = s.num_vertices * s.substeps * frames
    all_envs = np.arange(s.num_ropes, dtype=np.int32)

    def forward_frames(mode, variant=-1):
        b.reset_envs(all_envs)
        b.set_dispatch(mode, variant)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(frames):
            b.step(P0[f], V0[f], s.dt, s.substeps, root_pos_end=P1[f])
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    forward_frames(0)  # warm-up (module load, first launches)
    t_plain = forward_frames(0)
    b.set_graph(True)  # the same per-substep kernels replayed from a captured hipGraph
    forward_frames(0)
    t_plain_graph = forward_frames(0)
    b.set_graph(False)
    b.set_dispatch(1, -1)
    b.autotune(s.dt, s.substeps, True)
    variant = b.dispatch
    forward_frames(1, variant[1])  # the tuned variant, launched directly
    t_persist = forward_frames(1, variant[1])
    b.set_graph(True)  # the batching form: the tuned persistent kernel with the frame captured as a hipGraph
    forward_frames(1, variant[1])
    t_persist_graph = forward_frames(1, variant[1])
    b.set_graph(False)
    b.set_dispatch(*variant)

    w = np.asarray(b.inv_masses(), np.float64).reshape(-1)
    mass = torch.tensor(np.where(w > 0, 1.0 / np.where(w > 0, w, 1.0), 1.0), dtype=torch.float32, device="cuda")
    ea = torch.full((s.num_vertices,), s.k_stretch, device="cuda", requires_grad=True)
    ei = torch.full((s.num_vertices,), s.k_bend, device="cuda", requires_grad=True)
    mass.requires_grad_(True)
    drag = torch.tensor(s.drag, device="cuda", requires_grad=True)
    p1 = P1.clone().requires_grad_(True)
    tips = torch.as_tensor(s.offsets[1:] - 1, device="cuda", dtype=torch.long)
    # target trajectory: the same solver with hidden per-vertex EA / EI x U(0.5, 2); gradients are taken at the
    # nominal material
    ea_h, ei_h = (torch.as_tensor(a, device="cuda") for a in s.hidden_material())
    with torch.no_grad():
        target = ropesim.diff.rollout(b, ea_h, ei_h, mass.detach(), float(s.drag), P0, V0, P1, s.dt,
                                      s.substeps)[:, tips].clone()
    fwds, bwds = [], []
    for rep in range(5):  # two warm-ups (first-use allocations of torch's caching allocator), three timed
        for t in (ea, ei, mass, drag, p1):
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = ropesim.diff.rollout(b, ea, ei, mass, drag, P0, V0, p1, s.dt, s.substeps)
        loss = ((x[:, tips] - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if rep >= 2:
            fwds.append(t1 - t0)
            bwds.append(t2 - t1)
    assert all(bool(torch.isfinite(g).all()) for g in (ea.grad, ei.grad, mass.grad, drag.grad, p1.grad))
    fwd, bwd = float(np.median(fwds)), float(np.median(bwds))
    return {"workload": f"config4: {n_ropes} ropes x 65 vertices (config-1 material), plane + one static sphere SDF "
                        f"(48^3) per rope, root program, {frames} frames x {s.substeps} substeps x {s.num_itr} "
                        "iterations, tips tracking a hidden-material trajectory",
            "frames": frames, "vertex_substeps": units,
            "plain_kernels_ms_per_frame": 1e3 * t_plain / frames,
            "plain_kernels_graph_ms_per_frame": 1e3 * t_plain_graph / frames,
            "persistent_kernel_ms_per_frame": 1e3 * t_persist / frames,
            "persistent_kernel_graph_ms_per_frame": 1e3 * t_persist_graph / frames,
            "persistent_speedup_vs_plain": t_plain / t_persist,
            "persistent_graph_speedup_vs_plain": t_plain / t_persist_graph,
            "persistent_dispatch": list(variant),
            "rollout_forward_ms_per_frame": 1e3 * fwd / frames,
            "rollout_forward_vertex_substeps_per_s": units / fwd,
            "adjoint_ms_per_frame": 1e3 * bwd / frames,
            "adjoint_forward_ratio": bwd / fwd,
            "adjoint_vs_persistent_forward_ratio": bwd / t_persist,
            "loss": "sum over frames of squared tip distances to the trajectory of hidden per-vertex EA / EI x U(0.5, 2)",
            "timing": "median of 3 rollouts after 2 warm-ups"}


def scale_bench(frames, adj_frames):
    """Config-5 workload: forward throughput of the graph-captured persistent frame kernel with everything on,
    then a differentiable rollout over adj_frames frames of the same geometry with a tip-tracking loss, with
    rope-rope contact (the stated workload) and without it; device memory in use at the peak."""
    import numpy as np
    import torch

    import ropesim
    from ropesim import scenes
    t0 = time.perf_counter()
    s = scenes.config5(frames=max(frames, adj_frames) + 1)
    free0 = torch.cuda.mem_get_info()[0]
    low = [free0]
    b = s.build(ropesim)
    fis = s.frame_inputs(max(frames, adj_frames) + 1)
    dev = [{k: torch.as_tensor(v, device="cuda") for k, v in fi.items() if v is not None} for fi in fis]
    b.set_dispatch(1, -1)
    b.autotune(s.dt, s.substeps, True)
    b.set_graph(True)
    setup = time.perf_counter() - t0
    s.step(b, dev[0])  # warm-up: captures the graph
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for k in range(1, frames + 1):
        s.step(b, dev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    low.append(torch.cuda.mem_get_info()[0])
    assert bool(torch.isfinite(torch.as_tensor(b.positions(), device="cuda")).all())
    units = s.num_vertices * s.substeps
    out = {"workload": f"config5: {s.num_ropes} ropes x 96 vertices (config-1 material, +-20 % per-vertex EA/EI/mass), "
                       f"groups of 4 with rope-rope contact, {len(s.meshes)} static SDF meshes per environment "
                       "(96^3 sphere/box/capsule/torus lattices), plane, root program, 100 substeps x 8 iterations",
           "vertices": s.num_vertices, "dispatch": list(b.dispatch), "graph_captures": b.graph_captures,
           "forward_ms_per_frame": 1e3 * wall / frames, "forward_vertex_substeps_per_s": units * frames / wall,
           "setup_s": round(setup, 1)}
    del b
    # adjoint legs: same geometry and material, rope-rope contact on (the stated workload), then off
    rc = s.rope_contact
    out.update(scale_adjoint(s, dev, adj_frames, low, free0))
    s.rope_contact = None
    off = scale_adjoint(s, dev, adj_frames, low, free0)
    s.rope_contact = rc
    out["without_rope_contact"] = {k: off[k] for k in ("adjoint_workload", "rollout_forward_ms_per_frame",
                                                        "adjoint_ms_per_frame", "adjoint_forward_ratio",
                                                        "nonfinite_gradient_entries")}
    out["device_memory_in_use_peak_gb"] = (free0 - min(low)) / 1e9
    return out


def scale_adjoint(s, dev, F, low, free0):
    """A differentiable rollout over F frames of the config-5 scene with a tip-tracking loss, timed after a
    one-frame warm-up rollout, and its adjoint."""
    import numpy as np
    import torch

    import ropesim
    b = s.build(ropesim)
    b.set_dispatch(1, -1)
    b.autotune(s.dt, s.substeps, True)
    P0, V0, P1 = (torch.stack([dev[k][n] for k in range(F)]) for n in ("p0", "v0", "p1"))
    ea = torch.tensor(s.ea, device="cuda", requires_grad=True)
    ei = torch.tensor(s.ei, device="cuda", requires_grad=True)
    mass = torch.tensor(s.masses, device="cuda", requires_grad=True)
    drag = torch.tensor(0.0, device="cuda", requires_grad=True)
    p1 = P1.clone().requires_grad_(True)
    tips = torch.as_tensor(s.offsets[1:] - 1, device="cuda", dtype=torch.long)
    target = torch.as_tensor(s.positions[s.offsets[1:] - 1], device="cuda")[None] + 0.05
    # warm-up: a one-frame rollout and its adjoint allocate the frame history (its size does not depend on
    # the frame count), so the timed rollout measures the steady state, as the config-4 leg does
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    xw = ropesim.diff.rollout(b, ea, ei, mass, drag, P0[:1], V0[:1], p1[:1], s.dt, s.substeps)
    torch.cuda.synchronize()
    w1 = time.perf_counter()
    ((xw[:, tips] - target) ** 2).sum().backward()
    torch.cuda.synchronize()
    w2 = time.perf_counter()
    low.append(torch.cuda.mem_get_info()[0])
    del xw
    for g in (ea, ei, mass, drag, p1):
        g.grad = None
    b.reset_envs(np.arange(b.num_envs, dtype=np.int32))  # back to the construction state
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    x = ropesim.diff.rollout(b, ea, ei, mass, drag, P0, V0, p1, s.dt, s.substeps)
    loss = ((x[:, tips] - target) ** 2).sum()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    low.append(torch.cuda.mem_get_info()[0])
    loss.backward()
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    low.append(torch.cuda.mem_get_info()[0])
    nonfinite = {k: int((~torch.isfinite(g.grad)).sum()) for k, g in dict(ea=ea, ei=ei, mass=mass, p1=p1).items()}
    del b
    contact = "on" if s.rope_contact is not None else "off"
    return {"adjoint_workload": f"{F} frames, rope-rope contact {contact}, tip-tracking loss",
            "nonfinite_gradient_entries": nonfinite,
            "rollout_forward_ms_per_frame": 1e3 * (t3 - t2) / F, "adjoint_ms_per_frame": 1e3 * (t4 - t3) / F,
            "adjoint_forward_ratio": (t4 - t3) / (t3 - t2),
            "warmup_one_frame_rollout_ms": 1e3 * (w1 - w0), "warmup_one_frame_adjoint_ms": 1e3 * (w2 - w1),
            "timing": "one rollout after a one-frame warm-up rollout (which allocates the frame history)"}


def timed_window(scene, b, dev, W, K, kernel_timing=False):
    """W untimed frames, then K timed frames back to back. Each frame with rope-rope contact returns after its
    error word is read (one host synchronisation per frame, in the reference too), so the host clock between
    frames gives per-frame times without adding a synchronisation. Returns the wall time of the K frames, the
    per-frame times in ms and, with kernel_timing, the frame kernels' summed HIP-event time and launch count."""
    import torch
    for k in range(W):
        scene.step(b, dev[k], lambda a: a)
    torch.cuda.synchronize()
    if kernel_timing:
        b.set_kernel_timing(True)
        b.kernel_time()
    marks = [time.perf_counter()]
    for k in range(W, W + K):
        scene.step(b, dev[k], lambda a: a)
        marks.append(time.perf_counter())
    torch.cuda.synchronize()
    marks[-1] = time.perf_counter()
    kern = b.kernel_time() if kernel_timing else (0.0, 0)
    if kernel_timing:
        b.set_kernel_timing(False)
    x = torch.as_tensor(b.positions(), device="cuda")
    assert bool(torch.isfinite(x).all()), "non-finite state"
    ms = [1e3 * (b_ - a_) for a_, b_ in zip(marks[:-1], marks[1:])]
    return marks[-1] - marks[0], ms, kern


def progress(msg):
    """One line on stderr as each leg of the run ends (the JSON line goes to stdout at the very end)."""
    print(f"# {time.strftime('%H:%M:%S')} {msg}", file=sys.stderr, flush=True)


def series(ms):
    return {"first_ms": ms[0], "last_ms": ms[-1], "min_ms": min(ms), "max_ms": max(ms),
            "ms": [round(v, 3) for v in ms]}


def build_headline(args, frames, fast):
    import torch

    import ropesim
    from ropesim import scenes
    s = scenes.config3(n_ropes=args.ropes, frames=frames)
    drop = {w for w in args.without.split(",") if w}
    unknown = drop - {"rope_contact", "meshes", "mesh_motion", "planes"}
    if unknown:
        sys.exit(f"--without: unknown feature(s) {sorted(unknown)}")
    if "rope_contact" in drop:
        s.rope_contact = None
    if "meshes" in drop:
        s.meshes, s.mesh_pos0, s.mesh_quat0 = [], None, None
    if "planes" in drop:
        s.planes = None
    b = s.build(ropesim)
    b.fast_math = fast
    fis = s.frame_inputs(frames)
    if "mesh_motion" in drop:
        fis = [{k: v for k, v in fi.items() if not k.startswith("mesh_")} for fi in fis]
    dev = [{k: torch.as_tensor(v, device="cuda") for k, v in fi.items() if v is not None} for fi in fis]
    torch.cuda.synchronize()
    tuned_ms = None
    b.set_dispatch(args.mode, args.variant if args.mode == 1 else -1)
    if args.mode == 1 and args.variant < 0:
        tuned_ms = b.autotune(s.dt, s.substeps, True)
    b.set_graph(not args.no_graph)  # the frame's kernel sequence is captured once and replayed
    return s, b, dev, tuned_ms


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ropes", type=int, default=4096)
    ap.add_argument("--mode", type=int, default=1, help="0: per-substep kernels, 1: persistent frame kernel")
    ap.add_argument("--variant", type=int, default=-1, help="frame-kernel variant (-1: autotune)")
    ap.add_argument("--no-graph", action="store_true", help="launch each frame's kernels directly instead of "
                    "replaying the frame captured as a hipGraph")
    ap.add_argument("--ieee", action="store_true",
                    help="headline in correctly rounded, unfused arithmetic (the float32 oracle's) instead of the "
                         "default fast arithmetic (fused multiply-add, hardware reciprocal / square root / log)")
    ap.add_argument("--no-ieee-leg", action="store_true", help="skip the headline window rerun in the other arithmetic")
    ap.add_argument("--cpu-sample-ropes", type=int, default=512, help="config-3 ropes timed on the CPU oracle")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the config-1 / config-2 plain-method legs")
    ap.add_argument("--diff-ropes", type=int, default=256, help="config-4 ropes of the differentiation leg")
    ap.add_argument("--diff-frames", type=int, default=120, help="config-4 frames (0 skips the leg)")
    ap.add_argument("--scale-frames", type=int, default=3, help="config-5 forward frames timed (0 skips the leg)")
    ap.add_argument("--scale-adjoint-frames", type=int, default=30, help="config-5 frames of the adjoint leg")
    ap.add_argument("--no-readout", action="store_true", help="skip the contact-force readout leg")
    ap.add_argument("--no-bake", action="store_true", help="skip the mesh-distance-grid leg")
    ap.add_argument("--bvh-sizes", default="1024,8192,32768",
                    help="vertex counts of the long-rope self-contact leg (scan vs hierarchy; empty skips it)")
    ap.add_argument("--without", default="",
                    help="comma-separated features to leave out of the headline workload, for cost breakdowns: "
                         "rope_contact, meshes, mesh_motion (meshes stay where they start), planes")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: the rope batch runs on a single device (across GPUs the batch would "
                 "shard by environment, one contiguous range of environments per GPU); run with --gpus 1")

    import torch

    t_setup = time.perf_counter()
    K, W = args.steps, args.warmup
    RK = 0 if args.no_readout or K < 2 else min(K, 10)  # frames per readout mode (one warm-up), after the headline
    s, b, dev, tuned_ms = build_headline(args, K + W + 2 * RK, not args.ieee)
    setup_s = time.perf_counter() - t_setup

    if args.copy_calibration_mb > 0:  # a plain copy of known size: reads and writes n bytes each
        n = args.copy_calibration_mb << 20
        src = torch.ones(n // 4, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    wall, frame_ms, (kern_ms, launches) = timed_window(s, b, dev, W, K, kernel_timing=True)
    progress(f"headline: {1e3 * wall / K:.2f} ms per frame over {K} frames")

    readout = None
    if RK > 0:  # the same batch continues with the contact-force readout on: float sums, then fixed-point sums
        readout = {"workload": "the headline batch continued with enable_contact_forces(True): per-vertex forces "
                               "and per-(rope, mesh) wrenches averaged over each frame", "frames": RK,
                   "off_ms_per_frame": 1e3 * wall / K}
        k = W + K
        for name, det in (("float", False), ("fixed_point", True)):
            b.enable_contact_forces(True, deterministic=det)
            s.step(b, dev[k], lambda a: a)  # recapture with the readout compiled in
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for j in range(k + 1, k + RK):
                s.step(b, dev[j], lambda a: a)
            torch.cuda.synchronize()
            readout[name + "_ms_per_frame"] = 1e3 * (time.perf_counter() - t1) / (RK - 1)
            k += RK
        b.enable_contact_forces(False)
        for name in ("float", "fixed_point"):
            readout[name + "_overhead"] = readout[name + "_ms_per_frame"] / readout["off_ms_per_frame"] - 1.0
        readout["resting"] = resting_readout_bench(args.ropes)
        progress("readout leg done")
    dispatch, captures = list(b.dispatch), b.graph_captures
    del b

    units = s.num_vertices * s.substeps          # vertex-substeps per frame
    value = units * K / wall
    arith = "ieee" if args.ieee else "fast"
    line = {
        "metric": "vertex-substeps/s (4096 uneven ropes, contact on)",
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"config3: {args.ropes} uneven ropes (V~U{{16..256}}, sum V={s.num_vertices}), "
                               f"groups of 4 with rope-rope contact, 8 SDF meshes/env (64^3), plane, root program, "
                               f"{s.substeps} substeps x {s.num_itr} iterations per 1/60 s frame",
                   "ropes": args.ropes, "vertices": s.num_vertices, "substeps": s.substeps, "iterations": s.num_itr,
                   "dispatch": dispatch, "arithmetic": arith, "graph_captures": captures,
                   "autotune_ms": tuned_ms, "setup_s": round(setup_s, 1),
                   "timed_frames": f"frames {W}..{W + K - 1} of the run (frames slow as the ropes tangle)",
                   "without": args.without or None},
        "frame_times": series(frame_ms),
    }
    if launches > 0:
        per_launch_s = kern_ms / launches * 1e-3
        achieved = B_ALG * units / per_launch_s / 1e9
        default = args.ropes == 4096 and args.mode == 1 and args.variant < 0 and not args.ieee
        traffic = profiled_traffic() if default else None
        line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
                            "kernel": "k_cell (persistent frame kernel; the timed span includes the two per-frame precompute kernels "
                                      "launched ahead of it, < 1 % of it)", "kernel_ms": kern_ms / launches}
        fl = flops_per_vertex_substep(s.num_itr, planes=0 if s.planes is None else len(s.planes),
                                      meshes=len(s.meshes or ()))
        tf = fl * units / per_launch_s / 1e12
        line["valu"] = {"achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tf / FP32_PEAK_TFLOPS,
                        "flop_per_vertex_substep": fl,
                        "note": "flops counted on the oracle (oracle/opcount.cpp): solver, plane tests and mesh broad "
                                "phase; mesh-candidate and rope-rope pair terms excluded"}
    if not args.no_ieee_leg:
        other = "fast" if args.ieee else "ieee"
        s2, b2, dev2, _ = build_headline(args, K + W, args.ieee)
        wall2, ms2, _ = timed_window(s2, b2, dev2, W, K)
        line["other_arithmetic"] = {"arithmetic": other, "value": units * K / wall2, "ms_per_step": 1e3 * wall2 / K,
                                    "frame_times": series(ms2),
                                    "note": "the same workload and frame window in the other arithmetic mode: ieee "
                                            "is every operation correctly rounded and unfused (the float32 oracle's "
                                            "operations); fast is fused multiply-add with the hardware reciprocal, "
                                            "square root and log"}
        del b2
        progress(f"other arithmetic ({other}): {1e3 * wall2 / K:.2f} ms per frame")
    if readout is not None:
        line["readout"] = readout
    if not args.no_plain:
        line["plain"] = plain_legs()
        progress("plain-method legs done")
    if not args.no_bake:
        line["bake_mesh"] = bake_bench()
        progress("mesh bake leg done")
    if args.bvh_sizes:
        line["rope_bvh"] = rope_bvh_bench([int(v) for v in args.bvh_sizes.split(",")])
        progress("rope hierarchy leg done")
    if args.diff_frames > 0:
        d = differentiation_bench(args.diff_ropes, args.diff_frames)
        line["differentiation"] = d
        line["adjoint_forward_ratio"] = d["adjoint_forward_ratio"]
        progress(f"differentiation leg done: adjoint/forward {d['adjoint_forward_ratio']:.2f}")
    if args.scale_frames > 0:
        try:
            line["scale"] = scale_bench(args.scale_frames, args.scale_adjoint_frames)
        except Exception as e:  # the headline line must still print
            line["scale"] = {"error": f"{type(e).__name__}: {e}"}
        progress("scale leg done")
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_ropes, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
