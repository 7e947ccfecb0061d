"""The flop constants behind bench.py's VALU figure are counts, not estimates: oracle/opcount.cpp runs the
oracle's substep on an operation-counting scalar, phase by phase, and this test recounts them."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

KEYS = {
    "init_flop_per_vertex_substep": "init_per_vertex_substep",
    "position_sweep_flop_per_vertex_iteration": "position_sweep_per_vertex_iteration",
    "orientation_sweep_flop_per_vertex_iteration": "orientation_sweep_per_vertex_iteration",
    "end_flop_per_vertex_substep": "end_per_vertex_substep",
    "plane_flop_per_free_vertex_iteration": "plane_per_free_vertex_iteration",
    "mesh_broad_phase_flop_per_vertex_substep": "mesh_broad_phase_per_vertex_substep_per_mesh",
    "flop_per_mesh_candidate_iteration": "mesh_candidate_per_iteration",
    "flop_per_pair_test": "pair_test",
}


@pytest.fixture(scope="module")
def counts():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_build/opcount"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_build", "opcount")], check=True, capture_output=True,
                         text=True, timeout=600).stdout
    return {k.strip(): float(v) for k, v in (l.split("=") for l in out.splitlines() if "=" in l)}


def test_bench_constants_are_the_counts(counts):
    import bench
    for key, name in KEYS.items():
        assert abs(bench.FLOP[name] - counts[key]) < 1e-3, (name, bench.FLOP[name], counts[key])


def test_counted_scenes_exercise_every_term(counts):
    """Every term was actually evaluated: the pair scene tests every pair of its groups, the mesh scene has
    one candidate per free vertex and substep, and the per-unit figures are in a plausible range."""
    assert counts["pair_tests_per_substep"] > 1e6
    assert counts["mesh_candidates_per_iteration_per_frame"] == (counts["vertices"] - 64) * counts["substeps"]
    assert 2500 < counts["solver_flop_per_vertex_substep_at_8_iterations"] < 4500
    assert 50 < counts["flop_per_pair_test"] < 200 and 50 < counts["flop_per_mesh_candidate_iteration"] < 300
