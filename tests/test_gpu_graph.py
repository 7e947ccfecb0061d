"""Graph form of a frame (set_graph): the captured kernel sequence, replayed frame after frame with eager work
and synchronisations in between, gives bit-for-bit the positions, velocities and frames of eager launches."""
import numpy as np
import pytest

from conftest import to_dev

pytestmark = pytest.mark.gpu


def _state(b):
    import torch
    return [torch.as_tensor(v, device="cuda").clone() for v in (b.positions(), b.velocities(), b.quaternions())]


@pytest.mark.parametrize("which", ["config2", "config3"])
@pytest.mark.parametrize("mode", [0, 1])
def test_graph_replays_match_eager_launches(rs, which, mode):
    import torch
    from ropesim import scenes
    F = 5
    s = scenes.config2(n_ropes=16, V=33, frames=F) if which == "config2" else \
        scenes.config3(n_ropes=8, frames=F, vmax=64, grid_n=32)
    eager = s.build(rs)
    graph = s.build(rs)
    eager.set_dispatch(mode)
    graph.set_dispatch(mode)
    graph.set_graph(True)
    fis = s.frame_inputs(F)
    for f in range(F):
        s.step(eager, fis[f], to_dev)
        s.step(graph, fis[f], to_dev)
        a, g = _state(eager), _state(graph)
        scratch = torch.randn(1 << 16, device="cuda")  # unrelated eager work between replays
        float((scratch * 2).sum())
        torch.cuda.synchronize()
        for x, y in zip(a, g):
            assert torch.equal(x, y), f"frame {f}"
    # one capture for the first frame (the previous substep length differs), at most one more, then replays
    assert 1 <= graph.graph_captures <= 2, graph.graph_captures


def test_graph_recaptures_when_the_frame_changes(rs):
    """A different substep count is a different kernel sequence: captured again, still equal to eager."""
    import torch
    from ropesim import scenes
    s = scenes.config2(n_ropes=4, V=17, frames=4)
    eager, graph = s.build(rs), s.build(rs)
    graph.set_graph(True)
    fis = s.frame_inputs(4)
    for f, S in enumerate((100, 100, 50, 100)):
        for b in (eager, graph):
            b.step(to_dev(fis[f]["p0"]), to_dev(fis[f]["v0"]), s.dt, S, root_pos_end=to_dev(fis[f]["p1"]))
        assert torch.equal(_state(eager)[0], _state(graph)[0])
    assert graph.graph_captures >= 3
