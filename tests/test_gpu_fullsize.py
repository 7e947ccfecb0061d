"""The batching-form workloads at their full sizes, through a property that does not depend on size: the
persistent frame kernel (dispatch mode 1, autotuned cell variant) and the per-substep kernels (mode 0, whose
rope-rope detection runs through the per-group hierarchy at these group sizes) give bit-identical states.
The oracle parity of the same method is established on the smaller cases of test_gpu_parity.py.

- config 3: 4096 uneven ropes (sum V 561,616), groups of 4 with rope-rope contact, 8 moving 64^3 meshes per
  environment, plane, root program, 2 frames of 100 substeps;
- config 5: 16384 ropes x 96 vertices, groups of 4, 64 static 96^3 mesh slots per environment, plane, root
  program, 1 frame of 100 substeps.
"""
import numpy as np
import pytest

from conftest import host, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which,frames", [("config3", 2), ("config5", 1)])
def test_full_size_dispatch_forms_bitwise(rs, which, frames):
    from ropesim import scenes
    s = scenes.config3(frames=frames) if which == "config3" else scenes.config5(frames=frames)
    fis = s.frame_inputs(frames)
    states = []
    for mode in (1, 0):
        b = s.build(rs)
        b.set_dispatch(mode)
        for fi in fis:
            s.step(b, fi, to_dev)
        assert b.last_error == 0
        if mode == 0:
            assert b.rope_bvh_active
        states.append([host(v).copy() for v in (b.positions(), b.velocities(), b.quaternions())])
        del b
    for a, c in zip(*states):
        assert np.isfinite(a).all()
        assert np.array_equal(a.view(np.uint32), c.view(np.uint32))
    x = states[0][0].reshape(-1, 3)
    assert x.shape[0] == s.num_vertices
