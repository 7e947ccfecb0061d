"""Triangle-mesh distance grids: the float64 checker (oracle/mesh_sdf.py) pinned by closed forms, the host
side of ropesim.sdf.bake_mesh (validation, lattice, is_closed), and on a GPU the HIP baker against the checker.

Reference: ropesim/sdf.py:61-170 of the reference (bake_mesh through Warp, is_closed, the lattice). Warp is
absent here; the checker restates the exact closest-point and generalized-winding-number methods the Warp
query computes (its header names them), and these tests pin it:
- a box mesh gives sdf_box's exact distance inside and outside (every region of the closest-point routine
  is visited: faces, edges and corners of the box);
- an inscribed icosphere lies between the sphere's distance and the sphere's distance plus the faceting sag;
- the winding number of a closed mesh is 1 inside and 0 outside, in either orientation (|w|);
- the closest-point distance to random triangles agrees with dense barycentric sampling;
- an open square baked with a thickness is |z| - t/2 over its interior.
"""
import numpy as np
import pytest

from oracle import mesh_sdf as M


def test_closest_point_matches_dense_sampling():
    rng = np.random.default_rng(0)
    tris = rng.normal(size=(40, 3, 3))
    pts = rng.normal(size=(200, 3)) * 2
    d = M.closest_point_distance(pts, tris)
    n = 120
    u, v = np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), indexing="ij")
    keep = u + v <= 1
    u, v = u[keep], v[keep]
    for t in range(len(tris)):
        a, b, c = tris[t]
        s = a[None] + u[:, None] * (b - a)[None] + v[:, None] * (c - a)[None]
        dense = np.linalg.norm(pts[:, None, :] - s[None], axis=-1).min(1)
        res = max(np.linalg.norm(b - a), np.linalg.norm(c - a), np.linalg.norm(c - b)) / n
        assert (d[:, t] <= dense + 1e-12).all()
        assert (dense - d[:, t] <= res).all()


def test_box_mesh_equals_sdf_box():
    from ropesim import sdf
    half = np.array([0.3, 0.2, 0.1])
    v, t = M.box_mesh(half, centre=(0.05, -0.02, 0.01))
    origin, spacing, shape = np.array([-0.45, -0.35, -0.25], np.float32), 0.0371, (25, 19, 14)
    got = M.bake(v, t, origin, spacing, shape)
    pts = M.lattice_points(origin, spacing, shape).astype(np.float64)
    want = sdf.sdf_box(half, (0.05, -0.02, 0.01))(pts).reshape(shape)
    assert (want < 0).sum() > 100 and (want > 0).sum() > 100
    np.testing.assert_allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("flip", [False, True])
def test_icosphere_within_faceting_of_sphere(flip):
    r = 0.1
    v, t = M.icosphere(r, 3)
    if flip:
        t = t[:, ::-1]
    origin, spacing, shape = np.full(3, -0.15, np.float32), 0.3 / 22, (23, 23, 23)
    got = M.bake(v, t, origin, spacing, shape)
    pts = M.lattice_points(origin, spacing, shape).astype(np.float64)
    sphere = (np.linalg.norm(pts, axis=1) - r).reshape(shape)
    tri = v[t]
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    inner = np.abs((nrm * tri[:, 0]).sum(1)) / np.linalg.norm(nrm, axis=1)
    sag = r - inner.min()
    assert 0 < sag < 2e-3
    assert (got - sphere >= -1e-12).all()
    assert (got - sphere <= sag + 1e-12).all()
    assert (got[sphere < -sag] < 0).all() and (got[sphere > 0] > 0).all()


def test_winding_number_of_closed_mesh():
    rng = np.random.default_rng(1)
    v, t = M.icosphere(0.5, 2)
    p = rng.uniform(-1, 1, (500, 3))
    rr = np.linalg.norm(p, axis=1)
    p = p[(rr < 0.45) | (rr > 0.55)]
    w = M.winding_number(p, v[t])
    inside = np.linalg.norm(p, axis=1) < 0.45
    np.testing.assert_allclose(w[inside], 1.0, atol=1e-9)
    np.testing.assert_allclose(w[~inside], 0.0, atol=1e-9)
    np.testing.assert_allclose(M.winding_number(p, v[t[:, ::-1]]), -w, atol=1e-12)


def test_open_square_shell():
    v = np.array([[-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]], np.float64)
    t = np.array([[0, 1, 2], [0, 2, 3]])
    origin, spacing, shape = np.array([-0.4, -0.4, -0.1], np.float32), 0.05, (17, 17, 5)
    got = M.bake(v, t, origin, spacing, shape, thickness=0.1)
    pts = M.lattice_points(origin, spacing, shape).astype(np.float64)
    np.testing.assert_allclose(got.reshape(-1), np.abs(pts[:, 2]) - 0.05, atol=1e-12)


# ------------------------------------------------------------------------- host side of bake_mesh

def test_is_closed():
    from ropesim import sdf
    v, t = M.box_mesh([1, 1, 1])
    assert sdf.is_closed(t)
    assert not sdf.is_closed(t[:-1])
    v2, t2 = M.icosphere(1.0, 1)
    assert sdf.is_closed(t2)
    assert not sdf.is_closed(np.array([[0, 1, 2], [0, 2, 3]]))


def test_bake_mesh_validation_is_host_side():
    """The reference's argument checks (ropesim/sdf.py:125-144) raise before any device work."""
    from ropesim import sdf
    v, t = M.box_mesh([0.1, 0.1, 0.1])
    with pytest.raises(ValueError, match="triangle mesh"):
        sdf.bake_mesh(v[:2], t[:0], 0.01, 0.02)
    with pytest.raises(ValueError, match="out of range"):
        sdf.bake_mesh(v, t + 1, 0.01, 0.02)
    with pytest.raises(ValueError, match="open mesh"):
        sdf.bake_mesh(v, t[:-1], 0.01, 0.02)
    with pytest.raises(ValueError, match="thickness"):
        sdf.bake_mesh(v, t[:-1], 0.01, 0.02, thickness=0.015)
    with pytest.raises(ValueError, match="round"):
        sdf.bake_mesh(v, t, 0.01, 0.02, round=-1.0)


def test_bake_lattice_matches_reference_padding():
    """Padded box: margin + thickness / 2 + round beyond the vertices, dims ceil((hi - lo) / spacing) + 1."""
    from ropesim import sdf
    v, t = M.box_mesh([0.1, 0.2, 0.05])
    origin, dims = sdf.bake_lattice(v, 0.0123, 0.031, thickness=0.004, round=0.005)
    np.testing.assert_allclose(origin, v.min(0) - 0.038, atol=1e-7)
    assert tuple(dims) == tuple(np.maximum(np.ceil((v.max(0) - v.min(0) + 0.076) / 0.0123).astype(int) + 1, 2))
    assert sdf.bake_lattice(np.zeros((3, 3)), 0.01, 0.0)[1].tolist() == [2, 2, 2]


# ------------------------------------------------------------------------- the HIP baker

MESHES = {
    "box": lambda: M.box_mesh([0.07, 0.05, 0.03], centre=(0.01, 0.0, -0.005)),
    "icosphere": lambda: M.icosphere(0.06, 3),
    "icosphere-flipped": lambda: (lambda v, t: (v, t[:, ::-1].copy()))(*M.icosphere(0.05, 2, centre=(0.01, 0.02, 0))),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(MESHES))
def test_device_bake_matches_checker(rs, name):
    from ropesim import sdf
    v, t = MESHES[name]()
    g = sdf.bake_mesh(v, t, 0.004, 0.01)
    want = M.bake(v, t, g.origin, g.spacing, g.values.shape)
    assert g.values.dtype == np.float32
    assert (want < 0).sum() > 50 and (want > 0).sum() > 50
    np.testing.assert_allclose(g.values, want, atol=2e-6)


@pytest.mark.gpu
def test_device_bake_shell_and_round(rs):
    from ropesim import sdf
    v = np.array([[-0.1, -0.1, 0], [0.1, -0.1, 0.02], [0.1, 0.1, 0], [-0.1, 0.1, -0.01]], np.float64)
    t = np.array([[0, 1, 2], [0, 2, 3]])
    g = sdf.bake_mesh(v, t, 0.005, 0.01, thickness=0.012, round=0.002)
    want = M.bake(v, t, g.origin, g.spacing, g.values.shape, thickness=0.012, round=0.002)
    np.testing.assert_allclose(g.values, want, atol=2e-6)
    vb, tb = M.box_mesh([0.05, 0.05, 0.05])
    g = sdf.bake_mesh(vb, tb, 0.005, 0.01, thickness=0.02)  # hollow: only a shell of the closed mesh is solid
    want = M.bake(vb, tb, g.origin, g.spacing, g.values.shape, thickness=0.02)
    np.testing.assert_allclose(g.values, want, atol=2e-6)
    assert g.values[g.values.shape[0] // 2, g.values.shape[1] // 2, g.values.shape[2] // 2] > 0


@pytest.mark.gpu
def test_rope_rests_on_baked_box_as_on_closed_form(rs):
    """A rope dropped on a baked box mesh lands as on the closed-form box grid of the same lattice."""
    from ropesim import sdf
    half = np.array([0.3, 0.1, 0.05])
    v, t = M.box_mesh(half, centre=(0, 0, -0.05))
    g = sdf.bake_mesh(v, t, 1 / 128, 0.02)
    ref = sdf.from_function(sdf.sdf_box(half, (0, 0, -0.05)), g.origin,
                            g.origin + g.spacing * (np.array(g.values.shape) - 1), g.spacing)
    assert ref.values.shape == g.values.shape
    np.testing.assert_allclose(g.values, ref.values, atol=2e-6)
    out = []
    for grid in (g, ref):
        pos = np.zeros((1, 20, 3), np.float32)
        pos[0, :, 0] = np.arange(20) * 0.02 - 0.19
        pos[0, :, 2] = 0.01
        b = rs.RopeBatch.from_positions(pos, 0.1, 500.0, 1e-4, 1 / 3000, 8, pinned=np.zeros((1, 20), bool))
        b.set_gravity(np.array([0, 0, -9.81], np.float32))
        b.set_contact(0.004, mu_s=0.5, mu_k=0.4)
        b.add_mesh(*grid.as_args())
        for _ in range(30):
            b.step_free(1 / 60, 50)
        import torch
        out.append(torch.as_tensor(b.positions(), device="cuda").cpu().numpy())
    assert abs(out[0][0, :, 2].mean() - 0.004) < 5e-4
    np.testing.assert_allclose(out[0], out[1], atol=1e-4)
