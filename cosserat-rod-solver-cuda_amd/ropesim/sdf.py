"""Signed-distance grids for rod-mesh contact.

Same grid format as the reference (ropesim/sdf.py:20-58): phi sampled on a node lattice in the
object's own frame, node (i, j, k) at origin + (i, j, k) * spacing, values[i, j, k] = signed distance
(negative inside), float32, C order. Pass ``grid.as_args()`` to ``add_mesh``. Contact happens where
phi < rope radius; outside the lattice there is no contact.

Closed forms (box, sphere, infinite cylinder, bowl; plus capsule and torus) are evaluated on the
lattice by ``from_function``. ``sample`` is the host trilinear interpolation of the solver in float64,
for checking a scene. ``bake_mesh`` turns a triangle mesh into a grid on the GPU (the HIP library's
tiled closest-point and generalized-winding-number kernel, where the reference uses Warp).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SdfGrid:
    values: np.ndarray   # float32 [nx, ny, nz], C order
    origin: np.ndarray   # float32 [3], object frame
    spacing: float

    @property
    def shape(self):
        return self.values.shape

    def nodes(self):
        """Lattice node coordinates [nx, ny, nz, 3] in the object frame."""
        ax = [self.origin[d] + self.spacing * np.arange(self.values.shape[d]) for d in range(3)]
        return np.stack(np.meshgrid(*ax, indexing="ij"), -1).astype(np.float32)

    def as_args(self):
        """(values, origin, spacing) for add_mesh()."""
        return self.values, self.origin, float(self.spacing)

    def sample(self, points):
        """Trilinear phi at object-frame points [M, 3] in float64 and a mask of the points inside the
        lattice (phi is nan outside, where the solver applies no contact)."""
        p = np.asarray(points, np.float64).reshape(-1, 3)
        u = (p - self.origin.astype(np.float64)) / self.spacing
        dims = np.array(self.values.shape)
        inside = (u >= 0).all(1) & (u < dims - 1).all(1)
        phi = np.full(len(p), np.nan)
        if inside.any():
            ui = u[inside]
            i = np.floor(ui).astype(int)
            f = ui - i
            v = self.values.astype(np.float64)

            def c(dx, dy, dz):
                return v[i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz]

            lx = lambda a, b: a + f[:, 0] * (b - a)
            c00, c01 = lx(c(0, 0, 0), c(1, 0, 0)), lx(c(0, 0, 1), c(1, 0, 1))
            c10, c11 = lx(c(0, 1, 0), c(1, 1, 0)), lx(c(0, 1, 1), c(1, 1, 1))
            c0 = c00 + f[:, 1] * (c10 - c00)
            c1 = c01 + f[:, 1] * (c11 - c01)
            phi[inside] = c0 + f[:, 2] * (c1 - c0)
        return phi, inside


def _lattice(lo, hi, spacing):
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    dims = np.maximum(np.ceil((hi - lo) / spacing).astype(int) + 1, 2)
    return lo.astype(np.float32), dims


def from_function(phi, lo, hi, spacing, chunk_nodes=2_000_000):
    """Sample a closed-form signed distance phi(points [M, 3] float64) -> [M] on the box [lo, hi]
    (x-slabs of about `chunk_nodes` nodes at a time)."""
    origin, dims = _lattice(lo, hi, spacing)
    g = SdfGrid(np.empty(tuple(dims), np.float32), origin, float(spacing))
    nx, ny, nz = dims
    ax = [origin[d].astype(np.float64) + spacing * np.arange(dims[d]) for d in range(3)]
    step = max(1, chunk_nodes // (ny * nz))
    for i0 in range(0, nx, step):
        i1 = min(nx, i0 + step)
        pts = np.stack(np.meshgrid(ax[0][i0:i1], ax[1], ax[2], indexing="ij"), -1).reshape(-1, 3)
        g.values[i0:i1] = np.asarray(phi(pts), np.float32).reshape(i1 - i0, ny, nz)
    return g


def is_closed(triangles):
    """True when every edge is shared by exactly two triangles (a closed, manifold surface)."""
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    if len(t) == 0:
        return False
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * (int(t.max()) + 1) + e[:, 1]
    _, counts = np.unique(key, return_counts=True)
    return bool((counts == 2).all())


def bake_lattice(vertices, spacing, margin, thickness=0.0, round=0.0):
    """Origin (float32) and node counts of the lattice bake_mesh samples: the vertices' bounding box padded by
    margin + thickness / 2 + round on every side (reference ropesim/sdf.py:145-148)."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3)
    pad = np.float32(margin + 0.5 * thickness + round)  # the box corners in float32, as the reference forms them
    return _lattice(v.min(0) - pad, v.max(0) + pad, spacing)


def bake_mesh(vertices, triangles, spacing, margin, thickness=0.0, round=0.0, device=None):
    """Distance grid of a triangle mesh on its padded bounding box, computed on the GPU.

    Same contract as the reference's bake_mesh (ropesim/sdf.py:96-157 there). vertices [V, 3], triangles
    [T, 3]. thickness == 0: SIGNED distance, negative inside, for a closed mesh (every edge shared by exactly
    two triangles); inside is where the generalized winding number exceeds 1/2 in magnitude, so either
    consistent orientation works. thickness > 0: an open mesh (sheet, heightfield) becomes a two-sided shell,
    |distance| - thickness / 2, and thickness must be >= 2 * spacing so the interpolated distance has a
    gradient at both faces; on a closed mesh it hollows the solid. round > 0 offsets the surface outward
    (phi - round), rounding convex edges with that radius. margin pads the box on every side beyond the shell
    and rounding. device: the HIP device index (None: the current one).

    Every lattice node is tested against every triangle (the kernel stages triangles through LDS), so the cost
    is nodes x triangles: meshes of up to ~1e5 triangles on ~1e6-node lattices take well under a second.
    """
    vertices = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 3))
    triangles = np.ascontiguousarray(np.asarray(triangles, np.int32).reshape(-1, 3))
    if len(vertices) < 3 or len(triangles) < 1:
        raise ValueError("bake_mesh needs a triangle mesh")
    if triangles.min() < 0 or triangles.max() >= len(vertices):
        raise ValueError("triangle indices out of range")
    if thickness > 0:
        if thickness < 2 * spacing:
            raise ValueError("thickness must be >= 2 * spacing (one cell on each side of the mid-surface)")
    elif not is_closed(triangles):
        raise ValueError("open mesh (some edge is not shared by exactly two triangles): a signed distance has no "
                         "inside. Weld duplicate seam vertices first if the asset is watertight, or pass "
                         "thickness >= 2 * spacing to bake it as a two-sided shell")
    if round < 0:
        raise ValueError("round must be >= 0")
    origin, dims = bake_lattice(vertices, spacing, margin, thickness, round)
    from . import _ropesim
    values = _ropesim.bake_mesh_sdf(vertices, triangles, origin, float(spacing), [int(d) for d in dims],
                                    float(thickness), float(round), -1 if device is None else int(device))
    g = SdfGrid(values, origin, float(spacing))
    if not np.isfinite(g.values).all():
        raise ValueError("bake_mesh produced non-finite distances")
    return g


# ---- closed-form shapes (object frame) ------------------------------------------------

def sdf_box(half, centre=(0, 0, 0), round=0.0):
    """Axis-aligned box of half extents `half` at `centre`, edges rounded with radius `round`
    (the overall size stays `half`). Exact inside and outside."""
    half = np.asarray(half, np.float64) - round
    if (half < 0).any():
        raise ValueError("round must not exceed the smallest half extent")
    centre = np.asarray(centre, np.float64)

    def phi(p):
        d = np.abs(p - centre) - half
        return np.linalg.norm(np.maximum(d, 0), axis=1) + np.minimum(d.max(1), 0) - round
    return phi


def sdf_sphere(radius, centre=(0, 0, 0)):
    c = np.asarray(centre, np.float64)
    return lambda p: np.linalg.norm(p - c, axis=1) - radius


def sdf_cylinder(radius, axis=2):
    """Infinite cylinder of `radius` along coordinate `axis`."""
    keep = [d for d in range(3) if d != axis]
    return lambda p: np.linalg.norm(p[:, keep], axis=1) - radius


def sdf_bowl(radius, centre=(0, 0, 0)):
    """Spherical cavity: negative outside the sphere, so the rope lives inside it."""
    s = sdf_sphere(radius, centre)
    return lambda p: -s(p)


def sdf_capsule(radius, half_length, axis=2, centre=(0, 0, 0)):
    """Capsule: the points within `radius` of the segment of half length `half_length` along `axis`."""
    c = np.asarray(centre, np.float64)

    def phi(p):
        q = p - c
        t = np.clip(q[:, axis], -half_length, half_length)
        q = q.copy()
        q[:, axis] -= t
        return np.linalg.norm(q, axis=1) - radius
    return phi


def sdf_torus(major, minor, axis=2, centre=(0, 0, 0)):
    """Torus of tube radius `minor` around a circle of radius `major` normal to `axis`."""
    c = np.asarray(centre, np.float64)
    keep = [d for d in range(3) if d != axis]

    def phi(p):
        q = p - c
        ring = np.linalg.norm(q[:, keep], axis=1) - major
        return np.sqrt(ring * ring + q[:, axis] ** 2) - minor
    return phi
