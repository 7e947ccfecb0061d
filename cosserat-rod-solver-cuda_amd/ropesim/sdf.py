"""Signed-distance lattices for rope-object contact.

The grid format is the one the solver's add_mesh takes (reference ropesim/sdf.py:20-58): a float32 array
values[nx, ny, nz] in C order holding the signed distance phi (negative inside the object) at the lattice
nodes x_ijk = origin + spacing * (i, j, k), all in the object's own frame. ``SdfGrid.as_args()`` gives the
(values, origin, spacing) triple add_mesh expects. The solver interpolates phi trilinearly inside the lattice
and finds no contact outside it, so a lattice must reach past the surface by at least the rope radius
wherever a rope can go.

Grids come from closed forms (``from_function`` with ``sdf_box``, ``sdf_sphere``, ``sdf_cylinder``,
``sdf_bowl``, ``sdf_capsule``, ``sdf_torus``) or from triangle meshes (``bake_mesh``, computed by the HIP
library: exact closest points and a generalized winding number for the sign).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# corner offsets of a lattice cell, in the order the trilinear weights below are formed
_CORNERS = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], np.int64)


@dataclass
class SdfGrid:
    values: np.ndarray   # float32 [nx, ny, nz], C order
    origin: np.ndarray   # float32 [3], object frame
    spacing: float

    @property
    def shape(self):
        return self.values.shape

    def nodes(self):
        """Object-frame coordinates of every node, float32 [nx, ny, nz, 3]."""
        idx = np.indices(self.values.shape, dtype=np.float64)             # [3, nx, ny, nz]
        pts = self.origin.astype(np.float64)[:, None, None, None] + float(self.spacing) * idx
        return np.moveaxis(pts, 0, -1).astype(np.float32)

    def as_args(self):
        """The (values, origin, spacing) arguments of add_mesh."""
        return self.values, self.origin, float(self.spacing)

    def sample(self, points):
        """Trilinear interpolation of the lattice at object-frame points [M, 3], in float64, for checking a
        scene on the host. Returns (phi [M], inside [M]): a point is inside when its lattice coordinate
        u = (p - origin) / spacing satisfies 0 <= u < n - 1 on every axis; phi is nan elsewhere (the solver
        finds no contact there)."""
        p = np.asarray(points, np.float64).reshape(-1, 3)
        n = np.asarray(self.values.shape)
        u = (p - self.origin.astype(np.float64)) / float(self.spacing)
        inside = np.all((u >= 0.0) & (u < n - 1), axis=1)
        phi = np.full(p.shape[0], np.nan)
        if not inside.any():
            return phi, inside
        ui = u[inside]
        cell = np.floor(ui).astype(np.int64)
        frac = ui - cell
        vals = self.values.astype(np.float64)
        acc = np.zeros(ui.shape[0])
        for off in _CORNERS:
            # weight of a corner: the product over axes of frac (offset 1) or 1 - frac (offset 0)
            w = np.prod(np.where(off == 1, frac, 1.0 - frac), axis=1)
            c = cell + off
            acc += w * vals[c[:, 0], c[:, 1], c[:, 2]]
        phi[inside] = acc
        return phi, inside


def _grid_of_box(lo, hi, spacing):
    """Lattice covering the box [lo, hi]: float32 origin lo, and per axis the node count
    ceil((hi - lo) / spacing) + 1, at least 2 (the count the reference's lattice takes)."""
    lo64 = np.asarray(lo, np.float64)
    span = np.asarray(hi, np.float64) - lo64
    counts = np.ceil(span / float(spacing)).astype(np.int64) + 1
    return lo64.astype(np.float32), np.maximum(counts, 2)


def from_function(phi, lo, hi, spacing, chunk_nodes=2_000_000):
    """Sample a signed-distance function phi(points [M, 3] float64) -> [M] at the nodes of the lattice over
    [lo, hi], a block of x-planes (about `chunk_nodes` nodes) at a time. Node coordinates are formed in float64
    from the float32 origin."""
    origin, counts = _grid_of_box(lo, hi, spacing)
    nx, ny, nz = (int(c) for c in counts)
    values = np.empty((nx, ny, nz), np.float32)
    o = origin.astype(np.float64)
    ys = o[1] + spacing * np.arange(ny)
    zs = o[2] + spacing * np.arange(nz)
    planes = max(1, chunk_nodes // (ny * nz))
    for x0 in range(0, nx, planes):
        xs = o[0] + spacing * np.arange(x0, min(nx, x0 + planes))
        pts = np.empty((len(xs), ny, nz, 3))
        pts[..., 0] = xs[:, None, None]
        pts[..., 1] = ys[None, :, None]
        pts[..., 2] = zs[None, None, :]
        values[x0:x0 + len(xs)] = np.asarray(phi(pts.reshape(-1, 3)), np.float32).reshape(len(xs), ny, nz)
    return SdfGrid(values, origin, float(spacing))


def is_closed(triangles):
    """Whether the triangles form a closed surface: each undirected edge belongs to exactly two triangles
    (orientation is not checked)."""
    tri = np.asarray(triangles, np.int64).reshape(-1, 3)
    if tri.shape[0] == 0:
        return False
    a = tri.reshape(-1)
    b = np.roll(tri, -1, axis=1).reshape(-1)                       # edges (v0 v1), (v1 v2), (v2 v0)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    order = np.lexsort((hi, lo))
    lo, hi = lo[order], hi[order]
    starts = np.flatnonzero(np.r_[True, (lo[1:] != lo[:-1]) | (hi[1:] != hi[:-1])])
    runs = np.diff(np.r_[starts, lo.size])
    return bool(np.all(runs == 2))


def bake_lattice(vertices, spacing, margin, thickness=0.0, round=0.0):
    """Float32 origin and node counts of the lattice bake_mesh fills: the vertices' bounding box grown on every
    side by margin + thickness / 2 + round, the corners formed in float32 (reference ropesim/sdf.py:145-148)."""
    v = np.asarray(vertices, np.float32).reshape(-1, 3)
    grow = np.float32(margin + 0.5 * thickness + round)
    return _grid_of_box(v.min(axis=0) - grow, v.max(axis=0) + grow, spacing)


def _bake_input_problem(v, t, spacing, thickness, round):
    """Why bake_mesh cannot take these inputs, or None."""
    if v.shape[0] < 3 or t.shape[0] < 1:
        return "bake_mesh needs a triangle mesh (at least 3 vertices and 1 triangle)"
    if t.min() < 0 or t.max() >= v.shape[0]:
        return "triangle indices out of range of the vertex array"
    if round < 0:
        return "round must be >= 0"
    if thickness > 0:
        if thickness < 2 * spacing:
            return "a shell needs thickness >= 2 * spacing: one lattice cell on either side of the surface"
        return None
    if not is_closed(t):
        return ("open mesh: some edge is not shared by exactly two triangles, so the signed distance has no "
                "inside; weld seam vertices of a watertight asset, or bake a two-sided shell with "
                "thickness >= 2 * spacing")
    return None


def bake_mesh(vertices, triangles, spacing, margin, thickness=0.0, round=0.0, device=None):
    """Distance lattice of a triangle mesh over its grown bounding box, computed on the GPU.

    The reference's contract (ropesim/sdf.py:96-157 there): vertices [V, 3], triangles [T, 3] of vertex
    indices.
    - thickness = 0: the signed distance of a closed mesh (``is_closed``), negative inside; a node is inside
      when the mesh's generalized winding number there exceeds 1/2 in magnitude, which holds for either
      orientation of a consistently wound mesh.
    - thickness > 0: the two-sided shell |d| - thickness / 2 of any mesh (open sheets and height fields;
      a closed mesh becomes hollow). The shell must span at least two lattice cells (thickness >= 2 spacing)
      so that the interpolated field has a gradient on both faces.
    - round >= 0 is subtracted from the result: the surface moves outward and convex edges round off with
      that radius.
    - margin widens the lattice beyond the shell and the rounding on every side.
    - device: HIP device index, or None for the current one.
    The HIP kernel visits every (node, triangle) pair, so the time grows with nodes x triangles.
    """
    v = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles, np.int32).reshape(-1, 3))
    problem = _bake_input_problem(v, t, spacing, thickness, round)
    if problem:
        raise ValueError(problem)
    origin, counts = bake_lattice(v, spacing, margin, thickness, round)
    from . import _ropesim
    values = _ropesim.bake_mesh_sdf(v, t, origin, float(spacing), [int(c) for c in counts], float(thickness),
                                    float(round), -1 if device is None else int(device))
    if not np.all(np.isfinite(values)):
        raise ValueError("bake_mesh produced non-finite distances")
    return SdfGrid(values, origin, float(spacing))


# ---- closed forms in the object frame: functions of points [M, 3] (float64) -> distances [M] ----------------

def sdf_box(half, centre=(0, 0, 0), round=0.0):
    """Box of half extents `half` about `centre`, its edges and corners rounded with radius `round` while the
    outer size stays `half`. Exact Euclidean distance to the core box of half extents half - round, minus
    round: outside the core, the length of the per-axis excess; inside, its largest (negative) component."""
    core = np.asarray(half, np.float64) - float(round)
    if np.any(core < 0):
        raise ValueError("round is larger than a half extent of the box")
    c = np.asarray(centre, np.float64)

    def phi(p):
        q = np.abs(np.asarray(p, np.float64) - c) - core
        outside = np.sqrt(np.sum(np.square(np.clip(q, 0.0, None)), axis=1))
        inside = np.clip(np.max(q, axis=1), None, 0.0)
        return outside + inside - round
    return phi


def sdf_sphere(radius, centre=(0, 0, 0)):
    """Solid sphere: |p - centre| - radius."""
    c = np.asarray(centre, np.float64)

    def phi(p):
        return np.sqrt(np.sum(np.square(np.asarray(p, np.float64) - c), axis=1)) - radius
    return phi


def sdf_cylinder(radius, axis=2):
    """Solid cylinder of `radius`, unbounded along coordinate `axis` through the origin."""
    others = [d for d in range(3) if d != axis]

    def phi(p):
        r = np.asarray(p, np.float64)[:, others]
        return np.sqrt(np.sum(r * r, axis=1)) - radius
    return phi


def sdf_bowl(radius, centre=(0, 0, 0)):
    """Hollow sphere seen from within: radius - |p - centre|, negative outside, so a rope inside is held in."""
    ball = sdf_sphere(radius, centre)

    def phi(p):
        return -ball(p)
    return phi


def sdf_capsule(radius, half_length, axis=2, centre=(0, 0, 0)):
    """Capsule: points within `radius` of the segment from -half_length to +half_length along `axis`."""
    c = np.asarray(centre, np.float64)

    def phi(p):
        q = np.asarray(p, np.float64) - c
        nearest = np.zeros_like(q)
        nearest[:, axis] = np.clip(q[:, axis], -half_length, half_length)
        return np.sqrt(np.sum(np.square(q - nearest), axis=1)) - radius
    return phi


def sdf_torus(major, minor, axis=2, centre=(0, 0, 0)):
    """Torus: points within `minor` of the circle of radius `major` about `axis` through `centre`."""
    c = np.asarray(centre, np.float64)
    others = [d for d in range(3) if d != axis]

    def phi(p):
        q = np.asarray(p, np.float64) - c
        ring = np.sqrt(np.sum(np.square(q[:, others]), axis=1)) - major
        return np.sqrt(ring * ring + q[:, axis] * q[:, axis]) - minor
    return phi
