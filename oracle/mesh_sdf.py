"""Float64 numpy restatement of the triangle-mesh distance grid. TEST INFRASTRUCTURE ONLY.

Only tests/ import this module; the ropesim package and its HIP library never do.

The reference bakes a mesh with Warp (ropesim/sdf.py:96-170 of the reference): for each lattice node p,
the closest point on the mesh (`wp.mesh_query_point_sign_winding_number(mesh, p, 1e6, 2.0, 0.5)` then
`wp.mesh_eval_position`), the distance |p - cp|, negative where the generalized winding number of the
closed mesh exceeds 0.5. Warp (warp-lang, unpinned in the reference's pyproject.toml:25) is absent here, so
this restates the published methods the query computes exactly, by brute force over all triangles:
- closest point on a triangle: the Voronoi-region method of Ericson, Real-Time Collision Detection (2005),
  section 5.1.5;
- generalized winding number: Jacobson, Kavan and Sorkine-Hornung, "Robust inside-outside segmentation
  using generalized winding numbers" (SIGGRAPH 2013), the sum over triangles of the signed solid angle
  (Van Oosterom and Strackee 1983: tan(omega / 2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|),
  with a, b, c the corners relative to p) over 4 pi. Warp's fast-winding-number tree approximates this sum
  (accuracy 2.0); the exact sum is its limit.
Then the reference's post-processing (ropesim/sdf.py:150-155): thickness > 0 -> |d| - thickness / 2, and
phi - round. The lattice points are the reference's: origin + spacing * index in float64, then float32
(sdf.py:61-75, 149).

Pinned by tests/test_mesh_sdf.py against closed forms: a box mesh reproduces sdf_box exactly inside and
outside, an icosphere lies between the sphere's distance and its faceting bound, the winding number of a
closed mesh is 1 inside and 0 outside, an open square with thickness gives |z| - t/2 over its interior.
"""
from __future__ import annotations

import numpy as np


def closest_point_distance(p, tris):
    """Distance from points p [M, 3] to each triangle tris [T, 3, 3] -> [M, T] (float64)."""
    p = np.asarray(p, np.float64)[:, None, :]
    a, b, c = (np.asarray(tris, np.float64)[None, :, k, :] for k in range(3))
    ab, ac, ap = b - a, c - a, p - a
    d1 = (ab * ap).sum(-1)
    d2 = (ac * ap).sum(-1)
    bp = p - b
    d3 = (ab * bp).sum(-1)
    d4 = (ac * bp).sum(-1)
    cp = p - c
    d5 = (ab * cp).sum(-1)
    d6 = (ac * cp).sum(-1)
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2
    with np.errstate(divide="ignore", invalid="ignore"):
        # interior: barycentric (1 - v - w, v, w)
        den = va + vb + vc
        v = vb / den
        w = vc / den
        q = a + ab * v[..., None] + ac * w[..., None]
        # regions applied from the lowest precedence up, so the last match wins in Ericson's test order:
        # vertex a, vertex b, edge ab, vertex c, edge ac, edge bc, interior
        t_bc = (d4 - d3) / ((d4 - d3) + (d5 - d6))
        e_bc = (va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0)
        q = np.where(e_bc[..., None], b + (c - b) * t_bc[..., None], q)
        t_ac = d2 / (d2 - d6)
        e_ac = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
        q = np.where(e_ac[..., None], a + ac * t_ac[..., None], q)
        q = np.where(((d6 >= 0) & (d5 <= d6))[..., None], c, q)
        t_ab = d1 / (d1 - d3)
        e_ab = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
        q = np.where(e_ab[..., None], a + ab * t_ab[..., None], q)
    q = np.where(((d3 >= 0) & (d4 <= d3))[..., None], b, q)
    q = np.where(((d1 <= 0) & (d2 <= 0))[..., None], a, q)
    return np.linalg.norm(p - q, axis=-1)


def winding_number(p, tris):
    """Generalized winding number of the triangle soup at points p [M, 3] -> [M] (float64)."""
    p = np.asarray(p, np.float64)[:, None, :]
    a, b, c = (np.asarray(tris, np.float64)[None, :, k, :] - p for k in range(3))
    la, lb, lc = (np.linalg.norm(x, axis=-1) for x in (a, b, c))
    det = (a * np.cross(b, c)).sum(-1)
    den = la * lb * lc + (a * b).sum(-1) * lc + (a * c).sum(-1) * lb + (b * c).sum(-1) * la
    return (2.0 * np.arctan2(det, den)).sum(-1) / (4.0 * np.pi)


def lattice_points(origin, spacing, shape):
    """Node coordinates in the reference's order and rounding: float64 axes, then float32."""
    ax = [np.float64(origin[d]) + spacing * np.arange(shape[d]) for d in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def bake(vertices, triangles, origin, spacing, shape, thickness=0.0, round=0.0, chunk=4096):
    """Distance grid [nx, ny, nz] (float64) of the mesh on the given lattice: signed by the winding number
    (|w| > 0.5 inside, so either consistent orientation of a closed mesh works) when thickness == 0, else the
    two-sided shell |d| - thickness / 2; then minus round."""
    v = np.asarray(vertices, np.float64).reshape(-1, 3)
    t = np.asarray(triangles, np.int64).reshape(-1, 3)
    tris = v[t]
    pts = lattice_points(origin, spacing, shape)
    out = np.empty(len(pts))
    for i in range(0, len(pts), chunk):
        p = pts[i:i + chunk]
        d = closest_point_distance(p, tris).min(1)
        if thickness > 0:
            d = np.abs(d) - 0.5 * thickness
        else:
            d = np.where(np.abs(winding_number(p, tris)) > 0.5, -d, d)
        out[i:i + chunk] = d - round
    return out.reshape(shape)


# ---- test meshes ---------------------------------------------------------------------------------------

def box_mesh(half, centre=(0, 0, 0)):
    """Closed box of half extents `half`: 8 vertices, 12 outward-wound triangles."""
    h = np.asarray(half, np.float64)
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64) * h + centre
    # vertex index = 4 * (x > 0) + 2 * (y > 0) + (z > 0)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tri = []
    for a, b, c, d in quads:
        tri += [(a, b, c), (a, c, d)]
    return v, np.array(tri, np.int64)


def icosphere(radius, subdivisions=2, centre=(0, 0, 0)):
    """Closed icosphere: the icosahedron subdivided `subdivisions` times, vertices projected on the sphere."""
    g = (1 + 5 ** 0.5) / 2
    v = [(-1, g, 0), (1, g, 0), (-1, -g, 0), (1, -g, 0), (0, -1, g), (0, 1, g), (0, -1, -g), (0, 1, -g),
         (g, 0, -1), (g, 0, 1), (-g, 0, -1), (-g, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    for _ in range(subdivisions):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[k] = len(v) - 1
            return mid[k]
        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v) * radius + np.asarray(centre, np.float64), np.array(f, np.int64)
