"""rs_bake_mesh_sdf_noalloc (no counterpart in the reference): the mesh bake on device arrays with the triangle
corners staged in the library's fixed device array instead of an allocation. Its grids must equal the allocating
entry point rs_bake_mesh_sdf's bit for bit (same kernel, same node arithmetic), on two calls in a row, the second
smaller than the first (the staging array keeps the first call's corners beyond the second's), and the size and
index checks must reject what the fixed array or the vertex array cannot hold."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP, IP = C.POINTER(C.c_float), C.POINTER(C.c_int)


def _lib(rs):
    L = C.CDLL(os.path.join(os.path.dirname(rs.__file__), "librope_hip.so"))
    L.rs_last_error_message.restype = C.c_char_p
    L.rs_bake_mesh_sdf.argtypes = [FP, C.c_int, IP, C.c_int, FP, C.c_float, C.c_int, C.c_int, C.c_int, C.c_float,
                                   C.c_float, C.c_int, FP]
    L.rs_bake_mesh_sdf_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, FP, C.c_float, C.c_int, C.c_int,
                                           C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p]
    return L


def torus(R=0.2, r=0.06, nu=64, nv=32):
    u = 2 * np.pi * np.arange(nu) / nu
    v = 2 * np.pi * np.arange(nv) / nv
    U, W = np.meshgrid(u, v, indexing="ij")
    verts = np.stack([(R + r * np.cos(W)) * np.cos(U), (R + r * np.cos(W)) * np.sin(U), r * np.sin(W)], -1)
    idx = lambda i, j: (i % nu) * nv + (j % nv)
    tris = []
    for i in range(nu):
        for j in range(nv):
            tris.append([idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)])
            tris.append([idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)])
    return verts.reshape(-1, 3).astype(np.float32), np.array(tris, np.int32)


def box(h=(0.1, 0.07, 0.05)):
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float64) * np.array(h)
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = [[a, b, c_] for a, b, c_, d in faces] + [[a, c_, d] for a, b, c_, d in faces]
    return c.astype(np.float32), np.array(tris, np.int32)


def bake_host(L, v, t, origin, spacing, shape, thickness=0.0, rnd=0.0):
    out = np.zeros(int(np.prod(shape)), np.float32)
    o = np.asarray(origin, np.float32)
    rc = L.rs_bake_mesh_sdf(v.ctypes.data_as(FP), len(v), t.ctypes.data_as(IP), len(t), o.ctypes.data_as(FP),
                            spacing, *shape, thickness, rnd, -1, out.ctypes.data_as(FP))
    assert rc == 0, L.rs_last_error_message()
    return out


def bake_noalloc(L, v, t, origin, spacing, shape, thickness=0.0, rnd=0.0, stream=None, ntri=None):
    import torch
    dv = torch.as_tensor(v, device="cuda").contiguous()
    dt = torch.as_tensor(t, device="cuda").contiguous()
    out = torch.full((int(np.prod(shape)),), float("nan"), device="cuda")
    o = np.asarray(origin, np.float32)
    torch.cuda.synchronize()
    rc = L.rs_bake_mesh_sdf_noalloc(dv.data_ptr(), len(v), dt.data_ptr(), len(t) if ntri is None else ntri,
                                    o.ctypes.data_as(FP), spacing, *shape, thickness, rnd, stream, out.data_ptr())
    return rc, out.cpu().numpy()


def test_noalloc_bake_equals_allocating_bake_twice(rs):
    import torch
    L = _lib(rs)
    v1, t1 = torus()
    shape1, origin1, h1 = (40, 40, 16), (-0.29, -0.29, -0.08), 0.015
    v2, t2 = box()
    shape2, origin2, h2 = (12, 10, 9), (-0.13, -0.1, -0.08), 0.02
    want1 = bake_host(L, v1, t1, origin1, h1, shape1)
    want2 = bake_host(L, v2, t2, origin2, h2, shape2, rnd=0.01)
    rc, got1 = bake_noalloc(L, v1, t1, origin1, h1, shape1, stream=torch.cuda.current_stream().cuda_stream)
    assert rc == 0, L.rs_last_error_message()
    rc, got2 = bake_noalloc(L, v2, t2, origin2, h2, shape2, rnd=0.01)  # 12 triangles after 4096; null stream
    assert rc == 0, L.rs_last_error_message()
    assert np.array_equal(got1.view(np.uint32), want1.view(np.uint32))
    assert np.array_equal(got2.view(np.uint32), want2.view(np.uint32))
    # the grids are what the shapes say: inside negative, outside positive
    assert (want1 < 0).sum() > 100 and (want1 > 0).sum() > 100
    g2 = want2.reshape(shape2)
    assert g2[6, 5, 4] < 0 and g2[0, 0, 0] > 0
    # a shell (thickness > 0) as well
    rc, got3 = bake_noalloc(L, v2, t2, origin2, h2, shape2, thickness=0.01)
    assert rc == 0
    assert np.array_equal(got3.view(np.uint32), bake_host(L, v2, t2, origin2, h2, shape2, thickness=0.01).view(np.uint32))


def test_noalloc_bake_rejects_what_it_cannot_hold(rs):
    L = _lib(rs)
    v, t = box()
    rc, _ = bake_noalloc(L, v, t, (0, 0, 0), 0.02, (4, 4, 4), ntri=262144 + 1)
    assert rc == 1 and b"staging array" in L.rs_last_error_message()
    bad = t.copy()
    bad[5, 1] = len(v)  # one index past the vertex array
    rc, _ = bake_noalloc(L, v, bad, (0, 0, 0), 0.02, (4, 4, 4))
    assert rc == 1 and b"out of range" in L.rs_last_error_message()
    # the check word is cleared for the next call
    rc, got = bake_noalloc(L, v, t, (-0.13, -0.1, -0.08), 0.02, (12, 10, 9))
    assert rc == 0
    want = bake_host(L, v, t, (-0.13, -0.1, -0.08), 0.02, (12, 10, 9))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
