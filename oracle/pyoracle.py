"""ctypes front end of the CPU oracle (rope_oracle.hpp). TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke entry and the benchmark's CPU-baseline leg import this module; the
product (the ropesim package and its HIP library) never does. The class mirrors the
RopeBatch surface of the reference (csrc/bindings.cpp:735-1210) with the uneven-rope
extension (flat positions + CSR offsets), in float32 or float64 arithmetic.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# arith "ieee": every float operation correctly rounded and unfused (g++ -ffp-contract=off), the default;
# "fma": the same source with multiply-adds contracted into fused ones and correctly rounded division and
# square root, the arithmetic nvcc compiles the reference's kernels with by default (float32 only differs).
_LIB_PATHS = {"ieee": os.path.join(_HERE, "_build", "librope_oracle.so"),
              "fma": os.path.join(_HERE, "_build", "librope_oracle_fma.so")}
_LIB_PATH = _LIB_PATHS["ieee"]
_libs = {}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp; -ffp-contract=off, and the contracted variant)."""
    if force:
        for p in _LIB_PATHS.values():
            if os.path.exists(p):
                os.remove(p)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib(arith: str = "ieee"):
    if arith not in _libs:
        path = _LIB_PATHS[arith]
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, dp, ip, fp, u8p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_uint8)
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, ip, fp, u8p, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_num_verts.argtypes = [vp]
        L.orc_get.argtypes = [vp, C.c_char_p, dp]
        L.orc_set_state.argtypes = [vp, C.c_char_p, dp]
        L.orc_set_scalar.argtypes = [vp, C.c_char_p, C.c_double]
        L.orc_set_gravity.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.orc_set_planes.argtypes = [vp, C.c_int, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_set_contact.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_set_vertex.argtypes = [vp, C.c_int, dp]
        L.orc_add_mesh.argtypes = [vp, fp, C.c_int, C.c_int, C.c_int, dp, C.c_double, C.c_double, C.c_double, dp,
                                   dp, C.c_int, ip]
        L.orc_set_mesh_targets.argtypes = [vp, C.c_int, dp, dp, C.c_int, u8p, C.c_int]
        L.orc_set_mesh_friction.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        L.orc_clear_meshes.argtypes = [vp]
        L.orc_enable_forces.argtypes = [vp, C.c_int]
        L.orc_set_rope_contact.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int]
        L.orc_step_program.argtypes = [vp, dp, dp, dp, dp, dp, C.c_double, C.c_int]
        L.orc_step_const.argtypes = [vp, dp, dp, dp, C.c_double, C.c_int]
        L.orc_step_free.argtypes = [vp, C.c_double, C.c_int]
        L.orc_reset_envs.argtypes = [vp, ip, C.c_int]
        L.orc_detect.argtypes = [vp, dp, dp, ip, ip, dp]
        L.orc_broadphase.argtypes = [vp, dp, dp, ip, ip]
        L.orc_get_poses.argtypes = [vp, dp]
        _libs[arith] = L
    return _libs[arith]


def _d(a):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, np.float64))
    return a


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    pass


def _check(rc, L=None):
    L = lib() if L is None else L
    if rc == 1:
        raise ValueError(L.orc_last_error().decode())
    if rc != 0:
        raise OracleError(L.orc_last_error().decode())


def flatten_ropes(positions):
    """[N, V, 3] array or list of [V_i, 3] arrays -> (flat [sum V, 3] float32, offsets [N + 1] int32)."""
    if isinstance(positions, np.ndarray) and positions.ndim == 3:
        N, V = positions.shape[:2]
        return np.ascontiguousarray(positions.reshape(-1, 3), np.float32), np.arange(N + 1, dtype=np.int32) * V
    ropes = [np.asarray(r, np.float32).reshape(-1, 3) for r in positions]
    off = np.zeros(len(ropes) + 1, np.int32)
    off[1:] = np.cumsum([len(r) for r in ropes])
    return np.ascontiguousarray(np.concatenate(ropes, 0)), off


class OracleBatch:
    """CPU restatement of RopeBatch (deterministic sweeps) in float32 (prec=32) or float64 (prec=64), with
    unfused correctly rounded operations (arith="ieee") or contracted multiply-adds (arith="fma")."""

    def __init__(self, positions, density, k_stretch, k_bend, max_h, num_itr, fix_root_orientation=False,
                 pinned=None, straight_rest=False, prec=64, offsets=None, arith="ieee"):
        self._L = lib(arith)
        self.arith = arith
        if offsets is None:
            flat, off = flatten_ropes(positions)
        else:
            flat = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
            off = np.ascontiguousarray(np.asarray(offsets, np.int32))
        self.offsets = off
        self.nv = int(off[-1])
        self.nropes = len(off) - 1
        if pinned is None:
            pin = np.zeros(self.nv, np.uint8)
            pin[off[:-1]] = 1
        else:
            pin = np.ascontiguousarray(np.asarray(pinned, bool).reshape(-1).astype(np.uint8))
        self.prec = prec
        self._h = self._L.orc_create(prec, self.nropes, off.ctypes.data_as(C.POINTER(C.c_int)),
                                   flat.ctypes.data_as(C.POINTER(C.c_float)), pin.ctypes.data_as(C.POINTER(C.c_uint8)),
                                   float(density), float(k_stretch), float(k_bend), float(max_h), int(num_itr),
                                   int(bool(fix_root_orientation)), int(bool(straight_rest)))
        if not self._h:
            raise ValueError(self._L.orc_last_error().decode())
        self.num_meshes = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_destroy(h)
            self._h = None

    def _ck(self, rc):
        _check(rc, self._L)

    # -- state -------------------------------------------------------------
    def get(self, what, width=None):
        sizes = {"positions": 3, "velocities": 3, "last_velocities": 3, "dx": 3, "quaternions": 4,
                 "rest_quaternions": 4, "contact_forces": 3, "cand": 8, "col": 128}
        if what == "scalars":
            out = np.zeros(9)
        elif what == "wrenches":
            out = np.zeros(self.nropes * 64 * 6)
        else:
            out = np.zeros(self.nv * sizes.get(what, 1))
        self._ck(self._L.orc_get(self._h, what.encode(), _dp(out)))
        if what in sizes:
            return out.reshape(-1, sizes[what])
        if what == "wrenches":
            return out.reshape(self.nropes, 64, 6)
        return out

    def positions(self):
        return self.get("positions")

    def velocities(self):
        return self.get("velocities")

    def quaternions(self):
        return self.get("quaternions")

    def inv_masses(self):
        return self.get("inv_masses")

    def rest_lengths(self):
        return self.get("rest_lengths")

    @property
    def time(self):
        return float(self.get("scalars")[0])

    def set_state(self, what, arr):
        a = _d(arr).reshape(-1)
        self._ck(self._L.orc_set_state(self._h, what.encode(), _dp(a)))

    def set_scalar(self, name, v):
        self._ck(self._L.orc_set_scalar(self._h, name.encode(), float(v)))

    def __setattr__(self, name, value):
        if name in ("drag", "damping", "drag_quad", "drag_quad_t", "accel_ratio", "num_itr"):
            self.set_scalar(name, value)
        else:
            object.__setattr__(self, name, value)

    def set_gravity(self, g):
        g = np.asarray(g, np.float64)
        self._ck(self._L.orc_set_gravity(self._h, float(g[0]), float(g[1]), float(g[2])))

    def set_planes(self, planes, radius=0.0, mu_s=0.5, mu_k=None, v_slip=1e-3, kappa=2.0):
        p = _d(planes).reshape(-1, 4)
        if len(p) > 0:
            self._mu = (mu_s, mu_s if mu_k is None else mu_k)
        self._ck(self._L.orc_set_planes(self._h, len(p), _dp(p), radius, mu_s, mu_s if mu_k is None else mu_k, v_slip,
                                    kappa))

    def set_contact(self, radius, mu_s=0.5, mu_k=None, v_slip=1e-3, kappa=2.0):
        self._mu = (mu_s, mu_s if mu_k is None else mu_k)
        self._ck(self._L.orc_set_contact(self._h, radius, mu_s, mu_s if mu_k is None else mu_k, v_slip, kappa))

    def set_masses(self, m):
        self._ck(self._L.orc_set_vertex(self._h, 0, _dp(_d(m).reshape(-1))))

    def set_stretch_stiffnesses(self, ea):
        """Per-segment EA on the flat vertex index (entry at a rope's last vertex ignored)."""
        self._ck(self._L.orc_set_vertex(self._h, 1, _dp(_d(ea).reshape(-1))))

    def set_bend_stiffnesses(self, ei):
        self._ck(self._L.orc_set_vertex(self._h, 2, _dp(_d(ei).reshape(-1))))

    def add_mesh(self, values, origin, spacing, pos=(0, 0, 0), quat=(0, 0, 0, 1), mu_s=None, mu_k=None):
        """Friction defaults to the contact material's, as in the reference (csrc/bindings.cpp:474-480)."""
        cs, ck = getattr(self, "_mu", (0.0, 0.0))
        mu_k = (ck if mu_s is None else mu_s) if mu_k is None else mu_k
        mu_s = cs if mu_s is None else mu_s
        v = np.ascontiguousarray(np.asarray(values, np.float32))
        nx, ny, nz = v.shape
        p, q = _d(pos).reshape(-1, 3), _d(quat).reshape(-1, 4)
        idx = C.c_int(0)
        self._ck(self._L.orc_add_mesh(self._h, v.ctypes.data_as(C.POINTER(C.c_float)), nx, ny, nz, _dp(_d(origin)),
                                  float(spacing), mu_s, mu_k, _dp(p), _dp(q), len(p),
                                  C.byref(idx)))
        self.num_meshes += 1
        return idx.value

    def set_mesh_poses(self, index, pos, quat, teleport=False):
        p, q = _d(pos).reshape(-1, 3), _d(quat).reshape(-1, 4)
        tele_all, flags = self._tele(teleport)
        self._ck(self._L.orc_set_mesh_targets(self._h, index, _dp(p), _dp(q), len(p), flags, tele_all))

    def set_mesh_poses_all(self, pos, quat, teleport=False):
        p, q = _d(pos).reshape(-1, 3), _d(quat).reshape(-1, 4)
        tele_all, flags = self._tele(teleport)
        self._ck(self._L.orc_set_mesh_targets(self._h, -1, _dp(p), _dp(q), len(p), flags, tele_all))

    def _tele(self, teleport):
        if isinstance(teleport, (bool, np.bool_)):
            return int(bool(teleport)), None
        self._tele_buf = np.ascontiguousarray(np.asarray(teleport, bool).astype(np.uint8))
        return 0, self._tele_buf.ctypes.data_as(C.POINTER(C.c_uint8))

    def enable_contact_forces(self, on=True):
        self._ck(self._L.orc_enable_forces(self._h, int(bool(on))))

    def contact_forces(self):
        return self.get("contact_forces")

    def mesh_wrenches(self):
        return self.get("wrenches")[:, : self.num_meshes]

    def set_rope_contact(self, enabled, radius=0.0, barrier=0.0, k_collision=0.0, mu=0.0, k_friction=0.0,
                         detection_scaler=1.2, group_ropes=1):
        self._ck(self._L.orc_set_rope_contact(self._h, int(bool(enabled)), radius, barrier, k_collision, mu, k_friction,
                                          detection_scaler, group_ropes))

    # -- stepping ----------------------------------------------------------
    def step(self, root_pos, root_vel, dt, substeps, root_q=None, root_pos_end=None, root_q_end=None):
        p, v = _d(root_pos).reshape(-1), _d(root_vel).reshape(-1)
        if root_pos_end is None:
            q = _d(root_q)
            self._ck(self._L.orc_step_const(self._h, _dp(p), _dp(v), _dp(q.reshape(-1) if q is not None else None),
                                        float(dt), int(substeps)))
        else:
            p1 = _d(root_pos_end).reshape(-1)
            q0 = _d(root_q)
            q1 = _d(root_q_end)
            self._ck(self._L.orc_step_program(self._h, _dp(p), _dp(v), _dp(p1),
                                          _dp(q0.reshape(-1) if q0 is not None else None),
                                          _dp(q1.reshape(-1) if q1 is not None else None), float(dt), int(substeps)))

    def step_free(self, dt, substeps):
        self._ck(self._L.orc_step_free(self._h, float(dt), int(substeps)))

    def reset_envs(self, ids):
        ids = np.ascontiguousarray(np.asarray(ids, np.int32))
        self._ck(self._L.orc_reset_envs(self._h, ids.ctypes.data_as(C.POINTER(C.c_int)), len(ids)))

    # -- function-level checks --------------------------------------------
    def detect(self, last_pos, pads):
        lp, pd = _d(last_pos).reshape(-1), _d(pads).reshape(-1)
        counts = np.zeros(self.nv, np.int32)
        lists = np.zeros(self.nv * 128, np.int32)
        lim = np.zeros(self.nv)
        self._ck(self._L.orc_detect(self._h, _dp(lp), _dp(pd), counts.ctypes.data_as(C.POINTER(C.c_int)),
                                lists.ctypes.data_as(C.POINTER(C.c_int)), _dp(lim)))
        return counts, lists.reshape(self.nv, 128), lim

    def broadphase(self, pos, dx):
        p, d = _d(pos).reshape(-1), _d(dx).reshape(-1)
        counts = np.zeros(self.nv, np.int32)
        cand = np.zeros(self.nv * 8, np.int32)
        self._ck(self._L.orc_broadphase(self._h, _dp(p), _dp(d), counts.ctypes.data_as(C.POINTER(C.c_int)),
                                    cand.ctypes.data_as(C.POINTER(C.c_int))))
        return counts, cand.reshape(self.nv, 8)

    def poses(self):
        out = np.zeros(self.nropes * 64 * 19)
        self._ck(self._L.orc_get_poses(self._h, _dp(out)))
        return out.reshape(self.nropes, 64, 19)
