"""The HIP RopeBatch behind the CPU oracle's test surface (oracle/pyoracle.py OracleBatch), so the analytic
checks run unchanged on the device: numpy in and out, root programs handed over as device tensors.

Only what the analytic tests call is provided. Arithmetic is float32 on the device; `FP32` tells tests to
use their float32 tolerances, `FAST` which arithmetic mode the batches run (`IEEE` is the same surface with
every operation correctly rounded).
"""
import ctypes as C
import os
import types

import numpy as np

FP32 = True
FAST = True


def _flat(positions, offsets):
    if offsets is not None:
        return np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3)), np.asarray(offsets, np.int32)
    if isinstance(positions, np.ndarray) and positions.ndim == 3:
        N, V = positions.shape[:2]
        return np.ascontiguousarray(positions.reshape(-1, 3), np.float32), np.arange(N + 1, dtype=np.int32) * V
    ropes = [np.asarray(r, np.float32).reshape(-1, 3) for r in positions]
    off = np.zeros(len(ropes) + 1, np.int32)
    off[1:] = np.cumsum([len(r) for r in ropes])
    return np.ascontiguousarray(np.concatenate(ropes)), off


class OracleBatch:
    _fast = True

    def __init__(self, positions, density, k_stretch, k_bend, max_h, num_itr, fix_root_orientation=False,
                 pinned=None, straight_rest=False, prec=32, offsets=None):
        import ropesim
        self._rs = ropesim
        flat, off = _flat(positions, offsets)
        object.__setattr__(self, "offsets", off)
        object.__setattr__(self, "nv", int(off[-1]))
        object.__setattr__(self, "nropes", len(off) - 1)
        pin = None if pinned is None else np.asarray(pinned, bool).reshape(-1)
        b = ropesim.RopeBatch.from_positions(flat, density, k_stretch, k_bend, max_h, num_itr,
                                             fix_root_orientation=fix_root_orientation, pinned=pin,
                                             straight_rest=straight_rest, offsets=off)
        b.fast_math = self._fast
        object.__setattr__(self, "b", b)
        object.__setattr__(self, "prec", 32)

    def __setattr__(self, name, value):
        if name in ("drag", "damping", "drag_quad", "drag_quad_t", "num_itr"):
            setattr(self.b, name, value)
        else:
            object.__setattr__(self, name, value)

    # -- reads (host numpy, flat over vertices)
    def _host(self, view, w):
        import torch
        return torch.as_tensor(view, device="cuda").cpu().numpy().reshape(-1, w).astype(np.float64)

    def positions(self):
        return self._host(self.b.positions(), 3)

    def velocities(self):
        return self._host(self.b.velocities(), 3)

    def quaternions(self):
        return self._host(self.b.quaternions(), 4)

    def inv_masses(self):
        return np.asarray(self.b.inv_masses(), np.float64).reshape(-1)

    def rest_lengths(self):
        lib = C.CDLL(os.path.join(os.path.dirname(self._rs.__file__), "librope_hip.so"))
        out = np.zeros(self.nv, np.float32)
        assert lib.rs_get_vertex(C.c_void_p(self.b._handle), 4, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
        return out.astype(np.float64)

    def contact_forces(self):
        return np.asarray(self.b.contact_forces(), np.float64).reshape(-1, 3)

    def mesh_wrenches(self):
        w = np.asarray(self.b.mesh_wrenches(), np.float64)
        return w.reshape(self.nropes, -1, 6)

    # -- configuration
    def set_gravity(self, g):
        self.b.set_gravity(np.asarray(g, np.float32))

    def set_state(self, what, arr):
        assert what == "velocities"
        self.b.set_velocities(np.asarray(arr, np.float32).reshape(self.nv, 3))

    def set_planes(self, planes, radius=0.0, mu_s=0.5, mu_k=None, v_slip=1e-3, kappa=2.0):
        self.b.set_planes(np.asarray(planes, np.float32).reshape(-1, 4), radius=radius, mu_s=mu_s, mu_k=mu_k,
                          v_slip=v_slip, kappa=kappa)

    def set_contact(self, radius, mu_s=0.5, mu_k=None, v_slip=1e-3, kappa=2.0):
        self.b.set_contact(radius, mu_s=mu_s, mu_k=mu_k, v_slip=v_slip, kappa=kappa)

    def set_masses(self, m):
        self.b.set_masses(np.asarray(m, np.float32).reshape(-1))

    def set_stretch_stiffnesses(self, ea):
        self.b.set_stretch_stiffnesses(np.asarray(ea, np.float32).reshape(-1))

    def set_bend_stiffnesses(self, ei):
        self.b.set_bend_stiffnesses(np.asarray(ei, np.float32).reshape(-1))

    def add_mesh(self, values, origin, spacing, pos=(0, 0, 0), quat=(0, 0, 0, 1), mu_s=None, mu_k=None):
        return self.b.add_mesh(np.asarray(values, np.float32), np.asarray(origin, np.float32), float(spacing),
                               pos=np.asarray(pos, np.float32), quat=np.asarray(quat, np.float32), mu_s=mu_s,
                               mu_k=mu_k)

    def set_mesh_poses(self, index, pos, quat, teleport=False):
        self.b.set_mesh_poses(index, np.asarray(pos, np.float32), np.asarray(quat, np.float32), teleport=teleport)

    def enable_contact_forces(self, on=True):
        self.b.enable_contact_forces(on)

    def set_rope_contact(self, enabled, radius=0.0, barrier=0.0, k_collision=0.0, mu=0.0, k_friction=0.0,
                         detection_scaler=1.2, group_ropes=1):
        self.b.set_rope_contact(enabled, radius, barrier, k_collision, mu, k_friction, detection_scaler, group_ropes)

    # -- stepping
    def step(self, root_pos, root_vel, dt, substeps, root_q=None, root_pos_end=None, root_q_end=None):
        import torch
        d = lambda a: None if a is None else torch.as_tensor(np.asarray(a, np.float32).reshape(self.nropes, -1),
                                                               device="cuda").contiguous()
        self.b.step(d(root_pos), d(root_vel), dt, substeps, root_q=d(root_q), root_pos_end=d(root_pos_end),
                    root_q_end=d(root_q_end))

    def step_free(self, dt, substeps):
        self.b.step_free(dt, substeps)


class _OracleBatchIEEE(OracleBatch):
    _fast = False


IEEE = types.SimpleNamespace(OracleBatch=_OracleBatchIEEE, FP32=True, FAST=False)
