"""Float64 torch transcription of the deterministic substep. TEST INFRASTRUCTURE ONLY.

Only tests/ import this module. It is the differentiable reference for the hand-written adjoint:
torch autograd through this transcription gives the gradient of a rollout loss with respect to
per-vertex EA / EI / mass, linear drag and the root program. Branches are taken from the forward
values, so autograd differentiates the same frozen active set the adjoint uses.

It restates the same method as rope_oracle.hpp (same reference lines) for the parts the adjoint
covers. Meshes are static within a frame. Rope-rope contact follows the oracle's detection: the candidate
lists are formed from detached values each substep (a discrete decision), while the step limit (the
distance of each listed pair and the detection padding) and the barrier and friction terms carry gradients.
- root program: csrc/batch.cu:122-159
- initItr: YB/sim/iteration.cu:34-69
- cosseratItr without rope-rope terms: YB/sim/cosserat.cu:41-88,198-215
- quaternionLambdaItr: YB/sim/cosserat.cu:227-273
- endItr with linear and quadratic drag: YB/sim/iteration.cu:77-119
- plane and SDF contact: YB/sim/contact.cuh:27-163
- rope-rope detection and step limit: csrc/batch.cu:19-49, YB/sim/collision.cu:103-136
- rope-rope barrier and friction: YB/sim/cosserat.cu:90-158
- per-vertex mesh broad phase: YB/sim/contact.cuh:121-136
(YB = third_party/yarnball/KittenEngine/YarnBall)

The construction state comes from a float64 OracleBatch, so the two share the frames, masses and
rest shapes. Parameters enter as leaves:
- k = EA * l
- rest Darboux = (4 EI / l) * unit direction
- invMass = 1 / m on free vertices
"""
from __future__ import annotations

import numpy as np
import torch

HAS_PREV, HAS_NEXT, HAS_NEXT_ORIENT, FIX_ORIENT = 1, 2, 4, 8
MAX_CAND = 8
F64 = torch.float64


def _dot(a, b):
    return (a * b).sum(-1, keepdim=True)


def _cross(a, b):
    return torch.cross(a, b, dim=-1)


def qmul(a, b):
    av, aw = a[..., :3], a[..., 3:]
    bv, bw = b[..., :3], b[..., 3:]
    v = (aw * bv + bw * av) + _cross(av, bv)
    return torch.cat([v, aw * bw - _dot(av, bv)], -1)


def qconj(q):
    return torch.cat([-q[..., :3], q[..., 3:]], -1)


def qrot(q, v):
    u, w = q[..., :3], q[..., 3:]
    a = w * v + _cross(u, v)
    c = _dot(v, u)
    return (w * a + _cross(u, a)) + c * u


def _safe_len(v, ok):
    """|v| where ok, 1 elsewhere (keeps masked-out lanes out of sqrt(0) gradients)."""
    return torch.sqrt(_dot(torch.where(ok, v, torch.ones_like(v)), torch.where(ok, v, torch.ones_like(v))))


def _sym_full(Hp):
    """Packed [xx, yy, zz, xy, xz, yz] -> full [n, 3, 3]."""
    xx, yy, zz, xy, xz, yz = Hp.unbind(-1)
    return torch.stack([torch.stack([xx, xy, xz], -1), torch.stack([xy, yy, yz], -1),
                        torch.stack([xz, yz, zz], -1)], -2)


def _outer_packed(n):
    x, y, z = n.unbind(-1)
    return torch.stack([x * x, y * y, z * z, x * y, x * z, y * z], -1)


def _ssolve(Hp, f):
    """H^-1 f by the adjugate, same steps as the oracle's ssolve."""
    a, b, c, d, e, g = Hp.unbind(-1)
    c00 = b * c - g * g
    c01 = e * g - d * c
    c02 = d * g - b * e
    c11 = a * c - e * e
    c12 = d * e - a * g
    c22 = a * b - d * d
    det = (a * c00 + d * c01) + e * c02
    inv = 1.0 / det
    fx, fy, fz = f.unbind(-1)
    return torch.stack([(c00 * fx + c01 * fy + c02 * fz) * inv, (c01 * fx + c11 * fy + c12 * fz) * inv,
                        (c02 * fx + c12 * fy + c22 * fz) * inv], -1)


def seg_closest(a0, a1, b0, b1):
    """Closest-point parameters (u on a0-a1, v on b0-b1) by the reference's three steps (KittenEngine
    segmentClosestPoints as restated in rope_oracle.hpp): the infinite-line v, clamped; u from it, clamped;
    v again from u, clamped. A singular 2x2 system takes v = 0.5. The infinite-line step of nearly parallel
    segments (a c - b^2 <= 1e-4 a c) passes no gradient: float32 cannot resolve that v, and its derivative grows
    as 1 / sin^2 of the angle; the device adjoint holds it fixed as well."""
    A0, A1 = a1 - a0, b1 - b0
    a, b, c = _dot(A0, A0), _dot(A0, A1), _dot(A1, A1)
    diff = b0 - a0
    num = b * _dot(A0, diff) - a * _dot(A1, diff)
    den = a * c - b * b
    ok = den != 0
    y = num / torch.where(ok, den, torch.ones_like(den))
    ok = ok & torch.isfinite(y)
    y = torch.where(ok, y, torch.full_like(y, 0.5))
    y = torch.where(den > 1e-4 * (a * c), y, y.detach())
    y = torch.clamp(y, 0.0, 1.0)
    x = torch.clamp(_dot((b0 + y * A1) - a0, A0) / a, 0.0, 1.0)
    y = torch.clamp(_dot((a0 + x * A0) - b0, A1) / c, 0.0, 1.0)
    return x, y


class TorchRope:
    """Rollouts of a configured float64 OracleBatch in torch float64 (autograd-capable).

    mesh_specs: list of dicts {values [nx,ny,nz] float32, origin (3,), spacing, mu_s, mu_k} in the order the
    meshes were added; static poses are read from the oracle.
    """

    def __init__(self, ob, gravity, planes=(), contact=(0.0, 0.5, 0.5, 1e-3, 2.0), mesh_specs=(), drag_quad=0.0,
                 drag_quad_t=0.0, damping=0.0, accel=1.0, dtype=torch.float64, rope_contact=None):
        """dtype float64 is the reference; float32 measures what float32 arithmetic alone does to the same
        method (the construction state is still taken from the float64 oracle, rounded once)."""
        assert ob.prec == 64
        self.dtype = dtype
        t = lambda a: torch.as_tensor(np.asarray(a, np.float64), dtype=dtype)
        self.nv, self.nropes = ob.nv, ob.nropes
        self.off = np.asarray(ob.offsets, np.int64)
        self.x0 = t(ob.positions())
        self.v0 = t(ob.velocities())
        self.vl0 = t(ob.get("last_velocities"))
        self.q0 = t(ob.quaternions())
        self.w0 = t(ob.inv_masses())
        self.l = t(ob.rest_lengths())
        self.k0 = t(ob.get("k_stretch"))
        rq = ob.get("rest_quaternions")
        self.rho_unit = t(rq / np.linalg.norm(rq, axis=1, keepdims=True))
        self.rho_len0 = t(np.linalg.norm(rq, axis=1))
        fl = np.asarray(ob.get("flags"), np.int64)
        sc = ob.get("scalars")
        self.h_init = float(sc[1])
        m = lambda bit: torch.as_tensor((fl & bit) != 0).unsqueeze(-1)
        self.hasPrev, self.hasNext, self.hasNextO, self.fix = m(HAS_PREV), m(HAS_NEXT), m(HAS_NEXT_ORIENT), m(FIX_ORIENT)
        self.free = (self.w0 != 0).unsqueeze(-1)
        idx = torch.arange(self.nv)
        self.nxt = torch.clamp(idx + 1, max=self.nv - 1)
        self.prv = torch.clamp(idx - 1, min=0)
        self.roots = torch.as_tensor(self.off[:-1])
        self.ropeOf = torch.as_tensor(np.repeat(np.arange(self.nropes), np.diff(self.off)))
        self.g = t(gravity)
        self.planes = [(t(p[:3]) / float(np.linalg.norm(p[:3])), float(p[3])) for p in planes]
        self.cR, self.cMuS, self.cMuK, self.cVSlip, self.cKappa = [float(c) for c in contact]
        self.dragQuad, self.dragQuadT, self.damping, self.accel = drag_quad, drag_quad_t, damping, accel
        self.ropeContact = rope_contact is not None
        self.stats = {}  # rope-rope activity of the last rollouts: active barrier terms, friction terms, clamped steps
        if self.ropeContact:  # set_rope_contact keywords; uploadMeta (YB/YarnBall.cu:261-272)
            rc = rope_contact
            self.rcRadius, self.rcBarrier = float(rc["radius"]), float(rc["barrier"])
            self.kCollision, self.fricMu = float(rc["k_collision"]), float(rc["mu"])
            self.kFriction = float(rc["k_friction"])
            self.detR = self.rcRadius + 0.5 * self.rcBarrier
            self.scaledDetR = self.detR * float(rc.get("detection_scaler", 1.2))
            G = int(rc.get("group_ropes", 1))
            grp = np.arange(self.nropes) // G
            self.vgroup = grp[np.repeat(np.arange(self.nropes), np.diff(self.off))]
            self.ropeOfNp = np.repeat(np.arange(self.nropes), np.diff(self.off))
            self.hasNextNp = (fl & HAS_NEXT) != 0
        self.meshes = []
        if mesh_specs:
            poses = ob.poses()  # [nropes, 64, 19]: pos, q, vel, omega, lo, hi
            for mi, spec in enumerate(mesh_specs):
                vals = np.asarray(spec["values"], np.float32)
                nx, ny, nz = vals.shape
                P = poses[:, mi]
                self.meshes.append(dict(
                    v=t(vals.astype(np.float64).reshape(-1)), n=(nx, ny, nz), o=t(spec["origin"]),
                    h=float(spec["spacing"]), invH=1.0 / float(spec["spacing"]), muS=float(spec["mu_s"]),
                    muK=float(spec["mu_k"] if spec.get("mu_k") is not None else spec["mu_s"]),
                    pos=t(P[:, 0:3])[self.ropeOf], q=t(P[:, 3:7])[self.ropeOf],
                    lo=t(P[:, 13:16])[self.ropeOf], hi=t(P[:, 16:19])[self.ropeOf]))

    # ------------------------------------------------------------------ contact
    def _surface(self, n, gap, d, k, muS, muK, active, h, n_const):
        """surfaceContact (contact.cuh:27-46) on the active lanes; returns (df, dH packed)."""
        eps = self.cVSlip
        N = -k * gap
        f = N * n
        nn = _outer_packed(n)
        H = nn * k
        ut = d - _dot(n, d) * n
        ul = _safe_len(ut, active & (_dot(ut, ut) > 0))
        ul = torch.where(_dot(ut, ut) > 0, ul, torch.zeros_like(ul))
        vt = ul / h
        stick = vt <= eps
        s = 1.0 - torch.clamp(vt / eps - 1.0, max=1.0)
        ul_safe = torch.where(stick, torch.ones_like(ul), ul)
        kt = torch.where(stick, muS * N / (eps * h), (s * muS + (1.0 - s) * muK) * N / ul_safe)
        f = f - kt * ut
        eye = torch.tensor([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], dtype=self.dtype)
        H = H - (nn - eye) * kt
        z3 = torch.zeros_like(f)
        z6 = torch.zeros_like(H)
        return torch.where(active, f, z3), torch.where(active, H, z6)

    def _sdf_sample(self, M, p):
        n = M["n"]
        u = (p - M["o"]) * M["invH"]
        inside = ((u[:, 0] >= 0) & (u[:, 1] >= 0) & (u[:, 2] >= 0) & (u[:, 0] < n[0] - 1) & (u[:, 1] < n[1] - 1)
                  & (u[:, 2] < n[2] - 1))
        ijk = torch.stack([torch.clamp(u[:, a].detach().floor(), 0, n[a] - 2) for a in range(3)], -1)
        ijk = torch.where(inside.unsqueeze(-1), ijk, torch.zeros_like(ijk))
        fr = (p - (M["o"] + ijk * M["h"])) * M["invH"]
        fr = torch.clamp(fr, 0.0, 1.0)
        fx, fy, fz = fr[:, 0:1], fr[:, 1:2], fr[:, 2:3]
        ii = ijk.long()
        sy, sx = n[2], n[1] * n[2]
        base = ii[:, 0] * sx + ii[:, 1] * sy + ii[:, 2]
        g = lambda off: M["v"][base + off].unsqueeze(-1)
        c000, c001, c010, c011 = g(0), g(1), g(sy), g(sy + 1)
        c100, c101, c110, c111 = g(sx), g(sx + 1), g(sx + sy), g(sx + sy + 1)
        dx00, dx01, dx10, dx11 = c100 - c000, c101 - c001, c110 - c010, c111 - c011
        c00, c01, c10, c11 = c000 + fx * dx00, c001 + fx * dx01, c010 + fx * dx10, c011 + fx * dx11
        dy0, dy1 = c10 - c00, c11 - c01
        c0, c1 = c00 + fy * dy0, c01 + fy * dy1
        phi = c0 + fz * (c1 - c0)
        gx0, gx1 = dx00 + fy * (dx10 - dx00), dx01 + fy * (dx11 - dx01)
        grad = torch.cat([gx0 + fz * (gx1 - gx0), dy0 + fz * (dy1 - dy0), c1 - c0], -1) * M["invH"]
        return inside.unsqueeze(-1), phi, grad

    def _surfaces(self, X, d, k, h, cand):
        f = torch.zeros_like(X)
        H = torch.zeros(self.nv, 6, dtype=self.dtype)
        for n, w in self.planes:
            gap = ((_dot(X, n) - w) + _dot(d, n)) - self.cR
            act = self.free & (gap < 0)
            df, dH = self._surface(n.expand_as(X), gap, d, k, self.cMuS, self.cMuK, act, h, True)
            f, H = f + df, H + dH
        for mi, M in enumerate(self.meshes):
            pobj = qrot(qconj(M["q"]), (X - M["pos"]) + d)
            inside, phi, gr = self._sdf_sample(M, pobj)
            gap = phi - self.cR
            gl2 = _dot(gr, gr)
            act = self.free & inside & (gap < 0) & (gl2 > 0) & cand[:, mi:mi + 1]
            gl = _safe_len(gr, act)
            n = qrot(M["q"], gr / gl)
            df, dH = self._surface(n, gap, d, k, M["muS"], M["muK"], act, h, False)
            f, H = f + df, H + dH
        return f, H

    def _candidates(self, x, d):
        """sdfBuildCandidates (contact.cuh:121-136) for static meshes: mask [nv, M]; > 8 hits -> all."""
        if not self.meshes:
            return torch.zeros(self.nv, 0, dtype=torch.bool)
        padv = 2.0 * torch.sqrt(_dot(d, d)).detach() + 0.02
        hits = []
        for M in self.meshes:
            lo, hi = M["lo"], M["hi"]
            xd = x.detach()
            hit = ((xd >= lo - padv) & (xd <= hi + padv)).all(-1)
            hits.append(hit)
        hm = torch.stack(hits, -1)
        over = hm.sum(-1, keepdim=True) > MAX_CAND
        hm = torch.where(over, torch.ones_like(hm), hm)
        return hm & self.free  # pinned vertices keep no candidates

    # ------------------------------------------------------------------ rope-rope contact
    def _detect(self, X, y):
        """bruteDetectK (csrc/batch.cu:19-49) with the step limit (YB/sim/collision.cu:103-136). The lists are
        formed from detached float64 values in the oracle's order (ascending j per segment i); the step limit
        min(pad_i - detR, 0.45 |d_ij|), with pad = scaledDetR + |y|, keeps its gradient. Returns the listed
        pairs (i, j) and each vertex's limit min(limit_i, limit_{i-1})."""
        pad = self.scaledDetR + torch.sqrt(_dot(y, y))
        Xd, pd = X.detach(), pad.detach()
        hn = self.hasNextNp
        I, J = [], []
        for g in np.unique(self.vgroup):
            vs = np.nonzero((self.vgroup == g) & hn)[0]
            if len(vs) < 2:
                continue
            ii, jj = np.meshgrid(vs, vs, indexing="ij")
            ii, jj = ii.ravel(), jj.ravel()
            keep = ~((self.ropeOfNp[ii] == self.ropeOfNp[jj]) & (np.abs(jj - ii) <= 2))
            ii, jj = ii[keep], jj[keep]
            ti, tj = torch.as_tensor(ii), torch.as_tensor(jj)
            a0 = Xd[ti]
            a1, b0, b1 = Xd[ti + 1] - a0, Xd[tj] - a0, Xd[tj + 1] - a0
            with torch.no_grad():
                u, v = seg_closest(torch.zeros_like(a1), a1, b0, b1)
                dv = u * a1 - (b0 + v * (b1 - b0))
                r = torch.maximum(pd[ti], pd[ti + 1]) + torch.maximum(pd[tj], pd[tj + 1])
                hit = (_dot(dv, dv) < r * r).squeeze(-1).numpy()
            I.append(ii[hit])
            J.append(jj[hit])
        I = np.concatenate(I) if I else np.zeros(0, np.int64)
        J = np.concatenate(J) if J else np.zeros(0, np.int64)
        order = np.lexsort((J, I))
        I, J = I[order], J[order]
        if len(I) and np.bincount(I).max() > 128:
            raise RuntimeError("MAX_COLLISIONS_PER_SEGMENT exceeded")
        inf = torch.full((self.nv, 1), float("inf"), dtype=self.dtype)
        ms = torch.where(self.hasNext, pad - self.detR, inf)
        if len(I):
            ti, tj = torch.as_tensor(I), torch.as_tensor(J)
            a0 = X[ti]
            a1, b0, b1 = X[ti + 1] - a0, X[tj] - a0, X[tj + 1] - a0
            u, v = seg_closest(torch.zeros_like(a1), a1, b0, b1)
            u, v = u.detach(), v.detach()  # the distance is a minimum over (u, v): no term through them
            dv = u * a1 - (b0 + v * (b1 - b0))
            dist = 0.45 * torch.sqrt(_dot(dv, dv))
            ms = ms.scatter_reduce(0, ti.unsqueeze(-1), dist, reduce="amin", include_self=True)
        lim = torch.where(self.hasPrev, torch.minimum(ms, ms[self.prv]), ms)
        return (torch.as_tensor(I), torch.as_tensor(J)), lim

    def _pair_terms(self, X, d, pairs, dmp):
        """The barrier and friction terms of every listed pair (YB/sim/cosserat.cu:99-158): forces and packed
        Hessians on vertex i (f, H) and on i + 1 (f2, H2)."""
        I, J = pairs
        p0, p1, d0, p1dx = X[I], X[I + 1], d[I], d[I + 1]
        b0, b1, db0, db1 = X[J], X[J + 1], d[J], d[J + 1]
        u, v = seg_closest(torch.zeros_like(p0), (p1 - p0) + (p1dx - d0), (b0 - p0) + (db0 - d0), (b1 - p0) + (db1 - d0))
        dpos = (p0 + u * (p1 - p0)) - (b0 + v * (b1 - b0))
        ddpos = (d0 + u * (p1dx - d0)) - (db0 + v * (db1 - db0))
        normal = dpos + ddpos
        n2 = _dot(normal, normal)
        on = n2 > 0
        dd = _safe_len(normal, on)
        n = normal / dd
        wy, wx = u, 1.0 - u
        invb = 1.0 / self.rcBarrier
        kCol = self.kCollision * invb
        g = (dd - 2.0 * self.rcRadius) * invb
        on = on & (g <= 1.0)
        g = torch.clamp(g, min=1e-3)
        invd = 1.0 / g
        logd = torch.log(g)
        dH = ((-3.0 + (2.0 + invd) * invd) - 2.0 * logd) * kCol * invb
        ff = -(1.0 - g) * ((g - 1.0) + 2.0 * g * logd) * invd * kCol
        f = (ff * wx - dmp * dH * wx * wx * _dot(n, d0)) * n
        f2 = (ff * wy - dmp * dH * wy * wy * _dot(n, p1dx)) * n
        dH = dH * (1.0 + dmp)
        op = _outer_packed(n)
        H = op * (dH * wx * wx)
        H2 = op * (dH * wy * wy)
        U = ddpos - _dot(n, ddpos) * n
        fr = on & (_dot(U, U) > 0)
        ul = _safe_len(U, fr)
        f1 = torch.minimum(torch.full_like(ul, self.kFriction), self.fricMu * ff / ul)
        eye = torch.tensor([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], dtype=self.dtype)
        z3, z6 = torch.zeros_like(f), torch.zeros_like(H)
        f = f - torch.where(fr, f1 * wx * U, z3)
        H = H - torch.where(fr, (op - eye) * (wx * wx * f1), z6)
        f2 = f2 - torch.where(fr, f1 * wy * U, z3)
        H2 = H2 - torch.where(fr, (op - eye) * (wy * wy * f1), z6)
        st = self.stats
        st["barrier"] = st.get("barrier", 0) + int(on.sum())
        st["friction"] = st.get("friction", 0) + int(fr.sum())
        return (torch.where(on, f, z3), torch.where(on, H, z6), torch.where(on, f2, z3), torch.where(on, H2, z6))

    # ------------------------------------------------------------------ the substep
    def substep(self, st, prm, s, S, dt, lastH, h, prog, cand):
        x, v, vl, q = st
        k, rho, w, drag = prm
        if prog is not None:  # applyRootProgramK (batch.cu:122-159), position part
            p0, v0, p1 = prog
            A = (2.0 / (dt * dt)) * ((p1 - p0) - v0 * dt)
            tau = float(s) * h
            xr = (p0 + v0 * tau) + (0.5 * tau * tau) * A
            vr = v0 + (tau + 0.5 * h) * A
            rm = torch.zeros(self.nv, 1, dtype=torch.bool)
            rm[self.roots] = True
            x = torch.where(rm, torch.zeros_like(x).index_add(0, self.roots, xr), x)
            v = torch.where(rm, torch.zeros_like(v).index_add(0, self.roots, vr), v)
        free = self.free
        g = self.g
        # initItr
        d = h * v
        y = d + (h * h) * g
        g2 = float((g * g).sum())
        if g2 > 0:
            a = (v - vl) / lastH
            sg = torch.clamp(_dot(a, g) / g2, 0.0, 1.0)
            d = torch.where(free, d + (h * h * sg) * g, d)
        yv = torch.where(free, y, v)
        vl_new = v
        X = x
        dmp = self.damping / h
        if self.ropeContact:
            pairs, lim = self._detect(X, y)
        invw = torch.where(free, w, torch.ones_like(w))
        ex = torch.tensor([1.0, 0.0, 0.0], dtype=self.dtype)
        kap = self.cKappa / h
        for _ in range(self.numItr):
            # position sweep (cosseratItr)
            coef = 1.0 / (invw * h * h)
            f = torch.where(free, coef * (yv - d), torch.zeros_like(d))
            H = torch.where(free, torch.cat([coef.expand(-1, 3), torch.zeros(self.nv, 3, dtype=self.dtype)], -1),
                            torch.zeros(self.nv, 6, dtype=self.dtype))
            if self.planes or self.meshes:
                df, dH = self._surfaces(X, d, kap * kap / invw, h, cand)
                f, H = f + df, H + dH
            invl = 1.0 / self.l[:, None]
            c = ((X[self.nxt] - X) + (d[self.nxt] - d)) * invl - qrot(q, ex.expand_as(X))
            kk = k * invl
            dd = kk * invl
            hn = self.hasNext
            f = torch.where(hn, f + (kk * c - (dmp * dd) * d), f)
            f2 = torch.where(hn, -kk * c - (dmp * dd) * d[self.nxt], torch.zeros_like(f))
            dd = dd * (1.0 + dmp)
            diag = torch.cat([dd.expand(-1, 3), torch.zeros(self.nv, 3, dtype=self.dtype)], -1)
            H = torch.where(hn, H + diag, H)
            H2 = torch.where(hn, diag, torch.zeros_like(H))
            if self.ropeContact and pairs[0].numel() > 0:
                df, dH, df2, dH2 = self._pair_terms(X, d, pairs, dmp)
                I = pairs[0]
                f, H = f.index_add(0, I, df), H.index_add(0, I, dH)
                f2, H2 = f2.index_add(0, I, df2), H2.index_add(0, I, dH2)
            F = torch.where(self.hasPrev, f + f2[self.prv], f)
            Ht = torch.where(self.hasPrev, H + H2[self.prv], H)
            Hs = torch.where(free, Ht, torch.tensor([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], dtype=self.dtype).expand_as(Ht))
            dn = torch.where(free, d + self.accel * _ssolve(Hs, F), d)
            if self.ropeContact:  # step clamp (YB/sim/cosserat.cu:206-215)
                ln = torch.sqrt(_dot(dn, dn))
                cl = free & (ln > lim) & (ln > 0)
                self.stats["clamped"] = self.stats.get("clamped", 0) + int(cl.sum())
                dn = torch.where(cl, dn * (lim / torch.where(cl, ln, torch.ones_like(ln))), dn)
            # orientation sweep (quaternionLambdaItr)
            upd = (~self.fix) & self.hasNext
            e = ((X[self.nxt] - X) + (dn[self.nxt] - dn)) / self.l[:, None]
            e = e * (-2.0 * k)
            q0 = q
            qp, qn = q[self.prv], q[self.nxt]
            rp = rho[self.prv]
            s1 = torch.where(_dot(qmul(qconj(qp), q0), rp) > 0, 1.0, -1.0)
            s2 = torch.where(_dot(qmul(qconj(q0), qn), rho) > 0, 1.0, -1.0)
            b = torch.where(self.hasPrev, s1 * qmul(qp, rp), torch.zeros_like(q))
            b = b + torch.where(self.hasNextO, s2 * qmul(qn, qconj(rho)), torch.zeros_like(q))
            le = _safe_len(e, upd & (_dot(e, e) > 0))
            le = torch.where(_dot(e, e) > 0, le, torch.zeros_like(le))
            lb = _safe_len(b, upd & (_dot(b, b) > 0))
            lb = torch.where(_dot(b, b) > 0, lb, torch.zeros_like(lb))
            lam = le + lb
            exq = torch.tensor([1.0, 0.0, 0.0, 0.0], dtype=self.dtype).expand_as(q)
            prod = qmul(qmul(torch.cat([e, torch.zeros_like(e[:, :1])], -1), b), exq)
            Z = prod + lam * b
            Zs = torch.where(upd, Z, exq)
            qnew = Zs * (1.0 / torch.sqrt(_dot(Zs, Zs)))
            q = torch.where(upd, qnew, q)
            d = dn
        # endItr
        vv = d * (1.0 / h)
        if self.dragQuad > 0:
            tq = torch.where(self.hasNext, q, q[self.prv])
            t = qrot(tq, ex.expand_as(X))
            vt = t * _dot(t, vv)
            vn = vv - vt
            vn = vn / (1.0 + self.dragQuad * torch.sqrt(_dot(vn, vn) + 0.0) * h)
            vt = vt / (1.0 + self.dragQuadT * torch.sqrt(_dot(vt, vt) + 0.0) * h)
            vv = vn + vt
        v_new = torch.where(free, vv * (1.0 - drag * h), yv)
        x_new = X + d
        cand = self._candidates(x_new, d)
        return (x_new, v_new, vl_new, q), cand

    def rollout(self, ea, ei, mass, drag, p0, v0, p1, dt, S, num_itr, h_prev=None):
        """Frames of the root program from the construction state. ea / ei [nv] per segment (last vertex
        ignored), mass [nv] (pinned ignored), drag scalar, p0 / v0 / p1 [F, N, 3]. Returns positions
        [F, nv, 3] after each frame."""
        self.numItr = num_itr
        ea, ei, mass = [torch.as_tensor(a, dtype=self.dtype) for a in (ea, ei, mass)]
        p0, v0, p1 = [torch.as_tensor(a, dtype=self.dtype) if not torch.is_tensor(a) else a for a in (p0, v0, p1)]
        k = torch.where(self.hasNext.squeeze(-1), ea * self.l, self.k0).unsqueeze(-1)
        rho_len = torch.where(self.hasNext.squeeze(-1), 4.0 * ei / self.l, self.rho_len0).unsqueeze(-1)
        rho = rho_len * self.rho_unit
        w = torch.where(self.free.squeeze(-1), 1.0 / mass, self.w0).unsqueeze(-1)
        drag = torch.as_tensor(drag, dtype=self.dtype)
        st = (self.x0, self.v0, self.vl0, self.q0)
        h = self.h_init if h_prev is None else h_prev
        out = []
        for f in range(p0.shape[0]):
            lastH, h = h, dt / S
            cand = torch.ones(self.nv, len(self.meshes), dtype=torch.bool)
            for s in range(S):
                st, cand = self.substep(st, (k, rho, w, drag), s, S, dt, lastH, h, (p0[f], v0[f], p1[f]), cand)
            out.append(st[0])
        return torch.stack(out, 0)
