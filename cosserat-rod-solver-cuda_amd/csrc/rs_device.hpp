// Per-vertex device routines of the substep, shared by the per-substep kernels and the persistent
// frame kernel so both run exactly the same arithmetic. Each cites the reference code whose
// behaviour it reproduces (paths relative to the reference repository; YB =
// third_party/yarnball/KittenEngine/YarnBall).
#pragma once

#include "rs_math.hpp"
#include "rs_types.hpp"

namespace rs {

__device__ __forceinline__ f3 ld3(const float4& a) { return mk3(a.x, a.y, a.z); }
__device__ __forceinline__ q4 ldq(const float4& a) { return q4{a.x, a.y, a.z, a.w}; }
__device__ __forceinline__ float4 st3(f3 a) { return make_float4(a.x, a.y, a.z, 0.f); }
__device__ __forceinline__ float4 stq(q4 a) { return make_float4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ uint32_t flags_of(const float4& prm) { return __float_as_uint(prm.w); }

// World AABB of the object-frame lattice box under a pose (YB/sim/contact.cuh:79-87).
__device__ __forceinline__ void world_box(const Grid& g, f3 pos, q4 q, f3& lo, f3& hi) {
    lo = mk3(INFINITY, INFINITY, INFINITY);
    hi = mk3(-INFINITY, -INFINITY, -INFINITY);
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const f3 corner = mk3((c & 1) ? g.hx : g.ox, (c & 2) ? g.hy : g.oy, (c & 4) ? g.hz : g.oz);
        const f3 w = qrot(q, corner) + pos;
        lo = vmin(lo, w);
        hi = vmax(hi, w);
    }
}

// Pose of a kinematic mesh at the end of substep `sub` of `nSub` (YB/sim/iteration.cu:12-33): constant
// linear and angular velocity over the frame; a target equal to the current pose stays exactly static.
__device__ __forceinline__ Pose substep_pose(const Motion& M, const Grid& g, int sub, int nSub, float h) {
    const float a = float(sub + 1) / float(nSub);
    const float invDt = 1.f / (float(nSub) * h);
    const f3 rot = mk3(M.rx, M.ry, M.rz), dp = mk3(M.dpx, M.dpy, M.dpz), p0 = mk3(M.p0x, M.p0y, M.p0z);
    const q4 q0{M.q0x, M.q0y, M.q0z, M.q0w};
    const float th = len(rot);
    f3 pos, vel, om;
    q4 q;
    if (th > 0.f || dp.x != 0.f || dp.y != 0.f || dp.z != 0.f) {
        pos = p0 + a * dp;
        q4 dq{0.f, 0.f, 0.f, 1.f};
        if (th > 0.f) {
            const float half = (a * th) * 0.5f;
            dq = mkq(sinf(half) * (rot / th), cosf(half));
        }
        q = qnormalize(qmul(dq, q0));
        vel = dp * invDt;
        om = rot * invDt;
    } else {
        pos = p0;
        q = q0;
        vel = mk3(0.f, 0.f, 0.f);
        om = vel;
    }
    f3 lo, hi;
    world_box(g, pos, q, lo, hi);
    // the padding build_candidates adds for a moving mesh, once per pose instead of once per vertex
    const float bpad = h * (len(vel) + len(om) * 0.5f * len(hi - lo));
    return Pose{pos.x, pos.y, pos.z, q.x, q.y, q.z, q.w, vel.x, vel.y, vel.z, om.x, om.y, om.z,
                lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, bpad};
}

__device__ __forceinline__ Pose load_pose(const Pose* p) {
    const float4* s = reinterpret_cast<const float4*>(p);
    const float4 a = s[0], b = s[1], c = s[2], d = s[3], e = s[4];
    return Pose{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w, e.x, e.y, e.z, e.w};
}
__device__ __forceinline__ void store_pose(Pose* p, const Pose& P) {
    float4* s = reinterpret_cast<float4*>(p);
    s[0] = make_float4(P.px, P.py, P.pz, P.qx);
    s[1] = make_float4(P.qy, P.qz, P.qw, P.vx);
    s[2] = make_float4(P.vy, P.vz, P.wx, P.wy);
    s[3] = make_float4(P.wz, P.lox, P.loy, P.loz);
    s[4] = make_float4(P.hix, P.hiy, P.hiz, P.bpad);
}

// Penalty + regularised Coulomb contact at one surface (YB/sim/contact.cuh:27-46).
template <bool F>
__device__ __forceinline__ void surface_contact(const Params& P, f3 n, float gap, f3 d, float k, float muS, float muK,
                                                f3& f, s3& H) {
    using A = Ar<F>;
    using M = V<F>;
    const float eps = P.cVSlip;
    const float N = -k * gap;
    f = M::axpy(N, n, f);
    s3 nn = M::souter(n);
    M::sadd_scaled(H, nn, k);
    const f3 ut = M::axpy(-M::dot(n, d), n, d);
    const float ul = M::len(ut);
    const float vt = A::div(ul, P.h);
    float kt;
    if (vt <= eps) {
        kt = A::div(muS * N, eps * P.h);
    } else {
        const float s = 1.f - fminf(1.f, A::div(vt, eps) - 1.f);
        kt = A::div(A::fma(s, muS, (1.f - s) * muK) * N, ul);
    }
    f = M::axpy(-kt, ut, f);
    sadd_diag(nn, -1.f);
    M::sadd_scaled(H, nn, -kt);
}

// The header of a lattice as the three 16-byte rows a sample reads. Read field by field, the compiler sank the
// loads of the fields the short-circuit bounds test does not need first (the sizes, the node pointer, the
// spacing, the friction) below it, one dependent round each: a contact waited on three rounds of header loads
// before the node values and two after them. From the rows it issues most of them with the pose. Where it still
// splits a row and sinks a field (the frame kernel: the first size and the node pointer), forcing that field
// into the first round measured slower.
__device__ __forceinline__ Grid load_grid(const Grid* gp) {
    const uint4* r = reinterpret_cast<const uint4*>(gp);
    const uint4 a = r[0], b = r[1], c = r[2];
    Grid g;
    g.v = reinterpret_cast<const float*>((uint64_t)a.x | ((uint64_t)a.y << 32));
    g.nx = (int)a.z;
    g.ny = (int)a.w;
    g.nz = (int)b.x;
    g.ox = __uint_as_float(b.y);
    g.oy = __uint_as_float(b.z);
    g.oz = __uint_as_float(b.w);
    g.invH = __uint_as_float(c.x);
    g.h = __uint_as_float(c.y);
    g.muS = __uint_as_float(c.z);
    g.muK = __uint_as_float(c.w);
    g.hx = g.hy = g.hz = g.pad_ = 0.f;  // not loaded
    return g;
}

// Trilinear SDF sample and the exact gradient of the interpolant (YB/sim/contact.cuh:57-76). With MIXED the
// interpolant's mixed second derivatives (d2/dxdy, d2/dxdz, d2/dydz; the pure ones vanish) are returned too.
// `g` is a header from load_grid.
template <bool MIXED, bool F>
__device__ __forceinline__ bool sdf_sample_t(const Grid& g, f3 p, float& phi, f3& grad, f3& mixed) {
    using A = Ar<F>;
    const f3 o = mk3(g.ox, g.oy, g.oz);
    const float invH = g.invH;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const f3 u = (p - o) * invH;
    if (!(u.x >= 0.f && u.y >= 0.f && u.z >= 0.f && u.x < float(nx - 1) && u.y < float(ny - 1) && u.z < float(nz - 1)))
        return false;
    const int i = (int)u.x, j = (int)u.y, k = (int)u.z;
    const f3 fr = (p - (o + mk3(float(i), float(j), float(k)) * g.h)) * invH;
    const float fx = fminf(fmaxf(fr.x, 0.f), 1.f), fy = fminf(fmaxf(fr.y, 0.f), 1.f), fz = fminf(fmaxf(fr.z, 0.f), 1.f);
    const int sy = nz, sx = ny * nz;
    // The node values are device memory: global loads, which wait on the vector-memory counter only (a
    // generic pointer made them flat loads, which also wait on the LDS counter).
    const __attribute__((address_space(1))) float* c =
        (const __attribute__((address_space(1))) float*)g.v + (size_t)i * sx + (size_t)j * sy + k;
    const float c000 = c[0], c001 = c[1], c010 = c[sy], c011 = c[sy + 1];
    const float c100 = c[sx], c101 = c[sx + 1], c110 = c[sx + sy], c111 = c[sx + sy + 1];
    const float dx00 = c100 - c000, dx01 = c101 - c001, dx10 = c110 - c010, dx11 = c111 - c011;
    const float c00 = A::fma(fx, dx00, c000), c01 = A::fma(fx, dx01, c001), c10 = A::fma(fx, dx10, c010),
                c11 = A::fma(fx, dx11, c011);
    const float dy0 = c10 - c00, dy1 = c11 - c01;
    const float c0 = A::fma(fy, dy0, c00), c1 = A::fma(fy, dy1, c01);
    phi = A::fma(fz, c1 - c0, c0);
    const float gx0 = A::fma(fy, dx10 - dx00, dx00), gx1 = A::fma(fy, dx11 - dx01, dx01);
    grad = mk3(A::fma(fz, gx1 - gx0, gx0), A::fma(fz, dy1 - dy0, dy0), c1 - c0) * invH;
    if (MIXED) {
        const float i2 = invH * invH;
        mixed = mk3(A::fma(fz, (dx11 - dx01) - (dx10 - dx00), dx10 - dx00) * i2, (gx1 - gx0) * i2, (dy1 - dy0) * i2);
    }
    return true;
}
template <bool F>
__device__ __forceinline__ bool sdf_sample(const Grid& g, f3 p, float& phi, f3& grad) {
    f3 unused;
    return sdf_sample_t<false, F>(g, p, phi, grad, unused);
}

// Contact of a vertex with one posed mesh at the probe (pos - P.pos) + dx (YB/sim/contact.cuh:94-110). `g` is
// the mesh's header from load_grid. Reads only the pose fields it needs (the velocities only while meshes move,
// and only for a contact).
template <bool F>
__device__ __forceinline__ bool sdf_contact(const Params& Pm, const Grid& g, const Pose* P, f3 p, f3 d, f3& n,
                                            float& gap, f3& dRel) {
    using M = V<F>;
    const float4* s = reinterpret_cast<const float4*>(P);
    const float4 a = s[0], b = s[1];
    const q4 q{a.w, b.x, b.y, b.z};
    const f3 pp = mk3(a.x, a.y, a.z);
    float phi;
    f3 gr;
    if (!sdf_sample<F>(g, M::qrot(qconj(q), (p - pp) + d), phi, gr)) return false;
    gap = phi - Pm.cRadius;
    if (gap >= 0.f) return false;
    const float gl = M::len(gr);
    if (!(gl > 0.f)) return false;
    n = M::qrot(q, M::div(gr, gl));
    dRel = d;
    if (Pm.sdfMoving) {  // read here: issued with the rest of the pose, they made 4096 contact ropes 2 % slower
        const float4 c = s[2], e = s[3];
        dRel = dRel - Pm.h * (mk3(b.w, c.x, c.y) + M::cross(mk3(c.z, c.w, e.x), p - pp));
    }
    return true;
}

// Is the world-space probe x + dx near the world box of a posed mesh's lattice? Outside the box (grown by a
// margin far above the rounding of the box corners and of the probe) the probe is outside the lattice and
// sdf_contact finds no contact, so skipping the mesh changes no result. Used where every mesh is visited (a
// frame's first substep) to save the probe's rotation into the lattice frame.
__device__ __forceinline__ bool probe_near_box(const Pose* P, f3 q) {
    const float4* s = reinterpret_cast<const float4*>(P);
    const float4 e = s[3], f = s[4];
    const float mg = 1e-4f + 2e-6f * fmaxf(fabsf(q.x), fmaxf(fabsf(q.y), fabsf(q.z)));
    return q.x >= e.y - mg && q.x <= f.x + mg && q.y >= e.z - mg && q.y <= f.y + mg && q.z >= e.w - mg &&
           q.z <= f.z + mg;
}

// Broad-phase candidates are kept packed: 6 bits per mesh id, up to 8 ids in 64 bits.
__device__ __forceinline__ int cand_get(uint64_t packed, int c) { return (int)((packed >> (6 * c)) & 63u); }

// Planes, then candidate meshes (YB/sim/contact.cuh:141-163). `ncand` < 0: every mesh. HITS: bit c of *hits
// is set when candidate c (ncand >= 0) is in contact.
template <bool F, bool HITS = false>
__device__ __forceinline__ void surfaces_at(const Params& P, const Grid* grids, const Pose* envPoses, int ncand,
                                            uint64_t cand, f3 p, f3 d, float k, f3& f, s3& H,
                                            uint32_t* hits = nullptr) {
    for (int pl = 0; pl < P.numPlanes; pl++) {
        const f3 n = mk3(P.planes[pl][0], P.planes[pl][1], P.planes[pl][2]);
        const float gap = ((V<F>::dot(n, p) - P.planes[pl][3]) + V<F>::dot(n, d)) - P.cRadius;
        if (gap >= 0.f) continue;
        surface_contact<F>(P, n, gap, d, k, P.cMuS, P.cMuK, f, H);
    }
    if (P.numSdf > 0) {
        const int cnt = ncand < 0 ? P.numSdf : ncand;
        for (int c = 0; c < cnt; c++) {
            const int m = ncand < 0 ? c : cand_get(cand, c);
            if (ncand < 0 && !probe_near_box(envPoses + m, p + d)) continue;
            f3 n, dRel;
            float gap;
            const Grid g = load_grid(grids + m);
            if (!sdf_contact<F>(P, g, envPoses + m, p, d, n, gap, dRel)) continue;
            if constexpr (HITS) {
                if (ncand >= 0) *hits |= 1u << c;  // listed candidates only (c < MAX_CAND)
            }
            surface_contact<F>(P, n, gap, dRel, k, g.muS, g.muK, f, H);
        }
    }
}

// Wrench accumulation of the force readout. The wrench array is [ropes][MAX_SDF][6]; `key` = rope * MAX_SDF
// + mesh names one wrench. Lanes of a wave holding a contribution to the same wrench are summed across the
// wave first, and one lane issues the six global updates: a wave of vertices resting on one mesh costs six
// atomics instead of six per lane. Fixed-point accumulation (P.forcesFixed): every contribution is rounded
// to a 64-bit integer in units of 2^-WRENCH_FIX_BITS before any sum, and integer sums are exact, so the
// result depends neither on the order of the updates nor on which lanes share a wave: it is the same bit
// for bit in every dispatch form, cell variant and run.
constexpr int WRENCH_FIX_BITS = 32;
constexpr float WRENCH_FIX_MAX = 1.0e9f;  // |contribution| clamp, keeps 2^32 * value inside int64

__device__ __forceinline__ long long wrench_to_fixed(float w) {
    const float c = fminf(fmaxf(w, -WRENCH_FIX_MAX), WRENCH_FIX_MAX);
    return __double2ll_rn((double)c * 4294967296.0);
}

// Sums over each row of 16 lanes, left in every lane of the row (DPP: quad swaps, then the half-row and row
// mirrors). Every lane of the wave must be active.
__device__ __forceinline__ float row16_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));  // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));  // row_mirror
    return v;
}
template <int CTRL>
__device__ __forceinline__ long long mov_dpp64(long long v) {
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(unsigned long long)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)((unsigned long long)v >> 32), CTRL, 0xF, 0xF, false);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ long long row16_sum(long long v) {
    v += mov_dpp64<0xB1>(v);
    v += mov_dpp64<0x4E>(v);
    v += mov_dpp64<0x141>(v);
    v += mov_dpp64<0x140>(v);
    return v;
}
// Wave total from the row sums, as a wave-uniform value: ((row0 + row1) + row2) + row3.
__device__ __forceinline__ float wave_total(float v) {
    v = row16_sum(v);
    const int b = __float_as_int(v);
    return ((__int_as_float(__builtin_amdgcn_readlane(b, 0)) + __int_as_float(__builtin_amdgcn_readlane(b, 16))) +
            __int_as_float(__builtin_amdgcn_readlane(b, 32))) +
           __int_as_float(__builtin_amdgcn_readlane(b, 48));
}
__device__ __forceinline__ long long readlane64(long long v, int l) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(unsigned long long)v, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)v >> 32), l);
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ long long wave_total(long long v) {
    v = row16_sum(v);
    return readlane64(v, 0) + readlane64(v, 16) + readlane64(v, 32) + readlane64(v, 48);
}

// Lanes sharing a wrench sum across the wave when more than this many hold a contribution to it; fewer
// update it directly (their atomics hardly contend).
constexpr int WAVE_SUM_MIN_LANES = 3;

// Every lane of the wave calls this, in uniform control flow; `has` marks the lanes with a contribution.
__device__ __forceinline__ void wave_wrench_flush(const Params& P, const Dev& D, bool has, int key, const float (&w)[6]) {
    unsigned long long pend = __ballot(has);
    while (pend) {
        const int lead = __ffsll((long long)pend) - 1;
        const int lk = __builtin_amdgcn_readlane(key, lead);
        const bool mine = has && key == lk;
        const unsigned long long grp = __ballot(mine);
        if (__popcll(grp) < WAVE_SUM_MIN_LANES) {
            if (mine) {
                if (P.forcesFixed) {
                    unsigned long long* dst = D.wrenchFx + (size_t)lk * 6;
#pragma unroll
                    for (int j = 0; j < 6; j++) atomicAdd(dst + j, (unsigned long long)wrench_to_fixed(w[j]));
                } else {
                    float* dst = D.wrench + (size_t)lk * 6;
#pragma unroll
                    for (int j = 0; j < 6; j++) atomicAdd(dst + j, w[j]);
                }
            }
        } else if (P.forcesFixed) {
            unsigned long long* dst = D.wrenchFx + (size_t)lk * 6;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const long long t = wave_total(mine ? wrench_to_fixed(w[j]) : 0ll);
                if (__lane_id() == lead) atomicAdd(dst + j, (unsigned long long)t);
            }
        } else {
            float* dst = D.wrench + (size_t)lk * 6;
#pragma unroll
            for (int j = 0; j < 6; j++) {
                const float t = wave_total(mine ? w[j] : 0.f);
                if (__lane_id() == lead) atomicAdd(dst + j, t);
            }
        }
        has = has && !mine;
        pend &= ~grp;
    }
}

// Force readout at the end of a substep (YB/sim/contact.cuh:172-204): the surface law re-evaluated at the
// final displacement gives the force on the vertex (returned) and the wrench (-f, r x -f) the vertex applies
// to each mesh about the mesh origin, r = (x - P.pos) + dx - R n (contact.cuh:196). Wave-cooperative: every
// lane of the wave calls it in uniform control flow, `on` false for lanes without a free vertex; the mesh loop
// runs to the wave's largest candidate count so the wrench sums see every lane.
template <bool F>
__device__ __forceinline__ f3 surface_readout(const Params& P, const Dev& D, const Pose* envPoses, int rope, bool on,
                                              int ncand, uint64_t cand, f3 p, f3 d, float k) {
    s3 H = szero();
    f3 fv = mk3(0.f, 0.f, 0.f);
    if (on) {
        for (int pl = 0; pl < P.numPlanes; pl++) {
            const f3 n = mk3(P.planes[pl][0], P.planes[pl][1], P.planes[pl][2]);
            const float gap = ((V<F>::dot(n, p) - P.planes[pl][3]) + V<F>::dot(n, d)) - P.cRadius;
            if (gap >= 0.f) continue;
            surface_contact<F>(P, n, gap, d, k, P.cMuS, P.cMuK, fv, H);
        }
    }
    if (P.numSdf > 0) {
        const int cnt = on ? (ncand < 0 ? P.numSdf : ncand) : 0;
        for (int c = 0; __any(c < cnt); c++) {
            bool has = false;
            int key = 0;
            float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (c < cnt) {
                const int m = ncand < 0 ? c : cand_get(cand, c);
                f3 n, dRel;
                float gap;
                if ((ncand >= 0 || probe_near_box(envPoses + m, p + d)) &&
                    sdf_contact<F>(P, load_grid(D.grids + m), envPoses + m, p, d, n, gap, dRel)) {
                    f3 fm = mk3(0.f, 0.f, 0.f);
                    surface_contact<F>(P, n, gap, dRel, k, D.grids[m].muS, D.grids[m].muK, fm, H);
                    fv = fv + fm;
                    const Pose* Pp = envPoses + m;
                    const f3 r = ((p - mk3(Pp->px, Pp->py, Pp->pz)) + d) - P.cRadius * n;
                    const f3 tq = cross(r, -fm);
                    w[0] = -fm.x;
                    w[1] = -fm.y;
                    w[2] = -fm.z;
                    w[3] = tq.x;
                    w[4] = tq.y;
                    w[5] = tq.z;
                    has = true;
                    key = rope * MAX_SDF + m;
                }
            }
            wave_wrench_flush(P, D, has, key, w);
        }
    }
    return fv;
}

// Is probe p inside a pose's world box (rows e, f of the pose) grown by the vertex padding?
__device__ __forceinline__ bool in_padded_box(const Params& P, f3 p, float padv, float4 e, float4 f) {
    // moving meshes: + h (|v| + |w| |hi - lo| / 2), evaluated with the pose (substep_pose)
    const float pm = P.sdfMoving ? padv + f.w : padv;
    return (p.x >= e.y - pm) & (p.x <= f.x + pm) & (p.y >= e.z - pm) & (p.y <= f.y + pm) & (p.z >= e.w - pm) &
           (p.z <= f.z + pm);
}
// Candidate meshes of a vertex for the next substep (YB/sim/contact.cuh:121-136). Returns the count
// (-1 on overflow) and the packed ids.
__device__ __forceinline__ int build_candidates(const Params& P, const Pose* envPoses, f3 p, float padv, uint64_t& packed) {
    int n = 0;
    packed = 0;
    const float4* s = reinterpret_cast<const float4*>(envPoses);  // 5 float4 per pose; the box is rows 3 and 4
    int m = 0;
    // Four meshes' boxes are loaded before any is tested: one round of loads for four meshes, not one each.
    for (; m + 4 <= P.numSdf; m += 4) {
        float4 e[4], f[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            e[u] = s[5 * (m + u) + 3];
            f[u] = s[5 * (m + u) + 4];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const bool in = in_padded_box(P, p, padv, e[u], f[u]);
            packed |= (in && n < MAX_CAND) ? (uint64_t)(m + u) << (6 * n) : 0ull;
            n += in ? 1 : 0;
        }
    }
    for (; m < P.numSdf; m++) {
        const bool in = in_padded_box(P, p, padv, s[5 * m + 3], s[5 * m + 4]);
        packed |= (in && n < MAX_CAND) ? (uint64_t)m << (6 * n) : 0ull;
        n += in ? 1 : 0;
    }
    return n <= MAX_CAND ? n : -1;
}

// Mesh counts above which the frame kernels' broad phase splits the meshes over the lanes of a chunk.
constexpr int CHUNK_BROADPHASE_MIN_MESHES = 16;

// build_candidates over the meshes of `mask` only (ascending): the same list whenever `mask` holds every mesh
// the full scan would accept.
__device__ __forceinline__ int build_candidates_masked(const Params& P, const Pose* envPoses, f3 p, float padv,
                                                       uint64_t mask, uint64_t& packed) {
    int n = 0;
    packed = 0;
    while (mask) {
        const int m = __builtin_ctzll(mask);
        mask &= mask - 1;
        const float4* s = reinterpret_cast<const float4*>(envPoses + m);
        const float4 e = s[3], f = s[4];
        const f3 lo = mk3(e.y, e.z, e.w), hi = mk3(f.x, f.y, f.z);
        const float pm = P.sdfMoving ? padv + f.w : padv;
        if (p.x >= lo.x - pm && p.x <= hi.x + pm && p.y >= lo.y - pm && p.y <= hi.y + pm && p.z >= lo.z - pm &&
            p.z <= hi.z + pm) {
            if (n < MAX_CAND) packed |= (uint64_t)m << (6 * n);
            n++;
        }
    }
    return n <= MAX_CAND ? n : -1;
}

// Meshes that may accept some vertex of a chunk of 16 consecutive lanes whose vertices share one rope (and so
// one pose array): the 16 lanes split the meshes and test each against the chunk's box of the vertices'
// padded positions, grown by the moving-mesh padding and a margin far above the rounding of the per-vertex
// test; the 16 partial masks are OR-ed. Lanes with act false add nothing. Every lane of the wave calls it in
// uniform control flow.
__device__ __forceinline__ uint64_t chunk_mesh_mask(const Params& P, const Pose* envPoses, bool act, f3 p, float padv) {
    f3 lo = act ? mk3(p.x - padv, p.y - padv, p.z - padv) : mk3(INFINITY, INFINITY, INFINITY);
    f3 hi = act ? mk3(p.x + padv, p.y + padv, p.z + padv) : mk3(-INFINITY, -INFINITY, -INFINITY);
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        lo = vmin(lo, mk3(__shfl_xor(lo.x, o), __shfl_xor(lo.y, o), __shfl_xor(lo.z, o)));
        hi = vmax(hi, mk3(__shfl_xor(hi.x, o), __shfl_xor(hi.y, o), __shfl_xor(hi.z, o)));
    }
    uint32_t mlo = 0, mhi = 0;
    if (lo.x <= hi.x) {  // some lane of the chunk is active
        const float mg = 1e-4f + 2e-6f * fmaxf(fmaxf(fabsf(lo.x), fabsf(hi.x)),
                                               fmaxf(fmaxf(fabsf(lo.y), fabsf(hi.y)), fmaxf(fabsf(lo.z), fabsf(hi.z))));
        const float4* s = reinterpret_cast<const float4*>(envPoses);  // 5 float4 per pose; the box is rows 3 and 4
        // The chunk box overlaps mesh m's grown box: branch-free, and a lane loads its four meshes' boxes (m, m + 16,
        // m + 32, m + 48) before testing any, one round of loads instead of four.
        auto overlaps = [&](float4 e, float4 f) {
            const float bp = (P.sdfMoving ? f.w : 0.f) + mg;
            return (lo.x - bp <= f.x) & (lo.y - bp <= f.y) & (lo.z - bp <= f.z) & (hi.x + bp >= e.y) &
                   (hi.y + bp >= e.z) & (hi.z + bp >= e.w);
        };
        uint64_t bits = 0;
        int m = __lane_id() & 15;
        for (; m + 48 < P.numSdf; m += 64) {
            float4 e[4], f[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                e[u] = s[5 * (m + 16 * u) + 3];
                f[u] = s[5 * (m + 16 * u) + 4];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) bits |= overlaps(e[u], f[u]) ? 1ull << (m + 16 * u) : 0ull;
        }
        for (; m < P.numSdf; m += 16) bits |= overlaps(s[5 * m + 3], s[5 * m + 4]) ? 1ull << m : 0ull;
        mlo = (uint32_t)bits;
        mhi = (uint32_t)(bits >> 32);
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        mlo |= __shfl_xor(mlo, o);
        mhi |= __shfl_xor(mhi, o);
    }
    return ((uint64_t)mhi << 32) | mlo;
}

// The gravity-acceleration clamp of initItr: the fraction of gravity the last substep's acceleration
// carried, (a . g) / |g|^2 with a = (v - v_last) / lastH (YB/sim/iteration.cu:47-56); unclamped.
template <bool F>
__device__ __forceinline__ float gravity_fraction(const Params& P, f3 v, f3 vl, f3 g, float g2) {
    return Ar<F>::div(V<F>::dot(V<F>::div(v - vl, P.lastH), g), g2);
}

// initItr for one vertex (YB/sim/iteration.cu:34-69): displacement guess, inertial target y, padding.
template <bool F>
__device__ __forceinline__ void init_vertex(const Params& P, float invMass, f3 v, f3 vl, f3& d, f3& y, float& padding) {
    const f3 g = mk3(P.gx, P.gy, P.gz);
    d = P.h * v;
    y = d + (P.h * P.h) * g;
    padding = P.scaledDetR + V<F>::len(y);
    if (invMass != 0.f) {
        const float g2 = V<F>::len2(g);
        if (g2 > 0.f) {
            const float s = clampf(gravity_fraction<F>(P, v, vl, g, g2), 0.f, 1.f);
            d = d + (P.h * P.h * s) * g;
        }
    }
}

// One rope-rope barrier + friction term of segment (i, i+1) against segment (j, j+1)
// (YB/sim/cosserat.cu:99-158), split in two: collision_eval computes everything the term adds (it does not
// depend on the accumulators), collision_apply adds it to f, H (vertex i) and f2, H2 (vertex i+1). The pair
// of calls performs exactly the operations of the undivided term, so a kernel may evaluate terms on other
// lanes and apply them in list order on the owner lane without changing a bit.
struct PairTerm {
    f3 n;                // contact normal
    float s1, s2;        // normal force factors on i and i+1
    float h1, h2;        // normal Hessian factors on i and i+1
    f3 u;                // tangential relative displacement
    float a1, a2;        // friction force factors on i and i+1
    float b1, b2;        // friction Hessian factors on i and i+1
    int on;              // bit 0: the barrier is active; bit 1: friction acts
};

template <bool F>
__device__ __forceinline__ PairTerm collision_eval(const Params& P, float dmp, f3 p0, f3 p1, f3 d0, f3 p1dx, f3 b0,
                                                   f3 b1, f3 db0, f3 db1) {
    using A = Ar<F>;
    using M = V<F>;
    PairTerm T;
    T.on = 0;
    const float invb = A::rcp(P.barrier);
    const float rad2 = 2.f * P.radius;
    const float kCol = P.kCollision * invb;
    float u, v;
    M::seg_closest(mk3(0.f, 0.f, 0.f), (p1 - p0) + (p1dx - d0), (b0 - p0) + (db0 - d0), (b1 - p0) + (db1 - d0), u, v);
    const f3 dpos = M::mix(p0, p1, u) - M::mix(b0, b1, v);
    const f3 ddpos = M::mix(d0, p1dx, u) - M::mix(db0, db1, v);
    f3 normal = dpos + ddpos;
    float dd = M::len(normal);
    if (!(dd > 0.f)) return T;
    normal = normal * A::rcp(dd);
    const float wy = u;
    const float wx = 1.f - u;
    dd = dd - rad2;
    dd = dd * invb;
    if (dd > 1.f) return T;
    dd = fmaxf(dd, 1e-3f);
    const float invd = A::rcp(dd);
    const float logd = A::log(dd);
    float dH = A::fma(-2.f, logd, A::fma(2.f + invd, invd, -3.f)) * kCol * invb;
    const float ff = -(1.f - dd) * A::fma(2.f * dd, logd, dd - 1.f) * invd * kCol;
    T.n = normal;
    T.s1 = A::fma(ff, wx, -(dmp * dH * wx * wx * M::dot(normal, d0)));
    T.s2 = A::fma(ff, wy, -(dmp * dH * wy * wy * M::dot(normal, p1dx)));
    dH = dH * (1.f + dmp);
    T.h1 = dH * wx * wx;
    T.h2 = dH * wy * wy;
    T.u = M::axpy(-M::dot(normal, ddpos), normal, ddpos);
    T.on = 1;
    const float ul = M::len(T.u);
    if (ul > 0.f) {
        const float f1 = fminf(P.kFriction, A::div(P.fricMu * ff, ul));
        T.a1 = -(f1 * wx);
        T.b1 = -(wx * wx * f1);
        T.a2 = -(f1 * wy);
        T.b2 = -(wy * wy * f1);
        T.on = 3;
    }
    return T;
}

template <bool F>
__device__ __forceinline__ void collision_apply(const PairTerm& T, f3& f, s3& H, f3& f2, s3& H2) {
    using M = V<F>;
    if (!(T.on & 1)) return;
    f = M::axpy(T.s1, T.n, f);
    f2 = M::axpy(T.s2, T.n, f2);
    s3 op = M::souter(T.n);
    M::sadd_scaled(H, op, T.h1);
    M::sadd_scaled(H2, op, T.h2);
    if (T.on & 2) {
        sadd_diag(op, -1.f);
        f = M::axpy(T.a1, T.u, f);
        M::sadd_scaled(H, op, T.b1);
        f2 = M::axpy(T.a2, T.u, f2);
        M::sadd_scaled(H2, op, T.b2);
    }
}

template <bool F>
__device__ __forceinline__ void collision_term(const Params& P, float dmp, f3 p0, f3 p1, f3 d0, f3 p1dx, f3 b0, f3 b1,
                                               f3 db0, f3 db1, f3& f, s3& H, f3& f2, s3& H2) {
    collision_apply<F>(collision_eval<F>(P, dmp, p0, p1, d0, p1dx, b0, b1, db0, db1), f, H, f2, H2);
}

// Inertia + surface terms of vertex i and the stretch term of segment (i, i+1)
// (YB/sim/cosserat.cu:41-88). Rope-rope terms are added by the caller.
template <bool F, bool HITS = false>
__device__ __forceinline__ void vertex_terms(const Params& P, const Grid* grids, const Pose* envPoses, int ncand,
                                             uint64_t cand, float invMass, float lRest, float kStretch, uint32_t fl,
                                             f3 p0, f3 d0, f3 y, q4 qi, f3 p1, f3 p1dx, f3& f, s3& H, f3& f2, s3& H2,
                                             uint32_t* hits = nullptr) {
    using A = Ar<F>;
    const float h = P.h;
    const float dmp = A::div(P.damping, h);
    f = mk3(0.f, 0.f, 0.f);
    H = szero();
    f2 = f;
    H2 = H;
    if (invMass != 0.f) {
        H = sdiag(A::rcp(invMass * h * h));
        f = A::rcp(h * h * invMass) * (y - d0);
        if (P.numPlanes > 0 || P.numSdf > 0) {
            const float kap = A::div(P.cKappa, h);
            surfaces_at<F, HITS>(P, grids, envPoses, ncand, cand, p0, d0, A::div(kap * kap, invMass), f, H, hits);
        }
    }
    if (fl & HAS_NEXT) {
        const float invl = A::rcp(lRest);
        const f3 c = ((p1 - p0) + (p1dx - d0)) * invl - V<F>::qrot_ex(qi);
        const float k = kStretch * invl;
        float dd = k * invl;
        const float dd0 = dmp * dd;
        f = f + mk3(A::fma(k, c.x, -(dd0 * d0.x)), A::fma(k, c.y, -(dd0 * d0.y)), A::fma(k, c.z, -(dd0 * d0.z)));
        f2 = f2 + mk3(A::fma(-k, c.x, -(dd0 * p1dx.x)), A::fma(-k, c.y, -(dd0 * p1dx.y)), A::fma(-k, c.z, -(dd0 * p1dx.z)));
        dd = dd * (1.f + dmp);
        sadd_diag(H, dd);
        sadd_diag(H2, dd);
    }
}

// Local Newton update of a free vertex with the optional step clamp (YB/sim/cosserat.cu:206-215).
template <bool F>
__device__ __forceinline__ f3 solve_update(const Params& P, f3 d, const s3& H, f3 f, bool limit, float stepLimit) {
    d = V<F>::axpy(P.accel, V<F>::ssolve(H, f), d);
    if (limit) {
        const float l = V<F>::len(d);
        if (l > stepLimit && l > 0.f) d = d * Ar<F>::div(stepLimit, l);
    }
    return d;
}

// Signs of the rest-twist terms of the frame update (YB/sim/cosserat.cu:253-262): sign <conj(q_{i-1}) q_i,
// rho_{i-1}> and sign <conj(q_i) q_{i+1}, rho_i>. Left and right quaternion multiplications are transposed
// by the conjugate (<conj(a) b, c> = <b, a c>, <a c, d> = <a, d conj(c)>), so the fast mode takes the same
// numbers as <q_i, t> with t = q_{i-1} rho_{i-1} and t = q_{i+1} conj(rho_i), the products the update needs
// anyway; IEEE mode evaluates the reference's form.
template <bool F>
__device__ __forceinline__ float twist_sign_prev(q4 qPrev, q4 q0, q4 qrPrev, q4 t) {
    const float c = F ? V<F>::qdot(q0, t) : V<F>::qdot(V<F>::qmul(qconj(qPrev), q0), qrPrev);
    return c > 0.f ? 1.f : -1.f;
}
template <bool F>
__device__ __forceinline__ float twist_sign_next(q4 q0, q4 qNext, q4 qr, q4 t) {
    const float c = F ? V<F>::qdot(q0, t) : V<F>::qdot(V<F>::qmul(qconj(q0), qNext), qr);
    return c > 0.f ? 1.f : -1.f;
}

// Closed-form quasi-static frame update of segment i (YB/sim/cosserat.cu:240-272).
template <bool F>
__device__ __forceinline__ q4 orient_update(uint32_t fl, float lRest, float kStretch, f3 p0, f3 p1, f3 d, f3 p1dx,
                                            q4 qPrev, q4 q0, q4 qNext, q4 qrPrev, q4 qr) {
    using A = Ar<F>;
    using M = V<F>;
    f3 e = M::div((p1 - p0) + (p1dx - d), lRest);
    e = e * (-2.f * kStretch);
    q4 b{0.f, 0.f, 0.f, 0.f};
    if (fl & HAS_PREV) {
        const q4 t = M::qmul(qPrev, qrPrev);
        b = qadd(b, qscale(twist_sign_prev<F>(qPrev, q0, qrPrev, t), t));
    }
    if (fl & HAS_NEXT_ORIENT) {
        const q4 t = M::qmul(qNext, qconj(qr));
        b = qadd(b, qscale(twist_sign_next<F>(q0, qNext, qr, t), t));
    }
    const float lambda = M::len(e) + M::qlen(b);
    const q4 prod = M::qmul_ex(M::qmul_pure(e, b));
    return M::qnormalize(q4{A::fma(lambda, b.x, prod.x), A::fma(lambda, b.y, prod.y), A::fma(lambda, b.z, prod.z),
                            A::fma(lambda, b.w, prod.w)});
}

// Velocity update with quadratic and linear drag (YB/sim/iteration.cu:85-112). `t` is the segment
// tangent (the tip uses its one real segment).
template <bool F>
__device__ __forceinline__ f3 end_velocity(const Params& P, f3 d, f3 t) {
    using M = V<F>;
    const float invH = Ar<F>::rcp(P.h);
    f3 v = d * invH;
    const float kn = P.dragQuad;
    if (kn > 0.f) {
        f3 vt = t * M::dot(t, v);
        f3 vn = v - vt;
        vn = M::div(vn, 1.f + kn * M::len(vn) * P.h);
        vt = M::div(vt, 1.f + P.dragQuadT * M::len(vt) * P.h);
        v = vn + vt;
    }
    return v * (1.f - P.drag * P.h);
}

// Mesh poses of a rope during substep s of a persistent frame kernel: the precomputed per-substep poses while
// meshes move, the committed poses otherwise.
__device__ __forceinline__ const Pose* cell_poses(const Params& P, const Dev& D, int rope, int s) {
    return P.sdfMoving ? D.posesS + ((size_t)s * P.nropes + rope) * P.numSdf : D.poses + (size_t)rope * MAX_SDF;
}

// SLERP of the root program (csrc/batch.cu:106-114).
__device__ __forceinline__ q4 slerp4(q4 a, q4 b, float t) {
    float d = qdot(a, b);
    if (d < 0.f) {
        b = q4{-b.x, -b.y, -b.z, -b.w};
        d = -d;
    }
    if (d > 0.9995f) {
        const q4 m{a.x * (1.f - t) + b.x * t, a.y * (1.f - t) + b.y * t, a.z * (1.f - t) + b.z * t,
                   a.w * (1.f - t) + b.w * t};
        return qnormalize(m);
    }
    const float th = acosf(fminf(d, 1.f));
    const float s = sinf(th);
    return qadd(qscale(sinf((1.f - t) * th) / s, a), qscale(sinf(t * th) / s, b));
}

// Root of rope e at substep s of the second-order motion program (csrc/batch.cu:122-159).
__device__ __forceinline__ void root_program(const Drive& D, int e, int s, int nSub, float h, f3& x, f3& v, q4& q,
                                             bool& setQ) {
    const f3 P0 = mk3(D.p0[3 * e], D.p0[3 * e + 1], D.p0[3 * e + 2]);
    const f3 V0 = mk3(D.v0[3 * e], D.v0[3 * e + 1], D.v0[3 * e + 2]);
    const f3 P1 = mk3(D.p1[3 * e], D.p1[3 * e + 1], D.p1[3 * e + 2]);
    const float dt = D.dt;
    const f3 A = (2.f / (dt * dt)) * ((P1 - P0) - V0 * dt);
    const float tau = float(s) * h;
    x = (P0 + V0 * tau) + (0.5f * tau * tau) * A;
    v = V0 + (tau + 0.5f * h) * A;
    setQ = D.hasQ0 != 0;
    if (setQ) {
        const q4 qa = qunit_or_identity(q4{D.q0[4 * e], D.q0[4 * e + 1], D.q0[4 * e + 2], D.q0[4 * e + 3]});
        q4 r = qa;
        if (D.hasQ1) {
            const q4 qb = qunit_or_identity(q4{D.q1[4 * e], D.q1[4 * e + 1], D.q1[4 * e + 2], D.q1[4 * e + 3]});
            r = slerp4(qa, qb, float(s + 1) / float(nSub));
        }
        q = qunit_or_identity(r);
    }
}

// Exact rope-rope candidate test of segment (a0, a0 + a1) against (b0, b1) relative to a0, radius r
// (csrc/batch.cu:38-42). Returns the squared distance in `d2` for the step limit.
__device__ __forceinline__ bool pair_hit(f3 a1, f3 b0, f3 b1, float r, float& d2) {
    float u, v;
    seg_closest(mk3(0.f, 0.f, 0.f), a1, b0, b1, u, v);
    const f3 d = u * a1 - mix(b0, b1, v);
    d2 = len2(d);
    return d2 < r * r;
}

}  // namespace rs
