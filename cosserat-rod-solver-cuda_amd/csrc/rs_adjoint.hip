// Adjoint of a recorded frame, with its host launchers.
//
// One workgroup per rope (no rope-rope contact, so ropes are independent), two kernels per frame.
// Pass 1 (k_adjoint_forward) recomputes the frame from its checkpoint. It calls the same device
// routines in the same order as the forward kernels, so the results are identical. It writes every
// substep's inputs and every iteration's iterate d^k, frame q^k, Hessian H and Newton direction to a
// history buffer, and checks that the frame-end state equals the next checkpoint bit for bit.
// Pass 2 (k_adjoint_reverse) walks the history backwards: through the end update, the iterations
// (orientation sweep, then position sweep), the init and the root program. It accumulates the
// adjoints of
// - the state,
// - per-vertex k = EA l, the rest Darboux vector and the mass,
// - the linear drag,
// - the root program inputs.
//
// Notation: for a Newton update d' = d + a H^-1 F, the adjoint u = a H^-1 d̄' gives F̄ = u and
// H̄ = -u δ^T with δ = H^-1 F. The pair (u, δ) stands for H̄; every H term here is s I or s n n^T, so
// only tr H̄ = -u.δ, n^T H̄ n = -(u.n)(δ.n) and (H̄ + H̄^T) n = -(u (δ.n) + δ (u.n)) are needed.
#include <algorithm>

#include "rs_adjoint.hpp"
#include "rs_device.hpp"

namespace rs {

namespace {

constexpr int EX_O = 0;    // orientation pushes: q̄ to i-1 [4], q̄ to i+1 [4], d̄ / x̄ to i+1 [3], rest Darboux to i-1 [4]
constexpr int EX_U = 15;   // u [3], δ [3]
constexpr int EX_P = 21;   // position-sweep pushes: d̄ [3], x̄ [3] to i+1
constexpr int EX_N = 27;

__device__ __forceinline__ f3 ld3s(const float* a, int B, int t) { return mk3(a[t], a[B + t], a[2 * B + t]); }
__device__ __forceinline__ void st3s(float* a, int B, int t, f3 v) {
    a[t] = v.x;
    a[B + t] = v.y;
    a[2 * B + t] = v.z;
}
__device__ __forceinline__ q4 ldq4s(const float* a, int B, int t) { return q4{a[t], a[B + t], a[2 * B + t], a[3 * B + t]}; }
__device__ __forceinline__ void stq4s(float* a, int B, int t, q4 v) {
    a[t] = v.x;
    a[B + t] = v.y;
    a[2 * B + t] = v.z;
    a[3 * B + t] = v.w;
}
__device__ __forceinline__ q4 qsub(q4 a, q4 b) { return q4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }

// J^T c for r(q) = qrot(q, e_x) = (w^2 + x^2 - y^2 - z^2, 2(xy + wz), 2(xz - wy)).
__device__ __forceinline__ q4 rot_ex_adj(q4 q, f3 c) {
    return q4{2.f * ((q.x * c.x + q.y * c.y) + q.z * c.z), 2.f * ((-q.y * c.x + q.x * c.y) - q.w * c.z),
              2.f * ((-q.z * c.x + q.w * c.y) + q.x * c.z), 2.f * ((q.w * c.x + q.z * c.y) - q.y * c.z)};
}

// Adjoint of surface_contact (YB/sim/contact.cuh:27-46) at fixed branch: given f̄ and H̄ = -u δ^T, the
// adjoints of the relative displacement, the normal, the gap and the stiffness.
template <bool F>
__device__ __forceinline__ void surface_contact_adj(const Params& P, f3 n, float gap, f3 d, float k, float muS,
                                                    float muK, f3 fb, f3 u, f3 dl, f3& db, f3& nb, float& gapb,
                                                    float& kb) {
    using A = Ar<F>;
    using M = V<F>;
    const float eps = P.cVSlip, h = P.h;
    const float N = -k * gap;
    const f3 ut = M::axpy(-M::dot(n, d), n, d);
    const float ul = M::len(ut);
    const float vt = A::div(ul, h);
    const float un = dot(u, n), dn = dot(dl, n);
    const float trG = -dot(u, dl), nGn = -un * dn;
    float Nb = dot(fb, n);
    kb = nGn;
    const float ktb = -dot(fb, ut) + (trG - nGn);
    float kt;
    f3 utb;
    if (vt <= eps) {
        kt = A::div(muS * N, eps * h);
        Nb = Nb + ktb * A::div(muS, eps * h);
        utb = (-kt) * fb;
    } else {
        const float r = A::div(vt, eps) - 1.f;
        const float s = 1.f - fminf(1.f, r);
        const float C = A::fma(s, muS, (1.f - s) * muK);
        const float iul = A::rcp(ul);
        kt = A::div(C * N, ul);
        Nb = Nb + ktb * (C * iul);
        const float dC = r < 1.f ? (muS - muK) * (-A::rcp(eps * h)) : 0.f;
        const float ulb = ktb * ((dC * N) * iul - (C * N) * (iul * iul));
        utb = (-kt) * fb + (ulb * iul) * ut;
    }
    const float nutb = dot(n, utb);
    db = utb - nutb * n;
    nb = (N * fb - (k - kt) * (u * dn + dl * un)) - (dot(n, d) * utb + nutb * d);
    kb = kb - gap * Nb;
    gapb = -k * Nb;
}

// Adjoint of the surface terms of one free vertex (surfaces_at, YB/sim/contact.cuh:141-163): adds to x̄, d̄
// and returns the adjoint of the contact stiffness k.
// hits: the candidates pass 1 found in contact at this iteration (bit c); the others add nothing and are skipped.
template <bool F>
__device__ __forceinline__ float surfaces_adj(const Params& P, const Grid* grids, const Pose* envPoses, int ncand,
                                              uint64_t cand, uint32_t hits, f3 p, f3 d, float k, f3 fb, f3 u, f3 dl,
                                              f3& xb, f3& db) {
    using M = V<F>;
    float kcb = 0.f;
    for (int pl = 0; pl < P.numPlanes; pl++) {
        const f3 n = mk3(P.planes[pl][0], P.planes[pl][1], P.planes[pl][2]);
        const float gap = ((M::dot(n, p) - P.planes[pl][3]) + M::dot(n, d)) - P.cRadius;
        if (gap >= 0.f) continue;
        f3 dRb, nb;
        float gapb, kb;
        surface_contact_adj<F>(P, n, gap, d, k, P.cMuS, P.cMuK, fb, u, dl, dRb, nb, gapb, kb);
        db = db + (dRb + gapb * n);
        xb = xb + gapb * n;
        kcb = kcb + kb;
    }
    if (P.numSdf > 0) {
        const int cnt = ncand < 0 ? P.numSdf : ncand;
        for (int c = 0; c < cnt; c++) {
            if (ncand >= 0 && !((hits >> c) & 1u)) continue;
            const int m = ncand < 0 ? c : cand_get(cand, c);
            if (ncand < 0 && !probe_near_box(envPoses + m, p + d)) continue;  // as in surfaces_at
            const float4* s = reinterpret_cast<const float4*>(envPoses + m);
            const float4 a = s[0], b = s[1];
            const q4 q{a.w, b.x, b.y, b.z};
            const f3 pp = mk3(a.x, a.y, a.z);
            float phi;
            f3 gr, mx;
            if (!sdf_sample_t<true, F>(grids + m, M::qrot(qconj(q), (p - pp) + d), phi, gr, mx)) continue;
            const float gap = phi - P.cRadius;
            if (gap >= 0.f) continue;
            const float gl = M::len(gr);
            if (!(gl > 0.f)) continue;
            const f3 gh = M::div(gr, gl);
            const f3 n = M::qrot(q, gh);
            f3 dRel = d, om = mk3(0.f, 0.f, 0.f);
            if (P.sdfMoving) {
                const float4 c2 = s[2], e = s[3];
                om = mk3(c2.z, c2.w, e.x);
                dRel = dRel - P.h * (mk3(b.w, c2.x, c2.y) + M::cross(om, p - pp));
            }
            f3 dRb, nb;
            float gapb, kb;
            surface_contact_adj<F>(P, n, gap, dRel, k, grids[m].muS, grids[m].muK, fb, u, dl, dRb, nb, gapb, kb);
            db = db + dRb;
            if (P.sdfMoving) xb = xb + P.h * cross(om, dRb);
            const f3 ghb = M::qrot(qconj(q), nb);
            const f3 grb = M::div(ghb - dot(gh, ghb) * gh, gl);
            const f3 pob = gapb * gr + mk3(mx.x * grb.y + mx.y * grb.z, mx.x * grb.x + mx.z * grb.z,
                                           mx.y * grb.x + mx.z * grb.y);
            const f3 pwb = M::qrot(q, pob);
            xb = xb + pwb;
            db = db + pwb;
            kcb = kcb + kb;
        }
    }
    return kcb;
}

// History of one frame of the ropes of a launch. Local vertex li = i - A.hv0 of A.hn.
__device__ __forceinline__ size_t hs(const AdjDev& A, int s, int li) { return (size_t)s * A.hn + li; }
__device__ __forceinline__ size_t hk(const AdjDev& A, int I, int s, int k, int li) {
    return ((size_t)s * (I + 1) + k) * A.hn + li;
}
__device__ __forceinline__ size_t hh(const AdjDev& A, int I, int s, int k, int m, int li) {
    return (((size_t)s * I + k) * 9 + m) * A.hn + li;
}

// Vertex-constant data.
struct Vc {
    float w, l, kS;
    uint32_t fl;
    q4 qr, qrp;
};
__device__ __forceinline__ Vc load_vc(const Dev& D, int i, bool act) {
    Vc c{0.f, 1.f, 0.f, 0u, q4{0.f, 0.f, 0.f, 1.f}, q4{0.f, 0.f, 0.f, 1.f}};
    if (act) {
        const float4 prm = D.prm[i];
        c.w = prm.x;
        c.l = prm.y;
        c.kS = prm.z;
        c.fl = flags_of(prm);
        c.qr = ldq(D.qrest[i]);
        c.qrp = (c.fl & HAS_PREV) ? ldq(D.qrest[i - 1]) : c.qr;
    }
    return c;
}

// Pass 1: the frame of one rope recomputed from its checkpoint with the routines and operation order of
// the frame kernel (so the values, and every branch, are those of the recorded forward). Every
// substep's inputs and every iteration's d^k, q^k, total H and Newton direction δ = H^-1 F go to the
// history. The frame-end state is compared bit for bit with the next checkpoint.
template <int B, bool F>
__global__ __launch_bounds__(B, 3) void k_adjoint_forward(Params P, Dev D, AdjDev A, Drive Dr, int S) {
    __shared__ float4 sQ[2][B];
    __shared__ float sX[3 * B], sD[2][3 * B], sF2[9 * B];
    extern __shared__ Pose sPoseAll[];  // the poses of the workgroup's ropes, sized at launch
    const int I = P.numItr;
    const int rb0 = A.blk[blockIdx.x], rb1 = A.blk[blockIdx.x + 1], t = threadIdx.x;
    const int v0 = D.roff[rb0], n = D.roff[rb1] - v0;
    const bool act = t < n;
    const int i = v0 + t, li = i - A.hv0;
    const int r = act ? D.ropeOf[i] : rb0;  // the vertex's rope; consecutive ropes share the workgroup
    const bool root = act && i == D.roff[r];
    const Pose* sPose = sPoseAll + (size_t)(r - rb0) * P.numSdf;
    const int nPose = (rb1 - rb0) * P.numSdf;
    const float h = P.h;
    const Vc c = load_vc(D, i, act);
    if (P.numSdf > 0 && !P.sdfMoving)
        for (int m = t; m < nPose; m += B) sPoseAll[m] = A.fposes[(size_t)rb0 * P.numSdf + m];

    f3 x = mk3(0.f, 0.f, 0.f), v = x, vl = x;
    q4 q{0.f, 0.f, 0.f, 1.f};
    if (act) {
        x = ld3(A.cx[i]);
        v = ld3(A.cv[i]);
        vl = ld3(A.cvl[i]);
        q = ldq(A.cq[i]);
        if (Dr.mode == DRIVE_CONST && root) {
            x = mk3(Dr.p0[3 * r], Dr.p0[3 * r + 1], Dr.p0[3 * r + 2]);
            v = mk3(Dr.v0[3 * r], Dr.v0[3 * r + 1], Dr.v0[3 * r + 2]);
            if (Dr.hasQ0) q = qunit_or_identity(q4{Dr.q0[4 * r], Dr.q0[4 * r + 1], Dr.q0[4 * r + 2], Dr.q0[4 * r + 3]});
        }
    }
    int nc = P.numSdf > 0 ? -1 : 0;
    uint64_t cand = 0;
    for (int s = 0; s < S; s++) {
        if (Dr.mode == DRIVE_PROGRAM && root) {
            f3 xr, vr;
            q4 qr;
            bool setQ;
            root_program(Dr, r, s, S, h, xr, vr, qr, setQ);
            x = xr;
            v = vr;
            if (setQ) q = qr;
        }
        __syncthreads();  // readers of the previous substep are done with sX / sD / sQ / sPose
        if (P.numSdf > 0 && P.sdfMoving)
            for (int m = t; m < nPose; m += B)
                sPoseAll[m] = substep_pose(A.fmotion[(size_t)rb0 * P.numSdf + m], D.grids[m % P.numSdf], s, S, h);
        f3 y = mk3(0.f, 0.f, 0.f);
        if (act) {
            f3 d;
            float padding;
            init_vertex<F>(P, c.w, v, vl, d, y, padding);
            A.hX[hs(A, s, li)] = make_float4(x.x, x.y, x.z, __int_as_float(nc));
            A.hV[hs(A, s, li)] = st3(v);
            A.hVL[hs(A, s, li)] = st3(vl);
            A.hY[hs(A, s, li)] = st3(y);
            A.hC[hs(A, s, li)] = make_uint2((uint32_t)cand, (uint32_t)(cand >> 32));
            A.hD[hk(A, I, s, 0, li)] = st3(d);
            A.hQ[hk(A, I, s, 0, li)] = stq(q);
            vl = v;
            st3s(sX, B, t, x);
            st3s(sD[0], B, t, d);
            sQ[0][t] = stq(q);
        }
        __syncthreads();
        uint64_t hitBits = 0;
        for (int it = 0; it < I; it++) {
            const float* dk = sD[it & 1];
            float* dn = sD[(it + 1) & 1];
            f3 f, f2 = mk3(0.f, 0.f, 0.f);
            s3 H, H2 = szero();
            if (act) {
                const f3 p0 = ld3s(sX, B, t), d0 = ld3s(dk, B, t);
                f3 p1 = p0, p1dx = d0;
                if (c.fl & HAS_NEXT) {
                    p1 = ld3s(sX, B, t + 1);
                    p1dx = ld3s(dk, B, t + 1);
                }
                uint32_t hit = 0;
                vertex_terms<F, true>(P, D.grids, sPose, nc, cand, c.w, c.l, c.kS, c.fl, p0, d0, y,
                                      ldq(sQ[it & 1][t]), p1, p1dx, f, H, f2, H2, &hit);
                if (it < 8) hitBits |= (uint64_t)(hit & 0xFFu) << (8 * it);
                st3s(sF2, B, t, f2);
#pragma unroll
                for (int m = 0; m < 6; m++) sF2[(3 + m) * B + t] = H2.m[m];
            }
            __syncthreads();
            if (act) {
                if (c.fl & HAS_PREV) {
                    f = f + ld3s(sF2, B, t - 1);
                    s3 Hp;
#pragma unroll
                    for (int m = 0; m < 6; m++) Hp.m[m] = sF2[(3 + m) * B + t - 1];
                    sadd(H, Hp);
                }
                f3 d = ld3s(dk, B, t), dl = mk3(0.f, 0.f, 0.f);
                if (c.w != 0.f) {
                    dl = V<F>::ssolve(H, f);
                    d = solve_update<F>(P, d, H, f, false, INFINITY);
                }
                st3s(dn, B, t, d);
                A.hD[hk(A, I, s, it + 1, li)] = st3(d);
#pragma unroll
                for (int m = 0; m < 6; m++) A.hH[hh(A, I, s, it, m, li)] = H.m[m];
                A.hH[hh(A, I, s, it, 6, li)] = dl.x;
                A.hH[hh(A, I, s, it, 7, li)] = dl.y;
                A.hH[hh(A, I, s, it, 8, li)] = dl.z;
            }
            __syncthreads();
            if (act) {
                const q4 q0 = ldq(sQ[it & 1][t]);
                q4 qo = q0;
                if (!(c.fl & FIX_ORIENT) && (c.fl & HAS_NEXT)) {
                    const q4 qp = (c.fl & HAS_PREV) ? ldq(sQ[it & 1][t - 1]) : q0;
                    const q4 qn = (c.fl & HAS_NEXT_ORIENT) ? ldq(sQ[it & 1][t + 1]) : q0;
                    qo = orient_update<F>(c.fl, c.l, c.kS, ld3s(sX, B, t), ld3s(sX, B, t + 1), ld3s(dn, B, t),
                                       ld3s(dn, B, t + 1), qp, q0, qn, c.qrp, c.qr);
                }
                sQ[(it + 1) & 1][t] = stq(qo);
                A.hQ[hk(A, I, s, it + 1, li)] = stq(qo);
            }
            // the ring slots written here were last read two barriers ago
        }
        if (act) {
            const uint64_t mk = (nc < 0 || I > 8) ? ~0ull : hitBits;
            A.hM[hs(A, s, li)] = make_uint2((uint32_t)mk, (uint32_t)(mk >> 32));
        }
        __syncthreads();
        float padv = 0.f;
        if (act) {
            const f3 d = ld3s(sD[I & 1], B, t);
            q = ldq(sQ[I & 1][t]);
            if (c.w != 0.f) {
                f3 tg = mk3(1.f, 0.f, 0.f);
                if (P.dragQuad > 0.f) tg = V<F>::qrot_ex((c.fl & HAS_NEXT) ? q : ldq(sQ[I & 1][t - 1]));
                v = end_velocity<F>(P, d, tg);
            }
            x = x + d;
            padv = 2.f * len(d) + 0.02f;
        }
        if (P.numSdf > CHUNK_BROADPHASE_MIN_MESHES) {  // the frame kernel's chunk broad phase
            const bool fv = act && c.w != 0.f;
            int q0 = fv ? r : 0x7fffffff, q1 = fv ? r : -1;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                q0 = min(q0, __shfl_xor(q0, o));
                q1 = max(q1, __shfl_xor(q1, o));
            }
            const bool oneRope = q0 == q1;
            const Pose* cp = sPoseAll + (size_t)((oneRope ? q0 : rb0) - rb0) * P.numSdf;
            const uint64_t mask = chunk_mesh_mask(P, cp, fv && oneRope, x, padv);
            if (act)
                nc = !fv ? 0
                         : oneRope ? build_candidates_masked(P, sPose, x, padv, mask, cand)
                                   : build_candidates(P, sPose, x, padv, cand);
        } else if (act && P.numSdf > 0) {
            nc = c.w != 0.f ? build_candidates(P, sPose, x, padv, cand) : 0;
        }
    }
    if (act && A.nx) {
        const float4 a = A.nx[i], b = A.nv[i], e = A.nvl[i], g = A.nq[i];
        const bool same = a.x == x.x && a.y == x.y && a.z == x.z && b.x == v.x && b.y == v.y && b.z == v.z &&
                          e.x == vl.x && e.y == vl.y && e.z == vl.z && g.x == q.x && g.y == q.y && g.z == q.z &&
                          g.w == q.w;
        if (!same) atomicAdd(A.mismatch, 1);
    }
}

// Pass 2: backwards through the frame of one rope from the history (no recompute of the iterations).
// Per iteration three phases: orientation adjoint (pushes to the neighbours), Newton adjoint u = a H^-1 d̄,
// segment-term adjoint (pushes to the next vertex).
template <int B, bool F>
__global__ __launch_bounds__(B) void k_adjoint_reverse(Params P, Dev D, AdjDev A, Drive Dr, int S) {
    __shared__ float ex[EX_N * B];
    extern __shared__ Pose sPoseAll[];  // the poses of the workgroup's ropes, sized at launch
    const int I = P.numItr;
    const int rb0 = A.blk[blockIdx.x], rb1 = A.blk[blockIdx.x + 1], t = threadIdx.x;
    const int v0 = D.roff[rb0], n = D.roff[rb1] - v0;
    const bool act = t < n;
    const int i = v0 + t, li = i - A.hv0;
    const int r = act ? D.ropeOf[i] : rb0;  // the vertex's rope; consecutive ropes share the workgroup
    const bool root = act && i == D.roff[r];
    const Pose* sPose = sPoseAll + (size_t)(r - rb0) * P.numSdf;
    const int nPose = (rb1 - rb0) * P.numSdf;
    const float h = P.h;
    const Vc c = load_vc(D, i, act);
    if (P.numSdf > 0 && !P.sdfMoving)
        for (int m = t; m < nPose; m += B) sPoseAll[m] = A.fposes[(size_t)rb0 * P.numSdf + m];

    f3 ax = mk3(0.f, 0.f, 0.f), av = ax, avl = ax;
    q4 aq{0.f, 0.f, 0.f, 0.f};
    if (act) {
        ax = ld3(A.ax[i]);
        av = ld3(A.av[i]);
        avl = ld3(A.avl[i]);
        aq = ldq(A.aq[i]);
        if (A.dlx) ax = ax + mk3(A.dlx[3 * (size_t)i], A.dlx[3 * (size_t)i + 1], A.dlx[3 * (size_t)i + 2]);
    }
    // parameter adjoints summed in double: their terms cancel (inertia against the Hessian's mass term,
    // stretch against its own constraint), and the sums run over every iteration of the frame
    double gk = 0.0, gm = 0.0, gdrag = 0.0;
    double grho[4] = {0.0, 0.0, 0.0, 0.0};
    f3 gp0 = mk3(0.f, 0.f, 0.f), gv0 = gp0, gp1 = gp0;
    using Af = Ar<F>;
    using M = V<F>;
    const float dmp = Af::div(P.damping, h);
    const float kap = Af::div(P.cKappa, h);
    const bool upd = !(c.fl & FIX_ORIENT) && (c.fl & HAS_NEXT);
    const bool hn = (c.fl & HAS_NEXT) != 0, hp = (c.fl & HAS_PREV) != 0;
    const q4 EX{1.f, 0.f, 0.f, 0.f};

    for (int s = S - 1; s >= 0; s--) {
        __syncthreads();  // sPose and the exchange are free
        if (P.numSdf > 0 && P.sdfMoving)
            for (int m = t; m < nPose; m += B)
                sPoseAll[m] = substep_pose(A.fmotion[(size_t)rb0 * P.numSdf + m], D.grids[m % P.numSdf], s, S, h);
        f3 X0 = mk3(0.f, 0.f, 0.f), X1 = X0, y = X0, vs = X0, vls = X0;
        int nc = 0;
        uint64_t cand = 0, hitMask = ~0ull;
        if (act) {
            const float4 xs = A.hX[hs(A, s, li)];
            X0 = ld3(xs);
            nc = __float_as_int(xs.w);
            if (hn) X1 = ld3(A.hX[hs(A, s, li + 1)]);
            y = ld3(A.hY[hs(A, s, li)]);
            vs = ld3(A.hV[hs(A, s, li)]);
            vls = ld3(A.hVL[hs(A, s, li)]);
            const uint2 cc = A.hC[hs(A, s, li)];
            cand = (uint64_t)cc.x | ((uint64_t)cc.y << 32);
            const uint2 mm = A.hM[hs(A, s, li)];
            hitMask = (uint64_t)mm.x | ((uint64_t)mm.y << 32);
        }
        // end update: x' = X + d, v' = (d / h)(1 - drag h) on free vertices
        f3 aD = ax, aX = ax, aY = mk3(0.f, 0.f, 0.f);
        q4 aQ = aq;
        if (act && c.w != 0.f) {
            const f3 d = ld3(A.hD[hk(A, I, s, I, li)]);
            const float invH = Af::rcp(h);
            aD = aD + (invH * (1.f - P.drag * h)) * av;
            gdrag -= (double)h * (double)dot(d * invH, av);
        }
        __syncthreads();  // sPose of a moving frame is in place

        // History of iteration k: d^{k+1} of vertex t and t+1 (the previous iteration's d^k, already loaded),
        // d^k, the frames q^k of t-1..t+1, the Hessian and the Newton direction. Iteration k-1's loads are
        // issued at the top of iteration k and land while it computes (the barriers wait only for LDS).
        struct IterHist {
            f3 dn0, dn1, d0, d1;
            q4 q0, qp, qn;
            s3 H;
            f3 dl;
        };
        auto load_iter = [&](int k, f3 dn0, f3 dn1) {
            IterHist h;
            h.dn0 = dn0;
            h.dn1 = dn1;
            h.d0 = h.d1 = mk3(0.f, 0.f, 0.f);
            h.q0 = q4{0.f, 0.f, 0.f, 1.f};
            h.qp = h.qn = h.q0;
            h.H = szero();
            h.dl = mk3(0.f, 0.f, 0.f);
            if (act) {
                h.d0 = ld3(A.hD[hk(A, I, s, k, li)]);
                h.q0 = ldq(A.hQ[hk(A, I, s, k, li)]);
                if (hn) h.d1 = ld3(A.hD[hk(A, I, s, k, li + 1)]);
                h.qp = hp ? ldq(A.hQ[hk(A, I, s, k, li - 1)]) : h.q0;
                h.qn = (c.fl & HAS_NEXT_ORIENT) ? ldq(A.hQ[hk(A, I, s, k, li + 1)]) : h.q0;
#pragma unroll
                for (int m = 0; m < 6; m++) h.H.m[m] = A.hH[hh(A, I, s, k, m, li)];
                h.dl = mk3(A.hH[hh(A, I, s, k, 6, li)], A.hH[hh(A, I, s, k, 7, li)], A.hH[hh(A, I, s, k, 8, li)]);
            }
            return h;
        };
        IterHist cur;
        {
            f3 dI0 = mk3(0.f, 0.f, 0.f), dI1 = dI0;
            if (act) {
                dI0 = ld3(A.hD[hk(A, I, s, I, li)]);
                if (hn) dI1 = ld3(A.hD[hk(A, I, s, I, li + 1)]);
            }
            cur = load_iter(I - 1, dI0, dI1);
        }
        for (int k = I - 1; k >= 0; k--) {
            IterHist nxt = cur;
            if (k > 0) nxt = load_iter(k - 1, cur.d0, cur.d1);
            const f3 dn0 = cur.dn0, dn1 = cur.dn1, d0 = cur.d0, d1 = cur.d1;
            const q4 q0 = cur.q0, qp = cur.qp, qn = cur.qn;
            const s3 H = cur.H;
            const f3 dl = cur.dl;
            // -- orientation sweep k: q^{k+1}_t from d^{k+1} and q^k_{t-1..t+1} (plus the previous
            //    iteration's segment-term pushes)
            q4 keepQ = aQ;
            if (act) {
                if (k < I - 1 && hp) {
                    aD = aD + ld3s(ex + EX_P * B, B, t - 1);
                    aX = aX + ld3s(ex + (EX_P + 3) * B, B, t - 1);
                }
                q4 pQp{0.f, 0.f, 0.f, 0.f}, pQn = pQp, pRp = pQp;
                f3 pD = mk3(0.f, 0.f, 0.f);
                if (upd) {
                    keepQ = q4{0.f, 0.f, 0.f, 0.f};
                    const f3 er = M::div((X1 - X0) + (dn1 - dn0), c.l);
                    const f3 e = er * (-2.f * c.kS);
                    q4 b{0.f, 0.f, 0.f, 0.f};
                    float s1 = 0.f, s2 = 0.f;
                    if (hp) {
                        const q4 tp = M::qmul(qp, c.qrp);
                        s1 = twist_sign_prev<F>(qp, q0, c.qrp, tp);
                        b = qadd(b, qscale(s1, tp));
                    }
                    if (c.fl & HAS_NEXT_ORIENT) {
                        const q4 tn = M::qmul(qn, qconj(c.qr));
                        s2 = twist_sign_next<F>(q0, qn, c.qr, tn);
                        b = qadd(b, qscale(s2, tn));
                    }
                    const float le = M::len(e), lb = M::qlen(b);
                    const float lam = le + lb;
                    const q4 Z = qadd(M::qmul_ex(M::qmul_pure(e, b)), qscale(lam, b));
                    const float izl = Af::rsq(M::qdot(Z, Z));
                    const q4 qo = qscale(izl, Z);
                    const q4 Zb = qscale(izl, qsub(aQ, qscale(M::qdot(qo, aQ), qo)));
                    const float lamb = M::qdot(Zb, b);
                    q4 bb = qscale(lam, Zb);
                    const q4 Ab = M::qmul(Zb, qconj(EX));
                    bb = qadd(bb, M::qmul(qconj(mkq(e, 0.f)), Ab));
                    f3 eb = qvec(M::qmul(Ab, qconj(b)));
                    if (le > 0.f) eb = eb + Af::div(lamb, le) * e;
                    if (lb > 0.f) bb = qadd(bb, qscale(Af::div(lamb, lb), b));
                    gk -= 2.0 * (double)dot(eb, er);
                    pD = Af::div(-2.f * c.kS, c.l) * eb;
                    aD = aD - pD;
                    aX = aX - pD;
                    if (hp) {
                        pQp = qscale(s1, qmul(bb, qconj(c.qrp)));
                        pRp = qscale(s1, qmul(qconj(qp), bb));
                    }
                    if (c.fl & HAS_NEXT_ORIENT) {
                        pQn = qscale(s2, qmul(bb, c.qr));
                        const q4 gr = qconj(qscale(s2, qmul(qconj(qn), bb)));
                        grho[0] += gr.x;
                        grho[1] += gr.y;
                        grho[2] += gr.z;
                        grho[3] += gr.w;
                    }
                }
                stq4s(ex + EX_O * B, B, t, pQp);
                stq4s(ex + (EX_O + 4) * B, B, t, pQn);
                st3s(ex + (EX_O + 8) * B, B, t, pD);
                stq4s(ex + (EX_O + 11) * B, B, t, pRp);
            }
            __syncthreads();
            // -- gather the orientation pushes; Newton adjoint u = a H^-1 d̄^{k+1}
            f3 u = mk3(0.f, 0.f, 0.f);
            if (act) {
                aQ = keepQ;
                if (hp) {
                    aQ = qadd(aQ, ldq4s(ex + (EX_O + 4) * B, B, t - 1));
                    const f3 pd = ld3s(ex + (EX_O + 8) * B, B, t - 1);
                    aD = aD + pd;
                    aX = aX + pd;
                }
                if (hn) {
                    aQ = qadd(aQ, ldq4s(ex + EX_O * B, B, t + 1));
                    const q4 gr = ldq4s(ex + (EX_O + 11) * B, B, t + 1);
                    grho[0] += gr.x;
                    grho[1] += gr.y;
                    grho[2] += gr.z;
                    grho[3] += gr.w;
                }
                if (c.w != 0.f) u = P.accel * M::ssolve(H, aD);
                st3s(ex + EX_U * B, B, t, u);
                st3s(ex + (EX_U + 3) * B, B, t, dl);
            }
            __syncthreads();
            // -- adjoint of the segment terms of vertex t (inertia, surfaces, stretch of segment (t, t+1))
            if (act) {
                f3 pD = mk3(0.f, 0.f, 0.f), pX = pD;
                const float trG = -dot(u, dl);
                if (c.w != 0.f) {
                    const float coef = Af::rcp(h * h * c.w);
                    aY = aY + coef * u;
                    aD = aD - coef * u;
                    gm += ((double)dot(y - d0, u) + (double)trG) / ((double)h * (double)h);
                    if (P.numPlanes > 0 || P.numSdf > 0) {
                        const float kc = Af::div(kap * kap, c.w);
                        const uint32_t hits = k < 8 ? (uint32_t)(hitMask >> (8 * k)) & 0xFFu : 0xFFu;
                        const float kcb =
                            surfaces_adj<F>(P, D.grids, sPose, nc, cand, hits, X0, d0, kc, u, u, dl, aX, aD);
                        gm += (double)kcb * ((double)kap * (double)kap);
                    }
                }
                if (hn) {
                    const f3 u2 = ld3s(ex + EX_U * B, B, t + 1), dl2 = ld3s(ex + (EX_U + 3) * B, B, t + 1);
                    const float invl = Af::rcp(c.l);
                    const f3 cc = ((X1 - X0) + (d1 - d0)) * invl - M::qrot_ex(q0);
                    const float kk = c.kS * invl;
                    const float dd = kk * invl;
                    const f3 fdiff = u - u2;
                    const f3 cb = kk * fdiff;
                    const float kkb = dot(cc, fdiff);
                    const float ddb = -dmp * (dot(d0, u) + dot(d1, u2)) + (1.f + dmp) * (trG - dot(u2, dl2));
                    aD = aD - (dmp * dd) * u;
                    pD = pD - (dmp * dd) * u2;
                    const f3 Db = cb * invl;
                    pD = pD + Db;
                    pX = pX + Db;
                    aD = aD - Db;
                    aX = aX - Db;
                    aQ = qsub(aQ, rot_ex_adj(q0, cb));
                    gk += ((double)kkb + (double)ddb * (double)invl) * (double)invl;
                }
                st3s(ex + EX_P * B, B, t, pD);
                st3s(ex + (EX_P + 3) * B, B, t, pX);
            }
            __syncthreads();
            cur = nxt;
        }
        if (act && hp) {  // segment-term pushes of iteration 0
            aD = aD + ld3s(ex + EX_P * B, B, t - 1);
            aX = aX + ld3s(ex + (EX_P + 3) * B, B, t - 1);
        }

        // init: d^0 = h v (+ h^2 s g), y = h v + h^2 g, last velocity <- v
        if (act) {
            f3 vb = h * aD + avl, vlb = mk3(0.f, 0.f, 0.f);
            if (c.w != 0.f) {
                vb = vb + h * aY;
                const f3 g = mk3(P.gx, P.gy, P.gz);
                const float g2 = M::len2(g);
                if (g2 > 0.f) {
                    const float sr = gravity_fraction<F>(P, vs, vls, g, g2);  // the forward's clamp decision
                    if (sr > 0.f && sr < 1.f) {
                        const f3 ab = ((h * h) * dot(g, aD) / g2) * g;
                        vb = vb + ab / P.lastH;
                        vlb = vlb - ab / P.lastH;
                    }
                }
            } else {
                vb = vb + av;  // the velocity of a pinned vertex passes through the substep
            }
            ax = aX;
            av = vb;
            avl = vlb;
            aq = aQ;
            if (Dr.mode == DRIVE_PROGRAM && root) {  // applyRootProgramK (csrc/batch.cu:122-159)
                const float tau = float(s) * h, dt = Dr.dt;
                const float c2 = 2.f / (dt * dt);
                const float cx1 = 0.5f * tau * tau * c2, cv1 = (tau + 0.5f * h) * c2;
                gp1 = gp1 + (cx1 * ax + cv1 * av);
                gp0 = gp0 + ((1.f - cx1) * ax - cv1 * av);
                gv0 = gv0 + ((tau - cx1 * dt) * ax + (1.f - cv1 * dt) * av);
                ax = mk3(0.f, 0.f, 0.f);
                av = ax;
                if (Dr.hasQ0) aq = q4{0.f, 0.f, 0.f, 0.f};
            }
        }
    }
    if (Dr.mode == DRIVE_CONST && root) {  // applyRootsK (csrc/batch.cu:90-104)
        gp0 = gp0 + ax;
        gv0 = gv0 + av;
        ax = mk3(0.f, 0.f, 0.f);
        av = ax;
        if (Dr.hasQ0) aq = q4{0.f, 0.f, 0.f, 0.f};
    }
    if (act) {
        A.ax[i] = st3(ax);
        A.av[i] = st3(av);
        A.avl[i] = st3(avl);
        A.aq[i] = stq(aq);
        A.gk[i] += gk;
        A.grho[4 * (size_t)i] += grho[0];
        A.grho[4 * (size_t)i + 1] += grho[1];
        A.grho[4 * (size_t)i + 2] += grho[2];
        A.grho[4 * (size_t)i + 3] += grho[3];
        A.gm[i] += gm;
        if (root) {
            if (A.gp0) {
                A.gp0[3 * r] = gp0.x;
                A.gp0[3 * r + 1] = gp0.y;
                A.gp0[3 * r + 2] = gp0.z;
            }
            if (A.gv0) {
                A.gv0[3 * r] = gv0.x;
                A.gv0[3 * r + 1] = gv0.y;
                A.gv0[3 * r + 2] = gv0.z;
            }
            if (A.gp1) {
                A.gp1[3 * r] = gp1.x;
                A.gp1[3 * r + 1] = gp1.y;
                A.gp1[3 * r + 2] = gp1.z;
            }
        }
    }
    // drag adjoint of each rope, summed over its vertices in order by its root
    __syncthreads();
    double* red = reinterpret_cast<double*>(ex);
    red[t] = act ? gdrag : 0.0;
    __syncthreads();
    if (root) {
        double sum = 0.0;
        for (int j = t; j < D.roff[r + 1] - v0; j++) sum += red[j];
        A.gdrag[r] += sum;
    }
}

// Parameter gradients from the accumulated adjoints: dL/dEA = k̄ l, dL/dEI = (4 / l) ρ̄ . unit, dL/dm = m̄.
__global__ void k_adjoint_finish(Params P, Dev D, const float4* rhoUnit, const double* gk, const double* grho,
                                 const double* gm, float* gea, float* gei, float* gmass) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nv) return;
    const float4 prm = D.prm[i];
    const bool seg = (flags_of(prm) & HAS_NEXT) != 0;
    if (gea) gea[i] = seg ? (float)(gk[i] * (double)prm.y) : 0.f;
    if (gei) {
        const double* g = grho + 4 * (size_t)i;
        const float4 u = rhoUnit[i];
        const double dotgu = ((g[0] * u.x + g[1] * u.y) + g[2] * u.z) + g[3] * u.w;
        gei[i] = seg ? (float)((4.0 / (double)prm.y) * dotgu) : 0.f;
    }
    if (gmass) gmass[i] = prm.x != 0.f ? (float)gm[i] : 0.f;
}

}  // namespace

size_t adjoint_history_bytes(int numItr) {
    return 5 * sizeof(float4) + 2 * sizeof(uint2) + (size_t)(numItr + 1) * 2 * sizeof(float4) +
           (size_t)numItr * 9 * sizeof(float);
}

bool adjoint_block_supported(int block, bool reverse) {
    return block == 64 || block == 128 || block == 192 || block == 256 || (reverse && block == 512);
}

namespace {
// Waves per SIMD that a kernel's VGPR allocation allows (512 VGPRs per SIMD lane, granule 8, at most 8).
int waves_of(const void* fn) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, fn) != hipSuccess || a.numRegs <= 0) return 1;
    const int regs = (a.numRegs + 7) / 8 * 8;
    return std::max(1, std::min(8, 512 / regs));
}
template <int B>
const void* pass_fn(bool reverse, bool fast) {
    if (reverse)
        return fast ? reinterpret_cast<const void*>(&k_adjoint_reverse<B, true>)
                    : reinterpret_cast<const void*>(&k_adjoint_reverse<B, false>);
    return fast ? reinterpret_cast<const void*>(&k_adjoint_forward<B, true>)
                : reinterpret_cast<const void*>(&k_adjoint_forward<B, false>);
}
}  // namespace

int adjoint_waves_per_simd(int block, bool reverse, bool fast) {
    switch (block) {
        case 64: return waves_of(pass_fn<64>(reverse, fast));
        case 128: return waves_of(pass_fn<128>(reverse, fast));
        case 192: return waves_of(pass_fn<192>(reverse, fast));
        case 256: return waves_of(pass_fn<256>(reverse, fast));
        case 512:
            if (!reverse) return 1;
            return waves_of(fast ? reinterpret_cast<const void*>(&k_adjoint_reverse<512, true>)
                                 : reinterpret_cast<const void*>(&k_adjoint_reverse<512, false>));
        default: return 1;
    }
}

template <int B, bool F>
static hipError_t launch_pass_f(bool reverse, const Params& P, const Dev& D, const AdjDev& A, const Drive& Dr, int S,
                                int nblocks, size_t poseBytes, hipStream_t s) {
    if (reverse)
        hipLaunchKernelGGL((k_adjoint_reverse<B, F>), dim3(nblocks), dim3(B), poseBytes, s, P, D, A, Dr, S);
    else
        hipLaunchKernelGGL((k_adjoint_forward<B, F>), dim3(nblocks), dim3(B), poseBytes, s, P, D, A, Dr, S);
    return hipGetLastError();
}
template <int B>
static hipError_t launch_pass_t(bool reverse, const Params& P, const Dev& D, const AdjDev& A, const Drive& Dr, int S,
                                int nblocks, size_t poseBytes, hipStream_t s) {
    if (P.fastMath) return launch_pass_f<B, true>(reverse, P, D, A, Dr, S, nblocks, poseBytes, s);
    return launch_pass_f<B, false>(reverse, P, D, A, Dr, S, nblocks, poseBytes, s);
}

hipError_t launch_adjoint_pass(bool reverse, int block, const Params& P, const Dev& D, const AdjDev& A, const Drive& Dr,
                               int S, int nblocks, int ropesPerBlock, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (!adjoint_block_supported(block, reverse) || ropesPerBlock < 1 ||
        (size_t)ropesPerBlock * P.numSdf > (size_t)adjoint_pose_slots(block))
        return hipErrorInvalidValue;
    const size_t poseBytes = (size_t)ropesPerBlock * P.numSdf * sizeof(Pose);
    switch (block) {
        case 64: return launch_pass_t<64>(reverse, P, D, A, Dr, S, nblocks, poseBytes, s);
        case 128: return launch_pass_t<128>(reverse, P, D, A, Dr, S, nblocks, poseBytes, s);
        case 192: return launch_pass_t<192>(reverse, P, D, A, Dr, S, nblocks, poseBytes, s);
        case 256: return launch_pass_t<256>(reverse, P, D, A, Dr, S, nblocks, poseBytes, s);
        case 512:
            if (P.fastMath)
                hipLaunchKernelGGL((k_adjoint_reverse<512, true>), dim3(nblocks), dim3(512), poseBytes, s, P, D, A, Dr, S);
            else
                hipLaunchKernelGGL((k_adjoint_reverse<512, false>), dim3(nblocks), dim3(512), poseBytes, s, P, D, A, Dr,
                                   S);
            return hipGetLastError();
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint_finish(const Params& P, const Dev& D, const float4* rhoUnit, const double* gk,
                                 const double* grho, const double* gm, float* gea, float* gei, float* gmass,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_adjoint_finish, dim3((P.nv + 255) / 256), dim3(256), 0, s, P, D, rhoUnit, gk, grho, gm, gea,
                       gei, gmass);
    return hipGetLastError();
}

}  // namespace rs
