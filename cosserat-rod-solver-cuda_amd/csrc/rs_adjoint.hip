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
#include <cstdlib>

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
            const Grid g = load_grid(grids + m);
            if (!sdf_sample_t<true, F>(g, M::qrot(qconj(q), (p - pp) + d), phi, gr, mx)) continue;
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
            surface_contact_adj<F>(P, n, gap, dRel, k, g.muS, g.muK, fb, u, dl, dRb, nb, gapb, kb);
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

// ------------------------------------------------------------------ rope-rope contact adjoints
//
// Reverse of the closest-point parameters (u, v) = seg_closest(0, A, B0, B1) by its three steps (infinite-line
// v, clamped; u from it, clamped; v again from u, clamped): the adjoints of A, B0, B1 given those of u and v. A
// clamped step passes no gradient unless its value lies in [0, 1]; the singular branch (v = 0.5) passes none,
// and neither does the infinite-line step of nearly parallel segments (a c - b^2 <= 1e-4 a c, within 0.6
// degrees), whose v float32 cannot resolve and whose derivative grows as 1 / sin^2 of their angle.
__device__ void seg_closest_adj(f3 A, f3 B0, f3 B1, float ub, float vb, f3& Ab, f3& B0b, f3& B1b) {
    const f3 A1 = B1 - B0;
    const float a = dot(A, A), b = dot(A, A1), c = dot(A1, A1);
    const float pa = dot(A, B0), pb = dot(A1, B0);
    const float den = a * c - b * b;
    float y0 = (b * pa - a * pb) / den;
    const bool fin = isfinite(y0);
    if (!fin) y0 = 0.5f;
    const float y0c = clampf(y0, 0.f, 1.f);
    const f3 m1 = B0 + y0c * A1;
    const float x0 = dot(m1, A) / a;
    const float xc = clampf(x0, 0.f, 1.f);
    const f3 m2 = xc * A - B0;
    const float y1 = dot(m2, A1) / c;
    f3 A0b = mk3(0.f, 0.f, 0.f), A1b = A0b, Db = A0b;
    float ab = 0.f, bb = 0.f, cb = 0.f;
    const float y1b = (y1 >= 0.f && y1 <= 1.f) ? vb : 0.f;
    const float n2b = y1b / c;
    cb -= y1b * y1 / c;
    const f3 m2b = n2b * A1;
    A1b = A1b + n2b * m2;
    const float xcb = ub + dot(m2b, A);
    A0b = A0b + xc * m2b;
    Db = Db - m2b;
    const float x0b = (x0 >= 0.f && x0 <= 1.f) ? xcb : 0.f;
    const float n1b = x0b / a;
    ab -= x0b * x0 / a;
    const f3 m1b = n1b * A;
    A0b = A0b + n1b * m1;
    Db = Db + m1b;
    A1b = A1b + y0c * m1b;
    const bool ill = !(den > 1e-4f * (a * c));
    const float y0b = (fin && !ill && y0 >= 0.f && y0 <= 1.f) ? dot(m1b, A1) : 0.f;
    if (y0b != 0.f) {
        const float numb = y0b / den, denb = -y0b * y0 / den;
        ab += denb * c;
        cb += denb * a;
        bb -= 2.f * b * denb;
        bb += numb * pa;
        ab -= numb * pb;
        const float pab = numb * b, pbb = -numb * a;
        A0b = A0b + pab * B0;
        Db = Db + pab * A;
        A1b = A1b + pbb * B0;
        Db = Db + pbb * A1;
    }
    A0b = A0b + (2.f * ab) * A + bb * A1;
    A1b = A1b + bb * A + (2.f * cb) * A1;
    Ab = A0b;
    B1b = A1b;
    B0b = Db - A1b;
}

// Adjoint of one rope-rope barrier + friction term (collision_eval, YB/sim/cosserat.cu:99-158) at the forward's
// branches. Seeds: vertex i's Newton adjoint u1 and direction l1 (its F adjoint is u1, its H adjoint -u1 l1^T),
// vertex i+1's u2, l2. Returns the adjoints of the term's inputs in g[0..7]: p0, p1, d0, p1dx (segment i, i+1)
// and b0, b1, db0, db1 (segment j, j+1). The forward part repeats collision_eval's arithmetic, so every branch
// (contact, the 1e-3 floor of the gap, friction, the friction cap) is the one the forward took.
// The rope-rope material a pair term reads, passed by value: a reference to the kernel's Params would make the
// compiler keep a copy of the whole struct in per-lane scratch memory for the call.
struct PairMat {
    float barrier, radius, kCollision, fricMu, kFriction;
};
__device__ __forceinline__ PairMat pair_mat(const Params& P) {
    return PairMat{P.barrier, P.radius, P.kCollision, P.fricMu, P.kFriction};
}

template <bool F>
__device__ __forceinline__ void pair_term_adj(PairMat P, float dmp, f3 p0, f3 p1, f3 d0, f3 p1dx, f3 b0, f3 b1, f3 db0,
                                           f3 db1, f3 u1, f3 l1, f3 u2, f3 l2, f3* g) {
    using Af = Ar<F>;
    using M = V<F>;
    for (int k = 0; k < 8; k++) g[k] = mk3(0.f, 0.f, 0.f);
    const float invb = Af::rcp(P.barrier);
    const float rad2 = 2.f * P.radius;
    const float kCol = P.kCollision * invb;
    const f3 A = (p1 - p0) + (p1dx - d0), B0 = (b0 - p0) + (db0 - d0), B1 = (b1 - p0) + (db1 - d0);
    float u, v;
    M::seg_closest(mk3(0.f, 0.f, 0.f), A, B0, B1, u, v);
    const f3 dpos = M::mix(p0, p1, u) - M::mix(b0, b1, v);
    const f3 ddpos = M::mix(d0, p1dx, u) - M::mix(db0, db1, v);
    const f3 N = dpos + ddpos;
    const float dd = M::len(N);
    if (!(dd > 0.f)) return;
    const f3 n = N * Af::rcp(dd);
    const float wy = u, wx = 1.f - u;
    const float gr = (dd - rad2) * invb;
    if (gr > 1.f) return;
    const float gc = fmaxf(gr, 1e-3f);
    const float invd = Af::rcp(gc), logd = Af::log(gc);
    const float dH0 = Af::fma(-2.f, logd, Af::fma(2.f + invd, invd, -3.f)) * kCol * invb;
    const float Bf = Af::fma(2.f * gc, logd, gc - 1.f);
    const float ff = -(1.f - gc) * Bf * invd * kCol;
    const float nd0 = dot(n, d0), nd1 = dot(n, p1dx);
    const float s1 = ff * wx - dmp * dH0 * wx * wx * nd0, s2 = ff * wy - dmp * dH0 * wy * wy * nd1;
    const float dH = dH0 * (1.f + dmp);
    const float h1 = dH * wx * wx, h2 = dH * wy * wy;
    const f3 U = M::axpy(-M::dot(n, ddpos), n, ddpos);
    const float ul = M::len(U);
    const bool fric = ul > 0.f;
    float f1 = 0.f, q = 0.f;
    if (fric) {
        q = Af::div(P.fricMu * ff, ul);
        f1 = fminf(P.kFriction, q);
    }
    const float a1 = -(f1 * wx), a2 = -(f1 * wy), e1 = -(wx * wx * f1), e2 = -(wy * wy * f1);
    // reverse
    const float un1 = dot(u1, n), dn1 = dot(l1, n), un2 = dot(u2, n), dn2 = dot(l2, n);
    const float s1b = un1, s2b = un2, h1b = -un1 * dn1, h2b = -un2 * dn2;
    f3 nb = s1 * u1 + s2 * u2 - (h1 + e1) * (dn1 * u1 + un1 * l1) - (h2 + e2) * (dn2 * u2 + un2 * l2);
    f3 Ub = a1 * u1 + a2 * u2;
    float wxb = 0.f, wyb = 0.f, ffb = 0.f, dHb = 0.f;
    if (fric) {
        const float a1b = dot(u1, U), a2b = dot(u2, U);
        const float e1b = -un1 * dn1 + dot(u1, l1), e2b = -un2 * dn2 + dot(u2, l2);
        const float f1b = -(a1b * wx + e1b * wx * wx + a2b * wy + e2b * wy * wy);
        wxb += -f1 * a1b - 2.f * wx * f1 * e1b;
        wyb += -f1 * a2b - 2.f * wy * f1 * e2b;
        if (q < P.kFriction) {
            ffb += f1b * P.fricMu / ul;
            const float ulb = -f1b * q / ul;
            Ub = Ub + (ulb / ul) * U;
        }
    }
    const float nddp = dot(n, ddpos), nUb = dot(n, Ub);
    f3 ddposb = Ub - nUb * n;
    nb = nb - (nddp * Ub + nUb * ddpos);
    dHb += (h1b * wx * wx + h2b * wy * wy) * (1.f + dmp);
    wxb += h1b * dH * 2.f * wx;
    wyb += h2b * dH * 2.f * wy;
    ffb += s1b * wx + s2b * wy;
    wxb += s1b * (ff - 2.f * dmp * dH0 * wx * nd0);
    wyb += s2b * (ff - 2.f * dmp * dH0 * wy * nd1);
    dHb -= dmp * (s1b * wx * wx * nd0 + s2b * wy * wy * nd1);
    nb = nb - (s1b * dmp * dH0 * wx * wx) * d0 - (s2b * dmp * dH0 * wy * wy) * p1dx;
    f3 d0b = -(s1b * dmp * dH0 * wx * wx) * n, p1dxb = -(s2b * dmp * dH0 * wy * wy) * n;
    float grb = 0.f;
    if (gr >= 1e-3f) {
        const float A_ = 1.f - gc;
        const float dff = -kCol * ((A_ * (2.f * logd + 3.f) - Bf) * invd - A_ * Bf * invd * invd);
        const float ddH = kCol * invb * (-2.f * invd) * (1.f + invd + invd * invd);
        grb = ffb * dff + dHb * ddH;
    }
    const f3 Nb = (nb - dot(n, nb) * n) / dd + (grb * invb) * n;
    ddposb = ddposb + Nb;
    const f3 dposb = Nb;
    float ub = wyb - wxb + dot(dposb, p1 - p0) + dot(ddposb, p1dx - d0);
    const float vb = -dot(dposb, b1 - b0) - dot(ddposb, db1 - db0);
    f3 p0b = (1.f - u) * dposb, p1b = u * dposb, b0b = -(1.f - v) * dposb, b1b = -v * dposb;
    d0b = d0b + (1.f - u) * ddposb;
    p1dxb = p1dxb + u * ddposb;
    f3 db0b = -(1.f - v) * ddposb, db1b = -v * ddposb;
    f3 Ab, B0b, B1b;
    seg_closest_adj(A, B0, B1, ub, vb, Ab, B0b, B1b);
    const f3 base = Ab + B0b + B1b;
    g[0] = p0b - base;
    g[1] = p1b + Ab;
    g[2] = d0b - base;
    g[3] = p1dxb + Ab;
    g[4] = b0b + B0b;
    g[5] = b1b + B1b;
    g[6] = db0b + B0b;
    g[7] = db1b + B1b;
}

// Adjoint of a pair's step-limit distance 0.45 |u a1 - mix(b0, b1, v)| with a1, b0, b1 relative to xa0 and
// (u, v) = seg_closest(0, a1, b0, b1) (YB/sim/collision.cu:103-136, the detection's arithmetic), scaled by gb:
// the adjoints of the four endpoint positions. The distance between two segments is a minimum over (u, v),
// so its derivative has no term through them (at an interior minimum the distance is stationary in u and v,
// at a clamped one they do not move); (u, v) are held fixed.
__device__ __noinline__ void limit_dist_adj(f3 xa0, f3 xa1, f3 xb0, f3 xb1, float gb, f3* g) {
    const f3 a1 = xa1 - xa0, b0 = xb0 - xa0, b1 = xb1 - xa0;
    float u, v;
    seg_closest(mk3(0.f, 0.f, 0.f), a1, b0, b1, u, v);
    const f3 d = u * a1 - mix(b0, b1, v);
    const float l = len(d);
    for (int k = 0; k < 4; k++) g[k] = mk3(0.f, 0.f, 0.f);
    if (!(l > 0.f)) return;
    const f3 Db = (0.45f * gb / l) * d;
    const f3 a1b = u * Db, b0b = -(1.f - v) * Db, b1b = -v * Db;
    g[0] = -(a1b + b0b + b1b);
    g[1] = a1b;
    g[2] = b0b;
    g[3] = b1b;
}

// Rope-rope detection of a workgroup's contact groups at one substep, from the positions and paddings in LDS:
// the reference's per-segment test in ascending order (k_detect, csrc/batch.cu:19-49) with its step limit
// (YB/sim/collision.cu:103-136), reproduced operation for operation so the lists and limits are the forward's.
// The candidates come from skin lists: at a build each segment records, ascending, the partners whose box
// (padded by the build radius R = max(pad) + slack, the skin and a rounding margin) meets its own. Until some
// vertex has moved more than the skin from its build position or some padding has outgrown R, every pair the
// exact test can accept is on those lists (two segments approach by at most twice the skin). A list longer than
// SKIN_CAP falls back to a scan of the whole contact group. Lists hold workgroup-local slots.
constexpr int SKIN_CAP = 64;
// per local vertex: partner list, reverse list, skin list, and (reverse pass) the pair index of each reverse-list
// entry
constexpr int REV_IDX = 2 * MAX_COLL + SKIN_CAP;
constexpr int LIST_STRIDE = 3 * MAX_COLL + SKIN_CAP;

struct RcLds {
    float *x, *pad, *lim, *x0, *R, *sb, *cb;  // positions [3][B], paddings [B], limits [B], build positions [3][B],
                                              // build radii [B], build boxes [6][B], chunk boxes [6][B/16]
    float *dk, *mb, *mown, *mprev;            // (reverse pass) iterate d^k [3][B], step-limit adjoints [B] x 3
    int *fl, *rope, *arg, *nrev, *ncol, *vc;  // flags, rope, limit argument, list lengths, skin-list length (-1: scan)
    // (reverse pass with pairs) first pair index of each slot [B + 1] (then 8 wave sums), the pairs (owner slot |
    // partner slot << 16) [PAIR_TAB_OF(B)], the reverse lists of up to REV_SLOTS_OF(B) entries [B][REV_SLOTS_OF(B)],
    // and the pair mailbox [PAIR_ADJ_N][slots]
    int* poff;
    uint32_t* pairs;
    uint16_t* rq;
    float* mbx;
};
// Floats a pair's adjoint leaves in the mailbox: the adjoints of its 8 inputs (pair_term_adj).
constexpr int PAIR_ADJ_N = 24;
constexpr int POFF_N(int B) { return B + 16; }
// Pairs of a workgroup's substep indexed in LDS (more are found from the offsets and the lists in memory), and
// reverse-list entries per segment kept in LDS (a longer reverse list lives in memory). The 512-thread form keeps
// both in memory: its LDS is taken by the arrays it needs at every size.
constexpr int PAIR_TAB_OF(int B) { return B <= 384 ? 1024 : 0; }
constexpr int REV_SLOTS_OF(int B) { return B <= 384 ? 8 : 0; }
// LDS of the pair index arrays before the mailbox.
constexpr size_t pair_index_lds_bytes(int B) {
    return sizeof(int) * POFF_N(B) + sizeof(uint32_t) * PAIR_TAB_OF(B) + sizeof(uint16_t) * B * REV_SLOTS_OF(B);
}

// The detection arrays in dynamic LDS at `base` (after the mesh pose slots); x: positions already in LDS, or
// null for a region of their own.
template <int B>
__device__ __forceinline__ RcLds rc_lds(char* base, float* x) {
    RcLds L;
    float* f = reinterpret_cast<float*>(base);
    auto take = [&](int n) {
        float* p = f;
        f += n;
        return p;
    };
    L.pad = take(B);
    L.lim = take(B);
    L.x0 = take(3 * B);
    L.R = take(B);
    L.sb = take(6 * B);
    L.cb = take(6 * (B / 16));
    L.dk = take(3 * B);
    L.mb = take(B);
    L.mown = take(B);
    L.mprev = take(B);
    L.x = x ? x : take(3 * B);
    int* q = reinterpret_cast<int*>(f);
    L.fl = q;
    L.rope = q + B;
    L.arg = q + 2 * B;
    L.nrev = q + 3 * B;
    L.ncol = q + 4 * B;
    L.vc = q + 5 * B;
    L.poff = q + 6 * B;
    L.pairs = reinterpret_cast<uint32_t*>(L.poff + POFF_N(B));
    L.rq = reinterpret_cast<uint16_t*>(L.pairs + PAIR_TAB_OF(B));
    L.mbx = reinterpret_cast<float*>(L.rq + B * REV_SLOTS_OF(B));
    return L;
}

// The clamp-only kernels' arrays (no detection, no pairs): paddings and the step-limit adjoints, [B] each.
template <int B>
__device__ __forceinline__ RcLds rc_lds_clamp(char* base) {
    RcLds L{};
    float* f = reinterpret_cast<float*>(base);
    L.pad = f;
    L.mb = f + B;
    L.mown = f + 2 * B;
    L.mprev = f + 3 * B;
    return L;
}
constexpr size_t rc_lds_clamp_bytes(int B) { return sizeof(float) * 4 * (size_t)B; }

__device__ __forceinline__ bool box_meets(const float* bx, int stride, int k, f3 lo, f3 hi) {
    return !(bx[k] > hi.x || bx[stride + k] > hi.y || bx[2 * stride + k] > hi.z || bx[3 * stride + k] < lo.x ||
             bx[4 * stride + k] < lo.y || bx[5 * stride + k] < lo.z);
}

// The rebuild vote of slot t (every substep; the caller ORs it over the workgroup): its vertex moved past the
// skin since the build, or its padding outgrew the build radius of a segment it belongs to.
__device__ __forceinline__ bool rc_skin_stale(const Params& P, const RcLds& L, int B, int t, bool act, uint32_t fl) {
    if (!act) return false;
    const float skin = 2.f * P.scaledDetR;
    const f3 dx = ld3s(L.x, B, t) - ld3s(L.x0, B, t);
    const float pad = L.pad[t];
    return len2(dx) > skin * skin || ((fl & HAS_NEXT) && pad > L.R[t]) || ((fl & HAS_PREV) && pad > L.R[t - 1]);
}

// Skin-list build, stage 1 (all lanes; a barrier follows): build positions and radii, boxes, chunk boxes.
template <int B>
__device__ __forceinline__ void rc_skin_boxes(const Params& P, const RcLds& L, int t, bool act, uint32_t fl) {
    const float skin = 2.f * P.scaledDetR, slackR = 0.25f * P.scaledDetR;
    f3 lo = mk3(INFINITY, INFINITY, INFINITY), hi = mk3(-INFINITY, -INFINITY, -INFINITY);
    if (act) {
        const f3 a = ld3s(L.x, B, t);
        st3s(L.x0, B, t, a);
        if (fl & HAS_NEXT) {
            const f3 b = ld3s(L.x, B, t + 1);
            const float R = fmaxf(L.pad[t], L.pad[t + 1]) + slackR;
            L.R[t] = R;
            const float rb = R + skin + 1e-4f + 2e-6f * fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z)));
            lo = mk3(fminf(a.x, b.x) - rb, fminf(a.y, b.y) - rb, fminf(a.z, b.z) - rb);
            hi = mk3(fmaxf(a.x, b.x) + rb, fmaxf(a.y, b.y) + rb, fmaxf(a.z, b.z) + rb);
        }
    }
    L.sb[t] = lo.x;
    L.sb[B + t] = lo.y;
    L.sb[2 * B + t] = lo.z;
    L.sb[3 * B + t] = hi.x;
    L.sb[4 * B + t] = hi.y;
    L.sb[5 * B + t] = hi.z;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        lo.x = fminf(lo.x, __shfl_xor(lo.x, o));
        lo.y = fminf(lo.y, __shfl_xor(lo.y, o));
        lo.z = fminf(lo.z, __shfl_xor(lo.z, o));
        hi.x = fmaxf(hi.x, __shfl_xor(hi.x, o));
        hi.y = fmaxf(hi.y, __shfl_xor(hi.y, o));
        hi.z = fmaxf(hi.z, __shfl_xor(hi.z, o));
    }
    if ((t & 15) == 0) {
        constexpr int NCH = B / 16;
        const int c = t >> 4;
        L.cb[c] = lo.x;
        L.cb[NCH + c] = lo.y;
        L.cb[2 * NCH + c] = lo.z;
        L.cb[3 * NCH + c] = hi.x;
        L.cb[4 * NCH + c] = hi.y;
        L.cb[5 * NCH + c] = hi.z;
    }
}

// Skin-list build, stage 2: the partners of segment t over its group's local slots [g0, g1), ascending.
template <int B>
__device__ __forceinline__ void rc_skin_list(const RcLds& L, int t, uint32_t fl, int g0, int g1, uint16_t* skinRow) {
    constexpr int NCH = B / 16;
    int cnt = 0;
    if (fl & HAS_NEXT) {
        const f3 lo = mk3(L.sb[t], L.sb[B + t], L.sb[2 * B + t]);
        const f3 hi = mk3(L.sb[3 * B + t], L.sb[4 * B + t], L.sb[5 * B + t]);
        const int myRope = L.rope[t];
        for (int ch = g0 >> 4; ch <= (g1 - 1) >> 4; ch++) {
            if (!box_meets(L.cb, NCH, ch, lo, hi)) continue;
            const int j1 = min(ch * 16 + 16, g1);
            for (int j = max(ch * 16, g0); j < j1; j++) {
                if (!(L.fl[j] & HAS_NEXT)) continue;
                if (L.rope[j] == myRope && abs(j - t) <= 2) continue;
                if (!box_meets(L.sb, B, j, lo, hi)) continue;
                if (cnt < SKIN_CAP) skinRow[cnt] = (uint16_t)j;
                cnt++;
            }
        }
    }
    L.vc[t] = cnt <= SKIN_CAP ? cnt : -1;
}

// The exact test of segment t against candidate j (the forward's: does t list j). rt, rj: the segments' detection
// radii.
__device__ __forceinline__ void rc_pair_test(const RcLds& L, int B, int j, float rt, float rj, f3 a0, f3 a1, int& cnt,
                                             uint16_t* fwd, float& lim, int& arg) {
    float d2;
    if (pair_hit(a1, ld3s(L.x, B, j) - a0, ld3s(L.x, B, j + 1) - a0, rt + rj, d2)) {
        if (cnt < MAX_COLL) {
            fwd[cnt] = (uint16_t)j;
            const float dj = 0.45f * sqrtf(d2);
            if (dj < lim) {
                lim = dj;
                arg = j;
            }
        }
        cnt++;
    }
}

// The substep's lists of segment t: exact tests over its skin list (or over its whole group after a skin-list
// overflow), candidates whose current boxes (padded by the detection radii) miss its own skipped first. Writes
// the partner count, limit and the limit's slot (-1: the padding) to LDS, and zeroes the reverse count.
template <int B>
__device__ __forceinline__ void rc_lists(const Params& P, const RcLds& L, int t, uint32_t fl, int g0, int g1,
                                         uint16_t* row, int* err) {
    uint16_t* fwd = row;
    const uint16_t* skin = row + 2 * MAX_COLL;
    int cnt = 0, arg = -1;
    float lim = INFINITY;
    if (fl & HAS_NEXT) {
        lim = L.pad[t] - P.detR;
        const f3 a0 = ld3s(L.x, B, t), x1 = ld3s(L.x, B, t + 1);
        const f3 a1 = x1 - a0;
        const float rt = fmaxf(L.pad[t], L.pad[t + 1]);
        const float mg = 1e-4f + 2e-6f * fmaxf(fabsf(a0.x), fmaxf(fabsf(a0.y), fabsf(a0.z)));
        const float rm = rt + mg;
        const f3 lo = mk3(fminf(a0.x, x1.x) - rm, fminf(a0.y, x1.y) - rm, fminf(a0.z, x1.z) - rm);
        const f3 hi = mk3(fmaxf(a0.x, x1.x) + rm, fmaxf(a0.y, x1.y) + rm, fmaxf(a0.z, x1.z) + rm);
        const int vc = L.vc[t];
        const int myRope = L.rope[t];
        const int n = vc < 0 ? g1 - g0 : vc;
        for (int e = 0; e < n; e++) {
            const int j = vc < 0 ? g0 + e : (int)skin[e];
            if (vc < 0 && (!(L.fl[j] & HAS_NEXT) || (L.rope[j] == myRope && abs(j - t) <= 2))) continue;
            const f3 c0 = ld3s(L.x, B, j), c1 = ld3s(L.x, B, j + 1);
            const float rj = fmaxf(L.pad[j], L.pad[j + 1]);
            if (fminf(c0.x, c1.x) - rj > hi.x || fminf(c0.y, c1.y) - rj > hi.y || fminf(c0.z, c1.z) - rj > hi.z ||
                fmaxf(c0.x, c1.x) + rj < lo.x || fmaxf(c0.y, c1.y) + rj < lo.y || fmaxf(c0.z, c1.z) + rj < lo.z)
                continue;
            rc_pair_test(L, B, j, rt, rj, a0, a1, cnt, fwd, lim, arg);
        }
    }
    if (cnt > MAX_COLL) *err = ERR_MAX_COLLISIONS;
    L.ncol[t] = min(cnt, MAX_COLL);
    L.nrev[t] = 0;
    L.lim[t] = lim;
    L.arg[t] = arg;
}

// One substep's detection by every lane of the workgroup: the rebuild vote (a barrier; it also publishes what
// the caller wrote before), a skin-list build when any slot voted or `first`, then the exact tests; ends with a
// barrier after which every slot's lists, counts and limits are in place.
template <int B>
__device__ __forceinline__ void rc_detect(const Params& P, const RcLds& L, int t, bool act, uint32_t fl, int g0, int g1,
                                          uint16_t* row, int* err, bool first) {
    const bool stale = first || rc_skin_stale(P, L, B, t, act, fl);
    if (__syncthreads_or(stale)) {
        rc_skin_boxes<B>(P, L, t, act, fl);
        __syncthreads();
        if (act) rc_skin_list<B>(L, t, fl, g0, g1, row + 2 * MAX_COLL);
    }
    if (act) rc_lists<B>(P, L, t, fl, g0, g1, row, err);
    else {
        L.ncol[t] = 0;
        L.nrev[t] = 0;
        L.lim[t] = INFINITY;
        L.arg[t] = -1;
    }
    __syncthreads();
}

// Pass 1 records each substep's lists for pass 2 in one word per vertex, slots relative to the contact group's
// first (groups hold at most 512 vertices: 10 bits): x the step limit; y the limit's partner (0x3FF: the padding)
// in bits 0-9, the list length in bits 10-17 and partner 0 in bits 18-27; z and w partners 1-3 and 4-6, 10 bits
// each. A segment with a longer list makes pass 2 detect that substep itself.
constexpr int TAPE_PARTNERS = 7;
__device__ __forceinline__ uint4 rc_tape_word(const RcLds& L, const uint16_t* row, int t, int g0) {
    const int n = L.ncol[t], a = L.arg[t];
    uint32_t pr[TAPE_PARTNERS] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
    for (int e = 0; e < min(n, TAPE_PARTNERS); e++) pr[e] = (uint32_t)(row[e] - g0);
    return make_uint4(__float_as_uint(L.lim[t]), (uint32_t)(a < 0 ? 0x3FF : a - g0) | ((uint32_t)n << 10) | (pr[0] << 18),
                      pr[1] | (pr[2] << 10) | (pr[3] << 20), pr[4] | (pr[5] << 10) | (pr[6] << 20));
}

// Pass 2's lists of a substep: from the tape words pass 1 recorded, or, when some segment of the workgroup has
// more than tapeMax partners, by the detection itself (rc_detect: the same lists and limits). Its skin lists are
// built at the first such substep (skinBuilt) and rebuilt when stale; they stay valid across the substeps read
// from the tape, since staleness is measured from the build. Every lane calls it; returns whether it detected
// (the same for every lane). The lists, counts and limits are in place after it.
template <int B>
__device__ __forceinline__ bool rc_lists_from_tape(const Params& P, const RcLds& L, int t, bool act, uint32_t fl,
                                                   int g0, int g1, uint16_t* row, uint4 w, int tapeMax,
                                                   bool& skinBuilt, int* err) {
    const int n = (int)((w.y >> 10) & 0xFFu);
    if (__syncthreads_or(act && n > tapeMax)) {
        rc_detect<B>(P, L, t, act, fl, g0, g1, row, err, !skinBuilt);
        skinBuilt = true;
        return true;
    }
    if (act) {
        L.ncol[t] = n;
        L.nrev[t] = 0;
        L.lim[t] = __uint_as_float(w.x);
        const uint32_t a = w.y & 0x3FFu;
        L.arg[t] = a == 0x3FFu ? -1 : g0 + (int)a;
        const uint32_t pr[TAPE_PARTNERS] = {(w.y >> 18) & 0x3FFu, w.z & 0x3FFu, (w.z >> 10) & 0x3FFu, (w.z >> 20) & 0x3FFu,
                                            w.w & 0x3FFu, (w.w >> 10) & 0x3FFu, (w.w >> 20) & 0x3FFu};
        for (int e = 0; e < n; e++) row[e] = (uint16_t)(g0 + (int)pr[e]);
    } else {
        L.ncol[t] = 0;
        L.nrev[t] = 0;
        L.lim[t] = INFINITY;
        L.arg[t] = -1;
    }
    __syncthreads();
    return false;
}

// Pair indexing (after rc_detect): the workgroup's pairs (segment t, entry e of its list)
// numbered t-major, poff[t] the first index of slot t (an exclusive prefix sum of the list lengths), the pair table
// (owner and partner of the first PAIR_TAB_OF(B) pairs), and the reverse lists: for each segment j, the indices of the
// pairs whose partner it is, ascending (so by owner). They are the transpose of the partner lists, filled through
// LDS counters and then sorted; a list of up to REV_SLOTS_OF(B) entries is kept in LDS, a longer one in its row in
// memory (REV: the reverse pass). Returns the number of mailbox chunks the workgroup's pairs take. Every lane
// calls it. The offsets and the table are in place on return; the reverse lists are read only after the barriers
// of the iteration that follows.
template <int B, bool REV = true>
__device__ __forceinline__ int rc_pair_index(const RcLds& L, int t, bool act, uint16_t* lists, int slots, int* err) {
    int* wsum = L.poff + B + 1;
    const int lane = t & 63, w = t >> 6;
    const int v = L.ncol[t];
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; k++) base += wsum[k];
    const int first = base + x - v;
    L.poff[t] = first;
    if (t == B - 1) L.poff[B] = base + x;
    if (act) {  // L.nrev is zero after the detection
        const uint16_t* row = lists + (size_t)t * LIST_STRIDE;
        for (int e = 0; e < v; e++) {
            const int j = row[e], q = first + e;
            if (q < PAIR_TAB_OF(B)) L.pairs[q] = (uint32_t)t | ((uint32_t)j << 16);
            if (!REV) continue;
            const int pos = atomicAdd(&L.nrev[j], 1);
            if (pos < REV_SLOTS_OF(B)) L.rq[j * REV_SLOTS_OF(B) + pos] = (uint16_t)q;
            if (pos < MAX_COLL) lists[(size_t)j * LIST_STRIDE + REV_IDX + pos] = (uint16_t)q;
        }
    }
    __syncthreads();
    if (REV && act) {
        int n = L.nrev[t];
        if (n > MAX_COLL) {
            *err = ERR_MAX_COLLISIONS;
            n = MAX_COLL;
            L.nrev[t] = n;
        }
        uint16_t* ri = lists + (size_t)t * LIST_STRIDE + REV_IDX;
        if constexpr (REV_SLOTS_OF(B) > 0) {  // else the pointer stays global: global, not flat, memory operations
            if (n <= REV_SLOTS_OF(B)) ri = L.rq + t * REV_SLOTS_OF(B);
        }
        for (int a = 1; a < n; a++) {  // insertion sort (short lists)
            const uint16_t ii = ri[a];
            int b = a - 1;
            while (b >= 0 && ri[b] > ii) {
                ri[b + 1] = ri[b];
                b--;
            }
            ri[b + 1] = ii;
        }
    }
    return (L.poff[B] + slots - 1) / slots;
}

// The reverse list of slot j (pair indices, ascending) and its length.
template <int B>
__device__ __forceinline__ const uint16_t* rc_rev_list(const RcLds& L, const uint16_t* lists, int j, int& n) {
    n = L.nrev[j];
    if constexpr (REV_SLOTS_OF(B) == 0) return lists + (size_t)j * LIST_STRIDE + REV_IDX;  // global loads
    return n <= REV_SLOTS_OF(B) ? L.rq + j * REV_SLOTS_OF(B) : lists + (size_t)j * LIST_STRIDE + REV_IDX;
}

// Owner slot and partner slot of pair q.
template <int B>
__device__ __forceinline__ void rc_pair_of(const RcLds& L, const uint16_t* lists, int q, int& o, int& j) {
    if (q < PAIR_TAB_OF(B)) {
        const uint32_t pr = L.pairs[q];
        o = (int)(pr & 0xFFFFu);
        j = (int)(pr >> 16);
        return;
    }
    int lo = 0, hi = B - 1;  // the last slot o with poff[o] <= q
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (L.poff[m] <= q)
            lo = m;
        else
            hi = m - 1;
    }
    o = lo;
    j = lists[(size_t)o * LIST_STRIDE + (q - L.poff[o])];
}

// Rope-rope contact modes of the adjoint kernels. With rope contact a frame's contact groups split in two: those
// in which some segment had a rope-rope pair at some substep (RC_PAIRS: detection, pair terms, the step limit's
// pair distances) and the others (RC_CLAMP: their lists are empty all frame, so the step limit of a segment is
// its padding less the detection radius and only the step clamp remains). Both kernels run over the same
// workgroups; each workgroup keeps to the one its groups' flags name.
enum { RC_NONE = 0, RC_CLAMP = 1, RC_PAIRS = 2 };

// The adjoint of pair q of the workgroup (pair_term_adj of owner segment o and its listed partner j, with the
// Newton adjoints u and directions of o, o + 1 from the exchange), left in mailbox slot e.
template <int B, bool F>
__device__ __forceinline__ void rc_pair_eval(const PairMat& pm, float dmp, const RcLds& L, const float* ex,
                                             const uint16_t* lists, int q, int e, int slots) {
    int o, j;
    rc_pair_of<B>(L, lists, q, o, j);
    f3 g[8];
    pair_term_adj<F>(pm, dmp, ld3s(L.x, B, o), ld3s(L.x, B, o + 1), ld3s(L.dk, B, o), ld3s(L.dk, B, o + 1),
                     ld3s(L.x, B, j), ld3s(L.x, B, j + 1), ld3s(L.dk, B, j), ld3s(L.dk, B, j + 1),
                     ld3s(ex + EX_U * B, B, o), ld3s(ex + (EX_U + 3) * B, B, o), ld3s(ex + EX_U * B, B, o + 1),
                     ld3s(ex + (EX_U + 3) * B, B, o + 1), g);
#pragma unroll
    for (int m = 0; m < 8; m++) st3s(L.mbx + 3 * m * slots, slots, e, g[m]);
}

// Vertex t's parts of the pair adjoints in the mailbox, for the pairs [q0, q1) it holds: of the pairs of segment
// t - 1 (to its second vertex, t) and of segment t (to its first vertex), and of the pairs whose partner segment
// holds it (reverse lists of t, side 0, and of t - 1, side 1). The sources are merged in ascending pair index,
// which within a contact group is the group's own order, so the sums do not depend on how the pairs are chunked
// or on which groups share the workgroup. The cursors carry over from chunk to chunk.
template <int B>
__device__ __forceinline__ void rc_pairs_gather(const RcLds& L, const uint16_t* lists, int t, bool hp, int q0, int q1,
                                                int slots, int& qo, int& c0, int& c1, f3& aX, f3& aD) {
    const int mid = L.poff[t], own1 = L.poff[t + 1];
    int nr0 = 0, nr1 = 0;
    const uint16_t* ri0 = rc_rev_list<B>(L, lists, t, nr0);
    const uint16_t* ri1 = hp ? rc_rev_list<B>(L, lists, t - 1, nr1) : ri0;
    if (!hp) nr1 = 0;
    const float* mb = L.mbx;
    while (true) {  // next source in ascending pair index (each pair is in at most one of them)
        const int a = qo < min(own1, q1) ? qo : INT_MAX;
        const int b0 = c0 < nr0 ? (int)ri0[c0] : INT_MAX;
        const int b1 = c1 < nr1 ? (int)ri1[c1] : INT_MAX;
        const int q = min(a, min(b0, b1));
        if (q >= q1) break;
        const int e = q - q0;
        if (q == a) {
            if (q < mid) {  // a pair of segment t - 1: its second vertex's part
                aX = aX + ld3s(mb + 3 * slots, slots, e);
                aD = aD + ld3s(mb + 9 * slots, slots, e);
            } else {
                aX = aX + ld3s(mb, slots, e);
                aD = aD + ld3s(mb + 6 * slots, slots, e);
            }
            qo++;
        } else if (q == b0) {
            aX = aX + ld3s(mb + 12 * slots, slots, e);
            aD = aD + ld3s(mb + 18 * slots, slots, e);
            c0++;
        } else {
            aX = aX + ld3s(mb + 15 * slots, slots, e);
            aD = aD + ld3s(mb + 21 * slots, slots, e);
            c1++;
        }
    }
}

// The pair adjoints of one iteration of the reverse pass. Each pair is evaluated once (by the workgroup's lanes
// with EVAL, else by its pair waves, rc_pair_waves) and leaves its eight adjoints in the mailbox. When the
// substep's pairs fit in one mailbox chunk (nch == 1) they are only evaluated here: the vertices gather them at the
// start of the next iteration's Newton adjoint (or after the last iteration), so the evaluation overlaps the
// segment terms and the next orientation sweep. Otherwise the chunks go through here one by one, each gathered
// after a barrier. Every vertex lane calls it with the same nch.
template <int B, bool F, bool EVAL>
__device__ __forceinline__ void rc_pairs_adj(const Params& P, const RcLds& L, const float* ex, const uint16_t* lists,
                                             int t, bool act, bool hp, int nch, int slots, float dmp, f3& aX,
                                             f3& aD) {
    const int ntot = L.poff[B];
    const PairMat pm = pair_mat(P);
    if (nch == 1) {
        if (EVAL)
            for (int q = t; q < ntot; q += B) rc_pair_eval<B, F>(pm, dmp, L, ex, lists, q, q, slots);
        return;
    }
    int qo = (act && hp) ? L.poff[t - 1] : act ? L.poff[t] : 0, c0 = 0, c1 = 0;
    for (int ch = 0; ch < nch; ch++) {
        const int q0 = ch * slots, q1 = min(q0 + slots, ntot);
        if (EVAL)
            for (int q = q0 + t; q < q1; q += B) rc_pair_eval<B, F>(pm, dmp, L, ex, lists, q, q - q0, slots);
        __syncthreads();
        if (act) rc_pairs_gather<B>(L, lists, t, hp, q0, q1, slots, qo, c0, c1, aX, aD);
        if (ch + 1 < nch) __syncthreads();
    }
}

// Pass 1's pair terms: collision_eval of pair q (owner o, partner j, at the iterate dk) left in mailbox slot e as
// PAIR_TERM_N floats: n [3], s1, s2, h1, h2, u [3], a1, a2, b1, b2, unused, on. The owner adds the first-vertex part
// and vertex o + 1 the second-vertex part, each in list order, which is collision_apply's arithmetic split by
// accumulator (its f, H and f2, H2 updates never mix).
constexpr int PAIR_TERM_N = 16;
template <int B, bool F>
__device__ __forceinline__ void rc_term_eval(const Params& P, float dmp, const RcLds& L, const float* sX,
                                             const float* dk, const uint16_t* lists, int q, int e, int slots) {
    int o, j;
    rc_pair_of<B>(L, lists, q, o, j);
    const PairTerm T = collision_eval<F>(P, dmp, ld3s(sX, B, o), ld3s(sX, B, o + 1), ld3s(dk, B, o), ld3s(dk, B, o + 1),
                                         ld3s(sX, B, j), ld3s(sX, B, j + 1), ld3s(dk, B, j), ld3s(dk, B, j + 1));
    float* mb = L.mbx;
    mb[15 * slots + e] = __int_as_float(T.on);
    if (!(T.on & 1)) return;
    st3s(mb, slots, e, T.n);
    mb[3 * slots + e] = T.s1;
    mb[4 * slots + e] = T.s2;
    mb[5 * slots + e] = T.h1;
    mb[6 * slots + e] = T.h2;
    if (!(T.on & 2)) return;
    st3s(mb + 7 * slots, slots, e, T.u);
    mb[10 * slots + e] = T.a1;
    mb[11 * slots + e] = T.a2;
    mb[12 * slots + e] = T.b1;
    mb[13 * slots + e] = T.b2;
}
// second = 0: the owner's part (f, H); 1: the part of the owner's next vertex (f2, H2)
template <bool F>
__device__ __forceinline__ void rc_term_apply(const float* mb, int slots, int e, int second, f3& f, s3& H) {
    using M = V<F>;
    const int on = __float_as_int(mb[15 * slots + e]);
    if (!(on & 1)) return;
    const f3 n = ld3s(mb, slots, e);
    f = M::axpy(mb[(3 + second) * slots + e], n, f);
    s3 op = M::souter(n);
    M::sadd_scaled(H, op, mb[(5 + second) * slots + e]);
    if (on & 2) {
        sadd_diag(op, -1.f);
        f = M::axpy(mb[(10 + second) * slots + e], ld3s(mb + 7 * slots, slots, e), f);
        M::sadd_scaled(H, op, mb[(12 + second) * slots + e]);
    }
}

// Lanes of the reverse pass's pair waves (threads B .. B + PAIR_LANES - 1 of a workgroup with rope-rope pairs):
// they hold no vertex, so the pair adjoints they evaluate do not compete with the vertex lanes' state for
// registers, and they run while the vertex lanes take their segment terms.
constexpr int PAIR_LANES = 128;
template <int B, int RCM>
constexpr bool pair_waves() { return RCM == RC_PAIRS && B <= 384; }

// The pair waves' side of the reverse pass: the same sequence of barriers as the vertex lanes (substep head,
// the lists' vote, with a detection its vote and rebuild, the lists' end, the two of the pair indexing; per iteration the orientation
// sweep, the Newton adjoint and, with more than one mailbox chunk, the chunks'; after the iterations the two of
// the step-limit adjoints; the two of the drag sum), evaluating each iteration's pairs after the Newton adjoint's
// barrier.
template <int B, bool F>
__device__ void rc_pair_waves(const Params& P, const RcLds& L, const float* ex, const uint16_t* lists, int lane, int S,
                              int I, int slots) {
    const PairMat pm = pair_mat(P);
    const float dmp = Ar<F>::div(P.damping, P.h);
    for (int s = S - 1; s >= 0; s--) {
        __syncthreads();  // substep head
        if (__syncthreads_or(0)) {  // lists: the vertex lanes detect this substep (rc_detect's barriers)
            if (__syncthreads_or(0)) __syncthreads();
        }
        __syncthreads();  // lists in place
        __syncthreads();  // pair indexing: offsets
        __syncthreads();  // pair indexing: reverse lists
        const int ntot = L.poff[B];
        const int nch = (ntot + slots - 1) / slots;
        for (int k = I - 1; k >= 0; k--) {
            __syncthreads();  // orientation sweep
            __syncthreads();  // Newton adjoint u, direction, d^k in place
            if (nch == 1) {  // gathered after the next barrier
                for (int q = lane; q < ntot; q += PAIR_LANES) rc_pair_eval<B, F>(pm, dmp, L, ex, lists, q, q, slots);
                continue;
            }
            for (int ch = 0; ch < nch; ch++) {
                const int q0 = ch * slots, q1 = min(q0 + slots, ntot);
                for (int q = q0 + lane; q < q1; q += PAIR_LANES) rc_pair_eval<B, F>(pm, dmp, L, ex, lists, q, q - q0, slots);
                __syncthreads();  // mailbox chunk full
                if (ch + 1 < nch) __syncthreads();  // chunk gathered
            }
        }
        __syncthreads();  // after the iterations: pushes and pairs of the last one
        __syncthreads();  // step-limit adjoints
    }
    __syncthreads();  // drag sum
    __syncthreads();
}

__device__ __forceinline__ bool rc_wrong_class(const Params& P, const AdjDev& A, int RCM, int rb0, int rb1) {
    if (RCM == RC_NONE) return false;
    bool pairs = A.pairFlag == nullptr;
    for (int g = rb0 / P.groupRopes; !pairs && g <= (rb1 - 1) / P.groupRopes; g++) pairs = A.pairFlag[g] != 0;
    return pairs != (RCM == RC_PAIRS);
}

// Local slots [g0, g1) of the contact group of rope r in a workgroup whose first vertex is v0.
__device__ __forceinline__ void group_range(const Params& P, const Dev& D, int r, int v0, int& g0, int& g1) {
    const int G = P.groupRopes, g = r / G;
    g0 = D.roff[min(g * G, P.nropes)] - v0;
    g1 = D.roff[min((g + 1) * G, P.nropes)] - v0;
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
// Held at three waves per SIMD, except with rope-rope pairs: that form's LDS keeps it to one workgroup per
// compute unit, and the registers go to the pair terms.
template <int B, bool F, int RCM>
__global__ __launch_bounds__(B, RCM == RC_PAIRS ? 2 : 3) void k_adjoint_forward(Params P, Dev D, AdjDev A, Drive Dr,
                                                                               int S) {
    constexpr bool RC = RCM == RC_PAIRS;
    __shared__ float4 sQ[2][B];
    __shared__ float sX[3 * B], sD[2][3 * B], sF2[9 * B];
    extern __shared__ Pose sPoseAll[];  // the poses of the workgroup's ropes, sized at launch; then, with rope
                                        // contact, the detection arrays (rc_lds_bytes)
    const int I = P.numItr;
    const int rb0 = A.blk[blockIdx.x], rb1 = A.blk[blockIdx.x + 1], t = threadIdx.x;
    if (rc_wrong_class(P, A, RCM, rb0, rb1)) return;
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
    RcLds L{};
    if (RCM == RC_PAIRS) L = rc_lds<B>(reinterpret_cast<char*>(sPoseAll + nPose), sX);
    if (RCM == RC_CLAMP) L = rc_lds_clamp<B>(reinterpret_cast<char*>(sPoseAll + nPose));
    uint16_t* lists0 = RC ? A.lists + (size_t)(li - t) * LIST_STRIDE : nullptr;  // the workgroup's list rows
    const float dmp = Ar<F>::div(P.damping, h);
    if (RC) {
        L.fl[t] = act ? (int)c.fl : 0;
        L.rope[t] = r;
    }

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
            if (RCM != RC_NONE) L.pad[t] = padding;
        }
        int nch = 0;  // mailbox chunks of the substep's pair terms
        if (RC) {  // rope-rope lists and step limits of the substep (the forward's detection); its vote is the barrier
            int g0, g1;
            group_range(P, D, r, v0, g0, g1);
            rc_detect<B>(P, L, t, act, c.fl, g0, g1, A.lists + (size_t)(act ? li : 0) * LIST_STRIDE, D.err, s == 0);
            if (act) A.hR[hs(A, s, li)] = rc_tape_word(L, A.lists + (size_t)li * LIST_STRIDE, t, g0);
            nch = rc_pair_index<B, false>(L, t, act, lists0, A.mbSlots, D.err);
        } else {
            __syncthreads();
        }
        uint64_t hitBits = 0;
        const int ntot = RC ? L.poff[B] : 0;
        for (int it = 0; it < I; it++) {
            const float* dk = sD[it & 1];
            float* dn = sD[(it + 1) & 1];
            f3 f, f2 = mk3(0.f, 0.f, 0.f);
            s3 H, H2 = szero();
            // with one mailbox chunk, iteration 0's pair terms are evaluated here and later ones during the
            // previous iteration's orientation sweep
            if (RC && nch == 1 && it == 0)
                for (int q = t; q < ntot; q += B) rc_term_eval<B, F>(P, dmp, L, sX, dk, lists0, q, q, A.mbSlots);
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
            // the next vertex's stretch terms of segment t - 1, then (rope contact) the rope-rope terms in list
            // order as the forward adds them: segment t's to f, H, segment t - 1's to its next-vertex part
            f3 fp = mk3(0.f, 0.f, 0.f);
            s3 Hp = szero();
            if (RC && act && (c.fl & HAS_PREV)) {
                fp = ld3s(sF2, B, t - 1);
#pragma unroll
                for (int m = 0; m < 6; m++) Hp.m[m] = sF2[(3 + m) * B + t - 1];
            }
            if (RC && nch > 0) {
                const int o0 = act ? L.poff[t] : 0, o1 = act ? L.poff[t + 1] : 0;
                const int pp0 = (act && (c.fl & HAS_PREV)) ? L.poff[t - 1] : o0;
                for (int ch = 0; ch < nch; ch++) {
                    const int q0 = ch * A.mbSlots, q1 = min(q0 + A.mbSlots, ntot);
                    if (nch > 1) {
                        for (int q = q0 + t; q < q1; q += B) rc_term_eval<B, F>(P, dmp, L, sX, dk, lists0, q, q - q0, A.mbSlots);
                        __syncthreads();
                    }
                    for (int q = max(o0, q0); q < min(o1, q1); q++) rc_term_apply<F>(L.mbx, A.mbSlots, q - q0, 0, f, H);
                    for (int q = max(pp0, q0); q < min(o0, q1); q++) rc_term_apply<F>(L.mbx, A.mbSlots, q - q0, 1, fp, Hp);
                    if (ch + 1 < nch) __syncthreads();
                }
            }
            if (act) {
                if (c.fl & HAS_PREV) {
                    if (!RC) {
                        fp = ld3s(sF2, B, t - 1);
#pragma unroll
                        for (int m = 0; m < 6; m++) Hp.m[m] = sF2[(3 + m) * B + t - 1];
                    }
                    f = f + fp;
                    sadd(H, Hp);
                }
                f3 d = ld3s(dk, B, t), dl = mk3(0.f, 0.f, 0.f);
                if (c.w != 0.f) {
                    float lim = INFINITY;
                    if (RC) {  // the step limit of the forward: min(limit_t, limit_t-1)
                        lim = L.lim[t];
                        if (c.fl & HAS_PREV) lim = fminf(lim, L.lim[t - 1]);
                    } else if (RCM == RC_CLAMP) {  // no pairs: a segment's limit is its padding less detR
                        if (c.fl & HAS_NEXT) lim = L.pad[t] - P.detR;
                        if (c.fl & HAS_PREV) lim = fminf(lim, L.pad[t - 1] - P.detR);
                    }
                    dl = V<F>::ssolve(H, f);
                    d = solve_update<F>(P, d, H, f, RCM != RC_NONE, lim);
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
            if (RC && nch == 1 && it + 1 < I)  // the next iteration's pair terms (its d^k is in place)
                for (int q = t; q < ntot; q += B) rc_term_eval<B, F>(P, dmp, L, sX, dn, lists0, q, q, A.mbSlots);
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
// Per iteration three phases and two barriers: orientation adjoint (pushes to the neighbours), Newton adjoint
// u = a H^-1 d̄ (after gathering the orientation pushes, and the previous iteration's segment-term pushes and
// pair parts), segment-term adjoint (pushes to the next vertex, gathered after the next orientation sweep).
template <int B, bool F, int RCM>
__global__ __launch_bounds__(B + (pair_waves<B, RCM>() ? PAIR_LANES : 0)) void k_adjoint_reverse(Params P, Dev D,
                                                                                                  AdjDev A, Drive Dr,
                                                                                                  int S) {
    constexpr bool RC = RCM == RC_PAIRS;
    constexpr bool PW = pair_waves<B, RCM>();
    __shared__ float ex[EX_N * B];
    extern __shared__ Pose sPoseAll[];  // the poses of the workgroup's ropes, sized at launch; then, with rope
                                        // contact, the detection arrays (rc_lds_bytes)
    const int I = P.numItr;
    const int rb0 = A.blk[blockIdx.x], rb1 = A.blk[blockIdx.x + 1], t = threadIdx.x;
    if (rc_wrong_class(P, A, RCM, rb0, rb1)) return;
    const int v0 = D.roff[rb0], n = D.roff[rb1] - v0;
    if (PW && t >= B) {  // a pair wave
        const RcLds L = rc_lds<B>(reinterpret_cast<char*>(sPoseAll + (size_t)(rb1 - rb0) * P.numSdf), nullptr);
        rc_pair_waves<B, F>(P, L, ex, A.lists + (size_t)(v0 - A.hv0) * LIST_STRIDE, t - B, S, I, A.mbSlots);
        return;
    }
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
    RcLds L{};
    if (RCM == RC_PAIRS) L = rc_lds<B>(reinterpret_cast<char*>(sPoseAll + nPose), nullptr);
    if (RCM == RC_CLAMP) L = rc_lds_clamp<B>(reinterpret_cast<char*>(sPoseAll + nPose));
    if (RC) {
        L.fl[t] = act ? (int)c.fl : 0;
        L.rope[t] = r;
    }
    // list rows of slot k (the workgroup's vertices are consecutive local vertices): partners, then reverse
    auto list_row = [&](int k) { return A.lists + (size_t)(li - t + k) * LIST_STRIDE; };

    bool skinBuilt = false;  // (rope-rope pairs) pass 2's own skin lists exist
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
            if (!RC) {  // (with rope-rope pairs they are loaded where used, not held through the iterations)
                vs = ld3(A.hV[hs(A, s, li)]);
                vls = ld3(A.hVL[hs(A, s, li)]);
            }
            const uint2 cc = A.hC[hs(A, s, li)];
            cand = (uint64_t)cc.x | ((uint64_t)cc.y << 32);
            const uint2 mm = A.hM[hs(A, s, li)];
            hitMask = (uint64_t)mm.x | ((uint64_t)mm.y << 32);
        }
        // end update: x' = X + d, v' = (d / h)(1 - drag h) on free vertices
        f3 aD = ax, aX = ax, aY = mk3(0.f, 0.f, 0.f);
        if (RCM != RC_NONE) {
            L.mown[t] = 0.f;  // adjoints of this substep's step limits (own segment's, previous segment's)
            L.mprev[t] = 0.f;
            if (act) {
                if (RC) st3s(L.x, B, t, X0);
                L.pad[t] = P.scaledDetR + M::len(y);  // initItr's padding of this substep
            }
        }
        q4 aQ = aq;
        if (act && c.w != 0.f) {
            const f3 d = ld3(A.hD[hk(A, I, s, I, li)]);
            const float invH = Af::rcp(h);
            aD = aD + (invH * (1.f - P.drag * h)) * av;
            gdrag -= (double)h * (double)dot(d * invH, av);
        }
        int nch = 0;  // mailbox chunks of the substep's pairs
        if (RC) {  // the substep's rope-rope lists and step limits, with the lists of the segments listing each one
            int g0, g1;
            group_range(P, D, r, v0, g0, g1);
            const uint4 w = act ? A.hR[hs(A, s, li)] : make_uint4(0u, 0x3FFu, 0u, 0u);
            rc_lists_from_tape<B>(P, L, t, act, c.fl, g0, g1, list_row(act ? t : 0), w, A.tapeMax, skinBuilt, D.err);
            nch = rc_pair_index<B>(L, t, act, list_row(0), A.mbSlots, D.err);
        } else {
            __syncthreads();  // sPose of a moving frame is in place
        }

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
            // iteration k-1's history is loaded here and lands while k computes
            IterHist nxt = cur;
            if (k > 0) nxt = load_iter(k - 1, cur.d0, cur.d1);
            const f3 dn0 = cur.dn0, dn1 = cur.dn1, d0 = cur.d0, d1 = cur.d1;
            const q4 q0 = cur.q0, qp = cur.qp, qn = cur.qn;
            const s3 H = cur.H;
            const f3 dl = cur.dl;
            // -- orientation sweep k: q^{k+1}_t from d^{k+1} and q^k_{t-1..t+1}
            q4 keepQ = aQ;
            if (act) {
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
            // -- gather the orientation pushes, and the previous iteration's segment-term pushes and pair parts;
            //    Newton adjoint u = a H^-1 d̄^{k+1}
            f3 u = mk3(0.f, 0.f, 0.f);
            if (act) {
                if (k < I - 1) {
                    if (hp) {
                        aD = aD + ld3s(ex + EX_P * B, B, t - 1);
                        aX = aX + ld3s(ex + (EX_P + 3) * B, B, t - 1);
                    }
                    if (RC && nch == 1) {
                        int qo = hp ? L.poff[t - 1] : L.poff[t], c0 = 0, c1 = 0;
                        rc_pairs_gather<B>(L, list_row(0), t, hp, 0, L.poff[B], A.mbSlots, qo, c0, c1, aX, aD);
                    }
                }
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
                if (RCM != RC_NONE && c.w != 0.f) {  // the step clamp d' = z lim / |z| of the forward, z = d^k + a H^-1 F
                    const float lo = RC ? L.lim[t] : (hn ? L.pad[t] - P.detR : INFINITY);
                    const float lp = !hp ? INFINITY : RC ? L.lim[t - 1] : L.pad[t - 1] - P.detR;
                    const float lim = lp < lo ? lp : lo;  // fminf(limit_t, limit_t-1), and which one it is
                    const f3 z = M::axpy(P.accel, dl, d0);
                    const float lz = M::len(z);
                    if (lz > lim && lz > 0.f) {
                        const f3 zh = z * (1.f / lz);
                        const float limb = dot(zh, aD);
                        aD = (lim / lz) * (aD - limb * zh);
                        if (lp < lo)
                            L.mprev[t] += limb;
                        else
                            L.mown[t] += limb;
                    }
                }
                if (c.w != 0.f) u = P.accel * M::ssolve(H, aD);
                st3s(ex + EX_U * B, B, t, u);
                st3s(ex + (EX_U + 3) * B, B, t, dl);
                if (RC) st3s(L.dk, B, t, d0);
            }
            __syncthreads();
            // -- adjoint of the segment terms of vertex t (inertia, surfaces, stretch of segment (t, t+1))
            f3 pD = mk3(0.f, 0.f, 0.f), pX = pD;
            if (act) {
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
            }
            if (RC) rc_pairs_adj<B, F, !PW>(P, L, ex, list_row(0), t, act, hp, nch, A.mbSlots, dmp, aX, aD);
            if (act) {
                st3s(ex + EX_P * B, B, t, pD);
                st3s(ex + (EX_P + 3) * B, B, t, pX);
            }
            // no barrier here: the pushes and the pair parts are gathered after the next orientation sweep's
            cur = nxt;
        }
        __syncthreads();  // the last iteration's pushes and pair parts, every clamp adjoint of the substep
        if (act) {
            if (hp) {
                aD = aD + ld3s(ex + EX_P * B, B, t - 1);
                aX = aX + ld3s(ex + (EX_P + 3) * B, B, t - 1);
            }
            if (RC && nch == 1) {
                int qo = hp ? L.poff[t - 1] : L.poff[t], c0 = 0, c1 = 0;
                rc_pairs_gather<B>(L, list_row(0), t, hp, 0, L.poff[B], A.mbSlots, qo, c0, c1, aX, aD);
            }
        }
        f3 aYp = mk3(0.f, 0.f, 0.f);  // the adjoint of y through the detection padding
        if (RCM != RC_NONE) {  // step-limit adjoints: limit_t = min(pad_t - detR, 0.45 |d_tj| over the listed j)
            const float Mt = L.mown[t] + ((act && hn) ? L.mprev[t + 1] : 0.f);
            L.mb[t] = Mt;
            __syncthreads();
            if (act) {
                f3 g4[4];
                if (Mt != 0.f) {
                    const int a = RC ? L.arg[t] : -1;
                    if (a < 0) {
                        aYp = (Mt / len(y)) * y;
                    } else {
                        limit_dist_adj(X0, X1, ld3s(L.x, B, a), ld3s(L.x, B, a + 1), Mt, g4);
                        aX = aX + g4[0];
                    }
                }
                if (RC && hp) {
                    const float Mp = L.mb[t - 1];
                    const int a = L.arg[t - 1];
                    if (Mp != 0.f && a >= 0) {
                        limit_dist_adj(ld3s(L.x, B, t - 1), X0, ld3s(L.x, B, a), ld3s(L.x, B, a + 1), Mp, g4);
                        aX = aX + g4[1];
                    }
                }
                for (int side = 0; RC && side < 2; side++) {
                    const int jt = t - side;
                    if (side == 1 && !hp) break;
                    int nr;
                    const uint16_t* rr = rc_rev_list<B>(L, list_row(0), jt, nr);
                    for (int e = 0; e < nr; e++) {
                        int k, jj;
                        rc_pair_of<B>(L, list_row(0), rr[e], k, jj);
                        const float Mk = L.mb[k];
                        if (Mk == 0.f || L.arg[k] != jt) continue;
                        limit_dist_adj(ld3s(L.x, B, k), ld3s(L.x, B, k + 1), ld3s(L.x, B, jt), ld3s(L.x, B, jt + 1), Mk,
                                       g4);
                        aX = aX + g4[2 + side];
                    }
                }
            }
        }

        // init: d^0 = h v (+ h^2 s g), y = h v + h^2 g, last velocity <- v
        if (act) {
            f3 vb = h * aD + avl, vlb = mk3(0.f, 0.f, 0.f);
            if (RCM != RC_NONE) vb = vb + h * aYp;
            if (c.w != 0.f) {
                vb = vb + h * aY;
                const f3 g = mk3(P.gx, P.gy, P.gz);
                const float g2 = M::len2(g);
                if (g2 > 0.f) {
                    if (RC) {
                        vs = ld3(A.hV[hs(A, s, li)]);
                        vls = ld3(A.hVL[hs(A, s, li)]);
                    }
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

size_t adjoint_rc_lds_bytes(int block) {
    return sizeof(float) * ((size_t)22 * block + 6 * (block / 16)) + sizeof(int) * 6 * (size_t)block;
}

size_t adjoint_list_entries_per_vertex() { return LIST_STRIDE; }

size_t adjoint_history_bytes(int numItr) {
    return 5 * sizeof(float4) + 2 * sizeof(uint2) + (size_t)(numItr + 1) * 2 * sizeof(float4) +
           (size_t)numItr * 9 * sizeof(float);
}

bool adjoint_block_supported(int block, bool reverse) {
    (void)reverse;
    return block == 64 || block == 128 || block == 192 || block == 256 || block == 384 || block == 512;
}

namespace {
// Waves per SIMD that a kernel's VGPR allocation allows (512 VGPRs per SIMD lane, granule 8, at most 8).
int waves_of(const void* fn) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, fn) != hipSuccess || a.numRegs <= 0) return 1;
    const int regs = (a.numRegs + 7) / 8 * 8;
    return std::max(1, std::min(8, 512 / regs));
}
template <int B, int RC>
const void* pass_fn(bool reverse, bool fast) {
    if (reverse)
        return fast ? reinterpret_cast<const void*>(&k_adjoint_reverse<B, true, RC>)
                    : reinterpret_cast<const void*>(&k_adjoint_reverse<B, false, RC>);
    return fast ? reinterpret_cast<const void*>(&k_adjoint_forward<B, true, RC>)
                : reinterpret_cast<const void*>(&k_adjoint_forward<B, false, RC>);
}
template <int B>
const void* pass_fn(bool reverse, bool fast, bool rc) {
    return rc ? pass_fn<B, RC_PAIRS>(reverse, fast) : pass_fn<B, RC_NONE>(reverse, fast);
}
}  // namespace

int adjoint_waves_per_simd(int block, bool reverse, bool fast, bool ropeContact) {
    switch (block) {
        case 64: return waves_of(pass_fn<64>(reverse, fast, ropeContact));
        case 128: return waves_of(pass_fn<128>(reverse, fast, ropeContact));
        case 192: return waves_of(pass_fn<192>(reverse, fast, ropeContact));
        case 256: return waves_of(pass_fn<256>(reverse, fast, ropeContact));
        case 384: return waves_of(pass_fn<384>(reverse, fast, ropeContact));
        case 512: return waves_of(pass_fn<512>(reverse, fast, ropeContact));
        default: return 1;
    }
}

// LDS of a gfx950 compute unit, all of which one workgroup may take.
constexpr size_t LDS_BYTES = 160 * 1024;

template <int B, bool F, int RC>
static hipError_t launch_pass_f(bool reverse, const Params& P, const Dev& D, const AdjDev& A, const Drive& Dr, int S,
                                int nblocks, size_t lds, hipStream_t s) {
    // dynamic LDS past the default limit needs the kernel's attribute raised once: to all the LDS its static
    // arrays leave
    static size_t stat[2] = {0, 0};
    const void* fn = pass_fn<B, RC>(reverse, F);
    if (!stat[reverse]) {
        hipFuncAttributes a{};
        hipError_t e = hipFuncGetAttributes(&a, fn);
        if (e != hipSuccess) return e;
        if (a.sharedSizeBytes >= LDS_BYTES) return hipErrorInvalidValue;
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(LDS_BYTES - a.sharedSizeBytes));
        if (e != hipSuccess) return e;
        stat[reverse] = a.sharedSizeBytes > 0 ? a.sharedSizeBytes : 1;
    }
    if (RC == RC_CLAMP) lds = lds - adjoint_rc_lds_bytes(B) + rc_lds_clamp_bytes(B);  // the compact layout
    AdjDev Ak = A;
    Ak.mbSlots = 0;
    // ROPESIM_ADJ_TAPE_PARTNERS (0..7) lowers the list length pass 2 reads from pass 1's record, so that tests can
    // send substeps through pass 2's own detection; the gradients do not depend on it
    Ak.tapeMax = TAPE_PARTNERS;
    if (const char* tp = getenv("ROPESIM_ADJ_TAPE_PARTNERS")) Ak.tapeMax = std::max(0, std::min(TAPE_PARTNERS, atoi(tp)));
    if (RC == RC_PAIRS) {  // the pair mailbox takes what is left, up to two slots per thread
        const size_t per = (reverse ? PAIR_ADJ_N : PAIR_TERM_N) * sizeof(float);
        const size_t used = stat[reverse] + lds + pair_index_lds_bytes(B);
        const size_t fit = used < LDS_BYTES ? (LDS_BYTES - used) / per / 64 * 64 : 0;
        Ak.mbSlots = (int)std::min<size_t>(2 * B, fit);
        if (Ak.mbSlots < 64) return hipErrorInvalidValue;
        // ROPESIM_ADJ_MAILBOX_SLOTS caps the slots (at least 4), so that tests can send the pairs through several
        // mailbox chunks; the gradients do not depend on it
        if (const char* cap = getenv("ROPESIM_ADJ_MAILBOX_SLOTS")) {
            const int c = atoi(cap);
            if (c >= 4) Ak.mbSlots = std::min(Ak.mbSlots, c);
        }
        lds += pair_index_lds_bytes(B) + (size_t)Ak.mbSlots * per;
    }
    if (reverse)
        hipLaunchKernelGGL((k_adjoint_reverse<B, F, RC>), dim3(nblocks), dim3(B + (pair_waves<B, RC>() ? PAIR_LANES : 0)),
                           lds, s, P, D, Ak, Dr, S);
    else
        hipLaunchKernelGGL((k_adjoint_forward<B, F, RC>), dim3(nblocks), dim3(B), lds, s, P, D, Ak, Dr, S);
    return hipGetLastError();
}
template <int B>
static hipError_t launch_pass_t(bool reverse, const Params& P, const Dev& D, const AdjDev& A, const Drive& Dr, int S,
                                int nblocks, size_t lds, hipStream_t s, int rcClass) {
    if (!P.ropeContact) {
        if (P.fastMath) return launch_pass_f<B, true, RC_NONE>(reverse, P, D, A, Dr, S, nblocks, lds, s);
        return launch_pass_f<B, false, RC_NONE>(reverse, P, D, A, Dr, S, nblocks, lds, s);
    }
    // the clamp-only kernel and the full one over the same workgroups; each workgroup runs in one of them
    if (rcClass != RC_PAIRS) {
        const hipError_t e = P.fastMath ? launch_pass_f<B, true, RC_CLAMP>(reverse, P, D, A, Dr, S, nblocks, lds, s)
                                        : launch_pass_f<B, false, RC_CLAMP>(reverse, P, D, A, Dr, S, nblocks, lds, s);
        if (e != hipSuccess || rcClass == RC_CLAMP) return e;
    }
    return P.fastMath ? launch_pass_f<B, true, RC_PAIRS>(reverse, P, D, A, Dr, S, nblocks, lds, s)
                      : launch_pass_f<B, false, RC_PAIRS>(reverse, P, D, A, Dr, S, nblocks, lds, s);
}

hipError_t launch_adjoint_pass(bool reverse, int block, const Params& P, const Dev& D, const AdjDev& A, const Drive& Dr,
                               int S, int nblocks, int ropesPerBlock, hipStream_t s, int rcClass) {
    if (nblocks <= 0) return hipSuccess;
    if (!adjoint_block_supported(block, reverse) || ropesPerBlock < 1 ||
        (size_t)ropesPerBlock * P.numSdf > (size_t)adjoint_pose_slots(block) || (P.ropeContact && !A.lists))
        return hipErrorInvalidValue;
    const size_t lds = (size_t)ropesPerBlock * P.numSdf * sizeof(Pose) + (P.ropeContact ? adjoint_rc_lds_bytes(block) : 0);
    switch (block) {
        case 64: return launch_pass_t<64>(reverse, P, D, A, Dr, S, nblocks, lds, s, rcClass);
        case 128: return launch_pass_t<128>(reverse, P, D, A, Dr, S, nblocks, lds, s, rcClass);
        case 192: return launch_pass_t<192>(reverse, P, D, A, Dr, S, nblocks, lds, s, rcClass);
        case 256: return launch_pass_t<256>(reverse, P, D, A, Dr, S, nblocks, lds, s, rcClass);
        case 384: return launch_pass_t<384>(reverse, P, D, A, Dr, S, nblocks, lds, s, rcClass);
        case 512: return launch_pass_t<512>(reverse, P, D, A, Dr, S, nblocks, lds, s, rcClass);
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
