// CPU oracle for the batched Cosserat-rod substep. TEST INFRASTRUCTURE ONLY.
//
// This is a plain C++ restatement (templated on float / double, OpenMP over
// vertices) of the reference solver's DETERMINISTIC (snapshot / Jacobi) method:
// every sweep reads only the previous iterate and writes a new one, which is the
// only well-defined semantics of the reference's racy Gauss-Seidel kernels
// (reference: third_party/yarnball/KittenEngine/YarnBall/sim/cosserat.cu:216-223,
// 234-235, 252). It is the checker for the HIP path: only tests/, the smoke entry
// and the benchmark's CPU-baseline leg may load it, never the product.
//
// Each function cites the reference file:line it restates. Paths are relative to
// the reference repository root; YB = third_party/yarnball/KittenEngine/YarnBall,
// KE = third_party/yarnball/KittenEngine/KittenEngine/includes/modules.
//
// Floating-point conventions shared with the HIP kernels (so that the float32
// instantiation tracks the device closely): no FMA contraction (build with
// -ffp-contract=off), left-to-right sums, mix(a,b,t) = a*(1-t) + b*t,
// normalize(v) = v * (1/sqrt(dot(v,v))), 3x3 solve by adjugate / determinant.
// Third-party arithmetic the reference takes from GLM 0.9.9.8 (mat3 inverse, mix,
// normalize) is restated from its definition; the bit-level equality with GLM is
// not pinned by anything in the reference (parity unpinned at the ulp level; the
// physics checks in tests/ pin the behaviour).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- small math

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T len2(V3<T> a) { return dot(a, a); }
template <class T> inline T len(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) {
    return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
}
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) {
    return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
}
// GLM mix for floating types: x * (1 - a) + y * a.
template <class T> inline V3<T> mix(V3<T> a, V3<T> b, T t) { return a * (T(1) - t) + b * t; }
template <class T> inline T clampv(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }

// Quaternion in (x, y, z, w) order (KE/Rotor.h:11-19).
template <class T> struct Q4 {
    T x, y, z, w;
};
template <class T> inline V3<T> qv(Q4<T> q) { return {q.x, q.y, q.z}; }
template <class T> inline Q4<T> mkq(V3<T> v, T w) { return {v.x, v.y, v.z, w}; }
template <class T> inline Q4<T> qinv(Q4<T> q) { return {-q.x, -q.y, -q.z, q.w}; }  // Rotor.h:87-89
// Hamilton product (Rotor.h:185-188).
template <class T> inline Q4<T> qmul(Q4<T> a, Q4<T> b) {
    V3<T> v = (a.w * qv(b) + b.w * qv(a)) + cross(qv(a), qv(b));
    return mkq(v, a.w * b.w - dot(qv(a), qv(b)));
}
// Rotate a vector (Rotor.h:96-105).
template <class T> inline V3<T> qrot(Q4<T> q, V3<T> v) {
    V3<T> a = q.w * v + cross(qv(q), v);
    T c = dot(v, qv(q));
    return (q.w * a + cross(qv(q), a)) + c * qv(q);
}
template <class T> inline T qdot(Q4<T> a, Q4<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
template <class T> inline T qlen(Q4<T> a) { return std::sqrt(qdot(a, a)); }
template <class T> inline Q4<T> qscale(T s, Q4<T> a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
template <class T> inline Q4<T> qadd(Q4<T> a, Q4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline Q4<T> qnormalize(Q4<T> a) { return qscale(T(1) / std::sqrt(qdot(a, a)), a); }
// angleAxis (Rotor.h:23-26).
template <class T> inline Q4<T> qangle_axis(T rad, V3<T> axis) {
    rad = rad * T(0.5);
    return mkq(std::sin(rad) * axis, std::cos(rad));
}
// fromTo (Rotor.h:57-64) with orthoBasisX (Common.h:244-259) for the antipodal case.
template <class T> inline Q4<T> qfrom_to(V3<T> from, V3<T> to) {
    V3<T> h = (from + to) * T(0.5);
    T l = len2(h);
    if (l > 0) {
        h = h * (T(1) / std::sqrt(l));
    } else {
        V3<T> n = from;
        if (n.z >= n.y) {
            T a = T(1) / (T(1) + n.z);
            T b = -n.x * n.y * a;
            h = mk(b, T(1) - n.y * n.y * a, -n.y);
        } else {
            T a = T(1) / (T(1) + n.y);
            T b = -n.x * n.z * a;
            h = mk(T(1) - n.x * n.x * a, -n.x, b);
        }
    }
    return mkq(cross(from, h), dot(from, h));
}

// Packed symmetric 3x3: [xx, yy, zz, xy, xz, yz] (KE/SymMat.h:8-111).
template <class T> struct S3 {
    T m[6];
};
template <class T> inline S3<T> sdiag(T d) { return S3<T>{{d, d, d, 0, 0, 0}}; }
template <class T> inline S3<T> szero() { return S3<T>{{0, 0, 0, 0, 0, 0}}; }
template <class T> inline S3<T> souter(V3<T> v) {
    return S3<T>{{v.x * v.x, v.y * v.y, v.z * v.z, v.x * v.y, v.x * v.z, v.y * v.z}};
}
template <class T> inline void sadd(S3<T>& a, const S3<T>& b) {
    for (int i = 0; i < 6; i++) a.m[i] = a.m[i] + b.m[i];
}
template <class T> inline void ssub(S3<T>& a, const S3<T>& b) {
    for (int i = 0; i < 6; i++) a.m[i] = a.m[i] - b.m[i];
}
template <class T> inline S3<T> sscale(const S3<T>& a, T s) {
    S3<T> r;
    for (int i = 0; i < 6; i++) r.m[i] = a.m[i] * s;
    return r;
}
template <class T> inline void sadd_diag(S3<T>& a, T d) {
    a.m[0] = a.m[0] + d;
    a.m[1] = a.m[1] + d;
    a.m[2] = a.m[2] + d;
}
// H^-1 f via the adjugate (the reference calls GLM inverse(mat3), cosserat.cu:208).
template <class T> inline V3<T> ssolve(const S3<T>& H, V3<T> f) {
    const T a = H.m[0], b = H.m[1], c = H.m[2], d = H.m[3], e = H.m[4], g = H.m[5];
    const T c00 = b * c - g * g;
    const T c01 = e * g - d * c;
    const T c02 = d * g - b * e;
    const T c11 = a * c - e * e;
    const T c12 = d * e - a * g;
    const T c22 = a * b - d * d;
    const T det = (a * c00 + d * c01) + e * c02;
    const T inv = T(1) / det;
    return mk((c00 * f.x + c01 * f.y + c02 * f.z) * inv, (c01 * f.x + c11 * f.y + c12 * f.z) * inv,
              (c02 * f.x + c12 * f.y + c22 * f.z) * inv);
}

// Closest points of segments a0a1 / b0b1 as (u on a, v on b) (KE/Common.h:327-345).
template <class T> inline void seg_closest(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1, T& u, T& v) {
    V3<T> A0 = a1 - a0, A1 = b1 - b0;
    T a = len2(A0), b = dot(A0, A1), c = len2(A1);
    V3<T> diff = b0 - a0;
    T y = b * dot(A0, diff) - a * dot(A1, diff);
    y = y / (a * c - b * b);
    if (!std::isfinite(y)) y = T(0.5);
    y = clampv(y, T(0), T(1));
    T x = clampv(dot(mix(b0, b1, y) - a0, A0) / a, T(0), T(1));
    y = clampv(dot(mix(a0, a1, x) - b0, A1) / c, T(0), T(1));
    u = x;
    v = y;
}

// ---------------------------------------------------------------- constants

enum : uint32_t { HAS_PREV = 1, HAS_NEXT = 2, HAS_NEXT_ORIENT = 4, FIX_ORIENT = 8 };  // YB/YarnBall.h:52-58
constexpr int MAX_SDF = 64, MAX_CAND = 8;                                            // YB/YarnBall.h:153
constexpr int MAX_COLL = 128;                                                        // YB/YarnBall.h:38
enum { ERR_NONE = 0, ERR_MAX_COLLISIONS = 1, WARN_INTERPENETRATION = 2 };            // YB/YarnBall.h:204-208

template <class T> struct Grid {
    std::vector<float> v;  // C order, i slowest (the product stores float32 grids too)
    int nx, ny, nz;
    V3<T> origin, hiCorner;
    T h, invH;
    T muS, muK;
};
template <class T> struct Pose {  // YB/YarnBall.h:147
    V3<T> pos;
    Q4<T> q;
    V3<T> vel, omega, lo, hi;
};
template <class T> struct Pose7 {
    V3<T> pos;
    Q4<T> q;
};
template <class T> struct Motion {  // YB/YarnBall.h:152
    V3<T> p0;
    Q4<T> q0;
    V3<T> dp, rot;
};

// World AABB of a posed object-frame box (YB/sim/contact.cuh:79-87).
template <class T> inline void world_box(V3<T> lo_o, V3<T> hi_o, V3<T> pos, Q4<T> q, V3<T>& lo, V3<T>& hi) {
    const T inf = std::numeric_limits<T>::infinity();
    lo = mk(inf, inf, inf);
    hi = mk(-inf, -inf, -inf);
    for (int c = 0; c < 8; c++) {
        V3<T> corner = mk((c & 1) ? hi_o.x : lo_o.x, (c & 2) ? hi_o.y : lo_o.y, (c & 4) ? hi_o.z : lo_o.z);
        V3<T> w = qrot(q, corner) + pos;
        lo = vmin(lo, w);
        hi = vmax(hi, w);
    }
}

// Rotation vector log(q1 conj(q0)), shortest arc (csrc/sdf.cu:25-33).
template <class T> inline V3<T> rotvec(Q4<T> q0, Q4<T> q1) {
    V3<T> a = qv(q1), b = -qv(q0);
    V3<T> v = (q1.w * b + q0.w * a) + cross(a, b);
    T w = q1.w * q0.w - dot(a, b);
    if (w < 0) {
        v = -v;
        w = -w;
    }
    T s = len(v);
    if (!(s > T(1e-9))) return mk(T(0), T(0), T(0));
    return v * (T(2) * std::atan2(s, w) / s);
}

// Trilinear SDF sample and the exact gradient of the interpolant (YB/sim/contact.cuh:57-76).
template <class T> inline bool sdf_sample(const Grid<T>& g, V3<T> p, T& phi, V3<T>& grad) {
    V3<T> u = (p - g.origin) * g.invH;
    if (!(u.x >= 0 && u.y >= 0 && u.z >= 0 && u.x < T(g.nx - 1) && u.y < T(g.ny - 1) && u.z < T(g.nz - 1)))
        return false;
    const int i = (int)u.x, j = (int)u.y, k = (int)u.z;
    V3<T> fr = (p - (g.origin + mk(T(i), T(j), T(k)) * g.h)) * g.invH;
    const T fx = std::min(std::max(fr.x, T(0)), T(1));
    const T fy = std::min(std::max(fr.y, T(0)), T(1));
    const T fz = std::min(std::max(fr.z, T(0)), T(1));
    const size_t sy = (size_t)g.nz, sx = (size_t)g.ny * g.nz;
    const float* c = g.v.data() + (size_t)i * sx + (size_t)j * sy + k;
    const T c000 = c[0], c001 = c[1], c010 = c[sy], c011 = c[sy + 1];
    const T c100 = c[sx], c101 = c[sx + 1], c110 = c[sx + sy], c111 = c[sx + sy + 1];
    const T dx00 = c100 - c000, dx01 = c101 - c001, dx10 = c110 - c010, dx11 = c111 - c011;
    const T c00 = c000 + fx * dx00, c01 = c001 + fx * dx01, c10 = c010 + fx * dx10, c11 = c011 + fx * dx11;
    const T dy0 = c10 - c00, dy1 = c11 - c01;
    const T c0 = c00 + fy * dy0, c1 = c01 + fy * dy1;
    phi = c0 + fz * (c1 - c0);
    const T gx0 = dx00 + fy * (dx10 - dx00), gx1 = dx01 + fy * (dx11 - dx01);
    grad = mk(gx0 + fz * (gx1 - gx0), dy0 + fz * (dy1 - dy0), c1 - c0) * g.invH;
    return true;
}

// ---------------------------------------------------------------- the solver state

template <class T> struct Oracle {
    using V = V3<T>;
    using Q = Q4<T>;

    // layout: flat vertices, ropes as a CSR over them
    int nv = 0, nropes = 0;
    std::vector<int> roff;   // [nropes + 1]
    std::vector<int> ropeOf; // [nv]

    // per-vertex state (YB/YarnBall.h:70-101)
    std::vector<V> pos, vel, lastVel, dx, dxNext, lastPos, initPos;
    std::vector<Q> q, qPrev, qRest, initQ;
    std::vector<T> invMass, lRest, kStretch, pad, maxStep;
    std::vector<uint32_t> flags;
    std::vector<uint8_t> pinned;

    // global knobs (defaults: YB/YarnBall.cu:14-45)
    V gravity = mk(T(0), T(-9.8), T(0));
    int numItr = 8;
    T h = 0, lastH = 0, time = 0, maxH = T(1e-3);
    T damping = T(1e-6), drag = T(0.2), dragQuad = 0, dragQuadT = 0, accel = 1;
    // surface contact (bindings.cpp:51-80)
    int numPlanes = 0;
    T planes[8][4] = {};
    T cRadius = 0, cKappa = 2, cMuS = 0, cMuK = 0, cVSlip = T(1e-3);
    // rope-rope contact (batch.cu:367-385; YB/YarnBall.cu:33-43)
    bool ropeContact = false;
    int groupRopes = 1;
    T radius = T(1e-4), barrier = T(8e-4), kCollision = T(1e-5), detScaler = T(1.2), fricMu = T(0.1),
      kFriction = 5;
    T detR = 0, scaledDetR = 0;  // uploadMeta (YB/YarnBall.cu:261-272)
    T minSegLen = 0, maxSegLen = 0;
    std::vector<int> nCol, coll;  // [nv], [nv * MAX_COLL] row-major (ascending j)
    int errorCode = 0, lastError = 0, lastWarning = 0;

    // meshes (csrc/sdf.h:25-107), env = rope
    std::vector<Grid<T>> grids;
    std::vector<Pose7<T>> cur, tgt;  // [nropes * MAX_SDF]
    std::vector<uint8_t> tele;
    std::vector<Motion<T>> motion;
    std::vector<Pose<T>> poses;
    bool pendingSet = false, moving = false, sdfMoving = false, meshesTouched = false;
    std::vector<int> cand, nCand;  // [nv * MAX_CAND], [nv]
    bool forcesOn = false;
    std::vector<V> cForce;   // [nv]
    std::vector<T> wrench;   // [nropes * MAX_SDF * 6]
    int lastSubsteps = 1;

    // staged root program
    std::vector<T> P0, V0, P1, Q0, Q1;
    bool hasQ0 = false, hasQ1 = false;

    int ompThreshold = 4096;

    // ------------------------------------------------------------ construction
    // RopeBatch ctor + Sim::configure + setKStretch / setKBend
    // (csrc/batch.cu:193-276; YB/YarnBall.cu:112-259).
    void build(int nropes_, const int* offsets, const float* xyz, const uint8_t* pin, T density, T EA, T EI,
               T max_h, int num_itr, bool fixRootOrientation, bool straightRest) {
        nropes = nropes_;
        roff.assign(offsets, offsets + nropes + 1);
        nv = roff[nropes];
        ropeOf.resize(nv);
        for (int r = 0; r < nropes; r++)
            for (int i = roff[r]; i < roff[r + 1]; i++) ropeOf[i] = r;
        pos.resize(nv);
        vel.assign(nv, mk(T(0), T(0), T(0)));
        lastVel = vel;
        dx = vel;
        dxNext = vel;
        lastPos = vel;
        q.resize(nv);
        qPrev.resize(nv);
        qRest.assign(nv, Q{0, 0, 0, 5});  // YB/YarnBall.cu:55
        invMass.assign(nv, T(1));
        lRest.assign(nv, T(1));
        kStretch.assign(nv, T(100));
        pad.assign(nv, 0);
        maxStep.assign(nv, std::numeric_limits<T>::infinity());
        flags.assign(nv, HAS_NEXT);
        pinned.assign(pin, pin + nv);
        for (int i = 0; i < nv; i++) pos[i] = mk(T(xyz[3 * i]), T(xyz[3 * i + 1]), T(xyz[3 * i + 2]));
        for (int r = 0; r < nropes; r++) {
            flags[roff[r + 1] - 1] = 0;                          // batch.cu:215
            if (fixRootOrientation) flags[roff[r]] |= FIX_ORIENT;  // batch.cu:222-223
        }
        for (int i = 0; i < nv; i++)
            if (pinned[i]) invMass[i] = 0;  // batch.cu:218-220
        // dead collision radii scale with the shortest given segment (batch.cu:230-238)
        T minSeg = T(1e30);
        for (int r = 0; r < nropes; r++)
            for (int i = roff[r]; i + 1 < roff[r + 1]; i++) minSeg = std::min(minSeg, len(pos[i + 1] - pos[i]));
        radius = T(0.01) * minSeg;
        barrier = T(0.08) * minSeg;

        // Sim::configure (YB/YarnBall.cu:112-182)
        maxSegLen = 0;
        minSegLen = std::numeric_limits<T>::max();
        Q lastQ{0, 0, 0, 1};
        for (int i = 0; i < nv; i++) {
            if (i < nv - 1) {
                bool hn = (flags[i] & HAS_NEXT) != 0;
                flags[i + 1] = hn ? (flags[i + 1] | HAS_PREV) : (flags[i + 1] & ~HAS_PREV);
                bool hnn = hn && (flags[i + 1] & HAS_NEXT);
                flags[i] = hnn ? (flags[i] | HAS_NEXT_ORIENT) : (flags[i] & ~HAS_NEXT_ORIENT);
                if (!hn) flags[i] |= FIX_ORIENT;
            }
            if (!(flags[i] & HAS_PREV) && !(flags[i] & HAS_NEXT)) throw std::runtime_error("Dangling segment");
            lRest[i] = T(1) / T(nv);
            T mass = 0;
            if (flags[i] & HAS_PREV) mass = mass + lRest[i - 1];
            if (flags[i] & HAS_NEXT) {
                V seg0 = pos[i + 1] - pos[i];
                lRest[i] = len(seg0);
                if (lRest[i] == 0 || !std::isfinite(lRest[i])) throw std::runtime_error("0 length segment");
                V sd = seg0 * (T(1) / std::sqrt(dot(seg0, seg0)));
                Q qq = (sd.x < T(-0.999999)) ? Q{0, 0, 1, 0} : qfrom_to(mk(T(1), T(0), T(0)), sd);
                Q t = qmul(qinv(qq), lastQ);
                Q tw{t.x, 0, 0, t.w};
                t = (qdot(tw, tw) > T(1e-12)) ? qnormalize(tw) : Q{0, 0, 0, 1};
                lastQ = qmul(qq, t);
                mass = mass + lRest[i];
                maxSegLen = std::max(maxSegLen, lRest[i]);
                minSegLen = std::min(minSegLen, lRest[i]);
            }
            q[i] = lastQ;
            mass = mass * (T(0.5) * density);
            if (mass != 0)
                invMass[i] = invMass[i] * (T(1) / mass);
            else
                invMass[i] = 0;
        }
        for (int i = 0; i < nv - 1; i++) {
            Q rel = qmul(qinv(q[i]), q[i + 1]);
            qRest[i] = qscale(qlen(qRest[i]), rel);  // YB/YarnBall.cu:181-182
        }
        if (straightRest)  // batch.cu:245-247
            for (int r = 0; r < nropes; r++)
                for (int i = roff[r]; i + 1 < roff[r + 1]; i++) qRest[i] = Q{0, 0, 0, 1};
        for (int i = 0; i < nv; i++) kStretch[i] = EA * lRest[i];  // YB/YarnBall.cu:237-245
        const T kb = EI * T(4);                                      // YB/YarnBall.cu:247-259
        for (int i = 0; i < nv; i++) qRest[i] = qscale(kb / lRest[i], qnormalize(qRest[i]));
        maxH = max_h;
        h = max_h;  // batch.cu:251
        numItr = num_itr;
        initPos = pos;
        initQ = q;
        P0.assign(3 * nropes, 0);
        V0 = P0;
        P1 = P0;
        Q0.assign(4 * nropes, 0);
        Q1 = Q0;
        nCol.assign(nv, 0);
        coll.assign((size_t)nv * MAX_COLL, 0);
    }

    // Per-vertex setters (bindings.cpp:1121-1201): masses [nv] (pinned entries ignored), EA / EI per segment
    // [nv] (entries at a rope's last vertex ignored).
    void set_masses(const T* m) {
        for (int i = 0; i < nv; i++)
            if (invMass[i] != 0) invMass[i] = T(1) / m[i];
    }
    void set_stretch(const T* ea) {
        for (int r = 0; r < nropes; r++)
            for (int i = roff[r]; i + 1 < roff[r + 1]; i++) kStretch[i] = ea[i] * lRest[i];
    }
    void set_bend(const T* ei) {
        for (int r = 0; r < nropes; r++)
            for (int i = roff[r]; i + 1 < roff[r + 1]; i++) {
                T k = T(4) * ei[i] / lRest[i];
                qRest[i] = qscale(k, qnormalize(qRest[i]));
            }
    }

    // ------------------------------------------------------------ meshes (csrc/sdf.cu)
    void ensure_mesh_state() {
        if (meshesTouched) return;
        meshesTouched = true;
        const size_t total = (size_t)nropes * MAX_SDF;
        Pose7<T> ident{mk(T(0), T(0), T(0)), Q{0, 0, 0, 1}};
        cur.assign(total, ident);
        tgt.assign(total, ident);
        tele.assign(total, 0);
        motion.assign(total, Motion<T>{});
        poses.assign(total, Pose<T>{});
        cand.assign((size_t)nv * MAX_CAND, 0);
        nCand.assign(nv, -1);  // sdf.cu:130
        cForce.assign(nv, mk(T(0), T(0), T(0)));
        wrench.assign(total * 6, 0);
    }
    // SdfStore::add (sdf.cu:137-158) + commitMeshK (sdf.cu:183-194).
    int add_mesh(const float* values, int nx, int ny, int nz, const T* origin, T spacing, T muS, T muK,
                 const T* p, const T* qq, int rows) {
        ensure_mesh_state();
        if ((int)grids.size() >= MAX_SDF) throw std::invalid_argument("at most 64 meshes");
        Grid<T> g;
        g.v.assign(values, values + (size_t)nx * ny * nz);
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.origin = mk(origin[0], origin[1], origin[2]);
        g.h = spacing;
        g.invH = T(1) / spacing;
        g.muS = muS;
        g.muK = muK;
        g.hiCorner = g.origin + mk(T(nx - 1), T(ny - 1), T(nz - 1)) * spacing;
        grids.push_back(g);
        const int idx = (int)grids.size() - 1;
        const bool pending = pendingSet;
        set_targets(idx, p, qq, rows, nullptr, true);
        for (int e = 0; e < nropes; e++) commit_slot((size_t)e * MAX_SDF + idx, idx);
        pendingSet = pending;
        return idx;
    }
    // scatterTargetsK (sdf.cu:37-51): index < 0 -> dense [nropes, count] rows.
    void set_targets(int index, const T* p, const T* qq, int rows, const uint8_t* teleFlags, bool teleAll) {
        ensure_mesh_state();
        const int M = (int)grids.size();
        const int per = index < 0 ? M : 1;
        for (int i = 0; i < nropes * per; i++) {
            const int e = i / per, m = index < 0 ? i % per : index;
            const int r = index < 0 ? i : (rows == 1 ? 0 : e);
            Q t{qq[4 * r], qq[4 * r + 1], qq[4 * r + 2], qq[4 * r + 3]};
            const T l = qlen(t);
            t = l > 0 ? Q{t.x / l, t.y / l, t.z / l, t.w / l} : Q{0, 0, 0, 1};
            const size_t slot = (size_t)e * MAX_SDF + m;
            tgt[slot].pos = mk(p[3 * r], p[3 * r + 1], p[3 * r + 2]);
            tgt[slot].q = t;
            tele[slot] = teleAll ? 1 : (teleFlags ? teleFlags[e] : 0);
        }
        pendingSet = true;
    }
    void commit_slot(size_t i, int m) {  // sdf.cu:67-75
        cur[i] = tgt[i];
        tele[i] = 0;
        Pose<T> P;
        P.pos = cur[i].pos;
        P.q = cur[i].q;
        P.vel = mk(T(0), T(0), T(0));
        P.omega = P.vel;
        world_box(grids[m].origin, grids[m].hiCorner, P.pos, P.q, P.lo, P.hi);
        poses[i] = P;
    }
    void clear_meshes() {
        grids.clear();
        pendingSet = false;
        moving = false;
    }
    // SdfStore::begin_step (sdf.cu:217-232).
    void begin_step() {
        if (!meshesTouched) return;
        if (pendingSet) {
            for (size_t i = 0; i < cur.size(); i++) {  // stageMotionK (sdf.cu:55-63)
                if (tele[i]) {
                    cur[i] = tgt[i];
                    tele[i] = 0;
                }
                motion[i].p0 = cur[i].pos;
                motion[i].q0 = cur[i].q;
                motion[i].dp = tgt[i].pos - cur[i].pos;
                motion[i].rot = rotvec(cur[i].q, tgt[i].q);
            }
            sdfMoving = true;
            moving = true;
        } else {
            sdfMoving = false;
        }
        if (!grids.empty()) std::fill(nCand.begin(), nCand.end(), -1);
        if (forcesOn) {
            std::fill(cForce.begin(), cForce.end(), mk(T(0), T(0), T(0)));
            std::fill(wrench.begin(), wrench.end(), T(0));
        }
    }
    // SdfStore::end_step (sdf.cu:239-253).
    void end_step(int substeps) {
        lastSubsteps = substeps > 0 ? substeps : 1;
        if (!meshesTouched) return;
        if (moving) {
            const int M = (int)grids.size();
            for (int e = 0; e < nropes; e++)
                for (int m = 0; m < M; m++) commit_slot((size_t)e * MAX_SDF + m, m);
            moving = false;
            pendingSet = false;
        }
        if (forcesOn) {
            const T k = T(1) / T(lastSubsteps);
            for (auto& f : cForce) f = mk(f.x * k, f.y * k, f.z * k);
            for (auto& w : wrench) w = w * k;
        }
    }
    // Per-substep kinematic pose (YB/sim/iteration.cu:12-33).
    void substep_poses(int sub, int nSub) {
        if (!sdfMoving) return;
        const int M = (int)grids.size();
        for (int e = 0; e < nropes; e++)
            for (int m = 0; m < M; m++) {
                const size_t s = (size_t)e * MAX_SDF + m;
                const Motion<T>& Mo = motion[s];
                const T a = T(sub + 1) / T(nSub);
                const T invDt = T(1) / (T(nSub) * h);
                Pose<T> P;
                const T th = len(Mo.rot);
                if (th > 0 || Mo.dp.x != 0 || Mo.dp.y != 0 || Mo.dp.z != 0) {
                    P.pos = Mo.p0 + a * Mo.dp;
                    Q dq = th > 0 ? qangle_axis(a * th, Mo.rot / th) : Q{0, 0, 0, 1};
                    P.q = qnormalize(qmul(dq, Mo.q0));
                    P.vel = Mo.dp * invDt;
                    P.omega = Mo.rot * invDt;
                } else {
                    P.pos = Mo.p0;
                    P.q = Mo.q0;
                    P.vel = mk(T(0), T(0), T(0));
                    P.omega = P.vel;
                }
                world_box(grids[m].origin, grids[m].hiCorner, P.pos, P.q, P.lo, P.hi);
                poses[s] = P;
            }
    }

    // ------------------------------------------------------------ surface contact (YB/sim/contact.cuh)
    // surfaceContact (contact.cuh:27-46).
    void surface_contact(V n, T gap, V d, T k, T muS, T muK, V& f, S3<T>& H) const {
        const T eps = cVSlip;
        const T N = -k * gap;
        f = f + N * n;
        S3<T> nn = souter(n);
        sadd(H, sscale(nn, k));
        const V ut = d - dot(n, d) * n;
        const T ul = len(ut);
        const T vt = ul / h;
        T kt;
        if (vt <= eps) {
            kt = muS * N / (eps * h);
        } else {
            const T s = T(1) - std::min(T(1), vt / eps - T(1));
            kt = (s * muS + (T(1) - s) * muK) * N / ul;
        }
        f = f - kt * ut;
        sadd_diag(nn, T(-1));
        ssub(H, sscale(nn, kt));
    }
    // sdfContactAt (contact.cuh:94-110).
    bool sdf_contact(const Pose<T>& P, int m, V p, V d, V& n, T& gap, V& dRel) const {
        T phi;
        V g;
        if (!sdf_sample(grids[m], qrot(qinv(P.q), (p - P.pos) + d), phi, g)) return false;
        gap = phi - cRadius;
        if (gap >= 0) return false;
        const T gl = len(g);
        if (!(gl > 0)) return false;
        n = qrot(P.q, g / gl);
        dRel = d;
        if (sdfMoving) dRel = dRel - h * (P.vel + cross(P.omega, p - P.pos));
        return true;
    }
    int cand_count(int i) const { return nCand[i] < 0 ? (int)grids.size() : nCand[i]; }
    int cand_at(int i, int c) const { return nCand[i] < 0 ? c : cand[(size_t)i * MAX_CAND + c]; }
    // surfacesAt (contact.cuh:141-163).
    void surfaces_at(int i, V p, V d, T k, V& f, S3<T>& H) const {
        for (int pl = 0; pl < numPlanes; pl++) {
            const V n = mk(planes[pl][0], planes[pl][1], planes[pl][2]);
            const T gap = ((dot(n, p) - planes[pl][3]) + dot(n, d)) - cRadius;
            if (gap >= 0) continue;
            surface_contact(n, gap, d, k, cMuS, cMuK, f, H);
        }
        if (!grids.empty()) {
            const int env = ropeOf[i];
            const int cnt = cand_count(i);
            for (int c = 0; c < cnt; c++) {
                const int m = cand_at(i, c);
                const Pose<T>& P = poses[(size_t)env * MAX_SDF + m];
                V n, dRel;
                T gap;
                if (!sdf_contact(P, m, p, d, n, gap, dRel)) continue;
                surface_contact(n, gap, dRel, k, grids[m].muS, grids[m].muK, f, H);
            }
        }
    }
    // surfaceReadout (contact.cuh:172-204), serial so the wrench sums are deterministic here.
    void surface_readout(int i, V p, V d, T k) {
        S3<T> H = szero<T>();
        V fv = mk(T(0), T(0), T(0));
        for (int pl = 0; pl < numPlanes; pl++) {
            const V n = mk(planes[pl][0], planes[pl][1], planes[pl][2]);
            const T gap = ((dot(n, p) - planes[pl][3]) + dot(n, d)) - cRadius;
            if (gap >= 0) continue;
            surface_contact(n, gap, d, k, cMuS, cMuK, fv, H);
        }
        if (!grids.empty()) {
            const int env = ropeOf[i];
            const int cnt = cand_count(i);
            for (int c = 0; c < cnt; c++) {
                const int m = cand_at(i, c);
                const Pose<T>& P = poses[(size_t)env * MAX_SDF + m];
                V n, dRel;
                T gap;
                if (!sdf_contact(P, m, p, d, n, gap, dRel)) continue;
                V fm = mk(T(0), T(0), T(0));
                surface_contact(n, gap, dRel, k, grids[m].muS, grids[m].muK, fm, H);
                fv = fv + fm;
                const V r = ((p - P.pos) + d) - cRadius * n;
                const V tq = cross(r, -fm);
                T* w = &wrench[((size_t)env * MAX_SDF + m) * 6];
                w[0] += -fm.x;
                w[1] += -fm.y;
                w[2] += -fm.z;
                w[3] += tq.x;
                w[4] += tq.y;
                w[5] += tq.z;
            }
        }
        cForce[i] = cForce[i] + fv;
    }
    // sdfBuildCandidates (contact.cuh:121-136).
    void build_candidates(int i, V p, T padv) {
        const int env = ropeOf[i];
        int n = 0;
        const int M = (int)grids.size();
        for (int m = 0; m < M; m++) {
            const Pose<T>& P = poses[(size_t)env * MAX_SDF + m];
            const T pm = padv + (sdfMoving ? h * (len(P.vel) + len(P.omega) * T(0.5) * len(P.hi - P.lo)) : T(0));
            if (p.x >= P.lo.x - pm && p.x <= P.hi.x + pm && p.y >= P.lo.y - pm && p.y <= P.hi.y + pm &&
                p.z >= P.lo.z - pm && p.z <= P.hi.z + pm) {
                if (n < MAX_CAND) cand[(size_t)i * MAX_CAND + n] = m;
                n++;
            }
        }
        nCand[i] = n <= MAX_CAND ? n : -1;
    }

    // ------------------------------------------------------------ rope-rope detection
    int group_of(int i) const { return ropeOf[i] / groupRopes; }
    int group_begin(int g) const { return roff[std::min(g * groupRopes, nropes)]; }
    int group_end(int g) const { return roff[std::min((g + 1) * groupRopes, nropes)]; }
    // bruteDetectK (csrc/batch.cu:19-49) fused with recomputeStepLimitKernel (YB/sim/collision.cu:103-136):
    // the step limit recomputes exactly the closest-point vector the detection formed for each listed pair.
    void detect(int i, bool limit) {
        nCol[i] = 0;
        T minDist = std::numeric_limits<T>::infinity();
        if (!(flags[i] & HAS_NEXT)) {
            maxStep[i] = minDist;
            return;
        }
        const int g = group_of(i);
        const int g0 = group_begin(g), g1 = group_end(g);
        const V a0 = lastPos[i];
        const V a1 = lastPos[i + 1] - a0;
        const T ri = std::max(pad[i], pad[i + 1]);
        minDist = pad[i] - detR;
        int cnt = 0;
        for (int j = g0; j < g1; j++) {
            if (!(flags[j] & HAS_NEXT)) continue;
            if (ropeOf[j] == ropeOf[i] && std::abs(j - i) <= 2) continue;
            const V b0 = lastPos[j] - a0, b1 = lastPos[j + 1] - a0;
            T u, v;
            seg_closest(mk(T(0), T(0), T(0)), a1, b0, b1, u, v);
            const V d = u * a1 - mix(b0, b1, v);
            const T r = ri + std::max(pad[j], pad[j + 1]);
            if (len2(d) < r * r) {
                if (cnt < MAX_COLL) {
                    coll[(size_t)i * MAX_COLL + cnt] = j;
                    minDist = std::min(minDist, T(0.45) * len(d));
                }
                cnt++;
            }
        }
        if (cnt > MAX_COLL) errorCode = ERR_MAX_COLLISIONS;
        nCol[i] = std::min(cnt, MAX_COLL);
        maxStep[i] = minDist;
        (void)limit;
    }

    // ------------------------------------------------------------ the substep kernels
    // initItr (YB/sim/iteration.cu:34-69): the inertial target y is stored in the velocity slot.
    void init_vertex(int i) {
        const V g = gravity;
        const V v = vel[i];
        V d = h * v;
        const V lv = lastVel[i];
        lastVel[i] = v;
        T padding = scaledDetR;
        const V y = d + (h * h) * g;
        padding = padding + len(y);
        if (invMass[i] != 0) {
            vel[i] = y;
            const T g2 = len2(g);
            if (g2 > 0) {
                const V a = (v - lv) / lastH;
                const T s = clampv(dot(a, g) / g2, T(0), T(1));
                d = d + (h * h * s) * g;
            }
        }
        pad[i] = padding;
        dx[i] = d;
        lastPos[i] = pos[i];
    }

    // cosseratItr<LIMIT, SNAPSHOT=true> (YB/sim/cosserat.cu:23-225). Returns the segment (i, i+1) terms
    // that act on vertex i+1 through f2 / H2 and the vertex's own f / H.
    void segment_terms(int i, V& f, S3<T>& H, V& f2, S3<T>& H2) const {
        const T dmp = damping / h;
        const V p0 = lastPos[i];
        const V d0 = dx[i];
        f = mk(T(0), T(0), T(0));
        H = szero<T>();
        f2 = f;
        H2 = H;
        if (invMass[i] != 0) {
            H = sdiag(T(1) / (invMass[i] * h * h));
            f = (T(1) / (h * h * invMass[i])) * (vel[i] - d0);
            if (numPlanes > 0 || !grids.empty()) {
                const T kap = cKappa / h;
                surfaces_at(i, p0, d0, kap * kap / invMass[i], f, H);
            }
        }
        if (flags[i] & HAS_NEXT) {
            const V p1 = lastPos[i + 1], p1dx = dx[i + 1];
            const T invl = T(1) / lRest[i];
            const V c = ((p1 - p0) + (p1dx - d0)) * invl - qrot(q[i], mk(T(1), T(0), T(0)));
            const T k = kStretch[i] * invl;
            T d = k * invl;
            f = f + (k * c - (dmp * d) * d0);
            f2 = f2 + (-k * c - (dmp * d) * p1dx);
            d = d * (T(1) + dmp);
            sadd_diag(H, d);
            sadd_diag(H2, d);
            if (ropeContact) {
                const T invb = T(1) / barrier;
                const T rad2 = T(2) * radius;
                const T kCol = kCollision * invb;
                for (int c2 = 0; c2 < nCol[i]; c2++) {
                    const int j = coll[(size_t)i * MAX_COLL + c2];
                    const V b0 = lastPos[j], b1 = lastPos[j + 1], db0 = dx[j], db1 = dx[j + 1];
                    T u, v;
                    seg_closest(mk(T(0), T(0), T(0)), (p1 - p0) + (p1dx - d0), (b0 - p0) + (db0 - d0),
                                (b1 - p0) + (db1 - d0), u, v);
                    const V dpos = mix(p0, p1, u) - mix(b0, b1, v);
                    const V ddpos = mix(d0, p1dx, u) - mix(db0, db1, v);
                    V normal = dpos + ddpos;
                    T dd = len(normal);
                    if (!(dd > 0)) continue;
                    normal = normal * (T(1) / dd);
                    const T wy = u;
                    const T wx = T(1) - u;
                    dd = dd - rad2;
                    dd = dd * invb;
                    if (dd > 1) continue;
                    dd = std::max(dd, T(1e-3));
                    const T invd = T(1) / dd;
                    const T logd = std::log(dd);
                    T dH = ((T(-3) + (T(2) + invd) * invd) - T(2) * logd) * kCol * invb;
                    const T ff = -(T(1) - dd) * ((dd - T(1)) + T(2) * dd * logd) * invd * kCol;
                    f = f + (ff * wx - dmp * dH * wx * wx * dot(normal, d0)) * normal;
                    f2 = f2 + (ff * wy - dmp * dH * wy * wy * dot(normal, p1dx)) * normal;
                    dH = dH * (T(1) + dmp);
                    S3<T> op = souter(normal);
                    sadd(H, sscale(op, dH * wx * wx));
                    sadd(H2, sscale(op, dH * wy * wy));
                    const V uu = ddpos - dot(normal, ddpos) * normal;
                    const T ul = len(uu);
                    if (ul > 0) {
                        const T f1 = std::min(kFriction, fricMu * ff / ul);
                        sadd_diag(op, T(-1));
                        f = f - f1 * wx * uu;
                        ssub(H, sscale(op, wx * wx * f1));
                        f2 = f2 - f1 * wy * uu;
                        ssub(H2, sscale(op, wy * wy * f1));
                    }
                }
            }
        }
    }
    void position_sweep() {
        const bool limit = ropeContact;
        std::vector<V> F(nv), F2(nv);
        std::vector<S3<T>> HH(nv), HH2(nv);
#pragma omp parallel for schedule(static) if (nv > ompThreshold)
        for (int i = 0; i < nv; i++) segment_terms(i, F[i], HH[i], F2[i], HH2[i]);
#pragma omp parallel for schedule(static) if (nv > ompThreshold)
        for (int i = 0; i < nv; i++) {
            V f = F[i];
            S3<T> H = HH[i];
            T stepLimit = std::numeric_limits<T>::infinity();
            if (limit) stepLimit = maxStep[i];
            if (flags[i] & HAS_PREV) {
                if (limit) stepLimit = std::min(stepLimit, maxStep[i - 1]);
                f = f + F2[i - 1];
                sadd(H, HH2[i - 1]);
            }
            V d = dx[i];
            if (invMass[i] != 0) {
                d = d + accel * ssolve(H, f);
                if (limit) {
                    const T l = len(d);
                    if (l > stepLimit && l > 0) d = d * (stepLimit / l);
                }
            }
            dxNext[i] = d;
            qPrev[i] = q[i];
        }
    }
    // quaternionLambdaItr<SNAPSHOT=true> (YB/sim/cosserat.cu:227-273).
    void orientation_sweep() {
#pragma omp parallel for schedule(static) if (nv > ompThreshold)
        for (int i = 0; i < nv; i++) {
            dx[i] = dxNext[i];
            const uint32_t fl = flags[i];
            if (!(fl & FIX_ORIENT) && (fl & HAS_NEXT)) {
                const V d = dxNext[i];
                const V p1 = pos[i + 1], p1dx = dxNext[i + 1];
                V e = ((p1 - pos[i]) + (p1dx - d)) / lRest[i];
                e = e * (T(-2) * kStretch[i]);
                Q b{0, 0, 0, 0};
                const Q q0 = qPrev[i];
                if (fl & HAS_PREV) {
                    const Q qr = qRest[i - 1];
                    const Q qq = qPrev[i - 1];
                    const T s = qdot(qmul(qinv(qq), q0), qr) > 0 ? T(1) : T(-1);
                    b = qadd(b, qscale(s, qmul(qq, qr)));
                }
                if (fl & HAS_NEXT_ORIENT) {
                    const Q qr = qRest[i];
                    const Q qq = qPrev[i + 1];
                    const T s = qdot(qmul(qinv(q0), qq), qr) > 0 ? T(1) : T(-1);
                    b = qadd(b, qscale(s, qmul(qq, qinv(qr))));
                }
                const T lambda = len(e) + qlen(b);
                const Q ex{1, 0, 0, 0};  // Kit::Rotor(1) is e_x (KE/Rotor.h:71)
                const Q prod = qmul(qmul(mkq(e, T(0)), b), ex);
                q[i] = qnormalize(qadd(prod, qscale(lambda, b)));
            }
        }
    }
    // endItr (YB/sim/iteration.cu:77-129).
    void end_vertex(int i) {
        const T invH = T(1) / h;
        const V d = dx[i];
        if (invMass[i] != 0) {
            V v = d * invH;
            const T kn = dragQuad;
            if (kn > 0) {
                const V t = (flags[i] & HAS_NEXT) ? qrot(q[i], mk(T(1), T(0), T(0))) : qrot(q[i - 1], mk(T(1), T(0), T(0)));
                V vt = t * dot(t, v);
                V vn = v - vt;
                vn = vn / (T(1) + kn * len(vn) * h);
                vt = vt / (T(1) + dragQuadT * len(vt) * h);
                v = vn + vt;
            }
            vel[i] = v * (T(1) - drag * h);
            if (forcesOn && (numPlanes > 0 || !grids.empty())) {
                const T kap = cKappa / h;
                surface_readout(i, pos[i], d, kap * kap / invMass[i]);
            }
        }
        pos[i] = pos[i] + d;
        if (!grids.empty()) {
            if (invMass[i] != 0)
                build_candidates(i, pos[i], T(2) * len(d) + T(0.02));
            else
                nCand[i] = 0;
        }
    }
    // applyRootProgramK + slerp4 (csrc/batch.cu:106-159).
    static Q slerp4(Q a, Q b, T t) {
        T d = qdot(a, b);
        if (d < 0) {
            b = Q{-b.x, -b.y, -b.z, -b.w};
            d = -d;
        }
        if (d > T(0.9995)) {
            Q m{a.x * (T(1) - t) + b.x * t, a.y * (T(1) - t) + b.y * t, a.z * (T(1) - t) + b.z * t,
                a.w * (T(1) - t) + b.w * t};
            return qnormalize(m);
        }
        const T th = std::acos(std::min(d, T(1)));
        const T s = std::sin(th);
        return qadd(qscale(std::sin((T(1) - t) * th) / s, a), qscale(std::sin(t * th) / s, b));
    }
    static Q unit_or_identity(Q qq) {
        const T l2 = qdot(qq, qq);
        return l2 > 0 ? qscale(T(1) / std::sqrt(l2), qq) : Q{0, 0, 0, 1};
    }
    void root_program(int s, int nSub, T dt) {
        for (int e = 0; e < nropes; e++) {
            const int root = roff[e];
            const V p0 = mk(P0[3 * e], P0[3 * e + 1], P0[3 * e + 2]);
            const V v0 = mk(V0[3 * e], V0[3 * e + 1], V0[3 * e + 2]);
            const V p1 = mk(P1[3 * e], P1[3 * e + 1], P1[3 * e + 2]);
            const V A = (T(2) / (dt * dt)) * ((p1 - p0) - v0 * dt);
            const T tau = T(s) * h;
            pos[root] = (p0 + v0 * tau) + (T(0.5) * tau * tau) * A;
            vel[root] = v0 + (tau + T(0.5) * h) * A;
            if (hasQ0) {
                const Q qa = unit_or_identity(Q{Q0[4 * e], Q0[4 * e + 1], Q0[4 * e + 2], Q0[4 * e + 3]});
                Q r;
                if (hasQ1) {
                    const Q qb = unit_or_identity(Q{Q1[4 * e], Q1[4 * e + 1], Q1[4 * e + 2], Q1[4 * e + 3]});
                    r = slerp4(qa, qb, T(s + 1) / T(nSub));
                } else {
                    r = qa;
                }
                q[root] = unit_or_identity(r);
            }
        }
    }

    // uploadMeta validation (YB/YarnBall.cu:261-272).
    void prepare(T hsub) {
        lastH = h;
        h = hsub;
        detR = radius + T(0.5) * barrier;
        scaledDetR = detR * detScaler;
        if (detScaler <= 1) throw std::runtime_error("Detection scaler must be greater than 1.");
        if (T(2) * minSegLen < T(2) * radius + barrier)
            throw std::runtime_error("Use thinner yarn or use longer segments.");
    }
    void substep(int s, int nSub) {
        substep_poses(s, nSub);
#pragma omp parallel for schedule(static) if (nv > ompThreshold)
        for (int i = 0; i < nv; i++) init_vertex(i);
        if (ropeContact) {
#pragma omp parallel for schedule(dynamic, 64) if (nv > 256)
            for (int i = 0; i < nv; i++) detect(i, true);
        }
        for (int it = 0; it < numItr; it++) {
            position_sweep();
            orientation_sweep();
        }
        if (forcesOn) {  // serial: deterministic wrench sums
            for (int i = 0; i < nv; i++) end_vertex(i);
        } else {
#pragma omp parallel for schedule(static) if (nv > ompThreshold)
            for (int i = 0; i < nv; i++) end_vertex(i);
        }
    }
    void finish(T dt) {
        time = time + dt;
        if (errorCode == ERR_MAX_COLLISIONS) {
            lastError = errorCode;
            errorCode = 0;
            throw std::runtime_error("MAX_COLLISIONS_PER_SEGMENT exceeded");
        }
    }
    // RopeBatch::step_program (csrc/batch.cu:321-363); q0 / q1 may be null.
    void step_program(const T* p0, const T* v0, const T* p1, const T* q0, const T* q1, T dt, int S) {
        std::copy(p0, p0 + 3 * nropes, P0.begin());
        std::copy(v0, v0 + 3 * nropes, V0.begin());
        std::copy(p1, p1 + 3 * nropes, P1.begin());
        hasQ0 = q0 != nullptr;
        hasQ1 = q1 != nullptr;
        if (q0) std::copy(q0, q0 + 4 * nropes, Q0.begin());
        if (q1) std::copy(q1, q1 + 4 * nropes, Q1.begin());
        const T hs = dt / T(S);
        begin_step();
        prepare(hs);
        for (int s = 0; s < S; s++) {
            root_program(s, S, dt);
            substep(s, S);
        }
        finish(dt);
        end_step(S);
    }
    // RopeBatch::step, constant-velocity sweep (csrc/batch.cu:288-319, applyRootsK :90-104).
    void step_const(const T* rp, const T* rv, const T* rq, T dt, int S) {
        for (int e = 0; e < nropes; e++) {
            const int root = roff[e];
            pos[root] = mk(rp[3 * e], rp[3 * e + 1], rp[3 * e + 2]);
            vel[root] = mk(rv[3 * e], rv[3 * e + 1], rv[3 * e + 2]);
            if (rq) q[root] = unit_or_identity(Q{rq[4 * e], rq[4 * e + 1], rq[4 * e + 2], rq[4 * e + 3]});
        }
        const T hs = dt / T(S);
        begin_step();
        prepare(hs);
        for (int s = 0; s < S; s++) substep(s, S);
        finish(dt);
        end_step(S);
    }
    // RopeBatch::step_free (csrc/batch.cu:388-398).
    void step_free(T dt, int S) {
        const T hs = dt / T(S);
        begin_step();
        prepare(hs);
        for (int s = 0; s < S; s++) substep(s, S);
        finish(dt);
        end_step(S);
    }
    // RopeBatch::reset_envs (csrc/batch.cu:178-189, 400-413).
    void reset_envs(const int* ids, int count) {
        for (int c = 0; c < count; c++) {
            const int e = ids[c];
            for (int i = roff[e]; i < roff[e + 1]; i++) {
                pos[i] = initPos[i];
                q[i] = initQ[i];
                vel[i] = mk(T(0), T(0), T(0));
                lastVel[i] = vel[i];
            }
        }
    }
    // set_rope_contact (csrc/batch.cu:367-385).
    void set_rope_contact(bool enabled, T rad, T bar, T kc, T mu, T kf, T scaler, int gr) {
        if (enabled) {
            if (T(2) * minSegLen < T(2) * rad + bar)
                throw std::invalid_argument("rope contact needs 2 * radius + barrier <= 2 * shortest segment");
            groupRopes = gr > 0 ? gr : 1;
            radius = rad;
            barrier = bar;
            kCollision = kc;
            fricMu = mu;
            kFriction = kf;
            detScaler = scaler;
        } else {
            std::fill(nCol.begin(), nCol.end(), 0);
        }
        ropeContact = enabled;
    }
};

}  // namespace orc
