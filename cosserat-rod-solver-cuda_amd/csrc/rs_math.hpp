// Small vector / quaternion / packed-symmetric math for the rope kernels (host and device).
//
// Conventions (they fix the floating-point operation order so that the persistent frame kernel
// and the per-substep kernels produce bit-identical results, and the library is compiled with
// -ffp-contract=off): left-to-right sums, mix(a,b,t) = a*(1-t) + b*t, normalize(v) =
// v * (1/sqrt(dot(v,v))), symmetric 3x3 solves by adjugate / determinant. Quaternions are
// (x, y, z, w) Hamilton quaternions as in the reference's Kit::Rotor
// (third_party/yarnball/KittenEngine/KittenEngine/includes/modules/Rotor.h:7-216).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define RS_HD __host__ __device__ __forceinline__

namespace rs {

struct f3 {
    float x, y, z;
};
struct q4 {
    float x, y, z, w;
};
// packed symmetric 3x3 [xx, yy, zz, xy, xz, yz]
struct s3 {
    float m[6];
};

RS_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
RS_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
RS_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
RS_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
RS_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
RS_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
RS_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
RS_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RS_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
RS_HD float len2(f3 a) { return dot(a, a); }
RS_HD float len(f3 a) { return sqrtf(dot(a, a)); }
RS_HD f3 mix(f3 a, f3 b, float t) { return a * (1.f - t) + b * t; }
RS_HD float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
RS_HD f3 vmin(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
RS_HD f3 vmax(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }

RS_HD f3 qvec(q4 q) { return {q.x, q.y, q.z}; }
RS_HD q4 mkq(f3 v, float w) { return {v.x, v.y, v.z, w}; }
RS_HD q4 qconj(q4 q) { return {-q.x, -q.y, -q.z, q.w}; }
RS_HD q4 qmul(q4 a, q4 b) {
    const f3 v = (a.w * qvec(b) + b.w * qvec(a)) + cross(qvec(a), qvec(b));
    return mkq(v, a.w * b.w - dot(qvec(a), qvec(b)));
}
RS_HD f3 qrot(q4 q, f3 v) {
    const f3 a = q.w * v + cross(qvec(q), v);
    const float c = dot(v, qvec(q));
    return (q.w * a + cross(qvec(q), a)) + c * qvec(q);
}
RS_HD float qdot(q4 a, q4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
RS_HD float qlen(q4 a) { return sqrtf(qdot(a, a)); }
RS_HD q4 qscale(float s, q4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
RS_HD q4 qadd(q4 a, q4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
RS_HD q4 qnormalize(q4 a) { return qscale(1.f / sqrtf(qdot(a, a)), a); }
RS_HD q4 qunit_or_identity(q4 a) {
    const float l2 = qdot(a, a);
    return l2 > 0.f ? qscale(1.f / sqrtf(l2), a) : q4{0.f, 0.f, 0.f, 1.f};
}

RS_HD s3 sdiag(float d) { return s3{{d, d, d, 0.f, 0.f, 0.f}}; }
RS_HD s3 szero() { return s3{{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}}; }
RS_HD s3 souter(f3 v) { return s3{{v.x * v.x, v.y * v.y, v.z * v.z, v.x * v.y, v.x * v.z, v.y * v.z}}; }
RS_HD void sadd(s3& a, const s3& b) {
#pragma unroll
    for (int i = 0; i < 6; i++) a.m[i] = a.m[i] + b.m[i];
}
RS_HD void ssub(s3& a, const s3& b) {
#pragma unroll
    for (int i = 0; i < 6; i++) a.m[i] = a.m[i] - b.m[i];
}
RS_HD s3 sscale(const s3& a, float s) {
    s3 r;
#pragma unroll
    for (int i = 0; i < 6; i++) r.m[i] = a.m[i] * s;
    return r;
}
RS_HD void sadd_diag(s3& a, float d) {
    a.m[0] = a.m[0] + d;
    a.m[1] = a.m[1] + d;
    a.m[2] = a.m[2] + d;
}
// x = H^-1 f by the adjugate.
RS_HD f3 ssolve(const s3& H, f3 f) {
    const float a = H.m[0], b = H.m[1], c = H.m[2], d = H.m[3], e = H.m[4], g = H.m[5];
    const float c00 = b * c - g * g;
    const float c01 = e * g - d * c;
    const float c02 = d * g - b * e;
    const float c11 = a * c - e * e;
    const float c12 = d * e - a * g;
    const float c22 = a * b - d * d;
    const float det = (a * c00 + d * c01) + e * c02;
    const float inv = 1.f / det;
    return mk3((c00 * f.x + c01 * f.y + c02 * f.z) * inv, (c01 * f.x + c11 * f.y + c12 * f.z) * inv,
               (c02 * f.x + c12 * f.y + c22 * f.z) * inv);
}

// Closest points between segments a0a1 and b0b1 as parameters (u on a, v on b). Same steps as the
// reference's segmentClosestPoints (KittenEngine/includes/modules/Common.h:327-345): unclamped v from
// the 2x2 normal equations (0.5 when singular), clamp, reproject u, reproject v.
RS_HD void seg_closest(f3 a0, f3 a1, f3 b0, f3 b1, float& u, float& v) {
    const f3 A0 = a1 - a0, A1 = b1 - b0;
    const float a = len2(A0), b = dot(A0, A1), c = len2(A1);
    const f3 diff = b0 - a0;
    float y = b * dot(A0, diff) - a * dot(A1, diff);
    y = y / (a * c - b * b);
    if (!isfinite(y)) y = 0.5f;
    y = clampf(y, 0.f, 1.f);
    const float x = clampf(dot(mix(b0, b1, y) - a0, A0) / a, 0.f, 1.f);
    y = clampf(dot(mix(a0, a1, x) - b0, A1) / c, 0.f, 1.f);
    u = x;
    v = y;
}

// ------------------------------------------------------------------ arithmetic of the iteration routines
//
// The per-iteration routines (inertia, stretch, surfaces, rope-rope terms, the 3x3 solve, the frame
// update, the end-of-substep velocity) are written once against `Ar<FAST>` and compiled in two modes:
//   FAST = false  IEEE: every operation correctly rounded and unfused, exactly the operations of the
//                 functions above (fma(a, b, c) is a*b + c; the sums keep their order up to commuting
//                 two operands, which is exact), so the float32 checker reproduces them.
//   FAST = true   fused multiply-add (v_fma_f32, correctly rounded), and the hardware reciprocal,
//                 square root, reciprocal square root and base-2 log (v_rcp_f32, v_sqrt_f32, v_rsq_f32,
//                 v_log_f32, about 1 ulp on normal operands; a/b is a * rcp(b)).
// In both modes every operation is an explicit builtin or an unfused IEEE operation (the library is
// compiled with -ffp-contract=off), so the optimiser cannot rewrite them differently in different
// kernels: the per-substep kernels, the persistent frame kernel and the adjoint's forward recompute
// execute the same instructions and stay bit-identical within a mode.
template <bool FAST>
struct Ar;

template <>
struct Ar<false> {
    static RS_HD float fma(float a, float b, float c) { return a * b + c; }
    static RS_HD float rcp(float x) { return 1.f / x; }
    static RS_HD float div(float a, float b) { return a / b; }
    static RS_HD float sqrt(float x) { return sqrtf(x); }
    static RS_HD float rsq(float x) { return 1.f / sqrtf(x); }
    static RS_HD float log(float x) { return logf(x); }
};

template <>
struct Ar<true> {
#if defined(__HIP_DEVICE_COMPILE__)
    static __device__ __forceinline__ float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float log(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
#else
    static float fma(float a, float b, float c) { return fmaf(a, b, c); }
    static float rcp(float x) { return 1.f / x; }
    static float div(float a, float b) { return a / b; }
    static float sqrt(float x) { return sqrtf(x); }
    static float rsq(float x) { return 1.f / sqrtf(x); }
    static float log(float x) { return logf(x); }
#endif
};

// Vector, quaternion and 3x3 helpers of the iteration routines in either mode. With FAST = false each
// returns bit for bit what its IEEE namesake above returns.
template <bool F>
struct V {
    using A = Ar<F>;
    static RS_HD float dot(f3 a, f3 b) { return A::fma(a.z, b.z, A::fma(a.y, b.y, a.x * b.x)); }
    static RS_HD float len2(f3 a) { return dot(a, a); }
    static RS_HD float len(f3 a) { return A::sqrt(dot(a, a)); }
    static RS_HD f3 cross(f3 a, f3 b) {
        return {A::fma(a.y, b.z, -(a.z * b.y)), A::fma(a.z, b.x, -(a.x * b.z)), A::fma(a.x, b.y, -(a.y * b.x))};
    }
    // s*a + b
    static RS_HD f3 axpy(float s, f3 a, f3 b) { return {A::fma(s, a.x, b.x), A::fma(s, a.y, b.y), A::fma(s, a.z, b.z)}; }
    static RS_HD f3 div(f3 a, float s) {
        if (F) {
            const float r = A::rcp(s);
            return a * r;
        }
        return {A::div(a.x, s), A::div(a.y, s), A::div(a.z, s)};
    }
    // a*(1-t) + b*t
    static RS_HD f3 mix(f3 a, f3 b, float t) {
        const float s = 1.f - t;
        return {A::fma(b.x, t, a.x * s), A::fma(b.y, t, a.y * s), A::fma(b.z, t, a.z * s)};
    }
    static RS_HD float qdot(q4 a, q4 b) { return A::fma(a.w, b.w, A::fma(a.z, b.z, A::fma(a.y, b.y, a.x * b.x))); }
    static RS_HD float qlen(q4 a) { return A::sqrt(qdot(a, a)); }
    static RS_HD q4 qmul(q4 a, q4 b) {
        const f3 c = cross(qvec(a), qvec(b));
        const f3 v = mk3(A::fma(a.w, b.x, b.w * a.x) + c.x, A::fma(a.w, b.y, b.w * a.y) + c.y,
                         A::fma(a.w, b.z, b.w * a.z) + c.z);
        return mkq(v, A::fma(a.w, b.w, -dot(qvec(a), qvec(b))));
    }
    // (e, 0) * b: qmul with the zero real part dropped (the dropped terms are exact zeros)
    static RS_HD q4 qmul_pure(f3 e, q4 b) {
        const f3 c = cross(e, qvec(b));
        return mkq(mk3(b.w * e.x + c.x, b.w * e.y + c.y, b.w * e.z + c.z), -dot(e, qvec(b)));
    }
    // a * (1, 0, 0, 0): an exact signed permutation of a
    static RS_HD q4 qmul_ex(q4 a) { return q4{a.w, a.z, -a.y, -a.x}; }
    static RS_HD f3 qrot(q4 q, f3 v) {
        const f3 qv = qvec(q);
        const f3 a = axpy(q.w, v, cross(qv, v));
        const float c = dot(v, qv);
        return axpy(c, qv, axpy(q.w, a, cross(qv, a)));
    }
    // q e_x q*: the general rotation in IEEE mode, the closed form for a unit quaternion in FAST mode
    static RS_HD f3 qrot_ex(q4 q) {
        if (!F) return qrot(q, mk3(1.f, 0.f, 0.f));
        return mk3(A::fma(q.w, q.w, A::fma(q.x, q.x, -A::fma(q.y, q.y, q.z * q.z))),
                   2.f * A::fma(q.x, q.y, q.w * q.z), 2.f * A::fma(q.x, q.z, -(q.w * q.y)));
    }
    static RS_HD q4 qnormalize(q4 a) { return qscale(A::rsq(qdot(a, a)), a); }
    static RS_HD s3 souter(f3 v) { return s3{{v.x * v.x, v.y * v.y, v.z * v.z, v.x * v.y, v.x * v.z, v.y * v.z}}; }
    // a + s*b
    static RS_HD void sadd_scaled(s3& a, const s3& b, float s) {
#pragma unroll
        for (int i = 0; i < 6; i++) a.m[i] = A::fma(b.m[i], s, a.m[i]);
    }
    // x = H^-1 f by the adjugate
    static RS_HD f3 ssolve(const s3& H, f3 f) {
        const float a = H.m[0], b = H.m[1], c = H.m[2], d = H.m[3], e = H.m[4], g = H.m[5];
        const float c00 = A::fma(b, c, -(g * g));
        const float c01 = A::fma(e, g, -(d * c));
        const float c02 = A::fma(d, g, -(b * e));
        const float c11 = A::fma(a, c, -(e * e));
        const float c12 = A::fma(d, e, -(a * g));
        const float c22 = A::fma(a, b, -(d * d));
        const float det = A::fma(e, c02, A::fma(d, c01, a * c00));
        const float inv = A::rcp(det);
        return mk3(A::fma(c02, f.z, A::fma(c01, f.y, c00 * f.x)) * inv,
                   A::fma(c12, f.z, A::fma(c11, f.y, c01 * f.x)) * inv,
                   A::fma(c22, f.z, A::fma(c12, f.y, c02 * f.x)) * inv);
    }
    // seg_closest in this mode (the rope-rope force terms; detection keeps the IEEE one). The 2x2
    // determinant and numerator stay unfused in both modes: for parallel segments with equal edge vectors
    // (a == b == c) a*c - b*b is then exactly 0 and the singular branch (v = 0.5) is taken as in the
    // reference, where a fused a*c - round(b*b) would leave a rounding residue and jump to an endpoint.
    static RS_HD void seg_closest(f3 a0, f3 a1, f3 b0, f3 b1, float& u, float& v) {
        const f3 A0 = a1 - a0, A1 = b1 - b0;
        const float a = len2(A0), b = dot(A0, A1), c = len2(A1);
        const f3 diff = b0 - a0;
        float y = b * dot(A0, diff) - a * dot(A1, diff);
        y = A::div(y, a * c - b * b);
        if (!isfinite(y)) y = 0.5f;
        y = clampf(y, 0.f, 1.f);
        const float x = clampf(A::div(dot(mix(b0, b1, y) - a0, A0), a), 0.f, 1.f);
        y = clampf(A::div(dot(mix(a0, a1, x) - b0, A1), c), 0.f, 1.f);
        u = x;
        v = y;
    }
};

}  // namespace rs
