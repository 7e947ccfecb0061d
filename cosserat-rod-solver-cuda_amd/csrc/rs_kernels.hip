// HIP kernels for gfx950 of the batched Cosserat-rod substep.
//
// Two dispatch forms over the same per-vertex routines (rs_device.hpp):
//  * per-substep kernels: root program, init (+ kinematic mesh poses), rope-rope detection (+ step
//    limit), numItr x (position sweep, orientation sweep), end (+ broad phase), one launch each per
//    substep, state in HBM between launches (the reference's launch structure,
//    csrc/batch.cu:351-358 and sim/step.cpp:14-29, in snapshot / Jacobi semantics);
//  * the persistent frame kernel: one workgroup per cell (whole ropes / contact groups), all
//    substeps of a frame with the rope state in LDS and registers, one state round trip per frame.
// Both give bit-identical results (same arithmetic, no FP contraction, same list order).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rs_device.hpp"
#include "rs_kernels.hpp"

namespace rs {

// ------------------------------------------------------------------ small kernels

__global__ void k_fill_i32(int* p, int v, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_scale_f32(float* p, float s, long n) {  // sdf.cu:89-92
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = p[i] * s;
}
// Fixed-point wrench sums (units of 2^-WRENCH_FIX_BITS) to float, averaged over the frame's substeps.
__global__ void k_wrench_from_fixed(const unsigned long long* fx, float* out, float s, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)((double)(long long)fx[i] * (1.0 / 4294967296.0)) * s;
}

// Constant-velocity root drive once per frame (csrc/batch.cu:90-104).
__global__ void k_root_const(Params P, Dev D, Drive Dr) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.nropes) return;
    const int root = D.roff[e];
    D.x[root] = make_float4(Dr.p0[3 * e], Dr.p0[3 * e + 1], Dr.p0[3 * e + 2], 0.f);
    D.v[root] = make_float4(Dr.v0[3 * e], Dr.v0[3 * e + 1], Dr.v0[3 * e + 2], 0.f);
    if (Dr.hasQ0)
        D.q[root] = stq(qunit_or_identity(q4{Dr.q0[4 * e], Dr.q0[4 * e + 1], Dr.q0[4 * e + 2], Dr.q0[4 * e + 3]}));
}

// Second-order root program at substep s (csrc/batch.cu:122-159).
__global__ void k_root_program(Params P, Dev D, Drive Dr, float4* q, int s, int S) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= P.nropes) return;
    const int root = D.roff[e];
    f3 x, v;
    q4 qq;
    bool setQ;
    root_program(Dr, e, s, S, P.h, x, v, qq, setQ);
    D.x[root] = st3(x);
    D.v[root] = st3(v);
    if (setQ) q[root] = stq(qq);
}

// initItr (YB/sim/iteration.cu:10-69): kinematic mesh poses of this substep, then per vertex the
// displacement guess, the inertial target (stored in the velocity slot) and the detection padding.
template <bool F>
__global__ void k_init(Params P, Dev D, int sub, int nSub, float4* dxOut) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (P.sdfMoving && tid < P.nropes * MAX_SDF) {
        const int m = tid % MAX_SDF;
        if (m < P.numSdf) store_pose(D.poses + tid, substep_pose(D.motion[tid], D.grids[m], sub, nSub, P.h));
    }
    if (tid >= P.nv) return;
    const float4 prm = D.prm[tid];
    const f3 v = ld3(D.v[tid]);
    const f3 vl = ld3(D.vlast[tid]);
    D.vlast[tid] = st3(v);
    f3 d, y;
    float padding;
    init_vertex<F>(P, prm.x, v, vl, d, y, padding);
    if (prm.x != 0.f) D.v[tid] = st3(y);
    D.pad[tid] = padding;
    dxOut[tid] = st3(d);
}

// Rope-rope detection, one thread per segment scanning its contact group in ascending order
// (csrc/batch.cu:19-49), fused with the step limit (YB/sim/collision.cu:103-136).
__global__ void k_detect(Params P, Dev D) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nv) return;
    const uint32_t fl = flags_of(D.prm[i]);
    if (!(fl & HAS_NEXT)) {
        D.ncol[i] = 0;
        D.maxStep[i] = INFINITY;
        return;
    }
    const int rope = D.ropeOf[i];
    const int G = P.groupRopes;
    const int g = rope / G;
    const int g0 = D.roff[min(g * G, P.nropes)], g1 = D.roff[min((g + 1) * G, P.nropes)];
    const f3 a0 = ld3(D.x[i]);
    const f3 a1 = ld3(D.x[i + 1]) - a0;
    const float ri = fmaxf(D.pad[i], D.pad[i + 1]);
    float minDist = D.pad[i] - P.detR;
    int cnt = 0;
    for (int j = g0; j < g1; j++) {
        if (!(flags_of(D.prm[j]) & HAS_NEXT)) continue;
        if (D.ropeOf[j] == rope && abs(j - i) <= 2) continue;
        const f3 b0 = ld3(D.x[j]) - a0, b1 = ld3(D.x[j + 1]) - a0;
        const float r = ri + fmaxf(D.pad[j], D.pad[j + 1]);
        float d2;
        if (pair_hit(a1, b0, b1, r, d2)) {
            if (cnt < MAX_COLL) {
                D.coll[(size_t)i * MAX_COLL + cnt] = j;
                minDist = fminf(minDist, 0.45f * sqrtf(d2));
            }
            cnt++;
        }
    }
    if (cnt > MAX_COLL) D.err[0] = ERR_MAX_COLLISIONS;
    D.ncol[i] = min(cnt, MAX_COLL);
    D.maxStep[i] = minDist;
}

__device__ __forceinline__ uint64_t load_cand(const Dev& D, int i, int n) {
    uint64_t packed = 0;
    for (int c = 0; c < n && c < MAX_CAND; c++) packed |= (uint64_t)D.cand[(size_t)i * MAX_CAND + c] << (6 * c);
    return packed;
}

// Position sweep (YB/sim/cosserat.cu:23-225, snapshot mode): a block owns PLAIN_OWN vertices plus one
// halo vertex on its left whose segment terms it assembles for its first owned vertex.
constexpr int PLAIN_BLOCK = 256;
constexpr int PLAIN_OWN = PLAIN_BLOCK - 1;
template <bool LIMIT, bool F>
__global__ __launch_bounds__(PLAIN_BLOCK) void k_pos(Params P, Dev D, const float4* dxIn, const float4* qIn,
                                                     float4* dxOut) {
    __shared__ float sF[10][PLAIN_BLOCK];
    const int lt = threadIdx.x;
    const int i = blockIdx.x * PLAIN_OWN + lt - 1;
    const bool valid = i >= 0 && i < P.nv;
    f3 f = mk3(0.f, 0.f, 0.f), f2 = f, d0 = f;
    s3 H = szero(), H2 = szero();
    float4 prm = make_float4(0.f, 0.f, 0.f, 0.f);
    float stepLimit = INFINITY;
    if (valid) {
        prm = D.prm[i];
        const uint32_t fl = flags_of(prm);
        const f3 p0 = ld3(D.x[i]);
        d0 = ld3(dxIn[i]);
        f3 p1 = p0, p1dx = d0;
        if (fl & HAS_NEXT) {
            p1 = ld3(D.x[i + 1]);
            p1dx = ld3(dxIn[i + 1]);
        }
        int nc = -1;
        uint64_t cand = 0;
        const Pose* envPoses = D.poses + (size_t)D.ropeOf[i] * MAX_SDF;
        if (P.numSdf > 0) {
            nc = D.ncand[i];
            if (nc > 0) cand = load_cand(D, i, nc);
        }
        vertex_terms<F>(P, D.grids, envPoses, nc, cand, prm.x, prm.y, prm.z, fl, p0, d0, ld3(D.v[i]), ldq(qIn[i]), p1,
                        p1dx, f, H, f2, H2);
        if (LIMIT && (fl & HAS_NEXT)) {
            const float dmp = Ar<F>::div(P.damping, P.h);
            const int n = D.ncol[i];
            if (n > 0) D.pairFlag[D.ropeOf[i] / P.groupRopes] = 1;
            for (int c = 0; c < n; c++) {
                const int j = D.coll[(size_t)i * MAX_COLL + c];
                collision_term<F>(P, dmp, p0, p1, d0, p1dx, ld3(D.x[j]), ld3(D.x[j + 1]), ld3(dxIn[j]), ld3(dxIn[j + 1]),
                               f, H, f2, H2);
            }
        }
        if (LIMIT) stepLimit = D.maxStep[i];
    }
    sF[0][lt] = f2.x;
    sF[1][lt] = f2.y;
    sF[2][lt] = f2.z;
#pragma unroll
    for (int c = 0; c < 6; c++) sF[3 + c][lt] = H2.m[c];
    sF[9][lt] = stepLimit;
    __syncthreads();
    if (!valid || lt == 0) return;
    const uint32_t fl = flags_of(prm);
    if (fl & HAS_PREV) {
        if (LIMIT) stepLimit = fminf(stepLimit, sF[9][lt - 1]);
        f = f + mk3(sF[0][lt - 1], sF[1][lt - 1], sF[2][lt - 1]);
        s3 Hp;
#pragma unroll
        for (int c = 0; c < 6; c++) Hp.m[c] = sF[3 + c][lt - 1];
        sadd(H, Hp);
    }
    f3 d = d0;
    if (prm.x != 0.f) d = solve_update<F>(P, d0, H, f, LIMIT, stepLimit);
    dxOut[i] = st3(d);
}

// Orientation sweep (YB/sim/cosserat.cu:227-273, snapshot mode): reads the new displacements and the
// previous frames, writes the next frames (frames of other segments are copied through).
template <bool F>
__global__ void k_orient(Params P, Dev D, const float4* dxNew, const float4* qIn, float4* qOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nv) return;
    const float4 prm = D.prm[i];
    const uint32_t fl = flags_of(prm);
    const q4 q0 = ldq(qIn[i]);
    if (!(fl & FIX_ORIENT) && (fl & HAS_NEXT)) {
        const q4 qp = (fl & HAS_PREV) ? ldq(qIn[i - 1]) : q0;
        const q4 qn = (fl & HAS_NEXT_ORIENT) ? ldq(qIn[i + 1]) : q0;
        const q4 qrp = (fl & HAS_PREV) ? ldq(D.qrest[i - 1]) : q0;
        qOut[i] = stq(orient_update<F>(fl, prm.y, prm.z, ld3(D.x[i]), ld3(D.x[i + 1]), ld3(dxNew[i]), ld3(dxNew[i + 1]),
                                    qp, q0, qn, qrp, ldq(D.qrest[i])));
    } else {
        qOut[i] = stq(q0);
    }
}

// endItr (YB/sim/iteration.cu:77-129): velocity with drag, optional force readout, position commit,
// broad phase for the next substep.
template <bool F>
__global__ void k_end(Params P, Dev D, const float4* dxIn, const float4* qCur) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i0 < P.nv;  // every lane stays to the readout's wave-wide wrench sums
    const int i = valid ? i0 : P.nv - 1;
    const float4 prm = D.prm[i];
    const uint32_t fl = flags_of(prm);
    const f3 d = ld3(dxIn[i]);
    const f3 x = ld3(D.x[i]);
    const int rope = D.ropeOf[i];
    const Pose* envPoses = D.poses + (size_t)rope * MAX_SDF;
    int nc = -1;
    uint64_t cand = 0;
    if (P.numSdf > 0) {
        nc = D.ncand[i];
        if (nc > 0) cand = load_cand(D, i, nc);
    }
    const bool free = valid && prm.x != 0.f;
    if (free) {
        f3 t = mk3(1.f, 0.f, 0.f);
        if (P.dragQuad > 0.f) t = V<F>::qrot_ex(ldq((fl & HAS_NEXT) ? qCur[i] : qCur[i - 1]));
        D.v[i] = st3(end_velocity<F>(P, d, t));
    }
    if (P.forcesOn && (P.numPlanes > 0 || P.numSdf > 0)) {
        const float kap = Ar<F>::div(P.cKappa, P.h);
        const f3 fv = surface_readout<F>(P, D, envPoses, rope, free, nc, cand, x, d,
                                         free ? Ar<F>::div(kap * kap, prm.x) : 0.f);
        if (free) {
            float* fo = D.force + 3 * (size_t)i;
            fo[0] = fo[0] + fv.x;
            fo[1] = fo[1] + fv.y;
            fo[2] = fo[2] + fv.z;
        }
    }
    if (!valid) return;
    const f3 xn = x + d;
    D.x[i] = st3(xn);
    if (P.numSdf > 0) {
        if (prm.x != 0.f) {
            uint64_t packed;
            const int n = build_candidates(P, envPoses, xn, 2.f * len(d) + 0.02f, packed);
            D.ncand[i] = n;
            for (int c = 0; c < MAX_CAND; c++) D.cand[(size_t)i * MAX_CAND + c] = (n < 0 || c >= n) ? -1 : cand_get(packed, c);
        } else {
            D.ncand[i] = 0;
        }
    }
}

// State -> contiguous float views (csrc/batch.cu:163-175).
__global__ void k_gather(Params P, Dev D) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nv) return;
    const float4 x = D.x[i], v = D.v[i], q = D.q[i];
    D.outPos[3 * i] = x.x;
    D.outPos[3 * i + 1] = x.y;
    D.outPos[3 * i + 2] = x.z;
    D.outVel[3 * i] = v.x;
    D.outVel[3 * i + 1] = v.y;
    D.outVel[3 * i + 2] = v.z;
    D.outQuat[4 * i] = q.x;
    D.outQuat[4 * i + 1] = q.y;
    D.outQuat[4 * i + 2] = q.z;
    D.outQuat[4 * i + 3] = q.w;
}

// Restore construction state of the selected ropes (csrc/batch.cu:178-189): one block per rope.
__global__ void k_reset(Params P, Dev D, const float4* initX, const float4* initQ, const int* ids, int count) {
    const int c = blockIdx.x;
    if (c >= count) return;
    const int e = ids[c];
    for (int i = D.roff[e] + threadIdx.x; i < D.roff[e + 1]; i += blockDim.x) {
        D.x[i] = initX[i];
        D.q[i] = initQ[i];
        D.v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        D.vlast[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// ------------------------------------------------------------------ mesh pose state machine
//
// Every (rope, mesh) slot holds its committed pose, the target it reaches at the end of the next step, and that
// step's motion (start pose, displacement, rotation vector: the form of YB/YarnBall.h:152), which follows from
// the two. Setting a target writes the target and the motion together; a teleported slot starts its motion at
// the target, with no displacement or rotation. The commit after a step makes the targets the committed poses
// and the motions static again. The reference keeps a teleport flag and derives the motions at the start of the
// step (csrc/sdf.cu:37-87); the motions are the same, including the last-call-wins order of several target calls
// before one step, since a target call reads the committed pose, which only the commit changes.

__device__ __forceinline__ Motion static_motion(const Pose7& p) {
    return Motion{p.px, p.py, p.pz, p.qx, p.qy, p.qz, p.qw, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
}

// Motion from pose a to pose b. The rotation vector is the logarithm of the relative rotation q_b conj(q_a),
// taken on the w >= 0 sheet, and zero once the vector part is below 1e-9 (csrc/sdf.cu:25-33).
__device__ __forceinline__ Motion motion_between(const Pose7& a, const Pose7& b) {
    q4 r = V<false>::qmul(q4{b.qx, b.qy, b.qz, b.qw}, qconj(q4{a.qx, a.qy, a.qz, a.qw}));
    if (r.w < 0.f) r = q4{-r.x, -r.y, -r.z, -r.w};
    const float s = V<false>::len(qvec(r));
    const f3 rv = s > 1e-9f ? qvec(r) * (2.f * atan2f(s, r.w) / s) : mk3(0.f, 0.f, 0.f);
    return Motion{a.px, a.py, a.pz, a.qx, a.qy, a.qz, a.qw, b.px - a.px, b.py - a.py, b.pz - a.pz, rv.x, rv.y, rv.z, 0.f};
}

// Targets of rope e's slots: mesh `index` from row 0 (rows == 1) or row e, or with index < 0 every mesh from the
// dense [nropes, M] rows. Quaternions are normalised, a zero one becoming the identity. One thread per rope.
__global__ void k_mesh_targets(MeshSlots S, int index, const float* pos, const float* quat, int rows, int M,
                               const uint8_t* teleFlags, int teleAll, int nropes) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nropes) return;
    const bool tele = teleAll != 0 || (teleFlags != nullptr && teleFlags[e] != 0);
    const bool every = index < 0;
    for (int m = every ? 0 : index; m < (every ? M : index + 1); m++) {
        const size_t row = every ? (size_t)e * M + m : (rows == 1 ? 0 : (size_t)e);
        const float* qr = quat + 4 * row;
        const float l = qlen(q4{qr[0], qr[1], qr[2], qr[3]});
        Pose7 t{pos[3 * row], pos[3 * row + 1], pos[3 * row + 2], 0.f, 0.f, 0.f, 1.f, 0.f};
        if (l > 0.f) {
            t.qx = qr[0] / l;
            t.qy = qr[1] / l;
            t.qz = qr[2] / l;
            t.qw = qr[3] / l;
        }
        const size_t slot = (size_t)e * MAX_SDF + m;
        S.tgt[slot] = t;
        S.motion[slot] = tele ? static_motion(t) : motion_between(S.cur[slot], t);
    }
}

// Commit rope e's targets as static committed poses with their world boxes (csrc/sdf.cu:67-87): mesh `index`, or
// every mesh with index < 0. One thread per rope.
__global__ void k_mesh_commit(MeshSlots S, Pose* poses, const Grid* grids, int numSdf, int index, int nropes) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nropes) return;
    for (int m = index < 0 ? 0 : index; m < (index < 0 ? numSdf : index + 1); m++) {
        const size_t slot = (size_t)e * MAX_SDF + m;
        const Pose7 c = S.tgt[slot];
        S.cur[slot] = c;
        S.motion[slot] = static_motion(c);
        const f3 p = mk3(c.px, c.py, c.pz);
        const q4 q{c.qx, c.qy, c.qz, c.qw};
        f3 lo, hi;
        world_box(grids[m], p, q, lo, hi);
        store_pose(poses + slot, Pose{p.x, p.y, p.z, q.x, q.y, q.z, q.w, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, lo.x, lo.y,
                                      lo.z, hi.x, hi.y, hi.z, 0.f});
    }
}

// ------------------------------------------------------------------ function-level check kernels

// Broad phase on given end-of-substep positions / displacements with the current poses.
__global__ void k_check_broadphase(Params P, Dev D, const float4* xs, const float4* dxs, int* counts, int* cand) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nv) return;
    int n = 0;
    uint64_t packed = 0;
    if (D.prm[i].x != 0.f) {
        const f3 d = ld3(dxs[i]);
        n = build_candidates(P, D.poses + (size_t)D.ropeOf[i] * MAX_SDF, ld3(xs[i]), 2.f * len(d) + 0.02f, packed);
    }
    counts[i] = n;
    for (int c = 0; c < MAX_CAND; c++) cand[(size_t)i * MAX_CAND + c] = (n > 0 && c < n) ? cand_get(packed, c) : -1;
}

// ------------------------------------------------------------------ the persistent frame kernel
//
// One workgroup advances one cell (a contiguous run of whole ropes; a rope-rope contact group never
// straddles cells) through all S substeps of a frame. Thread t owns local vertex slots t + k * BLOCK,
// k < VPT. LDS holds what neighbours and contact partners read: positions, the displacement iterate, the
// frames, the segment terms exchanged with the next vertex, the detection paddings, the step limits, the
// flags and per-chunk boxes for the rope-rope broad phase. Per-vertex constants, velocities, inertial
// targets, the vertex's own frame and broad-phase candidates stay in registers.
//
// Jacobi semantics with one displacement and one frame buffer: per iteration
//   [terms(k): read x, dx^k, partners; write segment terms] | [solve(k): dx^k -> dx^(k+1) in place; publish
//   q^k of the previous orientation sweep] | [orientation(k): read dx^(k+1), neighbours' q^k -> own q^(k+1)
//   kept in a register] | ...
// so nobody reads a slot while its owner overwrites it, and the results equal the per-substep kernels'.
// Root program of every rope at every substep of the frame (csrc/batch.cu:122-159), ahead of the persistent
// kernel so its per-vertex threads only load the result.
__global__ void k_root_traj(Params P, Drive Dr, int S, float4* out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= S * P.nropes) return;
    const int s = idx / P.nropes, e = idx - s * P.nropes;
    f3 x, v;
    q4 q{0.f, 0.f, 0.f, 1.f};
    bool setQ;
    root_program(Dr, e, s, S, P.h, x, v, q, setQ);
    out[3 * (size_t)idx] = make_float4(x.x, x.y, x.z, setQ ? 1.f : 0.f);
    out[3 * (size_t)idx + 1] = st3(v);
    out[3 * (size_t)idx + 2] = stq(q);
}

// Pose of every (rope, mesh) slot at every substep of a frame with moving meshes (YB/sim/iteration.cu:12-33).
__global__ void k_pose_traj(Params P, Dev D, int S, Pose* out) {
    const int M = P.numSdf;
    const long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (idx >= (long long)S * P.nropes * M) return;
    const int m = (int)(idx % M);
    const long long sr = idx / M;
    const int r = (int)(sr % P.nropes), s = (int)(sr / P.nropes);
    store_pose(out + idx, substep_pose(D.motion[(size_t)r * MAX_SDF + m], D.grids[m], s, S, P.h));
}

template <int BLOCK, int VPT, bool FORCES, bool F>
__global__ __launch_bounds__(BLOCK) void k_cell(Params P, Dev D, Drive Dr, const Cell* cells, const int* order,
                                                int ncells, int S, unsigned long long* stats) {
    constexpr int CAP = BLOCK * VPT;
    constexpr int NCH = CAP / 16;
    static_assert(NCH <= 64 && BLOCK % 64 == 0, "chunk masks are 64 bits; waves are 64 lanes");
    extern __shared__ float4 smem4[];
    float4* sQ = smem4;                                  // [CAP] frames
    float4* sQr = smem4 + CAP;                           // [CAP] rest Darboux quaternions (read-only after setup)
    float* sX = reinterpret_cast<float*>(sQr + CAP);    // [3][CAP] positions at the substep start
    float* sD = sX + 3 * CAP;                            // [3][CAP] displacement iterate
    float* sF = sD + 3 * CAP;                            // [9][CAP] segment terms for the next vertex
    float* sPad = sF + 9 * CAP;                          // [CAP]
    uint32_t* sFl = reinterpret_cast<uint32_t*>(sPad + CAP);  // [CAP]
    // during a skin-list build sF holds each segment's padded box and, after it, the list lengths
    // ([CAP][3] pairs (lo.x, lo.y), (lo.z, hi.x), (hi.y, hi.z): one box is three 8-byte reads)
    float2* sBx = reinterpret_cast<float2*>(sF);  // [CAP][3]
    int* sNcol = reinterpret_cast<int*>(sF + 8 * CAP);  // [CAP]
    float* sBox = reinterpret_cast<float*>(sFl + CAP);   // [6][NCH]
    uint32_t* sOv = reinterpret_cast<uint32_t*>(sBox + 6 * NCH);  // [2][NCH] chunk-overlap bit masks (lo, hi words)
    int* sChG = reinterpret_cast<int*>(sOv + 2 * NCH);              // [2][NCH] group range of each chunk
    float* sX0 = reinterpret_cast<float*>(sChG + 2 * NCH);         // [3][CAP] positions at the last skin build
    float* sR = sX0 + 3 * CAP;                                      // [CAP] detection radius bound of the build
    // per-vertex state read once per substep or once per iteration lives here rather than in registers
    float* sVel = sR + CAP;      // [3][CAP] velocity
    float* sVl = sVel + 3 * CAP;  // [3][CAP] velocity at the previous substep start
    float* sY = sVl + 3 * CAP;    // [3][CAP] inertial target of the substep
    float* sLim = sY + 3 * CAP;   // [CAP] step limit of each segment this substep
    int* sWork = reinterpret_cast<int*>(sLim + CAP);  // [4]: next unclaimed chunk of a skin-list build, claimed cell
    auto ld3L = [&](const float* a, int sl) { return mk3(a[sl], a[CAP + sl], a[2 * CAP + sl]); };
    auto st3L = [&](float* a, int sl, f3 v) {
        a[sl] = v.x;
        a[CAP + sl] = v.y;
        a[2 * CAP + sl] = v.z;
    };

    // Cells come from a queue in launch order (longest first): a workgroup that finishes a cell claims the next
    // one at once, so no compute unit idles while cells remain. The grid is at most what the device holds at
    // once; without the queue every workgroup runs cell order[blockIdx.x].
    for (int pass = 0;; pass++) {
        if (D.cellQueue) {
            if (threadIdx.x == 0) sWork[1] = atomicAdd(D.cellQueue, 1);
            __syncthreads();
        }
        const int claim = D.cellQueue ? sWork[1] : (pass == 0 ? (int)blockIdx.x : ncells);
        if (claim >= ncells) break;
        const int cellIdx = order[claim];
        const Cell C = cells[cellIdx];
        const int t = threadIdx.x;
        const unsigned long long startTick = __builtin_amdgcn_s_memrealtime();

        int gi[VPT], rope[VPT], nc[VPT], ncol[VPT];
        uint32_t pairs[VPT];  // the first two rope-rope partners of the substep (16 bits each), kept for the iterations
        int grp[VPT];   // packed local [begin, end) of the contact group (16 bits each)
        int vcnt[VPT];  // skin-list length of the vertex's segment (-1: overflow, scan the whole group)
        float4 prm[VPT];
        q4 qo[VPT];
        float limEff[VPT];
        uint64_t cand[VPT];
        bool isRoot[VPT];

#pragma unroll
        for (int k = 0; k < VPT; k++) {
            const int sl = t + k * BLOCK;
            gi[k] = sl < C.nv ? D.cellMap[C.v0 + sl] : -1;
            nc[k] = P.numSdf > 0 ? -1 : 0;  // a frame starts by visiting every mesh (sdf.cu:226-227)
            cand[k] = 0;
            ncol[k] = 0;
            pairs[k] = 0;
            vcnt[k] = 0;
            limEff[k] = INFINITY;
            if (gi[k] >= 0) {
                const int i = gi[k];
                prm[k] = D.prm[i];
                sQr[sl] = D.qrest[i];
                const uint32_t fl = flags_of(prm[k]);
                rope[k] = D.ropeOf[i];
                isRoot[k] = D.roff[rope[k]] == i;
                {  // a rope contact group occupies consecutive slots, so slot = vertex - (i - sl) inside it
                    const int G = P.groupRopes;
                    const int g = rope[k] / G;
                    const int base = i - sl;
                    grp[k] = (D.roff[min(g * G, P.nropes)] - base) | ((D.roff[min((g + 1) * G, P.nropes)] - base) << 16);
                }
                f3 x = ld3(D.x[i]);
                q4 q = ldq(D.q[i]);
                f3 v = ld3(D.v[i]);
                st3L(sVl, sl, ld3(D.vlast[i]));
                if (Dr.mode == DRIVE_CONST && isRoot[k]) {
                    const int e = rope[k];
                    x = mk3(Dr.p0[3 * e], Dr.p0[3 * e + 1], Dr.p0[3 * e + 2]);
                    v = mk3(Dr.v0[3 * e], Dr.v0[3 * e + 1], Dr.v0[3 * e + 2]);
                    if (Dr.hasQ0) q = qunit_or_identity(q4{Dr.q0[4 * e], Dr.q0[4 * e + 1], Dr.q0[4 * e + 2], Dr.q0[4 * e + 3]});
                }
                sX[sl] = x.x;
                sX[CAP + sl] = x.y;
                sX[2 * CAP + sl] = x.z;
                st3L(sVel, sl, v);
                qo[k] = q;
                sQ[sl] = stq(q);
                sFl[sl] = fl | ((uint32_t)(grp[k] & 0xFFFF) << 4) | ((uint32_t)(grp[k] >> 16) << 15);
            } else {
                prm[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                rope[k] = 0;
                grp[k] = 0;
                isRoot[k] = false;
                const f3 z = mk3(0.f, 0.f, 0.f);
                st3L(sVel, sl, z);
                st3L(sVl, sl, z);
                st3L(sY, sl, z);
                qo[k] = q4{0.f, 0.f, 0.f, 1.f};
                sQr[sl] = make_float4(0.f, 0.f, 0.f, 1.f);
                sX[sl] = sX[CAP + sl] = sX[2 * CAP + sl] = 0.f;
                sQ[sl] = make_float4(0.f, 0.f, 0.f, 1.f);
                sFl[sl] = 0u;
            }
        }
        const float h = P.h;

        for (int s = 0; s < S; s++) {
            // root program of this substep and kinematic mesh poses
            if (Dr.mode == DRIVE_PROGRAM) {
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    if (gi[k] >= 0 && isRoot[k]) {
                        const int sl = t + k * BLOCK;
                        f3 x, v;
                        q4 q;
                        bool setQ;
                        const float4* rt = D.rootTraj + 3 * ((size_t)s * P.nropes + rope[k]);
                        const float4 r0 = rt[0];
                        x = ld3(r0);
                        v = ld3(rt[1]);
                        setQ = r0.w != 0.f;
                        if (setQ) q = ldq(rt[2]);
                        sX[sl] = x.x;
                        sX[CAP + sl] = x.y;
                        sX[2 * CAP + sl] = x.z;
                        st3L(sVel, sl, v);
                        if (setQ) {
                            qo[k] = q;
                            sQ[sl] = stq(q);
                        }
                    }
                }
            }
            // init
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                if (gi[k] >= 0) {
                    const int sl = t + k * BLOCK;
                    f3 d, y;
                    float padding;
                    const f3 v = ld3L(sVel, sl);
                    init_vertex<F>(P, prm[k].x, v, ld3L(sVl, sl), d, y, padding);
                    st3L(sY, sl, y);
                    st3L(sVl, sl, v);
                    sD[sl] = d.x;
                    sD[CAP + sl] = d.y;
                    sD[2 * CAP + sl] = d.z;
                    sPad[sl] = padding;
                }
            }
            if (!P.ropeContact) __syncthreads();  // with rope contact the rebuild vote below is this barrier

            if (P.ropeContact) {
                // Rope-rope broad phase with skin (Verlet) lists. At a rebuild every segment records, in ascending
                // order, the partners whose box (padded by the build radius R + skin) meets its own: a superset of
                // every pair the exact test can accept until some vertex has moved more than the skin from its
                // build position or some detection radius has grown past R (then the next substep rebuilds).
                // Each substep the exact test of the reference runs over that list in order, so the candidate
                // lists are the ones a full ascending scan of the contact group produces.
                const float skin = 2.f * P.scaledDetR, slackR = 0.25f * P.scaledDetR;
                // The vote reads only the vertex's own slots and the build radii (stable since the build): a
                // segment's radius max(pad_i, pad_i+1) outgrows its bound exactly when one of its end vertices
                // does, so each vertex checks its own padding against the segments on either side.
                bool viol = s == 0;
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    if (gi[k] >= 0 && s > 0) {
                        const f3 dxv = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]) - mk3(sX0[sl], sX0[CAP + sl], sX0[2 * CAP + sl]);
                        if (len2(dxv) > skin * skin) viol = true;
                        const uint32_t fl = flags_of(prm[k]);
                        const float pad = sPad[sl];
                        if ((fl & HAS_NEXT) && pad > sR[sl]) viol = true;
                        if ((fl & HAS_PREV) && pad > sR[sl - 1]) viol = true;
                    }
                }
                if (__syncthreads_or(viol)) {  // also publishes the init phase
                    // (1) build boxes padded by the build radius, chunk boxes and chunk group ranges
#pragma unroll
                    for (int k = 0; k < VPT; k++) {
                        const int sl = t + k * BLOCK;
                        f3 lo = mk3(INFINITY, INFINITY, INFINITY), hi = mk3(-INFINITY, -INFINITY, -INFINITY);
                        int cg0 = 1 << 30, cg1 = 0;
                        if (gi[k] >= 0) {
                            cg0 = grp[k] & 0xFFFF;
                            cg1 = grp[k] >> 16;
                            const f3 a = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]);
                            sX0[sl] = a.x;
                            sX0[CAP + sl] = a.y;
                            sX0[2 * CAP + sl] = a.z;
                            if (flags_of(prm[k]) & HAS_NEXT) {
                                const f3 b = mk3(sX[sl + 1], sX[CAP + sl + 1], sX[2 * CAP + sl + 1]);
                                const float R = fmaxf(sPad[sl], sPad[sl + 1]) + slackR;
                                sR[sl] = R;
                                const float rb = R + skin + 1e-4f + 2e-6f * fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z)));
                                lo = vmin(a, b) - mk3(rb, rb, rb);
                                hi = vmax(a, b) + mk3(rb, rb, rb);
                            }
                        }
                        sBx[3 * sl] = make_float2(lo.x, lo.y);
                        sBx[3 * sl + 1] = make_float2(lo.z, hi.x);
                        sBx[3 * sl + 2] = make_float2(hi.y, hi.z);
#pragma unroll
                        for (int o = 1; o < 16; o <<= 1) {
                            lo = vmin(lo, mk3(__shfl_xor(lo.x, o), __shfl_xor(lo.y, o), __shfl_xor(lo.z, o)));
                            hi = vmax(hi, mk3(__shfl_xor(hi.x, o), __shfl_xor(hi.y, o), __shfl_xor(hi.z, o)));
                            cg0 = min(cg0, __shfl_xor(cg0, o));
                            cg1 = max(cg1, __shfl_xor(cg1, o));
                        }
                        if ((t & 15) == 0) {
                            const int ch = sl >> 4;
                            sBox[ch] = lo.x;
                            sBox[NCH + ch] = lo.y;
                            sBox[2 * NCH + ch] = lo.z;
                            sBox[3 * NCH + ch] = hi.x;
                            sBox[4 * NCH + ch] = hi.y;
                            sBox[5 * NCH + ch] = hi.z;
                            sChG[ch] = cg0;
                            sChG[NCH + ch] = cg1;
                        }
                    }
                    if (t == 0) sWork[0] = BLOCK / 64;
                    __syncthreads();
                    const int nch = (C.nv + 15) >> 4;
                    // (2) chunk A meets chunk B (B within A's contact groups): 16 lanes per A, OR-reduced masks
                    for (int base = 0; base < nch * 16; base += BLOCK) {
                        const int A = (base + t) >> 4, sub = t & 15;
                        uint32_t mlo = 0, mhi = 0;
                        if (A < nch && sChG[A] < sChG[NCH + A]) {
                            const float lx = sBox[A], ly = sBox[NCH + A], lz = sBox[2 * NCH + A];
                            const float hx = sBox[3 * NCH + A], hy = sBox[4 * NCH + A], hz = sBox[5 * NCH + A];
                            const int b0 = sChG[A] >> 4, b1 = (sChG[NCH + A] - 1) >> 4;
                            for (int B = b0 + sub; B <= b1; B += 16) {
                                const bool ov = !(sBox[B] > hx || sBox[NCH + B] > hy || sBox[2 * NCH + B] > hz ||
                                                  sBox[3 * NCH + B] < lx || sBox[4 * NCH + B] < ly || sBox[5 * NCH + B] < lz);
                                if (ov) {
                                    if (B < 32)
                                        mlo |= 1u << B;
                                    else
                                        mhi |= 1u << (B - 32);
                                }
                            }
                        }
#pragma unroll
                        for (int o = 1; o < 16; o <<= 1) {
                            mlo |= __shfl_xor(mlo, o);
                            mhi |= __shfl_xor(mhi, o);
                        }
                        if (A < nch && sub == 0) {
                            sOv[A] = mlo;
                            sOv[NCH + A] = mhi;
                        }
                    }
                    __syncthreads();
                    // (3) one wave per chunk A (its own first, then the next unclaimed one): lane (a, g) holds
                    //     segment i = 16 A + a and tests partners 4g..4g+3 of every overlapping chunk B in ascending
                    //     order; the four lanes of a segment append their hits in partner order (cross-lane prefix),
                    //     so each skin list ascends. The four partner boxes are read before any is compared, so the
                    //     twelve reads of a step are in flight together.
                    {
                        const int lane = t & 63, wave = t >> 6;
                        const int a = lane & 15, g = lane >> 4;
                        for (int A = wave; A < nch;) {
                            const int i = A * 16 + a;
                            const uint32_t fi = i < C.nv ? sFl[i] : 0u;
                            const bool seg = (fi & HAS_NEXT) != 0;
                            const int g0 = (fi >> 4) & 0x7FF, g1 = (fi >> 15) & 0x7FF;
                            const bool nextHN = seg && (sFl[i + 1] & HAS_NEXT);
                            const bool prevHN = seg && (fi & HAS_PREV);
                            f3 lo = mk3(INFINITY, INFINITY, INFINITY), hi = -lo;
                            if (seg) {
                                const float2 b0 = sBx[3 * i], b1 = sBx[3 * i + 1], b2 = sBx[3 * i + 2];
                                lo = mk3(b0.x, b0.y, b1.x);
                                hi = mk3(b1.y, b2.x, b2.y);
                            }
                            int cnt = 0;
                            // A is the same on every lane, so the mask is too: the loop below runs on the scalar unit
                            uint64_t mask = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)sOv[NCH + A]) << 32) |
                                            (uint32_t)__builtin_amdgcn_readfirstlane((int)sOv[A]);
                            uint16_t* vl = D.vlist + (size_t)C.v0 * VERLET_CAP + i;  // entry e at vl[e * CAP]
                            while (mask) {
                                const int B = __builtin_ctzll(mask);
                                mask &= mask - 1;
                                float2 e0[4], e1[4], e2[4];
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    const int jc = min(B * 16 + 4 * g + u, C.nv - 1);
                                    e0[u] = sBx[3 * jc];
                                    e1[u] = sBx[3 * jc + 1];
                                    e2[u] = sBx[3 * jc + 2];
                                }
                                uint32_t hm = 0;
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    const int j = B * 16 + 4 * g + u;
                                    const bool apart = (e0[u].x > hi.x) | (e0[u].y > hi.y) | (e1[u].x > hi.z) |
                                                       (e1[u].y < lo.x) | (e2[u].x < lo.y) | (e2[u].y < lo.z);
                                    const int dj = j - i;
                                    const bool self = ((unsigned)(dj + 1) <= 2u) | ((dj == 2) & nextHN) | ((dj == -2) & prevHN);
                                    if (seg & !apart & !self & (j < C.nv) & (j >= g0) & (j < g1)) hm |= 1u << u;
                                }
                                const int c = __builtin_popcount(hm);
                                const int c0 = __shfl(c, a), c1 = __shfl(c, a + 16), c2 = __shfl(c, a + 32), c3 = __shfl(c, a + 48);
                                int idx = cnt + (g > 0 ? c0 : 0) + (g > 1 ? c1 : 0) + (g > 2 ? c2 : 0);
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    if (hm & (1u << u)) {
                                        if (idx < VERLET_CAP) vl[(size_t)idx * CAP] = (uint16_t)(B * 16 + 4 * g + u);
                                        idx++;
                                    }
                                }
                                cnt += c0 + c1 + c2 + c3;
                            }
                            if (g == 0 && i < C.nv) sNcol[i] = cnt <= VERLET_CAP ? cnt : -1;
                            if (stats && g == 0 && seg) atomicAdd(stats, (unsigned long long)cnt);
                            int claim = 0;
                            if (lane == 0) claim = atomicAdd(sWork, 1);
                            A = __builtin_amdgcn_readfirstlane(claim);
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int k = 0; k < VPT; k++)
                        if (gi[k] >= 0) vcnt[k] = sNcol[t + k * BLOCK];
                    if (stats && t == 0) atomicAdd(stats + 3, 1ull);
                }
                // exact tests over the skin list (or over the whole contact group after a skin-list overflow)
                float minD[VPT];
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    float minDist = INFINITY;
                    ncol[k] = 0;
                    pairs[k] = 0;
                    if (gi[k] >= 0 && (flags_of(prm[k]) & HAS_NEXT)) {
                        const int i = gi[k];
                        const f3 a0 = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]);
                        const f3 x1 = mk3(sX[sl + 1], sX[CAP + sl + 1], sX[2 * CAP + sl + 1]);
                        const f3 a1 = x1 - a0;
                        const float ri = fmaxf(sPad[sl], sPad[sl + 1]);
                        minDist = sPad[sl] - P.detR;
                        const float mg = 1e-4f + 2e-6f * fmaxf(fabsf(a0.x), fmaxf(fabsf(a0.y), fabsf(a0.z)));
                        const float rm = ri + mg;
                        const f3 lo = vmin(a0, x1) - mk3(rm, rm, rm), hi = vmax(a0, x1) + mk3(rm, rm, rm);
                        const uint16_t* vl = D.vlist + (size_t)C.v0 * VERLET_CAP + sl;
                        uint16_t* cl = D.clist + (size_t)C.v0 * MAX_COLL + sl;
                        const bool full = vcnt[k] < 0;
                        const int g0 = grp[k] & 0xFFFF, g1 = grp[k] >> 16;
                        int rs0 = 0, rs1 = 0;
                        if (full) {  // local range of the vertex's own rope
                            rs0 = D.roff[rope[k]] - (i - sl);
                            rs1 = D.roff[rope[k] + 1] - (i - sl);
                        }
                        const int n = full ? g1 - g0 : vcnt[k];
                        int cnt = 0, tests = 0;
                        // skin entries come from global memory in batches of four loads issued together, so a long
                        // list waits once per batch instead of once per entry
                        for (int cb = 0; cb < n; cb += 4) {
                            int jb[4];
#pragma unroll
                            for (int u = 0; u < 4; u++)
                                jb[u] = full ? g0 + cb + u : (cb + u < n ? (int)vl[(size_t)(cb + u) * CAP] : 0);
#pragma unroll
                            for (int u = 0; u < 4; u++) {
                                const int j = jb[u];
                                if (cb + u >= n) break;
                                if (full && (!(sFl[j] & HAS_NEXT) || (j >= rs0 && j < rs1 && abs(j - sl) <= 2))) continue;
                                const f3 c0 = mk3(sX[j], sX[CAP + j], sX[2 * CAP + j]);
                                const f3 c1 = mk3(sX[j + 1], sX[CAP + j + 1], sX[2 * CAP + j + 1]);
                                const float rj = fmaxf(sPad[j], sPad[j + 1]);
                                const f3 blo = vmin(c0, c1), bhi = vmax(c0, c1);
                                if (blo.x - rj > hi.x || blo.y - rj > hi.y || blo.z - rj > hi.z || bhi.x + rj < lo.x ||
                                    bhi.y + rj < lo.y || bhi.z + rj < lo.z)
                                    continue;
                                tests++;
                                float d2;
                                if (pair_hit(a1, c0 - a0, c1 - a0, ri + rj, d2)) {
                                    if (cnt < MAX_COLL) {
                                        cl[(size_t)cnt * CAP] = (uint16_t)j;
                                        minDist = fminf(minDist, 0.45f * sqrtf(d2));
                                    }
                                    if (cnt < 2) pairs[k] |= (uint32_t)j << (16 * cnt);
                                    cnt++;
                                }
                            }
                        }
                        if (stats) {  // statistics only: [skin entries at builds, exact tests, list entries, builds]
                            atomicAdd(stats + 1, (unsigned long long)tests);
                            atomicAdd(stats + 2, (unsigned long long)cnt);
                        }
                        if (cnt > MAX_COLL) D.err[0] = ERR_MAX_COLLISIONS;
                        ncol[k] = min(cnt, MAX_COLL);
                        if (cnt > 0) D.pairFlag[rope[k] / P.groupRopes] = 1;
                    }
                    minD[k] = minDist;
                }
                // sLim has a region of its own (last read in the previous substep's first iteration): the first
                // iteration's barrier publishes the limits, and its solve phase reads the neighbour's
#pragma unroll
                for (int k = 0; k < VPT; k++) sLim[t + k * BLOCK] = minD[k];
            }

            // numItr x (position sweep, orientation sweep). A wave with rope-rope pairs this substep issues ahead
            // of the others on its SIMD, and one with mesh candidates ahead of the rest: the barriers wait for them.
            {
                bool pairs = false, meshes = false;
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    pairs = pairs || (P.ropeContact && ncol[k] > 0);
                    meshes = meshes || (gi[k] >= 0 && nc[k] != 0);
                }
                if (__any(pairs))
                    __builtin_amdgcn_s_setprio(2);
                else if (__any(meshes))
                    __builtin_amdgcn_s_setprio(1);
                else
                    __builtin_amdgcn_s_setprio(0);
            }
            for (int it = 0; it < P.numItr; it++) {
                f3 fk[VPT];
                s3 Hk[VPT];
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    f3 f2 = mk3(0.f, 0.f, 0.f);
                    s3 H2 = szero();
                    fk[k] = f2;
                    Hk[k] = H2;
                    if (gi[k] >= 0) {
                        const uint32_t fl = flags_of(prm[k]);
                        const f3 p0 = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]);
                        const f3 d0 = mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl]);
                        f3 p1 = p0, p1dx = d0;
                        if (fl & HAS_NEXT) {
                            p1 = mk3(sX[sl + 1], sX[CAP + sl + 1], sX[2 * CAP + sl + 1]);
                            p1dx = mk3(sD[sl + 1], sD[CAP + sl + 1], sD[2 * CAP + sl + 1]);
                        }
                        vertex_terms<F>(P, D.grids, cell_poses(P, D, rope[k], s), nc[k], cand[k], prm[k].x, prm[k].y,
                                        prm[k].z, fl, p0, d0, ld3L(sY, sl), qo[k], p1, p1dx, fk[k], Hk[k], f2, H2);
                        if (P.ropeContact && (fl & HAS_NEXT)) {
                            const float dmp = Ar<F>::div(P.damping, h);
                            const uint16_t* lst = D.clist + (size_t)C.v0 * MAX_COLL + sl;
                            for (int c = 0; c < ncol[k]; c++) {
                                const int j = c < 2 ? (int)((pairs[k] >> (16 * c)) & 0xFFFFu) : (int)lst[(size_t)c * CAP];
                                collision_term<F>(P, dmp, p0, p1, d0, p1dx, mk3(sX[j], sX[CAP + j], sX[2 * CAP + j]),
                                                  mk3(sX[j + 1], sX[CAP + j + 1], sX[2 * CAP + j + 1]),
                                                  mk3(sD[j], sD[CAP + j], sD[2 * CAP + j]),
                                                  mk3(sD[j + 1], sD[CAP + j + 1], sD[2 * CAP + j + 1]), fk[k], Hk[k], f2, H2);
                            }
                        }
                    }
                    sF[sl] = f2.x;
                    sF[CAP + sl] = f2.y;
                    sF[2 * CAP + sl] = f2.z;
#pragma unroll
                    for (int c = 0; c < 6; c++) sF[(3 + c) * CAP + sl] = H2.m[c];
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    if (gi[k] >= 0) {
                        const uint32_t fl = flags_of(prm[k]);
                        if (P.ropeContact && it == 0) {  // step limit of the substep: own and previous segment
                            limEff[k] = sLim[sl];
                            if (fl & HAS_PREV) limEff[k] = fminf(limEff[k], sLim[sl - 1]);
                        }
                        f3 f = fk[k];
                        s3 H = Hk[k];
                        if (fl & HAS_PREV) {
                            f = f + mk3(sF[sl - 1], sF[CAP + sl - 1], sF[2 * CAP + sl - 1]);
                            s3 Hp;
#pragma unroll
                            for (int c = 0; c < 6; c++) Hp.m[c] = sF[(3 + c) * CAP + sl - 1];
                            sadd(H, Hp);
                        }
                        if (prm[k].x != 0.f) {
                            f3 d = mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl]);
                            d = solve_update<F>(P, d, H, f, P.ropeContact != 0, limEff[k]);
                            sD[sl] = d.x;
                            sD[CAP + sl] = d.y;
                            sD[2 * CAP + sl] = d.z;
                        }
                        // publish the frames of the previous orientation sweep (at it = 0 those of the previous substep
                        // or of the root program)
                        sQ[sl] = stq(qo[k]);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    if (gi[k] >= 0) {
                        const uint32_t fl = flags_of(prm[k]);
                        if (!(fl & FIX_ORIENT) && (fl & HAS_NEXT)) {
                            const q4 q0 = ldq(sQ[sl]);
                            const q4 qp = (fl & HAS_PREV) ? ldq(sQ[sl - 1]) : q0;
                            const q4 qn = (fl & HAS_NEXT_ORIENT) ? ldq(sQ[sl + 1]) : q0;
                            qo[k] = orient_update<F>(fl, prm[k].y, prm[k].z, mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]),
                                                  mk3(sX[sl + 1], sX[CAP + sl + 1], sX[2 * CAP + sl + 1]),
                                                  mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl]),
                                                  mk3(sD[sl + 1], sD[CAP + sl + 1], sD[2 * CAP + sl + 1]), qp, q0, qn,
                                                  ldq(sQr[(fl & HAS_PREV) ? sl - 1 : sl]), ldq(sQr[sl]));
                        }
                    }
                }
            }
            __syncthreads();  // the last sweep's neighbour reads are done before the end phase moves the positions
            if (P.dragQuad > 0.f) {  // the tips' drag tangent: the previous segment's frame after the last sweep
#pragma unroll
                for (int k = 0; k < VPT; k++)
                    if (gi[k] >= 0) sQ[t + k * BLOCK] = stq(qo[k]);
                __syncthreads();
            }

            // end of substep: force readout (every lane, for the wave-wide wrench sums), then velocity and commit
            if (FORCES) {
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    const bool on = gi[k] >= 0 && prm[k].x != 0.f;
                    const float kap = Ar<F>::div(P.cKappa, h);
                    const f3 fv = surface_readout<F>(P, D, cell_poses(P, D, rope[k], s), rope[k], on, nc[k], cand[k],
                                                     mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]),
                                                     mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl]),
                                                     on ? Ar<F>::div(kap * kap, prm[k].x) : 0.f);
                    if (on) {
                        float* fo = D.force + 3 * (size_t)gi[k];
                        fo[0] = fo[0] + fv.x;
                        fo[1] = fo[1] + fv.y;
                        fo[2] = fo[2] + fv.z;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < VPT; k++) {
                if (gi[k] >= 0) {
                    const int sl = t + k * BLOCK;
                    const uint32_t fl = flags_of(prm[k]);
                    const f3 d = mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl]);
                    const f3 x = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]);
                    if (prm[k].x != 0.f) {
                        f3 tg = mk3(1.f, 0.f, 0.f);
                        if (P.dragQuad > 0.f) tg = V<F>::qrot_ex((fl & HAS_NEXT) ? qo[k] : ldq(sQ[sl - 1]));
                        st3L(sVel, sl, end_velocity<F>(P, d, tg));
                    }
                    const f3 xn = x + d;
                    sX[sl] = xn.x;
                    sX[CAP + sl] = xn.y;
                    sX[2 * CAP + sl] = xn.z;
                }
            }
            // mesh broad phase for the next substep (YB/sim/contact.cuh:121-136). With many meshes, chunks of 16 lanes
            // on one rope split the meshes over their lanes first, then each vertex runs the exact test on the chunk's
            // survivors only; with a few, each vertex tests them all.
            if (P.numSdf > 0 && P.numSdf <= CHUNK_BROADPHASE_MIN_MESHES) {
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    if (gi[k] >= 0)
                        nc[k] = prm[k].x != 0.f ? build_candidates(P, cell_poses(P, D, rope[k], s),
                                                                   mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]),
                                                                   2.f * len(mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl])) +
                                                                       0.02f,
                                                                   cand[k])
                                                : 0;
                }
            } else if (P.numSdf > 0) {
#pragma unroll
                for (int k = 0; k < VPT; k++) {
                    const int sl = t + k * BLOCK;
                    const bool act = gi[k] >= 0 && prm[k].x != 0.f;
                    const f3 xn = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]);
                    const float padv = 2.f * len(mk3(sD[sl], sD[CAP + sl], sD[2 * CAP + sl])) + 0.02f;
                    int r0 = act ? rope[k] : 0x7fffffff, r1 = act ? rope[k] : -1;
#pragma unroll
                    for (int o = 1; o < 16; o <<= 1) {
                        r0 = min(r0, __shfl_xor(r0, o));
                        r1 = max(r1, __shfl_xor(r1, o));
                    }
                    const bool oneRope = r0 == r1;  // false also for a chunk without free vertices
                    const Pose* chunkPoses = cell_poses(P, D, oneRope ? r0 : 0, s);
                    const uint64_t mask = chunk_mesh_mask(P, chunkPoses, act && oneRope, xn, padv);
                    if (act)
                        nc[k] = oneRope ? build_candidates_masked(P, chunkPoses, xn, padv, mask, cand[k])
                                        : build_candidates(P, cell_poses(P, D, rope[k], s), xn, padv, cand[k]);
                    else if (gi[k] >= 0)
                        nc[k] = 0;
                }
            }
            // no barrier here: the end phase touched only each vertex's own slots, and the next substep reads a
            // neighbour's slot only after its init barrier
        }

        // write the state back, fill the views
#pragma unroll
        for (int k = 0; k < VPT; k++) {
            if (gi[k] >= 0) {
                const int sl = t + k * BLOCK;
                const int i = gi[k];
                const f3 x = mk3(sX[sl], sX[CAP + sl], sX[2 * CAP + sl]);
                const q4 q = qo[k];
                D.x[i] = st3(x);
                const f3 vel = ld3L(sVel, sl);
                D.v[i] = st3(vel);
                D.vlast[i] = st3(ld3L(sVl, sl));
                D.q[i] = stq(q);
                D.outPos[3 * i] = x.x;
                D.outPos[3 * i + 1] = x.y;
                D.outPos[3 * i + 2] = x.z;
                D.outVel[3 * i] = vel.x;
                D.outVel[3 * i + 1] = vel.y;
                D.outVel[3 * i + 2] = vel.z;
                D.outQuat[4 * i] = q.x;
                D.outQuat[4 * i + 1] = q.y;
                D.outQuat[4 * i + 2] = q.z;
                D.outQuat[4 * i + 3] = q.w;
                if (P.numSdf > 0) {
                    D.ncand[i] = nc[k];
                    for (int c = 0; c < MAX_CAND; c++)
                        D.cand[(size_t)i * MAX_CAND + c] = (nc[k] > 0 && c < nc[k]) ? cand_get(cand[k], c) : -1;
                }
            }
        }
        // the cell's duration orders the next frame's launch (k_rank_cells); the barrier also keeps the next
        // claim from overwriting LDS that this cell still reads
        __syncthreads();
        if (D.cellTime && t == 0) D.cellTime[cellIdx] = (unsigned)(__builtin_amdgcn_s_memrealtime() - startTick);
    }
}

// Launch order of the cells for the next frame, longest first by the durations of this frame: cell c goes
// to position #{c' : time[c'] > time[c], or time[c'] == time[c] and c' < c}, a permutation. Workgroups claim
// cells in this order, so the slowest cells start first and the short ones fill in behind them (longest
// processing time first). The order decides only where and when a cell runs, never its results.
__global__ __launch_bounds__(256) void k_rank_cells(const unsigned* time, int n, int* order) {
    __shared__ unsigned tile[256];
    const int c = blockIdx.x * 256 + threadIdx.x;
    const unsigned mine = c < n ? time[c] : 0u;
    int rank = 0;
    for (int base = 0; base < n; base += 256) {
        const int j = base + threadIdx.x;
        tile[threadIdx.x] = j < n ? time[j] : 0u;
        __syncthreads();
        const int m = min(256, n - base);
        for (int k = 0; k < m; k++) {
            const unsigned o = tile[k];
            rank += (o > mine || (o == mine && base + k < c)) ? 1 : 0;
        }
        __syncthreads();
    }
    if (c < n) order[rank] = c;
}

// ------------------------------------------------------------------ host launchers

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

hipError_t launch_fill_i32(int* p, int v, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_fill_f32(float* p, float v, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_wrench_from_fixed(const unsigned long long* fx, float* out, float scale, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_wrench_from_fixed, dim3(nblk(n, 256)), dim3(256), 0, s, fx, out, scale, n);
    return hipGetLastError();
}
hipError_t launch_scale_f32(float* p, float v, long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_f32, dim3(nblk(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t launch_root_const(const Params& P, const Dev& D, const Drive& Dr, hipStream_t s) {
    hipLaunchKernelGGL(k_root_const, dim3(nblk(P.nropes, 128)), dim3(128), 0, s, P, D, Dr);
    return hipGetLastError();
}
hipError_t launch_gather(const Params& P, const Dev& D, hipStream_t s) {
    hipLaunchKernelGGL(k_gather, dim3(nblk(P.nv, 256)), dim3(256), 0, s, P, D);
    return hipGetLastError();
}
hipError_t launch_reset(const Params& P, const Dev& D, const float4* initX, const float4* initQ, const int* ids,
                        int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reset, dim3(count), dim3(256), 0, s, P, D, initX, initQ, ids, count);
    return hipGetLastError();
}

// One substep of the per-substep method. dx / q ping-pong through D.dx/D.dx2 and D.q/D.q2; on return
// D.q holds the current frames (the pointers in D are swapped as needed).
hipError_t launch_plain_substep(const Params& P, Dev& D, const Drive& Dr, int s, int S, hipStream_t st,
                                const Bvh* bvh) {
    const int nvb = nblk(P.nv, 256);
    if (Dr.mode == DRIVE_PROGRAM)
        hipLaunchKernelGGL(k_root_program, dim3(nblk(P.nropes, 128)), dim3(128), 0, st, P, D, Dr, D.q, s, S);
    const int initThreads = P.sdfMoving ? max(P.nv, P.nropes * MAX_SDF) : P.nv;
    if (P.fastMath)
        hipLaunchKernelGGL(k_init<true>, dim3(nblk(initThreads, 256)), dim3(256), 0, st, P, D, s, S, D.dx);
    else
        hipLaunchKernelGGL(k_init<false>, dim3(nblk(initThreads, 256)), dim3(256), 0, st, P, D, s, S, D.dx);
    if (P.ropeContact) {
        if (bvh) {
            const hipError_t e = launch_bvh_detect(P, D, *bvh, s == 0, st);
            if (e != hipSuccess) return e;
        } else {
            hipLaunchKernelGGL(k_detect, dim3(nblk(P.nv, 128)), dim3(128), 0, st, P, D);
        }
    }
    float4* dIn = D.dx;
    float4* dOut = D.dx2;
    float4* qIn = D.q;
    float4* qOut = D.q2;
    const int posBlocks = nblk(P.nv + 1, PLAIN_OWN);
    for (int it = 0; it < P.numItr; it++) {
        if (P.fastMath) {
            if (P.ropeContact)
                hipLaunchKernelGGL((k_pos<true, true>), dim3(posBlocks), dim3(PLAIN_BLOCK), 0, st, P, D, dIn, qIn, dOut);
            else
                hipLaunchKernelGGL((k_pos<false, true>), dim3(posBlocks), dim3(PLAIN_BLOCK), 0, st, P, D, dIn, qIn, dOut);
            hipLaunchKernelGGL(k_orient<true>, dim3(nvb), dim3(256), 0, st, P, D, dOut, qIn, qOut);
        } else {
            if (P.ropeContact)
                hipLaunchKernelGGL((k_pos<true, false>), dim3(posBlocks), dim3(PLAIN_BLOCK), 0, st, P, D, dIn, qIn, dOut);
            else
                hipLaunchKernelGGL((k_pos<false, false>), dim3(posBlocks), dim3(PLAIN_BLOCK), 0, st, P, D, dIn, qIn, dOut);
            hipLaunchKernelGGL(k_orient<false>, dim3(nvb), dim3(256), 0, st, P, D, dOut, qIn, qOut);
        }
        float4* t = dIn;
        dIn = dOut;
        dOut = t;
        t = qIn;
        qIn = qOut;
        qOut = t;
    }
    if (P.fastMath)
        hipLaunchKernelGGL(k_end<true>, dim3(nvb), dim3(256), 0, st, P, D, dIn, qIn);
    else
        hipLaunchKernelGGL(k_end<false>, dim3(nvb), dim3(256), 0, st, P, D, dIn, qIn);
    if (qIn != D.q) {
        D.q2 = D.q;
        D.q = qIn;
    }
    return hipGetLastError();
}

hipError_t launch_rank_cells(const unsigned* time, int n, int* order, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rank_cells, dim3((n + 255) / 256), dim3(256), 0, st, time, n, order);
    return hipGetLastError();
}

size_t variant_lds_bytes(int variant) {
    const CellVariant& v = kCellVariants[variant];
    return v.strand ? strand_lds_bytes(v.vpt) : cell_lds_bytes(v.block, v.vpt);
}

size_t cell_lds_bytes(int block, int vpt) {
    const size_t cap = (size_t)block * vpt;
    return cap * (4 + 4 + 3 + 3 + 9 + 1 + 1 + 3 + 1 + 3 + 3 + 3 + 1) * sizeof(float) + (cap / 16) * (6 + 2 + 2) * sizeof(float) +
           4 * sizeof(int);
}

template <int B, int V, bool FORCES, bool F>
static hipError_t launch_cell_f(const Params& P, const Dev& D, const Drive& Dr, const Cell* cells, const int* order,
                                int ncells, int S, unsigned long long* stats, hipStream_t st) {
    const size_t lds = cell_lds_bytes(B, V);
    static bool attrSet = false;
    static int resident = 0;  // workgroups of this kernel the device holds at once
    if (!attrSet) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cell<B, V, FORCES, F>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        int dev = 0, cus = 0, perCu = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, reinterpret_cast<const void*>(&k_cell<B, V, FORCES, F>),
                                                             B, lds);
        if (e != hipSuccess) return e;
        resident = cus * perCu;
        attrSet = true;
    }
    int grid = ncells;
    if (D.cellQueue) {
        // the queue starts at the head of the launch order; a workgroup beyond the resident count would only
        // find it drained
        const hipError_t e = launch_fill_i32(D.cellQueue, 0, 1, st);
        if (e != hipSuccess) return e;
        if (resident > 0) grid = std::min(ncells, resident);
    }
    hipLaunchKernelGGL((k_cell<B, V, FORCES, F>), dim3(grid), dim3(B), lds, st, P, D, Dr, cells, order, ncells, S,
                       stats);
    return hipGetLastError();
}
// The per-frame inputs of the persistent kernels: the root program and the moving-mesh poses at every substep.
static hipError_t launch_frame_precompute(const Params& P, const Dev& D, const Drive& Dr, int S, hipStream_t st) {
    if (Dr.mode == DRIVE_PROGRAM) {
        hipLaunchKernelGGL(k_root_traj, dim3(nblk(S * P.nropes, 256)), dim3(256), 0, st, P, Dr, S,
                           const_cast<float4*>(D.rootTraj));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (P.numSdf > 0 && P.sdfMoving) {
        const long long n = (long long)S * P.nropes * P.numSdf;
        hipLaunchKernelGGL(k_pose_traj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P, D, S,
                           const_cast<Pose*>(D.posesS));
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int B, int V>
static hipError_t launch_cell_t(const Params& P, const Dev& D, const Drive& Dr, const Cell* cells, const int* order,
                                int ncells, int S, unsigned long long* stats, hipStream_t st) {
    const hipError_t pe = launch_frame_precompute(P, D, Dr, S, st);
    if (pe != hipSuccess) return pe;
    const bool forces = P.forcesOn && (P.numPlanes > 0 || P.numSdf > 0);
    if (P.fastMath) {
        if (forces) return launch_cell_f<B, V, true, true>(P, D, Dr, cells, order, ncells, S, stats, st);
        return launch_cell_f<B, V, false, true>(P, D, Dr, cells, order, ncells, S, stats, st);
    }
    if (forces) return launch_cell_f<B, V, true, false>(P, D, Dr, cells, order, ncells, S, stats, st);
    return launch_cell_f<B, V, false, false>(P, D, Dr, cells, order, ncells, S, stats, st);
}

const CellVariant kCellVariants[] = {{64, 1, 0},  {128, 1, 0}, {256, 1, 0}, {384, 1, 0}, {512, 1, 0}, {768, 1, 0},
                                     {1024, 1, 0}, {256, 2, 0},  {256, 4, 0}, {512, 2, 0}, {64, 1, 1},  {64, 2, 1},
                                     {64, 4, 1}};
const int kNumCellVariants = sizeof(kCellVariants) / sizeof(kCellVariants[0]);

hipError_t launch_cell(int variant, const Params& P, const Dev& D, const Drive& Dr, const Cell* cells, const int* order,
                       int ncells, int S, unsigned long long* stats, hipStream_t st) {
    if (ncells <= 0) return hipSuccess;
    if (variant >= 0 && variant < kNumCellVariants && kCellVariants[variant].strand) {
        const hipError_t pe = launch_frame_precompute(P, D, Dr, S, st);
        if (pe != hipSuccess) return pe;
        return launch_strand(kCellVariants[variant].vpt, P, D, Dr, order, ncells, S, st);
    }
    switch (variant) {
        case 0: return launch_cell_t<64, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 1: return launch_cell_t<128, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 2: return launch_cell_t<256, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 3: return launch_cell_t<384, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 4: return launch_cell_t<512, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 5: return launch_cell_t<768, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 6: return launch_cell_t<1024, 1>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 7: return launch_cell_t<256, 2>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 8: return launch_cell_t<256, 4>(P, D, Dr, cells, order, ncells, S, stats, st);
        case 9: return launch_cell_t<512, 2>(P, D, Dr, cells, order, ncells, S, stats, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mesh_targets(const MeshSlots& S, int index, const float* pos, const float* quat, int rows, int M,
                               const uint8_t* teleFlags, int teleAll, int nropes, hipStream_t s) {
    if (nropes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mesh_targets, dim3(nblk(nropes, 128)), dim3(128), 0, s, S, index, pos, quat, rows, M, teleFlags,
                       teleAll, nropes);
    return hipGetLastError();
}
hipError_t launch_mesh_commit(const MeshSlots& S, Pose* poses, const Grid* grids, int numSdf, int index, int nropes,
                              hipStream_t s) {
    if (nropes <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mesh_commit, dim3(nblk(nropes, 128)), dim3(128), 0, s, S, poses, grids, numSdf, index, nropes);
    return hipGetLastError();
}
hipError_t launch_check_detect(const Params& P, const Dev& D, const Bvh* bvh, hipStream_t s) {
    if (bvh) return launch_bvh_detect(P, D, *bvh, true, s);
    hipLaunchKernelGGL(k_detect, dim3(nblk(P.nv, 128)), dim3(128), 0, s, P, D);
    return hipGetLastError();
}
hipError_t launch_check_broadphase(const Params& P, const Dev& D, const float4* xs, const float4* dxs, int* counts,
                                   int* cand, hipStream_t s) {
    hipLaunchKernelGGL(k_check_broadphase, dim3(nblk(P.nv, 256)), dim3(256), 0, s, P, D, xs, dxs, counts, cand);
    return hipGetLastError();
}

}  // namespace rs
