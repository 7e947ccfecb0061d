// The strand form of the persistent frame kernel: one wavefront per rope.
//
// k_cell spreads a cell's vertices over up to 16 waves and puts every neighbour exchange of the Jacobi sweeps
// behind a workgroup barrier, two per iteration: the waves of a cell wait for the slowest one 1600 times a
// frame. Here a rope of up to 64 * NB vertices is one wave. Lane l holds the rope's vertices l, l + 64, ...,
// (block k = vertex / 64), and their state stays in registers across the frame: positions, iterates, frames,
// inertial targets, material and mesh candidates. Neighbour values (vertex i + 1's position, iterate and frame,
// vertex i - 1's segment terms and frame), velocities and rest Darboux vectors sit in a small wave-private LDS
// area. The lanes of one wave run in lockstep and LDS serves a wave's accesses in issue order, so an exchange
// needs no barrier, only a wave-scope fence that keeps the compiler from reordering the accesses. With no
// workgroup barrier left, a CU runs several ropes' waves side by side and their latencies hide each other.
//
// Arithmetic, operation order and list order are those of the per-substep kernels (the same device routines of
// rs_device.hpp), so the results are bit-identical to them and to k_cell. Ropes with rope-rope contact need
// their contact partners' iterates every iteration, and take the cell kernel.
#include <hip/hip_runtime.h>

#include "rs_device.hpp"
#include "rs_kernels.hpp"

namespace rs {

constexpr int WAVE = 64;
// floats per vertex of the exchange area: a vertex's iterate (3) and frame (4); a segment's terms for the next
// vertex, f2 (3) and the diagonal of H2 (1; without rope-rope pairs H2 is that multiple of the identity)
constexpr int STRAND_EX = 7;

// Orders the LDS accesses of a wave: every access issued before it completes before any issued after it.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

size_t strand_lds_bytes(int nb) { return (size_t)(STRAND_EX + 3 + 3 + 3 + 4) * WAVE * nb * sizeof(float); }

template <int NB, bool FORCES, bool F>
__global__ __launch_bounds__(WAVE) void k_strand(Params P, Dev D, Drive Dr, const int* order, int S) {
    constexpr int CAP = NB * WAVE;
    extern __shared__ float smem[];
    float* ex = smem;                // [STRAND_EX][CAP] exchange rows
    float* sx = ex + STRAND_EX * CAP;  // [3][CAP] positions at the substep start
    float* sv = sx + 3 * CAP;        // [3][CAP] velocities
    float* svl = sv + 3 * CAP;       // [3][CAP] velocities at the previous substep start
    float* sqr = svl + 3 * CAP;      // [4][CAP] rest Darboux vectors
    auto put3 = [&](float* a, int s, f3 v) {
        a[s] = v.x;
        a[CAP + s] = v.y;
        a[2 * CAP + s] = v.z;
    };
    auto get3 = [&](const float* a, int s) { return mk3(a[s], a[CAP + s], a[2 * CAP + s]); };
    auto put4 = [&](float* a, int s, q4 v) {
        a[s] = v.x;
        a[CAP + s] = v.y;
        a[2 * CAP + s] = v.z;
        a[3 * CAP + s] = v.w;
    };
    auto get4 = [&](const float* a, int s) { return q4{a[s], a[CAP + s], a[2 * CAP + s], a[3 * CAP + s]}; };

    const int rope = order[blockIdx.x];
    const int lane = threadIdx.x;
    const int v0 = D.roff[rope];
    const int nv = D.roff[rope + 1] - v0;
    const int nb = (nv + WAVE - 1) / WAVE;  // blocks this rope fills (uniform over the wave)

    f3 x[NB], d[NB], dn[NB], y[NB];
    q4 q[NB];
    float4 prm[NB];
    uint64_t cand[NB];
    int nc[NB];
    bool ok[NB];

#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int s = k * WAVE + lane;
        ok[k] = k < nb && s < nv;
        nc[k] = P.numSdf > 0 ? -1 : 0;  // a frame starts by visiting every mesh (sdf.cu:226-227)
        cand[k] = 0;
        x[k] = d[k] = dn[k] = y[k] = mk3(0.f, 0.f, 0.f);
        q[k] = q4{0.f, 0.f, 0.f, 1.f};
        prm[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ok[k]) {
            const int i = v0 + s;
            prm[k] = D.prm[i];
            x[k] = ld3(D.x[i]);
            q[k] = ldq(D.q[i]);
            f3 v = ld3(D.v[i]);
            if (Dr.mode == DRIVE_CONST && s == 0) {
                x[k] = mk3(Dr.p0[3 * rope], Dr.p0[3 * rope + 1], Dr.p0[3 * rope + 2]);
                v = mk3(Dr.v0[3 * rope], Dr.v0[3 * rope + 1], Dr.v0[3 * rope + 2]);
                if (Dr.hasQ0)
                    q[k] = qunit_or_identity(
                        q4{Dr.q0[4 * rope], Dr.q0[4 * rope + 1], Dr.q0[4 * rope + 2], Dr.q0[4 * rope + 3]});
            }
            put3(sv, s, v);
            put3(svl, s, ld3(D.vlast[i]));
            put4(sqr, s, ldq(D.qrest[i]));
        }
    }
    wave_sync();
    const float h = P.h;

    for (int st = 0; st < S; st++) {
        const Pose* poses = cell_poses(P, D, rope, st);
        // root program of this substep
        if (Dr.mode == DRIVE_PROGRAM && lane == 0) {
            const float4* rt = D.rootTraj + 3 * ((size_t)st * P.nropes + rope);
            const float4 r0 = rt[0];
            x[0] = ld3(r0);
            put3(sv, 0, ld3(rt[1]));
            if (r0.w != 0.f) q[0] = ldq(rt[2]);
        }
        // init, then the neighbours' positions and first iterates
#pragma unroll
        for (int k = 0; k < NB; k++) {
            if (ok[k]) {
                const int s = k * WAVE + lane;
                const f3 v = get3(sv, s);
                float padding;
                init_vertex<F>(P, prm[k].x, v, get3(svl, s), d[k], y[k], padding);
                put3(svl, s, v);
                put3(sx, s, x[k]);
                put3(ex, s, d[k]);
            }
        }
        wave_sync();
#pragma unroll
        for (int k = 0; k < NB; k++)
            if (ok[k]) dn[k] = (flags_of(prm[k]) & HAS_NEXT) ? get3(ex, k * WAVE + lane + 1) : d[k];
        wave_sync();

        for (int it = 0; it < P.numItr; it++) {
            f3 fk[NB];
            s3 Hk[NB];
            // position sweep: own terms, and the segment's terms for the next vertex through the exchange area
#pragma unroll
            for (int k = 0; k < NB; k++) {
                fk[k] = mk3(0.f, 0.f, 0.f);
                Hk[k] = szero();
                if (ok[k]) {
                    const int s = k * WAVE + lane;
                    const uint32_t fl = flags_of(prm[k]);
                    f3 p1 = x[k], p1dx = d[k];
                    if (fl & HAS_NEXT) {
                        p1 = get3(sx, s + 1);
                        p1dx = dn[k];
                    }
                    f3 f2 = mk3(0.f, 0.f, 0.f);
                    s3 H2 = szero();
                    vertex_terms<F>(P, D.grids, poses, nc[k], cand[k], prm[k].x, prm[k].y, prm[k].z, fl, x[k], d[k],
                                    y[k], q[k], p1, p1dx, fk[k], Hk[k], f2, H2);
                    put3(ex, s, f2);
                    ex[3 * CAP + s] = H2.m[0];
                }
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (ok[k]) {
                    const int s = k * WAVE + lane;
                    const uint32_t fl = flags_of(prm[k]);
                    f3 f = fk[k];
                    s3 H = Hk[k];
                    if (fl & HAS_PREV) {
                        f = f + get3(ex, s - 1);
                        s3 Hp = szero();
                        sadd_diag(Hp, ex[3 * CAP + s - 1]);
                        sadd(H, Hp);
                    }
                    if (prm[k].x != 0.f) d[k] = solve_update<F>(P, d[k], H, f, false, INFINITY);
                }
            }
            wave_sync();
            // orientation sweep: the neighbours' new iterates and current frames
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (ok[k]) {
                    const int s = k * WAVE + lane;
                    put3(ex, s, d[k]);
                    put4(ex + 3 * CAP, s, q[k]);
                }
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (ok[k]) {
                    const int s = k * WAVE + lane;
                    const uint32_t fl = flags_of(prm[k]);
                    if (fl & HAS_NEXT) dn[k] = get3(ex, s + 1);
                    if (!(fl & FIX_ORIENT) && (fl & HAS_NEXT)) {
                        const q4 qp = (fl & HAS_PREV) ? get4(ex + 3 * CAP, s - 1) : q[k];
                        const q4 qn = (fl & HAS_NEXT_ORIENT) ? get4(ex + 3 * CAP, s + 1) : q[k];
                        q[k] = orient_update<F>(fl, prm[k].y, prm[k].z, x[k], get3(sx, s + 1), d[k], dn[k], qp, q[k],
                                                qn, get4(sqr, (fl & HAS_PREV) ? s - 1 : s), get4(sqr, s));
                    }
                }
            }
            wave_sync();
        }

        // the tip's drag tangent is the previous segment's frame after the last sweep
        q4 qtip[NB];
        if (P.dragQuad > 0.f) {
#pragma unroll
            for (int k = 0; k < NB; k++)
                if (ok[k]) put4(ex, k * WAVE + lane, q[k]);
            wave_sync();
#pragma unroll
            for (int k = 0; k < NB; k++)
                qtip[k] = (ok[k] && !(flags_of(prm[k]) & HAS_NEXT)) ? get4(ex, k * WAVE + lane - 1) : q[k];
            wave_sync();
        }
        // end of substep: force readout (every lane, for the wave-wide wrench sums), then velocity and commit
        if (FORCES) {
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (k < nb) {
                    const bool on = ok[k] && prm[k].x != 0.f;
                    const float kap = Ar<F>::div(P.cKappa, h);
                    const f3 fv = surface_readout<F>(P, D, poses, rope, on, nc[k], cand[k], x[k], d[k],
                                                     on ? Ar<F>::div(kap * kap, prm[k].x) : 0.f);
                    if (on) {
                        float* fo = D.force + 3 * (size_t)(v0 + k * WAVE + lane);
                        fo[0] = fo[0] + fv.x;
                        fo[1] = fo[1] + fv.y;
                        fo[2] = fo[2] + fv.z;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
            if (ok[k]) {
                const uint32_t fl = flags_of(prm[k]);
                if (prm[k].x != 0.f) {
                    f3 tg = mk3(1.f, 0.f, 0.f);
                    if (P.dragQuad > 0.f) tg = V<F>::qrot_ex((fl & HAS_NEXT) ? q[k] : qtip[k]);
                    put3(sv, k * WAVE + lane, end_velocity<F>(P, d[k], tg));
                }
                x[k] = x[k] + d[k];
            }
        }
        // mesh broad phase for the next substep (YB/sim/contact.cuh:121-136); the 16 lanes of a chunk hold one rope
        if (P.numSdf > 0 && P.numSdf <= CHUNK_BROADPHASE_MIN_MESHES) {
#pragma unroll
            for (int k = 0; k < NB; k++)
                if (ok[k])
                    nc[k] = prm[k].x != 0.f ? build_candidates(P, poses, x[k], 2.f * len(d[k]) + 0.02f, cand[k]) : 0;
        } else if (P.numSdf > 0) {
#pragma unroll
            for (int k = 0; k < NB; k++) {
                if (k < nb) {
                    const bool act = ok[k] && prm[k].x != 0.f;
                    const float padv = 2.f * len(d[k]) + 0.02f;
                    const uint64_t mask = chunk_mesh_mask(P, poses, act, x[k], padv);
                    if (act)
                        nc[k] = build_candidates_masked(P, poses, x[k], padv, mask, cand[k]);
                    else if (ok[k])
                        nc[k] = 0;
                }
            }
        }
    }

    // write the state back, fill the views
#pragma unroll
    for (int k = 0; k < NB; k++) {
        if (ok[k]) {
            const int s = k * WAVE + lane;
            const int i = v0 + s;
            const f3 vel = get3(sv, s);
            D.x[i] = st3(x[k]);
            D.v[i] = st3(vel);
            D.vlast[i] = st3(get3(svl, s));
            D.q[i] = stq(q[k]);
            D.outPos[3 * i] = x[k].x;
            D.outPos[3 * i + 1] = x[k].y;
            D.outPos[3 * i + 2] = x[k].z;
            D.outVel[3 * i] = vel.x;
            D.outVel[3 * i + 1] = vel.y;
            D.outVel[3 * i + 2] = vel.z;
            D.outQuat[4 * i] = q[k].x;
            D.outQuat[4 * i + 1] = q[k].y;
            D.outQuat[4 * i + 2] = q[k].z;
            D.outQuat[4 * i + 3] = q[k].w;
            if (P.numSdf > 0) {
                D.ncand[i] = nc[k];
                for (int c = 0; c < MAX_CAND; c++)
                    D.cand[(size_t)i * MAX_CAND + c] = (nc[k] > 0 && c < nc[k]) ? cand_get(cand[k], c) : -1;
            }
        }
    }
}

template <int NB, bool FORCES, bool F>
static hipError_t launch_strand_f(const Params& P, const Dev& D, const Drive& Dr, const int* order, int nropes, int S,
                                  hipStream_t st) {
    const size_t lds = strand_lds_bytes(NB);
    hipLaunchKernelGGL((k_strand<NB, FORCES, F>), dim3(nropes), dim3(WAVE), lds, st, P, D, Dr, order, S);
    return hipGetLastError();
}

template <int NB>
static hipError_t launch_strand_t(const Params& P, const Dev& D, const Drive& Dr, const int* order, int nropes, int S,
                                  hipStream_t st) {
    const bool forces = P.forcesOn && (P.numPlanes > 0 || P.numSdf > 0);
    if (P.fastMath) {
        if (forces) return launch_strand_f<NB, true, true>(P, D, Dr, order, nropes, S, st);
        return launch_strand_f<NB, false, true>(P, D, Dr, order, nropes, S, st);
    }
    if (forces) return launch_strand_f<NB, true, false>(P, D, Dr, order, nropes, S, st);
    return launch_strand_f<NB, false, false>(P, D, Dr, order, nropes, S, st);
}

hipError_t launch_strand(int nb, const Params& P, const Dev& D, const Drive& Dr, const int* order, int nropes, int S,
                         hipStream_t st) {
    if (nropes <= 0) return hipSuccess;
    switch (nb) {
        case 1: return launch_strand_t<1>(P, D, Dr, order, nropes, S, st);
        case 2: return launch_strand_t<2>(P, D, Dr, order, nropes, S, st);
        case 4: return launch_strand_t<4>(P, D, Dr, order, nropes, S, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace rs
