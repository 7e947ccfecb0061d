// Host orchestration of the rope batch behind the C ABI (include/ropesim_hip.h): construction,
// device buffers, the mesh pose state machine, frame dispatch (per-substep kernels or the persistent
// frame kernel), validation and error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ropesim_hip.h"
#include "rs_adjoint.hpp"
#include "rs_kernels.hpp"
#include "rs_math.hpp"
#include "rs_types.hpp"

using namespace rs;

namespace {

thread_local std::string g_msg;
// Device memory for one frame's adjoint history: at least 2 GiB, up to 40 % of the free memory; larger batches
// are differentiated in chunks of ropes.
constexpr size_t kAdjointHistoryBudget = size_t(2) << 30;
// Contact groups above this many vertices detect rope-rope pairs through their hierarchy in the per-substep
// kernels (the full scan costs the square of the group size per substep; measured on MI355X the hierarchy is
// already 1.3x faster on one 256-vertex rope and 4x on 1024 groups of ~550 vertices).
constexpr int kBvhMinGroup = 128;
// Batches of up to this many cells order each frame's launch by the previous frame's cell durations (the
// ranking visits every pair of cells, so it stays well under a microsecond per cell up to this count).
constexpr int kRankMaxCells = 16384;

struct Invalid : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct DeviceErr : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct SolverErr : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) throw DeviceErr(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T> T* dalloc(size_t n, const char* what) {
    void* p = nullptr;
    if (n == 0) n = 1;
    hcheck(hipMalloc(&p, n * sizeof(T)), what);
    return static_cast<T*>(p);
}

template <class F> int guarded(F&& fn) {
    try {
        fn();
        return RS_OK;
    } catch (const Invalid& e) {
        g_msg = e.what();
        return RS_INVALID;
    } catch (const std::invalid_argument& e) {
        g_msg = e.what();
        return RS_INVALID;
    } catch (const SolverErr& e) {
        g_msg = e.what();
        return RS_SOLVER;
    } catch (const DeviceErr& e) {
        g_msg = e.what();
        return RS_DEVICE;
    } catch (const std::exception& e) {
        g_msg = e.what();
        return RS_DEVICE;
    }
}

float4 f4(f3 a, float w = 0.f) { return make_float4(a.x, a.y, a.z, w); }
float4 f4q(q4 a) { return make_float4(a.x, a.y, a.z, a.w); }

int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
    return n;
}
size_t device_max_lds() {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return 65536;
    return std::max<size_t>(p.sharedMemPerBlock, (size_t)p.maxSharedMemoryPerMultiProcessor > 0
                                                     ? std::min<size_t>(p.maxSharedMemoryPerMultiProcessor, 163840)
                                                     : 65536);
}

}  // namespace

struct rs_batch {
    std::mutex mtx;
    int N = 0, nv = 0;
    std::vector<int> roff, ropeOf;
    // host mirrors (authoritative before the first step; refreshed on demand afterwards)
    std::vector<float4> hx, hq, hprm, hqrest, hv;
    std::vector<float4> initX, initQ;
    // knobs (defaults: YB/YarnBall.cu:14-45)
    float gx = 0.f, gy = -9.8f, gz = 0.f;
    float drag = 0.2f, damping = 1e-6f, dragQuad = 0.f, dragQuadT = 0.f, accel = 1.f;
    int numItr = 8;
    bool fastMath = true;  // arithmetic of the iteration routines (Params::fastMath)
    float maxH = 1e-3f, h = 1e-3f, lastH = 0.f, time = 0.f;
    int numPlanes = 0;
    float planes[8][4] = {};
    float cRadius = 0.f, cKappa = 2.f, cMuS = 0.f, cMuK = 0.f, cVSlip = 1e-3f;
    bool ropeContact = false;
    int groupRopes = 1;
    float radius = 1e-4f, barrier = 8e-4f, kCollision = 1e-5f, detScaler = 1.2f, fricMu = 0.1f, kFriction = 5.f;
    float minSegLen = 0.f, maxSegLen = 0.f;
    bool stepped = false, fixRoot = false;
    int lastErrorCode = 0, lastWarningCode = 0;
    // device
    hipStream_t stream = nullptr;
    Dev D{};
    float4 *d_initX = nullptr, *d_initQ = nullptr;
    int* d_roff = nullptr;
    int* d_ids = nullptr;
    float* d_prog = nullptr;  // [17 N]
    // cells of the persistent frame kernel
    int mode = 1, variant = -1;
    int cellsFor = -2;
    std::vector<Cell> cells;
    Cell* d_cells = nullptr;
    int* d_order = nullptr;
    unsigned* d_cellTime = nullptr;  // [cells] last frame's duration of each cell (orders the next launch)
    int* d_cellQueue = nullptr;      // [1] the frame kernel's claim counter
    unsigned long long* d_hits = nullptr;  // rope-rope statistics [4]
    bool countHits = false;
    // rope-rope detection of the per-substep kernels: 0 the hierarchy for groups above kBvhMinGroup
    // vertices, 1 always the full scan, 2 always the hierarchy
    int broadPhase = 0;
    Bvh bvh;
    int bvhGroups = 0;
    // optional timing of the frame kernels on the solver stream (event pairs, summed on query)
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    // meshes (csrc/sdf.h)
    std::vector<Grid> grids;
    std::vector<float*> gridMem;
    Grid* d_grids = nullptr;
    Pose7 *d_cur = nullptr, *d_tgt = nullptr;
    Motion* d_motion = nullptr;
    float* d_stage = nullptr;
    uint8_t* d_stageTele = nullptr;
    bool pendingSet = false, moving = false, sdfMoving = false, forcesOn = false, meshesTouched = false;
    bool forcesFixed = false;  // readout wrenches summed as 64-bit fixed point (reproducible bit for bit)
    int lastSubsteps = 1;
    // differentiation: unit rest Darboux directions, the recorded rollout (per-frame checkpoints) and the
    // adjoint's buffers
    std::vector<float4> hRhoUnit;
    float4* d_rhoUnit = nullptr;
    struct Tape {
        bool on = false, closed = false;
        int maxFrames = 0, frames = 0, numSdf = 0, maxS = 0;
        int unitRopes = 1;         // ropes per adjoint unit (adjoint_unit_ropes at the first recorded frame)
        int* pairFlags = nullptr;  // [maxFrames][groups] each frame's per-group rope-rope pair flags (rope contact)
        int groups = 0;
        float4* ck = nullptr;      // [(maxFrames + 1)][4][nv]: x, v, v_last, q at each frame start (+ the end)
        float* prog = nullptr;     // [maxFrames][17 N] staged root programs
        Pose* poses = nullptr;     // [maxFrames][N][numSdf]
        Motion* motion = nullptr;  // [maxFrames][N][numSdf]
        std::vector<Params> P;
        std::vector<Drive> dr;
        std::vector<int> S;
    } tape;
    // per-substep inputs of the persistent kernel (root program, moving-mesh poses), sized for the frame
    size_t rootTrajCap = 0, posesSCap = 0, vlistCap = 0;
    void ensure_frame_buffers(int S) {
        const size_t nr = (size_t)S * N * 3;
        if (nr > rootTrajCap) {
            if (D.rootTraj) hipFree(const_cast<float4*>(D.rootTraj));
            D.rootTraj = dalloc<float4>(nr, "root trajectory");
            rootTrajCap = nr;
        }
        const size_t np = (size_t)S * N * grids.size();
        if (np > posesSCap) {
            if (D.posesS) hipFree(const_cast<Pose*>(D.posesS));
            D.posesS = dalloc<Pose>(np, "substep poses");
            posesSCap = np;
        }
    }
    // optional hipGraph form of the frame's kernel sequence
    bool graphOn = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<char> gkey;
    long long graphCaptures = 0;
    float4* d_adj = nullptr;      // [4][nv] state adjoint
    char* d_hist = nullptr;       // history of one frame (a chunk of ropes at a time)
    size_t histBytes = 0;
    double *d_gk = nullptr, *d_gm = nullptr, *d_grho = nullptr;
    float* d_gp = nullptr;
    double* d_gdrag = nullptr;
    int* d_mismatch = nullptr;
    int* d_blk = nullptr;  // the adjoint's workgroup -> rope ranges
    size_t blkCap = 0;
    uint16_t* d_adjLists = nullptr;  // the adjoint's rope-rope lists (adjoint_list_entries_per_vertex per vertex)
    hipStream_t adjStream = nullptr;  // the adjoint's second stream (groups with rope-rope pairs), default priority
    hipEvent_t adjFork = nullptr, adjJoin = nullptr;
    size_t adjListsCap = 0;

    ~rs_batch() {
        if (stream) hipStreamSynchronize(stream);
        drop_graph();
        for (auto& e : events) {
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
        void* ptrs[] = {D.x, D.v, D.vlast, D.q, D.q2, D.dx, D.dx2, D.prm, D.qrest, D.ropeOf, d_roff, D.pad, D.maxStep,
                        D.ncol, D.coll, D.vlist, D.clist, D.poses, (void*)D.motion, D.ncand, D.cand, D.force, D.wrench, D.wrenchFx, D.outPos,
                        D.outVel, D.outQuat, D.err, (void*)D.rootTraj, (void*)D.posesS, (void*)D.cellMap, d_initX, d_initQ, d_ids, d_prog, d_cells, d_order, d_cellTime, d_cellQueue, d_hits,
                        d_grids, d_cur, d_tgt, d_stage, d_stageTele, d_rhoUnit, tape.ck, tape.prog,
                        tape.poses, tape.motion, tape.pairFlags, d_adj, d_hist, d_gk, d_gm, d_gp, d_grho,
                        d_gdrag, d_mismatch, d_blk, d_adjLists, D.pairFlag};
        for (void* p : ptrs)
            if (p) hipFree(p);
        for (float* g : gridMem) hipFree(g);
        bvh_free(bvh);
        if (adjStream) hipStreamDestroy(adjStream);
        if (adjFork) hipEventDestroy(adjFork);
        if (adjJoin) hipEventDestroy(adjJoin);
        if (stream) hipStreamDestroy(stream);
    }

    // ---------------------------------------------------------------- construction
    // RopeBatch ctor + Sim::configure + setKStretch / setKBend (csrc/batch.cu:193-276,
    // YB/YarnBall.cu:112-259), in float32 like the reference's host code.
    void build(const rs_desc& d) {
        if (d.num_ropes < 1) throw Invalid("num_envs must be >= 1");
        if (!d.offsets || !d.positions) throw Invalid("offsets and positions are required");
        N = d.num_ropes;
        roff.assign(d.offsets, d.offsets + N + 1);
        if (roff[0] != 0) throw Invalid("offsets[0] must be 0");
        for (int r = 0; r < N; r++)
            if (roff[r + 1] - roff[r] < 3) throw Invalid("every rope needs >= 3 vertices");
        nv = roff[N];
        if (!(d.density > 0) || !(d.k_stretch > 0) || !(d.k_bend > 0) || !(d.max_h > 0) || d.num_itr < 1)
            throw Invalid("density, k_stretch, k_bend, max_h > 0 and num_itr >= 1 required");
        ropeOf.resize(nv);
        for (int r = 0; r < N; r++)
            for (int i = roff[r]; i < roff[r + 1]; i++) ropeOf[i] = r;
        std::vector<f3> pos(nv);
        for (int i = 0; i < nv; i++) {
            pos[i] = mk3(d.positions[3 * i], d.positions[3 * i + 1], d.positions[3 * i + 2]);
            if (!std::isfinite(pos[i].x) || !std::isfinite(pos[i].y) || !std::isfinite(pos[i].z))
                throw Invalid("positions must be finite");
        }
        for (int r = 0; r < N; r++)
            for (int i = roff[r]; i + 1 < roff[r + 1]; i++)
                if (!(len(pos[i + 1] - pos[i]) > 0.f)) throw Invalid("consecutive vertices must not coincide");
        std::vector<uint8_t> pin(nv, 0);
        if (d.pinned)
            for (int i = 0; i < nv; i++) pin[i] = d.pinned[i] ? 1 : 0;
        else
            for (int r = 0; r < N; r++) pin[roff[r]] = 1;
        fixRoot = d.fix_root_orientation != 0;
        if (fixRoot)
            for (int r = 0; r < N; r++)
                if (!pin[roff[r]]) throw Invalid("fix_root_orientation=True needs every root pinned");

        std::vector<uint32_t> flags(nv, HAS_NEXT);
        std::vector<float> invMass(nv, 1.f), lRest(nv, 1.f), kS(nv);
        std::vector<q4> q(nv), qRest(nv, q4{0.f, 0.f, 0.f, 5.f});
        for (int r = 0; r < N; r++) {
            flags[roff[r + 1] - 1] = 0;
            if (fixRoot) flags[roff[r]] |= FIX_ORIENT;
        }
        for (int i = 0; i < nv; i++)
            if (pin[i]) invMass[i] = 0.f;
        float minSeg = d.min_seg;
        if (!(minSeg > 0.f)) {
            minSeg = 1e30f;
            for (int r = 0; r < N; r++)
                for (int i = roff[r]; i + 1 < roff[r + 1]; i++) minSeg = std::min(minSeg, len(pos[i + 1] - pos[i]));
        }
        radius = 0.01f * minSeg;
        barrier = 0.08f * minSeg;

        maxSegLen = 0.f;
        minSegLen = 3.402823466e38f;
        q4 lastQ{0.f, 0.f, 0.f, 1.f};
        for (int i = 0; i < nv; i++) {
            if (i < nv - 1) {
                const bool hn = (flags[i] & HAS_NEXT) != 0;
                flags[i + 1] = hn ? (flags[i + 1] | HAS_PREV) : (flags[i + 1] & ~HAS_PREV);
                const bool hnn = hn && (flags[i + 1] & HAS_NEXT);
                flags[i] = hnn ? (flags[i] | HAS_NEXT_ORIENT) : (flags[i] & ~HAS_NEXT_ORIENT);
                if (!hn) flags[i] |= FIX_ORIENT;
            }
            lRest[i] = 1.f / float(nv);
            float mass = 0.f;
            if (flags[i] & HAS_PREV) mass = mass + lRest[i - 1];
            if (flags[i] & HAS_NEXT) {
                const f3 seg0 = pos[i + 1] - pos[i];
                lRest[i] = len(seg0);
                const f3 sd = seg0 * (1.f / sqrtf(dot(seg0, seg0)));
                q4 qq;
                if (sd.x < -0.999999f) {
                    qq = q4{0.f, 0.f, 1.f, 0.f};
                } else {  // fromTo(e_x, sd) (Rotor.h:57-64)
                    const f3 from = mk3(1.f, 0.f, 0.f);
                    f3 hh = (from + sd) * 0.5f;
                    const float l = len2(hh);
                    if (l > 0.f) {
                        hh = hh * (1.f / sqrtf(l));
                    } else {  // orthoBasisX(from)[1] (Common.h:244-259)
                        const f3 n = from;
                        if (n.z >= n.y) {
                            const float a = 1.f / (1.f + n.z), b = -n.x * n.y * a;
                            hh = mk3(b, 1.f - n.y * n.y * a, -n.y);
                        } else {
                            const float a = 1.f / (1.f + n.y), b = -n.x * n.z * a;
                            hh = mk3(1.f - n.x * n.x * a, -n.x, b);
                        }
                    }
                    qq = mkq(cross(from, hh), dot(from, hh));
                }
                q4 t = qmul(qconj(qq), lastQ);
                const q4 tw{t.x, 0.f, 0.f, t.w};
                t = (qdot(tw, tw) > 1e-12f) ? qnormalize(tw) : q4{0.f, 0.f, 0.f, 1.f};
                lastQ = qmul(qq, t);
                mass = mass + lRest[i];
                maxSegLen = std::max(maxSegLen, lRest[i]);
                minSegLen = std::min(minSegLen, lRest[i]);
            }
            q[i] = lastQ;
            mass = mass * (0.5f * d.density);
            invMass[i] = mass != 0.f ? invMass[i] * (1.f / mass) : 0.f;
        }
        for (int i = 0; i < nv - 1; i++) qRest[i] = qscale(qlen(qRest[i]), qmul(qconj(q[i]), q[i + 1]));
        if (d.straight_rest)
            for (int r = 0; r < N; r++)
                for (int i = roff[r]; i + 1 < roff[r + 1]; i++) qRest[i] = q4{0.f, 0.f, 0.f, 1.f};
        for (int i = 0; i < nv; i++) kS[i] = d.k_stretch * lRest[i];
        const float kb = d.k_bend * 4.f;
        for (int i = 0; i < nv; i++) qRest[i] = qscale(kb / lRest[i], qnormalize(qRest[i]));
        maxH = d.max_h;
        h = d.max_h;
        numItr = d.num_itr;

        hx.resize(nv);
        hq.resize(nv);
        hprm.resize(nv);
        hqrest.resize(nv);
        hv.assign(nv, make_float4(0.f, 0.f, 0.f, 0.f));
        for (int i = 0; i < nv; i++) {
            hx[i] = f4(pos[i]);
            hq[i] = f4q(q[i]);
            uint32_t fl = flags[i];
            float flf;
            std::memcpy(&flf, &fl, 4);
            hprm[i] = make_float4(invMass[i], lRest[i], kS[i], flf);
            hqrest[i] = f4q(qRest[i]);
        }
        initX = hx;
        initQ = hq;
        hRhoUnit.resize(nv);
        for (int i = 0; i < nv; i++) {
            const q4 u = qunit_or_identity(qRest[i]);
            hRhoUnit[i] = make_float4(u.x, u.y, u.z, u.w);
        }
        allocate();
        upload_all();
    }

    void allocate() {
        hcheck(hipStreamCreate(&stream), "hipStreamCreate");
        const size_t n = nv;
        D.x = dalloc<float4>(n, "x");
        D.v = dalloc<float4>(n, "v");
        D.vlast = dalloc<float4>(n, "vlast");
        D.q = dalloc<float4>(n, "q");
        D.q2 = dalloc<float4>(n, "q2");
        D.dx = dalloc<float4>(n, "dx");
        D.dx2 = dalloc<float4>(n, "dx2");
        D.prm = dalloc<float4>(n, "prm");
        D.qrest = dalloc<float4>(n, "qrest");
        D.ropeOf = dalloc<int>(n, "ropeOf");
        d_roff = dalloc<int>(N + 1, "roff");
        D.roff = d_roff;
        D.pad = dalloc<float>(n, "pad");
        D.maxStep = dalloc<float>(n, "maxStep");
        D.ncol = dalloc<int>(n, "ncol");
        D.outPos = dalloc<float>(3 * n, "outPos");
        D.outVel = dalloc<float>(3 * n, "outVel");
        D.outQuat = dalloc<float>(4 * n, "outQuat");
        D.err = dalloc<int>(2, "err");
        d_initX = dalloc<float4>(n, "initX");
        d_initQ = dalloc<float4>(n, "initQ");
        d_ids = dalloc<int>(N, "ids");
        d_prog = dalloc<float>(17 * (size_t)N, "prog");
        d_hits = dalloc<unsigned long long>(4, "stats");
        d_rhoUnit = dalloc<float4>(n, "rho unit");
        hcheck(hipMemset(D.err, 0, 2 * sizeof(int)), "memset err");
        hcheck(hipMemset(D.vlast, 0, n * sizeof(float4)), "memset vlast");
        hcheck(hipMemset(D.ncol, 0, n * sizeof(int)), "memset ncol");
        hcheck(hipMemset(d_hits, 0, 4 * sizeof(unsigned long long)), "memset stats");
    }

    void upload_all() {
        const size_t n = nv;
        hcheck(hipMemcpy(D.x, hx.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload x");
        hcheck(hipMemcpy(D.v, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload v");
        hcheck(hipMemcpy(D.q, hq.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload q");
        hcheck(hipMemcpy(D.prm, hprm.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload prm");
        hcheck(hipMemcpy(D.qrest, hqrest.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload qrest");
        hcheck(hipMemcpy(D.ropeOf, ropeOf.data(), n * sizeof(int), hipMemcpyHostToDevice), "upload ropeOf");
        hcheck(hipMemcpy(d_roff, roff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice), "upload roff");
        hcheck(hipMemcpy(d_initX, initX.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload initX");
        hcheck(hipMemcpy(d_initQ, initQ.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload initQ");
        hcheck(hipMemcpy(d_rhoUnit, hRhoUnit.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload rho unit");
        gather_nolock();
        hcheck(hipStreamSynchronize(stream), "construct");
    }

    // Host mirror of the dynamic state from the device.
    void download() {
        hcheck(hipStreamSynchronize(stream), "sync");
        const size_t n = nv;
        hcheck(hipMemcpy(hx.data(), D.x, n * sizeof(float4), hipMemcpyDeviceToHost), "download x");
        hcheck(hipMemcpy(hv.data(), D.v, n * sizeof(float4), hipMemcpyDeviceToHost), "download v");
        hcheck(hipMemcpy(hq.data(), D.q, n * sizeof(float4), hipMemcpyDeviceToHost), "download q");
    }

    // ---------------------------------------------------------------- parameters
    Params params() const {
        Params P{};
        P.nv = nv;
        P.nropes = N;
        P.numItr = numItr;
        P.h = h;
        P.lastH = lastH;
        P.gx = gx;
        P.gy = gy;
        P.gz = gz;
        P.damping = damping;
        P.drag = drag;
        P.dragQuad = dragQuad;
        P.dragQuadT = dragQuadT;
        P.accel = accel;
        P.numPlanes = numPlanes;
        std::memcpy(P.planes, planes, sizeof(planes));
        P.cRadius = cRadius;
        P.cKappa = cKappa;
        P.cMuS = cMuS;
        P.cMuK = cMuK;
        P.cVSlip = cVSlip;
        P.numSdf = (int)grids.size();
        P.sdfMoving = sdfMoving ? 1 : 0;
        P.forcesOn = forcesOn ? 1 : 0;
        P.forcesFixed = forcesFixed ? 1 : 0;
        P.ropeContact = ropeContact ? 1 : 0;
        P.groupRopes = groupRopes;
        P.radius = radius;
        P.barrier = barrier;
        P.kCollision = kCollision;
        P.fricMu = fricMu;
        P.kFriction = kFriction;
        P.detR = radius + 0.5f * barrier;  // uploadMeta (YB/YarnBall.cu:261-272)
        P.scaledDetR = P.detR * detScaler;
        P.fastMath = fastMath ? 1 : 0;
        return P;
    }

    // ---------------------------------------------------------------- meshes
    void ensure_mesh_state() {
        if (meshesTouched) return;
        meshesTouched = true;
        const size_t total = (size_t)N * MAX_SDF;
        d_grids = dalloc<Grid>(MAX_SDF, "grids");
        d_cur = dalloc<Pose7>(total, "sdf cur");
        d_tgt = dalloc<Pose7>(total, "sdf tgt");
        D.motion = dalloc<Motion>(total, "sdf motion");
        D.poses = dalloc<Pose>(total, "sdf poses");
        d_stage = dalloc<float>(total * 7, "sdf stage");
        d_stageTele = dalloc<uint8_t>(N, "sdf stage tele");
        D.ncand = dalloc<int>(nv, "ncand");
        D.cand = dalloc<int>((size_t)nv * MAX_CAND, "cand");
        D.force = dalloc<float>(3 * (size_t)nv, "contact force");
        D.wrench = dalloc<float>(total * 6, "wrench");
        D.grids = d_grids;
        std::vector<Pose7> ident(total, Pose7{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f});
        hcheck(hipMemcpy(d_cur, ident.data(), total * sizeof(Pose7), hipMemcpyHostToDevice), "upload sdf cur");
        hcheck(hipMemcpy(d_tgt, ident.data(), total * sizeof(Pose7), hipMemcpyHostToDevice), "upload sdf tgt");
        std::vector<Motion> still(total, Motion{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f});
        hcheck(hipMemcpy(const_cast<Motion*>(D.motion), still.data(), total * sizeof(Motion), hipMemcpyHostToDevice),
               "upload sdf motion");
        hcheck(hipMemset(D.poses, 0, total * sizeof(Pose)), "memset poses");
        hcheck(hipMemset(D.ncand, 0xFF, nv * sizeof(int)), "memset ncand");
        hcheck(hipMemset(D.force, 0, 3 * nv * sizeof(float)), "memset force");
        hcheck(hipMemset(D.wrench, 0, total * 6 * sizeof(float)), "memset wrench");
    }
    MeshSlots slots() { return MeshSlots{d_cur, d_tgt, const_cast<Motion*>(D.motion)}; }
    void upload_grids() {
        std::vector<Grid> all(MAX_SDF);
        std::memset(all.data(), 0, sizeof(Grid) * MAX_SDF);
        for (size_t i = 0; i < grids.size(); i++) all[i] = grids[i];
        hcheck(hipMemcpyAsync(d_grids, all.data(), sizeof(Grid) * MAX_SDF, hipMemcpyHostToDevice, stream),
               "upload grids");
        hcheck(hipStreamSynchronize(stream), "upload grids");
    }
    void set_targets_dev(int index, const float* pos, const float* q, int rows, const uint8_t* tele, bool teleAll) {
        hcheck(launch_mesh_targets(slots(), index, pos, q, rows, (int)grids.size(), tele, teleAll ? 1 : 0, N, stream),
               "set_mesh_poses");
        pendingSet = true;
    }
    void set_targets_host(int index, const float* pos, const float* q, int rows, const uint8_t* tele, bool teleAll) {
        const int n = index < 0 ? N * (int)grids.size() : rows;
        float* dq = d_stage + 3 * (size_t)N * MAX_SDF;
        hcheck(hipMemcpyAsync(d_stage, pos, sizeof(float) * 3 * n, hipMemcpyHostToDevice, stream), "upload target pos");
        hcheck(hipMemcpyAsync(dq, q, sizeof(float) * 4 * n, hipMemcpyHostToDevice, stream), "upload target quat");
        const uint8_t* tf = nullptr;
        if (tele && !teleAll) {
            hcheck(hipMemcpyAsync(d_stageTele, tele, N, hipMemcpyHostToDevice, stream), "upload teleport flags");
            tf = d_stageTele;
        }
        set_targets_dev(index, d_stage, dq, rows, tf, teleAll);
        hcheck(hipStreamSynchronize(stream), "set_mesh_poses");  // the host arrays may die on return
    }
    void begin_step() {  // SdfStore::begin_step (csrc/sdf.cu:217-232)
        if (!meshesTouched) return;
        // the motions were written with the targets; without a target call this step the meshes stay put
        sdfMoving = pendingSet;
        if (pendingSet) moving = true;
        if (!grids.empty()) hcheck(launch_fill_i32(D.ncand, -1, nv, stream), "reset candidates");
        if (forcesOn) {
            hcheck(launch_fill_f32(D.force, 0.f, 3L * nv, stream), "zero forces");
            if (forcesFixed)
                hcheck(launch_fill_i32(reinterpret_cast<int*>(D.wrenchFx), 0, 12L * N * MAX_SDF, stream),
                       "zero wrench accumulators");
            else
                hcheck(launch_fill_f32(D.wrench, 0.f, 6L * N * MAX_SDF, stream), "zero wrenches");
        }
    }
    void end_step(int substeps) {  // SdfStore::end_step (csrc/sdf.cu:239-253)
        lastSubsteps = substeps > 0 ? substeps : 1;
        if (!meshesTouched) return;
        if (moving) {
            hcheck(launch_mesh_commit(slots(), D.poses, d_grids, (int)grids.size(), -1, N, stream), "commit");
            moving = false;
            pendingSet = false;
        }
        if (forcesOn) {
            const float k = 1.f / (float)lastSubsteps;
            hcheck(launch_scale_f32(D.force, k, 3L * nv, stream), "average forces");
            if (forcesFixed)
                hcheck(launch_wrench_from_fixed(D.wrenchFx, D.wrench, k, 6L * N * MAX_SDF, stream), "average wrenches");
            else
                hcheck(launch_scale_f32(D.wrench, k, 6L * N * MAX_SDF, stream), "average wrenches");
        }
    }

    // ---------------------------------------------------------------- cells
    // Cells of the persistent kernel: whole units (a contact group, or a rope without rope-rope contact) packed
    // best-fit-decreasing into cells of the variant's capacity, so the fewest workgroups carry the batch. A cell's
    // units are laid out in vertex order; cellMap gives the vertex of every slot.
    // The strand kernel takes one rope per wave: a cell per rope, launched longest first.
    void build_strand_order(int var) {
        cells.assign(N, Cell{0, 0, 1, 0});
        std::vector<int> order(N);
        for (int r = 0; r < N; r++) {
            cells[r] = Cell{roff[r], roff[r + 1] - roff[r], 1, 0};
            order[r] = r;
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a].nv > cells[b].nv; });
        if (d_cells) hipFree(d_cells);
        if (d_order) hipFree(d_order);
        D.cellTime = nullptr;  // the strand kernel keeps its rope order and one workgroup per rope
        D.cellQueue = nullptr;
        d_cells = dalloc<Cell>(N, "cells");
        d_order = dalloc<int>(N, "rope order");
        hcheck(hipMemcpy(d_cells, cells.data(), N * sizeof(Cell), hipMemcpyHostToDevice), "upload cells");
        hcheck(hipMemcpy(d_order, order.data(), N * sizeof(int), hipMemcpyHostToDevice), "upload order");
        cellsFor = var;
    }
    void build_cells(int var) {
        if (kCellVariants[var].strand) return build_strand_order(var);
        const int cap = kCellVariants[var].block * kCellVariants[var].vpt;
        const int G = ropeContact ? groupRopes : 1;
        struct Unit {
            int v0, nv;
        };
        std::vector<Unit> units;
        for (int g0 = 0; g0 < N; g0 += G) {
            const int g1 = std::min(N, g0 + G);
            const Unit u{roff[g0], roff[g1] - roff[g0]};
            if (u.nv > cap) throw Invalid("a rope / contact group does not fit the cell");
            units.push_back(u);
        }
        std::vector<int> byNv(units.size());
        for (size_t i = 0; i < byNv.size(); i++) byNv[i] = (int)i;
        std::stable_sort(byNv.begin(), byNv.end(), [&](int a, int b) { return units[a].nv > units[b].nv; });
        std::vector<std::vector<int>> members;
        std::vector<int> used;
        std::multimap<int, int> freeCap;  // free slots -> cell
        for (int u : byNv) {
            const int need = units[u].nv;
            auto it = freeCap.lower_bound(need);
            int c;
            if (it == freeCap.end()) {
                c = (int)members.size();
                members.emplace_back();
                used.push_back(0);
            } else {
                c = it->second;
                freeCap.erase(it);
            }
            members[c].push_back(u);
            used[c] += need;
            if (cap - used[c] > 0) freeCap.emplace(cap - used[c], c);
        }
        const int nc = (int)members.size();
        cells.assign(nc, Cell{0, 0, 0, 0});
        std::vector<int> map((size_t)nc * cap, -1);
        for (int c = 0; c < nc; c++) {
            std::sort(members[c].begin(), members[c].end());
            int slot = 0;
            for (int u : members[c])
                for (int k = 0; k < units[u].nv; k++) map[(size_t)c * cap + slot++] = units[u].v0 + k;
            cells[c] = Cell{c * cap, slot, (int)members[c].size(), 0};
        }
        std::vector<int> order(nc);
        for (int i = 0; i < nc; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a].nv > cells[b].nv; });
        if (d_cells) hipFree(d_cells);
        if (d_order) hipFree(d_order);
        if (D.cellMap) hipFree(const_cast<int*>(D.cellMap));
        d_cells = dalloc<Cell>(nc, "cells");
        d_order = dalloc<int>(nc, "cell order");
        int* dm = dalloc<int>(map.size(), "cell map");
        D.cellMap = dm;
        hcheck(hipMemcpy(d_cells, cells.data(), nc * sizeof(Cell), hipMemcpyHostToDevice), "upload cells");
        hcheck(hipMemcpy(d_order, order.data(), nc * sizeof(int), hipMemcpyHostToDevice), "upload order");
        hcheck(hipMemcpy(dm, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice), "upload cell map");
        // the first frame runs the cells largest first; every later frame longest first by measured duration
        if (d_cellTime) hipFree(d_cellTime);
        d_cellTime = nullptr;
        D.cellTime = nullptr;
        if (nc <= kRankMaxCells) {
            d_cellTime = dalloc<unsigned>(nc, "cell durations");
            hcheck(hipMemset(d_cellTime, 0, nc * sizeof(unsigned)), "clear cell durations");
            D.cellTime = d_cellTime;
        }
        if (!d_cellQueue) d_cellQueue = dalloc<int>(1, "cell queue");
        D.cellQueue = d_cellQueue;
        // the persistent kernel's skin and candidate lists are indexed by cell slot
        if (ropeContact && map.size() > vlistCap) {
            if (D.vlist) hipFree(D.vlist);
            if (D.clist) hipFree(D.clist);
            D.vlist = nullptr;
            D.clist = nullptr;
            D.vlist = dalloc<uint16_t>(map.size() * VERLET_CAP, "skin lists");
            D.clist = dalloc<uint16_t>(map.size() * MAX_COLL, "cell candidate lists");
            vlistCap = map.size();
        }
        cellsFor = var;
    }
    bool tree_detect() const {
        return ropeContact && (broadPhase == 2 || (broadPhase == 0 && max_unit() > kBvhMinGroup));
    }
    // The hierarchy when this batch's detection uses it (allocated for the current vertex and group counts).
    const Bvh* detect_tree() {
        if (!tree_detect()) return nullptr;
        const int groups = (N + groupRopes - 1) / groupRopes;
        if (bvh.n != (int)nv || bvhGroups != groups) {
            hcheck(bvh_alloc(bvh, (int)nv, groups), "rope-rope hierarchy");
            bvhGroups = groups;
        }
        return &bvh;
    }
    int max_unit() const {
        const int G = ropeContact ? groupRopes : 1;
        int m = 0;
        for (int g0 = 0; g0 < N; g0 += G) m = std::max(m, roff[std::min(N, g0 + G)] - roff[g0]);
        return m;
    }
    bool variant_supported(int var) const {
        const CellVariant& cv = kCellVariants[var];
        if (variant_lds_bytes(var) > device_max_lds()) return false;
        if (cv.strand) {  // one rope per wave, no rope-rope contact
            if (ropeContact) return false;
            int m = 0;
            for (int r = 0; r < N; r++) m = std::max(m, roff[r + 1] - roff[r]);
            return m <= 64 * cv.vpt;
        }
        return cv.block * cv.vpt >= max_unit();
    }
    int default_variant() const {
        // the strand kernel with the fewest blocks per lane that holds every rope, else the smallest supported
        // cell that still fills the machine with work; the autotuner refines it
        for (int v = 0; v < kNumCellVariants; v++)
            if (kCellVariants[v].strand && variant_supported(v)) return v;
        int best = -1;
        for (int v = 0; v < kNumCellVariants; v++) {
            if (!variant_supported(v)) continue;
            if (kCellVariants[v].strand || kCellVariants[v].vpt != 1) continue;
            best = v;
            break;
        }
        if (best < 0)
            for (int v = 0; v < kNumCellVariants; v++)
                if (variant_supported(v)) return v;
        return best;
    }

    // ---------------------------------------------------------------- frames
    void prepare(float hs) {  // SimAccess::prepare + uploadMeta checks (csrc/batch.cu:57-62, YB/YarnBall.cu:261-272)
        if (!(detScaler > 1.f)) throw SolverErr("Detection scaler must be greater than 1.");
        if (2.f * minSegLen < 2.f * radius + barrier)
            throw SolverErr("Use thinner yarn or use longer segments. (2 * Min seg length must be at least 2 * radius + barrierThickness");
        lastH = h;
        h = hs;
    }
    void stage_inputs(const Drive& in, Drive& out) {
        out = in;
        float* P0 = d_prog;
        float* V0 = P0 + 3 * N;
        float* P1 = V0 + 3 * N;
        float* Q0 = P1 + 3 * N;
        float* Q1 = Q0 + 4 * N;
        if (in.p0) hcheck(hipMemcpyAsync(P0, in.p0, 12 * (size_t)N, hipMemcpyDeviceToDevice, stream), "stage p0");
        if (in.v0) hcheck(hipMemcpyAsync(V0, in.v0, 12 * (size_t)N, hipMemcpyDeviceToDevice, stream), "stage v0");
        if (in.p1) hcheck(hipMemcpyAsync(P1, in.p1, 12 * (size_t)N, hipMemcpyDeviceToDevice, stream), "stage p1");
        if (in.q0) hcheck(hipMemcpyAsync(Q0, in.q0, 16 * (size_t)N, hipMemcpyDeviceToDevice, stream), "stage q0");
        if (in.q1) hcheck(hipMemcpyAsync(Q1, in.q1, 16 * (size_t)N, hipMemcpyDeviceToDevice, stream), "stage q1");
        out.p0 = P0;
        out.v0 = V0;
        out.p1 = P1;
        out.q0 = Q0;
        out.q1 = Q1;
        out.hasQ0 = in.q0 != nullptr;
        out.hasQ1 = in.q1 != nullptr;
    }
    // The frame's kernel sequence: per-substep launches (var < 0) or the persistent frame kernel variant var.
    void enqueue_frame(const Params& P, const Drive& dr, int S, int var, const Bvh* tree) {
        if (P.ropeContact)  // which contact groups have rope-rope pairs this frame (read by the adjoint)
            hcheck(launch_fill_i32(D.pairFlag, 0, (N + groupRopes - 1) / groupRopes, stream), "clear pair flags");
        if (var < 0) {
            if (dr.mode == DRIVE_CONST) hcheck(launch_root_const(P, D, dr, stream), "root drive");
            for (int s = 0; s < S; s++) hcheck(launch_plain_substep(P, D, dr, s, S, stream, tree), "substep");
            hcheck(launch_gather(P, D, stream), "gather");
        } else {
            hcheck(launch_cell(var, P, D, dr, d_cells, d_order, (int)cells.size(), S, countHits ? d_hits : nullptr,
                               stream),
                   "frame kernel");
            if (D.cellTime && !kCellVariants[var].strand)
                hcheck(launch_rank_cells(D.cellTime, (int)cells.size(), d_order, stream), "cell order");
        }
    }
    void run_substeps(const Drive& dr, int S) {
        const Params P = params();
        // mode 1 runs the persistent kernel unless no cell holds the largest rope or contact group: then, like
        // mode 0, the per-substep kernels
        int var = -1;
        if (mode == 1) {
            var = variant;
            if (var < 0 || !variant_supported(var)) var = default_variant();
            if (var >= 0) {
                if (cellsFor != var) build_cells(var);
                ensure_frame_buffers(S);
            }
        }
        const Bvh* tree = var < 0 ? detect_tree() : nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timing && var >= 0) {
            hcheck(hipEventCreate(&e0), "event");
            hcheck(hipEventCreate(&e1), "event");
            hcheck(hipEventRecord(e0, stream), "event");
        }
        if (graphOn)
            launch_graph(P, dr, S, var, tree);
        else
            enqueue_frame(P, dr, S, var, tree);
        if (timing && var >= 0) {
            hcheck(hipEventRecord(e1, stream), "event");
            events.emplace_back(e0, e1);
        }
    }
    // Graph form of the frame: the kernel sequence is captured once into a hipGraph and replayed while
    // every argument it was captured with (parameters, buffers, staged drive, substeps, dispatch) is the same.
    void launch_graph(const Params& P, const Drive& dr, int S, int var, const Bvh* tree) {
        std::vector<char> key(sizeof(Params) + sizeof(Dev) + sizeof(Drive) + 8 * sizeof(long long));
        char* k = key.data();
        std::memcpy(k, &P, sizeof(Params));
        std::memcpy(k + sizeof(Params), &D, sizeof(Dev));
        std::memcpy(k + sizeof(Params) + sizeof(Dev), &dr, sizeof(Drive));
        const long long tail[8] = {mode, var, S, (long long)(uintptr_t)d_cells, (long long)cells.size(),
                                   countHits ? 1 : 0, (long long)(uintptr_t)(tree ? tree->key : nullptr),
                                   (long long)(uintptr_t)d_order};
        std::memcpy(k + sizeof(Params) + sizeof(Dev) + sizeof(Drive), tail, sizeof(tail));
        if (!gexec || key != gkey) {
            drop_graph();
            hcheck(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "begin capture");
            try {
                enqueue_frame(P, dr, S, var, tree);
            } catch (...) {
                hipGraph_t g = nullptr;
                hipStreamEndCapture(stream, &g);
                if (g) hipGraphDestroy(g);
                throw;
            }
            hcheck(hipStreamEndCapture(stream, &graph), "end capture");
            hcheck(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0), "instantiate graph");
            gkey.swap(key);
            graphCaptures++;
        }
        hcheck(hipGraphLaunch(gexec, stream), "graph launch");
    }
    void drop_graph() {
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        gkey.clear();
    }
    void frame(const Drive& in, float dt, int S) {
        if (!(dt > 0.f)) throw Invalid("dt must be > 0");
        if (S < 1) throw Invalid("substeps must be >= 1");
        Drive dr;
        stage_inputs(in, dr);
        dr.dt = dt;
        begin_step();
        prepare(dt / float(S));
        if (tape.on) tape_record(dr, S);
        stepped = true;
        run_substeps(dr, S);
        if (tape.on && tape.pairFlags && ropeContact && tape.groups == (N + groupRopes - 1) / groupRopes)
            hcheck(hipMemcpyAsync(tape.pairFlags + (size_t)(tape.frames - 1) * tape.groups, D.pairFlag,
                                  tape.groups * sizeof(int), hipMemcpyDeviceToDevice, stream),
                   "record pair flags");
        time = time + dt;
        check_errors();
        end_step(S);
    }
    void check_errors() {  // Sim::checkErrors (YB/YarnBall.cu:311-339): errors come only from rope-rope detection
        hcheck(hipGetLastError(), "launch");
        if (!ropeContact) return;
        int e[2] = {0, 0};
        hcheck(hipMemcpyAsync(e, D.err, sizeof(e), hipMemcpyDeviceToHost, stream), "read error word");
        hcheck(hipStreamSynchronize(stream), "frame");
        if (e[0] != 0 || e[1] != 0) hcheck(hipMemsetAsync(D.err, 0, sizeof(e), stream), "reset error word");
        if (e[1] != 0) lastWarningCode = e[1];
        if (e[0] != 0) {
            lastErrorCode = e[0];
            throw SolverErr(e[0] == ERR_MAX_COLLISIONS ? "MAX_COLLISIONS_PER_SEGMENT exceeded"
                            : e[0] == ERR_BVH_DEPTH    ? "rope-rope hierarchy deeper than its traversal stack"
                                                       : "Indescript error");
        }
    }
    void gather_nolock() { hcheck(launch_gather(params(), D, stream), "gather"); }

    // ---------------------------------------------------------------- differentiation
    // Ropes the adjoint keeps in one workgroup: one rope, or a whole contact group while rope-rope contact is on.
    int adjoint_unit_ropes() const { return ropeContact ? groupRopes : 1; }
    void tape_check() const {
        if (dragQuad > 0.f) throw Invalid("the adjoint does not cover quadratic drag; set drag_quad = 0 while recording");
        if ((int)grids.size() != tape.numSdf) throw Invalid("meshes were added or cleared while recording");
        if (tape.frames > 0 && adjoint_unit_ropes() != tape.unitRopes)
            throw Invalid("rope-rope contact or its group size changed while recording");
        const int G = adjoint_unit_ropes();
        int maxV = 0;
        for (int r = 0; r < N; r += G) maxV = std::max(maxV, roff[std::min(r + G, N)] - roff[r]);
        if (maxV > ADJ_MAX_BLOCK)
            throw Invalid(ropeContact ? "the adjoint supports rope contact groups of at most 512 vertices"
                                      : "the adjoint supports ropes of at most 512 vertices");
    }
    void copy_state(float4* dst) {
        const size_t n = nv;
        hcheck(hipMemcpyAsync(dst, D.x, n * 16, hipMemcpyDeviceToDevice, stream), "checkpoint");
        hcheck(hipMemcpyAsync(dst + n, D.v, n * 16, hipMemcpyDeviceToDevice, stream), "checkpoint");
        hcheck(hipMemcpyAsync(dst + 2 * n, D.vlast, n * 16, hipMemcpyDeviceToDevice, stream), "checkpoint");
        hcheck(hipMemcpyAsync(dst + 3 * n, D.q, n * 16, hipMemcpyDeviceToDevice, stream), "checkpoint");
    }
    void tape_record(const Drive& dr, int S) {
        tape_check();
        if (tape.frames >= tape.maxFrames) throw Invalid("the tape is full (max_frames of rs_tape_begin)");
        const int f = tape.frames;
        if (f == 0) tape.unitRopes = adjoint_unit_ropes();
        copy_state(tape.ck + (size_t)f * 4 * nv);
        hcheck(hipMemcpyAsync(tape.prog + (size_t)f * 17 * N, d_prog, 17 * (size_t)N * 4, hipMemcpyDeviceToDevice,
                              stream),
               "record program");
        const int M = tape.numSdf;
        if (M > 0) {
            hcheck(hipMemcpy2DAsync(tape.poses + (size_t)f * N * M, M * sizeof(Pose), D.poses, MAX_SDF * sizeof(Pose),
                                    M * sizeof(Pose), N, hipMemcpyDeviceToDevice, stream),
                   "record poses");
            if (sdfMoving)
                hcheck(hipMemcpy2DAsync(tape.motion + (size_t)f * N * M, M * sizeof(Motion), D.motion,
                                        MAX_SDF * sizeof(Motion), M * sizeof(Motion), N, hipMemcpyDeviceToDevice,
                                        stream),
                       "record motion");
        }
        tape.P.push_back(params());
        tape.dr.push_back(dr);
        tape.S.push_back(S);
        tape.maxS = std::max(tape.maxS, S);
        tape.frames++;
    }
    void tape_close() {
        if (!tape.on) return;
        copy_state(tape.ck + (size_t)tape.frames * 4 * nv);  // the state after the last recorded frame
        tape.on = false;
        tape.closed = true;
    }
    void tape_free() {
        for (void* p : {(void*)tape.ck, (void*)tape.prog, (void*)tape.poses, (void*)tape.motion, (void*)tape.pairFlags})
            if (p) hipFree(p);
        tape = Tape{};
    }
};

// ==================================================================== C ABI

void rs::set_error_message(const char* msg) { g_msg = msg; }

extern "C" {

const char* rs_last_error_message(void) { return g_msg.c_str(); }
int rs_abi_version(void) { return 1; }

int rs_create(const rs_desc* desc, rs_batch** out) {
    if (!desc || !out) {
        g_msg = "null argument";
        return RS_INVALID;
    }
    *out = nullptr;
    rs_batch* b = new rs_batch();
    const int rc = guarded([&] { b->build(*desc); });
    if (rc != RS_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return RS_OK;
}

void rs_destroy(rs_batch* b) { delete b; }
int rs_num_vertices(const rs_batch* b) { return b ? b->nv : 0; }
int rs_num_ropes(const rs_batch* b) { return b ? b->N : 0; }
int rs_offsets(const rs_batch* b, int* out) {
    return guarded([&] { std::copy(b->roff.begin(), b->roff.end(), out); });
}

int rs_get_param(const rs_batch* b, int which, double* out) {
    return guarded([&] {
        switch (which) {
            case RS_P_DRAG: *out = b->drag; break;
            case RS_P_DAMPING: *out = b->damping; break;
            case RS_P_DRAG_QUAD: *out = b->dragQuad; break;
            case RS_P_DRAG_QUAD_T: *out = b->dragQuadT; break;
            case RS_P_NUM_ITR: *out = b->numItr; break;
            case RS_P_ACCEL: *out = b->accel; break;
            case RS_P_MAX_H: *out = b->maxH; break;
            case RS_P_TIME: *out = b->time; break;
            case RS_P_STEPPED: *out = b->stepped ? 1 : 0; break;
            default: throw Invalid("unknown parameter");
        }
    });
}

int rs_set_param(rs_batch* b, int which, double value) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        const float v = (float)value;
        switch (which) {
            case RS_P_DRAG: b->drag = v; break;
            case RS_P_DAMPING: b->damping = v; break;
            case RS_P_DRAG_QUAD:
                if (!std::isfinite(v) || v < 0) throw Invalid("drag_quad must be finite and >= 0");
                b->dragQuad = v;
                break;
            case RS_P_DRAG_QUAD_T:
                if (!std::isfinite(v) || v < 0) throw Invalid("drag_quad_t must be finite and >= 0");
                b->dragQuadT = v;
                break;
            case RS_P_NUM_ITR:
                if ((int)value < 1) throw Invalid("num_itr must be >= 1");
                b->numItr = (int)value;
                break;
            case RS_P_ACCEL: b->accel = v; break;
            case RS_P_MAX_H:
                if (!(v > 0)) throw Invalid("max_h must be > 0");
                b->maxH = v;
                break;
            default: throw Invalid("parameter is read only or unknown");
        }
    });
}

int rs_set_gravity(rs_batch* b, const float* g3) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->gx = g3[0];
        b->gy = g3[1];
        b->gz = g3[2];
    });
}
int rs_get_gravity(const rs_batch* b, float* g3) {
    g3[0] = b->gx;
    g3[1] = b->gy;
    g3[2] = b->gz;
    return RS_OK;
}

int rs_get_vertex(rs_batch* b, int which, float* out) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        const int n = b->nv;
        if (which == RS_V_POSITIONS || which == RS_V_VELOCITIES || which == RS_V_QUATERNIONS ||
            which == RS_V_LAST_VELOCITIES) {
            b->download();
            if (which == RS_V_LAST_VELOCITIES) {
                std::vector<float4> t(n);
                hcheck(hipMemcpy(t.data(), b->D.vlast, n * sizeof(float4), hipMemcpyDeviceToHost), "download vlast");
                for (int i = 0; i < n; i++) {
                    out[3 * i] = t[i].x;
                    out[3 * i + 1] = t[i].y;
                    out[3 * i + 2] = t[i].z;
                }
                return;
            }
            const std::vector<float4>& src = which == RS_V_POSITIONS ? b->hx : (which == RS_V_VELOCITIES ? b->hv : b->hq);
            const int w = which == RS_V_QUATERNIONS ? 4 : 3;
            for (int i = 0; i < n; i++) {
                out[w * i] = src[i].x;
                out[w * i + 1] = src[i].y;
                out[w * i + 2] = src[i].z;
                if (w == 4) out[w * i + 3] = src[i].w;
            }
        } else if (which == RS_V_INV_MASSES) {
            for (int i = 0; i < n; i++) out[i] = b->hprm[i].x;
        } else if (which == RS_V_REST_LENGTHS) {
            for (int i = 0; i < n; i++) out[i] = b->hprm[i].y;
        } else {
            throw Invalid("field is not readable");
        }
    });
}

int rs_set_vertex(rs_batch* b, int which, const float* in) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        const int n = b->nv;
        if (which == RS_V_VELOCITIES) {  // set_velocities (csrc/bindings.cpp:1000-1013)
            for (int i = 0; i < 3 * n; i++)
                if (!std::isfinite(in[i])) throw Invalid("velocities must be finite");
            b->download();
            for (int i = 0; i < n; i++) b->hv[i] = make_float4(in[3 * i], in[3 * i + 1], in[3 * i + 2], 0.f);
            hcheck(hipMemcpy(b->D.v, b->hv.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload v");
            hcheck(launch_gather(b->params(), b->D, b->stream), "gather");
            return;
        }
        if (which == RS_V_INV_MASSES) {  // raw inverse masses (RopeSim.unpin_root, csrc/bindings.cpp:520-526)
            for (int i = 0; i < n; i++)
                if (!std::isfinite(in[i]) || in[i] < 0) throw Invalid("inverse masses must be finite and >= 0");
            for (int i = 0; i < n; i++) b->hprm[i].x = in[i];
            hcheck(hipMemcpy(b->D.prm, b->hprm.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload prm");
            return;
        }
        if (which == RS_V_MASSES || which == RS_V_EA || which == RS_V_EI) {
            for (int r = 0; r < b->N; r++)
                for (int i = b->roff[r]; i < b->roff[r + 1]; i++) {
                    const bool seg = i + 1 < b->roff[r + 1];
                    if (which == RS_V_MASSES) {
                        if (b->hprm[i].x != 0.f && (!std::isfinite(in[i]) || !(in[i] > 0)))
                            throw Invalid("node masses must be finite and > 0 (pinned columns ignored)");
                    } else if (seg && (!std::isfinite(in[i]) || !(in[i] > 0))) {
                        throw Invalid(which == RS_V_EA ? "EA values must be finite and > 0"
                                                       : "EI values must be finite and > 0");
                    }
                }
            for (int r = 0; r < b->N; r++)
                for (int i = b->roff[r]; i < b->roff[r + 1]; i++) {
                    const bool seg = i + 1 < b->roff[r + 1];
                    float4& p = b->hprm[i];
                    if (which == RS_V_MASSES) {
                        if (p.x != 0.f) p.x = 1.f / in[i];
                    } else if (which == RS_V_EA) {
                        if (seg) p.z = in[i] * p.y;
                    } else if (seg) {
                        const float k = 4.f * in[i] / p.y;
                        const float4 qr = b->hqrest[i];
                        const q4 nq = qscale(k, qnormalize(q4{qr.x, qr.y, qr.z, qr.w}));
                        b->hqrest[i] = make_float4(nq.x, nq.y, nq.z, nq.w);
                    }
                }
            hcheck(hipMemcpy(b->D.prm, b->hprm.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload prm");
            hcheck(hipMemcpy(b->D.qrest, b->hqrest.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload qrest");
            return;
        }
        throw Invalid("field is not settable");
    });
}

int rs_set_root_host(rs_batch* b, int rope, const float* pos3, const float* vel3, const float* q) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (rope < 0 || rope >= b->N) throw Invalid("rope index out of range");
        const int i = b->roff[rope];
        hcheck(hipStreamSynchronize(b->stream), "sync");
        const float4 x = make_float4(pos3[0], pos3[1], pos3[2], 0.f);
        const float4 v = make_float4(vel3[0], vel3[1], vel3[2], 0.f);
        hcheck(hipMemcpy(b->D.x + i, &x, sizeof(float4), hipMemcpyHostToDevice), "set root x");
        hcheck(hipMemcpy(b->D.v + i, &v, sizeof(float4), hipMemcpyHostToDevice), "set root v");
        if (q) {
            const q4 u = qunit_or_identity(q4{q[0], q[1], q[2], q[3]});
            const float4 qq = make_float4(u.x, u.y, u.z, u.w);
            hcheck(hipMemcpy(b->D.q + i, &qq, sizeof(float4), hipMemcpyHostToDevice), "set root q");
        }
        hcheck(launch_gather(b->params(), b->D, b->stream), "gather");
        hcheck(hipStreamSynchronize(b->stream), "sync");
    });
}

int rs_step_const(rs_batch* b, const float* pos, const float* vel, const float* q, float dt, int substeps) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (!pos || !vel) throw Invalid("root_pos and root_vel are required");
        Drive d{};
        d.mode = DRIVE_CONST;
        d.p0 = pos;
        d.v0 = vel;
        d.q0 = q;
        b->frame(d, dt, substeps);
    });
}

int rs_step_program(rs_batch* b, const float* p0, const float* v0, const float* p1, const float* q0, const float* q1,
                    float dt, int substeps) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (!p0 || !v0 || !p1) throw Invalid("root_pos, root_vel and root_pos_end are required");
        if (q1 && !q0) throw Invalid("root_q_end requires root_q");
        Drive d{};
        d.mode = DRIVE_PROGRAM;
        d.p0 = p0;
        d.v0 = v0;
        d.p1 = p1;
        d.q0 = q0;
        d.q1 = q1;
        b->frame(d, dt, substeps);
    });
}

int rs_step_free(rs_batch* b, float dt, int substeps) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        Drive d{};
        d.mode = DRIVE_FREE;
        b->frame(d, dt, substeps);
    });
}

int rs_reset_envs(rs_batch* b, const int* ids, int count) {
    return guarded([&] {
        if (count <= 0) return;
        std::lock_guard<std::mutex> g(b->mtx);
        if (count > b->N) throw Invalid("ids longer than num_envs");
        for (int i = 0; i < count; i++)
            if (ids[i] < 0 || ids[i] >= b->N) throw Invalid("env id out of range");
        hcheck(hipMemcpyAsync(b->d_ids, ids, sizeof(int) * count, hipMemcpyHostToDevice, b->stream), "upload ids");
        hcheck(launch_reset(b->params(), b->D, b->d_initX, b->d_initQ, b->d_ids, count, b->stream), "reset");
        b->gather_nolock();
        hcheck(hipStreamSynchronize(b->stream), "reset");
    });
}

int rs_gather(rs_batch* b) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->gather_nolock();
    });
}

int rs_views(rs_batch* b, rs_views_t* out) {
    return guarded([&] {
        out->positions = b->D.outPos;
        out->velocities = b->D.outVel;
        out->quaternions = b->D.outQuat;
        out->contact_forces = b->forcesOn ? b->D.force : nullptr;
        out->wrenches = b->forcesOn ? b->D.wrench : nullptr;
        out->num_vertices = b->nv;
        out->num_ropes = b->N;
    });
}

static void check_material(float radius, float mu_s, float mu_k, float v_slip, float kappa, const char* who) {
    if (!(radius > 0) || !(kappa > 0) || !(mu_s >= 0) || !(mu_k >= 0) || !(v_slip > 0))
        throw Invalid(std::string(who) + " needs radius > 0, kappa > 0, mu_s >= 0, mu_k >= 0, v_slip > 0");
}

int rs_set_planes(rs_batch* b, int n, const float* planes, float radius, float mu_s, float mu_k, float v_slip,
                  float kappa) {
    return guarded([&] {
        if (n < 0 || n > 8) throw Invalid("planes must be [P<=8, 4] rows (nx, ny, nz, w): half-spaces n.x >= w");
        if (n > 0) check_material(radius, mu_s, mu_k, v_slip, kappa, "plane contact");
        float rows[8][4];
        for (int i = 0; i < n; i++) {
            const f3 nrm = mk3(planes[4 * i], planes[4 * i + 1], planes[4 * i + 2]);
            const float l = len(nrm);
            if (!(l > 0) || !std::isfinite(l) || !std::isfinite(planes[4 * i + 3]))
                throw Invalid("plane rows need a finite, nonzero normal and a finite offset");
            rows[i][0] = nrm.x / l;
            rows[i][1] = nrm.y / l;
            rows[i][2] = nrm.z / l;
            rows[i][3] = planes[4 * i + 3] / l;
        }
        std::lock_guard<std::mutex> g(b->mtx);
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) b->planes[i][c] = rows[i][c];
        b->numPlanes = n;
        if (n > 0) {
            b->cRadius = radius;
            b->cMuS = mu_s;
            b->cMuK = mu_k;
            b->cVSlip = v_slip;
            b->cKappa = kappa;
        }
    });
}

int rs_set_contact(rs_batch* b, float radius, float mu_s, float mu_k, float v_slip, float kappa) {
    return guarded([&] {
        check_material(radius, mu_s, mu_k, v_slip, kappa, "surface contact");
        std::lock_guard<std::mutex> g(b->mtx);
        b->cRadius = radius;
        b->cMuS = mu_s;
        b->cMuK = mu_k;
        b->cVSlip = v_slip;
        b->cKappa = kappa;
    });
}

int rs_get_contact(const rs_batch* b, float* o) {
    o[0] = b->cRadius;
    o[1] = b->cMuS;
    o[2] = b->cMuK;
    o[3] = b->cVSlip;
    o[4] = b->cKappa;
    return RS_OK;
}

int rs_set_rope_contact(rs_batch* b, int enabled, float radius, float barrier, float k_collision, float mu,
                        float k_friction, float detection_scaler, int group_ropes) {
    return guarded([&] {
        if (enabled && (!(radius > 0) || !(barrier > 0) || !(k_collision > 0) || !(mu >= 0) || !(k_friction >= 0) ||
                        !(detection_scaler > 1) || group_ropes < 1))
            throw Invalid("rope contact needs radius, barrier, k_collision > 0, mu, k_friction >= 0, "
                          "detection_scaler > 1, group_ropes >= 1");
        std::lock_guard<std::mutex> g(b->mtx);
        if (enabled) {
            if (2.f * b->minSegLen < 2.f * radius + barrier)
                throw Invalid("rope contact needs 2 * radius + barrier <= 2 * shortest segment");
            if (!b->D.coll) b->D.coll = dalloc<int>((size_t)b->nv * MAX_COLL, "collision lists");
            if (!b->D.pairFlag) b->D.pairFlag = dalloc<int>(b->N, "pair flags");
            b->groupRopes = group_ropes;
            b->radius = radius;
            b->barrier = barrier;
            b->kCollision = k_collision;
            b->fricMu = mu;
            b->kFriction = k_friction;
            b->detScaler = detection_scaler;
        } else {
            hcheck(launch_fill_i32(b->D.ncol, 0, b->nv, b->stream), "clear collision counts");
        }
        b->ropeContact = enabled != 0;
        b->cellsFor = -2;
    });
}

int rs_add_mesh(rs_batch* b, const float* values, int nx, int ny, int nz, const float* origin3, float spacing,
                float mu_s, float mu_k, const float* pos, const float* quat, int rows, int* index_out) {
    return guarded([&] {
        if (nx < 2 || ny < 2 || nz < 2) throw Invalid("sdf values must be [nx, ny, nz] with every dim >= 2");
        if (!(spacing > 0) || !std::isfinite(spacing)) throw Invalid("sdf spacing must be > 0");
        const size_t n = (size_t)nx * ny * nz;
        for (size_t i = 0; i < n; i++)
            if (!std::isfinite(values[i])) throw Invalid("sdf values must be finite");
        if (rows != 1 && rows != b->N) throw Invalid("add_mesh: pose rows must be 1 or num_envs");
        std::lock_guard<std::mutex> g(b->mtx);
        if (!(b->cRadius > 0))
            throw Invalid("set the contact material first: set_contact(radius, ...) or set_planes(..., radius=)");
        if ((int)b->grids.size() >= MAX_SDF) throw std::invalid_argument("at most 64 meshes per sim");
        b->ensure_mesh_state();
        float* d = dalloc<float>(n, "sdf grid");
        const hipError_t e = hipMemcpy(d, values, n * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(d);
            hcheck(e, "upload sdf grid");
        }
        b->gridMem.push_back(d);
        Grid G{};
        G.v = d;
        G.nx = nx;
        G.ny = ny;
        G.nz = nz;
        G.ox = origin3[0];
        G.oy = origin3[1];
        G.oz = origin3[2];
        G.hx = origin3[0] + float(nx - 1) * spacing;
        G.hy = origin3[1] + float(ny - 1) * spacing;
        G.hz = origin3[2] + float(nz - 1) * spacing;
        G.h = spacing;
        G.invH = 1.f / spacing;
        G.muS = mu_s;
        G.muK = mu_k;
        b->grids.push_back(G);
        b->upload_grids();
        const int idx = (int)b->grids.size() - 1;
        const bool pending = b->pendingSet;  // other meshes' targets stay pending
        b->set_targets_host(idx, pos, quat, rows, nullptr, true);
        hcheck(launch_mesh_commit(b->slots(), b->D.poses, b->d_grids, (int)b->grids.size(), idx, b->N, b->stream),
               "commit mesh");
        hcheck(hipStreamSynchronize(b->stream), "add_mesh");
        b->pendingSet = pending;
        *index_out = idx;
    });
}

int rs_set_mesh_friction(rs_batch* b, int index, float mu_s, float mu_k) {
    return guarded([&] {
        if (!(mu_s >= 0) || !(mu_k >= 0)) throw Invalid("mu_s / mu_k must be >= 0");
        std::lock_guard<std::mutex> g(b->mtx);
        if (index < 0 || index >= (int)b->grids.size()) throw Invalid("mesh index out of range");
        b->grids[index].muS = mu_s;
        b->grids[index].muK = mu_k;
        b->upload_grids();
    });
}

int rs_set_mesh_targets(rs_batch* b, int index, const float* pos, const float* quat, int rows, const unsigned char* tele,
                        int tele_all, int on_device) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        const int M = (int)b->grids.size();
        if (index >= M || (index < 0 && M == 0)) throw Invalid("mesh index out of range");
        if (on_device)
            b->set_targets_dev(index, pos, quat, rows, tele_all ? nullptr : tele, tele_all != 0);
        else
            b->set_targets_host(index, pos, quat, rows, tele, tele_all != 0);
    });
}

int rs_clear_meshes(rs_batch* b) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        hcheck(hipStreamSynchronize(b->stream), "sync");
        for (float* p : b->gridMem) hipFree(p);
        b->gridMem.clear();
        b->grids.clear();
        b->pendingSet = false;
        b->moving = false;
        if (b->meshesTouched) b->upload_grids();
    });
}

int rs_num_meshes(const rs_batch* b) { return (int)b->grids.size(); }

int rs_enable_contact_forces(rs_batch* b, int on) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->ensure_mesh_state();
        b->forcesOn = on != 0;
    });
}

int rs_set_contact_force_mode(rs_batch* b, int fixed_point) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->ensure_mesh_state();
        if (fixed_point && !b->D.wrenchFx) {
            b->D.wrenchFx = dalloc<unsigned long long>((size_t)b->N * MAX_SDF * 6, "wrench accumulators");
            hcheck(hipMemset(b->D.wrenchFx, 0, (size_t)b->N * MAX_SDF * 6 * sizeof(unsigned long long)),
                   "memset wrench accumulators");
        }
        b->forcesFixed = fixed_point != 0;
    });
}

int rs_get_contact_force_mode(const rs_batch* b, int* fixed_point) {
    if (!b || !fixed_point) return RS_INVALID;
    *fixed_point = b->forcesFixed ? 1 : 0;
    return RS_OK;
}

int rs_contact_forces_host(rs_batch* b, float* out) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (!b->forcesOn) throw Invalid("enable_contact_forces(True) first");
        hcheck(hipStreamSynchronize(b->stream), "sync");
        hcheck(hipMemcpy(out, b->D.force, 3 * (size_t)b->nv * sizeof(float), hipMemcpyDeviceToHost), "download forces");
    });
}

int rs_mesh_wrenches_host(rs_batch* b, float* out) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (!b->forcesOn) throw Invalid("enable_contact_forces(True) first");
        const size_t total = (size_t)b->N * MAX_SDF * 6;
        std::vector<float> w(total);
        hcheck(hipStreamSynchronize(b->stream), "sync");
        hcheck(hipMemcpy(w.data(), b->D.wrench, total * sizeof(float), hipMemcpyDeviceToHost), "download wrenches");
        const int M = (int)b->grids.size();
        for (int e = 0; e < b->N; e++)
            for (int m = 0; m < M; m++)
                for (int c = 0; c < 6; c++) out[((size_t)e * M + m) * 6 + c] = w[((size_t)e * MAX_SDF + m) * 6 + c];
    });
}

int rs_reset_clock(rs_batch* b) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->time = 0.f;
        b->h = b->maxH;
        b->lastErrorCode = 0;
        b->lastWarningCode = 0;
    });
}

int rs_error_codes(const rs_batch* b, int* last_error, int* last_warning) {
    *last_error = b->lastErrorCode;
    *last_warning = b->lastWarningCode;
    return RS_OK;
}
int rs_clear_error_codes(rs_batch* b) {
    b->lastErrorCode = 0;
    b->lastWarningCode = 0;
    return RS_OK;
}

int rs_set_dispatch(rs_batch* b, int mode, int variant) {
    return guarded([&] {
        if (mode != 0 && mode != 1) throw Invalid("dispatch mode must be 0 or 1");
        if (variant >= kNumCellVariants || variant < -1) throw Invalid("unknown persistent-kernel variant");
        std::lock_guard<std::mutex> g(b->mtx);
        if (mode == 1 && variant >= 0 && !b->variant_supported(variant))
            throw Invalid("persistent-kernel variant too small for these ropes / groups or for the device's LDS");
        b->mode = mode;
        b->variant = variant;
    });
}
int rs_set_rope_broad_phase(rs_batch* b, int mode) {
    return guarded([&] {
        if (mode < 0 || mode > 2) throw Invalid("rope broad phase must be 0 (auto), 1 (scan) or 2 (bvh)");
        std::lock_guard<std::mutex> g(b->mtx);
        b->broadPhase = mode;
    });
}
int rs_get_rope_broad_phase(const rs_batch* b, int* mode, int* tree_active) {
    if (!b) return RS_INVALID;
    if (mode) *mode = b->broadPhase;
    if (tree_active) *tree_active = b->tree_detect() ? 1 : 0;
    return RS_OK;
}
int rs_get_dispatch(const rs_batch* b, int* mode, int* variant) {
    *mode = b->mode;
    *variant = b->variant >= 0 ? b->variant : (b->cellsFor >= 0 ? b->cellsFor : -1);
    return RS_OK;
}
int rs_num_cell_variants(void) { return kNumCellVariants; }
int rs_cell_variant_info(int variant, int* block, int* vpt, int* supported) {
    if (variant < 0 || variant >= kNumCellVariants) return RS_INVALID;
    *block = kCellVariants[variant].block;
    *vpt = kCellVariants[variant].vpt;
    if (supported) *supported = variant_lds_bytes(variant) <= device_max_lds() ? 1 : 0;
    return RS_OK;
}

int rs_autotune(rs_batch* b, float dt, int substeps, int program, float* best_ms) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        const size_t n = b->nv;
        // snapshot what a frame changes
        float4 *sx = dalloc<float4>(n, "tune x"), *sv = dalloc<float4>(n, "tune v"), *svl = dalloc<float4>(n, "tune vl"),
               *sq = dalloc<float4>(n, "tune q");
        hcheck(hipMemcpyAsync(sx, b->D.x, n * 16, hipMemcpyDeviceToDevice, b->stream), "snapshot");
        hcheck(hipMemcpyAsync(sv, b->D.v, n * 16, hipMemcpyDeviceToDevice, b->stream), "snapshot");
        hcheck(hipMemcpyAsync(svl, b->D.vlast, n * 16, hipMemcpyDeviceToDevice, b->stream), "snapshot");
        hcheck(hipMemcpyAsync(sq, b->D.q, n * 16, hipMemcpyDeviceToDevice, b->stream), "snapshot");
        const float sh = b->h, slh = b->lastH;
        // a representative drive: roots held where they are (program with p1 = p0, v0 = 0)
        std::vector<float> root(3 * b->N);
        {
            std::vector<float4> hx(n);
            hcheck(hipMemcpyAsync(hx.data(), b->D.x, n * 16, hipMemcpyDeviceToHost, b->stream), "roots");
            hcheck(hipStreamSynchronize(b->stream), "roots");
            for (int e = 0; e < b->N; e++) {
                root[3 * e] = hx[b->roff[e]].x;
                root[3 * e + 1] = hx[b->roff[e]].y;
                root[3 * e + 2] = hx[b->roff[e]].z;
            }
        }
        float* dRoot = dalloc<float>(3 * b->N, "tune roots");
        float* dZero = dalloc<float>(3 * b->N, "tune zero");
        hcheck(hipMemcpy(dRoot, root.data(), 12 * (size_t)b->N, hipMemcpyHostToDevice), "tune roots");
        hcheck(hipMemset(dZero, 0, 12 * (size_t)b->N), "tune zero");
        Drive in{};
        in.mode = program ? DRIVE_PROGRAM : DRIVE_FREE;
        in.p0 = dRoot;
        in.v0 = dZero;
        in.p1 = dRoot;
        Drive dr;
        b->stage_inputs(in, dr);
        dr.dt = dt;
        b->lastH = b->h;
        b->h = dt / float(substeps);
        b->ensure_frame_buffers(substeps);
        const int oldMode = b->mode, oldVar = b->variant;
        b->mode = 1;
        hipEvent_t e0, e1;
        hcheck(hipEventCreate(&e0), "event");
        hcheck(hipEventCreate(&e1), "event");
        int best = -1;
        float bestMs = 1e30f;
        std::string err;
        for (int v = 0; v < kNumCellVariants; v++) {
            if (!b->variant_supported(v)) continue;
            b->variant = v;
            b->build_cells(v);
            const Params P = b->params();
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {
                hcheck(hipEventRecord(e0, b->stream), "event");
                const hipError_t le = launch_cell(v, P, b->D, dr, b->d_cells, b->d_order, (int)b->cells.size(),
                                                  substeps, nullptr, b->stream);
                hcheck(hipEventRecord(e1, b->stream), "event");
                hcheck(hipEventSynchronize(e1), "event");
                if (le != hipSuccess) {
                    ms = 1e30f;
                    break;
                }
                hcheck(hipEventElapsedTime(&ms, e0, e1), "event");
                // restore the state for the next timing
                hcheck(hipMemcpyAsync(b->D.x, sx, n * 16, hipMemcpyDeviceToDevice, b->stream), "restore");
                hcheck(hipMemcpyAsync(b->D.v, sv, n * 16, hipMemcpyDeviceToDevice, b->stream), "restore");
                hcheck(hipMemcpyAsync(b->D.vlast, svl, n * 16, hipMemcpyDeviceToDevice, b->stream), "restore");
                hcheck(hipMemcpyAsync(b->D.q, sq, n * 16, hipMemcpyDeviceToDevice, b->stream), "restore");
            }
            if (ms < bestMs) {
                bestMs = ms;
                best = v;
            }
        }
        hcheck(hipMemsetAsync(b->D.err, 0, 2 * sizeof(int), b->stream), "reset error word");
        hcheck(hipStreamSynchronize(b->stream), "autotune");
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        for (void* p : {(void*)sx, (void*)sv, (void*)svl, (void*)sq, (void*)dRoot, (void*)dZero}) hipFree(p);
        b->h = sh;
        b->lastH = slh;
        b->gather_nolock();
        b->mode = oldMode == 0 ? 0 : 1;
        b->variant = best >= 0 ? best : oldVar;
        if (b->cellsFor != b->variant && b->variant >= 0) b->build_cells(b->variant);
        if (best_ms) *best_ms = bestMs;
    });
}

int rs_check_detect(rs_batch* b, const float* pos, const float* pads, int* counts, int* lists, float* max_step) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (!b->D.coll) throw Invalid("enable rope contact first");
        const size_t n = b->nv;
        std::vector<float4> x(n);
        for (size_t i = 0; i < n; i++) x[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
        Dev D = b->D;
        D.x = dalloc<float4>(n, "check x");
        hcheck(hipMemcpy(D.x, x.data(), n * 16, hipMemcpyHostToDevice), "check x");
        hcheck(hipMemcpy(D.pad, pads, n * 4, hipMemcpyHostToDevice), "check pads");
        hcheck(launch_fill_i32(D.coll, -1, (long)n * MAX_COLL, b->stream), "clear lists");
        hcheck(launch_check_detect(b->params(), D, b->detect_tree(), b->stream), "check detect");
        hcheck(hipStreamSynchronize(b->stream), "check detect");
        hcheck(hipMemcpy(counts, D.ncol, n * 4, hipMemcpyDeviceToHost), "counts");
        hcheck(hipMemcpy(lists, D.coll, n * MAX_COLL * 4, hipMemcpyDeviceToHost), "lists");
        hcheck(hipMemcpy(max_step, D.maxStep, n * 4, hipMemcpyDeviceToHost), "step limits");
        for (size_t i = 0; i < n; i++)
            for (int c = counts[i]; c < MAX_COLL; c++) lists[i * MAX_COLL + c] = -1;
        hipFree(D.x);
        hcheck(hipMemsetAsync(D.err, 0, 2 * sizeof(int), b->stream), "reset error word");
        hcheck(launch_fill_i32(D.ncol, 0, (long)n, b->stream), "clear counts");
        hcheck(hipStreamSynchronize(b->stream), "check detect");
    });
}

int rs_check_broadphase(rs_batch* b, const float* pos, const float* dx, int* counts, int* cand) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->ensure_mesh_state();
        const size_t n = b->nv;
        std::vector<float4> x(n), d(n);
        for (size_t i = 0; i < n; i++) {
            x[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.f);
            d[i] = make_float4(dx[3 * i], dx[3 * i + 1], dx[3 * i + 2], 0.f);
        }
        float4* dX = dalloc<float4>(n, "check x");
        float4* dD = dalloc<float4>(n, "check dx");
        int* dC = dalloc<int>(n, "check counts");
        int* dL = dalloc<int>(n * MAX_CAND, "check cand");
        hcheck(hipMemcpy(dX, x.data(), n * 16, hipMemcpyHostToDevice), "check x");
        hcheck(hipMemcpy(dD, d.data(), n * 16, hipMemcpyHostToDevice), "check dx");
        Params P = b->params();
        P.sdfMoving = 0;
        hcheck(launch_check_broadphase(P, b->D, dX, dD, dC, dL, b->stream), "check broadphase");
        hcheck(hipStreamSynchronize(b->stream), "check broadphase");
        hcheck(hipMemcpy(counts, dC, n * 4, hipMemcpyDeviceToHost), "counts");
        hcheck(hipMemcpy(cand, dL, n * MAX_CAND * 4, hipMemcpyDeviceToHost), "cand");
        for (void* p : {(void*)dX, (void*)dD, (void*)dC, (void*)dL}) hipFree(p);
    });
}

int rs_set_kernel_timing(rs_batch* b, int on) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->timing = on != 0;
    });
}

int rs_kernel_time(rs_batch* b, double* total_ms, int* launches) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        double t = 0.0;
        for (auto& e : b->events) {
            hcheck(hipEventSynchronize(e.second), "event");
            float ms = 0.f;
            hcheck(hipEventElapsedTime(&ms, e.first, e.second), "event");
            t += ms;
            hipEventDestroy(e.first);
            hipEventDestroy(e.second);
        }
        *launches = (int)b->events.size();
        *total_ms = t;
        b->events.clear();
    });
}

int rs_set_chunk_stats(rs_batch* b, int on, long long* out4) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        unsigned long long h[4] = {0, 0, 0, 0};
        hcheck(hipStreamSynchronize(b->stream), "sync");
        hcheck(hipMemcpy(h, b->d_hits, sizeof(h), hipMemcpyDeviceToHost), "stats");
        hcheck(hipMemset(b->d_hits, 0, sizeof(h)), "stats");
        if (out4)
            for (int i = 0; i < 4; i++) out4[i] = (long long)h[i];
        b->countHits = on != 0;
    });
}

int rs_set_graph(rs_batch* b, int on) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        hcheck(hipStreamSynchronize(b->stream), "sync");
        b->graphOn = on != 0;
        if (!b->graphOn) b->drop_graph();
    });
}

int rs_graph_captures(const rs_batch* b, long long* captures) {
    if (!b || !captures) return RS_INVALID;
    *captures = b->graphCaptures;
    return RS_OK;
}

int rs_set_fast_math(rs_batch* b, int on) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (b->tape.on) throw Invalid("the arithmetic mode cannot change while a rollout is being recorded");
        b->fastMath = on != 0;
    });
}
int rs_get_fast_math(const rs_batch* b, int* on) {
    if (!b || !on) return RS_INVALID;
    *on = b->fastMath ? 1 : 0;
    return RS_OK;
}

// ------------------------------------------------------------------ differentiation form (no reference counterpart)

int rs_set_material(rs_batch* b, const float* ea, const float* ei, const float* mass, int on_device) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        const int n = b->nv;
        if (b->tape.on) throw Invalid("the material cannot change while a rollout is being recorded");
        std::vector<float> hea, hei, hm;
        auto fetch = [&](const float* src, std::vector<float>& dst) -> const float* {
            if (!src) return nullptr;
            if (!on_device) return src;
            dst.resize(n);
            hcheck(hipStreamSynchronize(b->stream), "sync");
            hcheck(hipMemcpy(dst.data(), src, n * sizeof(float), hipMemcpyDeviceToHost), "read material");
            return dst.data();
        };
        const float* EA = fetch(ea, hea);
        const float* EI = fetch(ei, hei);
        const float* M = fetch(mass, hm);
        for (int r = 0; r < b->N; r++)
            for (int i = b->roff[r]; i < b->roff[r + 1]; i++) {
                const bool seg = i + 1 < b->roff[r + 1];
                if (M && b->hprm[i].x != 0.f && (!std::isfinite(M[i]) || !(M[i] > 0)))
                    throw Invalid("node masses must be finite and > 0 (pinned entries ignored)");
                if (seg && EA && (!std::isfinite(EA[i]) || !(EA[i] > 0))) throw Invalid("EA values must be finite and > 0");
                if (seg && EI && (!std::isfinite(EI[i]) || !(EI[i] > 0))) throw Invalid("EI values must be finite and > 0");
            }
        hcheck(hipStreamSynchronize(b->stream), "sync");
        for (int r = 0; r < b->N; r++)
            for (int i = b->roff[r]; i < b->roff[r + 1]; i++) {
                const bool seg = i + 1 < b->roff[r + 1];
                float4& p = b->hprm[i];
                if (M && p.x != 0.f) p.x = 1.f / M[i];
                if (seg && EA) p.z = EA[i] * p.y;
                if (seg && EI) {
                    const float k = 4.f * EI[i] / p.y;
                    const float4 u = b->hRhoUnit[i];
                    b->hqrest[i] = make_float4(k * u.x, k * u.y, k * u.z, k * u.w);
                }
            }
        hcheck(hipMemcpy(b->D.prm, b->hprm.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload prm");
        hcheck(hipMemcpy(b->D.qrest, b->hqrest.data(), n * sizeof(float4), hipMemcpyHostToDevice), "upload qrest");
    });
}

int rs_tape_begin(rs_batch* b, int max_frames) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        if (max_frames < 1) throw Invalid("max_frames must be >= 1");
        b->tape_free();
        b->tape.numSdf = (int)b->grids.size();
        b->tape_check();
        const size_t n = b->nv;
        b->tape.maxFrames = max_frames;
        b->tape.ck = dalloc<float4>((size_t)(max_frames + 1) * 4 * n, "tape checkpoints");
        b->tape.prog = dalloc<float>((size_t)max_frames * 17 * b->N, "tape programs");
        if (b->ropeContact) {
            b->tape.groups = (b->N + b->groupRopes - 1) / b->groupRopes;
            b->tape.pairFlags = dalloc<int>((size_t)max_frames * b->tape.groups, "tape pair flags");
            // a frame whose flags are not recorded counts every group as having pairs (the full adjoint)
            hcheck(launch_fill_i32(b->tape.pairFlags, 1, (long)max_frames * b->tape.groups, b->stream), "pair flags");
        }
        if (b->tape.numSdf > 0) {
            const size_t slots = (size_t)max_frames * b->N * b->tape.numSdf;
            b->tape.poses = dalloc<Pose>(slots, "tape poses");
            b->tape.motion = dalloc<Motion>(slots, "tape motion");
        }
        b->tape.on = true;
    });
}

int rs_tape_end(rs_batch* b) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->tape_close();
    });
}

int rs_tape_frames(const rs_batch* b) { return b ? b->tape.frames : 0; }

int rs_adjoint(rs_batch* b, const float* dl_dx, float* g_ea, float* g_ei, float* g_mass, double* g_drag, float* g_p0,
               float* g_v0, float* g_p1, int* mismatches) {
    return guarded([&] {
        std::lock_guard<std::mutex> g(b->mtx);
        b->tape_close();
        auto& T = b->tape;
        if (!T.closed || T.frames < 1) throw Invalid("no recorded rollout: call rs_tape_begin, then step");
        const int F = T.frames, N = b->N;
        const size_t n = b->nv;
        // A unit is what a workgroup never splits: one rope, or a whole contact group of U consecutive ropes when
        // the rollout has rope-rope contact (the group's lists and pair terms stay inside the workgroup).
        const int U = T.unitRopes;
        bool anyRC = false;
        for (const Params& P : T.P) anyRC = anyRC || P.ropeContact != 0;
        auto unit_end = [&](int r) { return std::min(r + U, N); };
        int maxV = 0;
        for (int r = 0; r < N; r = unit_end(r)) maxV = std::max(maxV, b->roff[unit_end(r)] - b->roff[r]);
        // Workgroups of each pass: consecutive units packed up to B vertices (and the LDS pose slots). The passes
        // share only the per-vertex history, so each takes its own packing.
        const int M = T.numSdf;
        const bool fastAdj = T.P.empty() ? b->fastMath : T.P[0].fastMath != 0;  // the recorded arithmetic
        auto pack = [&](int B, int ropesPerBlock, int r0, int r1, std::vector<int>* out) {
            long waves = 0;
            int r = r0;
            while (r < r1) {
                const int start = r;
                int nvb = 0;
                while (r < r1) {
                    const int e = std::min(unit_end(r), r1);
                    if (r > start && (e - start > ropesPerBlock || nvb + (b->roff[e] - b->roff[r]) > B)) break;
                    nvb += b->roff[e] - b->roff[r];
                    r = e;
                }
                if (out) out->push_back(start);
                waves += (nvb + 63) / 64;
            }
            if (out) out->push_back(r1);
            return waves;
        };
        // One rope per workgroup of the smallest size that holds the longest rope, unless packing consecutive ropes
        // costs less and still leaves every CU several workgroups (a small batch needs the parallelism more than the
        // lanes). Both passes are latency-bound: a workgroup takes about as long with a few busy lanes as with
        // all, so the cost is the number of rounds, workgroups over the workgroups a CU holds at the pass's
        // register-limited waves per SIMD.
        struct Packing {
            int blk = -1, perBlock = 1;
        };
        auto choose = [&](bool reverse) {
            Packing best;
            double bestCost = 0.0;
            for (int B : {64, 128, 192, 256, 384, 512}) {
                if (B < maxV || !adjoint_block_supported(B, reverse)) continue;
                // LDS: the pose slots and, with rope contact, the detection arrays
                if (M > 0 && (size_t)U * M > (size_t)adjoint_pose_slots(B)) continue;
                const int perCU = 4 * adjoint_waves_per_simd(B, reverse, fastAdj, anyRC) / (B / 64);
                if (perCU < 1) continue;
                if (best.blk < 0) {
                    std::vector<int> t;
                    pack(B, 1, 0, N, &t);
                    best.blk = B;
                    bestCost = (double)(t.size() - 1) / perCU;
                }
                const int maxRopes = M > 0 ? adjoint_pose_slots(B) / M : B;
                if (maxRopes < 2 * U) continue;
                std::vector<int> t;
                pack(B, maxRopes, 0, N, &t);
                const double c = (double)(t.size() - 1) / perCU;
                if (c < bestCost && (long)t.size() - 1 >= 8L * device_cus()) {
                    best.blk = B;
                    best.perBlock = maxRopes;
                    bestCost = c;
                }
            }
            if (best.blk < 0) throw Invalid("the adjoint supports ropes (rope contact groups) of at most 512 vertices");
            return best;
        };
        const Packing pk[2] = {choose(false), choose(true)};
        // the history of a frame is kept for a chunk of ropes at a time, within a fixed budget
        int maxI = 1;
        for (const Params& P : T.P) maxI = std::max(maxI, P.numItr);
        const size_t perV = adjoint_history_bytes(maxI) * (size_t)T.maxS;
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) freeB = 0;
        const size_t budget =
            std::max(std::max(kAdjointHistoryBudget, freeB / 5 * 2 + b->histBytes), perV * (size_t)maxV);
        struct Chunk {
            int rope0, nropes, v0, nv;
        };
        std::vector<Chunk> chunks;
        for (int r = 0; r < N;) {
            Chunk c{r, 0, b->roff[r], 0};
            while (r < N) {
                const int e = unit_end(r);
                if (c.nropes > 0 && (size_t)(c.nv + b->roff[e] - b->roff[r]) * perV > budget) break;
                c.nv += b->roff[e] - b->roff[r];
                c.nropes += e - r;
                r = e;
            }
            chunks.push_back(c);
        }
        // per pass and chunk: offset into the table and workgroups; per pass the most ropes any workgroup holds
        std::vector<int> blkTable, chunkBlk0[2], chunkBlks[2];
        int ropesPerBlock[2] = {1, 1};
        for (int pass = 0; pass < 2; pass++)
            for (const Chunk& c : chunks) {
                std::vector<int> t;
                pack(pk[pass].blk, pk[pass].perBlock, c.rope0, c.rope0 + c.nropes, &t);
                for (size_t k = 0; k + 1 < t.size(); k++)
                    ropesPerBlock[pass] = std::max(ropesPerBlock[pass], t[k + 1] - t[k]);
                chunkBlk0[pass].push_back((int)blkTable.size());
                chunkBlks[pass].push_back((int)t.size() - 1);
                blkTable.insert(blkTable.end(), t.begin(), t.end());
            }
        if (blkTable.size() > b->blkCap) {
            if (b->d_blk) hipFree(b->d_blk);
            b->d_blk = dalloc<int>(blkTable.size(), "adjoint workgroups");
            b->blkCap = blkTable.size();
        }
        hcheck(hipMemcpyAsync(b->d_blk, blkTable.data(), blkTable.size() * sizeof(int), hipMemcpyHostToDevice,
                              b->stream),
               "adjoint workgroups");
        size_t need = 0, maxChunkV = 0;
        for (const Chunk& c : chunks) {
            need = std::max(need, (size_t)c.nv * perV);
            maxChunkV = std::max(maxChunkV, (size_t)c.nv);
        }
        if (anyRC && !b->adjStream) {
            hcheck(hipStreamCreate(&b->adjStream), "adjoint stream");
            hcheck(hipEventCreateWithFlags(&b->adjFork, hipEventDisableTiming), "adjoint event");
            hcheck(hipEventCreateWithFlags(&b->adjJoin, hipEventDisableTiming), "adjoint event");
        }
        const size_t listEntries = maxChunkV * adjoint_list_entries_per_vertex();
        if (anyRC && listEntries > b->adjListsCap) {
            if (b->d_adjLists) hipFree(b->d_adjLists);
            b->d_adjLists = dalloc<uint16_t>(listEntries, "adjoint rope-rope lists");
            b->adjListsCap = listEntries;
        }
        if (need > b->histBytes) {
            if (b->d_hist) hipFree(b->d_hist);
            b->d_hist = dalloc<char>(need, "adjoint history");
            b->histBytes = need;
        }
        if (!b->d_adj) {
            b->d_adj = dalloc<float4>(4 * n, "adjoint state");
            b->d_gk = dalloc<double>(n, "dk");
            b->d_gm = dalloc<double>(n, "dm");
            b->d_grho = dalloc<double>(4 * n, "drho");
            b->d_gdrag = dalloc<double>(N, "ddrag");
            b->d_gp = dalloc<float>(9 * (size_t)N, "droot");
            b->d_mismatch = dalloc<int>(1, "mismatch");
        }
        hipStream_t st = b->stream;
        hcheck(hipMemsetAsync(b->d_adj, 0, 4 * n * 16, st), "zero adjoint");
        hcheck(hipMemsetAsync(b->d_gk, 0, n * 8, st), "zero");
        hcheck(hipMemsetAsync(b->d_gm, 0, n * 8, st), "zero");
        hcheck(hipMemsetAsync(b->d_grho, 0, n * 32, st), "zero");
        hcheck(hipMemsetAsync(b->d_gdrag, 0, N * 8, st), "zero");
        hcheck(hipMemsetAsync(b->d_mismatch, 0, 4, st), "zero");
        for (int f = F - 1; f >= 0; f--) {
            const float4* ck = T.ck + (size_t)f * 4 * n;
            const float4* nk = T.ck + (size_t)(f + 1) * 4 * n;
            const int S = T.S[f], I = T.P[f].numItr;
            Drive dr = T.dr[f];
            const float* prog = T.prog + (size_t)f * 17 * N;
            dr.p0 = prog;
            dr.v0 = prog + 3 * N;
            dr.p1 = prog + 6 * N;
            dr.q0 = prog + 9 * N;
            dr.q1 = prog + 13 * N;
            for (size_t ci = 0; ci < chunks.size(); ci++) {
                const Chunk& c = chunks[ci];
                AdjDev A{};
                A.cx = ck;
                A.cv = ck + n;
                A.cvl = ck + 2 * n;
                A.cq = ck + 3 * n;
                A.nx = nk;
                A.nv = nk + n;
                A.nvl = nk + 2 * n;
                A.nq = nk + 3 * n;
                A.rope0 = c.rope0;
                A.hv0 = c.v0;
                A.hn = c.nv;
                const size_t Sn = (size_t)S * c.nv;
                char* p = b->d_hist;
                A.hX = reinterpret_cast<float4*>(p);
                A.hV = A.hX + Sn;
                A.hVL = A.hV + Sn;
                A.hY = A.hVL + Sn;
                A.hD = A.hY + Sn;
                A.hQ = A.hD + Sn * (I + 1);
                A.hC = reinterpret_cast<uint2*>(A.hQ + Sn * (I + 1));
                A.hM = A.hC + Sn;
                A.hR = reinterpret_cast<uint4*>(A.hM + Sn);
                A.hH = reinterpret_cast<float*>(A.hR + Sn);
                A.ax = b->d_adj;
                A.av = A.ax + n;
                A.avl = A.av + n;
                A.aq = A.avl + n;
                A.dlx = dl_dx ? dl_dx + (size_t)f * n * 3 : nullptr;
                A.gk = b->d_gk;
                A.grho = b->d_grho;
                A.gm = b->d_gm;
                A.gdrag = b->d_gdrag;
                A.gp0 = g_p0 ? g_p0 + (size_t)f * 3 * N : b->d_gp;
                A.gv0 = g_v0 ? g_v0 + (size_t)f * 3 * N : b->d_gp + 3 * N;
                A.gp1 = g_p1 ? g_p1 + (size_t)f * 3 * N : b->d_gp + 6 * N;
                A.mismatch = b->d_mismatch;
                A.fposes = M > 0 ? T.poses + (size_t)f * N * M : nullptr;
                A.fmotion = (M > 0 && T.P[f].sdfMoving) ? T.motion + (size_t)f * N * M : nullptr;
                A.lists = anyRC ? b->d_adjLists : nullptr;
                A.pairFlag = T.pairFlags ? T.pairFlags + (size_t)f * T.groups : nullptr;
                // With rope contact the groups with rope-rope pairs run their two passes on a second stream while
                // the pair-free groups run theirs here (groups never interact): the pair groups' kernels hold few,
                // long workgroups, which would leave most compute units idle on their own.
                const bool split = T.P[f].ropeContact != 0;
                if (split) {
                    hcheck(hipEventRecord(b->adjFork, st), "adjoint fork");
                    hcheck(hipStreamWaitEvent(b->adjStream, b->adjFork, 0), "adjoint fork");
                }
                // the pair groups' (long) workgroups are issued first
                const int classes[2] = {split ? 2 : 0, 1};
                for (int ci2 = 0; ci2 < (split ? 2 : 1); ci2++) {
                    const int cls = classes[ci2];
                    hipStream_t sc = cls == 2 ? b->adjStream : st;
                    for (int pass = 0; pass < 2; pass++) {
                        A.blk = b->d_blk + chunkBlk0[pass][ci];
                        hcheck(launch_adjoint_pass(pass == 1, pk[pass].blk, T.P[f], b->D, A, dr, S, chunkBlks[pass][ci],
                                                   ropesPerBlock[pass], sc, cls),
                               pass ? "adjoint reverse pass" : "adjoint forward pass");
                    }
                }
                if (split) {
                    hcheck(hipEventRecord(b->adjJoin, b->adjStream), "adjoint join");
                    hcheck(hipStreamWaitEvent(st, b->adjJoin, 0), "adjoint join");
                }
            }
        }
        hcheck(launch_adjoint_finish(b->params(), b->D, b->d_rhoUnit, b->d_gk, b->d_grho, b->d_gm, g_ea, g_ei, g_mass,
                                     st),
               "adjoint finish");
        std::vector<double> gd(N);
        int mis = 0;
        hcheck(hipMemcpyAsync(gd.data(), b->d_gdrag, N * 8, hipMemcpyDeviceToHost, st), "read drag adjoint");
        hcheck(hipMemcpyAsync(&mis, b->d_mismatch, 4, hipMemcpyDeviceToHost, st), "read mismatch");
        int err[2] = {0, 0};
        if (anyRC) hcheck(hipMemcpyAsync(err, b->D.err, sizeof(err), hipMemcpyDeviceToHost, st), "read error word");
        hcheck(hipStreamSynchronize(st), "adjoint");
        if (err[0] != 0 || err[1] != 0) {
            hcheck(hipMemsetAsync(b->D.err, 0, sizeof(err), st), "reset error word");
            hcheck(hipStreamSynchronize(st), "reset error word");
            throw SolverErr("MAX_COLLISIONS_PER_SEGMENT exceeded in the adjoint's rope-rope lists");
        }
        double sum = 0.0;
        for (double v : gd) sum += v;
        if (g_drag) *g_drag = sum;
        if (mismatches) *mismatches = mis;
    });
}

}  // extern "C"

// ------------------------------------------------------------------ mesh distance grids

int rs_bake_mesh_sdf(const float* vertices, int num_vertices, const int* triangles, int num_triangles,
                     const float* origin3, float spacing, int nx, int ny, int nz, float thickness, float round,
                     int device, float* out) {
    return guarded([&] {
        if (!vertices || !triangles || !origin3 || !out) throw Invalid("bake_mesh: null pointer");
        if (num_vertices < 3 || num_triangles < 1) throw Invalid("bake_mesh needs a triangle mesh");
        if (!(spacing > 0.f) || nx < 2 || ny < 2 || nz < 2) throw Invalid("bake_mesh: bad lattice");
        if (!(thickness >= 0.f) || !(round >= 0.f)) throw Invalid("bake_mesh: thickness and round must be >= 0");
        std::vector<float4> tri(3 * (size_t)num_triangles);
        for (int t = 0; t < num_triangles; t++)
            for (int c = 0; c < 3; c++) {
                const int v = triangles[3 * t + c];
                if (v < 0 || v >= num_vertices) throw Invalid("triangle indices out of range");
                tri[3 * (size_t)t + c] = make_float4(vertices[3 * v], vertices[3 * v + 1], vertices[3 * v + 2], 0.f);
            }
        int prev = -1;
        if (device >= 0) {
            hcheck(hipGetDevice(&prev), "get device");
            hcheck(hipSetDevice(device), "set device");
        }
        const size_t nodes = (size_t)nx * ny * nz;
        float4* dTri = nullptr;
        float* dOut = nullptr;
        hipStream_t st = nullptr;
        try {
            hcheck(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
            dTri = dalloc<float4>(tri.size(), "bake triangles");
            dOut = dalloc<float>(nodes, "bake grid");
            hcheck(hipMemcpyAsync(dTri, tri.data(), tri.size() * sizeof(float4), hipMemcpyHostToDevice, st),
                   "upload triangles");
            const double o[3] = {origin3[0], origin3[1], origin3[2]};
            hcheck(launch_bake_mesh(dTri, num_triangles, o, (double)spacing, nx, ny, nz, thickness > 0.f ? 1 : 0,
                                    (float)(0.5 * (double)thickness), round, dOut, st),
                   "bake kernel");
            hcheck(hipMemcpyAsync(out, dOut, nodes * sizeof(float), hipMemcpyDeviceToHost, st), "download grid");
            hcheck(hipStreamSynchronize(st), "bake");
        } catch (...) {
            if (st) hipStreamSynchronize(st);
            if (dTri) hipFree(dTri);
            if (dOut) hipFree(dOut);
            if (st) hipStreamDestroy(st);
            if (prev >= 0) hipSetDevice(prev);
            throw;
        }
        hipFree(dTri);
        hipFree(dOut);
        hipStreamDestroy(st);
        if (prev >= 0) hipSetDevice(prev);
    });
}
