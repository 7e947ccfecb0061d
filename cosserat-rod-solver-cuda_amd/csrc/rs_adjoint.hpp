// Reverse-mode (adjoint) pass of the substep: device buffers and host launchers (rs_adjoint.hip).
//
// The reference has no adjoint. The backward pass differentiates exactly the arithmetic of the
// forward substep (rs_device.hpp), with the forward active set frozen:
// - contact on / off
// - friction stick / slide
// - the gravity acceleration clamp
// - the orientation sign tests
// - the broad-phase candidate lists
#pragma once

#include <hip/hip_runtime.h>

#include "rs_types.hpp"

namespace rs {

struct AdjDev {
    // frame checkpoint: state at the start of the frame [nv], and at the start of the next one (recompute check)
    const float4 *cx, *cv, *cvl, *cq;
    const float4 *nx, *nv, *nvl, *nq;
    // history of the frame for the vertices [hv0, hv0 + hn) of a launch, local vertex i - hv0
    int rope0, hv0, hn;
    // workgroup b of a launch advances the consecutive ropes [blk[b], blk[b + 1]) (at most the launch's block size
    // in vertices and adjoint_pose_slots(block) / numSdf ropes in all)
    const int* blk;
    float4 *hX, *hV, *hVL, *hY;  // [S][hn] substep inputs: x after the root program (w: broad-phase count bits),
                                 // v after the root program, last velocity before init, inertial target y
    uint2* hC;                   // [S][hn] packed broad-phase candidate ids
    uint2* hM;                   // [S][hn] mesh candidates in contact: bit 8k + c for iteration k < 8 and
                                 // candidate c (all bits set on scan-all substeps and past 8 iterations)
    uint4* hR;                   // [S][hn] (rope-rope pairs) the substep's step limit, limit partner and list of
                                 // the vertex's segment: rc_tape_word
    float4 *hD, *hQ;             // [S][I + 1][hn] iterates d^k and frames q^k
    float* hH;                   // [S][I][9][hn] total Hessian (packed 6) and Newton direction H^-1 F (3)
    // adjoint of the state at the end of the frame on entry, at its start on exit [nv]
    float4 *ax, *av, *avl, *aq;
    const float* dlx;  // dL/dx of this frame's output positions [nv * 3], may be null
    // accumulated parameter adjoints [nv]: k = EA l, rest Darboux vector [nv][4], mass; drag per rope [N]
    double* gk;
    double* grho;
    double* gm;
    double* gdrag;
    // this frame's root-program adjoints [N * 3] (p0, v0, p1)
    float *gp0, *gv0, *gp1;
    int* mismatch;  // count of vertices whose recomputed end state differs from the next checkpoint
    const Pose* fposes;     // frame-start poses [N][numSdf]
    const Motion* fmotion;  // frame motion [N][numSdf] while meshes move, else null
    // with rope-rope contact, per local vertex (adjoint_list_entries_per_vertex entries): the substep's partner
    // list (workgroup-local slots, ascending), the list of the segments whose lists hold it (MAX_COLL each), the
    // skin list its candidates come from and the pair index of each reverse-list entry; each pass rebuilds them
    uint16_t* lists;
    // with rope-rope contact, per contact group of this frame: nonzero when some segment of the group had a
    // rope-rope pair at some substep of the recorded forward (null: treat every group so)
    const int* pairFlag;
    int mbSlots;  // (set by the launcher) pair-mailbox slots of a workgroup with rope-rope pairs
    int tapeMax;  // (set by the launcher) longest list pass 2 takes from pass 1's record (at most 7)
};

// Longest rope (or rope contact group) the backward kernels support; pose slots a workgroup keeps in LDS (its
// ropes times the meshes); history bytes per vertex-substep.
constexpr int ADJ_MAX_BLOCK = 512;
constexpr int adjoint_pose_slots(int block) { return block > MAX_SDF ? block : MAX_SDF; }
size_t adjoint_history_bytes(int numItr);
bool adjoint_block_supported(int block, bool reverse);  // 64, 128, 192, 256, 384, 512
// Waves per SIMD the pass's register allocation allows at this block size (the device's occupancy limit).
int adjoint_waves_per_simd(int block, bool reverse, bool fast, bool ropeContact);
// Dynamic LDS of the rope-rope detection arrays of one workgroup (beside its pose slots).
size_t adjoint_rc_lds_bytes(int block);
// Entries (uint16) of AdjDev::lists per local vertex: partner list, reverse list, skin list, pair indices.
size_t adjoint_list_entries_per_vertex();
// One pass (forward recompute, or reverse) over `nblocks` workgroups of `block` threads, workgroup b on the
// ropes [A.blk[b], A.blk[b + 1]), at most ropesPerBlock of them (their mesh poses take ropesPerBlock * numSdf
// LDS slots, at most adjoint_pose_slots(block)). The two passes of a frame may use different packings: they
// share only the per-vertex history.
// With rope contact, rcClass picks the workgroups' class: 1 pair-free groups (the clamp-only kernel), 2 groups
// with rope-rope pairs (the full kernel), 0 both one after the other. Without rope contact it is ignored.
hipError_t launch_adjoint_pass(bool reverse, int block, const Params& P, const Dev& D, const AdjDev& A,
                               const Drive& Dr, int S, int nblocks, int ropesPerBlock, hipStream_t s, int rcClass = 0);
hipError_t launch_adjoint_finish(const Params& P, const Dev& D, const float4* rhoUnit, const double* gk,
                                 const double* grho, const double* gm, float* gea, float* gei, float* gmass,
                                 hipStream_t s);

}  // namespace rs
