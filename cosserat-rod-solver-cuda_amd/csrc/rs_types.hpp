// Device data layout of the rope batch: flat structure-of-arrays over all vertices, ropes as a
// CSR over them (uneven lengths are native), kernel parameters passed by value.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rs {

// Vertex flags, same meaning as the reference's VertexFlags
// (third_party/yarnball/KittenEngine/YarnBall/YarnBall.h:52-58).
enum : uint32_t { HAS_PREV = 1u, HAS_NEXT = 2u, HAS_NEXT_ORIENT = 4u, FIX_ORIENT = 8u };
constexpr int MAX_SDF = 64;    // meshes per batch (YarnBall.h:153 kMaxSdf)
constexpr int MAX_CAND = 8;    // broad-phase candidates per vertex (YarnBall.h:153 kMaxSdfCand)
constexpr int MAX_COLL = 128;  // rope-rope candidates per segment (YarnBall.h:38)
constexpr int VERLET_CAP = 64; // skin-list entries per segment in the persistent kernel
enum { ERR_NONE = 0, ERR_MAX_COLLISIONS = 1, ERR_BVH_DEPTH = 2 };

// A signed-distance grid: phi on an (nx, ny, nz) node lattice, C order (i slowest), at
// origin + (i, j, k) * h in the object's frame (YarnBall.h:141). Four 16-byte rows; a sample reads the first
// three (load_grid), and each of them holds a field its bounds test needs, so all three loads are issued at
// once ahead of that test.
struct alignas(16) Grid {
    const float* v;    // row 0: the node values and the first two sizes
    int nx, ny;
    int nz;            // row 1
    float ox, oy, oz;  // origin
    float invH, h;     // row 2: the spacing and the friction coefficients
    float muS, muK;
    float hx, hy, hz;  // row 3: far lattice corner origin + (n - 1) * h (world boxes only)
    float pad_;
};
static_assert(sizeof(Grid) == 64, "Grid is four 16-byte rows");

// Pose of (rope, mesh) at the end of the current substep plus its velocity over it and the world
// AABB of its lattice (YarnBall.h:147). 20 floats so a slot is five float4 loads.
struct Pose {
    float px, py, pz, qx, qy, qz, qw, vx, vy, vz, wx, wy, wz, lox, loy, loz, hix, hiy, hiz;
    float bpad;  // the broad phase's moving-mesh padding h (|v| + |w| |hi - lo| / 2) of this pose
};
// One frame of kinematic motion (YarnBall.h:152).
struct Motion {
    float p0x, p0y, p0z, q0x, q0y, q0z, q0w, dpx, dpy, dpz, rx, ry, rz, pad_;
};
// Target / current pose (csrc/sdf.h:32).
struct Pose7 {
    float px, py, pz, qx, qy, qz, qw, pad_;
};

// Solver knobs and contact material for one launch (what the reference keeps in MetaData,
// YarnBall.h:103-192). Passed by value.
struct Params {
    int nv, nropes;
    int numItr;
    float h, lastH;
    float gx, gy, gz;
    float damping, drag, dragQuad, dragQuadT, accel;
    int numPlanes;
    float planes[8][4];
    float cRadius, cKappa, cMuS, cMuK, cVSlip;
    int numSdf, sdfMoving, forcesOn;
    int forcesFixed;  // readout wrenches accumulated as 64-bit fixed point (order-independent sums)
    int ropeContact, groupRopes;
    float radius, barrier, kCollision, fricMu, kFriction, detR, scaledDetR;
    int fastMath;  // arithmetic of the iteration routines: 0 IEEE, 1 fused multiply-add + hardware rcp/sqrt/rsq/log
};

// Device buffers.
struct Dev {
    // state
    float4* x;      // positions (w unused)
    float4* v;      // velocities; the inertial target y during a substep (plain kernels)
    float4* vlast;  // velocity at the previous substep start
    float4* q;      // segment frames (current)
    float4* q2;     // frame ping-pong buffer (plain kernels)
    float4* dx;     // displacement iterate
    float4* dx2;    // displacement ping-pong buffer
    // per-vertex constants
    float4* prm;    // (invMass, lRest, kStretch, flags as bits)
    float4* qrest;  // rest Darboux quaternion scaled by 4 EI / l
    int* ropeOf;    // rope of each vertex
    const int* roff;  // [nropes + 1]
    // rope-rope contact
    float* pad;       // detection padding of each vertex this substep
    float* maxStep;   // step limit of each segment this substep
    int* ncol;        // [nv]
    int* coll;        // [nv * MAX_COLL] row-major, ascending partner index
    // Lists of the persistent kernel, per cell of `cap` slots, entry-major ([entry][slot], so the lanes of a
    // wave reading entry e of their own lists touch consecutive addresses), cell-local slots as 16 bits:
    uint16_t* vlist;  // skin lists, [cells][VERLET_CAP][cap]
    uint16_t* clist;  // rope-rope candidate lists of the substep, [cells][MAX_COLL][cap]
    // meshes
    const Grid* grids;     // [MAX_SDF]
    Pose* poses;           // [nropes * MAX_SDF]
    const Motion* motion;  // [nropes * MAX_SDF]
    int* ncand;            // [nv] (-1 = visit every mesh)
    int* cand;             // [nv * MAX_CAND]
    float* force;          // [nv * 3] contact-force accumulator
    float* wrench;         // [nropes * MAX_SDF * 6]
    unsigned long long* wrenchFx;  // [nropes * MAX_SDF * 6] fixed-point accumulator (forcesFixed)
    // outputs (gathered views) and the error word
    float* outPos;   // [nv * 3]
    float* outVel;   // [nv * 3]
    float* outQuat;  // [nv * 4]
    int* err;        // [2]
    // per-frame inputs of the persistent kernel, precomputed per substep
    const float4* rootTraj;  // [S][nropes][3]: root program x (w = 1 when it sets the frame), v, q
    const Pose* posesS;      // [S][nropes][numSdf] mesh poses while meshes move
    const int* cellMap;      // [cells * cell capacity] vertex of each cell slot (-1 past the cell's end)
    unsigned* cellTime;      // [cells] each cell's duration in its last frame (100 MHz ticks), or null
    int* cellQueue;          // [1] next unclaimed position of the cell launch order (cleared before each frame)
    int* pairFlag;           // [contact groups] set when a segment of the group has a rope-rope pair at some substep
                             // of the frame (cleared before each frame while rope contact is on); the adjoint reads
                             // it to run pair-free groups without the pair machinery
};

// What drives the roots during one frame.
enum { DRIVE_FREE = 0, DRIVE_CONST = 1, DRIVE_PROGRAM = 2 };
struct Drive {
    int mode;
    int hasQ0, hasQ1;
    float dt;
    const float* p0;  // [nropes * 3] (CONST: position; PROGRAM: start position)
    const float* v0;  // [nropes * 3]
    const float* p1;  // [nropes * 3] PROGRAM end position
    const float* q0;  // [nropes * 4]
    const float* q1;  // [nropes * 4]
};

// A cell: whole ropes (whole contact groups) that one workgroup advances through every substep of a frame.
// Its slots [0, nv) map to vertices through Dev::cellMap[v0 + slot]; each rope and each contact group
// occupies consecutive slots.
struct Cell {
    int v0, nv, units, pad_;
};

}  // namespace rs
