/* C ABI of the MI355X rope solver (librope_hip.so).
 *
 * The drop-in boundary for the reference's RopeBatch / RopeSim hot path. Plain pointers and sizes
 * only; "device" pointers are HIP device addresses (for example a torch tensor's data_ptr()),
 * "host" pointers are ordinary host memory. Every entry point returns an rs_status; the message of
 * the last failure on the calling thread is rs_last_error_message().
 *
 * Reference interfaces each entry point replaces are cited as path:line in the reference repository
 * (csrc/ = the ropesim extension, YB = third_party/yarnball/KittenEngine/YarnBall).
 */
#ifndef ROPESIM_HIP_H
#define ROPESIM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RS_OK = 0,
    RS_INVALID = 1,  /* bad argument or call order -> ValueError */
    RS_DEVICE = 2,   /* HIP runtime failure -> RuntimeError */
    RS_SOLVER = 3    /* solver error (rope-rope candidate overflow) -> RuntimeError */
} rs_status;

typedef struct rs_batch rs_batch;

/* Construction input: ropes as a CSR over flat vertices (uneven lengths allowed). */
typedef struct {
    int num_ropes;
    const int* offsets;          /* host [num_ropes + 1], offsets[0] == 0, every rope >= 3 vertices */
    const float* positions;      /* host [offsets[num_ropes] * 3] initial (and rest) shape */
    const unsigned char* pinned; /* host [num_vertices] kinematic vertices, or NULL = each rope's vertex 0 */
    float density;               /* rho [kg/m] */
    float k_stretch;             /* EA [N] */
    float k_bend;                /* EI [N m^2], > 0 */
    float max_h;                 /* seeds the previous substep length of the first frame */
    int num_itr;
    int fix_root_orientation;    /* 1: every rope's vertex-0 frame is kinematic */
    int straight_rest;           /* 1: straight untwisted rest shape */
    float min_seg;               /* > 0: length that scales the inactive rope-rope radii (uniform
                                    constructor); <= 0: the shortest given segment */
} rs_desc;

/* Solver knobs (rs_get_param / rs_set_param). */
typedef enum {
    RS_P_DRAG = 0,        /* linear drag [1/s] */
    RS_P_DAMPING = 1,
    RS_P_DRAG_QUAD = 2,   /* quadratic normal drag [1/m], >= 0 */
    RS_P_DRAG_QUAD_T = 3, /* quadratic tangential drag [1/m], >= 0 */
    RS_P_NUM_ITR = 4,
    RS_P_ACCEL = 5,       /* acceleration (over-relaxation) ratio */
    RS_P_MAX_H = 6,
    RS_P_TIME = 7,        /* read only */
    RS_P_STEPPED = 8      /* read only: 1 after the first step */
} rs_param;

/* Host-side per-vertex arrays (rs_get_vertex / rs_set_vertex), all [num_vertices * width]. */
typedef enum {
    RS_V_POSITIONS = 0,    /* width 3 */
    RS_V_VELOCITIES = 1,   /* width 3 */
    RS_V_QUATERNIONS = 2,  /* width 4, (x, y, z, w) */
    RS_V_INV_MASSES = 3,   /* width 1 (set: raw inverse masses, 0 = kinematic) */
    RS_V_REST_LENGTHS = 4, /* width 1 (entry at a rope's last vertex is padding) */
    RS_V_MASSES = 5,       /* width 1, set only (pinned vertices ignored) */
    RS_V_EA = 6,           /* width 1 per segment at its first vertex, set only */
    RS_V_EI = 7,           /* width 1 per segment at its first vertex, set only */
    RS_V_LAST_VELOCITIES = 8 /* width 3, read only */
} rs_vertex_field;

typedef struct {
    float* positions;      /* device [num_vertices * 3] */
    float* velocities;     /* device [num_vertices * 3] */
    float* quaternions;    /* device [num_vertices * 4] */
    float* contact_forces; /* device [num_vertices * 3] or NULL before rs_enable_contact_forces */
    float* wrenches;       /* device [num_ropes * 64 * 6] or NULL */
    int num_vertices, num_ropes;
} rs_views_t;

const char* rs_last_error_message(void);
int rs_abi_version(void);

/* RopeBatch ctor / from_positions (csrc/batch.cu:193-276; csrc/bindings.cpp:735-767, 889-934). */
int rs_create(const rs_desc* desc, rs_batch** out);
void rs_destroy(rs_batch* b);
int rs_num_vertices(const rs_batch* b);
int rs_num_ropes(const rs_batch* b);
int rs_offsets(const rs_batch* b, int* out); /* host [num_ropes + 1] */

/* Knobs (csrc/bindings.cpp:1161-1191 properties; YB/YarnBall.h:103-192). */
int rs_get_param(const rs_batch* b, int which, double* out);
int rs_set_param(rs_batch* b, int which, double value);
int rs_set_gravity(rs_batch* b, const float* g3);
int rs_get_gravity(const rs_batch* b, float* g3);

/* Per-vertex host access (csrc/bindings.cpp:1121-1201 setters; RopeSim accessors :354-400). Setters of
 * masses / EA / EI are legal at any time here; RopeBatch's Python layer refuses them after the first step as the
 * reference does. */
int rs_get_vertex(rs_batch* b, int which, float* host_out);
int rs_set_vertex(rs_batch* b, int which, const float* host_in);
/* Overwrite one rope's root position / velocity / frame from the host (RopeSim.set_root,
 * csrc/bindings.cpp:318-350). q may be NULL. */
int rs_set_root_host(rs_batch* b, int rope, const float* pos3, const float* vel3, const float* q4);

/* One control frame (csrc/batch.cu:288-398). Device arrays [num_ropes * 3] / [num_ropes * 4]; q pointers
 * may be NULL. step_const: constant-velocity sweep; step_program: quadratic position program to p1 with
 * SLERP q0 -> q1; step_free: no root drive. Inputs are consumed before the call returns (staged). */
int rs_step_const(rs_batch* b, const float* pos, const float* vel, const float* q, float dt, int substeps);
int rs_step_program(rs_batch* b, const float* p0, const float* v0, const float* p1, const float* q0, const float* q1,
                    float dt, int substeps);
int rs_step_free(rs_batch* b, float dt, int substeps);

/* RopeBatch.reset_envs / gather (csrc/batch.cu:400-425). ids are host ints. */
int rs_reset_envs(rs_batch* b, const int* ids, int count);
int rs_gather(rs_batch* b);
int rs_views(rs_batch* b, rs_views_t* out);

/* Surface contact (csrc/bindings.cpp:51-80 write_contact / write_planes). planes host [n * 4]. */
int rs_set_planes(rs_batch* b, int n, const float* planes, float radius, float mu_s, float mu_k, float v_slip,
                  float kappa);
int rs_set_contact(rs_batch* b, float radius, float mu_s, float mu_k, float v_slip, float kappa);
int rs_get_contact(const rs_batch* b, float* radius_mus_muk_vslip_kappa5);

/* Rope-rope contact (csrc/batch.cu:367-385). */
int rs_set_rope_contact(rs_batch* b, int enabled, float radius, float barrier, float k_collision, float mu,
                        float k_friction, float detection_scaler, int group_ropes);

/* Meshes (csrc/sdf.cu:137-296). values host [nx*ny*nz] C order; pos/quat host rows (1 or num_ropes). */
int rs_add_mesh(rs_batch* b, const float* values, int nx, int ny, int nz, const float* origin3, float spacing,
                float mu_s, float mu_k, const float* pos, const float* quat, int rows, int* index_out);
int rs_set_mesh_friction(rs_batch* b, int index, float mu_s, float mu_k);
/* Targets reached at the END of the next step. index < 0: every mesh from dense [num_ropes, M] rows.
 * on_device: 1 when pos/quat/tele are device pointers. tele may be NULL; tele_all replaces every pose. */
int rs_set_mesh_targets(rs_batch* b, int index, const float* pos, const float* quat, int rows,
                        const unsigned char* tele, int tele_all, int on_device);
int rs_clear_meshes(rs_batch* b);
int rs_num_meshes(const rs_batch* b);
int rs_enable_contact_forces(rs_batch* b, int on);
int rs_contact_forces_host(rs_batch* b, float* out); /* [num_vertices * 3] averaged over the last step */
int rs_mesh_wrenches_host(rs_batch* b, float* out);  /* [num_ropes * num_meshes * 6] */
/* Wrench summation of the readout. The reference sums with float atomics (contact.cuh:188-203, sdf.cu:89-92),
 * whose order, and so whose last bits, change from run to run. fixed_point = 1 rounds every contribution to
 * a 64-bit integer in units of 2^-32 N (N m) and sums exactly: the wrenches are then the same bit for bit in
 * every run, dispatch form and cell variant. Contributions are clamped to |1e9|, and a frame's sum per
 * component must stay below 2^31 N (N m). 0 (default) keeps float sums. No reference counterpart. */
int rs_set_contact_force_mode(rs_batch* b, int fixed_point);
int rs_get_contact_force_mode(const rs_batch* b, int* fixed_point);

/* Distance grid of a triangle mesh on a node lattice, computed on the GPU (reference ropesim/sdf.py:96-157,
 * which calls Warp's mesh_query_point_sign_winding_number per node). Host arrays: vertices [num_vertices*3],
 * triangles [num_triangles*3]; node (i, j, k) at origin + (i, j, k) * spacing (in double, then float); out
 * [nx*ny*nz] C order. thickness > 0: |distance| - thickness / 2 (two-sided shell); else the distance signed
 * negative where the generalized winding number exceeds 1/2 in magnitude. Then minus round. device < 0: the
 * current device. Blocks until the grid is on the host. */
int rs_bake_mesh_sdf(const float* vertices, int num_vertices, const int* triangles, int num_triangles,
                     const float* origin3, float spacing, int nx, int ny, int nz, float thickness, float round,
                     int device, float* out);

/* No counterpart in the reference: rs_bake_mesh_sdf for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc; it synchronises its stream, so not code under graph capture). vertices [num_vertices*3],
 * triangles [num_triangles*3] and out [nx*ny*nz] are DEVICE pointers; origin3 is host. The triangle corners are
 * staged in a fixed device array of the library (12 MiB per GPU, held until the process ends), so at most
 * RS_BAKE_NOALLOC_MAX_TRIANGLES triangles are accepted (RS_INVALID beyond). Same kernel and node arithmetic as
 * rs_bake_mesh_sdf, so the grids are identical bit for bit. stream: a hipStream_t, or NULL for the null stream;
 * the call blocks until the grid is written (it reads back the index check: RS_INVALID if an index is outside
 * [0, num_vertices), and out is then unspecified). Calls that overlap on one GPU share the staging array and
 * race; calls on different GPUs do not. */
#define RS_BAKE_NOALLOC_MAX_TRIANGLES 262144
int rs_bake_mesh_sdf_noalloc(const float* vertices, int num_vertices, const int* triangles, int num_triangles,
                             const float* origin3, float spacing, int nx, int ny, int nz, float thickness, float round,
                             void* stream, float* out);

/* RopeSim.reset bookkeeping (csrc/bindings.cpp:303-316): time 0, previous substep = max_h, codes cleared. */
int rs_reset_clock(rs_batch* b);

/* last_error / last_warning codes (YB/YarnBall.h:204-208). */
int rs_error_codes(const rs_batch* b, int* last_error, int* last_warning);
int rs_clear_error_codes(rs_batch* b);

/* Dispatch: mode 0 = one launch per kernel per substep, mode 1 = one persistent kernel per frame
 * (variant -1: autotuned on the first frame; when no cell holds the largest rope or contact group, mode 1
 * runs the per-substep kernels). No counterpart in the reference. */
int rs_set_dispatch(rs_batch* b, int mode, int variant);
int rs_get_dispatch(const rs_batch* b, int* mode, int* variant);
/* Rope-rope detection of the per-substep kernels (replaces YarnBall's LBVH path,
 * third_party/yarnball/KittenEngine/YarnBall/sim/collision.cu:10-101 and KittenGpuLBVH/lbvh.cu, reached
 * through RopeSim(collisions=True), csrc/bindings.cpp:225-235): 0 = auto (a per-group bounding-volume
 * hierarchy for groups above 128 vertices, the full scan below), 1 = always the full scan, 2 = always the
 * hierarchy. Both give the same candidate lists (ascending, lowest 128) and step limits. tree_active reports
 * whether the hierarchy is used for the batch as configured. */
int rs_set_rope_broad_phase(rs_batch* b, int mode);
int rs_get_rope_broad_phase(const rs_batch* b, int* mode, int* tree_active);
int rs_num_cell_variants(void);
int rs_cell_variant_info(int variant, int* block, int* vpt, int* supported);
/* Time one frame per supported variant on the current state (restored afterwards) and keep the fastest. */
int rs_autotune(rs_batch* b, float dt, int substeps, int program, float* best_ms);
/* Graph form of a frame (the role of the reference's rebuildCUDAGraph, YB/sim/step.cpp:5-57). When on,
 * each frame's kernel sequence (the per-substep kernels of mode 0, the persistent kernel of mode 1) is
 * captured once into a hipGraph. The graph is replayed while the frame's parameters, buffers, staged drive
 * and substep count are unchanged, and captured again when they change. rs_graph_captures counts the
 * captures. */
int rs_set_graph(rs_batch* b, int on);
int rs_graph_captures(const rs_batch* b, long long* captures);
/* Arithmetic of the per-iteration routines, no counterpart in the reference (which compiles its kernels
 * with nvcc's defaults: fused multiply-add, correctly rounded division and square root).
 *   on = 1 (default): fused multiply-add and the hardware reciprocal, square root, reciprocal square root
 *          and log (about 1 ulp each);
 *   on = 0: every operation correctly rounded and unfused, the operations of the float32 oracle.
 * Detection, broad phase, the root program and the mesh poses are IEEE in both modes. Both dispatch forms
 * and the adjoint produce bit-identical results within a mode. Refused while a rollout is recorded. */
int rs_set_fast_math(rs_batch* b, int on);
int rs_get_fast_math(const rs_batch* b, int* on);

/* Function-level checks, no counterpart in the reference: run the device rope-rope detection on given
 * substep-start positions and paddings (host [nv*3], [nv]) -> counts [nv], lists [nv*128] (-1 padded),
 * step limits [nv]; and the mesh broad phase on given positions / displacements -> counts [nv],
 * candidates [nv*8] (-1 padded) with the current poses. */
int rs_check_detect(rs_batch* b, const float* pos, const float* pads, int* counts, int* lists, float* max_step);
int rs_check_broadphase(rs_batch* b, const float* pos, const float* dx, int* counts, int* cand);
/* Timing of the persistent frame kernel with HIP events on the solver stream (no counterpart in the
 * reference): rs_kernel_time returns the summed duration and count of the frame kernels launched since
 * the last query, synchronising on them. */
int rs_set_kernel_timing(rs_batch* b, int on);
int rs_kernel_time(rs_batch* b, double* total_ms, int* launches);
/* Rope-rope broad-phase statistics of the persistent kernel (no counterpart in the reference): returns
 * and clears [skin-list entries written at list builds, exact segment-pair tests, candidate-list entries, list
 * builds per cell] counted since the last call, and turns counting on or off for the next frames. */
int rs_set_chunk_stats(rs_batch* b, int on, long long* out4);

/* ---- Differentiation form. The reference has no adjoint, so these entry points have no counterpart in it.
 * A recorded rollout keeps one checkpoint per frame (x, v, v_last, q, the staged root program and the mesh
 * poses). rs_adjoint recomputes each frame from its checkpoint and runs the hand-written adjoint of the
 * substep backwards. The forward active set is frozen: contact on/off, friction branch, gravity clamp,
 * orientation signs, broad-phase lists and each substep's rope-rope lists. Covered: root program (or constant
 * drive), init, solver iterations, end update with linear drag, plane and SDF contact, rope-rope contact
 * (barrier, friction, step limit). Not covered: quadratic drag; recording refuses it, and ropes or rope-rope
 * contact groups above 512 vertices (RS_INVALID).
 *
 * rs_set_material sets per-vertex EA [N], EI [N m^2] (per segment, entry at a rope's last vertex ignored) and
 * node masses [kg] (pinned entries ignored). Any pointer may be null to keep that field. Unlike the
 * reference's per-vertex setters it is allowed after stepping. The arrays are device memory when on_device is
 * nonzero, host memory otherwise. */
int rs_set_material(rs_batch* b, const float* ea, const float* ei, const float* mass, int on_device);
/* Start recording: each following step records a checkpoint, up to max_frames. */
int rs_tape_begin(rs_batch* b, int max_frames);
/* Stop recording and keep the state after the last recorded frame. rs_adjoint does this by itself. */
int rs_tape_end(rs_batch* b);
int rs_tape_frames(const rs_batch* b);
/* Backward pass over the recorded frames.
 * In (device): dl_dx [frames, num_vertices, 3], dL/d(positions after each frame).
 * Out (device, each may be null): dL/dEA, dL/dEI, dL/dmass [num_vertices]; dL/dp0, dL/dv0, dL/dp1
 * [frames, num_ropes, 3] of each frame's root program (constant drive: p0 = position, v0 = velocity).
 * Out (host): dL/d(linear drag) in *g_drag. In *mismatches, the number of vertices whose recomputed
 * frame-end state differs from the next checkpoint; it is 0 when the material and meshes are unchanged
 * since recording. */
int rs_adjoint(rs_batch* b, const float* dl_dx, float* g_ea, float* g_ei, float* g_mass, double* g_drag,
               float* g_p0, float* g_v0, float* g_p1, int* mismatches);

#ifdef __cplusplus
}
#endif

#endif /* ROPESIM_HIP_H */
