// Host-side launchers of the rope kernels (rs_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "rs_bvh.hpp"
#include "rs_types.hpp"

namespace rs {

struct CellVariant {
    int block, vpt;  // workgroup size and vertex slots per thread: a cell holds up to block * vpt vertices
    int strand;      // 1: the strand kernel, one wave per rope of up to 64 * vpt vertices (block 64)
};
extern const CellVariant kCellVariants[];
extern const int kNumCellVariants;
size_t cell_lds_bytes(int block, int vpt);
size_t strand_lds_bytes(int nb);
size_t variant_lds_bytes(int variant);
// Longest-first launch order of `n` cells from their last durations (k_cell writes Dev::cellTime).
hipError_t launch_rank_cells(const unsigned* time, int n, int* order, hipStream_t st);
// The strand kernel (rs_strand.hip): one wave per rope, `order` [nropes] the launch order of the ropes.
hipError_t launch_strand(int nb, const Params& P, const Dev& D, const Drive& Dr, const int* order, int nropes, int S,
                         hipStream_t st);

hipError_t launch_fill_i32(int* p, int v, long n, hipStream_t s);
hipError_t launch_fill_f32(float* p, float v, long n, hipStream_t s);
hipError_t launch_scale_f32(float* p, float v, long n, hipStream_t s);
hipError_t launch_wrench_from_fixed(const unsigned long long* fx, float* out, float scale, long n, hipStream_t s);
hipError_t launch_root_const(const Params& P, const Dev& D, const Drive& Dr, hipStream_t s);
hipError_t launch_gather(const Params& P, const Dev& D, hipStream_t s);
hipError_t launch_reset(const Params& P, const Dev& D, const float4* initX, const float4* initQ, const int* ids,
                        int count, hipStream_t s);
// bvh: rope-rope detection through the per-group hierarchy (rebuilt at s = 0, refitted after), else the full scan
hipError_t launch_plain_substep(const Params& P, Dev& D, const Drive& Dr, int s, int S, hipStream_t st,
                                const Bvh* bvh = nullptr);
hipError_t launch_cell(int variant, const Params& P, const Dev& D, const Drive& Dr, const Cell* cells, const int* order,
                       int ncells, int S, unsigned long long* stats, hipStream_t st);
// Mesh pose slots [nropes * MAX_SDF]: committed pose, target of the next step, and that step's motion.
struct MeshSlots {
    Pose7* cur;
    Pose7* tgt;
    Motion* motion;
};
// Targets of mesh `index` (rows 1 or nropes of pos [3] / quat [4]) or, index < 0, of every mesh from dense
// [nropes, M] rows; teleported ropes (teleAll, or teleFlags[rope] != 0) start the step at their target.
hipError_t launch_mesh_targets(const MeshSlots& S, int index, const float* pos, const float* quat, int rows, int M,
                               const uint8_t* teleFlags, int teleAll, int nropes, hipStream_t s);
// Targets become the committed poses (mesh `index`, or every mesh with index < 0), with world boxes in `poses`.
hipError_t launch_mesh_commit(const MeshSlots& S, Pose* poses, const Grid* grids, int numSdf, int index, int nropes,
                              hipStream_t s);
hipError_t launch_check_detect(const Params& P, const Dev& D, const Bvh* bvh, hipStream_t s);
hipError_t launch_bake_mesh(const float4* tri, int ntri, const double* origin, double spacing, int nx, int ny, int nz,
                            int shell, float halfThick, float rnd, float* out, hipStream_t s);
// Message returned by rs_last_error_message() on the calling thread.
void set_error_message(const char* msg);
hipError_t launch_check_broadphase(const Params& P, const Dev& D, const float4* xs, const float4* dxs, int* counts,
                                   int* cand, hipStream_t s);

}  // namespace rs
