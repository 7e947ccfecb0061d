// Distance grid of a triangle mesh: what the reference bakes with Warp (ropesim/sdf.py:96-170 of the
// reference: `wp.mesh_query_point_sign_winding_number` + `wp.mesh_eval_position` per lattice node), here as
// one brute-force pass of every node against every triangle.
//
// One thread per lattice node; triangles stream through LDS in tiles (each lane reads the same triangle, an
// LDS broadcast), N-body style. Per (node, triangle):
// - squared distance to the closest point on the triangle, by the Voronoi-region method of Ericson,
//   Real-Time Collision Detection (2005), section 5.1.5;
// - for a signed grid, the triangle's signed solid angle at the node (Van Oosterom and Strackee 1983:
//   tan(omega / 2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (a.c)|b| + (b.c)|a|)), summed into the generalized
//   winding number (Jacobson, Kavan and Sorkine-Hornung 2013). |sum| > 2 pi (|w| > 1/2) is inside.
// Warp's tree approximates the winding number; this sum is exact. Node coordinates follow the reference:
// origin + spacing * index in double, rounded to float (sdf.py:61-75, 149).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "ropesim_hip.h"
#include "rs_kernels.hpp"
#include "rs_math.hpp"

namespace rs {

constexpr int BAKE_BLOCK = 256;
constexpr int BAKE_TILE = 256;  // triangles per LDS tile: 3 float4 each, 12 KB

__device__ __forceinline__ float closest_dist2(f3 p, f3 a, f3 b, f3 c) {
    const f3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.f && d2 <= 0.f) return dot(ap, ap);  // vertex a
    const f3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.f && d4 <= d3) return dot(bp, bp);  // vertex b
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) {  // edge ab
        const f3 r = ap - (d1 / (d1 - d3)) * ab;
        return dot(r, r);
    }
    const f3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.f && d5 <= d6) return dot(cp, cp);  // vertex c
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) {  // edge ac
        const f3 r = ap - (d2 / (d2 - d6)) * ac;
        return dot(r, r);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {  // edge bc
        const f3 r = bp - ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
        return dot(r, r);
    }
    const float den = 1.f / (va + vb + vc);  // interior: barycentric (1 - v - w, v, w)
    const f3 r = ap - (vb * den) * ab - (vc * den) * ac;
    return dot(r, r);
}

// Signed solid angle of triangle (a, b, c) seen from the origin (corners already relative to the node).
__device__ __forceinline__ float solid_angle(f3 a, f3 b, f3 c) {
    const float la = len(a), lb = len(b), lc = len(c);
    const float det = dot(a, cross(b, c));
    const float den = la * lb * lc + dot(a, b) * lc + dot(a, c) * lb + dot(b, c) * la;
    return 2.f * atan2f(det, den);
}

__global__ __launch_bounds__(BAKE_BLOCK) void k_bake_mesh(const float4* tri, int ntri, double ox, double oy, double oz,
                                                       double h, int nx, int ny, int nz, int shell, float halfThick,
                                                       float rnd, float* out) {
    __shared__ float4 sT[3 * BAKE_TILE];
    const long long n = (long long)blockIdx.x * BAKE_BLOCK + threadIdx.x;
    const long long total = (long long)nx * ny * nz;
    const bool valid = n < total;
    const long long m = valid ? n : 0;
    const int i = (int)(m / ((long long)ny * nz));
    const int r = (int)(m - (long long)i * ny * nz);
    const int j = r / nz, k = r - (r / nz) * nz;
    const f3 p = mk3((float)(ox + h * i), (float)(oy + h * j), (float)(oz + h * k));
    float best = INFINITY, wsum = 0.f;
    for (int t0 = 0; t0 < ntri; t0 += BAKE_TILE) {
        const int cnt = min(BAKE_TILE, ntri - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < 3 * cnt; e += BAKE_BLOCK) sT[e] = tri[3 * (size_t)t0 + e];
        __syncthreads();
        if (valid) {
            for (int t = 0; t < cnt; t++) {
                const float4 A = sT[3 * t], B = sT[3 * t + 1], C = sT[3 * t + 2];
                const f3 a = mk3(A.x, A.y, A.z), b = mk3(B.x, B.y, B.z), c = mk3(C.x, C.y, C.z);
                best = fminf(best, closest_dist2(p, a, b, c));
                if (!shell) wsum += solid_angle(a - p, b - p, c - p);
            }
        }
    }
    if (!valid) return;
    float d = sqrtf(best);
    if (shell)
        d = fabsf(d) - halfThick;
    else if (fabsf(wsum) > 6.283185307179586f)
        d = -d;
    out[n] = d - rnd;
}

hipError_t launch_bake_mesh(const float4* tri, int ntri, const double* origin, double spacing, int nx, int ny, int nz,
                            int shell, float halfThick, float rnd, float* out, hipStream_t s) {
    const long long total = (long long)nx * ny * nz;
    if (total <= 0) return hipSuccess;
    const long long blocks = (total + BAKE_BLOCK - 1) / BAKE_BLOCK;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_bake_mesh, dim3((unsigned)blocks), dim3(BAKE_BLOCK), 0, s, tri, ntri, origin[0], origin[1],
                       origin[2], spacing, nx, ny, nz, shell, halfThick, rnd, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ the bake without allocation
//
// rs_bake_mesh_sdf_noalloc stages the triangle corners in this fixed array instead of a buffer of its own.
// Entry 0 is the index-check word (an int in its first lane), entries 1 .. 3 * RS_BAKE_NOALLOC_MAX_TRIANGLES the
// corners, three per triangle: 12 MiB of every GPU the library runs on, held until the process ends. Host code
// reaches it only through the address hipGetSymbolAddress returns on each device.
constexpr int kNoAllocCorners = 3 * RS_BAKE_NOALLOC_MAX_TRIANGLES;
static __device__ __align__(256) float4 scratch[1 + kNoAllocCorners];

__global__ void k_clear_word(int* w) {
    if (blockIdx.x == 0 && threadIdx.x == 0) w[0] = 0;
}

// Gather the corners of every triangle from device vertex and index arrays (one thread per corner). An index
// outside [0, nv) marks the check word and reads vertex 0 instead, so every access stays in bounds.
__global__ void k_bake_stage(const float* vtx, int nv, const int* tri, int ntri, float4* corners, int* bad) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 3LL * ntri) return;
    int v = tri[e];
    if (v < 0 || v >= nv) {
        bad[0] = 1;
        v = 0;
    }
    corners[e] = make_float4(vtx[3 * (size_t)v], vtx[3 * (size_t)v + 1], vtx[3 * (size_t)v + 2], 0.f);
}

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
float4* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU, looked up once per device

float4* scratch_address(int dev, std::string& err) {
    if (dev < 0 || dev >= kMaxDevices) {
        err = "bake_mesh_noalloc: device index beyond " + std::to_string(kMaxDevices);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("bake_mesh_noalloc: staging array address: ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch[dev] = static_cast<float4*>(p);
    }
    return g_scratch[dev];
}

}  // namespace

}  // namespace rs

extern "C" int rs_bake_mesh_sdf_noalloc(const float* vertices, int num_vertices, const int* triangles, int num_triangles,
                                        const float* origin3, float spacing, int nx, int ny, int nz, float thickness,
                                        float round, void* stream, float* out) {
    using namespace rs;
    if (!vertices || !triangles || !origin3 || !out) {
        set_error_message("bake_mesh_noalloc: null pointer");
        return RS_INVALID;
    }
    if (num_vertices < 3 || num_triangles < 1) {
        set_error_message("bake_mesh_noalloc needs a triangle mesh");
        return RS_INVALID;
    }
    if (num_triangles > RS_BAKE_NOALLOC_MAX_TRIANGLES) {
        set_error_message(("bake_mesh_noalloc: " + std::to_string(num_triangles) + " triangles exceed the fixed "
                           "staging array of " + std::to_string(RS_BAKE_NOALLOC_MAX_TRIANGLES) +
                           "; use rs_bake_mesh_sdf").c_str());
        return RS_INVALID;
    }
    if (!(spacing > 0.f) || nx < 2 || ny < 2 || nz < 2 ||
        (long long)nx * ny * nz > 0x7fffffffLL * (long long)BAKE_BLOCK) {
        set_error_message("bake_mesh_noalloc: bad lattice");
        return RS_INVALID;
    }
    if (!(thickness >= 0.f) || !(round >= 0.f)) {
        set_error_message("bake_mesh_noalloc: thickness and round must be >= 0");
        return RS_INVALID;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    std::string err;
    float4* base = e == hipSuccess ? scratch_address(dev, err) : nullptr;
    if (!base) {
        set_error_message(e != hipSuccess ? (std::string("bake_mesh_noalloc: ") + hipGetErrorString(e)).c_str()
                                          : err.c_str());
        return RS_DEVICE;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    int* bad = reinterpret_cast<int*>(base);
    float4* corners = base + 1;
    const long long nc = 3LL * num_triangles;
    hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, st, bad);
    hipLaunchKernelGGL(k_bake_stage, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, vertices, num_vertices,
                       triangles, num_triangles, corners, bad);
    e = hipGetLastError();
    if (e == hipSuccess) {
        const double o[3] = {origin3[0], origin3[1], origin3[2]};
        e = launch_bake_mesh(corners, num_triangles, o, (double)spacing, nx, ny, nz, thickness > 0.f ? 1 : 0,
                             (float)(0.5 * (double)thickness), round, out, st);
    }
    int flag = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        set_error_message((std::string("bake_mesh_noalloc: ") + hipGetErrorString(e)).c_str());
        return RS_DEVICE;
    }
    if (flag) {
        set_error_message("triangle indices out of range");
        return RS_INVALID;
    }
    return RS_OK;
}
