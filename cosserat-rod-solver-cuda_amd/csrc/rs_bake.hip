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

}  // namespace rs
