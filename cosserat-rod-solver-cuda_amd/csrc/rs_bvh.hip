// Rope-rope detection through a bounding-volume hierarchy per contact group (the reference's LBVH path:
// YarnBall sim/collision.cu:10-101 builds padded segment boxes, queries the LBVH of lbvh.cu for overlapping
// pairs and keeps those within the summed paddings).
//
// Build, once per frame:
// 1. scene box: partial min/max per block, reduced again by every block of the key kernel;
// 2. keys: (group index << 30) | 30-bit Morton code of the segment midpoint, value = vertex;
// 3. stable LSD radix sort of the keys, 8 bits per pass (digit histograms per tile, a scan per digit row,
//    scatter ranked by wave ballots). The group bits keep every group at its own position range;
// 4. one binary radix tree per group over its sorted keys, by Karras, "Maximizing parallelism in the
//    construction of BVHs, octrees, and k-d trees" (HPG 2012), section 4: internal node p of the group
//    starting at position lo covers a key range ending at p, the root is node lo; equal keys are told apart
//    by their position (the augmented-key rule of the paper).
// Every substep, the boxes: the leaf boxes from the current positions and paddings in sorted order are level 0
// of a segment tree of box unions (each workgroup takes 256 entries of a level up 8 levels in LDS); the box of
// internal node p is the union over its contiguous leaf range, O(log n) segment-tree entries (no atomics, no
// ordering between workgroups). Then one traversal per segment, threads in sorted (Morton) order so that
// the lanes of a wave walk nearby parts of the tree.
//
// The lists are the full scan's lists: the traversal visits every leaf whose box meets the segment's box (a
// superset of the pairs within the summed paddings, widened by a rounding margin), the exact test is the
// scan's own, and partners are inserted in ascending order keeping the lowest MAX_COLL. The tree only
// decides how fast.
#include <hip/hip_runtime.h>

#include <cmath>

#include "rs_bvh.hpp"
#include "rs_device.hpp"
#include "rs_kernels.hpp"

namespace rs {

using u64 = unsigned long long;

namespace {

constexpr float BVH_MARGIN_ABS = 1e-6f;  // leaf boxes are widened so float rounding of the box corners
constexpr float BVH_MARGIN_REL = 1e-5f;  // never drops a pair the exact test keeps

inline int nblocks(long n, int b) { return (int)((n + b - 1) / b); }

__device__ __forceinline__ void group_range(const Params& P, const Dev& D, int i, int& lo, int& hi) {
    const int G = P.groupRopes;
    const int g = D.ropeOf[i] / G;
    lo = D.roff[min(g * G, P.nropes)];
    hi = D.roff[min((g + 1) * G, P.nropes)];
}

// Padded box of segment (i, i + 1); empty (lo > hi) at a rope's last vertex.
__device__ __forceinline__ void segment_box(const Dev& D, int i, f3& lo, f3& hi) {
    lo = mk3(INFINITY, INFINITY, INFINITY);
    hi = mk3(-INFINITY, -INFINITY, -INFINITY);
    if (!(flags_of(D.prm[i]) & HAS_NEXT)) return;
    const f3 a = ld3(D.x[i]), b = ld3(D.x[i + 1]);
    lo = mk3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z));
    hi = mk3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z));
    const float mag = fmaxf(fmaxf(fmaxf(fabsf(lo.x), fabsf(lo.y)), fabsf(lo.z)),
                            fmaxf(fmaxf(fabsf(hi.x), fabsf(hi.y)), fabsf(hi.z)));
    const float w = fmaxf(D.pad[i], D.pad[i + 1]) + (BVH_MARGIN_ABS + BVH_MARGIN_REL * mag);
    lo = mk3(lo.x - w, lo.y - w, lo.z - w);
    hi = mk3(hi.x + w, hi.y + w, hi.z + w);
}

__device__ __forceinline__ bool boxes_meet(f3 alo, f3 ahi, float4 blo, float4 bhi) {
    return alo.x <= bhi.x && blo.x <= ahi.x && alo.y <= bhi.y && blo.y <= ahi.y && alo.z <= bhi.z && blo.z <= ahi.z;
}

// 10 bits spread to every third bit of 30.
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

}  // namespace

// Partial scene boxes of the vertex positions (grid-stride, one box per block).
__global__ __launch_bounds__(256) void k_bvh_bounds(int n, const float4* x, float4* part) {
    __shared__ float s[6][256];
    const int t = threadIdx.x;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * 256 + t; i < n; i += gridDim.x * 256) {
        const float4 p = x[i];
        lo[0] = fminf(lo[0], p.x);
        lo[1] = fminf(lo[1], p.y);
        lo[2] = fminf(lo[2], p.z);
        hi[0] = fmaxf(hi[0], p.x);
        hi[1] = fmaxf(hi[1], p.y);
        hi[2] = fmaxf(hi[2], p.z);
    }
    for (int c = 0; c < 3; c++) {
        s[c][t] = lo[c];
        s[3 + c][t] = hi[c];
    }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o)
            for (int c = 0; c < 3; c++) {
                s[c][t] = fminf(s[c][t], s[c][t + o]);
                s[3 + c][t] = fmaxf(s[3 + c][t], s[3 + c][t + o]);
            }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = make_float4(s[0][0], s[1][0], s[2][0], 0.f);
        part[2 * blockIdx.x + 1] = make_float4(s[3][0], s[4][0], s[5][0], 0.f);
    }
}

// Sort keys of every vertex: its group above the Morton code of its segment's midpoint (its own position
// at a rope end, whose box is empty) in the scene box.
__global__ __launch_bounds__(256) void k_bvh_keys(Params P, Dev D, Bvh T) {
    static_assert(BVH_BOUNDS_BLOCKS == 256, "one partial box per thread");
    __shared__ float s[6][256];
    const int t = threadIdx.x;
    const float4 a = T.bounds[2 * t], b = T.bounds[2 * t + 1];
    s[0][t] = a.x;
    s[1][t] = a.y;
    s[2][t] = a.z;
    s[3][t] = b.x;
    s[4][t] = b.y;
    s[5][t] = b.z;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (t < o)
            for (int c = 0; c < 3; c++) {
                s[c][t] = fminf(s[c][t], s[c][t + o]);
                s[3 + c][t] = fmaxf(s[3 + c][t], s[3 + c][t + o]);
            }
        __syncthreads();
    }
    const int i = blockIdx.x * 256 + t;
    if (i >= P.nv) return;
    f3 c = ld3(D.x[i]);
    if (flags_of(D.prm[i]) & HAS_NEXT) c = 0.5f * (c + ld3(D.x[i + 1]));
    uint32_t m = 0;
    const float lo[3] = {s[0][0], s[1][0], s[2][0]};
    const float hi[3] = {s[3][0], s[4][0], s[5][0]};
    const float cc[3] = {c.x, c.y, c.z};
    for (int k = 0; k < 3; k++) {
        const float ext = hi[k] - lo[k];
        const float sc = ext > 0.f && ext < INFINITY ? 1024.f / ext : 0.f;
        const float q = fminf(fmaxf((cc[k] - lo[k]) * sc, 0.f), 1023.f);  // NaN -> 0
        m |= spread3((uint32_t)q) << (2 - k);
    }
    const u64 g = (u64)(D.ropeOf[i] / P.groupRopes);
    T.key[i] = (g << 30) | m;
    T.val[i] = i;
}

// Digit counts of one tile: hist[digit * tiles + tile].
__global__ __launch_bounds__(BVH_SORT_THREADS) void k_sort_hist(const u64* key, int n, int shift, int* hist,
                                                                int tiles) {
    __shared__ int cnt[256];
    const int t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    const int base = blockIdx.x * BVH_SORT_TILE;
    for (int r = 0; r < BVH_SORT_ITEMS; r++) {
        const int i = base + r * BVH_SORT_THREADS + t;
        if (i < n) atomicAdd(&cnt[(int)((key[i] >> shift) & 255u)], 1);
    }
    __syncthreads();
    hist[t * tiles + blockIdx.x] = cnt[t];
}

// Exclusive prefix of each digit's row of tile counts, in place (one workgroup per digit, rows in chunks of
// 256 tiles), and the row's total.
__global__ __launch_bounds__(256) void k_sort_scan(int* hist, int tiles, int* digitTotal) {
    __shared__ int part[256];
    const int t = threadIdx.x;
    int* row = hist + (size_t)blockIdx.x * tiles;
    int carry = 0;
    for (int c0 = 0; c0 < tiles; c0 += 256) {
        const int j = c0 + t;
        const int v = j < tiles ? row[j] : 0;
        part[t] = v;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int u = t >= o ? part[t - o] : 0;
            __syncthreads();
            part[t] += u;
            __syncthreads();
        }
        if (j < tiles) row[j] = carry + part[t] - v;
        carry += part[255];
        __syncthreads();
    }
    if (t == 0) digitTotal[blockIdx.x] = carry;
}

// Stable scatter of one tile: items go out in (round, wave, lane) order, which is their input order.
// A lane's rank among the earlier lanes of its wave with the same digit comes from eight ballots. A digit's
// base is the exclusive prefix of the digit totals (formed here) plus this tile's row prefix.
__global__ __launch_bounds__(BVH_SORT_THREADS) void k_sort_scatter(const u64* kin, const int* vin, u64* kout,
                                                                   int* vout, int n, int shift, const int* hist,
                                                                   int tiles, const int* digitTotal) {
    static_assert(BVH_SORT_THREADS == 256, "one digit per thread");
    constexpr int W = BVH_SORT_THREADS / 64;
    __shared__ int run[256];
    __shared__ int wcnt[W][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    {
        const int tot = digitTotal[t];
        run[t] = tot;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            const int u = t >= o ? run[t - o] : 0;
            __syncthreads();
            run[t] += u;
            __syncthreads();
        }
        run[t] = run[t] - tot + hist[t * tiles + blockIdx.x];
    }
    const u64 below = (1ull << lane) - 1ull;
    const int base = blockIdx.x * BVH_SORT_TILE;
    for (int r = 0; r < BVH_SORT_ITEMS; r++) {
        const int i = base + r * BVH_SORT_THREADS + t;
        const bool valid = i < n;
        u64 k = 0;
        int v = 0, dg = 0;
        if (valid) {
            k = kin[i];
            v = vin[i];
            dg = (int)((k >> shift) & 255u);
        }
#pragma unroll
        for (int c = 0; c < W; c++) wcnt[c][t] = 0;
        __syncthreads();
        u64 peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const int on = (dg >> bit) & 1;
            const u64 bb = __ballot(on);
            peers &= on ? bb : ~bb;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) wcnt[w][dg] = __popcll(peers);
        __syncthreads();
        if (valid) {
            int off = run[dg] + rank;
            for (int c = 0; c < w; c++) off += wcnt[c][dg];
            kout[off] = k;
            vout[off] = v;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int c = 0; c < W; c++) add += wcnt[c][t];
        run[t] += add;
    }
}

namespace {
// Length of the common prefix of the augmented keys at sorted positions a and b (-1 outside [lo, hi)).
__device__ __forceinline__ int bvh_delta(const u64* key, int a, int b, int lo, int hi) {
    if (b < lo || b >= hi) return -1;
    const u64 ka = key[a], kb = key[b];
    if (ka != kb) return __clzll((long long)(ka ^ kb));
    return 64 + __clz(a ^ b);
}
}  // namespace

// Internal node p of its group's tree (positions lo .. hi - 2; the root is lo).
__global__ __launch_bounds__(256) void k_bvh_build(Params P, Dev D, Bvh T) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P.nv) return;
    int lo, hi;
    group_range(P, D, p, lo, hi);
    if (p >= hi - 1) return;  // position hi - 1 has no internal node (nor has a group of one vertex)
    const u64* key = T.key;
    const int d = bvh_delta(key, p, p + 1, lo, hi) >= bvh_delta(key, p, p - 1, lo, hi) ? 1 : -1;
    const int dmin = bvh_delta(key, p, p - d, lo, hi);
    int lmax = 2;
    while (bvh_delta(key, p, p + lmax * d, lo, hi) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (bvh_delta(key, p, p + (l + t) * d, lo, hi) > dmin) l += t;
    const int j = p + l * d;
    const int dnode = bvh_delta(key, p, j, lo, hi);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (bvh_delta(key, p, p + (s + t) * d, lo, hi) > dnode) s += t;
    } while (t > 1);
    const int gamma = p + s * d + min(d, 0);
    const int left = min(p, j) == gamma ? ~gamma : gamma;
    const int right = max(p, j) == gamma + 1 ? ~(gamma + 1) : gamma + 1;
    T.child[p] = make_int2(left, right);
    T.range[p] = make_int2(min(p, j), max(p, j));
}

namespace {
__device__ __forceinline__ float4 lo4(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.f);
}
__device__ __forceinline__ float4 hi4(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.f);
}
constexpr float4 EMPTY_LO = {INFINITY, INFINITY, INFINITY, 0.f};
constexpr float4 EMPTY_HI = {-INFINITY, -INFINITY, -INFINITY, 0.f};

// Levels k0 + 1 .. k0 + 8 above this workgroup's BVH_SEG_BLOCK entries (a, b) of level k0.
__device__ __forceinline__ void seg_up(const Bvh& T, int k0, float4 a, float4 b, float4* slo, float4* shi) {
    const int t = threadIdx.x;
    slo[t] = a;
    shi[t] = b;
    for (int step = 1; step <= 8 && k0 + step < T.levels; step++) {
        const int half = BVH_SEG_BLOCK >> step;
        const bool on = t < half;
        __syncthreads();
        float4 l = EMPTY_LO, h = EMPTY_HI;
        if (on) {
            l = lo4(slo[2 * t], slo[2 * t + 1]);
            h = hi4(shi[2 * t], shi[2 * t + 1]);
        }
        __syncthreads();
        if (on) {
            slo[t] = l;
            shi[t] = h;
            const int k = k0 + step, j = blockIdx.x * half + t;
            if (j < T.cnt[k]) {
                T.seg[2 * (T.off[k] + j)] = l;
                T.seg[2 * (T.off[k] + j) + 1] = h;
            }
        }
    }
}
}  // namespace

// Leaf boxes in sorted order (segment-tree level 0) and levels 1..8 above them.
__global__ __launch_bounds__(BVH_SEG_BLOCK) void k_bvh_leaves(Params P, Dev D, Bvh T) {
    __shared__ float4 slo[BVH_SEG_BLOCK], shi[BVH_SEG_BLOCK];
    const int p = blockIdx.x * BVH_SEG_BLOCK + threadIdx.x;
    float4 a = EMPTY_LO, b = EMPTY_HI;
    if (p < P.nv) {
        f3 lo, hi;
        segment_box(D, T.val[p], lo, hi);
        a = make_float4(lo.x, lo.y, lo.z, 0.f);
        b = make_float4(hi.x, hi.y, hi.z, 0.f);
        T.seg[2 * p] = a;
        T.seg[2 * p + 1] = b;
    }
    seg_up(T, 0, a, b, slo, shi);
}

// Levels k0 + 1 .. k0 + 8 of the segment tree from level k0.
__global__ __launch_bounds__(BVH_SEG_BLOCK) void k_bvh_levels(Bvh T, int k0) {
    __shared__ float4 slo[BVH_SEG_BLOCK], shi[BVH_SEG_BLOCK];
    const int j = blockIdx.x * BVH_SEG_BLOCK + threadIdx.x;
    float4 a = EMPTY_LO, b = EMPTY_HI;
    if (j < T.cnt[k0]) {
        a = T.seg[2 * (T.off[k0] + j)];
        b = T.seg[2 * (T.off[k0] + j) + 1];
    }
    seg_up(T, k0, a, b, slo, shi);
}

// Box of internal node p: the union of the leaf boxes of its range, from the segment tree (half-open range
// [l, r) at level k: an odd l is a right child whose sibling lies outside, an odd r leaves a left child).
__global__ __launch_bounds__(256) void k_bvh_nodes(Params P, Dev D, Bvh T) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P.nv) return;
    int lo, hi;
    group_range(P, D, p, lo, hi);
    if (p >= hi - 1) return;
    const int2 rg = T.range[p];
    int l = rg.x, r = rg.y + 1, k = 0;
    float4 blo = EMPTY_LO, bhi = EMPTY_HI;
    while (l < r) {
        if (l & 1) {
            blo = lo4(blo, T.seg[2 * (T.off[k] + l)]);
            bhi = hi4(bhi, T.seg[2 * (T.off[k] + l) + 1]);
            l++;
        }
        if (r & 1) {
            r--;
            blo = lo4(blo, T.seg[2 * (T.off[k] + r)]);
            bhi = hi4(bhi, T.seg[2 * (T.off[k] + r) + 1]);
        }
        l >>= 1;
        r >>= 1;
        k++;
    }
    T.box[2 * p] = blo;
    T.box[2 * p + 1] = bhi;
}

// Candidate list and step limit of segment i (k_detect's results): traverse the group's tree, run the
// scan's exact test on every leaf whose box meets the segment's, insert hits in ascending order.
__global__ __launch_bounds__(128) void k_bvh_query(Params P, Dev D, Bvh T) {
    const int q = blockIdx.x * 128 + threadIdx.x;
    if (q >= P.nv) return;
    const int i = T.val[q];
    if (!(flags_of(D.prm[i]) & HAS_NEXT)) {
        D.ncol[i] = 0;
        D.maxStep[i] = INFINITY;
        return;
    }
    const int rope = D.ropeOf[i];
    int lo, hi;
    group_range(P, D, i, lo, hi);
    f3 qlo, qhi;
    segment_box(D, i, qlo, qhi);
    const f3 a0 = ld3(D.x[i]);
    const f3 a1 = ld3(D.x[i + 1]) - a0;
    const float ri = fmaxf(D.pad[i], D.pad[i + 1]);
    int* row = D.coll + (size_t)i * MAX_COLL;
    int cnt = 0;
    int stack[BVH_STACK];  // private (scratch): an LDS stack of 256 B per lane would cap a CU at 10 waves
    int sp = 0;
    int node = hi - lo >= 2 ? lo : ~lo;
    for (;;) {
        if (node < 0) {
            const int j = T.val[~node];
            if ((flags_of(D.prm[j]) & HAS_NEXT) && !(D.ropeOf[j] == rope && abs(j - i) <= 2)) {
                const f3 b0 = ld3(D.x[j]) - a0, b1 = ld3(D.x[j + 1]) - a0;
                const float r = ri + fmaxf(D.pad[j], D.pad[j + 1]);
                float d2;
                if (pair_hit(a1, b0, b1, r, d2)) {
                    const int m = min(cnt, MAX_COLL);
                    if (m < MAX_COLL || j < row[MAX_COLL - 1]) {
                        int s = m < MAX_COLL ? m : MAX_COLL - 1;
                        while (s > 0 && row[s - 1] > j) {
                            row[s] = row[s - 1];
                            s--;
                        }
                        row[s] = j;
                    }
                    cnt++;
                }
            }
        } else {
            const int2 c = T.child[node];
            const float4* ba = c.x < 0 ? T.seg + 2 * ~c.x : T.box + 2 * c.x;
            const float4* bb = c.y < 0 ? T.seg + 2 * ~c.y : T.box + 2 * c.y;
            const bool ha = boxes_meet(qlo, qhi, ba[0], ba[1]);
            const bool hb = boxes_meet(qlo, qhi, bb[0], bb[1]);
            if (ha && hb) {
                if (sp == BVH_STACK) {  // deeper than any tree over these keys can be
                    D.err[0] = ERR_BVH_DEPTH;
                    break;
                }
                stack[sp++] = c.y;
                node = c.x;
                continue;
            }
            if (ha || hb) {
                node = ha ? c.x : c.y;
                continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp];
    }
    // step limit over the kept partners (the scan takes it over the same first MAX_COLL hits)
    float minDist = D.pad[i] - P.detR;
    const int m = min(cnt, MAX_COLL);
    for (int c = 0; c < m; c++) {
        const int j = row[c];
        const f3 b0 = ld3(D.x[j]) - a0, b1 = ld3(D.x[j + 1]) - a0;
        const float r = ri + fmaxf(D.pad[j], D.pad[j + 1]);
        float d2;
        pair_hit(a1, b0, b1, r, d2);
        minDist = fminf(minDist, 0.45f * sqrtf(d2));
    }
    if (cnt > MAX_COLL) D.err[0] = ERR_MAX_COLLISIONS;
    D.ncol[i] = m;
    D.maxStep[i] = minDist;
}

hipError_t bvh_alloc(Bvh& T, int nv, int numGroups) {
    bvh_free(T);
    if (nv <= 0) return hipSuccess;
    int gbits = 0;
    while ((1ll << gbits) < (long long)numGroups) gbits++;
    T.n = nv;
    T.tiles = nblocks(nv, BVH_SORT_TILE);
    T.sortBits = 30 + gbits;
    T.levels = 0;
    T.segEntries = 0;
    for (int c = nv;; c = (c + 1) / 2) {
        T.cnt[T.levels] = c;
        T.off[T.levels] = T.segEntries;
        T.segEntries += c;
        T.levels++;
        if (c == 1) break;
    }
    const size_t n = nv;
    hipError_t e = hipSuccess;
    auto get = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    get((void**)&T.key, n * 8);
    get((void**)&T.key2, n * 8);
    get((void**)&T.val, n * 4);
    get((void**)&T.val2, n * 4);
    get((void**)&T.hist, (size_t)256 * T.tiles * 4);
    get((void**)&T.digitTotal, 256 * 4);
    get((void**)&T.child, n * 8);
    get((void**)&T.range, n * 8);
    get((void**)&T.box, 2 * n * 16);
    get((void**)&T.seg, (size_t)2 * T.segEntries * 16);
    get((void**)&T.bounds, 2 * BVH_BOUNDS_BLOCKS * 16);
    if (e != hipSuccess) bvh_free(T);
    return e;
}

void bvh_free(Bvh& T) {
    for (void* p : {(void*)T.key, (void*)T.key2, (void*)T.val, (void*)T.val2, (void*)T.hist, (void*)T.digitTotal, (void*)T.child,
                    (void*)T.range, (void*)T.box, (void*)T.seg, (void*)T.bounds})
        if (p) (void)hipFree(p);
    T = Bvh{};
}

hipError_t launch_bvh_detect(const Params& P, const Dev& D, const Bvh& T, bool rebuild, hipStream_t st) {
    const int n = P.nv;
    if (n <= 0) return hipSuccess;
    if (T.n != n) return hipErrorInvalidValue;
    if (rebuild) {
        hipLaunchKernelGGL(k_bvh_bounds, dim3(BVH_BOUNDS_BLOCKS), dim3(256), 0, st, n, (const float4*)D.x, T.bounds);
        hipLaunchKernelGGL(k_bvh_keys, dim3(nblocks(n, 256)), dim3(256), 0, st, P, D, T);
        int passes = (T.sortBits + 7) / 8;
        passes += passes & 1;  // an even count leaves the sorted keys in T.key / T.val
        const u64* kin = T.key;
        const int* vin = T.val;
        u64* kout = T.key2;
        int* vout = T.val2;
        for (int pass = 0; pass < passes; pass++) {
            const int shift = 8 * pass;
            hipLaunchKernelGGL(k_sort_hist, dim3(T.tiles), dim3(BVH_SORT_THREADS), 0, st, kin, n, shift, T.hist,
                               T.tiles);
            hipLaunchKernelGGL(k_sort_scan, dim3(256), dim3(256), 0, st, T.hist, T.tiles, T.digitTotal);
            hipLaunchKernelGGL(k_sort_scatter, dim3(T.tiles), dim3(BVH_SORT_THREADS), 0, st, kin, vin, kout, vout, n,
                               shift, (const int*)T.hist, T.tiles, (const int*)T.digitTotal);
            const u64* k2 = kout;
            const int* v2 = vout;
            kout = const_cast<u64*>(kin);
            vout = const_cast<int*>(vin);
            kin = k2;
            vin = v2;
        }
        hipLaunchKernelGGL(k_bvh_build, dim3(nblocks(n, 256)), dim3(256), 0, st, P, D, T);
    }
    hipLaunchKernelGGL(k_bvh_leaves, dim3(nblocks(n, BVH_SEG_BLOCK)), dim3(BVH_SEG_BLOCK), 0, st, P, D, T);
    for (int k0 = 8; k0 < T.levels - 1; k0 += 8)
        hipLaunchKernelGGL(k_bvh_levels, dim3(nblocks(T.cnt[k0], BVH_SEG_BLOCK)), dim3(BVH_SEG_BLOCK), 0, st, T, k0);
    hipLaunchKernelGGL(k_bvh_nodes, dim3(nblocks(n, 256)), dim3(256), 0, st, P, D, T);
    hipLaunchKernelGGL(k_bvh_query, dim3(nblocks(n, 128)), dim3(128), 0, st, P, D, T);
    return hipGetLastError();
}

}  // namespace rs
