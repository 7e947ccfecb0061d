// Bounding-volume hierarchy for rope-rope detection in large contact groups: what YarnBall builds with its
// Morton-code LBVH (third_party/yarnball/KittenEngine/YarnBall/sim/collision.cu:10-101,
// KittenEngine/KittenGpuLBVH/lbvh.cu), here one radix tree per contact group over 64-bit keys
// (group, 30-bit Morton code), sorted by a stable LSD radix sort of this library (no library sort: the
// frame may be captured into a hipGraph) and queried into the same ascending candidate lists a full scan
// of the group gives.
#pragma once

#include <hip/hip_runtime.h>

#include "rs_types.hpp"

namespace rs {

constexpr int BVH_SORT_THREADS = 256;
constexpr int BVH_SORT_ITEMS = 8;
constexpr int BVH_SORT_TILE = BVH_SORT_THREADS * BVH_SORT_ITEMS;
constexpr int BVH_BOUNDS_BLOCKS = 256;  // partial scene boxes reduced by every key block
constexpr int BVH_STACK = 64;           // traversal stack; a tree path has at most 62 internal nodes
constexpr int BVH_SEG_BLOCK = 256;      // segment-tree entries reduced through 8 levels per workgroup
constexpr int BVH_MAX_LEVELS = 32;      // levels of the box segment tree (n < 2^31 leaves)

struct Bvh {
    int n = 0;          // vertices of the batch (leaves: one per vertex, empty boxes at rope ends)
    int tiles = 0;      // radix-sort tiles of BVH_SORT_TILE keys
    int sortBits = 0;   // key bits the sort covers: 30 Morton bits + the group index bits
    int levels = 0;     // levels of the segment tree (level 0: the leaves)
    int cnt[BVH_MAX_LEVELS] = {};  // entries of each level: cnt[k + 1] = ceil(cnt[k] / 2)
    int off[BVH_MAX_LEVELS] = {};  // first entry of each level in `seg`
    int segEntries = 0;            // entries of all levels
    unsigned long long *key = nullptr, *key2 = nullptr;  // [n] sort keys (ping-pong)
    int *val = nullptr, *val2 = nullptr;                 // [n] sorted position -> vertex (ping-pong)
    int* hist = nullptr;      // [256 * tiles] digit counts per tile, then their prefix within each digit's row
    int* digitTotal = nullptr;  // [256] keys per digit
    int2* child = nullptr;    // [n] children of internal node p (~x: leaf at sorted position x)
    int2* range = nullptr;    // [n] first and last sorted position under internal node p
    float4* box = nullptr;    // [2n] (lo, hi) of internal node p at [2p]
    float4* seg = nullptr;    // [2 * segEntries] (lo, hi) segment tree of box unions over sorted positions:
                              // entry j of level k + 1 = entries 2j, 2j + 1 of level k; level 0 = leaf boxes
    float4* bounds = nullptr; // [2 * BVH_BOUNDS_BLOCKS] partial scene boxes
};

hipError_t bvh_alloc(Bvh& T, int nv, int numGroups);
void bvh_free(Bvh& T);
// Rope-rope detection of every segment through the hierarchy: ncol / coll / maxStep exactly as the full scan
// of the group writes them. `rebuild` re-sorts and rebuilds the trees from the current positions; otherwise
// the boxes are refitted on the trees of the last rebuild (the lists do not depend on the tree).
hipError_t launch_bvh_detect(const Params& P, const Dev& D, const Bvh& T, bool rebuild, hipStream_t st);

}  // namespace rs
