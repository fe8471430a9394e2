// bvh_host.cpp — host BVH builder: binned-SAH binary tree, collapsed into 128-byte BVH4 nodes.
//
// No reference counterpart (the reference intersects its 5 spheres by brute force,
// src/gpu/primaryRay.comp:26-32). Binary stage: centroid binning with `bins` bins on each
// axis, SAH cost Ct*A + Ci*(nL*AL + nR*AR) with Ct = Ci = 1, leaves of at most `max_leaf`
// triangles (a node of <= max_leaf triangles becomes a leaf unless a split is cheaper),
// median split by index when all centroids coincide, stable partitions. Subtrees above
// 16k triangles are built on separate threads; the result is independent of the thread
// count because every partition is stable and the wide layout is emitted in BFS order.
// Collapse: each wide node takes its binary node's two children and repeatedly opens the
// internal candidate of largest surface area until it holds four.
#include "internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();
constexpr uint32_t kMaxBins = 64;

struct AABB {
    float lo[3], hi[3];
    void clear() { for (int k = 0; k < 3; ++k) { lo[k] = kInf; hi[k] = -kInf; } }
    void grow(const AABB& o) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], o.lo[k]); hi[k] = std::max(hi[k], o.hi[k]); }
    }
    void grow_point(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BinNode {
    AABB box;
    uint32_t left;          // internal: index of left child (right = left + 1)
    uint32_t first, count;  // leaf: range in prim order
};

struct BinaryBuilder {
    const float* verts;
    uint32_t max_leaf, bins;
    std::vector<AABB> pbox;
    std::vector<float> cent;       // 3 per prim
    std::vector<uint32_t> prim;    // permutation
    std::vector<uint32_t> scratch; // partition buffer (same size as prim)
    std::vector<BinNode> nodes;
    std::atomic<uint32_t> next_node{1};
    std::atomic<int> live_threads{0};
    int max_threads = 1;

    uint32_t alloc_pair() { return next_node.fetch_add(2); }

    void build(uint32_t ni, uint32_t begin, uint32_t end) {
        AABB bounds, cb;
        bounds.clear(); cb.clear();
        for (uint32_t i = begin; i < end; ++i) {
            uint32_t p = prim[i];
            bounds.grow(pbox[p]);
            cb.grow_point(&cent[3 * p]);
        }
        BinNode& node = nodes[ni];
        node.box = bounds;
        const uint32_t n = end - begin;
        auto leaf = [&]() { node.left = 0; node.first = begin; node.count = n; };
        if (n == 1) { leaf(); return; }

        float best = kInf; int best_axis = -1, best_bin = -1;
        AABB bbox[kMaxBins];
        uint32_t bcnt[kMaxBins];
        float right_area[kMaxBins];
        uint32_t right_cnt[kMaxBins];
        for (int axis = 0; axis < 3; ++axis) {
            const float extent = cb.hi[axis] - cb.lo[axis];
            if (!(extent > 0.0f)) continue;
            const float scale = (float)bins / extent;
            if (!(scale <= std::numeric_limits<float>::max())) continue;  // extent below ~bins / FLT_MAX
            for (uint32_t b = 0; b < bins; ++b) { bbox[b].clear(); bcnt[b] = 0; }
            for (uint32_t i = begin; i < end; ++i) {
                uint32_t p = prim[i];
                int b = (int)((cent[3 * p + axis] - cb.lo[axis]) * scale);
                b = std::min(std::max(b, 0), (int)bins - 1);
                bcnt[b]++;
                bbox[b].grow(pbox[p]);
            }
            AABB acc; acc.clear();
            uint32_t c = 0;
            for (int b = (int)bins - 1; b > 0; --b) {
                acc.grow(bbox[b]); c += bcnt[b];
                right_area[b] = acc.area(); right_cnt[b] = c;
            }
            acc.clear(); c = 0;
            for (uint32_t b = 0; b + 1 < bins; ++b) {
                acc.grow(bbox[b]); c += bcnt[b];
                if (c == 0 || right_cnt[b + 1] == 0) continue;
                float cost = acc.area() * (float)c + right_area[b + 1] * (float)right_cnt[b + 1];
                if (cost < best) { best = cost; best_axis = axis; best_bin = (int)b; }
            }
        }
        const float area = bounds.area();
        if (n <= max_leaf) {
            if (best_axis < 0 || (float)n * area <= area + best) { leaf(); return; }
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = begin + n / 2;
        } else {
            const float scale = (float)bins / (cb.hi[best_axis] - cb.lo[best_axis]);
            uint32_t l = begin, r = begin;
            for (uint32_t i = begin; i < end; ++i) {
                uint32_t p = prim[i];
                int b = (int)((cent[3 * p + best_axis] - cb.lo[best_axis]) * scale);
                b = std::min(std::max(b, 0), (int)bins - 1);
                if (b <= best_bin) prim[l++] = p; else scratch[r++] = p;
            }
            std::memcpy(&prim[l], &scratch[begin], (size_t)(r - begin) * sizeof(uint32_t));
            mid = l;
        }
        const uint32_t left = alloc_pair();
        node.left = left; node.count = 0;
        const bool fork = (end - begin) > 16384 && live_threads.load() < max_threads;
        if (fork) {
            live_threads++;
            std::thread th([this, left, begin, mid]() { build(left, begin, mid); live_threads--; });
            build(left + 1, mid, end);
            th.join();
        } else {
            build(left, begin, mid);
            build(left + 1, mid, end);
        }
    }
};

}  // namespace

void build_bvh4(const float* verts, uint32_t ntri, uint32_t max_leaf, uint32_t bins, HostBVH& out) {
    auto t0 = std::chrono::steady_clock::now();
    out = HostBVH();
    BinaryBuilder bb;
    bb.verts = verts;
    bb.max_leaf = max_leaf;
    bb.bins = bins;
    bb.max_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    bb.pbox.resize(ntri);
    bb.cent.resize((size_t)ntri * 3);
    bb.prim.resize(ntri);
    bb.scratch.resize(ntri);
    for (uint32_t t = 0; t < ntri; ++t) {
        AABB b; b.clear();
        for (int v = 0; v < 3; ++v) b.grow_point(verts + (size_t)t * 9 + 3 * v);
        // + 0.0f turns -0 into +0, so that every later min/max is independent of the order
        // in which it meets +0 and -0 (the device builder reduces in another order).
        for (int k = 0; k < 3; ++k) { b.lo[k] = b.lo[k] + 0.0f; b.hi[k] = b.hi[k] + 0.0f; }
        bb.pbox[t] = b;
        for (int k = 0; k < 3; ++k) bb.cent[3 * (size_t)t + k] = (b.lo[k] + b.hi[k]) * 0.5f + 0.0f;
        bb.prim[t] = t;
    }
    WNode empty_node;
    for (int c = 0; c < 4; ++c) {
        empty_node.lox[c] = empty_node.loy[c] = empty_node.loz[c] = kInf;
        empty_node.hix[c] = empty_node.hiy[c] = empty_node.hiz[c] = -kInf;
        empty_node.child[c] = kEmptyChild;
        empty_node.meta[c] = 0;
    }
    if (ntri == 0) {
        out.nodes.push_back(empty_node);
        out.depth = 1;
        return;
    }
    bb.nodes.resize((size_t)ntri * 2);
    bb.build(0, 0, ntri);
    out.binary_nodes = bb.next_node.load();

    // Collapse to BVH4 in BFS order.
    struct Item { uint32_t bin; uint32_t depth; };
    std::vector<Item> queue;
    queue.push_back(Item{0, 1});
    out.nodes.reserve(out.binary_nodes / 2 + 1);
    out.tri4.reserve((size_t)ntri * 12);
    uint32_t tri_cursor = 0;
    auto emit_leaf = [&](const BinNode& leaf) -> uint32_t {
        uint32_t first = tri_cursor;
        for (uint32_t i = 0; i < leaf.count; ++i) {
            uint32_t id = bb.prim[leaf.first + i];
            const float* v = verts + (size_t)id * 9;
            float rec[12] = {v[0], v[1], v[2], rt::u2f(id), v[3], v[4], v[5], 0.0f, v[6], v[7], v[8], 0.0f};
            out.tri4.insert(out.tri4.end(), rec, rec + 12);
            ++tri_cursor;
        }
        out.leaves++;
        return kLeafBit | (first << 3) | (leaf.count - 1);
    };
    for (size_t qi = 0; qi < queue.size(); ++qi) {
        const Item it = queue[qi];
        const BinNode& bn = bb.nodes[it.bin];
        out.depth = std::max(out.depth, it.depth);
        uint32_t cand[4];
        int nc = 0;
        if (bn.count) {
            cand[nc++] = it.bin;  // the whole tree is one leaf
        } else {
            cand[nc++] = bn.left;
            cand[nc++] = bn.left + 1;
            while (nc < 4) {
                int pick = -1; float best_area = -1.0f;
                for (int c = 0; c < nc; ++c) {
                    const BinNode& x = bb.nodes[cand[c]];
                    if (x.count) continue;
                    float a = x.box.area();
                    if (a > best_area) { best_area = a; pick = c; }
                }
                if (pick < 0) break;
                uint32_t opened = cand[pick];
                cand[pick] = bb.nodes[opened].left;
                cand[nc++] = bb.nodes[opened].left + 1;
            }
        }
        WNode w = empty_node;
        for (int c = 0; c < nc; ++c) {
            const BinNode& x = bb.nodes[cand[c]];
            w.lox[c] = x.box.lo[0]; w.hix[c] = x.box.hi[0];
            w.loy[c] = x.box.lo[1]; w.hiy[c] = x.box.hi[1];
            w.loz[c] = x.box.lo[2]; w.hiz[c] = x.box.hi[2];
            if (x.count) {
                w.child[c] = emit_leaf(x);
            } else {
                w.child[c] = (uint32_t)queue.size();  // wide index = BFS queue position
                queue.push_back(Item{cand[c], it.depth + 1});
            }
        }
        out.nodes.push_back(w);  // out.nodes.size() == qi before this push
    }
    auto t1 = std::chrono::steady_clock::now();
    out.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
}
