// bvh_build.hip — device BVH builder. No reference counterpart (the reference intersects its
// five spheres by brute force, src/gpu/primaryRay.comp:26-32).
//
// It computes the host builder's tree (bvh_host.cpp: binned SAH with `bins` centroid bins per
// axis, SAH cost A*n_L*A_L + ... with Ct = Ci = 1, leaves of at most max_leaf triangles, the
// median split by position when all centroids coincide, stable partitions, then the BVH4
// collapse that opens the largest-area internal candidate, in BFS order) and emits the same
// bytes: every bound is a min/max (exact in any order once -0 is folded into +0), every count
// an integer, every SAH sweep runs in the host's order, and every partition is stable.
//
// Binary stage, one pass per tree level over the level's segments (a node and its contiguous
// range of the triangle permutation), by segment size:
//   big   (> 4096)  binned by several workgroups into partial bins, reduced per segment by one
//                   workgroup that also evaluates the SAH; stable partition by chunk counts,
//                   a per-segment scan of them and a scatter;
//   mid   (65..4096) one workgroup per segment: bins in LDS, SAH, stable partition via a block
//                   scan;
//   small (<= 64)   one wave finishes the whole subtree: one triangle per lane, bounds by wave
//                   reductions, bins in LDS, the SAH with one lane per (axis, bin) using
//                   segmented prefix/suffix scans, partitions by ballot ranks.
// Children's centroid bounds come from the parent's bins; every node's box is reduced from its
// own triangles. Node slots and next-level lists are assigned by exclusive scans in segment
// order, so the binary tree (and its numbering) is deterministic.
// Collapse stage, one pass per wide level: each wide node picks its <= 4 candidates, scans of
// (internal children, leaf triangles) over the level give the BFS indices and the leaf-ordered
// triangle positions, and the wide nodes and triangle records are written.
#include "internal.h"

#include <chrono>
#include <vector>

namespace bvhdev {

constexpr int kThreads = 256;
constexpr uint32_t kBig = 4096;
constexpr uint32_t kSmall = 64;
constexpr uint32_t kChunk = 4096;
constexpr int kMaxBins = 64;
constexpr int kBinWords = 13;                 // lo[3] hi[3] clo[3] chi[3] count
constexpr uint32_t kKeyPosInf = 0xFF800000u;  // fkey(+inf)
constexpr uint32_t kKeyNegInf = 0x007FFFFFu;  // fkey(-inf)

// Order-preserving map of floats (no NaN) to unsigned keys, so min/max run as integer atomics.
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float kfloat(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

struct BNode {           // binary node: count > 0 leaf over [first, first + count), else children left, left+1
    float lo[3], hi[3];
    uint32_t left, first, count;
    uint32_t pad[3];
};
struct Seg {             // a node to process and its range of the permutation
    uint32_t node, begin, end, pad;
    float clo[3], chi[3];  // centroid bounds
};
struct Dec {             // split decision of a big or mid segment
    int axis, bin;       // axis < 0: median split
    uint32_t nl, pad;
    float lclo[3], lchi[3], rclo[3], rchi[3];  // children's centroid bounds
};
struct Item { uint32_t seg, begin, end, pad; };  // chunk of a big segment

__device__ __forceinline__ float area6(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Triangle box (folded zeros) and centroid of prim p.
__device__ __forceinline__ void prim_box(const float4* blo, const float4* bhi, uint32_t p, float lo[3], float hi[3], float c[3]) {
    const float4 a = blo[p], b = bhi[p];
    lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
    hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    for (int k = 0; k < 3; ++k) c[k] = (lo[k] + hi[k]) * 0.5f + 0.0f;
}

// Axis scales of a segment: valid iff extent > 0 and bins / extent is finite.
struct Axes { float clo[3], scale[3]; bool valid[3]; };
__device__ __forceinline__ Axes make_axes(const float clo[3], const float chi[3], uint32_t bins) {
    Axes a;
    for (int k = 0; k < 3; ++k) {
        a.clo[k] = clo[k];
        const float extent = chi[k] - clo[k];
        a.valid[k] = extent > 0.0f;
        a.scale[k] = a.valid[k] ? (float)bins / extent : 0.0f;
        if (!(a.scale[k] <= 3.40282347e38f)) a.valid[k] = false;
    }
    return a;
}
__device__ __forceinline__ int bin_of(float c, float clo, float scale, uint32_t bins) {
    int b = (int)((c - clo) * scale);
    return min(max(b, 0), (int)bins - 1);
}

__global__ __launch_bounds__(kThreads) void k_prim_boxes(const float* __restrict__ verts, uint32_t n, float4* __restrict__ blo,
                                                         float4* __restrict__ bhi) {
    const uint32_t t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= n) return;
    const float* v = verts + (size_t)t * 9;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = fminf(fminf(v[k], v[3 + k]), v[6 + k]) + 0.0f;
        hi[k] = fmaxf(fmaxf(v[k], v[3 + k]), v[6 + k]) + 0.0f;
    }
    blo[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    bhi[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
}

// Root centroid bounds: per-block key min/max, then one block folds them into the root segment.
__global__ __launch_bounds__(kThreads) void k_root_partial(const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                           uint32_t n, uint32_t* __restrict__ part, uint32_t* __restrict__ prim) {
    __shared__ uint32_t s[6];
    if (threadIdx.x < 6) s[threadIdx.x] = threadIdx.x < 3 ? kKeyPosInf : kKeyNegInf;
    __syncthreads();
    uint32_t mn[3] = {kKeyPosInf, kKeyPosInf, kKeyPosInf}, mx[3] = {kKeyNegInf, kKeyNegInf, kKeyNegInf};
    for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
        prim[t] = t;
        float lo[3], hi[3], c[3];
        prim_box(blo, bhi, t, lo, hi, c);
        for (int k = 0; k < 3; ++k) { mn[k] = min(mn[k], fkey(c[k])); mx[k] = max(mx[k], fkey(c[k])); }
    }
    for (int k = 0; k < 3; ++k) { atomicMin(&s[k], mn[k]); atomicMax(&s[3 + k], mx[k]); }
    __syncthreads();
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x];
}
__global__ __launch_bounds__(64) void k_root_final(const uint32_t* __restrict__ part, uint32_t nparts, uint32_t n,
                                                   Seg* __restrict__ seg) {
    if (threadIdx.x >= 6) return;
    uint32_t v = threadIdx.x < 3 ? kKeyPosInf : kKeyNegInf;
    for (uint32_t b = 0; b < nparts; ++b) {
        const uint32_t x = part[b * 6 + threadIdx.x];
        v = threadIdx.x < 3 ? min(v, x) : max(v, x);
    }
    if (threadIdx.x == 0) { seg->node = 0; seg->begin = 0; seg->end = n; seg->pad = 0; }
    if (threadIdx.x < 3) seg->clo[threadIdx.x] = kfloat(v); else seg->chi[threadIdx.x - 3] = kfloat(v);
}

// ---- exclusive scan of uint4 (component-wise), tiles of 1024 ------------------------------
constexpr uint32_t kTile = 1024;
__device__ __forceinline__ uint4 add4(uint4 a, uint4 b) { return make_uint4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }

// Block-wide exclusive scan of one uint4 per thread; returns the exclusive prefix, total in *tot.
__device__ uint4 block_excl4(uint4 v, uint4* s, uint4* tot) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
        uint4 o = make_uint4(0, 0, 0, 0);
        if ((int)threadIdx.x >= d) o = s[threadIdx.x - d];
        __syncthreads();
        s[threadIdx.x] = add4(s[threadIdx.x], o);
        __syncthreads();
    }
    const uint4 incl = s[threadIdx.x];
    if (tot) *tot = s[kThreads - 1];
    __syncthreads();
    return make_uint4(incl.x - v.x, incl.y - v.y, incl.z - v.z, incl.w - v.w);
}

__global__ __launch_bounds__(kThreads) void k_scan_tiles(const uint4* __restrict__ in, uint32_t n, uint4* __restrict__ sums) {
    __shared__ uint4 s[kThreads];
    uint4 acc = make_uint4(0, 0, 0, 0);
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * 4;
    for (int k = 0; k < 4; ++k)
        if (base + k < n) acc = add4(acc, in[base + k]);
    uint4 tot;
    block_excl4(acc, s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}
// One block: exclusive scan of the tile sums in place; grand total to *total.
__global__ __launch_bounds__(kThreads) void k_scan_sums(uint4* __restrict__ sums, uint32_t ntiles, uint4* __restrict__ total) {
    __shared__ uint4 s[kThreads];
    __shared__ uint4 carry;
    if (threadIdx.x == 0) carry = make_uint4(0, 0, 0, 0);
    __syncthreads();
    for (uint32_t base = 0; base < ntiles; base += kThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint4 v = i < ntiles ? sums[i] : make_uint4(0, 0, 0, 0);
        uint4 tot;
        const uint4 ex = block_excl4(v, s, &tot);
        const uint4 c = carry;
        if (i < ntiles) sums[i] = add4(ex, c);
        __syncthreads();
        if (threadIdx.x == 0) carry = add4(c, tot);
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}
__global__ __launch_bounds__(kThreads) void k_scan_apply(const uint4* __restrict__ in, uint32_t n, const uint4* __restrict__ sums,
                                                         uint4* __restrict__ out) {
    __shared__ uint4 s[kThreads];
    const uint32_t base = blockIdx.x * kTile + threadIdx.x * 4;
    uint4 v[4];
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : make_uint4(0, 0, 0, 0);
        acc = add4(acc, v[k]);
    }
    uint4 run = add4(block_excl4(acc, s, nullptr), sums[blockIdx.x]);
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run = add4(run, v[k]);
    }
}

// ---- level bookkeeping -----------------------------------------------------------------------
// v1 = (items, isBig, isMid, isSmall), v2 = (splits, slots, 0, 0)
__global__ __launch_bounds__(kThreads) void k_classify(const Seg* __restrict__ segs, uint32_t s, uint4* __restrict__ v1,
                                                       uint4* __restrict__ v2) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= s) return;
    const uint32_t n = segs[i].end - segs[i].begin;
    const uint32_t big = n > kBig, mid = !big && n > kSmall, small = n <= kSmall;
    v1[i] = make_uint4(big ? (n + kChunk - 1) / kChunk : 0u, big, mid, small);
    v2[i] = make_uint4(big | mid, (big | mid) ? 2u : (n >= 2 ? 2u * n - 2u : 0u), 0u, 0u);
}

__global__ __launch_bounds__(kThreads) void k_lists(const Seg* __restrict__ segs, uint32_t s, const uint4* __restrict__ v1,
                                                    const uint4* __restrict__ e1, uint32_t* __restrict__ big_list,
                                                    uint32_t* __restrict__ big_item0, Item* __restrict__ items,
                                                    uint32_t* __restrict__ mid_list, uint32_t* __restrict__ small_list) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= s) return;
    const uint4 f = v1[i], e = e1[i];
    const Seg g = segs[i];
    if (f.y) {
        big_list[e.y] = i;
        big_item0[e.y] = e.x;
        for (uint32_t k = 0; k < f.x; ++k) {
            Item it;
            it.seg = i;
            it.begin = g.begin + k * kChunk;
            it.end = min(g.end, it.begin + kChunk);
            it.pad = 0;
            items[e.x + k] = it;
        }
    }
    if (f.z) mid_list[e.z] = i;
    if (f.w) small_list[e.w] = i;
}

// ---- SAH over float bins (the host's sweeps, in its order) -------------------------------
// bins: [axis][bin] -> lo[3] hi[3] (floats), cnt[axis][bin]. Returns best cost, axis, bin.
__device__ void sah_serial(const float* __restrict__ bbox, const uint32_t* __restrict__ bcnt, const Axes& ax, uint32_t bins,
                           float* best_cost, int* best_axis, int* best_bin) {
    float best = __builtin_huge_valf();
    int ba = -1, bb = -1;
    float right_area[kMaxBins];
    uint32_t right_cnt[kMaxBins];
    for (int axis = 0; axis < 3; ++axis) {
        if (!ax.valid[axis]) continue;
        const float* bx = bbox + (size_t)axis * bins * 6;
        const uint32_t* bc = bcnt + (size_t)axis * bins;
        float lo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
        float hi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
        uint32_t c = 0;
        for (int b = (int)bins - 1; b > 0; --b) {
            for (int k = 0; k < 3; ++k) { lo[k] = fminf(lo[k], bx[b * 6 + k]); hi[k] = fmaxf(hi[k], bx[b * 6 + 3 + k]); }
            c += bc[b];
            right_area[b] = area6(lo, hi);
            right_cnt[b] = c;
        }
        for (int k = 0; k < 3; ++k) { lo[k] = __builtin_huge_valf(); hi[k] = -__builtin_huge_valf(); }
        c = 0;
        for (uint32_t b = 0; b + 1 < bins; ++b) {
            for (int k = 0; k < 3; ++k) { lo[k] = fminf(lo[k], bx[b * 6 + k]); hi[k] = fmaxf(hi[k], bx[b * 6 + 3 + k]); }
            c += bc[b];
            if (c == 0 || right_cnt[b + 1] == 0) continue;
            const float cost = area6(lo, hi) * (float)c + right_area[b + 1] * (float)right_cnt[b + 1];
            if (cost < best) { best = cost; ba = axis; bb = (int)b; }
        }
    }
    *best_cost = best; *best_axis = ba; *best_bin = bb;
}

// Decision from key bins (LDS, [axis][bin][13]) for a big or mid segment (n > max_leaf, so it
// always splits): node box into nodes[node], Dec for the children.
__device__ void decide_from_keys(const uint32_t* kb, const Axes& ax, uint32_t bins, uint32_t n, float* fbox, uint32_t* fcnt,
                                 Dec* dec) {
    for (uint32_t i = 0; i < 3 * bins; ++i) {
        for (int k = 0; k < 6; ++k) fbox[i * 6 + k] = kfloat(kb[i * kBinWords + k]);
        fcnt[i] = kb[i * kBinWords + 12];
    }
    float best; int ba, bb;
    sah_serial(fbox, fcnt, ax, bins, &best, &ba, &bb);
    Dec d;
    d.axis = ba; d.bin = bb; d.pad = 0;
    if (ba < 0) {
        d.nl = n / 2;
        for (int k = 0; k < 3; ++k) { d.lclo[k] = d.rclo[k] = ax.clo[k]; d.lchi[k] = d.rchi[k] = 0.0f; }
    } else {
        uint32_t nl = 0;
        uint32_t l[6] = {kKeyPosInf, kKeyPosInf, kKeyPosInf, kKeyNegInf, kKeyNegInf, kKeyNegInf};
        uint32_t r[6] = {kKeyPosInf, kKeyPosInf, kKeyPosInf, kKeyNegInf, kKeyNegInf, kKeyNegInf};
        for (uint32_t b = 0; b < bins; ++b) {
            const uint32_t* w = kb + ((size_t)ba * bins + b) * kBinWords;
            uint32_t* t = (int)b <= bb ? l : r;
            for (int k = 0; k < 3; ++k) { t[k] = min(t[k], w[6 + k]); t[3 + k] = max(t[3 + k], w[9 + k]); }
            if ((int)b <= bb) nl += w[12];
        }
        d.nl = nl;
        for (int k = 0; k < 3; ++k) {
            d.lclo[k] = kfloat(l[k]); d.lchi[k] = kfloat(l[3 + k]);
            d.rclo[k] = kfloat(r[k]); d.rchi[k] = kfloat(r[3 + k]);
        }
    }
    *dec = d;
}

__device__ __forceinline__ void init_key_bins(uint32_t* kb, uint32_t words) {
    for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) {
        const uint32_t r = w % kBinWords;
        kb[w] = r == 12 ? 0u : ((r < 3 || (r >= 6 && r < 9)) ? kKeyPosInf : kKeyNegInf);
    }
}
// Bin one triangle into key bins [axis][bin][13] and the 7-word total (lo hi count).
__device__ __forceinline__ void bin_prim(uint32_t* kb, uint32_t* tot, const Axes& ax, uint32_t bins, const float lo[3],
                                         const float hi[3], const float c[3]) {
    for (int k = 0; k < 3; ++k) { atomicMin(&tot[k], fkey(lo[k])); atomicMax(&tot[3 + k], fkey(hi[k])); }
    atomicAdd(&tot[6], 1u);
    for (int a = 0; a < 3; ++a) {
        if (!ax.valid[a]) continue;
        uint32_t* w = kb + ((size_t)a * bins + bin_of(c[a], ax.clo[a], ax.scale[a], bins)) * kBinWords;
        for (int k = 0; k < 3; ++k) {
            atomicMin(&w[k], fkey(lo[k])); atomicMax(&w[3 + k], fkey(hi[k]));
            atomicMin(&w[6 + k], fkey(c[k])); atomicMax(&w[9 + k], fkey(c[k]));
        }
        atomicAdd(&w[12], 1u);
    }
}
__device__ __forceinline__ void write_box(BNode* node, const uint32_t* tot) {
    for (int k = 0; k < 3; ++k) { node->lo[k] = kfloat(tot[k]); node->hi[k] = kfloat(tot[3 + k]); }
    node->count = 0;
    node->first = 0;
}

// ---- big segments -------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_big_bin(const Item* __restrict__ items, const Seg* __restrict__ segs,
                                                      const uint32_t* __restrict__ prim, const float4* __restrict__ blo,
                                                      const float4* __restrict__ bhi, uint32_t bins, uint32_t* __restrict__ part) {
    __shared__ uint32_t kb[3 * kMaxBins * kBinWords + 8];
    const uint32_t words = 3 * bins * kBinWords;
    uint32_t* tot = kb + words;
    init_key_bins(kb, words);
    if (threadIdx.x < 7) tot[threadIdx.x] = threadIdx.x < 3 ? kKeyPosInf : (threadIdx.x < 6 ? kKeyNegInf : 0u);
    __syncthreads();
    const Item it = items[blockIdx.x];
    const Seg g = segs[it.seg];
    const Axes ax = make_axes(g.clo, g.chi, bins);
    for (uint32_t j = it.begin + threadIdx.x; j < it.end; j += kThreads) {
        float lo[3], hi[3], c[3];
        prim_box(blo, bhi, prim[j], lo, hi, c);
        bin_prim(kb, tot, ax, bins, lo, hi, c);
    }
    __syncthreads();
    uint32_t* out = part + (size_t)blockIdx.x * (words + 7);
    for (uint32_t w = threadIdx.x; w < words + 7; w += kThreads) out[w] = kb[w];
}

// One workgroup per big segment: reduce its items' partial bins, then thread 0 decides.
__global__ __launch_bounds__(kThreads) void k_big_reduce(const uint32_t* __restrict__ big_list, const uint32_t* __restrict__ big_item0,
                                                         const Seg* __restrict__ segs, const uint32_t* __restrict__ part,
                                                         uint32_t bins, BNode* __restrict__ nodes, Dec* __restrict__ decs) {
    __shared__ uint32_t kb[3 * kMaxBins * kBinWords + 8];
    __shared__ float fbox[3 * kMaxBins * 6];
    __shared__ uint32_t fcnt[3 * kMaxBins];
    const uint32_t si = big_list[blockIdx.x];
    const Seg g = segs[si];
    const uint32_t n = g.end - g.begin;
    const uint32_t nitems = (n + kChunk - 1) / kChunk;
    const uint32_t first = big_item0[blockIdx.x];
    const uint32_t words = 3 * bins * kBinWords;
    for (uint32_t w = threadIdx.x; w < words + 7; w += kThreads) {
        const uint32_t r = w < words ? w % kBinWords : (w - words < 3 ? 0u : (w - words < 6 ? 3u : 12u));
        const bool is_min = r < 3 || (r >= 6 && r < 9);
        uint32_t v = r == 12 ? 0u : (is_min ? kKeyPosInf : kKeyNegInf);
        for (uint32_t k = 0; k < nitems; ++k) {
            const uint32_t x = part[(size_t)(first + k) * (words + 7) + w];
            v = r == 12 ? v + x : (is_min ? min(v, x) : max(v, x));
        }
        kb[w] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const Axes ax = make_axes(g.clo, g.chi, bins);
        write_box(&nodes[g.node], kb + words);
        decide_from_keys(kb, ax, bins, n, fbox, fcnt, &decs[si]);
    }
}

__device__ __forceinline__ bool goes_left(const Dec& d, const Axes& ax, uint32_t bins, const float c[3], uint32_t pos_in_seg) {
    if (d.axis < 0) return pos_in_seg < d.nl;
    return bin_of(c[d.axis], ax.clo[d.axis], ax.scale[d.axis], bins) <= d.bin;
}

__global__ __launch_bounds__(kThreads) void k_big_count(const Item* __restrict__ items, const Seg* __restrict__ segs,
                                                        const Dec* __restrict__ decs, const uint32_t* __restrict__ prim,
                                                        const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                        uint32_t bins, uint32_t* __restrict__ item_left) {
    __shared__ uint32_t cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const Item it = items[blockIdx.x];
    const Seg g = segs[it.seg];
    const Dec d = decs[it.seg];
    const Axes ax = make_axes(g.clo, g.chi, bins);
    uint32_t mine = 0;
    for (uint32_t j = it.begin + threadIdx.x; j < it.end; j += kThreads) {
        float lo[3], hi[3], c[3];
        prim_box(blo, bhi, prim[j], lo, hi, c);
        mine += goes_left(d, ax, bins, c, j - g.begin) ? 1u : 0u;
    }
    atomicAdd(&cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0) item_left[blockIdx.x] = cnt;
}

// Exclusive scan of each big segment's item left-counts (in item order).
__global__ __launch_bounds__(kThreads) void k_big_offsets(const uint32_t* __restrict__ big_list, const uint32_t* __restrict__ big_item0,
                                                          const Seg* __restrict__ segs, const uint32_t* __restrict__ item_left,
                                                          uint32_t* __restrict__ item_loff) {
    __shared__ uint4 s[kThreads];
    __shared__ uint32_t carry;
    const Seg g = segs[big_list[blockIdx.x]];
    const uint32_t nitems = (g.end - g.begin + kChunk - 1) / kChunk;
    const uint32_t first = big_item0[blockIdx.x];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nitems; base += kThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nitems ? item_left[first + i] : 0u;
        uint4 tot;
        const uint4 ex = block_excl4(make_uint4(v, 0, 0, 0), s, &tot);
        const uint32_t c = carry;
        if (i < nitems) item_loff[first + i] = ex.x + c;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot.x;
        __syncthreads();
    }
}

// Stable scatter of each item's elements into prim_tmp.
__global__ __launch_bounds__(kThreads) void k_big_scatter(const Item* __restrict__ items, const Seg* __restrict__ segs,
                                                          const Dec* __restrict__ decs, const uint32_t* __restrict__ prim,
                                                          const float4* __restrict__ blo, const float4* __restrict__ bhi,
                                                          uint32_t bins, const uint32_t* __restrict__ item_loff,
                                                          uint32_t* __restrict__ prim_tmp) {
    __shared__ uint4 s[kThreads];
    const Item it = items[blockIdx.x];
    const Seg g = segs[it.seg];
    const Dec d = decs[it.seg];
    const Axes ax = make_axes(g.clo, g.chi, bins);
    const uint32_t loff = item_loff[blockIdx.x];
    const uint32_t roff = (it.begin - g.begin) - loff;
    uint32_t run_l = 0, run_r = 0;
    for (uint32_t base = it.begin; base < it.end; base += kThreads) {
        const uint32_t j = base + threadIdx.x;
        const bool in = j < it.end;
        uint32_t p = 0;
        bool left = false;
        if (in) {
            p = prim[j];
            float lo[3], hi[3], c[3];
            prim_box(blo, bhi, p, lo, hi, c);
            left = goes_left(d, ax, bins, c, j - g.begin);
        }
        uint4 tot;
        const uint4 ex = block_excl4(make_uint4(in && left ? 1u : 0u, in && !left ? 1u : 0u, 0, 0), s, &tot);
        if (in) {
            const uint32_t dst = left ? g.begin + loff + run_l + ex.x : g.begin + d.nl + roff + run_r + ex.y;
            if (dst < g.end) prim_tmp[dst] = p;  // always true: d.nl counts exactly the left elements
        }
        run_l += tot.x;
        run_r += tot.y;
    }
}

__global__ __launch_bounds__(kThreads) void k_big_copy(const Item* __restrict__ items, const uint32_t* __restrict__ prim_tmp,
                                                       uint32_t* __restrict__ prim) {
    const Item it = items[blockIdx.x];
    for (uint32_t j = it.begin + threadIdx.x; j < it.end; j += kThreads) prim[j] = prim_tmp[j];
}

// ---- mid segments: one workgroup each ---------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_mid(const uint32_t* __restrict__ mid_list, const Seg* __restrict__ segs,
                                                  uint32_t* __restrict__ prim, const float4* __restrict__ blo,
                                                  const float4* __restrict__ bhi, uint32_t bins, BNode* __restrict__ nodes,
                                                  Dec* __restrict__ decs) {
    __shared__ uint32_t ids[kBig];
    __shared__ uint32_t kb[3 * kMaxBins * kBinWords + 8];
    __shared__ float fbox[3 * kMaxBins * 6];
    __shared__ uint32_t fcnt[3 * kMaxBins];
    __shared__ uint4 s[kThreads];
    __shared__ Dec sd;
    const uint32_t si = mid_list[blockIdx.x];
    const Seg g = segs[si];
    const uint32_t n = g.end - g.begin;
    const uint32_t words = 3 * bins * kBinWords;
    uint32_t* tot = kb + words;
    init_key_bins(kb, words);
    if (threadIdx.x < 7) tot[threadIdx.x] = threadIdx.x < 3 ? kKeyPosInf : (threadIdx.x < 6 ? kKeyNegInf : 0u);
    for (uint32_t j = threadIdx.x; j < n; j += kThreads) ids[j] = prim[g.begin + j];
    __syncthreads();
    const Axes ax = make_axes(g.clo, g.chi, bins);
    for (uint32_t j = threadIdx.x; j < n; j += kThreads) {
        float lo[3], hi[3], c[3];
        prim_box(blo, bhi, ids[j], lo, hi, c);
        bin_prim(kb, tot, ax, bins, lo, hi, c);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        write_box(&nodes[g.node], tot);
        decide_from_keys(kb, ax, bins, n, fbox, fcnt, &sd);
        decs[si] = sd;
    }
    __syncthreads();
    const Dec d = sd;
    uint32_t run_l = 0, run_r = 0;
    for (uint32_t base = 0; base < n; base += kThreads) {
        const uint32_t j = base + threadIdx.x;
        const bool in = j < n;
        bool left = false;
        if (in) {
            float lo[3], hi[3], c[3];
            prim_box(blo, bhi, ids[j], lo, hi, c);
            left = goes_left(d, ax, bins, c, j);
        }
        uint4 t4;
        const uint4 ex = block_excl4(make_uint4(in && left ? 1u : 0u, in && !left ? 1u : 0u, 0, 0), s, &t4);
        const uint32_t dst = left ? run_l + ex.x : d.nl + run_r + ex.y;
        if (in && dst < n) prim[g.begin + dst] = ids[j];  // dst < n always: d.nl counts exactly the left elements
        run_l += t4.x;
        run_r += t4.y;
    }
}

// Children of big and mid segments: link, and append to the next level's list.
__global__ __launch_bounds__(kThreads) void k_next(const Seg* __restrict__ segs, uint32_t s, const uint4* __restrict__ v2,
                                                   const uint4* __restrict__ e2, const Dec* __restrict__ decs, uint32_t slot_base,
                                                   BNode* __restrict__ nodes, Seg* __restrict__ next) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= s || !v2[i].x) return;
    const Seg g = segs[i];
    const Dec d = decs[i];
    const uint32_t left = slot_base + e2[i].y;
    nodes[g.node].left = left;
    Seg a, b;
    a.node = left; a.begin = g.begin; a.end = g.begin + d.nl; a.pad = 0;
    b.node = left + 1; b.begin = g.begin + d.nl; b.end = g.end; b.pad = 0;
    for (int k = 0; k < 3; ++k) {
        if (d.axis < 0) {  // all centroids coincide: the children keep the parent's (point) bounds
            a.clo[k] = b.clo[k] = g.clo[k]; a.chi[k] = b.chi[k] = g.chi[k];
        } else {
            a.clo[k] = d.lclo[k]; a.chi[k] = d.lchi[k]; b.clo[k] = d.rclo[k]; b.chi[k] = d.rchi[k];
        }
    }
    next[2 * e2[i].x] = a;
    next[2 * e2[i].x + 1] = b;
}

// ---- small segments: one wave finishes the subtree ---------------------------------------
constexpr int kWaves = kThreads / 64;
struct WaveLds {
    uint32_t kb[3 * kMaxBins * 7];  // lo[3] hi[3] count
    uint32_t stack[64 * 3];
    uint32_t perm[64 * 7];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ float wmin(float v) {
    for (int d = 32; d > 0; d >>= 1) v = fminf(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ float wmax(float v) {
    for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

__global__ __launch_bounds__(kThreads) void k_small(const uint32_t* __restrict__ small_list, uint32_t nsmall,
                                                    const Seg* __restrict__ segs, const uint4* __restrict__ e2, uint32_t slot_base,
                                                    uint32_t* __restrict__ prim, const float4* __restrict__ blo,
                                                    const float4* __restrict__ bhi, uint32_t bins, uint32_t max_leaf,
                                                    BNode* __restrict__ nodes) {
    __shared__ WaveLds lds_all[kWaves];
    const uint32_t w = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nsmall) return;  // uniform per wave
    WaveLds& L = lds_all[threadIdx.x >> 6];
    const uint32_t si = small_list[w];
    const Seg g = segs[si];
    const uint32_t n = g.end - g.begin;
    uint32_t slot = slot_base + e2[si].y;
    const float inf = __builtin_huge_valf();
    // this lane's triangle
    uint32_t p = 0;
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf}, c[3] = {0.0f, 0.0f, 0.0f};
    if ((uint32_t)lane < n) {
        p = prim[g.begin + lane];
        prim_box(blo, bhi, p, lo, hi, c);
    }
    if (lane == 0) { L.stack[0] = g.node; L.stack[1] = 0; L.stack[2] = n; }
    int sp = 1;
    wave_sync();
    const uint32_t nb = 3 * bins;
    // at most 2n - 1 nodes, each popped once; the bound only guarantees termination
    for (uint32_t iter = 0; sp > 0 && iter < 2 * kSmall; ++iter) {
        --sp;
        const uint32_t node = L.stack[sp * 3], lb = L.stack[sp * 3 + 1], le = L.stack[sp * 3 + 2];
        wave_sync();
        if (le <= lb) continue;  // never pushed: both sides of a split are non-empty
        const uint32_t m = le - lb;
        const bool act = (uint32_t)lane >= lb && (uint32_t)lane < le;
        float blo3[3], bhi3[3], clo3[3], chi3[3];
        for (int k = 0; k < 3; ++k) {
            blo3[k] = wmin(act ? lo[k] : inf);
            bhi3[k] = wmax(act ? hi[k] : -inf);
            clo3[k] = wmin(act ? c[k] : inf);
            chi3[k] = wmax(act ? c[k] : -inf);
        }
        if (lane == 0) {
            for (int k = 0; k < 3; ++k) { nodes[node].lo[k] = blo3[k]; nodes[node].hi[k] = bhi3[k]; }
        }
        if (m == 1) {
            if (lane == 0) { nodes[node].left = 0; nodes[node].first = g.begin + lb; nodes[node].count = 1; }
            continue;
        }
        const Axes ax = make_axes(clo3, chi3, bins);
        // bins
        for (uint32_t q = lane; q < nb * 7; q += 64) {
            const uint32_t r = q % 7;
            L.kb[q] = r == 6 ? 0u : (r < 3 ? kKeyPosInf : kKeyNegInf);
        }
        wave_sync();
        if (act) {
            for (int a = 0; a < 3; ++a) {
                if (!ax.valid[a]) continue;
                uint32_t* wb = L.kb + ((uint32_t)a * bins + bin_of(c[a], ax.clo[a], ax.scale[a], bins)) * 7;
                for (int k = 0; k < 3; ++k) { atomicMin(&wb[k], fkey(lo[k])); atomicMax(&wb[3 + k], fkey(hi[k])); }
                atomicAdd(&wb[6], 1u);
            }
        }
        wave_sync();
        // SAH: lanes = (axis, bin) in groups of `per` axes; prefix (left) and suffix (right) unions.
        float best = inf;
        int best_l = 0x7FFFFFFF;
        const uint32_t per = bins <= 21 ? 3u : 1u;
        for (uint32_t a0 = 0; a0 < 3; a0 += per) {
            const uint32_t a = a0 + (uint32_t)lane / bins;
            const uint32_t b = (uint32_t)lane % bins;
            const bool mine = (uint32_t)lane < per * bins && a < 3;
            float l6[6], r6[6];
            uint32_t lc = 0, rc = 0;
            if (mine) {
                const uint32_t* wb = L.kb + (a * bins + b) * 7;
                for (int k = 0; k < 6; ++k) l6[k] = r6[k] = kfloat(wb[k]);
                lc = rc = wb[6];
            } else {
                for (int k = 0; k < 3; ++k) { l6[k] = r6[k] = inf; l6[3 + k] = r6[3 + k] = -inf; }
            }
            for (uint32_t d = 1; d < bins; d <<= 1) {
                float ul[6], dr[6];
                for (int k = 0; k < 6; ++k) { ul[k] = __shfl_up(l6[k], d, 64); dr[k] = __shfl_down(r6[k], d, 64); }
                const uint32_t ulc = __shfl_up(lc, d, 64), drc = __shfl_down(rc, d, 64);
                if (b >= d) {
                    for (int k = 0; k < 3; ++k) { l6[k] = fminf(l6[k], ul[k]); l6[3 + k] = fmaxf(l6[3 + k], ul[3 + k]); }
                    lc += ulc;
                }
                if (b + d < bins) {
                    for (int k = 0; k < 3; ++k) { r6[k] = fminf(r6[k], dr[k]); r6[3 + k] = fmaxf(r6[3 + k], dr[3 + k]); }
                    rc += drc;
                }
            }
            // right side of split b is the suffix starting at b + 1
            float nr6[6];
            for (int k = 0; k < 6; ++k) nr6[k] = __shfl_down(r6[k], 1, 64);
            const uint32_t nrc = __shfl_down(rc, 1, 64);
            float cost = inf;
            if (mine && ax.valid[a] && b + 1 < bins && lc != 0 && nrc != 0) {
                const float la = area6(l6, l6 + 3), ra = area6(nr6, nr6 + 3);
                cost = la * (float)lc + ra * (float)nrc;
            }
            int key = (mine && cost < inf) ? (int)(a * bins + b) : 0x7FFFFFFF;
            if (!(cost < inf)) cost = inf;
            // argmin over the wave: smaller cost, then smaller (axis, bin)
            for (int d = 32; d > 0; d >>= 1) {
                const float oc = __shfl_xor(cost, d, 64);
                const int ok = __shfl_xor(key, d, 64);
                if (oc < cost || (oc == cost && ok < key)) { cost = oc; key = ok; }
            }
            if (cost < best) { best = cost; best_l = key; }
        }
        const int ba = best_l == 0x7FFFFFFF ? -1 : best_l / (int)bins;
        const int bb = best_l == 0x7FFFFFFF ? -1 : best_l % (int)bins;
        const float area = area6(blo3, bhi3);
        if (m <= max_leaf && (ba < 0 || (float)m * area <= area + best)) {
            if (lane == 0) { nodes[node].left = 0; nodes[node].first = g.begin + lb; nodes[node].count = m; }
            continue;
        }
        // stable partition of lanes [lb, le)
        bool left = false;
        if (act) left = ba < 0 ? ((uint32_t)lane - lb) < m / 2 : bin_of(c[ba], ax.clo[ba], ax.scale[ba], bins) <= bb;
        const uint64_t lmask = __ballot(act && left), rmask = __ballot(act && !left);
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const uint32_t nl = (uint32_t)__popcll(lmask);
        if (act) {
            const uint32_t dst = left ? lb + (uint32_t)__popcll(lmask & below) : lb + nl + (uint32_t)__popcll(rmask & below);
            uint32_t* q = L.perm + dst * 7;
            q[0] = p;
            for (int k = 0; k < 3; ++k) { q[1 + k] = __float_as_uint(lo[k]); q[4 + k] = __float_as_uint(hi[k]); }
        }
        wave_sync();
        if (act) {
            const uint32_t* q = L.perm + (uint32_t)lane * 7;
            p = q[0];
            for (int k = 0; k < 3; ++k) { lo[k] = __uint_as_float(q[1 + k]); hi[k] = __uint_as_float(q[4 + k]); }
            for (int k = 0; k < 3; ++k) c[k] = (lo[k] + hi[k]) * 0.5f + 0.0f;
        }
        const uint32_t lch = slot;
        slot += 2;
        if (lane == 0) {
            nodes[node].left = lch; nodes[node].first = 0; nodes[node].count = 0;
            L.stack[sp * 3] = lch + 1; L.stack[sp * 3 + 1] = lb + nl; L.stack[sp * 3 + 2] = le;
            L.stack[sp * 3 + 3] = lch; L.stack[sp * 3 + 4] = lb; L.stack[sp * 3 + 5] = lb + nl;
        }
        sp += 2;
        wave_sync();
    }
    if ((uint32_t)lane < n) prim[g.begin + lane] = p;
}

// ---- collapse to BVH4 -------------------------------------------------------------------------
__device__ __forceinline__ float node_area(const BNode& x) { return area6(x.lo, x.hi); }

__global__ __launch_bounds__(kThreads) void k_col_count(const uint32_t* __restrict__ level, uint32_t m,
                                                        const BNode* __restrict__ nodes, uint4* __restrict__ cand_out,
                                                        uint32_t* __restrict__ ncand, uint4* __restrict__ v) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const uint32_t bn = level[i];
    uint32_t cand[4] = {0, 0, 0, 0};
    int nc = 0;
    if (nodes[bn].count) {
        cand[nc++] = bn;
    } else {
        cand[nc++] = nodes[bn].left;
        cand[nc++] = nodes[bn].left + 1;
        while (nc < 4) {
            int pick = -1;
            float best_area = -1.0f;
            for (int c = 0; c < nc; ++c) {
                const BNode& x = nodes[cand[c]];
                if (x.count) continue;
                const float a = node_area(x);
                if (a > best_area) { best_area = a; pick = c; }
            }
            if (pick < 0) break;
            const uint32_t opened = cand[pick];
            cand[pick] = nodes[opened].left;
            cand[nc++] = nodes[opened].left + 1;
        }
    }
    uint32_t ni = 0, nt = 0, nlf = 0;
    for (int c = 0; c < nc; ++c) {
        const uint32_t cnt = nodes[cand[c]].count;
        if (cnt) { nt += cnt; ++nlf; } else { ++ni; }
    }
    cand_out[i] = make_uint4(cand[0], cand[1], cand[2], cand[3]);
    ncand[i] = (uint32_t)nc;
    v[i] = make_uint4(ni, nt, nlf, 0);
}

__global__ __launch_bounds__(kThreads) void k_col_emit(const uint32_t* __restrict__ level, uint32_t m, uint32_t level_base,
                                                       uint32_t next_base, uint32_t tri_base, const BNode* __restrict__ nodes,
                                                       const uint4* __restrict__ cand_in, const uint32_t* __restrict__ ncand,
                                                       const uint4* __restrict__ e, const uint32_t* __restrict__ prim,
                                                       const float* __restrict__ verts, WNode* __restrict__ wnodes,
                                                       float4* __restrict__ tris, uint32_t* __restrict__ next_level) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= m) return;
    const uint4 cv = cand_in[i];
    const uint32_t cand[4] = {cv.x, cv.y, cv.z, cv.w};
    const int nc = (int)ncand[i];
    uint32_t run_i = e[i].x, run_t = tri_base + e[i].y;
    WNode wn;
    const float inf = __builtin_huge_valf();
    for (int c = 0; c < 4; ++c) {
        wn.lox[c] = wn.loy[c] = wn.loz[c] = inf;
        wn.hix[c] = wn.hiy[c] = wn.hiz[c] = -inf;
        wn.child[c] = kEmptyChild;
        wn.meta[c] = 0;
    }
    for (int c = 0; c < nc; ++c) {
        const BNode x = nodes[cand[c]];
        wn.lox[c] = x.lo[0]; wn.hix[c] = x.hi[0];
        wn.loy[c] = x.lo[1]; wn.hiy[c] = x.hi[1];
        wn.loz[c] = x.lo[2]; wn.hiz[c] = x.hi[2];
        if (x.count) {
            wn.child[c] = kLeafBit | (run_t << 3) | (x.count - 1);
            for (uint32_t k = 0; k < x.count; ++k) {
                const uint32_t id = prim[x.first + k];
                const float* v = verts + (size_t)id * 9;
                float4* r = tris + (size_t)(run_t + k) * 3;
                r[0] = make_float4(v[0], v[1], v[2], __uint_as_float(id));
                r[1] = make_float4(v[3], v[4], v[5], 0.0f);
                r[2] = make_float4(v[6], v[7], v[8], 0.0f);
            }
            run_t += x.count;
        } else {
            wn.child[c] = next_base + run_i;
            next_level[run_i] = cand[c];
            ++run_i;
        }
    }
    wnodes[level_base + i] = wn;
}

}  // namespace bvhdev

namespace {

template <class T>
hipError_t dalloc(T** p, size_t count, std::vector<void*>& owned) {
    hipError_t e = hipMalloc((void**)p, (count ? count : 1) * sizeof(T));
    if (e == hipSuccess) owned.push_back(*p);
    return e;
}

struct Scanner {
    uint4* sums = nullptr;
    uint4* total = nullptr;
    uint32_t cap_tiles = 0;
};

hipError_t scan4(hipStream_t st, Scanner& sc, const uint4* in, uint32_t n, uint4* out) {
    using namespace bvhdev;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    if (tiles > sc.cap_tiles) return hipErrorInvalidValue;
    if (n == 0) return hipMemsetAsync(sc.total, 0, sizeof(uint4), st);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kThreads), 0, st, in, n, sc.sums);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, st, sc.sums, tiles, sc.total);
    hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(kThreads), 0, st, in, n, sc.sums, out);
    return hipGetLastError();
}

inline uint32_t blocks(uint32_t n) { return (n + bvhdev::kThreads - 1) / bvhdev::kThreads; }

}  // namespace

// Builds the BVH4 of n_tris triangles already on the device (d_verts, 9 floats each) with the
// host builder's rules; writes exact-size node and triangle arrays (caller frees them).
hipError_t build_bvh4_device(mrt_ctx* ctx, const float* d_verts, uint32_t ntri, uint32_t max_leaf, uint32_t bins,
                             DeviceBVH& out) {
    using namespace bvhdev;
    auto t0 = std::chrono::steady_clock::now();
    out = DeviceBVH();
    hipStream_t st = ctx->stream;
    std::vector<void*> owned;
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    auto cleanup = [&]() { for (void* q : owned) hipFree(q); owned.clear(); };
    if (ntri == 0) return hipErrorInvalidValue;

    const uint32_t seg_cap = ntri / 32 + 4;              // next levels hold children of segments > 64
    const uint32_t item_cap = 2 * (ntri / kChunk) + 4;
    const uint32_t words = 3 * bins * kBinWords + 7;
    const uint32_t tiles_cap = (std::max(seg_cap, ntri) + kTile - 1) / kTile + 1;
    float4 *blo, *bhi;
    uint32_t *prim, *prim_tmp, *part, *root_part, *big_list, *big_item0, *mid_list, *small_list, *item_left, *item_loff;
    BNode* nodes;
    Seg *segs[2];
    Dec* decs;
    Item* items;
    uint4 *v1, *v2, *e1, *e2;
    Scanner sc;
    sc.cap_tiles = tiles_cap;
    chk(dalloc(&blo, ntri, owned)); chk(dalloc(&bhi, ntri, owned));
    chk(dalloc(&prim, ntri, owned)); chk(dalloc(&prim_tmp, ntri, owned));
    chk(dalloc(&nodes, (size_t)2 * ntri, owned));
    chk(dalloc(&segs[0], seg_cap, owned)); chk(dalloc(&segs[1], seg_cap, owned));
    chk(dalloc(&decs, seg_cap, owned));
    chk(dalloc(&items, item_cap, owned));
    chk(dalloc(&part, (size_t)item_cap * words, owned));
    chk(dalloc(&root_part, 1024 * 6, owned));
    chk(dalloc(&big_list, seg_cap, owned)); chk(dalloc(&big_item0, seg_cap, owned));
    chk(dalloc(&mid_list, seg_cap, owned)); chk(dalloc(&small_list, seg_cap, owned));
    chk(dalloc(&item_left, item_cap, owned)); chk(dalloc(&item_loff, item_cap, owned));
    chk(dalloc(&v1, std::max(seg_cap, ntri), owned)); chk(dalloc(&v2, seg_cap, owned));
    chk(dalloc(&e1, std::max(seg_cap, ntri), owned)); chk(dalloc(&e2, seg_cap, owned));
    chk(dalloc(&sc.sums, tiles_cap, owned)); chk(dalloc(&sc.total, 2, owned));
    if (e != hipSuccess) { cleanup(); return e; }

    // boxes, identity permutation, root centroid bounds
    hipLaunchKernelGGL(k_prim_boxes, dim3(blocks(ntri)), dim3(kThreads), 0, st, d_verts, ntri, blo, bhi);
    const uint32_t rparts = std::min(1024u, blocks(ntri));
    hipLaunchKernelGGL(k_root_partial, dim3(rparts), dim3(kThreads), 0, st, blo, bhi, ntri, root_part, prim);
    hipLaunchKernelGGL(k_root_final, dim3(1), dim3(64), 0, st, root_part, rparts, ntri, segs[0]);
    chk(hipGetLastError());

    // binary levels
    uint32_t s = 1, cur = 0, next_node = 1, levels = 0;
    uint4 tot[2];
    while (s > 0 && e == hipSuccess) {
        // every split shrinks its segment, so the depth is below ntri; this only bounds the loop
        if (++levels > ntri + 1) { e = hipErrorInvalidValue; break; }
        Seg* sg = segs[cur];
        hipLaunchKernelGGL(k_classify, dim3(blocks(s)), dim3(kThreads), 0, st, sg, s, v1, v2);
        chk(scan4(st, sc, v1, s, e1));
        chk(hipMemcpyAsync(&tot[0], sc.total, sizeof(uint4), hipMemcpyDeviceToHost, st));
        chk(scan4(st, sc, v2, s, e2));
        chk(hipMemcpyAsync(&tot[1], sc.total, sizeof(uint4), hipMemcpyDeviceToHost, st));
        chk(hipStreamSynchronize(st));
        if (e != hipSuccess) break;
        const uint32_t n_items = tot[0].x, n_big = tot[0].y, n_mid = tot[0].z, n_small = tot[0].w;
        const uint32_t n_split = tot[1].x, n_slots = tot[1].y;
        if (n_items > item_cap || 2 * n_split > seg_cap || next_node + n_slots > 2 * ntri) { e = hipErrorInvalidValue; break; }
        hipLaunchKernelGGL(k_lists, dim3(blocks(s)), dim3(kThreads), 0, st, sg, s, v1, e1, big_list, big_item0, items,
                           mid_list, small_list);
        const uint32_t slot_base = next_node;
        next_node += n_slots;
        if (n_big) {
            hipLaunchKernelGGL(k_big_bin, dim3(n_items), dim3(kThreads), 0, st, items, sg, prim, blo, bhi, bins, part);
            hipLaunchKernelGGL(k_big_reduce, dim3(n_big), dim3(kThreads), 0, st, big_list, big_item0, sg, part, bins, nodes, decs);
            hipLaunchKernelGGL(k_big_count, dim3(n_items), dim3(kThreads), 0, st, items, sg, decs, prim, blo, bhi, bins, item_left);
            hipLaunchKernelGGL(k_big_offsets, dim3(n_big), dim3(kThreads), 0, st, big_list, big_item0, sg, item_left, item_loff);
            hipLaunchKernelGGL(k_big_scatter, dim3(n_items), dim3(kThreads), 0, st, items, sg, decs, prim, blo, bhi, bins,
                               item_loff, prim_tmp);
            hipLaunchKernelGGL(k_big_copy, dim3(n_items), dim3(kThreads), 0, st, items, prim_tmp, prim);
        }
        if (n_mid)
            hipLaunchKernelGGL(k_mid, dim3(n_mid), dim3(kThreads), 0, st, mid_list, sg, prim, blo, bhi, bins, nodes, decs);
        if (n_small)
            hipLaunchKernelGGL(k_small, dim3((n_small + kWaves - 1) / kWaves), dim3(kThreads), 0, st, small_list, n_small, sg,
                               e2, slot_base, prim, blo, bhi, bins, max_leaf, nodes);
        if (n_split)
            hipLaunchKernelGGL(k_next, dim3(blocks(s)), dim3(kThreads), 0, st, sg, s, v2, e2, decs, slot_base, nodes,
                               segs[cur ^ 1]);
        chk(hipGetLastError());
        s = 2 * n_split;
        cur ^= 1;
    }
    if (e != hipSuccess) { cleanup(); return e; }
    out.binary_nodes = next_node;

    // collapse, level by level (BFS order)
    uint32_t* lvl[2];
    uint4* cand;
    uint32_t* ncand;
    WNode* wtmp;
    float4* tris;
    const uint32_t wide_cap = next_node / 2 + 2;  // wide nodes <= internal binary nodes (+ a leaf root)
    chk(dalloc(&lvl[0], wide_cap, owned)); chk(dalloc(&lvl[1], wide_cap, owned));
    chk(dalloc(&cand, wide_cap, owned)); chk(dalloc(&ncand, wide_cap, owned));
    chk(dalloc(&wtmp, wide_cap, owned));
    if (e == hipSuccess) {
        e = hipMalloc((void**)&tris, (size_t)ntri * 3 * sizeof(float4));
        if (e != hipSuccess) tris = nullptr;
    }
    if (e != hipSuccess) { cleanup(); return e; }
    const uint32_t root = 0;
    chk(hipMemcpyAsync(lvl[0], &root, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    uint32_t m = 1, lc = 0, level_base = 0, tri_base = 0, depth = 0, leaves = 0;
    while (m > 0 && e == hipSuccess) {
        if (level_base + m > wide_cap || depth > next_node) { e = hipErrorInvalidValue; break; }
        hipLaunchKernelGGL(k_col_count, dim3(blocks(m)), dim3(kThreads), 0, st, lvl[lc], m, nodes, cand, ncand, v1);
        chk(scan4(st, sc, v1, m, e1));
        chk(hipMemcpyAsync(&tot[0], sc.total, sizeof(uint4), hipMemcpyDeviceToHost, st));
        hipLaunchKernelGGL(k_col_emit, dim3(blocks(m)), dim3(kThreads), 0, st, lvl[lc], m, level_base, level_base + m, tri_base,
                           nodes, cand, ncand, e1, prim, d_verts, wtmp, tris, lvl[lc ^ 1]);
        chk(hipGetLastError());
        chk(hipStreamSynchronize(st));
        if (e != hipSuccess) break;
        ++depth;
        leaves += tot[0].z;
        tri_base += tot[0].y;
        level_base += m;
        m = tot[0].x;
        lc ^= 1;
    }
    if (e == hipSuccess && tri_base != ntri) e = hipErrorInvalidValue;
    WNode* wn = nullptr;
    if (e == hipSuccess) chk(hipMalloc((void**)&wn, (size_t)level_base * sizeof(WNode)));
    if (e == hipSuccess) chk(hipMemcpyAsync(wn, wtmp, (size_t)level_base * sizeof(WNode), hipMemcpyDeviceToDevice, st));
    if (e == hipSuccess) chk(hipStreamSynchronize(st));
    cleanup();
    if (e != hipSuccess) {
        if (wn) hipFree(wn);
        if (tris) hipFree(tris);
        return e;
    }
    out.d_nodes = wn;
    out.d_tris = tris;
    out.n_nodes = level_base;
    out.depth = depth;
    out.leaves = leaves;
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return hipSuccess;
}
