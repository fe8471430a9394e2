// trace.hip — closest-hit and any-hit BVH4 traversal for gfx950.
//
// Replaces the brute-force closest-hit loops of the reference (tracePrimaryRay,
// src/gpu/primaryRay.comp:23-36; the inner loop of traceSecondaryRays,
// src/gpu/secondaryRays.comp:77-83) by a traversal of the 128-byte BVH4 built on the host.
//
// Execution model (CDNA4, wave64): persistent waves, one ray per lane. Wave w first takes rays
// [64 w, 64 w + 64) with no atomic; further rays come from a global queue counter that a
// small kernel clears before each launch. (Device-scope atomics on one address serialize
// across the XCDs: a grid of 5k waves spending two atomics each cost ~250 us per launch, even
// with no rays to trace.) The v3x8 variants split the rays into 8 contiguous partitions, one
// per XCD (workgroup b runs on XCD b % 8), each with its own counter, so that each XCD's L2
// serves one spatially coherent range. Idle lanes are refilled as soon as REFILL lanes of the
// wave are idle; the traversal is while-while with postponed leaves (Aila & Laine 2009), in
// branch-light code (unconditional node loads, non-short-circuit slab tests, one 8-byte LDS word
// per stack entry). The traversal stack lives in LDS, one column per lane (entry k of lane l at
// [k * BLOCK + l], conflict-free), spilling to a global per-thread area past LDS_DEPTH
// entries. Children are visited nearest first; an entry is skipped when popped if its entry
// distance is beyond the closest hit found so far.
//
// Box test (rtrav::slab_dir): the planes ordered by the ray's direction signs, each plane
// distance one FMA against o * inv moved outward by 2^-21 |o * inv|, and the interval widened by
// a relative 2^-15, so that a box is never culled for a triangle the watertight test accepts
// (the triangle results then do not depend on the tree; the tree only changes the work done).
#include "internal.h"
#include "traverse.h"

namespace {

using rt::v3;
using rtrav::kSlack;

constexpr int kModeClosest = 0, kModeAny = 1, kModeMixed = 2;  // k_trace3's MODE
// A wave leaves its node loop (for the leaf loop, then a refill of its finished lanes) once
// fewer than this many lanes are still descending: the few lanes of a long subtree otherwise hold
// the rest of the wave idle until they finish. Measured on incoherent bounce-1 rays: 12 to 24
// are equal and 7-8% faster than never leaving (on the 1M- and the 16.4M-triangle scenes); 40
// and more leave too often.
constexpr uint32_t kMinDescending = 16;

// Number of set bits of m in the lanes below this one (v_mbcnt).
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

struct TraceArgs {
    const WNode* __restrict__ nodes;
    const float4* __restrict__ tris;
    const mrt_ray* __restrict__ rays;
    mrt_hit* __restrict__ hits;
    uint32_t* __restrict__ occluded;
    const mrt_ray* __restrict__ rays2;      // mixed launches: closest-hit rays (rays: any-hit rays)
    const uint32_t* __restrict__ perm;      // optional processing order (ray index per item)
    const uint32_t* __restrict__ count_dev; // optional device-side ray count (<= n)
    const uint32_t* __restrict__ ray_root;  // optional per-ray root node (batched scenes)
    uint32_t n;
    uint32_t* queue;        // items taken beyond the static first batches, one counter per partition
                            // at queue[p * kQueueStride] (zeroed before a launch)
    uint32_t* overflow;     // spill stack: entry k of thread g at overflow[(k*threads + g)*2]
    uint32_t overflow_depth;
    uint32_t camera;        // 1: no ray array; ray i is the camera ray of pixel i (rt::camera_ray:
    rt::CamBasis cam;       //    origin, tmin 0, unit direction, tmax +inf)
};

// ---- the traversal kernel -------------------------------------------------------------------
// Every node is fetched as eight unconditional 16-byte loads; the four slab tests are
// evaluated without short-circuit branches; the widened closest distance best_t(1 + 2^-15)
// is kept in a register; stack entries are one 8-byte LDS word {node, entry distance} with a
// global spill area addressed directly from the kernel argument.

// PARTS > 1: the rays are split into PARTS contiguous partitions and workgroup b works on
// partition b % PARTS first. Workgroups are dispatched to the XCDs round-robin, so each XCD
// (each with its own L2) takes one contiguous range of rays, whose origins are spatially
// coherent when the rays come in pixel or sorted order; each partition has its own counter,
// and a wave whose partition is exhausted continues with the next ones.
__device__ __forceinline__ uint32_t part_begin(uint32_t n, uint32_t p, uint32_t parts) {
    return p >= parts ? n : (uint32_t)(((uint64_t)n * p / parts) & ~63ull);
}

// amdgpu_waves_per_eu(6): a register budget of 80 VGPRs keeps six waves per SIMD in flight.
// MODE: kModeClosest, kModeAny, or kModeMixed: items [0, n) are any-hit rays (a.rays, results in
// a.occluded) and items [n, 2n) closest-hit rays (a.rays2 through the optional a.perm, results
// in a.hits), both indexed by the same slot, in one launch (one tail instead of two).
template <int BLOCK, int LDS_DEPTH, int MODE, int REFILL, int PARTS>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_trace3(TraceArgs a) {
    __shared__ uint2 s_stack[LDS_DEPTH * BLOCK];
    const int tid = threadIdx.x;
    const uint32_t lane = (uint32_t)(tid & 63);
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t nthreads = gridDim.x * BLOCK;
    uint2* __restrict__ spill = reinterpret_cast<uint2*>(a.overflow);
    constexpr bool kAnyOnly = MODE == kModeAny, kMixed = MODE == kModeMixed;
    const uint32_t n_rays = a.count_dev ? min(a.n, *a.count_dev) : a.n;
    // Items. Mixed: 64-item blocks alternate between the any-hit and the closest-hit rays of the
    // same 64 slots (block 2k: any-hit slots 64k.., block 2k+1: closest-hit slots 64k..), so every
    // partition holds both kinds over one range of slots; items past n_rays in the last pair of
    // blocks hold no ray.
    const uint32_t n = kMixed ? 2u * ((n_rays + 63u) & ~63u) : n_rays;
    bool lane_any = kAnyOnly;  // the lane's current ray is an any-hit ray

    int sp = 0;
    // spill index (entry - LDS_DEPTH) * nthreads + gtid: both factors are below 2^24 (at most
    // 3 * depth + 1 entries, at most 2048 threads per CU) and the sum below 2^32
    auto push = [&](uint32_t node, float t) {
        const uint2 e = make_uint2(node, __float_as_uint(t));
        if (sp < LDS_DEPTH) s_stack[sp * BLOCK + tid] = e;
        else spill[__umul24((uint32_t)(sp - LDS_DEPTH), nthreads) + gtid] = e;
        ++sp;
    };
    auto pop_live = [&](float tlim) -> uint32_t {
        while (sp > 0) {
            --sp;
            const uint2 e = rtrav::stack_entry<BLOCK, LDS_DEPTH>(s_stack, spill, sp, nthreads, gtid, tid);
            if (__uint_as_float(e.y) <= tlim) return e.x;
        }
        return kEmptyChild;
    };

    bool has = false, queue_open = true, first = true;
    // partition of this workgroup, its static range (the first 64 items of each of its waves)
    const uint32_t my_part = blockIdx.x % PARTS;
    const uint32_t local_block = blockIdx.x / PARTS;
    uint32_t cp = my_part, tries = 0;  // partition the wave currently draws from
    uint32_t ray = 0, cur = kEmptyChild, leaf = kEmptyChild;
    v3 o = rt::make_v3(0, 0, 0), inv = o;
    rtrav::DirRay dr = rtrav::make_dir_ray(o, rt::make_v3(1.0f, 1.0f, 1.0f));
    rt::RayShear sh;
    sh.o = o; sh.kx = 0; sh.ky = 1; sh.kz = 2; sh.Sx = sh.Sy = sh.Sz = 0.0f; sh.tmin = 0.0f; sh.valid = false;
    float tmin = 0.0f, best_t = 0.0f, tlim = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t best_id = kMissId;

    for (;;) {
        const uint64_t idle = __ballot(!has);
        if (queue_open && __popcll(idle) >= REFILL) {
            uint32_t base = 0, end = 0;
            if (first) {  // the wave's own first 64 items: no atomic
                base = part_begin(n, my_part, PARTS) + ((local_block * BLOCK + (uint32_t)tid) & ~63u);
                end = part_begin(n, my_part + 1, PARTS);
                first = false;
            } else {
                // blocks of partition p: those with b % PARTS == p; their static items come first
                const uint32_t nblk = (gridDim.x + PARTS - 1 - cp) / PARTS;
                const uint32_t dyn = part_begin(n, cp, PARTS) + nblk * BLOCK;  // first item of the shared queue
                end = part_begin(n, cp + 1, PARTS);
                if (lane == 0) {
                    // No atomic on a partition whose static ranges cover it, nor, once the wave steals
                    // from other partitions (the end of the launch), on one whose counter (which only
                    // grows) already shows it drained; a stale read only means the atomic is tried.
                    // The home partition takes the atomic directly: a read first would add a round
                    // trip to every refill.
                    uint32_t* q = &a.queue[cp * kQueueStride];
                    if (dyn >= end ||
                        (tries > 0 && dyn + __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= end))
                        base = end;
                    else
                        base = dyn + atomicAdd(q, (uint32_t)__popcll(idle));
                }
                base = __builtin_amdgcn_readfirstlane(base);  // every lane is active here: lane 0's value
                if (base + (uint32_t)__popcll(idle) >= end) {  // this partition is exhausted from now on
                    cp = (cp + 1) % PARTS;
                    if (++tries >= PARTS) queue_open = false;
                }
            }
            if (!has) {
                const uint32_t item = base + rank_below(idle);
                const uint32_t slot = kMixed ? (((item >> 7) << 6) | (item & 63u)) : item;
                if (item < end && (!kMixed || slot < n_rays)) {
                    const mrt_ray* src = a.rays;
                    const uint32_t idx = slot;
                    if (kMixed) {
                        lane_any = ((item >> 6) & 1u) == 0u;
                        if (!lane_any) src = a.rays2;
                    }
                    ray = (a.perm && !lane_any) ? a.perm[idx] : idx;
                    v3 d;
                    if (a.camera) {
                        rt::camera_ray(a.cam, ray % a.cam.width, ray / a.cam.width, o, d);
                        tmin = 0.0f;
                        best_t = __builtin_huge_valf();
                    } else {
                        const float4* rp = reinterpret_cast<const float4*>(src + ray);
                        const float4 r0 = rp[0], r1 = rp[1];
                        o = rt::make_v3(r0.x, r0.y, r0.z);
                        d = rt::make_v3(r1.x, r1.y, r1.z);
                        tmin = r0.w;
                        best_t = r1.w;
                    }
                    sh = rt::make_shear(o, d, tmin);
                    inv = rt::make_v3(rt::safe_rcp(d.x), rt::safe_rcp(d.y), rt::safe_rcp(d.z));
                    dr = rtrav::make_dir_ray(o, inv);
                    tlim = best_t + fabsf(best_t) * kSlack;
                    best_id = kMissId;
                    bu = bv = 0.0f;
                    sp = 0;
                    leaf = kEmptyChild;
                    cur = sh.valid ? (a.ray_root ? a.ray_root[ray] : 0u) : kEmptyChild;
                    has = true;
                }
            }
        }
        if (__ballot(has) == 0) {
            if (!queue_open) break;
            continue;
        }
        if (has) {
            while (cur != kEmptyChild && !(cur & kLeafBit)) {
                const rtrav::DirRows w = rtrav::load_dir_rows(reinterpret_cast<const char*>(a.nodes), cur, dr);
                const float4 nx = w.nx, fx = w.fx, ny = w.ny, fy = w.fy, nz = w.nz, fz = w.fz;
                const float4 cf = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(a.nodes) + ((cur << 7) + 96u));
                const uint32_t cw0 = __float_as_uint(cf.x), cw1 = __float_as_uint(cf.y);
                const uint32_t cw2 = __float_as_uint(cf.z), cw3 = __float_as_uint(cf.w);
                float t0, t1, t2, t3;
                const bool h0 = rtrav::slab_dir(nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, dr, tmin, tlim, t0) & (cw0 != kEmptyChild);
                const bool h1 = rtrav::slab_dir(nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, dr, tmin, tlim, t1) & (cw1 != kEmptyChild);
                const bool h2 = rtrav::slab_dir(nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, dr, tmin, tlim, t2) & (cw2 != kEmptyChild);
                const bool h3 = rtrav::slab_dir(nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, dr, tmin, tlim, t3) & (cw3 != kEmptyChild);
                const float kFar = __builtin_huge_valf();
                float k0 = h0 ? t0 : kFar, k1 = h1 ? t1 : kFar, k2 = h2 ? t2 : kFar, k3 = h3 ? t3 : kFar;
                uint32_t c0 = cw0, c1 = cw1, c2 = cw2, c3 = cw3;
                const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
#define CSWAP(ka, ca, kb, cb) { const bool sw = kb < ka; const float tk = sw ? kb : ka; kb = sw ? ka : kb; ka = tk; \
                                const uint32_t tc = sw ? cb : ca; cb = sw ? ca : cb; ca = tc; }
                CSWAP(k0, c0, k1, c1); CSWAP(k2, c2, k3, c3); CSWAP(k0, c0, k2, c2); CSWAP(k1, c1, k3, c3); CSWAP(k1, c1, k2, c2);
#undef CSWAP
                if (nh > 3) push(c3, k3);
                if (nh > 2) push(c2, k2);
                if (nh > 1) push(c1, k1);
                cur = nh > 0 ? c0 : pop_live(tlim);
                if (cur != kEmptyChild && (cur & kLeafBit) && leaf == kEmptyChild) {
                    leaf = cur;
                    cur = pop_live(tlim);
                }
                if (__ballot(leaf == kEmptyChild) == 0) break;
                if (__popcll(__ballot(true)) < kMinDescending) break;  // __ballot(true): the lanes still here
            }
            while (leaf != kEmptyChild) {
                const uint32_t first = (leaf >> 3) & 0x0FFFFFFFu;
                const uint32_t count = (leaf & 7u) + 1u;
                const float4* tp = a.tris + (size_t)first * 3;
                bool stop = false;
                for (uint32_t k = 0; k < count; ++k, tp += 3) {
                    const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
                    if (rt::watertight_hit(sh, rt::make_v3(p0.x, p0.y, p0.z), rt::make_v3(p1.x, p1.y, p1.z),
                                           rt::make_v3(p2.x, p2.y, p2.z), __float_as_uint(p0.w), best_t, best_id, bu, bv)) {
                        tlim = best_t + fabsf(best_t) * kSlack;
                        if (lane_any) { stop = true; break; }
                    }
                }
                if ((kAnyOnly || kMixed) && stop) { leaf = kEmptyChild; cur = kEmptyChild; sp = 0; break; }
                leaf = kEmptyChild;
                if (cur != kEmptyChild && (cur & kLeafBit)) {
                    leaf = cur;
                    cur = pop_live(tlim);
                }
            }
            if (cur == kEmptyChild && leaf == kEmptyChild) {
                if (lane_any) {
                    a.occluded[ray] = best_id != kMissId ? 1u : 0u;
                } else {
                    const float4 out = best_id != kMissId
                                           ? make_float4(best_t, __uint_as_float(best_id), bu, bv)
                                           : make_float4(__builtin_huge_valf(), __uint_as_float(kMissId), 0.0f, 0.0f);
                    *reinterpret_cast<float4*>(a.hits + ray) = out;
                }
                has = false;
            }
        }
    }
}

// ---- variant table (launch shapes the autotuner chooses from) ------------------------------
struct Variant {
    const char* name;
    int block, lds_depth, refill;
    const void* closest;
    const void* any;
    const void* mixed;  // any-hit and closest-hit rays in one launch
};

#define V3(B, D, R) {"v3_b" #B "_s" #D "_r" #R, B, D, R, (const void*)&k_trace3<B, D, kModeClosest, R, 1>, \
                     (const void*)&k_trace3<B, D, kModeAny, R, 1>, (const void*)&k_trace3<B, D, kModeMixed, R, 1>}
#define V3X(B, D, R) {"v3x8_b" #B "_s" #D "_r" #R, B, D, R, (const void*)&k_trace3<B, D, kModeClosest, R, 8>, \
                      (const void*)&k_trace3<B, D, kModeAny, R, 8>, (const void*)&k_trace3<B, D, kModeMixed, R, 8>}
// Entry 0 is the default launch shape: among the fastest measured on MI355X for incoherent
// bounce-1 rays of the 1,024,000-triangle scene and for whole 1080p frames (the best shapes are
// within a few percent of each other and trade places from run to run).
const Variant kVariants[] = {
    V3X(256, 12, 8), V3X(128, 12, 8), V3X(256, 8, 16), V3X(256, 8, 32), V3X(256, 8, 24), V3X(128, 12, 24), V3X(128, 12, 16),
    V3X(128, 12, 4), V3X(128, 16, 8), V3(128, 12, 16), V3(128, 12, 8),
    V3X(128, 12, 32), V3X(128, 16, 32), V3X(64, 16, 16),
    V3(128, 16, 32), V3(64, 16, 16), V3(128, 8, 32), V3(128, 12, 32), V3(256, 8, 32), V3(128, 16, 48),
    V3(64, 8, 16), V3(128, 24, 32),
};
#undef V3X
#undef V3
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
static_assert(kNumVariants <= (int)(sizeof(mrt_ctx::trace_occupancy[0]) / sizeof(int)), "context occupancy table");
constexpr int kDefaultVariant = 0;
constexpr int kMinLdsDepth = 8;

}  // namespace

int trace_variant_count() { return kNumVariants; }
const char* trace_variant_name(int v) { return (v >= 0 && v < kNumVariants) ? kVariants[v].name : nullptr; }
int trace_min_lds_depth() { return kMinLdsDepth; }

// Clears a traversal queue counter before a launch (a kernel, not a memset, so that captured
// frames replay it).
__global__ void k_reset_queue(uint32_t* q) {
    if (threadIdx.x < kMaxParts) q[threadIdx.x * kQueueStride] = 0u;
}

hipError_t launch_trace(mrt_ctx* ctx, const TraceTarget& s, const mrt_ray* rays, uint32_t n,
                        mrt_hit* hits, uint32_t* occluded, int any, const uint32_t* perm,
                        const uint32_t* count_dev, const uint32_t* ray_root, const rt::CamBasis* cam) {
    if (n == 0) return hipSuccess;
    hipStream_t stream = ctx->stream;
    TraceArgs a;
    a.camera = cam ? 1u : 0u;
    if (cam) a.cam = *cam;
    a.perm = perm;
    a.count_dev = count_dev;
    a.ray_root = ray_root;
    a.nodes = s.nodes;
    a.tris = s.tris;
    a.rays = rays;
    a.rays2 = nullptr;
    a.hits = hits;
    a.occluded = occluded;
    a.n = n;
    a.queue = ctx->d_counters + (any ? kMaxParts * kQueueStride : 0);
    a.overflow = s.overflow;
    a.overflow_depth = (uint32_t)s.overflow_entries;
    int vi = any ? ctx->trace_variant_any : ctx->trace_variant_closest;
    if (vi < 0 || vi >= kNumVariants) vi = kDefaultVariant;
    const Variant& v = kVariants[vi];
    const void* fn = any ? v.any : v.closest;
    int& per_cu = ctx->trace_occupancy[any ? 1 : 0][vi];
    if (per_cu == 0) {
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, v.block, 0);
        if (e != hipSuccess) { per_cu = 0; return e; }
        if (per_cu < 1) per_cu = 1;
    }
    uint32_t blocks = (uint32_t)per_cu * (uint32_t)ctx->num_cus;
    const uint32_t need = (n + v.block - 1) / v.block;
    if (blocks > need) blocks = need;
    if ((size_t)blocks * v.block > s.overflow_threads) blocks = (uint32_t)(s.overflow_threads / v.block);
    if (blocks == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, stream, a.queue);
    ProfScope prof(ctx, any ? "trace_any" : "trace_closest", stream);
    void* args[] = {&a};
    hipError_t e = hipLaunchKernel(fn, dim3(blocks), dim3(v.block), args, 0, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_trace_pair(mrt_ctx* ctx, const TraceTarget& s, const mrt_ray* shadow, uint32_t* occluded,
                             const mrt_ray* rays, mrt_hit* hits, const uint32_t* perm, uint32_t n,
                             const uint32_t* count_dev, const uint32_t* ray_root, bool queue_clear) {
    int vi = ctx->trace_variant_closest;
    if (vi < 0 || vi >= kNumVariants) vi = kDefaultVariant;
    const Variant& v = kVariants[vi];
    if (n == 0) return hipSuccess;
    TraceArgs a;
    a.camera = 0u;
    a.perm = perm;
    a.count_dev = count_dev;
    a.ray_root = ray_root;
    a.nodes = s.nodes;
    a.tris = s.tris;
    a.rays = shadow;
    a.occluded = occluded;
    a.rays2 = rays;
    a.hits = hits;
    a.n = n;
    a.queue = ctx->d_counters;
    a.overflow = s.overflow;
    a.overflow_depth = (uint32_t)s.overflow_entries;
    int& per_cu = ctx->trace_occupancy[2][vi];
    if (per_cu == 0) {
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.mixed, v.block, 0);
        if (e != hipSuccess) { per_cu = 0; return e; }
        if (per_cu < 1) per_cu = 1;
    }
    uint32_t blocks = (uint32_t)per_cu * (uint32_t)ctx->num_cus;
    const uint32_t need = (uint32_t)((2ull * n + v.block - 1) / v.block);
    if (blocks > need) blocks = need;
    if ((size_t)blocks * v.block > s.overflow_threads) blocks = (uint32_t)(s.overflow_threads / v.block);
    if (blocks == 0) return hipErrorInvalidValue;
    if (!queue_clear) hipLaunchKernelGGL(k_reset_queue, dim3(1), dim3(64), 0, ctx->stream, a.queue);
    ProfScope prof(ctx, "trace_pair");
    void* args[] = {&a};
    hipError_t e = hipLaunchKernel(v.mixed, dim3(blocks), dim3(v.block), args, 0, ctx->stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_trace(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, uint32_t n,
                        mrt_hit* hits, uint32_t* occluded, int any, const uint32_t* perm,
                        const uint32_t* count_dev) {
    TraceTarget t{s->d_nodes, s->d_tris, s->d_overflow, s->overflow_entries, s->overflow_threads, nullptr};
    return launch_trace(ctx, t, rays, n, hits, occluded, any, perm, count_dev, nullptr, nullptr);
}

hipError_t launch_trace_camera(mrt_ctx* ctx, const mrt_scene* s, const rt::CamBasis& cam, mrt_hit* hits) {
    TraceTarget t{s->d_nodes, s->d_tris, s->d_overflow, s->overflow_entries, s->overflow_threads, nullptr};
    return launch_trace(ctx, t, nullptr, cam.width * cam.height, hits, nullptr, 0, nullptr, nullptr, nullptr, &cam);
}
