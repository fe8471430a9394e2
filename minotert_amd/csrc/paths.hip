// paths.hip — wavefront diffuse path tracing over a table of scenes, scene batches and
// graph-captured frames.
//
// Reference: Pathtracer::secondaryRays (src/gfx/modules/pathtracer.ixx:118-195) running
// secondaryRays.comp, a per-pixel loop over samples (:124-134) and bounces (:64-100). Here the
// per-pixel loop becomes one launch per stage and bounce over SoA path queues:
//     init -> [trace primary] -> { shade_k -> trace_any(shadow) -> trace_closest(extension) }_k -> resolve
// A frame may cover several scenes ("batch"): their BVHs are concatenated, every path looks
// its scene up in a per-frame table (path range, root node, triangle base, camera, output),
// every ray carries its scene's root node, and each bounce is ONE persistent traversal launch
// over the shared queue of all scenes. A single scene is a table of one. A path's random
// numbers come from its own counter-seeded PCG state, so compaction order (wave ballots + one
// atomic per wave) never changes any path's result, and a scene renders bit-identically
// alone or inside a batch.
#include "internal.h"
#include "sky.h"
#include "traverse.h"

#include <cstring>
#include <new>

namespace {

using rt::v3;

constexpr uint32_t kMaxScenes = 24;

struct SceneEntry {
    uint32_t path_base;   // first path id of the scene
    uint32_t npix;        // width * height
    uint32_t tri_base;    // offset of the scene's triangles in verts/material
    uint32_t root;        // root wide node in the (concatenated) BVH
    uint32_t pix_base;    // first pixel of the scene in the batch's pixel range
    uint32_t pad[3];
    rt::CamBasis cam;
    float* rgb;                // resolve output, npix * 3 floats
    const mrt_hit* gb_hits;    // primary hits from a G-buffer, or null: traced in the wavefront
};

struct SceneTable {
    uint32_t n, total_paths, total_pix, spp;
    SceneEntry e[kMaxScenes];
};

struct PathBuffers {
    mrt_ray* ray;
    uint32_t* pid;
    uint32_t* rng;
    float4* thr;    // xyz throughput, w = 1 while the path may spawn rays, 0 once only NEE is pending
    float4* L;      // xyz accumulated radiance
    float4* pend;   // xyz NEE contribution waiting for its shadow ray
};

struct ShadeParams {
    v3 sky, sun_dir, sun_e;
    uint32_t bounces, rr_start, bounce, capture;
    uint32_t nee_on;   // 0 when sun_e is zero: no shadow rays
    uint32_t sky_on;   // escaped rays take the physical sky's radiance (skyd)
    uint32_t init;     // vertex 0 of a frame with G-buffer hits: k_shade builds the initial path state
    uint32_t sample_base, frame_seed;  // (init) the RNG seeding of k_paths_init
    SkyDev skyd;
};

constexpr uint32_t kCacheOff = 0xFFFFFFFFu;

// Cache termination state of a frame. Training paths are drawn among the paths that reach the
// cache vertex on a surface: path pid is one when train != 0 and pcg(pid ^ tseed) < *tsel, the
// threshold that keeps each such path with probability min(1, target / count) for the frame's
// count of them (k_train_select; 0xFFFFFFFF keeps all). The predicate is only asked of paths at
// or past the vertex.
struct CacheIO {
    mrt_cache_sample* q;   // one query per path that reached `vertex` on a surface
    uint32_t* q_pid;
    float4* q_L;           // radiance gathered before the vertex
    float4* q_thr;         // throughput arriving at the vertex
    uint32_t* q_count;
    float4* train_L;       // final radiance of the training paths, by path id
    const uint32_t* bad;   // cache not finite: the frame renders plain
    const uint32_t* limit; // a budget of the frame (queries, loss) exceeded: the frame renders plain
    const uint32_t* tsel;  // selection threshold of the training paths (device)
    uint32_t vertex;       // kCacheOff: no cache
    uint32_t train, tseed;
};

__device__ __forceinline__ bool training_path(const CacheIO& C, uint32_t pid) {
    if (C.train == 0u) return false;
    const uint32_t t = *C.tsel;
    return t == 0xFFFFFFFFu || rt::pcg_hash(pid ^ C.tseed) < t;
}

__device__ __forceinline__ uint32_t scene_of_path(const SceneTable& T, uint32_t p) {
    uint32_t lo = 0, hi = T.n;  // last entry with path_base <= p
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (T.e[mid].path_base <= p) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t scene_of_pixel(const SceneTable& T, uint32_t q) {
    uint32_t lo = 0, hi = T.n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (T.e[mid].pix_base <= q) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Compacted append: every active lane of the wave must call it; returns the slot for lanes
// with keep = true.
__device__ __forceinline__ uint32_t wave_append(bool keep, uint32_t* counter) {
    const uint64_t m = __ballot(keep);
    const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
    const uint32_t total = (uint32_t)__popcll(m);
    const uint64_t active = __ballot(true);
    const uint32_t leader = (uint32_t)(__ffsll((long long)active) - 1);
    uint32_t base = 0;
    if (__lane_id() == leader && total) base = atomicAdd(counter, total);
    base = __shfl(base, (int)leader, 64);
    return base + rank;
}

// The same for a whole block with one atomic: every thread of the block must call it (the
// waves' totals meet in LDS). Appends from all blocks go to one address, and such atomics are
// served one after another, so one per block instead of one per wave shortens a launch whose
// waves are short.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_append(bool keep, uint32_t* counter) {
    __shared__ uint32_t s_tot[BLOCK / 64 + 1];
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(keep);
    const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
    if (__lane_id() == 0) s_tot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            const uint32_t t = s_tot[w];
            s_tot[w] = acc;
            acc += t;
        }
        s_tot[BLOCK / 64] = acc ? atomicAdd(counter, acc) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_tot[BLOCK / 64] + s_tot[wave] + rank;
    __syncthreads();  // s_tot is reused by the next call
    return slot;
}
constexpr int kShadeBlock = 1024;

__device__ __forceinline__ void tri_vertices(const float* __restrict__ verts, uint32_t tri, v3& v0, v3& v1, v3& v2) {
    const float* p = verts + (size_t)tri * 9;
    v0 = rt::make_v3(p[0], p[1], p[2]);
    v1 = rt::make_v3(p[3], p[4], p[5]);
    v2 = rt::make_v3(p[6], p[7], p[8]);
}

__global__ __launch_bounds__(256) void k_paths_init(SceneTable T, uint32_t sample_base, uint32_t frame_seed,
                                                    const uint32_t* __restrict__ frame_dev, PathBuffers cur,
                                                    mrt_hit* __restrict__ hits, uint32_t* __restrict__ root,
                                                    uint32_t* __restrict__ counts, uint32_t ncounts) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p == 0) {
        counts[0] = T.total_paths;
        for (uint32_t k = 1; k < ncounts; ++k) counts[k] = 0;
    }
    if (p >= T.total_paths) return;
    const uint32_t seed = frame_dev ? (frame_seed ^ *frame_dev) : frame_seed;
    const SceneEntry& e = T.e[T.n == 1 ? 0u : scene_of_path(T, p)];
    const uint32_t local = p - e.path_base;
    const uint32_t pixel = local % e.npix, j = local / e.npix;
    v3 o, d;
    rt::camera_ray(e.cam, pixel % e.cam.width, pixel / e.cam.width, o, d);
    float4* r = reinterpret_cast<float4*>(cur.ray + p);
    r[0] = make_float4(o.x, o.y, o.z, 0.0f);
    r[1] = make_float4(d.x, d.y, d.z, __builtin_huge_valf());
    if (e.gb_hits) hits[p] = e.gb_hits[pixel];
    root[p] = e.root;
    cur.pid[p] = p;
    cur.rng[p] = rt::path_seed(pixel, sample_base + j, seed);
    cur.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    cur.L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    cur.pend[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// State of one path between two vertices (the payload of the reference's per-pixel loop,
// secondaryRays.comp:64-100, with the RNG state travelling with the path).
struct PathState {
    uint32_t pid, rng, root;
    float4 thr;           // xyz throughput, w = 1 while the path may spawn rays
    float4 L;             // xyz radiance gathered so far
    float4 pend;          // NEE contribution of the last vertex, waiting for its shadow ray
    float4 ray0, ray1;    // extension ray (origin, tmin | direction, tmax)
    float4 sh0, sh1;      // shadow ray of the pending NEE
};

struct QueryRec {
    bool query;
    float4 qs[4], qL, qthr;
};

// Shading of vertex P.bounce of one path (the body of traceSecondaryRays, secondaryRays.comp:
// 67-98, for one iteration) with next-event estimation of the directional light: first the
// pending NEE of the previous vertex is added when its shadow ray was unoccluded, then the hit of
// the current ray (direction d) is shaded. With a cache (cache_on), a path that reaches vertex
// C.vertex on a surface records a cache query and ends there unless it is a training path.
// Returns true when the path is finished. Shared by the wavefront and the megakernel, so both
// evaluate the same operation sequence. Training paths record their query and continue past
// the vertex.
__device__ __forceinline__ bool shade_vertex(const ShadeParams& P, const SceneTable& T, const CacheIO& C, bool cache_on,
                                             const float* __restrict__ verts, const uint32_t* __restrict__ material,
                                             const float* __restrict__ albedo, bool occluded, const mrt_hit& h, v3 d,
                                             PathState& S, QueryRec& Q) {
    if (P.bounce > 0 && !occluded) {
        S.L.x = S.L.x + S.pend.x; S.L.y = S.L.y + S.pend.y; S.L.z = S.L.z + S.pend.z;
    }
    S.pend = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    S.ray0 = S.ray1 = S.sh0 = S.sh1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool alive = S.thr.w != 0.0f;
    if (alive && h.tri == kMissId) {
        v3 env = P.sky;
        if (P.sky_on) {
            const float3 e = sky::radiance(P.skyd, make_float3(d.x, d.y, d.z));
            env = rt::make_v3(e.x, e.y, e.z);
        }
        S.L.x = S.L.x + S.thr.x * env.x; S.L.y = S.L.y + S.thr.y * env.y; S.L.z = S.L.z + S.thr.z * env.z;
        return true;
    }
    if (!(alive && P.bounce < P.bounces)) return true;
    const SceneEntry& e = T.e[T.n == 1 ? 0u : scene_of_path(T, S.pid)];
    S.root = e.root;
    const uint32_t tri = e.tri_base + h.tri;
    v3 v0, v1, v2, p, nrm;
    tri_vertices(verts, tri, v0, v1, v2);
    rt::surface_frame(v0, v1, v2, h.u, h.v, d, p, nrm);
    const float* a3 = albedo + (size_t)material[tri] * 3;
    const v3 a = rt::make_v3(a3[0], a3[1], a3[2]);
    if (cache_on && P.bounce == C.vertex) {
        Q.query = true;
        Q.qs[0] = make_float4(p.x, p.y, p.z, 0.0f);
        Q.qs[1] = make_float4(d.x, d.y, d.z, 0.0f);
        Q.qs[2] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
        Q.qs[3] = make_float4(a.x, a.y, a.z, 0.0f);
        Q.qL = make_float4(S.L.x, S.L.y, S.L.z, 0.0f);
        Q.qthr = make_float4(S.thr.x, S.thr.y, S.thr.z, 0.0f);
        if (!training_path(C, S.pid)) return true;
    }
    const v3 po = rt::offset_origin(p, nrm);
    const float cosl = rt::vdot(nrm, P.sun_dir);
    bool nee = false;
    if (cosl > 0.0f && P.nee_on) {
        nee = true;
        S.pend = make_float4(((S.thr.x * a.x) * cosl) * P.sun_e.x, ((S.thr.y * a.y) * cosl) * P.sun_e.y,
                             ((S.thr.z * a.z) * cosl) * P.sun_e.z, 0.0f);
        S.sh0 = make_float4(po.x, po.y, po.z, 0.0f);
        S.sh1 = make_float4(P.sun_dir.x, P.sun_dir.y, P.sun_dir.z, __builtin_huge_valf());
    }
    S.thr.x = S.thr.x * a.x; S.thr.y = S.thr.y * a.y; S.thr.z = S.thr.z * a.z;
    bool spawn = true;
    if (P.bounce + 1 >= P.rr_start) {
        const float q = fminf(0.95f, fmaxf(S.thr.x, fmaxf(S.thr.y, S.thr.z)));
        const float xi = rt::next_float(S.rng);
        if (!(xi < q)) spawn = false;
        else { S.thr.x = S.thr.x / q; S.thr.y = S.thr.y / q; S.thr.z = S.thr.z / q; }
    }
    if (spawn) {
        const float xi1 = rt::next_float(S.rng);
        const float xi2 = rt::next_float(S.rng);
        const v3 nd = rt::cosine_direction(nrm, xi1, xi2);
        S.ray0 = make_float4(po.x, po.y, po.z, 0.0f);
        S.ray1 = make_float4(nd.x, nd.y, nd.z, __builtin_huge_valf());
    }
    S.thr.w = spawn ? 1.0f : 0.0f;
    return !spawn && !nee;
}

// Where a finished path's radiance goes: a query path gets its radiance from the cache later;
// a training path that passed the cache vertex keeps its full estimate for the training target.
__device__ __forceinline__ void finish_path(const ShadeParams& P, const CacheIO& C, bool cache_on, const PathState& S,
                                            bool query, float4* __restrict__ out_L) {
    if (cache_on && (query || P.bounce > C.vertex) && training_path(C, S.pid)) {
        C.train_L[S.pid] = S.L;
    } else if (!query) {
        out_L[S.pid] = S.L;
    }
}

// One shading step for every path of bounce k, then compaction of the live paths (one atomic
// per block) into the next queue, their shadow rays and, with a cache, the query list.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_shade(ShadeParams P, SceneTable T, PathBuffers cur, PathBuffers nxt,
                                               const mrt_hit* __restrict__ hits, const uint32_t* __restrict__ occ,
                                               mrt_ray* __restrict__ shadow, uint32_t* __restrict__ root,
                                               const float* __restrict__ verts, const uint32_t* __restrict__ material,
                                               const float* __restrict__ albedo, float4* __restrict__ out_L,
                                               uint32_t* __restrict__ counts, CacheIO C,
                                               const uint32_t* __restrict__ frame_dev) {
    const uint32_t s = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t n = P.init ? T.total_paths : counts[P.bounce];
    const bool active = s < n;
    // Before the cache vertex the cache plays no part (and its flag may still be being written by
    // training steps running beside this frame): not read there.
    const bool cache_on = C.vertex != kCacheOff && P.bounce >= C.vertex && *C.bad == 0u && *C.limit == 0u;
    QueryRec Q;
    Q.query = false;
    PathState S;
    S.pid = S.rng = S.root = 0;
    S.thr = S.L = S.pend = S.ray0 = S.ray1 = S.sh0 = S.sh1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool keep = false;
    if (active) {
        mrt_hit h;
        v3 d;
        if (P.init) {  // the state k_paths_init would have written for path s, and its G-buffer hit
            const SceneEntry& e = T.e[T.n == 1 ? 0u : scene_of_path(T, s)];
            const uint32_t local = s - e.path_base;
            const uint32_t pixel = local % e.npix, j = local / e.npix;
            const uint32_t seed = frame_dev ? (P.frame_seed ^ *frame_dev) : P.frame_seed;
            S.pid = s;
            S.rng = rt::path_seed(pixel, P.sample_base + j, seed);
            S.thr = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            v3 o;
            rt::camera_ray(e.cam, pixel % e.cam.width, pixel / e.cam.width, o, d);
            h = e.gb_hits[pixel];
        } else {
            S.pid = cur.pid[s];
            S.rng = cur.rng[s];
            S.thr = cur.thr[s];
            S.L = cur.L[s];
            S.pend = cur.pend[s];
            const float4 r1 = reinterpret_cast<const float4*>(cur.ray + s)[1];
            d = rt::make_v3(r1.x, r1.y, r1.z);
            h = hits[s];
        }
        const bool finished = shade_vertex(P, T, C, cache_on, verts, material, albedo, P.bounce > 0 && occ[s] != 0u,
                                           h, d, S, Q);
        keep = !finished;
        if (finished && !P.capture) finish_path(P, C, cache_on, S, Q.query, out_L);
    }
    if (C.vertex != kCacheOff) {  // uniform over the grid
        const uint32_t qslot = block_append<BLOCK>(Q.query, C.q_count);
        if (Q.query) {
            float4* q = reinterpret_cast<float4*>(C.q + qslot);
            q[0] = Q.qs[0]; q[1] = Q.qs[1]; q[2] = Q.qs[2]; q[3] = Q.qs[3];
            C.q_pid[qslot] = S.pid;
            C.q_L[qslot] = Q.qL;
            C.q_thr[qslot] = Q.qthr;
        }
    }
    const uint32_t slot = block_append<BLOCK>(keep, &counts[P.bounce + 1]);
    if (keep) {
        nxt.pid[slot] = S.pid;
        nxt.rng[slot] = S.rng;
        nxt.thr[slot] = S.thr;
        nxt.L[slot] = S.L;
        nxt.pend[slot] = S.pend;
        root[slot] = S.root;
        float4* r = reinterpret_cast<float4*>(nxt.ray + slot);
        r[0] = S.ray0; r[1] = S.ray1;
        float4* sr = reinterpret_cast<float4*>(shadow + slot);
        sr[0] = S.sh0; sr[1] = S.sh1;
    }
}

// Megakernel path loop: every path is carried by a pair of lanes (2i, 2i+1 of a wave) from the
// state the wavefront left after vertex P0.bounce - 1 (its extension ray, pending NEE and
// shadow ray) through all remaining vertices. Both lanes hold the same path state. At each
// vertex the even lane traces the pending shadow ray (any hit) and the odd lane the extension
// ray (closest hit) in the same traversal loop, so a vertex costs the longer of the two
// traversals instead of their sum (the tail of a frame is latency-bound: few paths, one wave
// per SIMD); the lanes exchange their results and both shade the vertex with shade_vertex,
// identically. The traversal and shading are those of the wavefront, so each path ends with
// the same radiance bit for bit. It runs after the cache vertex of a cached frame, so it issues
// no cache queries; there it carries the training paths (and, when the cache is not finite,
// every path). The launch bounds hold it to four waves per SIMD.
// Work distribution: each lane pair runs one path at a time, one vertex per iteration of its
// wave's loop; a pair whose path ended takes the next path at the top of the loop (first the
// static one, gtid >> 1, then from *queue past the grid's pairs; *queue is zero at the launch),
// so the pairs of a wave do not wait for the longest path of a fixed group of paths.
template <int BLOCK, int LDS_DEPTH>
__global__ __launch_bounds__(BLOCK, 4) void k_paths_mega(ShadeParams P0, SceneTable T, PathBuffers cur,
                                                      const mrt_ray* __restrict__ shadow, const uint32_t* __restrict__ root,
                                                      const uint32_t* __restrict__ count, uint32_t nmax,
                                                      const WNode* __restrict__ nodes, const float4* __restrict__ tris,
                                                      uint32_t* __restrict__ overflow, const float* __restrict__ verts,
                                                      const uint32_t* __restrict__ material, const float* __restrict__ albedo,
                                                      float4* __restrict__ out_L, CacheIO C, uint32_t* __restrict__ queue) {
    __shared__ uint2 s_stack[LDS_DEPTH * BLOCK];
    const int tid = threadIdx.x;
    const uint32_t lane = (uint32_t)(tid & 63);
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t nthreads = gridDim.x * BLOCK;
    const bool shadow_lane = (tid & 1) == 0;
    // the lanes below this lane's pair (the rank of the pair among the idle ones)
    const uint64_t pairs_below = (lane & ~1u) == 0u ? 0ull : (~0ull >> (64u - (lane & ~1u)));
    uint2* spill = reinterpret_cast<uint2*>(overflow);
    const uint32_t n = min(*count, nmax);
    const bool cache_on = C.vertex != kCacheOff && *C.bad == 0u && *C.limit == 0u;
    PathState S;
    ShadeParams P = P0;
    auto load = [&](uint32_t s) {  // both lanes of a pair load the same path
        S.pid = cur.pid[s];
        S.rng = cur.rng[s];
        S.root = root[s];
        S.thr = cur.thr[s];
        S.L = cur.L[s];
        S.pend = cur.pend[s];
        const float4* r = reinterpret_cast<const float4*>(cur.ray + s);
        S.ray0 = r[0]; S.ray1 = r[1];
        const float4* sr = reinterpret_cast<const float4*>(shadow + s);
        S.sh0 = sr[0]; S.sh1 = sr[1];
        P = P0;
    };
    // The paths are split into kMaxParts contiguous partitions and workgroup b works on partition
    // b % kMaxParts first: workgroups are dispatched to the XCDs round-robin, so each XCD (with its
    // own L2) takes one range of paths, whose pixels are close together. A pair takes its first
    // path statically, later ones from its partition's counter (queue[p * kQueueStride]); a wave
    // whose partition is exhausted continues with the next ones.
    const uint32_t my_part = blockIdx.x % kMaxParts;
    const uint32_t local_block = blockIdx.x / kMaxParts;
    auto part_at = [&](uint32_t p) { return (uint32_t)((uint64_t)n * p / kMaxParts); };
    uint32_t cp = my_part, tries = 0;  // partition the wave currently draws from
    const uint32_t first = part_at(my_part) + local_block * (uint32_t)(BLOCK / 2) + ((uint32_t)tid >> 1);
    bool active = first < part_at(my_part + 1);
    bool open = true;  // some partition may hold paths past its static ones
    if (active) load(first);
    for (;;) {
        if (open) {
            const uint64_t idle = __ballot(!active) & 0x5555555555555555ull;  // even lanes of idle pairs
            if (idle != 0ull) {
                const uint32_t k = (uint32_t)__popcll(idle);
                // partition cp: its static paths are those of the workgroups b with b % kMaxParts == cp
                const uint32_t nblk = (gridDim.x + kMaxParts - 1 - cp) / kMaxParts;
                const uint32_t dyn = part_at(cp) + nblk * (uint32_t)(BLOCK / 2);
                const uint32_t end = part_at(cp + 1);
                uint32_t base = 0;
                if (lane == 0) base = dyn >= end ? end : dyn + atomicAdd(&queue[cp * kQueueStride], k);
                base = __builtin_amdgcn_readfirstlane(base);  // every lane is active here: lane 0's value
                if (!active) {
                    const uint32_t item = base + (uint32_t)__popcll(idle & pairs_below);
                    if (item < end) {
                        load(item);
                        active = true;
                    }
                }
                if (base + k >= end) {  // this partition is exhausted from now on
                    cp = (cp + 1) % kMaxParts;
                    if (++tries >= kMaxParts) open = false;
                }
            }
        }
        // A refill either gives every idle pair a path or moves on to the next partition, so a
        // wave with no busy pair retries at most kMaxParts times before its queue is closed.
        if (__ballot(active) == 0ull) {
            if (!open) break;
            continue;
        }
        if (active) {
            // traces of the previous vertex: its shadow ray (even lane), its extension ray (odd lane);
            // a zero-length shadow ray means no pending NEE, thr.w == 0 no extension ray
            const v3 d = rt::make_v3(S.ray1.x, S.ray1.y, S.ray1.z);
            const bool traced = shadow_lane ? S.sh1.w != 0.0f : S.thr.w != 0.0f;
            float bt = shadow_lane ? S.sh1.w : S.ray1.w, bu = 0.0f, bv = 0.0f;
            uint32_t bid = kMissId;
            if (traced) {
                const float4 ro = shadow_lane ? S.sh0 : S.ray0;
                const v3 rd = shadow_lane ? rt::make_v3(S.sh1.x, S.sh1.y, S.sh1.z) : d;
                rtrav::traverse<BLOCK, LDS_DEPTH>(nodes, tris, S.root, rt::make_v3(ro.x, ro.y, ro.z), rd, ro.w, bt, bid, bu,
                                                  bv, shadow_lane, s_stack, spill, nthreads, gtid, tid);
            }
            const uint32_t pbid = __shfl_xor(bid, 1, 64);
            const float pbt = __shfl_xor(bt, 1, 64), pbu = __shfl_xor(bu, 1, 64), pbv = __shfl_xor(bv, 1, 64);
            const bool occluded = (shadow_lane ? bid : pbid) != kMissId;
            mrt_hit h;
            h.t = __builtin_huge_valf(); h.tri = kMissId; h.u = h.v = 0.0f;
            const uint32_t cbid = shadow_lane ? pbid : bid;
            if (cbid != kMissId) {
                h.t = shadow_lane ? pbt : bt;
                h.tri = cbid;
                h.u = shadow_lane ? pbu : bu;
                h.v = shadow_lane ? pbv : bv;
            }
            QueryRec Q;
            Q.query = false;
            if (shade_vertex(P, T, C, cache_on, verts, material, albedo, occluded, h, d, S, Q)) {
                if (shadow_lane) finish_path(P, C, cache_on, S, false, out_L);
                active = false;
            } else {
                P.bounce += 1;
            }
        }
    }
}

// ---- frame variants: the shading block (the queue-append chunk: one atomic per block) and the
// megakernel's block and LDS stack depth, chosen per context (the autotuner's knobs besides the
// traversal shape). Every variant evaluates the same operations per path, so images are equal.
struct FrameVariant {
    const char* name;
    int shade_block, mega_block, mega_lds;
};
// Entry 0 is the default: 1024-thread shading blocks (one append atomic per block, which kept the
// first shading launch of a 1080p frame off the atomic rate) and 128-thread megakernel blocks
// with a 16-entry stack (24 or 32 entries cost occupancy, 64-thread blocks were equal).
const FrameVariant kFrameVariants[] = {
    {"s1024_m128x16", 1024, 128, 16}, {"s512_m128x16", 512, 128, 16}, {"s256_m128x16", 256, 128, 16},
    {"s1024_m64x16", 1024, 64, 16},   {"s1024_m256x8", 1024, 256, 8}, {"s1024_m128x12", 1024, 128, 12},
    {"s512_m64x16", 512, 64, 16},
};
constexpr int kNumFrameVariants = (int)(sizeof(kFrameVariants) / sizeof(kFrameVariants[0]));
static_assert(kNumFrameVariants <= kMaxFrameVariants, "context occupancy table");
inline const FrameVariant& frame_variant(const mrt_ctx* ctx) {
    const int v = ctx->frame_variant;
    return kFrameVariants[(v >= 0 && v < kNumFrameVariants) ? v : 0];
}

// Exports the extension rays of the live paths (capture mode).
__global__ __launch_bounds__(256) void k_capture(PathBuffers nxt, const uint32_t* __restrict__ count,
                                                 mrt_ray* __restrict__ out_rays, uint32_t* __restrict__ out_pid,
                                                 uint32_t* __restrict__ out_count) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    const bool live = s < *count && nxt.thr[s].w != 0.0f;
    const uint32_t slot = wave_append(live, out_count);
    if (live) {
        const float4* r = reinterpret_cast<const float4*>(nxt.ray + s);
        float4* o = reinterpret_cast<float4*>(out_rays + slot);
        o[0] = r[0]; o[1] = r[1];
        out_pid[slot] = nxt.pid[s];
    }
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// A frame's counters in one launch: the queue counts, the closest-hit / mixed traversal queue
// (d_counters word 0 on) and both megakernel path queues.
__global__ __launch_bounds__(64) void k_frame_clear(uint32_t* counts, uint32_t ncounts, uint32_t* d_counters) {
    const uint32_t i = threadIdx.x;
    if (i < ncounts) counts[i] = 0u;
    if (i < kMaxParts) {
        d_counters[i * kQueueStride] = 0u;
        d_counters[kMegaQueue + i * kQueueStride] = 0u;
        d_counters[kMegaQueueSide + i * kQueueStride] = 0u;
    }
}

// The partition counters of one megakernel path queue.
__global__ __launch_bounds__(64) void k_clear_parts(uint32_t* queue) {
    if (threadIdx.x < kMaxParts) queue[threadIdx.x * kQueueStride] = 0u;
}

// rgb[pixel] = (sum over samples j in order of L[pixel + j*npix]) / spp  (secondaryRays.comp:131-133);
// advances the device frame counter of a captured frame.
__global__ __launch_bounds__(256) void k_resolve(SceneTable T, const float4* __restrict__ out_L,
                                                 uint32_t* __restrict__ frame_dev) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q == 0 && frame_dev) *frame_dev += 1u;
    if (q >= T.total_pix) return;
    const SceneEntry& e = T.e[T.n == 1 ? 0u : scene_of_pixel(T, q)];
    const uint32_t pixel = q - e.pix_base;
    float4 acc = out_L[e.path_base + pixel];
    for (uint32_t j = 1; j < T.spp; ++j) {
        const float4 v = out_L[e.path_base + pixel + j * e.npix];
        acc.x = acc.x + v.x; acc.y = acc.y + v.y; acc.z = acc.z + v.z;
    }
    const float fs = (float)T.spp;
    e.rgb[3 * pixel + 0] = acc.x / fs;
    e.rgb[3 * pixel + 1] = acc.y / fs;
    e.rgb[3 * pixel + 2] = acc.z / fs;
}

// Adds the node / triangle offsets of a scene to its child words inside a batch.
__global__ __launch_bounds__(256) void k_rebase_nodes(WNode* __restrict__ nodes, uint32_t n, uint32_t node_base,
                                                      uint32_t tri_base) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n * 4) return;
    uint32_t& w = nodes[i >> 2].child[i & 3];
    if (w == kEmptyChild) return;
    if (w & kLeafBit) {
        const uint32_t first = ((w >> 3) & 0x0FFFFFFFu) + tri_base;
        w = kLeafBit | (first << 3) | (w & 7u);
    } else {
        w += node_base;
    }
}

__global__ __launch_bounds__(256) void k_add_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t add) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] += add;
}

// out_L[pid] = L_before + thr * max(0, cache) for every query; optional per-pixel export.
__global__ __launch_bounds__(256) void k_cache_apply(CacheIO C, const float4* __restrict__ y, float4* __restrict__ out_L,
                                                     float* __restrict__ cache_rgb, uint32_t npix) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    if (s >= *C.q_count) return;
    const uint32_t pid = C.q_pid[s];
    const float4 v = y[s], Lb = C.q_L[s], th = C.q_thr[s];
    const float cx = fmaxf(v.x, 0.0f), cy = fmaxf(v.y, 0.0f), cz = fmaxf(v.z, 0.0f);
    out_L[pid] = make_float4(Lb.x + th.x * cx, Lb.y + th.y * cy, Lb.z + th.z * cz, 0.0f);
    if (cache_rgb) {
        const uint32_t px = pid % npix;
        cache_rgb[3 * px + 0] = cx;
        cache_rgb[3 * px + 1] = cy;
        cache_rgb[3 * px + 2] = cz;
    }
}

// Selection threshold of the frame's training paths: the paths that reach the cache vertex on a
// surface are counted (hits of the vertex's rays: the G-buffer at vertex 0, the traced extension
// rays after it), and each is kept with probability min(1, target / count): thr = target 2^32 /
// count, or 0xFFFFFFFF (all) when count <= target. Two launches: the count (one atomic per block
// into sel[0], cleared with the frame's counters), then one thread writes sel[1].
// A bounded grid strides over the paths (appends to one counter from many blocks are served one
// after another).
__global__ __launch_bounds__(256) void k_train_count(const mrt_hit* __restrict__ hits, const uint32_t* __restrict__ n_dev,
                                                     uint32_t nmax, const mrt_hit* __restrict__ gb, uint32_t npix,
                                                     uint32_t* __restrict__ sel) {
    const uint32_t n = gb ? nmax : min(*n_dev, nmax);  // G-buffer: path s at pixel s % npix
    uint32_t c = 0;
    for (uint32_t s = blockIdx.x * 256u + threadIdx.x; s < n; s += gridDim.x * 256u)
        c += (gb ? gb[s % npix].tri : hits[s].tri) != kMissId ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    __shared__ uint32_t s_sum;
    if (threadIdx.x == 0) s_sum = 0u;
    __syncthreads();
    if (__lane_id() == 0 && c) atomicAdd(&s_sum, c);
    __syncthreads();
    if (threadIdx.x == 0 && s_sum) atomicAdd(&sel[0], s_sum);
}

__global__ void k_train_threshold(uint32_t* __restrict__ sel, uint32_t target) {
    const uint32_t count = sel[0];
    sel[1] = count <= target ? 0xFFFFFFFFu : (uint32_t)(((uint64_t)target << 32) / count);
}

// Training batch of the frame: the queries of training paths with target
// (L_final - L_before) / thr per channel.
// Sample j of the batch (in append order, also the order of the optional export) goes to
// optimizer step j % steps, row j / steps of that step's share (shares start chunk rows apart),
// so the steps train on equal shares whatever the batch size. Samples past the steps * chunk rows
// of the batch (which the selection makes vanishingly rare) are dropped.
__global__ __launch_bounds__(kShadeBlock) void k_cache_targets(CacheIO C, mrt_cache_sample* __restrict__ ts,
                                                       float4* __restrict__ tt, uint32_t* __restrict__ tcount,
                                                       uint32_t steps, uint32_t chunk,
                                                       mrt_cache_sample* __restrict__ ex_s, float4* __restrict__ ex_t,
                                                       uint32_t* __restrict__ ex_pid) {
    const uint32_t s = blockIdx.x * kShadeBlock + threadIdx.x;
    const bool act = s < *C.q_count;
    const uint32_t pid = act ? C.q_pid[s] : 0u;
    const bool t = act && training_path(C, pid);
    const uint32_t slot = block_append<kShadeBlock>(t, tcount);  // one atomic per block
    if (t && slot < steps * chunk) {
        const uint32_t row = (slot % steps) * chunk + slot / steps;
        const float4 Lf = C.train_L[pid], Lb = C.q_L[s], th = C.q_thr[s];
        const float4 tg = make_float4(th.x > 0.0f ? (Lf.x - Lb.x) / th.x : 0.0f, th.y > 0.0f ? (Lf.y - Lb.y) / th.y : 0.0f,
                                      th.z > 0.0f ? (Lf.z - Lb.z) / th.z : 0.0f, 0.0f);
        const float4* src = reinterpret_cast<const float4*>(C.q + s);
        const float4 a = src[0], b = src[1], c = src[2], d = src[3];
        float4* dst = reinterpret_cast<float4*>(ts + row);
        dst[0] = a; dst[1] = b; dst[2] = c; dst[3] = d;
        tt[row] = tg;
        if (ex_s) {
            float4* e = reinterpret_cast<float4*>(ex_s + slot);
            e[0] = a; e[1] = b; e[2] = c; e[3] = d;
        }
        if (ex_t) ex_t[slot] = tg;
        if (ex_pid) ex_pid[slot] = pid;
    }
}

// Sizes of the per-step shares of the training batch: out[i] = #{j < n : j % steps == i}, with n
// at most the steps * chunk rows of the batch.
__global__ void k_split_counts(const uint32_t* __restrict__ n, uint32_t steps, uint32_t chunk, uint32_t* __restrict__ out,
                               uint32_t* __restrict__ export_n) {
    const uint32_t i = threadIdx.x;
    const uint32_t total = min(*n, steps * chunk);
    if (i < steps) out[i] = total > i ? (total - i + steps - 1) / steps : 0u;
    if (i == 0 && export_n) *export_n = total;
}

inline uint32_t blocks_for(size_t n) { return (uint32_t)((n + 255) / 256); }

struct Geometry {
    TraceTarget trace;
    const float* verts;
    const uint32_t* material;
    const float* albedo;
    const float* bounds;
};

// Rows of a frame's training batch for a target of n samples: n + n/8 + 256 (the selected count
// is binomial with mean at most n; this is over 30 standard deviations above it).
inline size_t train_capacity(uint32_t n) { return (size_t)n + n / 8 + 256; }

struct CacheFrame {
    mrt_cache* cache;
    CacheIO io;
    uint32_t target;         // training samples per frame
    uint32_t steps, chunk;   // optimizer steps, rows per step
    float* cache_rgb;
    uint32_t npix;
    mrt_cache_sample* ex_s;
    float* ex_t;
    uint32_t* ex_pid;
    uint32_t* ex_count;
    bool defer;              // leave this frame's training steps for the next cached frame
    bool wait_train;         // launch the previous frame's steps on ctx->side, wait before the cache vertex
    uint32_t limits, query_budget;  // MRT_CACHE_LIMIT_* checks of the frame
    float loss_budget;
};

// The frame's budgets, checked before the cache vertex is shaded: *limit = 1 when more paths
// reach the vertex on a surface than query_budget (at_vertex: their count) or when the mean loss
// per sample of the cache's last training pass (parts[0..blocks) summed, over the count at
// parts[kLossCountSlot]) exceeds loss_budget. One block of 256 threads.
__global__ __launch_bounds__(256) void k_cache_limits(const uint32_t* __restrict__ at_vertex, uint32_t limits,
                                                      uint32_t query_budget, float loss_budget,
                                                      const float* __restrict__ parts, uint32_t blocks,
                                                      uint32_t* __restrict__ limit) {
    __shared__ double s_sum[4];
    double sum = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += 256u) sum += (double)parts[b];
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (__lane_id() == 0) s_sum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x != 0) return;
    bool over = false;
    if (limits & MRT_CACHE_LIMIT_QUERIES) over |= *at_vertex > query_budget;
    if ((limits & MRT_CACHE_LIMIT_LOSS) && blocks) {
        const uint32_t n = __float_as_uint(parts[kLossCountSlot]);
        const double mean = (((s_sum[0] + s_sum[1]) + s_sum[2]) + s_sum[3]) / (double)(n ? n : 1u);
        over |= !(mean <= (double)loss_budget);  // a non-finite loss counts as a spike
    }
    *limit = over ? 1u : 0u;
}

// Launches the context's deferred training steps (those of cf->cache) on ctx->side, after the end
// of the frame that formed their batch (ev_pending), and records ev_train after them. The steps
// advance the cache's step counter, so the frame's non-finite flag address is taken again.
hipError_t launch_deferred(mrt_ctx* ctx, CacheFrame* cf, CacheIO& C) {
    hipError_t e = hipStreamWaitEvent(ctx->side, ctx->ev_pending, 0);
    if (e != hipSuccess) return e;
    std::swap(ctx->stream, ctx->side);  // the cache's launches go to ctx->stream
    e = flush_training(ctx, cf->cache);
    std::swap(ctx->stream, ctx->side);
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(ctx->ev_train, ctx->side)) != hipSuccess) return e;
    if ((e = cache_check(ctx, cf->cache, &cf->io.bad)) != hipSuccess) return e;
    C.bad = cf->io.bad;
    return hipSuccess;
}

// The frame's optimizer steps over the training batch in the workspace (shares of chunk rows,
// sizes at c_counts[kStepCounts + i]), as one group of steps, on ctx->stream.
hipError_t train_steps(mrt_ctx* ctx, mrt_cache* cache, uint32_t steps, uint32_t chunk) {
    Workspace& w = ctx->ws;
    for (uint32_t i = 0; i < steps; ++i) {
        const hipError_t e = cache_train_batch(ctx, cache, w.ct_samples + (size_t)i * chunk,
                                               reinterpret_cast<const float*>(w.ct_targets + (size_t)i * chunk),
                                               w.c_counts + kStepCounts + i, chunk, 1, i, steps);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// side: launch on ctx->side with the second spill area (the caller orders the streams).
hipError_t launch_mega(mrt_ctx* ctx, const Geometry& g, const SceneTable& T, ShadeParams P, const PathBuffers& cur,
                       Workspace& w, uint32_t bounce, uint32_t cap, const CacheIO& C, bool side = false,
                       bool queue_clear = false) {
    if (side && !g.trace.overflow_side) return hipErrorInvalidValue;
    hipStream_t stream = side ? ctx->side : ctx->stream;
    const FrameVariant& fv = frame_variant(ctx);
    const void* kernel = nullptr;
#define MRT_MEGA_FN(MB, ML) if (fv.mega_block == MB && fv.mega_lds == ML) kernel = (const void*)&k_paths_mega<MB, ML>;
    MRT_MEGA_FN(128, 16) MRT_MEGA_FN(64, 16) MRT_MEGA_FN(256, 8) MRT_MEGA_FN(128, 12)
#undef MRT_MEGA_FN
    if (!kernel) return hipErrorInvalidValue;
    if (fv.mega_lds < trace_min_lds_depth()) return hipErrorInvalidValue;  // the spill area assumes this depth
    int& per_cu = ctx->mega_occupancy[ctx->frame_variant >= 0 && ctx->frame_variant < kNumFrameVariants ? ctx->frame_variant : 0];
    if (per_cu == 0) {
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, fv.mega_block, 0);
        if (e != hipSuccess) { per_cu = 0; return e; }
        if (per_cu < 1) per_cu = 1;
    }
    // two lanes per path
    const uint32_t mb = (uint32_t)fv.mega_block;
    uint32_t blocks = (uint32_t)std::min<size_t>((size_t)per_cu * ctx->num_cus, (2 * (size_t)cap + mb - 1) / mb);
    blocks = (uint32_t)std::min<size_t>(blocks, g.trace.overflow_threads / mb);
    if (blocks == 0) return hipErrorInvalidValue;
    P.bounce = bounce;
    P.init = 0u;
    // the path queue of this stream's megakernel (a kernel clears it: graph capture replays that)
    uint32_t* queue = ctx->d_counters + (side ? kMegaQueueSide : kMegaQueue);
    if (!queue_clear) hipLaunchKernelGGL(k_clear_parts, dim3(1), dim3(64), 0, stream, queue);
    ProfScope p(ctx, "paths_mega", stream);
    const uint32_t* count = w.counts + bounce;
    uint32_t* ovf = side ? g.trace.overflow_side : g.trace.overflow;
    void* args[] = {&P, const_cast<SceneTable*>(&T), const_cast<PathBuffers*>(&cur), &w.shadow, &w.root, &count, &cap,
                    const_cast<WNode**>(&g.trace.nodes), const_cast<float4**>(&g.trace.tris), &ovf,
                    const_cast<float**>(&g.verts), const_cast<uint32_t**>(&g.material), const_cast<float**>(&g.albedo),
                    &w.out_L, const_cast<CacheIO*>(&C), &queue};
    hipError_t e = hipLaunchKernel(kernel, dim3(blocks), dim3(mb), args, 0, stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// One shading launch of the context's frame variant.
hipError_t launch_shade(mrt_ctx* ctx, size_t cap, ShadeParams& P, SceneTable& T, PathBuffers& cur, PathBuffers& nxt,
                        Workspace& w, const Geometry& g, CacheIO& C, const uint32_t* frame_dev) {
    const FrameVariant& fv = frame_variant(ctx);
    const void* kernel = nullptr;
    if (fv.shade_block == 1024) kernel = (const void*)&k_shade<1024>;
    if (fv.shade_block == 512) kernel = (const void*)&k_shade<512>;
    if (fv.shade_block == 256) kernel = (const void*)&k_shade<256>;
    if (!kernel) return hipErrorInvalidValue;
    ProfScope p(ctx, "shade");
    const mrt_hit* hit = w.hit;
    const uint32_t* occ = w.occ;
    void* args[] = {&P, &T, &cur, &nxt, &hit, &occ, &w.shadow, &w.root, const_cast<float**>(&g.verts),
                    const_cast<uint32_t**>(&g.material), const_cast<float**>(&g.albedo), &w.out_L, &w.counts, &C,
                    const_cast<uint32_t**>(&frame_dev)};
    const uint32_t sb = (uint32_t)fv.shade_block;
    hipError_t e = hipLaunchKernel(kernel, dim3((uint32_t)((cap + sb - 1) / sb)), dim3(sb), args, 0, ctx->stream);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t run_paths(mrt_ctx* ctx, const Geometry& g, const SceneTable& T, const mrt_render_opts& o,
                     uint32_t* frame_dev, CacheFrame* cf = nullptr) {
    const size_t cap = T.total_paths;
    Workspace& w = ctx->ws;
    hipError_t e = w.ensure_paths(cap);
    if (e != hipSuccess) return e;
    int c = 0;
    PathBuffers buf[2];
    for (int b = 0; b < 2; ++b) buf[b] = PathBuffers{w.ray[b], w.pid[b], w.rng[b], w.thr[b], w.L[b], w.pend[b]};
    const bool traced_primary = T.e[0].gb_hits == nullptr;
    bool all_gbuffer = true;
    for (uint32_t i = 0; i < T.n; ++i) all_gbuffer = all_gbuffer && T.e[i].gb_hits != nullptr;
    // Queues this frame's first traversal pair launch and its megakernel launches find cleared
    // (by k_frame_clear, when no launch uses them before): they then skip their own clear launch.
    bool pair_clear = false, mega_clear = false, mega_clear_side = false;
    if (all_gbuffer) {
        // every path starts from a G-buffer hit: the first shading launch builds the initial path
        // state itself (no 190 MB of state written and read back at 1080p); only the counters
        // are cleared here
        hipLaunchKernelGGL(k_frame_clear, dim3(1), dim3(64), 0, ctx->stream, w.counts, 64u, ctx->d_counters);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        pair_clear = mega_clear = mega_clear_side = true;
    } else {
        ProfScope p(ctx, "paths_init");
        hipLaunchKernelGGL(k_paths_init, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, T, o.sample_base, o.frame_seed,
                           frame_dev, buf[0], w.hit, w.root, w.counts, 64u);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (traced_primary) {
        if ((e = launch_trace(ctx, g.trace, buf[0].ray, (uint32_t)cap, w.hit, nullptr, 0, nullptr, w.counts, w.root)) !=
            hipSuccess)
            return e;
    }
    if (o.capture_bounce) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, ctx->stream, o.capture_count_dev, 1u);
    }
    CacheIO C;
    std::memset(&C, 0, sizeof(C));
    C.vertex = kCacheOff;
    // Vertex from which one megakernel launch finishes the live paths. By default the wavefront
    // runs vertices 0 and 1 (the large ray batches, where the persistent traversal and compaction
    // pay) and the megakernel the rest: past vertex 1 few paths are live, and each wavefront
    // trace launch over them would cost the latency of its slowest wave (~100 us) however few
    // rays it holds. A captured bounce is reached by the wavefront first.
    uint32_t mega_from = (o.flags & MRT_RENDER_MEGAKERNEL) ? 1u : 2u;
    if (o.capture_bounce) mega_from = std::max(mega_from, o.capture_bounce);
    if (cf) {
        C = cf->io;
        mega_from = C.vertex + 1;  // past the cache vertex (launched where the vertex is shaded, below)
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, ctx->stream, w.c_counts, 64u);
    }
    ShadeParams P;
    P.sky = rt::make_v3(o.sky[0], o.sky[1], o.sky[2]);
    P.sun_dir = rt::make_v3(o.sun_dir[0], o.sun_dir[1], o.sun_dir[2]);
    P.sun_e = rt::make_v3(o.sun_e[0], o.sun_e[1], o.sun_e[2]);
    P.nee_on = (o.sun_e[0] != 0.0f || o.sun_e[1] != 0.0f || o.sun_e[2] != 0.0f) ? 1u : 0u;
    std::memset(&P.skyd, 0, sizeof(P.skyd));
    P.sky_on = o.sky_lut && sky_device(o.sky_lut, &P.skyd) ? 1u : 0u;
    P.bounces = o.bounces;
    P.rr_start = o.rr_start;
    P.sample_base = o.sample_base;
    P.frame_seed = o.frame_seed;
    P.bounce = P.capture = P.init = 0u;
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    if ((e = hipStreamIsCapturing(ctx->stream, &capture)) != hipSuccess) return e;
    const bool fork = g.trace.overflow_side && capture == hipStreamCaptureStatusNone;
    for (uint32_t k = 0; k <= o.bounces; ++k) {
        P.bounce = k;
        P.init = (k == 0 && all_gbuffer) ? 1u : 0u;
        const bool capture_now = o.capture_bounce && k + 1 == o.capture_bounce;
        P.capture = capture_now ? 1u : 0u;
        if (cf && k == C.vertex && (C.train || (cf->limits & MRT_CACHE_LIMIT_QUERIES))) {
            // paths at the cache vertex, and the selection threshold of the training paths
            ProfScope p(ctx, "train_select");
            const mrt_hit* gb = P.init ? T.e[0].gb_hits : nullptr;
            hipLaunchKernelGGL(k_train_count, dim3((uint32_t)std::min<size_t>(blocks_for(cap), 2u * ctx->num_cus)),
                               dim3(256), 0, ctx->stream, w.hit, w.counts + k,
                               (uint32_t)cap, gb, T.e[0].npix, w.c_counts + 48);
            hipLaunchKernelGGL(k_train_threshold, dim3(1), dim3(1), 0, ctx->stream, w.c_counts + 48, cf->target);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if (cf && k == C.vertex && cf->wait_train) {
            if ((e = launch_deferred(ctx, cf, C)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(ctx->stream, ctx->ev_train, 0)) != hipSuccess) return e;
            cf->wait_train = false;
        }
        if (cf && k == C.vertex && cf->limits) {  // after the deferred steps: the loss is the last step's
            const float* parts = nullptr;
            uint32_t blocks = 0;
            cache_loss(cf->cache, &parts, &blocks);
            hipLaunchKernelGGL(k_cache_limits, dim3(1), dim3(256), 0, ctx->stream, w.c_counts + 48, cf->limits,
                               cf->query_budget, cf->loss_budget, parts, blocks, w.c_counts + 50);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if ((e = launch_shade(ctx, cap, P, const_cast<SceneTable&>(T), buf[c], buf[c ^ 1], w, g, C, frame_dev)) !=
            hipSuccess)
            return e;
        if (cf && k == C.vertex) {
            // Past the cache vertex only the training paths live (every path when the cache is not
            // finite): one megakernel launch finishes them, on the side stream beside the
            // inference when that is available. The queries of this bounce are inferred and the
            // results added to the paths' radiance.
            const bool more = !capture_now && k < o.bounces;
            const bool side = more && fork;
            if (side) {
                if ((e = hipEventRecord(ctx->ev_fork, ctx->stream)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0)) != hipSuccess) return e;
                if ((e = launch_mega(ctx, g, T, P, buf[c ^ 1], w, k + 1, (uint32_t)cap, C, true, mega_clear_side)) !=
                    hipSuccess)
                    return e;
                mega_clear_side = false;
                if ((e = hipEventRecord(ctx->ev_join, ctx->side)) != hipSuccess) return e;
            }
            const float* y = nullptr;
            if ((e = cache_query_batch(ctx, cf->cache, w.cq, C.q_count, (uint32_t)cap, &y)) != hipSuccess) return e;
            {
                ProfScope p(ctx, "cache_apply");
                hipLaunchKernelGGL(k_cache_apply, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, C,
                                   reinterpret_cast<const float4*>(y), w.out_L, cf->cache_rgb, cf->npix);
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
            if (more && !side) {
                if ((e = launch_mega(ctx, g, T, P, buf[c ^ 1], w, k + 1, (uint32_t)cap, C, false, mega_clear)) !=
                    hipSuccess)
                    return e;
                mega_clear = false;
            }
            if (side && (e = hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0)) != hipSuccess) return e;
            if (more) break;
        }
        if (capture_now) {
            hipLaunchKernelGGL(k_capture, dim3(blocks_for(cap)), dim3(256), 0, ctx->stream, buf[c ^ 1], w.counts + k + 1,
                               o.capture_rays_dev, o.capture_pid_dev, o.capture_count_dev);
            return hipGetLastError();
        }
        if (k == o.bounces) break;
        if (k + 1 == mega_from) {  // the rest of every live path in one launch
            if ((e = launch_mega(ctx, g, T, P, buf[c ^ 1], w, k + 1, (uint32_t)cap, C, false, mega_clear)) != hipSuccess)
                return e;
            mega_clear = false;
            break;
        }
        c ^= 1;
        // The shadow rays and the extension rays of vertex k are independent. They are traced in
        // one launch (64-item blocks alternating between the two kinds over the same slots), so
        // the GPU sees one tail of slow waves instead of two.
        uint32_t* order = nullptr;
        if (o.flags & MRT_RENDER_SORT_RAYS) {
            if ((e = launch_ray_sort(ctx, g.bounds, buf[c].ray, (uint32_t)cap, w.counts + k + 1, &order)) != hipSuccess)
                return e;
        }
        if ((e = launch_trace_pair(ctx, g.trace, w.shadow, w.occ, buf[c].ray, w.hit, order, (uint32_t)cap,
                                   w.counts + k + 1, w.root, pair_clear)) != hipSuccess)
            return e;
        pair_clear = false;
    }
    {
        ProfScope p(ctx, "resolve");
        hipLaunchKernelGGL(k_resolve, dim3(blocks_for(T.total_pix)), dim3(256), 0, ctx->stream, T, w.out_L, frame_dev);
    }
    if (cf && C.train) {  // training batch of the frame, then the optimizer steps
        {
            ProfScope p(ctx, "cache_targets");
            hipLaunchKernelGGL(k_cache_targets, dim3((uint32_t)((cap + kShadeBlock - 1) / kShadeBlock)), dim3(kShadeBlock), 0,
                               ctx->stream, C, w.ct_samples,
                               w.ct_targets, w.c_counts + 1, cf->steps, cf->chunk, cf->ex_s,
                               reinterpret_cast<float4*>(cf->ex_t), cf->ex_pid);
            hipLaunchKernelGGL(k_split_counts, dim3(1), dim3(64), 0, ctx->stream, w.c_counts + 1, cf->steps, cf->chunk,
                               w.c_counts + kStepCounts, cf->ex_count);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        if (cf->steps && cf->defer) {
            ctx->deferred.cache = cf->cache;
            ctx->deferred.steps = cf->steps;
            ctx->deferred.chunk = cf->chunk;
            if ((e = hipEventRecord(ctx->ev_pending, ctx->stream)) != hipSuccess) return e;
        } else if ((e = train_steps(ctx, cf->cache, cf->steps, cf->chunk)) != hipSuccess) {
            return e;
        }
    }
    return hipGetLastError();
}

void fill_entry(SceneEntry& e, const mrt_camera& cam, uint32_t path_base, uint32_t pix_base, uint32_t tri_base,
                uint32_t root, float* rgb, const mrt_hit* gb) {
    std::memset(&e, 0, sizeof(e));
    e.path_base = path_base;
    e.npix = cam.width * cam.height;
    e.tri_base = tri_base;
    e.root = root;
    e.pix_base = pix_base;
    e.cam = camera_basis(cam);
    e.rgb = rgb;
    e.gb_hits = gb;
}

}  // namespace

// ---- single scene (Pathtracer::secondaryRays) ---------------------------------------------
hipError_t run_secondary(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_gbuffer& gb,
                         const mrt_render_opts& o, float* rgb) {
    SceneTable T;
    std::memset(&T, 0, sizeof(T));
    T.n = 1;
    T.spp = o.spp;
    T.total_pix = cam.width * cam.height;
    T.total_paths = T.total_pix * o.spp;
    fill_entry(T.e[0], cam, 0, 0, 0, 0, rgb, gb.hits_dev);
    Geometry g{TraceTarget{s->d_nodes, s->d_tris, s->d_overflow, s->overflow_entries, s->overflow_threads,
                           s->d_overflow + s->overflow_entries * s->overflow_threads * 2}, s->d_verts,
               s->d_material, s->d_albedo, s->bounds};
    return run_paths(ctx, g, T, o, nullptr);
}

// ---- cache-terminated frames ------------------------------------------------------------------
hipError_t flush_training(mrt_ctx* ctx, const mrt_cache* cache) {
    if (!ctx->deferred.cache || (cache && ctx->deferred.cache != cache)) return hipSuccess;
    mrt_cache* c = ctx->deferred.cache;
    ctx->deferred.cache = nullptr;
    return train_steps(ctx, c, ctx->deferred.steps, ctx->deferred.chunk);
}

void drop_training(mrt_ctx* ctx, const mrt_cache* cache) {
    if (ctx->deferred.cache == cache) ctx->deferred.cache = nullptr;
}

hipError_t run_cached(mrt_ctx* ctx, const mrt_scene* s, mrt_cache* cache, const mrt_camera& cam, const mrt_gbuffer& gb,
                      const mrt_render_opts& o, const mrt_cache_render_opts& co, float* rgb) {
    SceneTable T;
    std::memset(&T, 0, sizeof(T));
    T.n = 1;
    T.spp = o.spp;
    T.total_pix = cam.width * cam.height;
    T.total_paths = T.total_pix * o.spp;
    fill_entry(T.e[0], cam, 0, 0, 0, 0, rgb, gb.hits_dev);
    Geometry g{TraceTarget{s->d_nodes, s->d_tris, s->d_overflow, s->overflow_entries, s->overflow_threads,
                           s->d_overflow + s->overflow_entries * s->overflow_threads * 2}, s->d_verts,
               s->d_material, s->d_albedo, s->bounds};
    const size_t cap = T.total_paths;
    const uint32_t ntrain = std::min<uint32_t>(co.train_samples, T.total_paths);
    CacheFrame cf;
    std::memset(&cf, 0, sizeof(cf));
    cf.cache = cache;
    cf.io.vertex = co.vertex;
    cf.io.train = ntrain ? 1u : 0u;
    cf.io.tseed = rt::pcg_hash(o.frame_seed ^ rt::pcg_hash(o.sample_base ^ 0x7A1Eu));
    cf.target = ntrain;
    // rows of the training batch: the expected count plus a margin of many standard deviations
    const size_t rows = ntrain ? std::min<size_t>(T.total_paths, train_capacity(ntrain)) : 1;
    cf.steps = ntrain ? std::max<uint32_t>(1u, co.train_steps) : 0u;
    cf.chunk = ntrain ? (uint32_t)(((rows + cf.steps - 1) / cf.steps + 31) & ~(size_t)31) : 0u;
    cf.cache_rgb = co.cache_rgb_dev;
    cf.npix = T.total_pix;
    cf.ex_s = co.train_samples_dev;
    cf.ex_t = co.train_targets_dev;
    cf.ex_pid = co.train_pid_dev;
    cf.ex_count = co.train_count_dev;
    Workspace& w = ctx->ws;
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(ctx->stream, &capture);
    if (e != hipSuccess) return e;
    // Steps left by the previous cached frame run beside this frame's traversal when they belong
    // to this cache and their batch survives the workspace sizing below; otherwise first, in order.
    const size_t trows = (size_t)cf.chunk * std::max<uint32_t>(cf.steps, 1u);
    const bool keeps = cap <= w.cq_cap && (ntrain ? cap : 1) <= w.ct_cap && trows <= w.ct_rows;
    const bool side_ok = ctx->side && capture == hipStreamCaptureStatusNone;
    if (ctx->deferred.cache && !(ctx->deferred.cache == cache && keeps && side_ok) &&
        (e = flush_training(ctx, nullptr)) != hipSuccess)
        return e;
    cf.defer = co.overlap_training && side_ok && cf.steps > 0;
    e = w.ensure_cache(cap, ntrain ? cap : 1, trows);
    if (e == hipSuccess) e = cache_reserve(cache, cap, cf.chunk);
    if (e != hipSuccess) return e;
    // This cache's steps left by the previous frame go to the second stream once this frame's work
    // up to the cache vertex is queued (run_paths): they start at once, beside it.
    cf.wait_train = ctx->deferred.cache != nullptr;
    if (co.train_count_dev && !ntrain)
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, ctx->stream, co.train_count_dev, 1u);
    cf.io.q = w.cq;
    cf.io.q_pid = w.cq_pid;
    cf.io.q_L = w.cq_L;
    cf.io.q_thr = w.cq_thr;
    cf.io.q_count = w.c_counts;
    cf.io.train_L = w.ct_L;
    cf.io.tsel = w.c_counts + 49;
    cf.io.limit = w.c_counts + 50;  // cleared with the frame's counters; set by k_cache_limits
    cf.limits = co.limits & (MRT_CACHE_LIMIT_QUERIES | MRT_CACHE_LIMIT_LOSS);
    cf.query_budget = co.query_budget;
    cf.loss_budget = co.loss_budget;
    // With steps still to run, the flag is taken after they are queued (launch_deferred; no kernel
    // before the cache vertex reads it).
    if (!cf.wait_train && (e = cache_check(ctx, cache, &cf.io.bad)) != hipSuccess) return e;
    e = run_paths(ctx, g, T, o, nullptr, &cf);
    // A frame that returned before its cache vertex (capture) has not launched the steps yet.
    if (e == hipSuccess && cf.wait_train) e = flush_training(ctx, cache);
    return e;
}

// ---- batches ---------------------------------------------------------------------------------
struct mrt_batch {
    uint32_t n = 0;
    std::vector<uint32_t> root, tri_base;
    WNode* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float* d_verts = nullptr;
    uint32_t* d_material = nullptr;
    float* d_albedo = nullptr;
    uint32_t* d_overflow = nullptr;
    size_t overflow_entries = 0, overflow_threads = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};
    uint64_t device_bytes = 0;
    uint64_t total_tris = 0, total_nodes = 0;
};

struct mrt_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* d_frame = nullptr;
    std::vector<float*> outputs;
};

static void free_batch(mrt_batch* b) {
    if (!b) return;
    hipFree(b->d_nodes); hipFree(b->d_tris); hipFree(b->d_verts); hipFree(b->d_material); hipFree(b->d_albedo);
    hipFree(b->d_overflow);
    delete b;
}

static mrt_status build_table(mrt_ctx* ctx, const mrt_batch* b, const mrt_batch_view* views, uint32_t n_views,
                              const mrt_render_opts* o, SceneTable& T) {
    if (!views || n_views != b->n) return api_fail(ctx, MRT_ERR_INVALID, "batch: %u views for %u scenes", n_views, b->n);
    if (!o || o->spp == 0 || o->spp > 4096) return api_fail(ctx, MRT_ERR_INVALID, "batch: spp must be in 1..4096");
    if (o->rr_start < 2 || o->bounces > 60) return api_fail(ctx, MRT_ERR_INVALID, "batch: bad bounces / rr_start");
    if (o->capture_bounce) return api_fail(ctx, MRT_ERR_UNSUPPORTED, "batch: capture is single-scene only");
    SkyDev sd;
    if (o->sky_lut && !sky_device(o->sky_lut, &sd))
        return api_fail(ctx, MRT_ERR_STATE, "batch: sky_lut has no view yet (mrt_sky_update_view)");
    std::memset(&T, 0, sizeof(T));
    T.n = b->n;
    T.spp = o->spp;
    uint64_t paths = 0, pix = 0;
    for (uint32_t j = 0; j < b->n; ++j) {
        const mrt_camera& c = views[j].camera;
        if (c.width == 0 || c.height == 0 || !views[j].rgb_dev)
            return api_fail(ctx, MRT_ERR_INVALID, "batch: view %u has no viewport or output", j);
        fill_entry(T.e[j], c, (uint32_t)paths, (uint32_t)pix, b->tri_base[j], b->root[j], views[j].rgb_dev, nullptr);
        pix += (uint64_t)c.width * c.height;
        paths += (uint64_t)c.width * c.height * o->spp;
    }
    if (paths > (1ull << 28)) return api_fail(ctx, MRT_ERR_INVALID, "batch: %llu paths exceed 2^28",
                                              (unsigned long long)paths);
    T.total_paths = (uint32_t)paths;
    T.total_pix = (uint32_t)pix;
    return MRT_OK;
}

static Geometry batch_geometry(const mrt_batch* b) {
    return Geometry{TraceTarget{b->d_nodes, b->d_tris, b->d_overflow, b->overflow_entries, b->overflow_threads,
                                b->d_overflow + b->overflow_entries * b->overflow_threads * 2},
                    b->d_verts, b->d_material, b->d_albedo, b->bounds};
}

extern "C" {

uint32_t mrt_frame_variant_count(void) { return (uint32_t)kNumFrameVariants; }

const char* mrt_frame_variant_name(uint32_t v) { return v < (uint32_t)kNumFrameVariants ? kFrameVariants[v].name : nullptr; }

mrt_status mrt_frame_variant_set(mrt_ctx* ctx, uint32_t v) {
    if (!ctx) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_frame_variant_set: null context");
    if (v >= (uint32_t)kNumFrameVariants)
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_frame_variant_set: variant %u of %d", v, kNumFrameVariants);
    ctx->frame_variant = (int)v;
    return MRT_OK;
}

mrt_status mrt_frame_variant_get(mrt_ctx* ctx, uint32_t* v) {
    if (!ctx || !v) return api_fail(ctx, MRT_ERR_INVALID, "mrt_frame_variant_get: null argument");
    *v = (uint32_t)ctx->frame_variant;
    return MRT_OK;
}

mrt_status mrt_batch_create(mrt_ctx* ctx, const mrt_scene* const* scenes, uint32_t n, mrt_batch** out) {
    if (!ctx || !scenes || !out) return api_fail(ctx, MRT_ERR_INVALID, "mrt_batch_create: null argument");
    *out = nullptr;
    if (n == 0 || n > kMaxScenes) return api_fail(ctx, MRT_ERR_INVALID, "mrt_batch_create: %u scenes not in 1..%u", n,
                                                  kMaxScenes);
    uint64_t nodes = 0, tris = 0, mats = 0;
    size_t ovf = 1;
    for (uint32_t j = 0; j < n; ++j) {
        if (!scenes[j]) return api_fail(ctx, MRT_ERR_INVALID, "mrt_batch_create: scene %u is null", j);
        nodes += scenes[j]->n_nodes;
        tris += scenes[j]->n_tris;
        mats += scenes[j]->n_materials;
        ovf = std::max(ovf, scenes[j]->overflow_entries);
    }
    if (mrt_check_size(tris, nodes) != MRT_OK)
        return api_fail(ctx, MRT_ERR_INVALID, "mrt_batch_create: the concatenated batch: %s", mrt_last_error(nullptr));
    auto* b = new (std::nothrow) mrt_batch();
    if (!b) return api_fail(ctx, MRT_ERR_OOM, "mrt_batch_create: host allocation failed");
    b->n = n;
    b->overflow_entries = ovf;
    b->overflow_threads = (size_t)ctx->num_cus * 2048;
    b->total_tris = tris;
    b->total_nodes = nodes;
    hipError_t e = hipSuccess;
    auto A = [&](auto** p, size_t bytes) {
        if (e == hipSuccess) { e = hipMalloc((void**)p, bytes ? bytes : 256); b->device_bytes += bytes; }
    };
    A(&b->d_nodes, nodes * sizeof(WNode));
    A(&b->d_tris, tris * 48);
    A(&b->d_verts, tris * 36);
    A(&b->d_material, tris * 4);
    A(&b->d_albedo, mats * 12);
    A(&b->d_overflow, 2 * ovf * b->overflow_threads * 8);  // two halves: main and side stream
    for (int k = 0; k < 3; ++k) { b->bounds[k] = 3.4e38f; b->bounds[3 + k] = -3.4e38f; }
    uint64_t nb = 0, tb = 0, mb = 0;
    for (uint32_t j = 0; j < n && e == hipSuccess; ++j) {
        const mrt_scene* s = scenes[j];
        b->root.push_back((uint32_t)nb);
        b->tri_base.push_back((uint32_t)tb);
        for (int k = 0; k < 3; ++k) {
            b->bounds[k] = std::min(b->bounds[k], s->bounds[k]);
            b->bounds[3 + k] = std::max(b->bounds[3 + k], s->bounds[3 + k]);
        }
        auto cp = [&](void* dst, const void* src, size_t bytes) {
            if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
        };
        cp(b->d_nodes + nb, s->d_nodes, (size_t)s->n_nodes * sizeof(WNode));
        cp(b->d_tris + tb * 3, s->d_tris, (size_t)s->n_tris * 48);
        cp(b->d_verts + tb * 9, s->d_verts, (size_t)s->n_tris * 36);
        cp(b->d_material + tb, s->d_material, (size_t)s->n_tris * 4);
        cp(b->d_albedo + mb * 3, s->d_albedo, (size_t)s->n_materials * 12);
        if (e == hipSuccess && s->n_nodes) {
            hipLaunchKernelGGL(k_rebase_nodes, dim3(blocks_for((size_t)s->n_nodes * 4)), dim3(256), 0, ctx->stream,
                               b->d_nodes + nb, s->n_nodes, (uint32_t)nb, (uint32_t)tb);
            e = hipGetLastError();
        }
        if (e == hipSuccess && s->n_tris && mb) {
            hipLaunchKernelGGL(k_add_u32, dim3(blocks_for(s->n_tris)), dim3(256), 0, ctx->stream, b->d_material + tb,
                               s->n_tris, (uint32_t)mb);
            e = hipGetLastError();
        }
        nb += s->n_nodes;
        tb += s->n_tris;
        mb += s->n_materials;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_batch(b);
        return api_fail(ctx, e == hipErrorOutOfMemory ? MRT_ERR_OOM : MRT_ERR_HIP, "mrt_batch_create: %s",
                        hipGetErrorString(e));
    }
    *out = b;
    return MRT_OK;
}

mrt_status mrt_batch_destroy(mrt_ctx* ctx, mrt_batch* b) {
    if (!ctx) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_batch_destroy: null context");
    if (!b) return MRT_OK;
    hipStreamSynchronize(ctx->stream);
    free_batch(b);
    return MRT_OK;
}

mrt_status mrt_batch_info(mrt_ctx* ctx, const mrt_batch* b, uint64_t info[4]) {
    if (!ctx || !b || !info) return api_fail(ctx, MRT_ERR_INVALID, "mrt_batch_info: null argument");
    info[0] = b->n;
    info[1] = b->total_tris;
    info[2] = b->total_nodes;
    info[3] = b->device_bytes;
    return MRT_OK;
}

mrt_status mrt_batch_render(mrt_ctx* ctx, mrt_batch* b, const mrt_batch_view* views, uint32_t n_views,
                            const mrt_render_opts* o) {
    if (!ctx || !b) return api_fail(ctx, MRT_ERR_INVALID, "mrt_batch_render: null context or batch");
    SceneTable T;
    mrt_status st = build_table(ctx, b, views, n_views, o, T);
    if (st != MRT_OK) return st;
    hipError_t e = run_paths(ctx, batch_geometry(b), T, *o, nullptr);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_batch_render: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_graph_create(mrt_ctx* ctx, mrt_batch* b, const mrt_batch_view* views, uint32_t n_views,
                            const mrt_render_opts* o, mrt_graph** out) {
    if (!ctx || !b || !out) return api_fail(ctx, MRT_ERR_INVALID, "mrt_graph_create: null argument");
    *out = nullptr;
    SceneTable T;
    mrt_status st = build_table(ctx, b, views, n_views, o, T);
    if (st != MRT_OK) return st;
    if (o->flags & MRT_RENDER_SORT_RAYS) return api_fail(ctx, MRT_ERR_UNSUPPORTED, "mrt_graph_create: sorting not captured");
    auto* g = new (std::nothrow) mrt_graph();
    if (!g) return api_fail(ctx, MRT_ERR_OOM, "mrt_graph_create: host allocation failed");
    hipError_t e = ctx->ws.ensure_paths(T.total_paths);
    if (e == hipSuccess) e = hipMalloc((void**)&g->d_frame, 256);
    if (e == hipSuccess) e = hipMemsetAsync(g->d_frame, 0, 256, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    const bool prof = ctx->profiling;
    ctx->profiling = false;  // no event records inside the captured graph
    if (e == hipSuccess) e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t er = run_paths(ctx, batch_geometry(b), T, *o, g->d_frame);
        hipError_t ec = hipStreamEndCapture(ctx->stream, &g->graph);
        e = er != hipSuccess ? er : ec;
    }
    ctx->profiling = prof;
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        hipFree(g->d_frame);
        delete g;
        return api_fail(ctx, MRT_ERR_HIP, "mrt_graph_create: %s", hipGetErrorString(e));
    }
    *out = g;
    return MRT_OK;
}

mrt_status mrt_graph_set_frame(mrt_ctx* ctx, mrt_graph* g, uint32_t frame) {
    if (!ctx || !g) return api_fail(ctx, MRT_ERR_INVALID, "mrt_graph_set_frame: null argument");
    hipError_t e = hipMemcpyAsync(g->d_frame, &frame, 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_graph_set_frame: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_graph_launch(mrt_ctx* ctx, mrt_graph* g) {
    if (!ctx || !g) return api_fail(ctx, MRT_ERR_INVALID, "mrt_graph_launch: null argument");
    ProfScope p(ctx, "graph_frame");
    hipError_t e = hipGraphLaunch(g->exec, ctx->stream);
    if (e != hipSuccess) return api_fail(ctx, MRT_ERR_HIP, "mrt_graph_launch: %s", hipGetErrorString(e));
    return MRT_OK;
}

mrt_status mrt_graph_destroy(mrt_ctx* ctx, mrt_graph* g) {
    if (!ctx) return api_fail(nullptr, MRT_ERR_INVALID, "mrt_graph_destroy: null context");
    if (!g) return MRT_OK;
    hipStreamSynchronize(ctx->stream);
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    hipFree(g->d_frame);
    delete g;
    return MRT_OK;
}

}  // extern "C"
