// internal.h — private structures of the minotert library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>
#include <cmath>

#include "../../include/minotert.h"
#include "rt_math.h"

// ---- wide BVH layout ----------------------------------------------------------------------
// One BVH4 node is exactly 128 bytes (one aligned 128-B line): the six child-box planes as
// float[4] rows (SoA over the 4 children, so each row is one dwordx4 load), the four child
// words and four spare words.
//   child word: internal child   -> wide node index (bit 31 clear)
//               leaf child       -> 0x80000000 | first_tri << 3 | (count - 1), count 1..8,
//                                   first_tri indexes the leaf-ordered triangle array
//               empty slot       -> 0xFFFFFFFF
struct alignas(128) WNode {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t meta[4];
};
static_assert(sizeof(WNode) == 128, "wide node must be one 128-byte line");
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMissId = 0xFFFFFFFFu;
constexpr int kMaxFrameVariants = 16;

// Leaf triangle record: 48 bytes = 3 x float4 {v0, id-bits}, {v1, 0}, {v2, 0}.
struct HostBVH {
    std::vector<WNode> nodes;
    std::vector<float> tri4;        // leaf-ordered triangles, 12 floats each
    uint32_t depth = 0;             // wide tree depth (root = 1)
    uint32_t leaves = 0;
    uint32_t binary_nodes = 0;
    double build_ms = 0.0;
};

// Host builders (bvh_host.cpp).
void build_bvh4(const float* verts, uint32_t ntri, uint32_t max_leaf, uint32_t bins, HostBVH& out);

// Device builder (bvh_build.hip): the same tree from triangles on the device; the node and
// triangle arrays are exact-size device allocations owned by the caller.
struct DeviceBVH {
    WNode* d_nodes = nullptr;
    float4* d_tris = nullptr;
    uint32_t n_nodes = 0, depth = 0, leaves = 0, binary_nodes = 0;
    double build_ms = 0.0;
};
struct mrt_ctx;
hipError_t build_bvh4_device(mrt_ctx* ctx, const float* d_verts, uint32_t ntri, uint32_t max_leaf, uint32_t bins,
                             DeviceBVH& out);

// Camera basis (scene_host.cpp).
rt::CamBasis camera_basis(const mrt_camera& cam);

// Device workspace owned by a context: wavefront path queues (double-buffered SoA state)
// and per-pixel scratch, grown on demand.
struct Workspace {
    size_t path_cap = 0;
    mrt_ray* ray[2] = {nullptr, nullptr};
    uint32_t* pid[2] = {nullptr, nullptr};
    uint32_t* rng[2] = {nullptr, nullptr};
    float4* thr[2] = {nullptr, nullptr};
    float4* L[2] = {nullptr, nullptr};
    float4* pend[2] = {nullptr, nullptr};
    mrt_hit* hit = nullptr;
    mrt_ray* shadow = nullptr;
    uint32_t* occ = nullptr;
    float4* out_L = nullptr;
    uint32_t* root = nullptr;       // per queue slot: root wide node of the path's scene
    uint32_t* counts = nullptr;
    size_t pix_cap = 0;
    mrt_ray* cam_rays = nullptr;
    mrt_hit* gb_hits = nullptr;
    float4* contrib = nullptr;
    uint32_t* pix_occ = nullptr;
    hipError_t ensure_paths(size_t cap);
    void release_paths();
    hipError_t ensure_pixels(size_t cap);
    void release_pixels();
    // cache-terminated rendering (paths.hip)
    size_t cq_cap = 0, ct_cap = 0, ct_rows = 0;
    mrt_cache_sample* cq = nullptr;        // queries at the terminating vertex
    uint32_t* cq_pid = nullptr;
    float4* cq_L = nullptr;                // radiance gathered before the vertex
    float4* cq_thr = nullptr;              // throughput arriving at the vertex
    float4* ct_L = nullptr;                // final radiance of the training paths, by path id
    mrt_cache_sample* ct_samples = nullptr;
    float4* ct_targets = nullptr;
    uint32_t* c_counts = nullptr;          // [0] queries, [1] training samples, [48] paths at the cache
                                           // vertex, [49] training selection threshold, [50] budget
                                           // exceeded (cleared per frame);
                                           // [kStepCounts + i] the size of training step i's share, kept
                                           // until the steps run (possibly in the next cached frame)
    hipError_t ensure_cache(size_t queries, size_t train_paths, size_t train_rows);
    void release_cache();
};

// Radix-sort scratch owned by a context (sort.hip).
struct SortSpace {
    size_t cap = 0;
    uint32_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;
    hipError_t ensure(size_t cap_items);
    void release();
};

struct ProfEvent { std::string name; hipEvent_t a, b; };
struct ProfAccum { double ms = 0.0; uint64_t launches = 0; };

struct mrt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;              // second stream for a trace that overlaps one on `stream`
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_train = nullptr;           // end of deferred training steps on `side`
    hipEvent_t ev_pending = nullptr;         // end of the frame that formed the deferred steps' batch
    // Training steps of a cached frame rendered with overlap_training, not launched yet: they run
    // on `side` at the start of the next cached frame with the same cache, or on `stream` when
    // anything else reads or changes the cache or the training batch (flush_training).
    struct { mrt_cache* cache = nullptr; uint32_t steps = 0, chunk = 0; } deferred;
    std::string last_error;
    int num_cus = 0;
    bool profiling = false;
    std::vector<ProfEvent> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, ProfAccum> prof;
    std::map<std::string, hipEvent_t> region_open;  // start events of open mrt_profile_region regions
    uint32_t* d_counters = nullptr;          // persistent-kernel work queues (self-resetting): the
                                             // traversal's at [0, 1024), the megakernel's at
                                             // kMegaQueue (main stream) and kMegaQueueSide
    double* d_partials = nullptr;            // 1024 partial sums (relative MSE)
    int trace_variant_closest = 0;           // index into the traversal variant table
    int trace_variant_any = 0;
    int trace_occupancy[3][64] = {};         // blocks per CU of each traversal variant (0 = not queried):
                                             // closest, any, mixed
    int frame_variant = 0;                   // shading block and megakernel shape (paths.hip)
    int mega_occupancy[kMaxFrameVariants] = {};  // blocks per CU of the megakernel, by frame variant
                                             // forms (0 = not queried)
    Workspace ws;
    SortSpace sort;
    uint32_t* noalloc_scratch = nullptr;  // this device's scratch array of the allocation-free entry point
};

struct mrt_scene {
    uint32_t n_tris = 0, n_nodes = 0, n_materials = 0, depth = 0, leaves = 0;
    WNode* d_nodes = nullptr;
    float4* d_tris = nullptr;           // leaf order, 3 float4 per triangle
    float* d_verts = nullptr;           // original order, 9 floats per triangle
    uint32_t* d_material = nullptr;     // per triangle
    float* d_albedo = nullptr;          // per material, 3 floats
    uint32_t* d_overflow = nullptr;     // traversal stack spill area: two halves (main and side stream)
    size_t overflow_entries = 0;        // per thread
    size_t overflow_threads = 0;        // per half
    double build_ms = 0.0;
    size_t device_bytes = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};  // scene AABB lo.xyz, hi.xyz
};

// Profiling scope: records HIP events around a launch on the context stream (or the given
// stream) when enabled.
struct ProfScope {
    mrt_ctx* ctx; const char* name; hipStream_t stream = nullptr; hipEvent_t a = nullptr, b = nullptr;
    ProfScope(mrt_ctx* c, const char* n, hipStream_t s = nullptr);
    ~ProfScope();
};

// Kernel launchers (trace.hip).
// Geometry a traversal runs over (one scene, or the concatenated BVHs of a batch).
struct TraceTarget {
    const WNode* nodes;
    const float4* tris;
    uint32_t* overflow;
    size_t overflow_entries;
    size_t overflow_threads;
    uint32_t* overflow_side;  // spill area of a launch on ctx->side (null: no side launches)
};
// ray_root (optional): per-ray root wide node (batched scenes); default root is node 0.
// cam (optional, closest hit, no perm): no ray array; ray i is pixel i's camera ray.
hipError_t launch_trace(mrt_ctx* ctx, const TraceTarget& t, const mrt_ray* rays, uint32_t n,
                        mrt_hit* hits, uint32_t* occluded, int any, const uint32_t* perm,
                        const uint32_t* count_dev, const uint32_t* ray_root, const rt::CamBasis* cam = nullptr);
// One launch over the any-hit rays `shadow` (results in `occluded`) and the closest-hit rays
// `rays` (results in `hits`, optional processing order `perm`), both of device-side count
// *count_dev (<= n) and indexed by the same slots. queue_clear: the caller has zeroed the
// launch's queue counters (d_counters word 0 on) since the last launch that used them.
hipError_t launch_trace_pair(mrt_ctx* ctx, const TraceTarget& t, const mrt_ray* shadow, uint32_t* occluded,
                             const mrt_ray* rays, mrt_hit* hits, const uint32_t* perm, uint32_t n,
                             const uint32_t* count_dev, const uint32_t* ray_root, bool queue_clear = false);
// perm (optional): item i traces ray perm[i] and writes its result at perm[i].
// count_dev (optional): device-side number of items (<= n), read by the kernel.
hipError_t launch_trace(mrt_ctx* ctx, const mrt_scene* s, const mrt_ray* rays, uint32_t n,
                        mrt_hit* hits, uint32_t* occluded, int any, const uint32_t* perm = nullptr,
                        const uint32_t* count_dev = nullptr);
// Closest hits of the camera rays of every pixel (hits[i] for pixel i), without a ray array.
hipError_t launch_trace_camera(mrt_ctx* ctx, const mrt_scene* s, const rt::CamBasis& cam, mrt_hit* hits);

// Traversal variant table (trace.hip).
int trace_variant_count();
const char* trace_variant_name(int v);
int trace_min_lds_depth();

// Render stages (render.hip).
hipError_t run_primary(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_camera& prev,
                       const mrt_gbuffer& gb);
hipError_t run_secondary(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_gbuffer& gb,
                         const mrt_render_opts& o, float* rgb);
// Ray reordering and key/value radix sort (sort.hip). launch_ray_sort leaves the order in
// context scratch and returns a pointer to it.
hipError_t launch_ray_sort(mrt_ctx* ctx, const float bounds[6], const mrt_ray* rays, uint32_t cap,
                           const uint32_t* count_dev, uint32_t** order_out);
// The allocation-free form: its scratch is a file-scope device array of sort.hip, whose address
// on the current device noalloc_scratch_init looks up (once per device; mrt_create stores it in
// the context). At most noalloc_max_rays() rays; the order goes straight into `order`.
hipError_t noalloc_scratch_init(int device, uint32_t** out);
uint32_t noalloc_max_rays();
hipError_t launch_ray_sort_noalloc(mrt_ctx* ctx, const float bounds[6], const mrt_ray* rays, uint32_t n,
                                   uint32_t* order);

// Records an error message on the context (and the thread) and returns the status (api.cpp).
mrt_status api_fail(mrt_ctx* ctx, mrt_status st, const char* fmt, ...);
hipError_t launch_sort_pairs(mrt_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t bits);

// Radiance cache (cache.hip): device-count driven batches for the path tracer.
hipError_t cache_reserve(mrt_cache* c, size_t query_cap, size_t train_cap);
hipError_t cache_query_batch(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* qs, const uint32_t* count,
                             uint32_t nmax, const float** y);
// group_pos, group_len: with apply, the step's place in a group of consecutive steps whose table
// average is applied in closed form (one group per frame; 0, 1 for a step of its own).
hipError_t cache_train_batch(mrt_ctx* ctx, mrt_cache* c, const mrt_cache_sample* qs, const float* target,
                             const uint32_t* count, uint32_t nmax, int apply, uint32_t group_pos = 0,
                             uint32_t group_len = 1);
hipError_t cache_check(mrt_ctx* ctx, mrt_cache* c, const uint32_t** flag);
// The last training pass's loss: per-block sums parts[0..blocks) and its sample count at
// parts[kLossCountSlot] (uint32 bits); blocks = 0 before the first pass.
constexpr uint32_t kLossCountSlot = 1024;
void cache_loss(const mrt_cache* c, const float** parts, uint32_t* blocks);
// Cache-terminated path tracing (paths.hip).
// d_counters (2048 words): a traversal launch's queue has one counter per partition, kQueueStride
// words apart (closest-hit or mixed launches at word 0, any-hit launches at kMaxParts *
// kQueueStride); the megakernel's path queues, partitioned the same way, at kMegaQueue (main
// stream) and kMegaQueueSide.
constexpr uint32_t kQueueStride = 64;  // 256 bytes between partition counters
constexpr uint32_t kMaxParts = 8;
constexpr uint32_t kMegaQueue = 2 * kMaxParts * kQueueStride, kMegaQueueSide = 3 * kMaxParts * kQueueStride;
static_assert(kMegaQueueSide + kMaxParts * kQueueStride <= 2048, "d_counters layout");
constexpr uint32_t kStepCounts = 64;  // first per-step share size in Workspace::c_counts
// Launches deferred training steps on ctx->stream when they belong to `cache` (any cache when
// null); drop_training forgets those of a cache being destroyed.
hipError_t flush_training(mrt_ctx* ctx, const mrt_cache* cache);
void drop_training(mrt_ctx* ctx, const mrt_cache* cache);
hipError_t run_cached(mrt_ctx* ctx, const mrt_scene* s, mrt_cache* cache, const mrt_camera& cam,
                      const mrt_gbuffer& gb, const mrt_render_opts& o, const mrt_cache_render_opts& co, float* rgb);
// Progressive accumulation and relative MSE (util.hip).
hipError_t launch_accumulate(mrt_ctx* ctx, float* sum, const float* rgb, size_t n, uint32_t frames, float* mean);
hipError_t launch_rel_mse(mrt_ctx* ctx, const float* img, const float* ref, size_t n, double eps, double* out_host);

hipError_t run_lambert(mrt_ctx* ctx, const mrt_scene* s, const mrt_camera& cam, const mrt_gbuffer& gb,
                       const float* light_dir, const float* light_e, float* rgb);
