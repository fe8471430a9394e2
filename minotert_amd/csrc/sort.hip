// sort.hip — ray reordering for coherence: 30-bit keys (direction octant, then a 27-bit
// Morton code of the origin inside the scene bounds) sorted with a stable LSD radix sort.
//
// No reference counterpart (the reference traces one ray per pixel in pixel order). The sort
// only changes the order in which the traversal processes rays; results are written back at
// each ray's own index, so they are identical with and without sorting.
//
// Radix sort, 8 bits per pass over tiles of 256 threads x 8 items:
//   hist    per-tile 256-bin digit histogram (LDS atomics), stored digit-major
//   scan    per digit, one workgroup scans that digit's tile counts (coalesced); a second
//           pass adds the exclusive scan of the 256 digit totals -> scatter bases
//   scatter stable rank inside the tile: per round of 256 items, lanes of a wave with equal
//           digits are matched with 8 ballots, wave counts are combined through LDS in wave
//           order, and a running per-digit count carries ranks across rounds
// Item counts may live on the device (count_dev); grids are sized for the capacity.
#include "internal.h"

#include <mutex>

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kTile = kThreads * kItems;

__device__ __forceinline__ uint32_t part1by2(uint32_t v) {  // 9 bits -> every third bit
    v &= 0x1FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(kThreads) void k_ray_keys(const mrt_ray* __restrict__ rays, const uint32_t* __restrict__ perm_in,
                                                       const uint32_t* __restrict__ count_dev, uint32_t cap,
                                                       float3 lo, float3 scale, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
    const uint32_t n = count_dev ? min(cap, *count_dev) : cap;
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = perm_in ? perm_in[i] : i;
    const float4* p = reinterpret_cast<const float4*>(rays + r);
    const float4 a = p[0], b = p[1];
    const uint32_t oct = (b.x < 0.0f ? 4u : 0u) | (b.y < 0.0f ? 2u : 0u) | (b.z < 0.0f ? 1u : 0u);
    const float qx = fminf(fmaxf((a.x - lo.x) * scale.x, 0.0f), 511.0f);
    const float qy = fminf(fmaxf((a.y - lo.y) * scale.y, 0.0f), 511.0f);
    const float qz = fminf(fmaxf((a.z - lo.z) * scale.z, 0.0f), 511.0f);
    const uint32_t m = (part1by2((uint32_t)qx) << 2) | (part1by2((uint32_t)qy) << 1) | part1by2((uint32_t)qz);
    keys[i] = (oct << 27) | m;
    vals[i] = r;
}

__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ count_dev,
                                                         uint32_t cap, uint32_t shift, uint32_t mask, uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_h[256];
    const uint32_t n = count_dev ? min(cap, *count_dev) : cap;
    const uint32_t tid = threadIdx.x, nblocks = gridDim.x;
    s_h[tid] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; ++r) {
        const uint32_t i = base + r * kThreads + tid;
        if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[tid * nblocks + blockIdx.x] = s_h[tid];
}

// Block-wide inclusive scan of one value per thread (kThreads threads).
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t* s) {
    const uint32_t tid = threadIdx.x;
    s[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kThreads; off <<= 1) {
        const uint32_t add = tid >= off ? s[tid - off] : 0u;
        __syncthreads();
        s[tid] += add;
        __syncthreads();
    }
    const uint32_t r = s[tid];
    __syncthreads();
    return r;
}

// Level 1: workgroup d exclusive-scans the nblocks counts of digit d in place (coalesced
// chunks of kThreads) and writes the digit total.
__global__ __launch_bounds__(kThreads) void k_scan_digit(uint32_t* __restrict__ hist, uint32_t nblocks,
                                                         uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[kThreads];
    __shared__ uint32_t s_total;
    uint32_t* row = hist + (size_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += kThreads) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nblocks ? row[i] : 0u;
        const uint32_t incl = block_scan_incl(v, s);
        if (i < nblocks) row[i] = carry + incl - v;
        if (threadIdx.x == kThreads - 1) s_total = incl;  // the chunk total
        __syncthreads();
        carry += s_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Level 2: exclusive scan of the 256 digit totals, added to every entry of each digit row.
__global__ __launch_bounds__(kThreads) void k_scan_add(uint32_t* __restrict__ hist, uint32_t nblocks,
                                                       const uint32_t* __restrict__ totals) {
    __shared__ uint32_t s[kThreads];
    __shared__ uint32_t base;
    const uint32_t v = totals[threadIdx.x];
    const uint32_t incl = block_scan_incl(v, s);
    if (threadIdx.x == blockIdx.x) base = incl - v;
    __syncthreads();
    uint32_t* row = hist + (size_t)blockIdx.x * nblocks;
    for (uint32_t i = threadIdx.x; i < nblocks; i += kThreads) row[i] += base;
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                            const uint32_t* __restrict__ vals_in,
                                                            uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                            const uint32_t* __restrict__ count_dev, uint32_t cap,
                                                            uint32_t shift, uint32_t mask, const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t s_run[256];
    __shared__ uint32_t s_wave[kThreads / 64][256];
    __shared__ uint32_t s_base[256];
    const uint32_t n = count_dev ? min(cap, *count_dev) : cap;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u, nblocks = gridDim.x;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    s_run[tid] = 0;
    s_base[tid] = offsets[tid * nblocks + blockIdx.x];
    const uint32_t base = blockIdx.x * kTile;
    for (int r = 0; r < kItems; ++r) {
        for (uint32_t w = 0; w < kThreads / 64; ++w) s_wave[w][tid] = 0;
        __syncthreads();
        const uint32_t i = base + r * kThreads + tid;
        const bool valid = i < n;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? vals_in[i] : 0u;
        const uint32_t digit = (key >> shift) & mask;
        uint64_t m = __ballot(valid);
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t mb = __ballot((digit >> bit) & 1u);
            m &= ((digit >> bit) & 1u) ? mb : ~mb;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0) s_wave[wave][digit] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (uint32_t w = 0; w < wave; ++w) before += s_wave[w][digit];
            const uint32_t pos = s_base[digit] + s_run[digit] + before + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (uint32_t w = 0; w < kThreads / 64; ++w) add += s_wave[w][tid];
        s_run[tid] += add;
        __syncthreads();
    }
}

}  // namespace

hipError_t SortSpace::ensure(size_t cap_items) {
    if (cap_items <= cap) return hipSuccess;
    release();
    const size_t blocks = (cap_items + kTile - 1) / kTile;
    hipError_t e = hipSuccess;
    auto A = [&](uint32_t** p, size_t words) { if (e == hipSuccess) e = hipMalloc((void**)p, (words ? words : 64) * 4); };
    A(&keys[0], cap_items); A(&keys[1], cap_items); A(&vals[0], cap_items); A(&vals[1], cap_items);
    A(&hist, blocks * 256 + 256);  // digit-major counts + 256 digit totals
    if (e != hipSuccess) { release(); return e; }
    cap = cap_items;
    return hipSuccess;
}

void SortSpace::release() {
    for (int b = 0; b < 2; ++b) { hipFree(keys[b]); hipFree(vals[b]); keys[b] = vals[b] = nullptr; }
    hipFree(hist); hist = nullptr;
    cap = 0;
}

// The key kernel and the four 8-bit passes over explicit buffers: keys0 / vals0 receive the keys
// and ray indices, the passes ping-pong with keys1 / vals1, and after the fourth (30-bit keys)
// the sorted order is back in vals0. hist holds ceil(cap / kTile) * 256 + 256 words. Nothing in
// these buffers needs to start at zero: every pass writes the whole histogram it reads.
static void ray_sort_passes(hipStream_t stream, const float bounds[6], const mrt_ray* rays, uint32_t cap,
                            const uint32_t* count_dev, uint32_t* keys0, uint32_t* keys1, uint32_t* vals0,
                            uint32_t* vals1, uint32_t* hist) {
    const uint32_t blocks = (uint32_t)((cap + kTile - 1) / kTile);
    const uint32_t kblocks = (uint32_t)((cap + kThreads - 1) / kThreads);
    float3 lo = make_float3(bounds[0], bounds[1], bounds[2]);
    float3 ext = make_float3(bounds[3] - bounds[0], bounds[4] - bounds[1], bounds[5] - bounds[2]);
    float3 scale = make_float3(ext.x > 0 ? 512.0f / ext.x : 0.0f, ext.y > 0 ? 512.0f / ext.y : 0.0f,
                               ext.z > 0 ? 512.0f / ext.z : 0.0f);
    hipLaunchKernelGGL(k_ray_keys, dim3(kblocks), dim3(kThreads), 0, stream, rays, nullptr, count_dev, cap, lo, scale,
                       keys0, vals0);
    uint32_t* kb[2] = {keys0, keys1};
    uint32_t* vb[2] = {vals0, vals1};
    int c = 0;
    for (uint32_t shift = 0; shift < 30; shift += 8) {
        const uint32_t mask = (1u << std::min(8u, 30u - shift)) - 1u;
        hipLaunchKernelGGL(k_radix_hist, dim3(blocks), dim3(kThreads), 0, stream, kb[c], count_dev, cap, shift, mask, hist);
        hipLaunchKernelGGL(k_scan_digit, dim3(256), dim3(kThreads), 0, stream, hist, blocks, hist + (size_t)blocks * 256);
        hipLaunchKernelGGL(k_scan_add, dim3(256), dim3(kThreads), 0, stream, hist, blocks, hist + (size_t)blocks * 256);
        hipLaunchKernelGGL(k_radix_scatter, dim3(blocks), dim3(kThreads), 0, stream, kb[c], vb[c], kb[c ^ 1], vb[c ^ 1],
                           count_dev, cap, shift, mask, hist);
        c ^= 1;
    }
    static_assert(((30 + 7) / 8) % 2 == 0, "an even number of passes leaves the order in vals0");
}

// Computes in ctx->sort.vals[0] the processing order of the (count_dev or cap) rays.
hipError_t launch_ray_sort(mrt_ctx* ctx, const float bounds[6], const mrt_ray* rays, uint32_t cap,
                           const uint32_t* count_dev, uint32_t** order_out) {
    SortSpace& w = ctx->sort;
    hipError_t e = w.ensure(cap);
    if (e != hipSuccess) return e;
    {
        ProfScope prof(ctx, "ray_sort");
        ray_sort_passes(ctx->stream, bounds, rays, cap, count_dev, w.keys[0], w.keys[1], w.vals[0], w.vals[1], w.hist);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    *order_out = w.vals[0];
    return hipSuccess;
}

// ---- allocation-free ray order ------------------------------------------------------------
// For callers that cannot allocate (graph capture, serving loops): the keys, the second value
// buffer and the histograms live in one file-scope device array, 26.2 MB for at most
// kNoAllocMaxRays rays; the caller's order buffer is the first value buffer. Each device holds its
// own copy of the array (at its own address, looked up once per device at context creation);
// calls on one device share it, so calls that overlap on one device race.
constexpr uint32_t kNoAllocMaxRays = 1u << 21;
constexpr size_t kNoAllocBlocks = (kNoAllocMaxRays + kTile - 1) / kTile;
constexpr size_t kScratchWords = 3 * (size_t)kNoAllocMaxRays + kNoAllocBlocks * 256 + 256;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (call with the device current). The name
// `scratch` is a host-side stand-in in host code; only this address may reach a kernel.
hipError_t noalloc_scratch_init(int device, uint32_t** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

uint32_t noalloc_max_rays() { return kNoAllocMaxRays; }

// The order of n <= kNoAllocMaxRays rays into order (device, n words), with no allocation.
hipError_t launch_ray_sort_noalloc(mrt_ctx* ctx, const float bounds[6], const mrt_ray* rays, uint32_t n,
                                   uint32_t* order) {
    if (n > kNoAllocMaxRays || !ctx->noalloc_scratch) return hipErrorInvalidValue;
    uint32_t* base = ctx->noalloc_scratch;
    uint32_t* keys0 = base;
    uint32_t* keys1 = base + kNoAllocMaxRays;
    uint32_t* vals1 = base + 2 * (size_t)kNoAllocMaxRays;
    uint32_t* hist = base + 3 * (size_t)kNoAllocMaxRays;
    ray_sort_passes(ctx->stream, bounds, rays, n, nullptr, keys0, keys1, order, vals1, hist);
    return hipGetLastError();
}

// Sorts (key, value) pairs given by the caller (stable LSD, `bits` low key bits).
hipError_t launch_sort_pairs(mrt_ctx* ctx, uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t bits) {
    SortSpace& w = ctx->sort;
    hipError_t e = w.ensure(n);
    if (e != hipSuccess) return e;
    const uint32_t blocks = (uint32_t)((n + kTile - 1) / kTile);
    uint32_t* kb[2] = {keys, w.keys[1]};
    uint32_t* vb[2] = {vals, w.vals[1]};
    int c = 0;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        const uint32_t mask = (1u << std::min(8u, bits - shift)) - 1u;
        hipLaunchKernelGGL(k_radix_hist, dim3(blocks), dim3(kThreads), 0, ctx->stream, kb[c], nullptr, n, shift, mask,
                           w.hist);
        hipLaunchKernelGGL(k_scan_digit, dim3(256), dim3(kThreads), 0, ctx->stream, w.hist, blocks, w.hist + (size_t)blocks * 256);
        hipLaunchKernelGGL(k_scan_add, dim3(256), dim3(kThreads), 0, ctx->stream, w.hist, blocks, w.hist + (size_t)blocks * 256);
        hipLaunchKernelGGL(k_radix_scatter, dim3(blocks), dim3(kThreads), 0, ctx->stream, kb[c], vb[c], kb[c ^ 1],
                           vb[c ^ 1], nullptr, n, shift, mask, w.hist);
        c ^= 1;
    }
    if (c == 1) {  // result sits in the scratch pair: copy back
        if ((e = hipMemcpyAsync(keys, kb[1], (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(vals, vb[1], (size_t)n * 4, hipMemcpyDeviceToDevice, ctx->stream)) != hipSuccess) return e;
    }
    return hipGetLastError();
}
