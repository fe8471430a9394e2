// util.hip — small stream-ordered utility kernels: a plain byte copy (also the reference
// point for converting rocprof FETCH_SIZE/WRITE_SIZE counters into bytes) and a word fill
// (used instead of hipMemsetAsync so that the library's own clears are graph-capturable).
#include "internal.h"

namespace {

__global__ __launch_bounds__(256) void k_copy_bytes(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    size_t n16, const uint8_t* __restrict__ src_tail,
                                                    uint8_t* __restrict__ dst_tail, uint32_t tail) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
    if (blockIdx.x == 0 && threadIdx.x < tail) dst_tail[threadIdx.x] = src_tail[threadIdx.x];
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ dst, size_t n, uint32_t value) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = value;
}

}  // namespace

hipError_t launch_copy_bytes(mrt_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    const bool aligned = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0);
    size_t n16 = aligned ? bytes / 16 : 0;
    uint32_t tail = (uint32_t)(bytes - n16 * 16);
    if (tail > 256) return hipErrorInvalidValue;  // unaligned buffers larger than 256 B
    uint32_t blocks = (uint32_t)std::min<size_t>((n16 + 255) / 256, (size_t)ctx->num_cus * 16);
    if (blocks == 0) blocks = 1;
    ProfScope prof(ctx, "copy_bytes");
    hipLaunchKernelGGL(k_copy_bytes, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4*)src, (uint4*)dst, n16,
                       (const uint8_t*)src + n16 * 16, (uint8_t*)dst + n16 * 16, tail);
    return hipGetLastError();
}

hipError_t launch_fill_u32(mrt_ctx* ctx, uint32_t* dst, size_t n, uint32_t value) {
    if (n == 0) return hipSuccess;
    uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 8);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, ctx->stream, dst, n, value);
    return hipGetLastError();
}
