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

// sum += rgb; mean = sum / frames (progressive accumulation of 1-spp frames).
__global__ __launch_bounds__(256) void k_accumulate(float* __restrict__ sum, const float* __restrict__ rgb, size_t n,
                                                    float frames, float* __restrict__ mean) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = sum[i] + rgb[i];
        sum[i] = v;
        if (mean) mean[i] = v / frames;
    }
}

// Per-block partial sums of (x - r)^2 / (r^2 + eps) in double.
__global__ __launch_bounds__(256) void k_rel_mse(const float* __restrict__ img, const float* __restrict__ ref, size_t n,
                                                 double eps, double* __restrict__ partial) {
    __shared__ double s_sum[256];
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double d = (double)img[i] - (double)ref[i];
        acc += d * d / ((double)ref[i] * (double)ref[i] + eps);
    }
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s_sum[0];
}

}  // namespace

hipError_t launch_accumulate(mrt_ctx* ctx, float* sum, const float* rgb, size_t n, uint32_t frames, float* mean) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, (size_t)ctx->num_cus * 8);
    ProfScope prof(ctx, "accumulate");
    hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(256), 0, ctx->stream, sum, rgb, n, (float)frames, mean);
    return hipGetLastError();
}

// The partial sums go through a small device buffer owned by the context and are added on the host.
hipError_t launch_rel_mse(mrt_ctx* ctx, const float* img, const float* ref, size_t n, double eps, double* out_host) {
    *out_host = 0.0;
    if (n == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 1024);
    if (!ctx->d_partials) {
        hipError_t e = hipMalloc((void**)&ctx->d_partials, 1024 * sizeof(double));
        if (e != hipSuccess) return e;
    }
    {
        ProfScope prof(ctx, "rel_mse");
        hipLaunchKernelGGL(k_rel_mse, dim3(blocks), dim3(256), 0, ctx->stream, img, ref, n, eps, ctx->d_partials);
    }
    double host[1024];
    hipError_t e = hipMemcpyAsync(host, ctx->d_partials, blocks * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return e;
    double sum = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) sum += host[b];
    *out_host = sum / (double)n;
    return hipSuccess;
}

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
