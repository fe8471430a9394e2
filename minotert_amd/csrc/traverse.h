// traverse.h — per-lane BVH4 traversal of one ray, for kernels in which every lane carries its
// own path (the megakernel path loop). Same node layout, slab test and triangle test as the
// wavefront traversal kernels of trace.hip, so the results are identical: the box test is
// conservative (interval widened by a relative 2^-15) and the closest hit is decided by the
// watertight test's (t, smaller triangle ID) rule, neither of which depends on visiting order.
#pragma once
#include "internal.h"

namespace rtrav {

using rt::v3;

constexpr float kSlack = 1.0f / 32768.0f;

// Box test with the planes ordered by the ray's direction signs. Per ray: sel = which of the
// lo / hi rows is the near plane on each axis, off_near / off_far = o * inv moved outward by
// 2^-21 |o * inv|. A plane distance is then one FMA, p * inv - off, whose only error beyond the
// final rounding is the rounding of o * inv (at most 2^-24 |o * inv|), which the margin covers:
// computed near distances never exceed, and far distances never fall below, (p - o) * inv. The
// far distance is then widened by a relative kSlack, so the test stays conservative.
struct DirRay {
    v3 inv, off_near, off_far;
    uint32_t nx, ny, nz;  // byte offsets in a node of the near rows (lo row, or hi row when the
                          // axis direction is negative)
    uint32_t fx, fy, fz;  // and of the far rows (the other one); all below 128, so a row's byte
                          // address is (node << 7) | offset
};
__device__ __forceinline__ DirRay make_dir_ray(v3 o, v3 inv) {
    DirRay r;
    r.inv = inv;
    const float ox = o.x * inv.x, oy = o.y * inv.y, oz = o.z * inv.z;
    const float ex = fabsf(ox) * (1.0f / 2097152.0f), ey = fabsf(oy) * (1.0f / 2097152.0f), ez = fabsf(oz) * (1.0f / 2097152.0f);
    r.off_near = rt::make_v3(ox + ex, oy + ey, oz + ez);
    r.off_far = rt::make_v3(ox - ex, oy - ey, oz - ez);
    r.nx = inv.x < 0.0f ? 16u : 0u;   // rows: lo.x 0, hi.x 16, lo.y 32, hi.y 48, lo.z 64, hi.z 80
    r.ny = inv.y < 0.0f ? 48u : 32u;
    r.nz = inv.z < 0.0f ? 80u : 64u;
    r.fx = 16u - r.nx;
    r.fy = 80u - r.ny;
    r.fz = 144u - r.nz;
    return r;
}
// The six rows of node `cur` in near / far order for this ray (base = the node array).
struct DirRows { float4 nx, fx, ny, fy, nz, fz; };
__device__ __forceinline__ DirRows load_dir_rows(const char* __restrict__ base, uint32_t cur, const DirRay& r) {
    const uint32_t b = cur << 7;
    DirRows w;
    w.nx = *reinterpret_cast<const float4*>(base + (b | r.nx));
    w.fx = *reinterpret_cast<const float4*>(base + (b | r.fx));
    w.ny = *reinterpret_cast<const float4*>(base + (b | r.ny));
    w.fy = *reinterpret_cast<const float4*>(base + (b | r.fy));
    w.nz = *reinterpret_cast<const float4*>(base + (b | r.nz));
    w.fz = *reinterpret_cast<const float4*>(base + (b | r.fz));
    return w;
}
__device__ __forceinline__ bool slab_dir(float nx, float fx, float ny, float fy, float nz, float fz, const DirRay& r,
                                         float tmin, float tlim, float& tnear) {
    const float tn = fmaxf(fmaxf(fmaf(nx, r.inv.x, -r.off_near.x), fmaf(ny, r.inv.y, -r.off_near.y)),
                           fmaf(nz, r.inv.z, -r.off_near.z));
    const float tf = fminf(fminf(fmaf(fx, r.inv.x, -r.off_far.x), fmaf(fy, r.inv.y, -r.off_far.y)),
                           fmaf(fz, r.inv.z, -r.off_far.z));
    const float tfw = fmaf(fabsf(tf), kSlack, tf);  // tf widened by 2^-15 |tf|
    tnear = tn;
    return (tn <= fminf(tfw, tlim)) & (tfw >= tmin);
}

// Stack entry sp {node, entry t} of lane tid: the LDS copy at a clamped index, replaced by the
// spilled entry past LDS_DEPTH. The LDS read is unconditional and the spill read nontemporal, so
// the compiler cannot fold the two into one generic (flat) load, which would wait on both memory
// counters at every pop.
template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ uint2 stack_entry(const uint2* s_stack, const uint2* __restrict__ spill, int sp,
                                             uint32_t nthreads, uint32_t gtid, int tid) {
    uint2 e = s_stack[min(sp, LDS_DEPTH - 1) * BLOCK + tid];
    if (sp >= LDS_DEPTH) {
        const unsigned long long v = __builtin_nontemporal_load(
            reinterpret_cast<const unsigned long long*>(spill) + (__umul24((uint32_t)(sp - LDS_DEPTH), nthreads) + gtid));
        e = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    }
    return e;
}

// Stack: entry k of lane `tid` at s_stack[k * BLOCK + tid] for k < LDS_DEPTH, then at
// spill[(k - LDS_DEPTH) * nthreads + gtid]. best_t holds tmax on entry; on return (best_t,
// best_id, bu, bv) is the closest hit (best_id = kMissId on a miss), or any hit when `any`.
// `any` is per lane: lanes of one wave may trace closest-hit and any-hit rays in the same loop.
template <int BLOCK, int LDS_DEPTH>
__device__ __forceinline__ void traverse(const WNode* __restrict__ nodes, const float4* __restrict__ tris, uint32_t root,
                                         v3 o, v3 d, float tmin, float& best_t, uint32_t& best_id, float& bu, float& bv,
                                         bool any, uint2* s_stack, uint2* __restrict__ spill, uint32_t nthreads,
                                         uint32_t gtid, int tid) {
    best_id = kMissId;
    bu = bv = 0.0f;
    const rt::RayShear sh = rt::make_shear(o, d, tmin);
    if (!sh.valid) return;
    const v3 inv = rt::make_v3(rt::safe_rcp(d.x), rt::safe_rcp(d.y), rt::safe_rcp(d.z));
    const DirRay dr = make_dir_ray(o, inv);
    float tlim = best_t + fabsf(best_t) * kSlack;
    int sp = 0;
    uint32_t cur = root;
    for (;;) {
        if (cur == kEmptyChild) {
            while (sp > 0) {
                --sp;
                const uint2 e = stack_entry<BLOCK, LDS_DEPTH>(s_stack, spill, sp, nthreads, gtid, tid);
                if (__uint_as_float(e.y) <= tlim) { cur = e.x; break; }
            }
            if (cur == kEmptyChild) return;
        }
        if (cur & kLeafBit) {
            const uint32_t first = (cur >> 3) & 0x0FFFFFFFu;
            const uint32_t count = (cur & 7u) + 1u;
            const float4* tp = tris + (size_t)first * 3;
            for (uint32_t k = 0; k < count; ++k, tp += 3) {
                const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
                if (rt::watertight_hit(sh, rt::make_v3(p0.x, p0.y, p0.z), rt::make_v3(p1.x, p1.y, p1.z),
                                       rt::make_v3(p2.x, p2.y, p2.z), __float_as_uint(p0.w), best_t, best_id, bu, bv)) {
                    if (any) return;
                    tlim = best_t + fabsf(best_t) * kSlack;
                }
            }
            cur = kEmptyChild;
            continue;
        }
        const DirRows w = load_dir_rows(reinterpret_cast<const char*>(nodes), cur, dr);
        const float4 nx = w.nx, fx = w.fx, ny = w.ny, fy = w.fy, nz = w.nz, fz = w.fz;
        const float4 cf = reinterpret_cast<const float4*>(nodes + cur)[6];
        const uint32_t cw0 = __float_as_uint(cf.x), cw1 = __float_as_uint(cf.y);
        const uint32_t cw2 = __float_as_uint(cf.z), cw3 = __float_as_uint(cf.w);
        float t0, t1, t2, t3;
        const bool h0 = slab_dir(nx.x, fx.x, ny.x, fy.x, nz.x, fz.x, dr, tmin, tlim, t0) & (cw0 != kEmptyChild);
        const bool h1 = slab_dir(nx.y, fx.y, ny.y, fy.y, nz.y, fz.y, dr, tmin, tlim, t1) & (cw1 != kEmptyChild);
        const bool h2 = slab_dir(nx.z, fx.z, ny.z, fy.z, nz.z, fz.z, dr, tmin, tlim, t2) & (cw2 != kEmptyChild);
        const bool h3 = slab_dir(nx.w, fx.w, ny.w, fy.w, nz.w, fz.w, dr, tmin, tlim, t3) & (cw3 != kEmptyChild);
        const float kFar = __builtin_huge_valf();
        float k0 = h0 ? t0 : kFar, k1 = h1 ? t1 : kFar, k2 = h2 ? t2 : kFar, k3 = h3 ? t3 : kFar;
        uint32_t c0 = cw0, c1 = cw1, c2 = cw2, c3 = cw3;
        const int nh = (int)h0 + (int)h1 + (int)h2 + (int)h3;
#define RTRAV_CSWAP(ka, ca, kb, cb) { const bool sw = kb < ka; const float tk = sw ? kb : ka; kb = sw ? ka : kb; ka = tk; \
                                      const uint32_t tc = sw ? cb : ca; cb = sw ? ca : cb; ca = tc; }
        RTRAV_CSWAP(k0, c0, k1, c1); RTRAV_CSWAP(k2, c2, k3, c3); RTRAV_CSWAP(k0, c0, k2, c2);
        RTRAV_CSWAP(k1, c1, k3, c3); RTRAV_CSWAP(k1, c1, k2, c2);
#undef RTRAV_CSWAP
        auto push = [&](uint32_t node, float t) {
            const uint2 e = make_uint2(node, __float_as_uint(t));
            if (sp < LDS_DEPTH) s_stack[sp * BLOCK + tid] = e;
            else spill[__umul24((uint32_t)(sp - LDS_DEPTH), nthreads) + gtid] = e;
            ++sp;
        };
        if (nh > 3) push(c3, k3);
        if (nh > 2) push(c2, k2);
        if (nh > 1) push(c1, k1);
        cur = nh > 0 ? c0 : kEmptyChild;
    }
}

}  // namespace rtrav
