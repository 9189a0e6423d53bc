// reorder.hip -- Morton (Z-order) ordering of the particles (gfx950).
//
// The sweep and the neighbour passes map slot t to particle perm[t] and
// read the particles through slot-ordered copies (gather_kernel: positions,
// velocities, densities).  Along a Z-order curve consecutive slots are
// spatial neighbours, so the 64 lanes of a wavefront share most of their
// SPH partners, consecutive j of the sweep are close together (few j chunks
// hold neighbour candidates), and the list passes' gathers of neighbour data
// hit the same cache lines.  The order changes the summation order of the
// (compensated) sums, not which pairs are summed.
#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"

namespace grasph {
namespace {

constexpr int kBlock = 256;
constexpr int kBoxBlocks = 256;

__global__ __launch_bounds__(kBlock) void bbox_partial(const float4* __restrict__ posm, const int n,
                                                       float* __restrict__ part)
{
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float4 p = posm[i];
        mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
    }
    __shared__ float sh[6][kBlock];
    for (int d = 0; d < 3; ++d) {
        sh[d][threadIdx.x] = mn[d];
        sh[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                sh[d][threadIdx.x] = fminf(sh[d][threadIdx.x], sh[d][threadIdx.x + s]);
                sh[3 + d][threadIdx.x] = fmaxf(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// one block of kBlock threads folds the nparts (<= kBlock) block boxes
__global__ __launch_bounds__(kBlock) void bbox_final(float* __restrict__ part, const int nparts)
{
    __shared__ float sh[6][kBlock / 64];
    const int b = threadIdx.x, lane = b & 63, w = b >> 6;
    float v[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) v[d] = (b < nparts) ? part[b * 6 + d] : ((d < 3) ? INFINITY : -INFINITY);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const float o = __shfl_xor(v[d], off, 64);
            v[d] = (d < 3) ? fminf(v[d], o) : fmaxf(v[d], o);
        }
    if (lane == 0)
#pragma unroll
        for (int d = 0; d < 6; ++d) sh[d][w] = v[d];
    __syncthreads();
    if (b < 6) {
        float r = sh[b][0];
        for (int q = 1; q < kBlock / 64; ++q) r = (b < 3) ? fminf(r, sh[b][q]) : fmaxf(r, sh[b][q]);
        part[nparts * 6 + b] = r;
    }
}

__device__ inline uint32_t spread10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(kBlock) void morton_keys(const float4* __restrict__ posm, const int n,
                                                      const float* __restrict__ box, uint32_t* __restrict__ keys,
                                                      int* __restrict__ vals)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 p = posm[i];
    uint32_t q[3];
    const float c[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d) {
        const float lo = box[d], hi = box[3 + d];
        const float ext = hi - lo;
        float u = (ext > 0.0f) ? (c[d] - lo) / ext : 0.0f;
        if (!(u >= 0.0f)) u = 0.0f;  // NaN and negatives
        if (u > 1.0f) u = 1.0f;
        q[d] = (uint32_t)(u * 1023.0f);
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = i;
}

// one block per instance: bounding box of its particles
__global__ __launch_bounds__(kBlock) void batch_bbox(const float4* __restrict__ posm,
                                                     const BatchInst* __restrict__ insts, float* __restrict__ box)
{
    const BatchInst I = insts[blockIdx.x];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = I.off + threadIdx.x; i < I.off + I.n; i += kBlock) {
        const float4 p = posm[i];
        mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
    }
    __shared__ float sh[6][kBlock];
    for (int d = 0; d < 3; ++d) {
        sh[d][threadIdx.x] = mn[d];
        sh[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                sh[d][threadIdx.x] = fminf(sh[d][threadIdx.x], sh[d][threadIdx.x + s]);
                sh[3 + d][threadIdx.x] = fmaxf(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) box[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void batch_keys(const float4* __restrict__ posm, const int* __restrict__ inst_of,
                                                     const int n, const float* __restrict__ box,
                                                     uint64_t* __restrict__ keys, int* __restrict__ vals)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int k = inst_of[i];
    const float4 p = posm[i];
    const float c[3] = {p.x, p.y, p.z};
    uint32_t q[3];
    for (int d = 0; d < 3; ++d) {
        const float lo = box[6 * k + d], hi = box[6 * k + 3 + d];
        const float ext = hi - lo;
        float u = (ext > 0.0f) ? (c[d] - lo) / ext : 0.0f;
        if (!(u >= 0.0f)) u = 0.0f;
        if (u > 1.0f) u = 1.0f;
        q[d] = (uint32_t)(u * 1023.0f);
    }
    keys[i] = ((uint64_t)k << 30) | (uint64_t)((spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]));
    vals[i] = i;
}

// slot-ordered copies: xs[t] = posm[perm[t]] (when xs is given) and the
// slot records' {x y z m} and {vx vy vz rho}; with mass_flags, also the
// first half of the equal-mass check: flag b is set when a mass of block b
// differs from particle 0's, bit for bit
__global__ __launch_bounds__(kBlock) void gather_kernel(const int* __restrict__ perm, const int n,
                                                        const float4* __restrict__ posm, const float4* __restrict__ vel,
                                                        const float* __restrict__ density, float4* __restrict__ xs,
                                                        SlotRec* __restrict__ rec, int* __restrict__ mass_flags)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    int bad = 0;
    if (t < n) {
        const int i = perm ? perm[t] : t;
        const float4 p = posm[i];
        if (xs) xs[t] = p;
        if (rec) {
            const float4 v = vel ? vel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            rec[t].pm = p;
            rec[t].vr = make_float4(v.x, v.y, v.z, density ? density[i] : 0.0f);
        }
        if (mass_flags) bad = __float_as_int(p.w) != __float_as_int(posm[0].w);
    }
    if (mass_flags) {  // uniform: the whole workgroup reaches the barrier
        bad = __syncthreads_or(bad);
        if (threadIdx.x == 0) mass_flags[blockIdx.x] = bad;
    }
}

}  // namespace

hipError_t launch_gather(const int* perm, int n, const float4* posm, const float4* vel, const float* density,
                         float4* xs, SlotRec* rec, hipStream_t s, int* mass_flags)
{
    if (n <= 0) return hipSuccess;
    if (!posm || (xs && !perm)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, perm, n, posm, vel,
                       density, xs, rec, mass_flags);
    return hipGetLastError();
}

size_t batch_order_tmp_bytes(int ntotal, int key_bits)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (uint64_t*)nullptr, (uint64_t*)nullptr, (int*)nullptr,
                                    (int*)nullptr, ntotal, 0u, (unsigned)key_bits, (hipStream_t)0);
    return bytes;
}

hipError_t launch_batch_order(const BatchOrder& B, hipStream_t s)
{
    if (B.ntotal <= 0 || B.n_inst <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_bbox, dim3(B.n_inst), dim3(kBlock), 0, s, B.posm, B.insts, B.bbox);
    hipLaunchKernelGGL(batch_keys, dim3((B.ntotal + kBlock - 1) / kBlock), dim3(kBlock), 0, s, B.posm, B.inst_of,
                       B.ntotal, B.bbox, B.keys, B.vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = B.sort_tmp_bytes;
    return rocprim::radix_sort_pairs(B.sort_tmp, bytes, B.keys, B.keys_alt, B.vals, B.perm, B.ntotal, 0u,
                                     (unsigned)B.key_bits, s);
}

size_t reorder_tmp_bytes(int n)
{
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (int*)nullptr,
                                    (int*)nullptr, n, 0u, 30u, (hipStream_t)0);
    return bytes;
}

hipError_t launch_reorder(const float4* posm, int n, Ctx& c, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bbox_partial, dim3(kBoxBlocks), dim3(kBlock), 0, s, posm, n, c.bbox);
    static_assert(kBoxBlocks <= kBlock, "bbox_final folds at most one box per thread");
    hipLaunchKernelGGL(bbox_final, dim3(1), dim3(kBlock), 0, s, c.bbox, kBoxBlocks);
    const int blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(morton_keys, dim3(blocks), dim3(kBlock), 0, s, posm, n, c.bbox + kBoxBlocks * 6, c.keys,
                       c.vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = c.sort_tmp_bytes;
    return rocprim::radix_sort_pairs(c.sort_tmp, bytes, c.keys, c.keys_alt, c.vals, c.perm, n, 0u, 30u, s);
}

}  // namespace grasph
