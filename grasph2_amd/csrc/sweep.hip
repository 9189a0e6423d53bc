// sweep.hip -- the O(N^2) all-pairs sweep (gfx950).
//
// One sweep serves the three pair jobs of src/computeDerivatives.h: it sums
// the softened gravity of calcAcceleration (:229-235) over every pair and
// records the neighbour list (0 < r^2 <= H^2, :72,122,238) that the O(N k)
// passes of nbr_passes.hip walk for the SPH sums (density, strain rate,
// SPH acceleration).  It replaces the tiled device loop
// do_for_each_pair_fast_impl (src/particles/algorithms.h:270-309).
//
// Mapping: NI i particles per lane (slots t and t + 64 for NI = 2).  The
// particle stream xs is in slot order: with a Morton order it is the
// positions gathered as xs[t] = posm[perm[t]] (a wavefront's particles are
// spatial neighbours, so the neighbour branch is coherent, and consecutive
// j are close too, so few j chunks hold candidates), without one it is
// posm itself.  Results go to particle perm[t]; the lists hold slot
// indices.  All lanes visit the same j at the same time, so the j stream
// (float4 x,y,z,m) is read with scalar loads into SGPRs -- no LDS, no
// barriers -- in chunks of 8, two SGPR buffers in ping-pong so the next
// chunk loads while this one computes.
// Per pair: 3 sub, 3 FMA for r^2 + H^2, one v_rsq_f32, 2 mul (3 with
// unequal masses), 3 FMA, and a v_min towards one wave-uniform
// neighbour-candidate test per chunk (a second transcendental per pair in
// place of a multiply, rsqrt * rcp, measured 12% slower); only
// chunks where some lane has a candidate enter the (rare) neighbour path.
// Sums are plain FP32 over blocks of 128 j, then compensated (TwoSum) into
// the running total.
//
// The batched kernel runs the same code persistently over a queue of work
// items (instance, 64 NI slots) ordered longest instance first; each
// wavefront takes the next item with one atomic add.
#include <algorithm>

#include "sweep_core.hpp"

namespace grasph {
namespace {

constexpr int kMassBlocks = 256;  // blocks of the single-instance mass check (flags after umass[0..1])

template <int NI, bool GRAV, bool NBR, bool POT, bool SOFT>
__global__ __launch_bounds__(kBlock) void sweep_kernel(const float4* __restrict__ xs, const int* __restrict__ perm,
                                                       const int n, float4* __restrict__ acc, int* __restrict__ nbr,
                                                       int* __restrict__ nbr_cnt, float* __restrict__ pot,
                                                       const int kmax, const int finish,
                                                       const float4* __restrict__ vel, float4* __restrict__ xvel,
                                                       float* __restrict__ maxvsig, const DevParams P,
                                                       const int slice_len, float4* __restrict__ part,
                                                       const float* __restrict__ umass)
{
    const int wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int s0 = wave * 64 * NI + lane;
    // slice blockIdx.y of the j range; its list and counts have their own planes
    const int sl = blockIdx.y;
    const int j0 = sl * slice_len;
    int j1 = min(n, j0 + slice_len);
    if (!(GRAV || NBR || POT)) j1 = j0;  // nothing pairwise to do: skip the j loop
    const Rare R = make_rare(nbr + (size_t)sl * list_plane(n, kmax), kmax, P.H2);
    float4* const pp = part ? part + 2 * (size_t)sl * n : nullptr;
    if constexpr (GRAV && !POT) {
        if (umass && umass[0] != 0.0f) {  // {uniform, mass} of the partners, wave-uniform
            sweep_slots<NI, GRAV, NBR, POT, SOFT, true>(xs, perm, s0, n, j0, j1, acc, nbr_cnt + (size_t)sl * n, pot,
                                                        finish, vel, xvel, maxvsig, P, P.H2, P.cand, R, umass[1], pp);
            return;
        }
    }
    sweep_slots<NI, GRAV, NBR, POT, SOFT, false>(xs, perm, s0, n, j0, j1, acc, nbr_cnt + (size_t)sl * n, pot,
                                                 finish, vel, xvel, maxvsig, P, P.H2, P.cand, R, 1.0f, pp);
}

// {1, m} when every particle of [off, off + n) has mass m (bit for bit),
// else {0, m}; one block per instance (blockIdx.x; insts null: one instance
// [0, n)).  Written every evaluation, so no reset is needed.
__global__ __launch_bounds__(1024) void mass_check_kernel(const float4* __restrict__ posm,
                                                          const BatchInst* __restrict__ insts, const int n,
                                                          float* __restrict__ umass)
{
    const int k = blockIdx.x;
    const int off = insts ? insts[k].off : 0, cnt = insts ? insts[k].n : n;
    if (cnt <= 0) {
        if (threadIdx.x == 0) { umass[2 * k] = 0.0f; umass[2 * k + 1] = 0.0f; }
        return;  // the whole block leaves: no barrier below is reached
    }
    const float m0 = posm[off].w;
    int bad = 0;
    // eight loads in flight per thread (one at a time, each waited for
    // before the next, 16384 masses took 7 us)
    int i = threadIdx.x;
    for (; i + 7 * (int)blockDim.x < cnt; i += 8 * (int)blockDim.x) {
        float w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = posm[off + i + u * (int)blockDim.x].w;
#pragma unroll
        for (int u = 0; u < 8; ++u) bad |= __float_as_int(w[u]) != __float_as_int(m0);
    }
    for (; i < cnt; i += blockDim.x) bad |= __float_as_int(posm[off + i].w) != __float_as_int(m0);
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        umass[2 * k] = bad ? 0.0f : 1.0f;
        umass[2 * k + 1] = m0;
    }
}

// Concatenate the per-slice neighbour lists of a j-split sweep into slice
// 0's list, in slice order (so slot indices stay increasing), and leave the
// total in slice 0's count (merge_counts_kernel, afterwards).  A total above
// kmax (some slice overflowed, or the slices together do) marks the particle
// for the exact-test scan of its whole instance; its list is left as it is.
// One wavefront per tile of 64 slots and per gridDim.y-th of the rows (one
// for the default capacity, up to 8 for long lists):
// lane t fills row r of its merged list (r past its slice-0 entries) from
// the slice whose entries land there, so every store is one whole 256-byte
// row of the tile (stores of single entries into rows that other
// wavefronts fill too left partial lines in several caches; one thread per
// slot copying slice after slice took 2.6 ms for 1e8 entries).  Reads come
// from slices >= 1 and writes go to rows of slice 0 past its own entries,
// so nothing is read after it is overwritten.
template <int kMax>  // slices, rounded up to a power of two (at most 32)
__global__ __launch_bounds__(64) void merge_lists_kernel(int* __restrict__ nbr, const int* __restrict__ cnt,
                                                         const int n, const int kmax, const int nslices)
{
    const int lane = threadIdx.x, t = blockIdx.x * 64 + lane, set = blockIdx.y, nsets = gridDim.y;
    const size_t ns = (size_t)n;
    int start[kMax + 1];  // rows of slice s: [start[s], start[s + 1])
    int c0 = 0, total = 0;
    if (t < n) {
        c0 = cnt[t];
        total = c0;
    }
    start[0] = 0;
    start[1] = c0;
#pragma unroll
    for (int s = 1; s < kMax; ++s) {
        if (s < nslices && t < n) total += cnt[(size_t)s * ns + t];
        start[s + 1] = total;
    }
    const bool fits = t < n && total <= kmax;
    int r_lo = fits ? c0 : INT32_MAX, r_hi = fits ? total : 0;
    // wave-uniform bounds of the rows to fill
    for (int o = 32; o > 0; o >>= 1) {
        r_lo = min(r_lo, __shfl_xor(r_lo, o, 64));
        r_hi = max(r_hi, __shfl_xor(r_hi, o, 64));
    }
    const size_t plane = list_plane(n, kmax);
    int* const dst = nbr + list_entry(t, 0, kmax);  // lane t's column of its tile, row 0
    for (int r0 = r_lo - r_lo % nsets + set; r0 < r_hi; r0 += 8 * nsets) {
        int v[8];
        bool w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int r = r0 + u * nsets;
            w[u] = fits && r >= c0 && r < total;
            v[u] = 0;
            if (w[u]) {
                int sl = 1;
#pragma unroll
                for (int s = 2; s < kMax; ++s)
                    if (s < nslices && r >= start[s]) sl = s;
                v[u] = nbr[(size_t)sl * plane + list_entry(t, r - start[sl], kmax)];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (w[u]) dst[64 * (size_t)(r0 + u * nsets)] = v[u];
    }
}

__global__ __launch_bounds__(kBlock) void merge_counts_kernel(int* __restrict__ cnt, const int n, const int nslices)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    int total = 0;
    for (int s = 0; s < nslices; ++s) total += cnt[(size_t)s * n + t];
    cnt[t] = total;
}

// The same check for one large instance: block b flags a mismatch in its
// grid-stride share, then one block folds the flags (every flag is written
// each time, so nothing needs clearing).
__global__ __launch_bounds__(kBlock) void mass_partial_kernel(const float4* __restrict__ posm, const int n,
                                                              int* __restrict__ partial)
{
    const float m0 = posm[0].w;
    int bad = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        bad |= __float_as_int(posm[i].w) != __float_as_int(m0);
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) partial[blockIdx.x] = bad;
}

__global__ __launch_bounds__(kBlock) void mass_final_kernel(const float4* __restrict__ posm,
                                                            const int* __restrict__ partial, const int nparts,
                                                            float* __restrict__ umass)
{
    int bad = 0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) bad |= partial[b];
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        umass[0] = bad ? 0.0f : 1.0f;
        umass[1] = posm[0].w;
    }
}

// Sum the slices' partial gravity in slice order (compensated) and finish
// the acceleration as a one-slice sweep would.
__global__ __launch_bounds__(kBlock) void combine_kernel(const float4* __restrict__ part, const int nslices,
                                                         const int n, const float4* __restrict__ posm,
                                                         float4* __restrict__ acc, const int finish,
                                                         const float4* __restrict__ vel, float4* __restrict__ xvel,
                                                         float* __restrict__ maxvsig, const DevParams P)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    CSum x{0.0f, 0.0f}, y{0.0f, 0.0f}, z{0.0f, 0.0f};
    for (int s = 0; s < nslices; ++s) {
        const float4 a = part[2 * ((size_t)s * n + i)], c = part[2 * ((size_t)s * n + i) + 1];
        x.add(a.x); y.add(a.y); z.add(a.z);
        x.c += c.x; y.c += c.y; z.c += c.z;
    }
    float ax = x.value(), ay = y.value(), az = z.value();
    if (finish) {
        environment_acceleration(P, ax, ay, az);
        if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (maxvsig) maxvsig[i] = 0.0f;
    }
    acc[i] = make_float4(ax, ay, az, 0.0f);
}

// Persistent sweep over a batch of independent instances.  Item w =
// (inst, slot0, ni) covers slots [slot0, slot0 + 64 ni) of instance inst
// with ni particles per lane (chosen per size bucket); the instance's
// particles are [off, off + n) and its slots are the same range of perm.
// Every wavefront leaves the loop once the counter passes n_items.  NIMAX
// is the largest ni of the items: the two-per-lane code needs about twice
// the registers, so a batch without such items gets the leaner kernel.
template <int NIMAX, bool GRAV, bool NBR, bool SOFT>
__global__ __launch_bounds__(kBlock) void batch_sweep_kernel(const float4* __restrict__ xs,
                                                             const int* __restrict__ perm,
                                                             const int4* __restrict__ items, const int n_items,
                                                             const BatchInst* __restrict__ insts,
                                                             int* __restrict__ queue, float4* __restrict__ acc,
                                                             int* __restrict__ nbr, int* __restrict__ nbr_cnt,
                                                             const int ntotal, const int kmax, const DevParams P,
                                                             const float* __restrict__ umass)
{
    const int lane = threadIdx.x & 63;
    const Rare R0 = make_rare(nbr, kmax, 0.0f);
    for (;;) {
        int w = 0;
        if (lane == 0) w = atomicAdd(queue, 1);
        w = __builtin_amdgcn_readfirstlane(w);
        if (w >= n_items) break;
        const int4 it = items[w];
        const BatchInst I = insts[it.x];
        Rare R = R0;
        R.h2 = vgpr_f32(I.H2);
        const int s0 = it.y + lane, e = I.off + I.n;
        if (GRAV && umass && umass[2 * it.x] != 0.0f) {
            const float m = umass[2 * it.x + 1];
            if (NIMAX == 2 && it.z == 2)
                sweep_slots<NIMAX, GRAV, NBR, false, SOFT, GRAV>(xs, perm, s0, e, I.off, e, acc, nbr_cnt, nullptr, 0,
                                                             nullptr, nullptr, nullptr, P, I.H2, I.cand, R, m);
            else
                sweep_slots<1, GRAV, NBR, false, SOFT, GRAV>(xs, perm, s0, e, I.off, e, acc, nbr_cnt, nullptr, 0,
                                                             nullptr, nullptr, nullptr, P, I.H2, I.cand, R, m);
        } else if (NIMAX == 2 && it.z == 2) {
            sweep_slots<NIMAX, GRAV, NBR, false, SOFT, false>(xs, perm, s0, e, I.off, e, acc, nbr_cnt, nullptr, 0,
                                                          nullptr, nullptr, nullptr, P, I.H2, I.cand, R, 1.0f);
        } else {
            sweep_slots<1, GRAV, NBR, false, SOFT, false>(xs, perm, s0, e, I.off, e, acc, nbr_cnt, nullptr, 0,
                                                          nullptr, nullptr, nullptr, P, I.H2, I.cand, R, 1.0f);
        }
    }
}

// Zero the work counters of the batched sweep (a kernel, so that a captured
// graph holds no memset node).
__global__ void reset_queue_kernel(int* __restrict__ queue, const int n)
{
    const int i = threadIdx.x;
    if (i < n) queue[i] = 0;
}

template <int NI, bool GRAV, bool NBR, bool POT, bool SOFT>
hipError_t launch_t(const SweepLaunch& L, const DevParams& P, hipStream_t s)
{
    const int waves = (L.n + 64 * NI - 1) / (64 * NI);
    const int blocks = (waves + kBlock / 64 - 1) / (kBlock / 64);
    const int ns = L.nslices > 1 ? L.nslices : 1;
    hipLaunchKernelGGL((sweep_kernel<NI, GRAV, NBR, POT, SOFT>), dim3(blocks, ns), dim3(kBlock), 0, s,
                       L.xs ? L.xs : L.posm, L.perm,
                       L.n, L.acc, L.nbr, L.nbr_cnt, L.pot, L.kmax, L.finish ? 1 : 0, L.vel, L.xvel, L.maxvsig, P,
                       ns > 1 ? L.slice_len : L.n, ns > 1 ? L.part : nullptr, L.umass);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ns == 1 || !(GRAV || L.finish)) return e;
    hipLaunchKernelGGL(combine_kernel, dim3((L.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, L.part, ns, L.n, L.posm,
                       L.acc, L.finish ? 1 : 0, L.vel, L.xvel, L.maxvsig, P);
    return hipGetLastError();
}

template <int NI>
hipError_t launch_ni(const SweepLaunch& L, const DevParams& P, hipStream_t s)
{
    const bool soft = P.H2 > 0.0f;
    const int key = (L.grav ? 8 : 0) | (L.want_nbr ? 4 : 0) | (L.want_pot ? 2 : 0) | (soft ? 1 : 0);
    switch (key) {
        case 0: return launch_t<NI, false, false, false, false>(L, P, s);
        case 1: return launch_t<NI, false, false, false, true>(L, P, s);
        case 2: return launch_t<NI, false, false, true, false>(L, P, s);
        case 3: return launch_t<NI, false, false, true, true>(L, P, s);
        case 4: return launch_t<NI, false, true, false, false>(L, P, s);
        case 5: return launch_t<NI, false, true, false, true>(L, P, s);
        case 8: return launch_t<NI, true, false, false, false>(L, P, s);
        case 9: return launch_t<NI, true, false, false, true>(L, P, s);
        case 12: return launch_t<NI, true, true, false, false>(L, P, s);
        case 13: return launch_t<NI, true, true, false, true>(L, P, s);
        default: return hipErrorInvalidValue;  // the potential sweep runs alone
    }
}

template <bool GRAV, bool NBR, bool SOFT>
hipError_t launch_b(const BatchSweepLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.ni_max == 2)
        hipLaunchKernelGGL((batch_sweep_kernel<2, GRAV, NBR, SOFT>), dim3(L.blocks), dim3(64 * L.waves_per_block), 0,
                           s, L.xs, L.perm, L.items, L.n_items, L.insts, L.queue, L.acc, L.nbr, L.nbr_cnt, L.ntotal,
                           L.kmax, P, L.umass);
    else
        hipLaunchKernelGGL((batch_sweep_kernel<1, GRAV, NBR, SOFT>), dim3(L.blocks), dim3(64 * L.waves_per_block), 0,
                           s, L.xs, L.perm, L.items, L.n_items, L.insts, L.queue, L.acc, L.nbr, L.nbr_cnt, L.ntotal,
                           L.kmax, P, L.umass);
    return hipGetLastError();
}

// The largest neighbour count of a list (plane 0 after the merge: totals,
// overflowed ones included), one workgroup: every thread strides the
// counts, then a tree in LDS; the result is written, never accumulated, so
// nothing needs clearing.
__global__ __launch_bounds__(1024) void count_max_kernel(const int* __restrict__ cnt, const int n, int* __restrict__ out)
{
    __shared__ int red[1024];
    int m = 0;
    for (int t = threadIdx.x; t < n; t += 1024) m = max(m, cnt[t]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

}  // namespace

hipError_t launch_sweep(const SweepLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    return L.ni == 2 ? launch_ni<2>(L, P, s) : launch_ni<1>(L, P, s);
}

hipError_t launch_mass_check(const float4* posm, const BatchInst* insts, int n_inst, int n, float* umass,
                             hipStream_t s)
{
    if (n_inst <= 0) return hipSuccess;
    if (!insts && n > 16384) {  // one large instance: spread the scan over blocks, then fold their flags
        const int blocks = std::min(kMassBlocks, (n + 4095) / 4096);
        int* partial = (int*)(umass + 2);
        hipLaunchKernelGGL(mass_partial_kernel, dim3(blocks), dim3(kBlock), 0, s, posm, n, partial);
        hipLaunchKernelGGL(mass_final_kernel, dim3(1), dim3(kBlock), 0, s, posm, partial, blocks, umass);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(mass_check_kernel, dim3(n_inst), dim3(1024), 0, s, posm, insts, n, umass);
    return hipGetLastError();
}

hipError_t launch_mass_final(const float4* posm, const int* mass_flags, int nflags, float* umass, hipStream_t s)
{
    if (nflags <= 0) return hipSuccess;
    hipLaunchKernelGGL(mass_final_kernel, dim3(1), dim3(kBlock), 0, s, posm, mass_flags, nflags, umass);
    return hipGetLastError();
}

hipError_t launch_merge_lists(int* nbr, int* cnt, int n, int kmax, int nslices, hipStream_t s)
{
    if (n <= 0 || nslices <= 1) return hipSuccess;
    if (nslices > 32) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((n + 63) / 64), (unsigned)std::min(8, std::max(1, kmax / 256)));
    if (nslices <= 2)
        hipLaunchKernelGGL(merge_lists_kernel<2>, grid, dim3(64), 0, s, nbr, cnt, n, kmax, nslices);
    else if (nslices <= 4)
        hipLaunchKernelGGL(merge_lists_kernel<4>, grid, dim3(64), 0, s, nbr, cnt, n, kmax, nslices);
    else if (nslices <= 8)
        hipLaunchKernelGGL(merge_lists_kernel<8>, grid, dim3(64), 0, s, nbr, cnt, n, kmax, nslices);
    else if (nslices <= 16)
        hipLaunchKernelGGL(merge_lists_kernel<16>, grid, dim3(64), 0, s, nbr, cnt, n, kmax, nslices);
    else
        hipLaunchKernelGGL(merge_lists_kernel<32>, grid, dim3(64), 0, s, nbr, cnt, n, kmax, nslices);
    hipLaunchKernelGGL(merge_counts_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cnt, n,
                       nslices);
    return hipGetLastError();
}

hipError_t launch_count_max(const int* cnt, int n, int* out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(count_max_kernel, dim3(1), dim3(1024), 0, s, cnt, n, out);
    return hipGetLastError();
}

hipError_t launch_reset_queue(int* queue, int n, hipStream_t s)
{
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, s, queue, n);
    return hipGetLastError();
}

hipError_t launch_batch_sweep(const BatchSweepLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n_items <= 0) return hipSuccess;
    if (L.waves_per_block < 1 || L.waves_per_block > kBlock / 64) return hipErrorInvalidValue;
    if (L.ni_max != 1 && L.ni_max != 2) return hipErrorInvalidValue;
    const int key = (L.grav ? 4 : 0) | (L.want_nbr ? 2 : 0) | (L.soft ? 1 : 0);
    switch (key) {
        case 1: return launch_b<false, false, true>(L, P, s);
        case 3: return launch_b<false, true, true>(L, P, s);
        case 4: return launch_b<true, false, false>(L, P, s);
        case 5: return launch_b<true, false, true>(L, P, s);
        case 7: return launch_b<true, true, true>(L, P, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace grasph
