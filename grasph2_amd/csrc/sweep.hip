// sweep.hip -- the O(N^2) all-pairs sweep (gfx950).
//
// One sweep serves the three pair jobs of src/computeDerivatives.h: it sums
// the softened gravity of calcAcceleration (:229-235) over every pair, the
// SPH density of calcDensity (:68-80) over neighbour pairs, and records the
// neighbour list (0 < r^2 <= H^2, :72,122,238) that the two O(N k) passes of
// nbr_passes.hip walk.  It replaces the tiled device loop
// do_for_each_pair_fast_impl (src/particles/algorithms.h:270-309).
//
// Mapping: one i particle per lane, i taken in Morton order (perm) so that a
// wavefront's 64 particles are spatial neighbours and the neighbour branch is
// coherent.  All lanes visit the same j at the same time, so the j stream
// (float4 x,y,z,m) is read with scalar loads into SGPRs -- no LDS, no
// barriers -- prefetched one chunk of 8 ahead.  Per pair: 3 sub, 3 FMA for
// r^2 + H^2, one v_rsq_f32, 3 mul, 3 FMA; neighbour candidates take a
// wave-uniform branch.  Sums are plain FP32 over blocks of 64 j and then
// compensated (TwoSum) into the running total.
#include "internal.hpp"

namespace grasph {
namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 8;   // j per prefetch chunk
constexpr int kFold = 64;   // j per plain partial sum before the compensated add

struct Acc {
    float ax, ay, az, rho, phi;
    CSum sx, sy, sz, srho, sphi;
    int cnt;
};

template <bool GRAV, bool BRANCH, bool POT, bool SOFT>
__device__ __forceinline__ void pair(const float4 pj, const int j, const float xi, const float yi, const float zi,
                                     Acc& a, int* __restrict__ nbr, const size_t nstride, const int slot,
                                     const int kmax, const bool dens, const bool want_nbr, const DevParams& P)
{
    const float dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
    float d2;
    if (SOFT)
        d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, P.H2)));
    else
        d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float inv = 0.0f;
    if (GRAV || POT) inv = __builtin_amdgcn_rsqf(d2);
    if (GRAV) {
        float s = pj.w * inv;
        s = s * (inv * inv);
        if (!SOFT) s = (d2 > 0.0f) ? s : 0.0f;  // r^2 > 0 (computeDerivatives.h:226)
        a.ax = fmaf(dx, s, a.ax);
        a.ay = fmaf(dy, s, a.ay);
        a.az = fmaf(dz, s, a.az);
    }
    float pc = 0.0f;
    if (POT) pc = (d2 > 0.0f) ? pj.w * inv : 0.0f;
    if (BRANCH) {
        if (d2 <= P.cand) {
            // rij = pi - pj, r2 = dot(rij, rij) exactly as the oracle rounds it
            const float r2 = dist2_exact(-dx, -dy, -dz);
            if (POT && r2 == 0.0f) pc = 0.0f;
            if (is_neighbour(r2, P.H2)) {
                if (dens) {
                    const float r = sqrtf(r2);
                    a.rho += pj.w * w_spline(r, P.H, P.Wpref);
                }
                if (want_nbr) {
                    if (a.cnt < kmax) nbr[(size_t)a.cnt * nstride + slot] = j;
                    ++a.cnt;
                }
            }
        }
    }
    if (POT) a.phi += pc;
}

__device__ __forceinline__ void fold(Acc& a)
{
    a.sx.add(a.ax); a.sy.add(a.ay); a.sz.add(a.az); a.srho.add(a.rho); a.sphi.add(a.phi);
    a.ax = a.ay = a.az = a.rho = a.phi = 0.0f;
}

// Sweep of i slots [t, t+1) (one per lane) against j in [j0, j1).
template <bool GRAV, bool BRANCH, bool POT, bool SOFT>
__device__ __forceinline__ void sweep_range(const float4* __restrict__ posm, const int j0, const int j1,
                                            const float xi, const float yi, const float zi, Acc& a,
                                            int* __restrict__ nbr, const size_t nstride, const int slot,
                                            const int kmax, const bool dens, const bool want_nbr,
                                            const DevParams& P)
{
    const int nfull = j0 + ((j1 - j0) / kChunk) * kChunk;
    float4 cur[kChunk];
    if (nfull > j0) {
#pragma unroll
        for (int c = 0; c < kChunk; ++c) cur[c] = posm[j0 + c];
    }
    for (int jb = j0; jb < nfull; jb += kChunk) {
        float4 nxt[kChunk];
        const int jn = (jb + kChunk < nfull) ? (jb + kChunk) : jb;
#pragma unroll
        for (int c = 0; c < kChunk; ++c) nxt[c] = posm[jn + c];
#pragma unroll
        for (int c = 0; c < kChunk; ++c)
            pair<GRAV, BRANCH, POT, SOFT>(cur[c], jb + c, xi, yi, zi, a, nbr, nstride, slot, kmax, dens, want_nbr, P);
        if (((jb - j0) & (kFold - 1)) == kFold - kChunk) fold(a);
#pragma unroll
        for (int c = 0; c < kChunk; ++c) cur[c] = nxt[c];
    }
    for (int j = nfull; j < j1; ++j)
        pair<GRAV, BRANCH, POT, SOFT>(posm[j], j, xi, yi, zi, a, nbr, nstride, slot, kmax, dens, want_nbr, P);
    fold(a);
}

template <bool GRAV, bool BRANCH, bool POT, bool SOFT>
__global__ __launch_bounds__(kBlock) void sweep_kernel(const float4* __restrict__ posm, const int* __restrict__ perm,
                                                       const int n, float4* __restrict__ acc,
                                                       float* __restrict__ density, int* __restrict__ nbr,
                                                       int* __restrict__ nbr_cnt, float* __restrict__ pot,
                                                       const int kmax, const int dens, const int want_nbr,
                                                       const int finish, const float4* __restrict__ vel,
                                                       float4* __restrict__ xvel, float* __restrict__ maxvsig,
                                                       const DevParams P)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const bool active = t < n;
    const int i = active ? (perm ? perm[t] : t) : 0;
    float xi = __int_as_float(0x7fc00000), yi = xi, zi = xi;  // NaN: inactive lanes never branch
    if (active) {
        const float4 p = posm[i];
        xi = p.x; yi = p.y; zi = p.z;
    }
    Acc a;
    a.ax = a.ay = a.az = a.rho = a.phi = 0.0f;
    a.sx = a.sy = a.sz = a.srho = a.sphi = CSum{0.0f, 0.0f};
    a.cnt = 0;
    sweep_range<GRAV, BRANCH, POT, SOFT>(posm, 0, n, xi, yi, zi, a, nbr, (size_t)n, t, kmax, dens != 0,
                                         want_nbr != 0, P);
    if (!active) return;
    if (GRAV) {
        float ax = a.sx.value(), ay = a.sy.value(), az = a.sz.value();
        if (finish) {
            const float4 v = vel ? vel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
            environment_acceleration(P, xi, yi, zi, v.x, v.y, v.z, ax, ay, az);
        }
        acc[i] = make_float4(ax, ay, az, 0.0f);
    } else if (finish && acc) {
        float ax = 0.f, ay = 0.f, az = 0.f;
        const float4 v = vel ? vel[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        environment_acceleration(P, xi, yi, zi, v.x, v.y, v.z, ax, ay, az);
        acc[i] = make_float4(ax, ay, az, 0.0f);
    }
    if (finish) {
        if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (maxvsig) maxvsig[i] = 0.0f;
    }
    if (BRANCH && dens) {
        float rho = a.srho.value();
        // computeDerivatives.h:85-88
        if ((P.flags & GRASPH_DEAL_WITH_NO_PARTNERS) && rho < 0.001f) rho = P.rho0 * 0.01f;
        density[i] = rho;
    }
    if (BRANCH && want_nbr) nbr_cnt[t] = a.cnt;
    if (POT) pot[i] = -a.sphi.value();
}

template <bool GRAV, bool BRANCH, bool POT, bool SOFT>
hipError_t launch_t(const SweepLaunch& L, const DevParams& P, hipStream_t s)
{
    const int blocks = (L.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((sweep_kernel<GRAV, BRANCH, POT, SOFT>), dim3(blocks), dim3(kBlock), 0, s, L.posm, L.perm, L.n,
                       L.acc, L.density, L.nbr, L.nbr_cnt, L.pot, L.kmax, L.dens ? 1 : 0, L.want_nbr ? 1 : 0,
                       L.finish ? 1 : 0, L.vel, L.xvel, L.maxvsig, P);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_sweep(const SweepLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    const bool branch = L.dens || L.want_nbr || L.want_pot;
    const bool soft = P.H2 > 0.0f;
    const int key = (L.grav ? 8 : 0) | (branch ? 4 : 0) | (L.want_pot ? 2 : 0) | (soft ? 1 : 0);
    switch (key) {
        case 0: return launch_t<false, false, false, false>(L, P, s);
        case 1: return launch_t<false, false, false, true>(L, P, s);
        case 4: return launch_t<false, true, false, false>(L, P, s);
        case 5: return launch_t<false, true, false, true>(L, P, s);
        case 6: return launch_t<false, true, true, false>(L, P, s);
        case 7: return launch_t<false, true, true, true>(L, P, s);
        case 8: return launch_t<true, false, false, false>(L, P, s);
        case 9: return launch_t<true, false, false, true>(L, P, s);
        case 12: return launch_t<true, true, false, false>(L, P, s);
        case 13: return launch_t<true, true, false, true>(L, P, s);
        case 14: return launch_t<true, true, true, false>(L, P, s);
        case 15: return launch_t<true, true, true, true>(L, P, s);
        default: return hipErrorInvalidValue;  // POT always takes the branch
    }
}

}  // namespace grasph
