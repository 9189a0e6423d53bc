// sweep_core.hpp -- the device-side core of the all-pairs sweep, shared by
// the sweep kernels (sweep.hip) and the near-field kernel of the symmetric
// sweep (sym_sweep.hip): the chunked pair loop with its neighbour-candidate
// test and rare path, the compensated folds, and the per-lane slot driver.
// Each including file gets its own copy (anonymous namespace).
#pragma once

#include "internal.hpp"

namespace grasph {
namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 8;    // j per prefetch chunk and per neighbour-candidate test
constexpr int kFold = 128;   // j per plain partial sum before the compensated add

template <int NI>
struct Acc {
    float ax[NI], ay[NI], az[NI], phi[NI];
    CSum sx[NI], sy[NI], sz[NI], sphi[NI];
    int cnt[NI];
};

// Constants of the rare path, held in VGPRs: values the compiler keeps in
// SGPRs would, under the SGPR pressure of the j buffers, be re-read from
// the kernel arguments, and that wait would also stall on the j prefetch.
struct Rare {
    int* nbr;  // the list plane
    int kmax;
    float h2;
};

__device__ __forceinline__ int vgpr_i32(int x)
{
    int r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}
__device__ __forceinline__ float vgpr_f32(float x) { return __int_as_float(vgpr_i32(__float_as_int(x))); }
__device__ __forceinline__ uint64_t vgpr_u64(uint64_t x)
{
    const uint32_t lo = (uint32_t)vgpr_i32((int)(uint32_t)x), hi = (uint32_t)vgpr_i32((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ Rare make_rare(int* nbr, int kmax, float h2)
{
    Rare r;
    r.nbr = (int*)vgpr_u64((uint64_t)nbr);
    r.kmax = vgpr_i32(kmax);
    r.h2 = vgpr_f32(h2);
    return r;
}

// Rare path: the pair may be an SPH neighbour (d2 <= cand for this lane).
// Only the exact test and the list append live here.
template <bool NBR, bool POT>
__device__ __forceinline__ void neighbour_pair(const float4 pj, const int j, const float xi, const float yi,
                                               const float zi, const float d2, float& phi, int& cnt,
                                               const int slot, const Rare& R)
{
    // rij = pi - pj, r2 = dot(rij, rij) exactly as the oracle rounds it
    const float r2 = dist2_exact(xi - pj.x, yi - pj.y, zi - pj.z);
    if (POT && r2 == 0.0f) phi -= pj.w * __builtin_amdgcn_rsqf(d2);  // the self pair was added on the fast path
    if (NBR && is_neighbour(r2, R.h2)) {
        if (cnt < R.kmax) R.nbr[list_entry(slot, cnt, R.kmax)] = j;
        ++cnt;
    }
}

// One chunk of kChunk j against the lane's NI particles: gravity (and
// potential) on every pair, then one wave-uniform candidate test.  UM (all
// partners of equal mass): the mass is factored out of the sum and applied
// at the fold, one multiply per pair fewer.
template <int NI, bool GRAV, bool NBR, bool POT, bool SOFT, bool UM>
__device__ __forceinline__ void chunk(const float4 (&pj)[kChunk], const int jb, const float (&xi)[NI],
                                      const float (&yi)[NI], const float (&zi)[NI], Acc<NI>& a, const int slot,
                                      const float H2, const float cand, const Rare& R)
{
    if constexpr (NBR && !GRAV && !POT) {
        // Lists alone: the exact r^2 on the fast path (two instructions more
        // per pair than the softened d2), so a candidate costs only the test
        // and the append -- in the near field of small dense instances
        // nearly every chunk holds one.  Same test, same pairs, same order
        // as the rare path below.
        float r2[NI][kChunk];
#pragma unroll
        for (int c = 0; c < kChunk; ++c)
#pragma unroll
            for (int q = 0; q < NI; ++q) r2[q][c] = dist2_exact(xi[q] - pj[c].x, yi[q] - pj[c].y, zi[q] - pj[c].z);
        float mn = r2[0][0];
#pragma unroll
        for (int q = 0; q < NI; ++q)
#pragma unroll
            for (int c = 0; c < kChunk; ++c) mn = fminf(mn, r2[q][c]);
        if (mn <= R.h2) {
#pragma unroll
            for (int c = 0; c < kChunk; ++c)
#pragma unroll
                for (int q = 0; q < NI; ++q)
                    if (is_neighbour(r2[q][c], R.h2)) {
                        if (a.cnt[q] < R.kmax) R.nbr[list_entry(slot + 64 * q, a.cnt[q], R.kmax)] = jb + c;
                        ++a.cnt[q];
                    }
        }
        return;
    }
    constexpr bool BRANCH = NBR || (POT && SOFT);
    float d2[NI][kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; ++c) {
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const float dx = pj[c].x - xi[q], dy = pj[c].y - yi[q], dz = pj[c].z - zi[q];
            if (SOFT)
                d2[q][c] = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, H2)));
            else
                d2[q][c] = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            float inv = 0.0f;
            if (GRAV || POT) inv = __builtin_amdgcn_rsqf(d2[q][c]);
            if (GRAV) {
                float s;
                if (UM) {
                    s = inv * inv;
                    s = s * inv;
                } else {
                    s = pj[c].w * inv;
                    s = s * (inv * inv);
                }
                if (!SOFT) s = (d2[q][c] > 0.0f) ? s : 0.0f;  // r^2 > 0 (computeDerivatives.h:226)
                a.ax[q] = fmaf(dx, s, a.ax[q]);
                a.ay[q] = fmaf(dy, s, a.ay[q]);
                a.az[q] = fmaf(dz, s, a.az[q]);
            }
            if (POT) a.phi[q] += SOFT ? pj[c].w * inv : ((d2[q][c] > 0.0f) ? pj[c].w * inv : 0.0f);
        }
    }
    if (BRANCH) {
        float mn = d2[0][0];
#pragma unroll
        for (int q = 0; q < NI; ++q)
#pragma unroll
            for (int c = 0; c < kChunk; ++c) mn = fminf(mn, d2[q][c]);
        if (mn <= cand) {
#pragma unroll
            for (int c = 0; c < kChunk; ++c)
#pragma unroll
                for (int q = 0; q < NI; ++q)
                    if (d2[q][c] <= cand)
                        neighbour_pair<NBR, POT && SOFT>(pj[c], jb + c, xi[q], yi[q], zi[q], d2[q][c], a.phi[q],
                                                         a.cnt[q], slot + 64 * q, R);
        }
    }
}

template <int NI, bool GRAV, bool POT, bool UM>
__device__ __forceinline__ void fold(Acc<NI>& a, const float m)
{
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        if (GRAV) {
            if (UM) {
                a.sx[q].add(a.ax[q] * m); a.sy[q].add(a.ay[q] * m); a.sz[q].add(a.az[q] * m);
            } else {
                a.sx[q].add(a.ax[q]); a.sy[q].add(a.ay[q]); a.sz[q].add(a.az[q]);
            }
            a.ax[q] = a.ay[q] = a.az[q] = 0.0f;
        }
        if (POT) {
            a.sphi[q].add(a.phi[q]);
            a.phi[q] = 0.0f;
        }
    }
}

// Sweep of the lane's NI slots against j in [j0, j1); m is the common
// partner mass when UM.
template <int NI, bool GRAV, bool NBR, bool POT, bool SOFT, bool UM>
__device__ __forceinline__ void sweep_range(const float4* __restrict__ posm, const int j0, const int j1,
                                            const float (&xi)[NI], const float (&yi)[NI], const float (&zi)[NI],
                                            Acc<NI>& a, const int slot, const float H2, const float cand,
                                            const Rare& R, const float m)
{
    // Two SGPR buffers in ping-pong: while one chunk is computed the next one
    // loads into the other buffer, so no register copies are needed.
    constexpr int kFoldChunks = kFold / kChunk;
    const int nchunks = (j1 - j0) / kChunk;
    const int nfull = j0 + nchunks * kChunk;
    float4 A[kChunk], B[kChunk];
    if (nchunks > 0) {
#pragma unroll
        for (int c = 0; c < kChunk; ++c) A[c] = posm[j0 + c];
    }
    int k = 0;
    for (; k + 1 < nchunks; k += 2) {
        const int jb = j0 + k * kChunk;
#pragma unroll
        for (int c = 0; c < kChunk; ++c) B[c] = posm[jb + kChunk + c];
        chunk<NI, GRAV, NBR, POT, SOFT, UM>(A, jb, xi, yi, zi, a, slot, H2, cand, R);
        if ((k % kFoldChunks) == kFoldChunks - 1) fold<NI, GRAV, POT, UM>(a, m);
        // next A: chunk k + 2, or chunk k + 1 again when k + 2 is past the end
        const int ja = j0 + ((k + 2 < nchunks) ? (k + 2) : (k + 1)) * kChunk;
#pragma unroll
        for (int c = 0; c < kChunk; ++c) A[c] = posm[ja + c];
        chunk<NI, GRAV, NBR, POT, SOFT, UM>(B, jb + kChunk, xi, yi, zi, a, slot, H2, cand, R);
        if (((k + 1) % kFoldChunks) == kFoldChunks - 1) fold<NI, GRAV, POT, UM>(a, m);
    }
    if (k < nchunks) chunk<NI, GRAV, NBR, POT, SOFT, UM>(A, j0 + k * kChunk, xi, yi, zi, a, slot, H2, cand, R);
    // tail: pad the last partial chunk with far-away massless partners
    if (nfull < j1) {
        float4 t[kChunk];
#pragma unroll
        for (int c = 0; c < kChunk; ++c)
            t[c] = (nfull + c < j1) ? posm[nfull + c] : make_float4(3.0e18f, 3.0e18f, 3.0e18f, 0.0f);
        chunk<NI, GRAV, NBR, POT, SOFT, UM>(t, nfull, xi, yi, zi, a, slot, H2, cand, R);
    }
    fold<NI, GRAV, POT, UM>(a, m);
}

// Load the lane's particles (slots s0 + 64 q below s_end), run the sweep
// over [j0, j1) and store the results -- or, with part set (one slice of a
// j-split sweep), the compensated partial sums (sum, compensation) of the
// slice, combined later in slice order.
template <int NI, bool GRAV, bool NBR, bool POT, bool SOFT, bool UM>
__device__ __forceinline__ void sweep_slots(const float4* __restrict__ xs, const int* __restrict__ perm,
                                            const int s0, const int s_end, const int j0, const int j1,
                                            float4* __restrict__ acc, int* __restrict__ nbr_cnt,
                                            float* __restrict__ pot, const int finish,
                                            const float4* __restrict__ vel, float4* __restrict__ xvel,
                                            float* __restrict__ maxvsig, const DevParams& P, const float H2,
                                            const float cand, const Rare& R, const float m,
                                            float4* __restrict__ part = nullptr)
{
    float xi[NI], yi[NI], zi[NI];
    int ii[NI];
    bool act[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int t = s0 + 64 * q;
        act[q] = t < s_end;
        ii[q] = act[q] ? (perm ? perm[t] : t) : 0;  // the particle of slot t
        xi[q] = yi[q] = zi[q] = __int_as_float(0x7fc00000);  // NaN: inactive slots never branch
        if (act[q]) {
            const float4 p = xs[t];
            xi[q] = p.x; yi[q] = p.y; zi[q] = p.z;
        }
    }
    Acc<NI> a;
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        a.ax[q] = a.ay[q] = a.az[q] = a.phi[q] = 0.0f;
        a.sx[q] = a.sy[q] = a.sz[q] = a.sphi[q] = CSum{0.0f, 0.0f};
        a.cnt[q] = 0;
    }
    sweep_range<NI, GRAV, NBR, POT, SOFT, UM>(xs, j0, j1, xi, yi, zi, a, s0, H2, cand, R, m);
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        if (!act[q]) continue;
        const int i = ii[q];
        if (part) {
            if (GRAV) {
                part[2 * (size_t)i] = make_float4(a.sx[q].s, a.sy[q].s, a.sz[q].s, 0.0f);
                part[2 * (size_t)i + 1] = make_float4(a.sx[q].c, a.sy[q].c, a.sz[q].c, 0.0f);
            }
            if (NBR) nbr_cnt[s0 + 64 * q] = a.cnt[q];
            continue;
        }
        if (GRAV || (finish && acc)) {
            float ax = GRAV ? a.sx[q].value() : 0.0f, ay = GRAV ? a.sy[q].value() : 0.0f;
            float az = GRAV ? a.sz[q].value() : 0.0f;
            if (finish) {
                environment_acceleration(P, ax, ay, az);
            }
            acc[i] = make_float4(ax, ay, az, 0.0f);
        }
        if (finish) {
            if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (maxvsig) maxvsig[i] = 0.0f;
        }
        if (NBR) nbr_cnt[s0 + 64 * q] = a.cnt[q];
        if (POT) pot[i] = -a.sphi[q].value();
    }
}

}  // namespace
}  // namespace grasph
