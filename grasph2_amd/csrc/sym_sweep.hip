// sym_sweep.hip -- the single-instance all-pairs sweep evaluated once per
// unordered pair (gfx950).
//
// Same results as the sweep of sweep.hip (softened gravity of
// calcAcceleration, src/computeDerivatives.h:229-235, plus the neighbour
// lists of :72,122,238), organised by groups of 1024 consecutive slots:
//  - near field, one wavefront per block of 64 slots, partner blocks in
//    increasing slot order through the candidate sweep of sweep_core.hpp:
//    the blocks of its own group with gravity, and (when lists are wanted)
//    every other block whose bounding box comes within H of its own,
//    distance test only -- every SPH neighbour of a slot lies in such a
//    block (the box distance, rounded like the pair distance, never exceeds
//    it), so the lists come out complete and in slot order, exactly as the
//    one-range sweep writes them;
//  - far field: every pair of slots in different groups is computed once,
//    for both sides (Newton's third law).  One wavefront owns a group (16
//    slots per lane).  The later groups' particles come 16 at a time, one
//    per lane of each 16-lane DPP row; rotation k (row_ror:k) pairs a lane's
//    slots with the partner k lanes away in its row.  A lane sums its own
//    slots' gravity in registers, and the partner side of each rotation
//    over its 16 slots in one accumulator per rotation, so nothing crosses
//    lanes until the 16 x 16 x 16 pairs of the chunk are done; 15 DPP adds
//    per component then gather each partner's rotations in its own lane,
//    two swaps add the four rows, and the values are stored as the
//    (group, partner) partial straight away;
//  - a combine kernel adds, per slot and in a fixed order, the near sum,
//    the far own-side fragments and the far partner-side partials
//    (compensated), then finishes the acceleration as the one-range sweep.
// Per far pair: 3 sub, 3 FMA for r^2 + H^2, one v_rsq_f32, 2 mul, 3 FMA for
// the own side and 3 FMA for the partner side (2 more mul with unequal
// masses), plus, per chunk and lane, 45 DPP moves of the partners' x, y, z
// (a DPP move with several uses is not folded into them) and 51 for the
// reduction, over 256 pairs: ~15.4 VALU instructions per wavefront and
// partner for two interactions per lane and slot, against ~12 for one
// interaction in the one-sided sweep.
// Every sum is in a fixed order, so results do not vary from run to run.
#include <algorithm>
#include <utility>
#include <vector>

#include "sweep_core.hpp"
#include "sym_geometry.hpp"

namespace grasph {
namespace {

constexpr int kFarThreads = 64;  // one wavefront owns a group
static_assert(kGroupSB == 1 && kSB == kGroup, "the far field assumes one wavefront per group");

// ------------------------------------------------------------------ cross-lane reduction

__device__ __forceinline__ float swap32_sum(float p, float q)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(p), __float_as_uint(q), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap16_sum(float p, float q)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(p), __float_as_uint(q), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// ------------------------------------------------------------------ kernels

// Bounding box of every block of 64 slots (one wavefront per block).
__global__ __launch_bounds__(kBlock) void block_box_kernel(const float4* __restrict__ xs, const int n, const int nb,
                                                           float4* __restrict__ box)
{
    const int blk = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (blk >= nb) return;  // whole wavefronts leave together
    const int t = 64 * blk + lane;
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (t < n) {
        const float4 p = xs[t];
        v[0] = v[3] = p.x; v[1] = v[4] = p.y; v[2] = v[5] = p.z;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = fminf(v[c], __shfl_xor(v[c], o, 64));
            v[c + 3] = fmaxf(v[c + 3], __shfl_xor(v[c + 3], o, 64));
        }
    }
    if (lane == 0) {
        box[2 * blk] = make_float4(v[0], v[1], v[2], 0.0f);
        box[2 * blk + 1] = make_float4(v[3], v[4], v[5], 0.0f);
    }
}

// near[I][J] (bit J % 32 of word I * nwords + J / 32) for blocks of 64:
// boxes within H.  The per-axis gaps and their squared sum round like the
// exact pair distance, so a pair with r^2 <= H^2 never sits in two blocks
// whose box distance exceeds H^2.  Symmetric by construction.
__global__ __launch_bounds__(kBlock) void near_bits_kernel(const float4* __restrict__ box, const int nb,
                                                           const int nwords, const float H2,
                                                           uint32_t* __restrict__ bits)
{
    const int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (idx >= (int64_t)nb * nwords) return;
    const int I = (int)(idx / nwords), w = (int)(idx - (int64_t)I * nwords);
    const float4 lo = box[2 * I], hi = box[2 * I + 1];
    uint32_t word = 0;
    for (int b = 0; b < 32; ++b) {
        const int J = 32 * w + b;
        if (J >= nb) break;
        const float4 l2 = box[2 * J], h2 = box[2 * J + 1];
        const float gx = fmaxf(0.0f, fmaxf(l2.x - hi.x, lo.x - h2.x));
        const float gy = fmaxf(0.0f, fmaxf(l2.y - hi.y, lo.y - h2.y));
        const float gz = fmaxf(0.0f, fmaxf(l2.z - hi.z, lo.z - h2.z));
        if (dist2_exact(gx, gy, gz) <= H2) word |= 1u << b;
    }
    bits[idx] = word;
}

// Near field of the 64 slots of block I of the instance whose slots are
// [base, base + n), share `sl` of `ns`: the partner
// blocks are the blocks of its group (gravity, and the distance test when
// lists are wanted) and, with bits, the other blocks near it (distance test
// only); in increasing order they are cut into ns runs of equal count, and
// this wavefront takes run sl.  A block of the distance test only is swept
// only if some slot of block I comes within H of its box (box: the
// instance's block boxes): the point-to-box gaps and their squared sum
// round like the exact pair distance (lo <= x <= hi, rounding is monotonic),
// so they never exceed the distance to a particle in the box, and a skipped
// block holds no neighbour.  Blocks whose own box is large (a jump of the
// slot order) skip most of their box neighbours this way.  Writes the slot's compensated gravity
// partial (slot-indexed, two float4) and its list and count -- all in the
// planes of share sl, which later fold in share order.
template <bool GRAV, bool NBR, bool SOFT, bool UM>
__device__ __forceinline__ void near_blocks(const float4* __restrict__ xs, const int base, const int n, const int nb,
                                            const int nwords, const uint32_t* __restrict__ bits,
                                            const float4* __restrict__ box, const int I,
                                            const int sl, const int ns, const int lane, int* __restrict__ nbr_cnt,
                                            float4* __restrict__ near_part, const DevParams& P, const Rare& R,
                                            const float m)
{
    constexpr int kGroupBlocks = kGroup / 64;  // 16
    const int t = base + 64 * I + lane;  // slots of the instance: [base, base + n)
    const bool act = 64 * I + lane < n;
    float xi[1], yi[1], zi[1];
    xi[0] = yi[0] = zi[0] = __int_as_float(0x7fc00000);  // NaN: inactive slots never branch
    if (act) {
        const float4 p = xs[t];
        xi[0] = p.x; yi[0] = p.y; zi[0] = p.z;
    }
    Acc<1> a;
    a.ax[0] = a.ay[0] = a.az[0] = a.phi[0] = 0.0f;
    a.sx[0] = a.sy[0] = a.sz[0] = a.sphi[0] = CSum{0.0f, 0.0f};
    a.cnt[0] = 0;
    const int G0 = (I / kGroupBlocks) * kGroupBlocks, G1 = min(nb, G0 + kGroupBlocks);  // own group's blocks
    const bool use_bits = NBR && bits;
    const uint32_t* row = use_bits ? bits + (size_t)I * nwords : nullptr;
    // the own group's blocks lie in one word (groups of 16 blocks are 16-aligned)
    const int gw = G0 >> 5;
    const uint32_t gmask = (uint32_t)(((1ull << (G1 - G0)) - 1) << (G0 - 32 * gw));
    // the row through the constant address space: a uniform address there
    // becomes a scalar load (a global pointer gives a vector load and a wait
    // per word)
    const __attribute__((address_space(4))) uint32_t* crow = (const __attribute__((address_space(4))) uint32_t*)row;
    auto word_of = [&](int w) -> uint32_t {
        uint32_t word = use_bits ? crow[w] : 0u;
        if ((GRAV || !use_bits) && w == gw) word |= gmask;
        return word;
    };
    // the boxes through the constant address space too (scalar loads)
    const __attribute__((address_space(4))) float* cbox = (const __attribute__((address_space(4))) float*)box;
    auto lane_near = [&](int J) -> bool {
        const __attribute__((address_space(4))) float* c = cbox + 8 * J;  // lo.xyzw, hi.xyzw
        const float gx = fmaxf(0.0f, fmaxf(c[0] - xi[0], xi[0] - c[4]));
        const float gy = fmaxf(0.0f, fmaxf(c[1] - yi[0], yi[0] - c[5]));
        const float gz = fmaxf(0.0f, fmaxf(c[2] - zi[0], zi[0] - c[6]));
        return __builtin_amdgcn_ballot_w64(act && dist2_exact(gx, gy, gz) <= P.H2) != 0;
    };
    const int w0 = use_bits ? 0 : gw, w1 = use_bits ? nwords : gw + 1;
    int total = 0;
#pragma unroll 8
    for (int w = w0; w < w1; ++w) total += __builtin_popcount(word_of(w));
    const int r0 = (int)((int64_t)sl * total / ns), r1 = (int)((int64_t)(sl + 1) * total / ns);
    // blocks of this share, consecutive ones of one kind (own group or not) swept as one range
    int rank = 0, run0 = -1, run1 = -1;
    bool run_grav = false;
    auto flush = [&]() {
        if (run0 < 0) return;
        const int j0 = base + 64 * run0, j1 = base + min(n, 64 * run1);
        if (GRAV && run_grav)
            sweep_range<1, GRAV, NBR, false, SOFT, UM>(xs, j0, j1, xi, yi, zi, a, t, P.H2, P.cand, R, m);
        else
            sweep_range<1, false, NBR, false, SOFT, UM>(xs, j0, j1, xi, yi, zi, a, t, P.H2, P.cand, R, m);
        run0 = -1;
    };
    for (int w = w0; w < w1 && rank < r1; ++w) {
        uint32_t word = word_of(w);
        const int c = __builtin_popcount(word);
        if (rank + c <= r0) {
            rank += c;
            continue;
        }
        while (word && rank < r1) {
            const int J = 32 * w + __builtin_ctz(word);
            word &= word - 1;
            if (rank++ < r0) continue;
            const bool g = J >= G0 && J < G1;
            if (use_bits && !(GRAV && g) && !lane_near(J)) continue;
            if (run0 >= 0 && (J != run1 || g != run_grav)) flush();
            if (run0 < 0) {
                run0 = J;
                run_grav = g;
            }
            run1 = J + 1;
        }
    }
    flush();
    if (!act) return;
    if (GRAV) {
        near_part[2 * (size_t)t] = make_float4(a.sx[0].s, a.sy[0].s, a.sz[0].s, 0.0f);
        near_part[2 * (size_t)t + 1] = make_float4(a.sx[0].c, a.sy[0].c, a.sz[0].c, 0.0f);
    }
    if (NBR) nbr_cnt[t] = a.cnt[0];
}

template <bool GRAV, bool NBR, bool SOFT>
__global__ __launch_bounds__(kBlock) void near_kernel(const float4* __restrict__ xs, const int n, const int nb,
                                                      const int nwords, const uint32_t* __restrict__ bits,
                                                      const float4* __restrict__ box, int* __restrict__ nbr,
                                                      int* __restrict__ nbr_cnt,
                                                      const int kmax, float4* __restrict__ near_part,
                                                      const DevParams P, const float* __restrict__ umass,
                                                      const int prio)
{
    // issue priority over the far field's wavefronts, which run beside this
    // kernel on the other stream (see the list_prio option of SymLaunch)
    if (prio) __builtin_amdgcn_s_setprio(3);
    const int I = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * kBlock + threadIdx.x) >> 6));
    const int lane = threadIdx.x & 63;
    if (I >= nb) return;
    // share blockIdx.y of gridDim.y: its list, count and partial planes
    const int sl = blockIdx.y, ns = gridDim.y;
    const Rare R = make_rare(nbr + (size_t)sl * list_plane(n, kmax), kmax, P.H2);
    int* const cnt = nbr_cnt + (size_t)sl * n;
    float4* const part = near_part + 2 * (size_t)sl * n;
    if constexpr (GRAV) {
        if (umass && umass[0] != 0.0f) {
            near_blocks<GRAV, NBR, SOFT, true>(xs, 0, n, nb, nwords, bits, box, I, sl, ns, lane, cnt, part, P, R, umass[1]);
            return;
        }
    }
    near_blocks<GRAV, NBR, SOFT, false>(xs, 0, n, nb, nwords, bits, box, I, sl, ns, lane, cnt, part, P, R, 1.0f);
}

// Own slots of a lane in the far field: position, mass and the plain
// gravity sums of the current unit.  At the end of a unit those go into
// the wavefront's LDS sums (lsum[1024 c + slot], component c: plain sums of
// at most kFoldsPerFlush units), and every kFoldsPerFlush units and at the
// end of the row the LDS sums go into the row's own-side fragment in global
// memory (float4 sums at frag[frag_elem(slot)], their corrections 64
// further), compensated.  Neither level takes registers from the pair loop,
// and the fragment is written about once per row.
struct FarOwn {
    float x[kNI], y[kNI], z[kNI], m[kNI];
    float ax[kNI], ay[kNI], az[kNI];
    float* lsum;
    float4* frag;
    int folds;   // units in lsum since the last flush (0: lsum holds nothing)
    bool fresh;  // nothing flushed into the fragment yet: the first flush writes it
    bool plain;  // every fragment is flushed once: no correction planes
};

// ------------------------------------------------------------------ rotated far segment
// The partners of a chunk sit one per lane of a DPP row (16 partners, the
// same 16 in all four rows).  Rotation k pairs lane l's slots with the
// partner of row lane (l - k) & 15: DPP row_ror:k reads it straight into the
// subtraction, so no instruction moves partner data.  The partner side of
// rotation k is summed over the lane's slots into its own accumulator
// (P[k]: lane l holds partner (l - k) & 15's sum), so nothing crosses lanes
// until the chunk's 16 x 16 x 4 pairs are done; then 15 DPP adds per
// component bring partner p's rotations to row lane p (row_ror:16-k reads
// lane (p + k) & 15), and two swaps add the four rows.  Per chunk and lane:
// 64 pairs at 15 instructions, plus 51 for the reduction (0.8 per pair,
// against 54 / 32 for a reduction of 8 broadcast partners).

template <int K>
__device__ __forceinline__ float ror(float v)
{
    if constexpr (K == 0)
        return v;
    else
        return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x120 + K, 0xf, 0xf, true));
}

constexpr int kRow = 16;  // partners per rotated chunk (one DPP row)

// Rotation K of a chunk: the lane's kNI slots against partner (l - K) & 15
// of its row; own side into o.a*, partner side into P[K].
template <int K, bool UM, bool SOFT>
__device__ __forceinline__ void far_rotation(const float4 pj, FarOwn& o, const float H2, float (&P)[kRow][3])
{
    const float jx = ror<K>(pj.x), jy = ror<K>(pj.y), jz = ror<K>(pj.z);
    const float jm = UM ? 1.0f : ror<K>(pj.w);
#pragma unroll
    for (int q = 0; q < kNI; ++q) {
        const float dx = jx - o.x[q], dy = jy - o.y[q], dz = jz - o.z[q];
        const float d2 = SOFT ? fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, H2))) : fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        const float inv = __builtin_amdgcn_rsqf(d2);
        float s = inv * inv;
        s = s * inv;
        if (!SOFT) s = (d2 > 0.0f) ? s : 0.0f;  // coincident particles exert nothing (computeDerivatives.h:226)
        float sa = s, sb = s;
        if (!UM) {
            sa = jm * s;
            sb = s * o.m[q];
        }
        o.ax[q] = fmaf(dx, sa, o.ax[q]); o.ay[q] = fmaf(dy, sa, o.ay[q]); o.az[q] = fmaf(dz, sa, o.az[q]);
        if (q == 0) {
            P[K][0] = dx * sb; P[K][1] = dy * sb; P[K][2] = dz * sb;
        } else {
            P[K][0] = fmaf(dx, sb, P[K][0]); P[K][1] = fmaf(dy, sb, P[K][1]); P[K][2] = fmaf(dz, sb, P[K][2]);
        }
    }
}

// partner p's rotations to row lane p: P[K] of lane (p + K) & 15 holds them
template <int K>
__device__ __forceinline__ void fold_rotation(const float (&P)[kRow][3], float (&T)[3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) T[c] = ror<kRow - K>(P[K][c]) + T[c];
}

template <bool UM, bool SOFT, int... K>
__device__ __forceinline__ void far_rotations(const float4 pj, FarOwn& o, const float H2, float (&P)[kRow][3],
                                              std::integer_sequence<int, K...>)
{
    (far_rotation<K, UM, SOFT>(pj, o, H2, P), ...);
}
template <int... K>
__device__ __forceinline__ void fold_rotations(const float (&P)[kRow][3], float (&T)[3],
                                               std::integer_sequence<int, K...>)
{
    (fold_rotation<K + 1>(P, T), ...);
}

// One chunk of 16 partners (pj: row lane l holds partner l) against the
// lane's kNI slots.  Own side into o.a*; returns the partner side: lanes
// 0-15 hold partner l's x, lanes 16-31 partner l - 16's z, lanes 32-47
// partner l - 32's y (UM: without the common mass; else with the slots' masses).
template <bool UM, bool SOFT>
__device__ __forceinline__ float far_chunk_rot(const float4 pj, FarOwn& o, const float H2)
{
    float P[kRow][3];
    far_rotations<UM, SOFT>(pj, o, H2, P, std::make_integer_sequence<int, kRow>{});
    float T[3] = {P[0][0], P[0][1], P[0][2]};
    fold_rotations(P, T, std::make_integer_sequence<int, kRow - 1>{});
    // rows 0 + 2 and 1 + 3: x to lanes 0-31, y to lanes 32-63; z everywhere
    const float xy = swap32_sum(T[0], T[1]);
    const float zz = swap32_sum(T[2], T[2]);
    // row 0: x, row 1: z, row 2: y (row 3: z again)
    return swap16_sum(xy, zz);
}

// partner b + (lane & 15) of an instance of n particles (clamped to n - 1:
// a look-ahead past the end loads values nobody uses).  UM loads x y z
// alone: a loaded register left unused would be reused at once, and that
// write would wait for the load, a chunk too early.
template <bool UM>
__device__ __forceinline__ float4 load_partner_row(const float4* __restrict__ xs, const int b, const int n,
                                                   const int lane)
{
    const int jc = min(b + (lane & (kRow - 1)), n - 1);
    if (UM) {
        const float3 p = *reinterpret_cast<const float3*>(xs + jc);
        return make_float4(p.x, p.y, p.z, 0.0f);
    }
    return xs[jc];
}
// BOUNDED: partners at or past n are far away and massless (applied where
// the chunk is used, so the load's wait stays there)
template <bool BOUNDED>
__device__ __forceinline__ float4 bound_partner_row(const float4 v, const int b, const int n, const int lane)
{
    if (BOUNDED && b + (lane & (kRow - 1)) >= n) return make_float4(3.0e18f, 3.0e18f, 3.0e18f, 0.0f);
    return v;
}

// component of the partner side this lane holds after far_chunk_rot: row 0
// x, row 2 y, rows 1 and 3 z (the same value, bit for bit, stored to the
// same place: every lane stores, so an unbounded chunk's store needs no
// branch, behind which the compiler would sink the chunk's own-side sums)
__device__ __forceinline__ int partner_comp_rot(const int lane)
{
    const int row = lane >> 4;
    return row == 0 ? 0 : row == 2 ? 1 : 2;
}

// Partners [b0, b0 + 16 nchunks) of group J against the wavefront's slots;
// each chunk's partner side (summed over the whole group: one wavefront
// owns it) goes straight to row g's partials bp (bp[3 p + c]: partner p of
// group J, component c; UM: times the common mass m).  The store has no
// condition: partners past n (only in the instance's last group, whose
// partials are allocated whole) get values nobody reads, and a store behind
// a branch makes the compiler spill the chunk's registers.
template <bool UM, bool SOFT, bool BOUNDED>
__device__ __forceinline__ void far_segment_rot(const float4* __restrict__ xs, const int n, const int b0,
                                                const int nchunks, const int J, FarOwn& o, const float H2,
                                                const int lane, float* __restrict__ bp, const float m)
{
    const int comp = partner_comp_rot(lane);
    const int p0 = b0 - kGroup * J + (lane & (kRow - 1));  // this lane's partner within group J, chunk 0
    float4 A = load_partner_row<UM>(xs, b0, n, lane);
    for (int q = 0; q < nchunks; ++q) {
        const float4 B = load_partner_row<UM>(xs, b0 + kRow * (q + 1), n, lane);  // unused after the last
        const float v = far_chunk_rot<UM, SOFT>(bound_partner_row<BOUNDED>(A, b0 + kRow * q, n, lane), o, H2);
        const int p = p0 + kRow * q;
        bp[3 * p + comp] = UM ? v * m : v;  // p past n: in the last group's partials, never read
        A = B;
    }
}

// The own slots of row g (slots 1024 g + 64 q + lane, q < kNI) of an
// instance of n particles at xs (slots past n sit far away and massless),
// the wavefront's LDS sums lsum, and the row's fragment of workgroup k:
// frag (fresh until the first flush).
__device__ __forceinline__ void far_load_row(const float4* __restrict__ xs, const int n, const int64_t g,
                                             const int lane, float* lsum, float4* frag, const bool plain, FarOwn& o)
{
#pragma unroll
    for (int q = 0; q < kNI; ++q) {
        const int t = kSB * (int)g + 64 * q + lane;
        o.x[q] = o.y[q] = o.z[q] = -3.0e18f;
        o.m[q] = 0.0f;
        if (t < n) {
            const float4 p = xs[t];
            o.x[q] = p.x; o.y[q] = p.y; o.z[q] = p.z; o.m[q] = p.w;
        }
        o.ax[q] = o.ay[q] = o.az[q] = 0.0f;
    }
    o.lsum = lsum;
    o.frag = frag;
    o.folds = 0;
    o.fresh = true;
    o.plain = plain;
}

// The lane's slot of own slot q, through an empty asm statement: indexed
// by it, the LDS sums and the fragment are never kept in registers across
// a row's units by the compiler (96 more live values, which the pair loop
// cannot spare).
__device__ __forceinline__ int own_lane(const int lane)
{
    int l = lane;
    __asm__ volatile("" : "+v"(l));
    return l;
}

// The LDS sums of the lane's slots added (compensated) into the row's
// fragment; the first flush stores them (and zero corrections unless the
// fragments are plain).
__device__ __forceinline__ void far_flush_own(FarOwn& o, const int lane)
{
    if (o.folds == 0) return;
    const int l = own_lane(lane);
    float4* const fr = o.frag;
#pragma unroll
    for (int q = 0; q < kNI; ++q) {
        const int slot = 64 * q + l, e = 128 * q + l;  // e: frag_elem(slot)
        float4 sv = make_float4(0.f, 0.f, 0.f, 0.f), cv = sv;
        if (!o.fresh) {
            sv = fr[e];
            cv = fr[e + 64];
        }
        CSum x{sv.x, cv.x}, y{sv.y, cv.y}, z{sv.z, cv.z};
        x.add(o.lsum[slot]);
        y.add(o.lsum[kGroup + slot]);
        z.add(o.lsum[2 * kGroup + slot]);
        fr[e] = make_float4(x.s, y.s, z.s, 0.0f);
        if (!o.plain) fr[e + 64] = make_float4(x.c, y.c, z.c, 0.0f);
    }
    o.fresh = false;
    o.folds = 0;
}

// The lane's plain own-side sums of a unit added to its LDS sums, then
// cleared; every kFoldsPerFlush units the LDS sums go to the fragment.
template <bool UM>
__device__ __forceinline__ void far_fold_own(FarOwn& o, const float m, const int lane)
{
    const int l = own_lane(lane);
    const bool first = o.folds == 0;
#pragma unroll
    for (int q = 0; q < kNI; ++q) {
        const int slot = 64 * q + l;
        const float ax = UM ? o.ax[q] * m : o.ax[q];
        const float ay = UM ? o.ay[q] * m : o.ay[q];
        const float az = UM ? o.az[q] * m : o.az[q];
        o.lsum[slot] = first ? ax : o.lsum[slot] + ax;
        o.lsum[kGroup + slot] = first ? ay : o.lsum[kGroup + slot] + ay;
        o.lsum[2 * kGroup + slot] = first ? az : o.lsum[2 * kGroup + slot] + az;
        o.ax[q] = o.ay[q] = o.az[q] = 0.0f;
    }
    if (++o.folds == kFoldsPerFlush) far_flush_own(o, lane);
}

// One far unit of row g: piece `piece` of Q (kSB / Q partners, Q >= 2) of
// partner group J against the wavefront's 1024 slots; the partials of group
// J start at element 3 (bpart_group_base(J) - boff) of bpart (boff: the
// groups before the column batch).  The partner side is
// stored per chunk as row g's partials; the own side is folded at the end,
// so a slot adds at most 512 terms in plain FP32 before they join its LDS
// sum (and at most kFoldsPerFlush such sums before they join the
// compensated fragment).
template <bool UM, bool SOFT, int QC = 0>
__device__ __forceinline__ void far_unit(const float4* __restrict__ xs, const int n, const int64_t g, const int J,
                                         const int piece, const int Qr, FarOwn& o, float* __restrict__ bpart,
                                         const int64_t boff, const float H2, const float m, const int lane)
{
    const int Q = QC ? QC : Qr;  // QC: the split known at compile time (2 for the largest counts)
    const int len = kSB / Q, b0 = kSB * J + len * piece, nchunks = len / kRow;
    float* const bp = bpart + 3 * (bpart_group_base(J) - boff + g * kGroup);
    if (b0 + len > n)
        far_segment_rot<UM, SOFT, true>(xs, n, b0, nchunks, J, o, H2, lane, bp, m);
    else
        far_segment_rot<UM, SOFT, false>(xs, n, b0, nchunks, J, o, H2, lane, bp, m);
    far_fold_own<UM>(o, m, lane);
}

// Far field of one instance, or (COL) column batch [c0, c1) of its partner
// groups.  Workgroup k of W takes the units [k U / W, (k + 1) U / W) of the
// rows g (group g against its later groups, or those in the batch, each cut
// into Q pieces) laid end to end; its own-side sums of row g go to fragment
// g + k (plain: no workgroup has more than kFoldsPerFlush units).  The
// whole far field keeps a kernel of its own (far_kernel), so the row
// arithmetic of the batches does not change its code.
template <bool UM, bool SOFT, int QC, bool COL>
__device__ __forceinline__ void far_work(const float4* __restrict__ xs, const int n, const int64_t nsb, const int c0,
                                         const int c1, const int64_t U, const int W, const int Qr,
                                         float4* __restrict__ apart, const bool plain, float* __restrict__ bpart,
                                         const float H2, const float m, float* lsum)
{
    const int Q = QC ? QC : Qr;
    const int k = blockIdx.x, lane = threadIdx.x;
    const int64_t u0 = (int64_t)k * U / W, u1 = (int64_t)(k + 1) * U / W;
    if (u0 >= u1) return;
    const int64_t boff = COL ? bpart_group_base(c0) : 0;
    int64_t g = COL ? col_row_of_unit(u0, c0, c1, Q) : row_of_unit(u0, nsb, Q);
    int64_t s = u0 - (COL ? col_row_prefix(g, c0, c1, Q) : row_prefix(g, nsb, Q));
    FarOwn o;
    far_load_row(xs, n, g, lane, lsum, apart + (size_t)kFragElems * (g + k), plain, o);
    for (int64_t u = u0; u < u1; ++u) {
        if (s == (COL ? col_row_units(g, c0, c1, Q) : row_units(g, nsb, Q))) {
            far_flush_own(o, lane);
            ++g;
            s = 0;
            far_load_row(xs, n, g, lane, lsum, apart + (size_t)kFragElems * (g + k), plain, o);
        }
        const int J = (int)((COL ? col_first_group(g, c0) : g + 1) + s / Q);
        far_unit<UM, SOFT, QC>(xs, n, g, J, (int)(s % Q), Q, o, bpart, boff, H2, m, lane);
        ++s;
    }
    far_flush_own(o, lane);
}

template <bool SOFT, bool COL>
__device__ __forceinline__ void far_dispatch(const float4* __restrict__ xs, const int n, const int64_t nsb,
                                             const int c0, const int c1, const int64_t U, const int W, const int Q,
                                             float4* __restrict__ apart, const int plain,
                                             float* __restrict__ bpart, const float H2,
                                             const float* __restrict__ umass, float* lsum)
{
    const bool um = umass[0] != 0.0f;
    if (Q == 2) {
        if (um)
            far_work<true, SOFT, 2, COL>(xs, n, nsb, c0, c1, U, W, 2, apart, plain, bpart, H2, umass[1], lsum);
        else
            far_work<false, SOFT, 2, COL>(xs, n, nsb, c0, c1, U, W, 2, apart, plain, bpart, H2, 1.0f, lsum);
    } else {
        if (um)
            far_work<true, SOFT, 0, COL>(xs, n, nsb, c0, c1, U, W, Q, apart, plain, bpart, H2, umass[1], lsum);
        else
            far_work<false, SOFT, 0, COL>(xs, n, nsb, c0, c1, U, W, Q, apart, plain, bpart, H2, 1.0f, lsum);
    }
}

template <bool SOFT>
__global__ __launch_bounds__(kFarThreads, 2) void far_kernel(const float4* __restrict__ xs, const int n,
                                                            const int64_t nsb, const int64_t U, const int W,
                                                            const int Q, float4* __restrict__ apart,
                                                            const int plain, float* __restrict__ bpart,
                                                            const float H2, const float* __restrict__ umass)
{
    __shared__ float lsum[3 * kGroup];
    far_dispatch<SOFT, false>(xs, n, nsb, 1, (int)nsb, U, W, Q, apart, plain, bpart, H2, umass, lsum);
}

// one column batch [c0, c1) of the far field
template <bool SOFT>
__global__ __launch_bounds__(kFarThreads, 2) void far_column_kernel(const float4* __restrict__ xs, const int n,
                                                                   const int c0, const int c1, const int64_t U,
                                                                   const int W, const int Q,
                                                                   float4* __restrict__ apart, const int plain,
                                                                   float* __restrict__ bpart, const float H2,
                                                                   const float* __restrict__ umass)
{
    __shared__ float lsum[3 * kGroup];
    far_dispatch<SOFT, true>(xs, n, 0, c0, c1, U, W, Q, apart, plain, bpart, H2, umass, lsum);
}

// Per slot, in a fixed order: the near partials (share order), the far own-side fragments
// (workgroup order), the far partner-side partials (group order, negated:
// they were summed as seen from the other side), compensated; then the
// acceleration is finished as the one-range sweep does.  With PARTS > 1
// (small counts: the slots alone would not fill the GPU, and a slot's row
// has up to hundreds of fragments), PARTS threads share a slot: part p
// takes every PARTS-th term of each list from the p-th on, and part 0 adds
// the parts' compensated sums in part order.  PLAIN: the fragments hold
// sums alone (their corrections are zero and were not stored).
template <int PARTS, bool PLAIN>
__global__ __launch_bounds__(kBlock) void sym_combine_kernel(
    const float4* __restrict__ near_part, const float4* __restrict__ apart, const float* __restrict__ bpart,
    const int n, const int nshares, const int64_t nsb, const int64_t U, const int W, const int Q, const int grav,
    const int* __restrict__ perm,
    const float4* __restrict__ xs, float4* __restrict__ acc, const int finish, const int add_acc,
    const float4* __restrict__ vel, float4* __restrict__ xvel, float* __restrict__ maxvsig, const DevParams P)
{
    constexpr int kSlots = kBlock / PARTS;  // slots per block
    const int p = (int)threadIdx.x / kSlots;
    const int t = blockIdx.x * kSlots + ((int)threadIdx.x - p * kSlots);
    CSum x{0.0f, 0.0f}, y{0.0f, 0.0f}, z{0.0f, 0.0f};
    if (grav && t < n) {
        for (int sl = p; sl < nshares; sl += PARTS) {
            const size_t e = 2 * ((size_t)sl * n + t);
            const float4 a = near_part[e], c = near_part[e + 1];
            x.add(a.x); y.add(a.y); z.add(a.z);
            x.c += c.x; y.c += c.y; z.c += c.z;
        }
        const int64_t g = t / kGroup, local = t - g * kGroup;
        if (U > 0 && g < rows_with_units(nsb)) {
            // one fragment per workgroup that touched the row (every workgroup
            // has a share: W <= U); eight fragments' loads in flight before
            // their ordered adds
            const int64_t k0 = wg_of_unit(row_prefix(g, nsb, Q), U, W);
            const int64_t k1 = wg_of_unit(row_prefix(g + 1, nsb, Q) - 1, U, W);
            int64_t k = k0 + p;
            const int el = frag_elem((int)local);
            for (; k + 7 * PARTS <= k1; k += 8 * PARTS) {
                float4 f[8], fc[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float4* fr = apart + (size_t)kFragElems * (g + k + u * PARTS) + el;
                    f[u] = fr[0];
                    if (!PLAIN) fc[u] = fr[64];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    x.add(f[u].x); y.add(f[u].y); z.add(f[u].z);
                    if (!PLAIN) { x.c += fc[u].x; y.c += fc[u].y; z.c += fc[u].z; }
                }
            }
            for (; k <= k1; k += PARTS) {
                const float4* fr = apart + (size_t)kFragElems * (g + k) + el;
                const float4 f = fr[0];
                x.add(f.x); y.add(f.y); z.add(f.z);
                if (!PLAIN) {
                    const float4 fc = fr[64];
                    x.c += fc.x; y.c += fc.y; z.c += fc.z;
                }
            }
        }
        const int64_t base = bpart_group_base(g) + local;
        const int ng = (int)g;
        // eight groups' loads in flight before their (ordered) adds
        int gp = p;
        for (; gp + 7 * PARTS < ng; gp += 8 * PARTS) {
            float v[8][3];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const size_t e = 3 * (size_t)(base + (int64_t)(gp + u * PARTS) * kGroup);
                v[u][0] = bpart[e]; v[u][1] = bpart[e + 1]; v[u][2] = bpart[e + 2];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                x.add(-v[u][0]); y.add(-v[u][1]); z.add(-v[u][2]);
            }
        }
        for (; gp < ng; gp += PARTS) {
            const size_t e = 3 * (size_t)(base + (int64_t)gp * kGroup);
            x.add(-bpart[e]); y.add(-bpart[e + 1]); z.add(-bpart[e + 2]);
        }
    }
    if constexpr (PARTS > 1) {
        __shared__ float red[6][kBlock];  // every thread reaches the barrier
        if (p > 0) {
            red[0][threadIdx.x] = x.s; red[1][threadIdx.x] = x.c;
            red[2][threadIdx.x] = y.s; red[3][threadIdx.x] = y.c;
            red[4][threadIdx.x] = z.s; red[5][threadIdx.x] = z.c;
        }
        __syncthreads();
        if (p > 0) return;
#pragma unroll
        for (int q = 1; q < PARTS; ++q) {
            const int o = (int)threadIdx.x + q * kSlots;
            x.add(red[0][o]); x.c += red[1][o];
            y.add(red[2][o]); y.c += red[3][o];
            z.add(red[4][o]); z.c += red[5][o];
        }
    }
    if (t >= n) return;
    const int i = perm ? perm[t] : t;
    if (add_acc) {  // the SPH terms (and environment), summed before the far field was done
        const float4 a = acc[i];
        x.add(a.x); y.add(a.y); z.add(a.z);
    }
    float ax = x.value(), ay = y.value(), az = z.value();
    if (finish) {
        environment_acceleration(P, ax, ay, az);
        if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (maxvsig) maxvsig[i] = 0.0f;
    }
    acc[i] = make_float4(ax, ay, az, 0.0f);
}

// After column batch [c0, c1) of the far field (column-batched far field,
// one thread per slot): the slot's running compensated sums (accsum: 2
// float4 per slot, sums then corrections; none before the first batch)
// take the batch's own-side fragments of the slot's row and, for slots in
// the batch's groups, their partner-side partials, complete now (every row
// before the group was in the batch).  After the last batch the near
// partials join them and the acceleration is finished as in the combine.
__global__ __launch_bounds__(kBlock) void sym_column_combine_kernel(
    const float4* __restrict__ near_part, const float4* __restrict__ apart, const float* __restrict__ bpart,
    float4* __restrict__ accsum, const int n, const int nshares, const int c0, const int c1, const int64_t U,
    const int W, const int Q, const int plain, const int first, const int last, const int* __restrict__ perm,
    float4* __restrict__ acc, const int finish, const int add_acc, float4* __restrict__ xvel,
    float* __restrict__ maxvsig, const DevParams P)
{
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    CSum x{0.0f, 0.0f}, y{0.0f, 0.0f}, z{0.0f, 0.0f};
    if (!first) {
        const float4 a = accsum[2 * (size_t)t], c = accsum[2 * (size_t)t + 1];
        x = CSum{a.x, c.x}; y = CSum{a.y, c.y}; z = CSum{a.z, c.z};
    }
    const int g = t / kGroup, local = t - g * kGroup;
    if (g < col_rows(c1) && U > 0) {
        // the fragments of the workgroups that worked on row g in this batch
        const int64_t k0 = wg_of_unit(col_row_prefix(g, c0, c1, Q), U, W);
        const int64_t k1 = wg_of_unit(col_row_prefix(g + 1, c0, c1, Q) - 1, U, W);
        const int el = frag_elem(local);
        for (int64_t k = k0; k <= k1; ++k) {
            const float4* fr = apart + (size_t)kFragElems * (g + k) + el;
            const float4 f = fr[0];
            x.add(f.x); y.add(f.y); z.add(f.z);
            if (!plain) {
                const float4 fc = fr[64];
                x.c += fc.x; y.c += fc.y; z.c += fc.z;
            }
        }
    }
    if (g >= c0 && g < c1) {
        const int64_t base = col_bpart_base(g, c0) + local;
        int gp = 0;  // eight rows' loads in flight before their (ordered) adds
        for (; gp + 8 <= g; gp += 8) {
            float v[8][3];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const size_t e = 3 * (size_t)(base + (int64_t)(gp + u) * kGroup);
                v[u][0] = bpart[e]; v[u][1] = bpart[e + 1]; v[u][2] = bpart[e + 2];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                x.add(-v[u][0]); y.add(-v[u][1]); z.add(-v[u][2]);
            }
        }
        for (; gp < g; ++gp) {
            const size_t e = 3 * (size_t)(base + (int64_t)gp * kGroup);
            x.add(-bpart[e]); y.add(-bpart[e + 1]); z.add(-bpart[e + 2]);
        }
    }
    if (!last) {
        accsum[2 * (size_t)t] = make_float4(x.s, y.s, z.s, 0.0f);
        accsum[2 * (size_t)t + 1] = make_float4(x.c, y.c, z.c, 0.0f);
        return;
    }
    for (int sl = 0; sl < nshares; ++sl) {
        const size_t e = 2 * ((size_t)sl * n + t);
        const float4 a = near_part[e], c = near_part[e + 1];
        x.add(a.x); y.add(a.y); z.add(a.z);
        x.c += c.x; y.c += c.y; z.c += c.z;
    }
    const int i = perm ? perm[t] : t;
    if (add_acc) {  // the SPH terms (and environment), summed before the far field was done
        const float4 a = acc[i];
        x.add(a.x); y.add(a.y); z.add(a.z);
    }
    float ax = x.value(), ay = y.value(), az = z.value();
    if (finish) {
        environment_acceleration(P, ax, ay, az);
        if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (maxvsig) maxvsig[i] = 0.0f;
    }
    acc[i] = make_float4(ax, ay, az, 0.0f);
}

// ------------------------------------------------------------------ batches
// The same sweep over the larger instances of a batch (slot indices are
// the batch's; the smaller instances stay with the one-sided persistent
// sweep).  Differences: no shares in the near field (the batch has
// wavefronts to spare); the far kernel is persistent and takes whole rows
// (instance, group) from a counter, longest first, so each row's own-side
// sums leave as one fragment, numbered by the instance's first row + g.

// Bounding boxes of the 64-slot blocks of every instance (the batch's
// 256-slot block table, one wavefront per 64 slots).
__global__ __launch_bounds__(kBlock) void sym_batch_box_kernel(const float4* __restrict__ xs,
                                                               const int4* __restrict__ blk,
                                                               const SymInst* __restrict__ sinst,
                                                               float4* __restrict__ box)
{
    const int4 bl = blk[blockIdx.x];
    const int lane = threadIdx.x & 63, first = bl.y + 64 * (threadIdx.x >> 6);
    if (first >= bl.z) return;  // whole wavefronts leave together
    const SymInst S = sinst[bl.x];
    const int I = (first - S.off) >> 6, t = first + lane;
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (t < bl.z) {
        const float4 p = xs[t];
        v[0] = v[3] = p.x; v[1] = v[4] = p.y; v[2] = v[5] = p.z;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = fminf(v[c], __shfl_xor(v[c], o, 64));
            v[c + 3] = fmaxf(v[c + 3], __shfl_xor(v[c + 3], o, 64));
        }
    }
    if (lane == 0) {
        box[2 * (size_t)(S.box0 + I)] = make_float4(v[0], v[1], v[2], 0.0f);
        box[2 * (size_t)(S.box0 + I) + 1] = make_float4(v[3], v[4], v[5], 0.0f);
    }
}

// near-block bits of instance blockIdx.y (as near_bits_kernel, per instance)
__global__ __launch_bounds__(kBlock) void sym_batch_bits_kernel(const float4* __restrict__ box,
                                                                const SymInst* __restrict__ sinst,
                                                                const BatchInst* __restrict__ insts,
                                                                uint32_t* __restrict__ bits)
{
    const SymInst S = sinst[blockIdx.y];
    const int idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= S.nb * S.nwords) return;
    const float H2 = insts[blockIdx.y].H2;
    const int I = idx / S.nwords, w = idx - I * S.nwords;
    const float4* bx = box + 2 * (size_t)S.box0;
    const float4 lo = bx[2 * I], hi = bx[2 * I + 1];
    uint32_t word = 0;
    for (int b = 0; b < 32; ++b) {
        const int J = 32 * w + b;
        if (J >= S.nb) break;
        const float4 l2 = bx[2 * J], h2 = bx[2 * J + 1];
        const float gx = fmaxf(0.0f, fmaxf(l2.x - hi.x, lo.x - h2.x));
        const float gy = fmaxf(0.0f, fmaxf(l2.y - hi.y, lo.y - h2.y));
        const float gz = fmaxf(0.0f, fmaxf(l2.z - hi.z, lo.z - h2.z));
        if (dist2_exact(gx, gy, gz) <= H2) word |= 1u << b;
    }
    bits[S.bits0 + idx] = word;
}

template <bool GRAV, bool NBR, bool SOFT>
__global__ __launch_bounds__(kBlock) void sym_batch_near_kernel(
    const float4* __restrict__ xs, const int4* __restrict__ blk, const SymInst* __restrict__ sinst,
    const uint32_t* __restrict__ bits, const float4* __restrict__ box, int* __restrict__ nbr,
    int* __restrict__ nbr_cnt, const int kmax, const int ntotal, float4* __restrict__ near_part,
    const DevParams* __restrict__ prm, const float* __restrict__ umass)
{
    // the lists issue ahead of the far field's wavefronts beside them: the
    // passes wait for them (ensemble of 512: 16.25-16.30 -> 15.98-16.03 ms
    // per step, with the passes' own priority)
    if (NBR) __builtin_amdgcn_s_setprio(3);
    const int4 bl = blk[xcd_block((int)blockIdx.x, (int)gridDim.x)];  // XCD-contiguous (xcd_block)
    const int lane = threadIdx.x & 63;
    const int first = __builtin_amdgcn_readfirstlane(bl.y + 64 * (int)(threadIdx.x >> 6));
    if (first >= bl.z) return;
    const int k = bl.x;
    const SymInst S = sinst[k];
    const int I = (first - S.off) >> 6;
    const DevParams P = prm[k];
    const Rare R = make_rare(nbr, kmax, P.H2);
    const uint32_t* b = NBR ? bits + S.bits0 : nullptr;
    const float4* bx = NBR ? box + 2 * (size_t)S.box0 : nullptr;
    if constexpr (GRAV) {
        if (umass[2 * k] != 0.0f) {
            near_blocks<GRAV, NBR, SOFT, true>(xs, S.off, S.n, S.nb, S.nwords, b, bx, I, 0, 1, lane, nbr_cnt, near_part, P,
                                               R, umass[2 * k + 1]);
            return;
        }
    }
    near_blocks<GRAV, NBR, SOFT, false>(xs, S.off, S.n, S.nb, S.nwords, b, bx, I, 0, 1, lane, nbr_cnt, near_part, P, R,
                                        1.0f);
}

// one piece of a far row of an instance of the batch: group R.g against
// R.units partner groups from j0 on, two pieces of 512 partners each
template <bool UM, bool SOFT>
__device__ __forceinline__ void batch_far_row(const float4* __restrict__ xk, const int n, const SymRow R, const int j0,
                                              float* lsum, float4* frag, const bool plain, float* __restrict__ bk,
                                              const float H2, const float m, const int lane)
{
    FarOwn o;
    far_load_row(xk, n, R.g, lane, lsum, frag, plain, o);
    for (int s = 0; s < 2 * R.units; ++s)
        far_unit<UM, SOFT, 2>(xk, n, R.g, j0 + (s >> 1), s & 1, 2, o, bk, 0, H2, m, lane);
    far_flush_own(o, lane);
}

// Persistent far field over the batch's rows; every wavefront leaves once
// the counter passes nrows.  plain: no piece has more than kFoldsPerFlush
// units.
template <bool SOFT>
__global__ __launch_bounds__(kFarThreads, 2) void sym_batch_far_kernel(
    const float4* __restrict__ xs, const SymInst* __restrict__ sinst, const BatchInst* __restrict__ insts,
    const SymRow* __restrict__ rows, const int nrows, int* __restrict__ queue, float4* __restrict__ apart,
    const int plain, float* __restrict__ bpart, const float* __restrict__ umass)
{
    __shared__ float lsum[3 * kGroup];
    const int lane = threadIdx.x;
    for (;;) {
        int r = 0;
        if (lane == 0) r = atomicAdd(queue, 1);
        r = __builtin_amdgcn_readfirstlane(__shfl(r, 0, 64));
        if (r >= nrows) break;
        const SymRow R = rows[r];
        const SymInst S = sinst[R.inst];
        const float H2 = insts[R.inst].H2;
        const float4* xk = xs + S.off;
        float* const bk = bpart + S.bpart0;
        const bool um = umass[2 * R.inst] != 0.0f;
        const float m = umass[2 * R.inst + 1];
        const int j0 = R.g + 1 + (int)batch_piece_begin(row_units(R.g, S.nsb), R.piece);
        float4* const frag = apart + (size_t)kFragElems * ((size_t)kBatchPieces * (S.row0 + R.g) + R.piece);
        if (um)
            batch_far_row<true, SOFT>(xk, S.n, R, j0, lsum, frag, plain, bk, H2, m, lane);
        else
            batch_far_row<false, SOFT>(xk, S.n, R, j0, lsum, frag, plain, bk, H2, 1.0f, lane);
    }
}

// Per slot of the instances in the sweep (256-slot block table), as
// sym_combine_kernel: near partial, the row's one own-side fragment, the
// partner-side partials (group order, negated).
template <bool PLAIN>
__global__ __launch_bounds__(kBlock) void sym_batch_combine_kernel(
    const float4* __restrict__ near_part, const float4* __restrict__ apart, const float* __restrict__ bpart,
    const SymInst* __restrict__ sinst, const int4* __restrict__ blk, const int grav, const int* __restrict__ perm,
    const float4* __restrict__ xs, float4* __restrict__ acc, const int finish, const int add_acc,
    const float4* __restrict__ vel, float4* __restrict__ xvel, float* __restrict__ maxvsig,
    const DevParams* __restrict__ prm)
{
    const int4 bl = blk[blockIdx.x];
    const int t = bl.y + threadIdx.x;
    if (t >= bl.z) return;
    const int k = bl.x;
    const SymInst S = sinst[k];
    CSum x{0.0f, 0.0f}, y{0.0f, 0.0f}, z{0.0f, 0.0f};
    if (grav) {
        const float4 a = near_part[2 * (size_t)t], c = near_part[2 * (size_t)t + 1];
        x.add(a.x); y.add(a.y); z.add(a.z);
        x.c += c.x; y.c += c.y; z.c += c.z;
        const int local = t - S.off;
        const int g = local / kGroup, lg = local - g * kGroup;
        if (g < S.nrows) {
            const int pieces = batch_row_pieces(row_units(g, S.nsb));
            for (int h = 0; h < pieces; ++h) {
                const float4* fr = apart + (size_t)kFragElems * ((size_t)kBatchPieces * (S.row0 + g) + h) + frag_elem(lg);
                const float4 f = fr[0];
                x.add(f.x); y.add(f.y); z.add(f.z);
                if (!PLAIN) {
                    const float4 fc = fr[64];
                    x.c += fc.x; y.c += fc.y; z.c += fc.z;
                }
            }
        }
        const float* bk = bpart + S.bpart0;
        const int64_t base = bpart_group_base(g) + lg;
        int gp = 0;  // eight groups' loads in flight before their (ordered) adds
        for (; gp + 8 <= g; gp += 8) {
            float v[8][3];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const size_t e = 3 * (size_t)(base + (int64_t)(gp + u) * kGroup);
                v[u][0] = bk[e]; v[u][1] = bk[e + 1]; v[u][2] = bk[e + 2];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                x.add(-v[u][0]); y.add(-v[u][1]); z.add(-v[u][2]);
            }
        }
        for (; gp < g; ++gp) {
            const size_t e = 3 * (size_t)(base + (int64_t)gp * kGroup);
            x.add(-bk[e]); y.add(-bk[e + 1]); z.add(-bk[e + 2]);
        }
    }
    const int i = perm ? perm[t] : t;
    if (add_acc) {  // the SPH terms (and environment), summed before the far field was done
        const float4 a = acc[i];
        x.add(a.x); y.add(a.y); z.add(a.z);
    }
    float ax = x.value(), ay = y.value(), az = z.value();
    if (finish) {
        environment_acceleration(prm[k], ax, ay, az);
        if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (maxvsig) maxvsig[i] = 0.0f;
    }
    acc[i] = make_float4(ax, ay, az, 0.0f);
}

template <bool GRAV, bool NBR, bool SOFT>
void launch_batch_near(const SymBatchLaunch& L, hipStream_t s)
{
    hipLaunchKernelGGL((sym_batch_near_kernel<GRAV, NBR, SOFT>), dim3(L.n_blk), dim3(kBlock), 0, s, L.xs, L.blk,
                       L.sinst, L.bits, L.box, L.nbr, L.nbr_cnt, L.kmax, L.ntotal, L.near_part, L.prm, L.umass);
}

template <bool GRAV, bool NBR, bool SOFT>
void launch_near(const SymLaunch& L, const SymGeom& G, const DevParams& P, hipStream_t s)
{
    const int blocks = (G.nb + kBlock / 64 - 1) / (kBlock / 64);
    hipLaunchKernelGGL((near_kernel<GRAV, NBR, SOFT>), dim3(blocks, L.nshares), dim3(kBlock), 0, s, L.xs, L.n, G.nb,
                       G.nwords,
                       L.want_nbr ? L.bits : nullptr, L.want_nbr ? L.box : nullptr, L.nbr, L.nbr_cnt, L.kmax,
                       L.near_part, P, L.umass, L.list_prio ? 1 : 0);
}

}  // namespace

int sym_default_workgroups(int cus)
{
    return std::max(1, 8 * cus / kGroupSB);  // 2 wavefronts per SIMD
}

// one column batch's units, far workgroups (of `wanted`) and fragment kind
SymCol sym_column(int64_t nsb, int split, int wanted, int c0, int c1)
{
    SymCol C;
    C.c0 = c0;
    C.c1 = c1;
    C.units = col_row_prefix(col_rows(c1), c0, c1, split);
    C.wgs = far_workgroups(wanted, C.units);
    C.plain = fragments_plain((C.units + C.wgs - 1) / C.wgs);
    (void)nsb;
    return C;
}

SymGeom sym_geometry(int64_t n, int wgs, bool want_nbr, int nshares, int64_t col_elems)
{
    SymGeom G;
    G.nb = (int)((n + 63) / 64);
    G.nsb = (n + kSB - 1) / kSB;
    G.groups = (int)((G.nsb + kGroupSB - 1) / kGroupSB);
    G.split = far_split(n, wgs);
    G.units = row_prefix(rows_with_units(G.nsb), G.nsb, G.split);
    G.wgs = far_workgroups(wgs, G.units);
    G.nwords = (G.nb + 31) / 32;
    G.box_elems = 2 * (int64_t)G.nb;
    G.bits_elems = want_nbr ? (int64_t)G.nb * G.nwords : 0;
    G.near_elems = 2 * n * nshares;
    G.apart_elems = (int64_t)kFragElems * (G.groups + wgs);
    G.plain = fragments_plain((G.units + G.wgs - 1) / G.wgs);
    G.bpart_elems = 3 * bpart_group_base(G.groups);
    G.ncols = 1;
    if (G.bpart_elems > col_elems) {
        const std::vector<int> c = sym_columns(G.nsb, col_elems);
        G.ncols = (int)c.size() - 1;
        G.bpart_elems = 0;
        for (int b = 0; b < G.ncols; ++b) G.bpart_elems = std::max(G.bpart_elems, 3 * col_bpart_base(c[b + 1], c[b]));
    }
    return G;
}

// near field, neighbour lists: boxes, near bits, the distance tests
hipError_t launch_sym_lists(const SymLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0 || !L.want_nbr) return hipSuccess;
    if (L.list_shares < 1) return hipErrorInvalidValue;
    SymLaunch M = L;
    M.nshares = L.list_shares;  // the list kernel's shares (list planes), not the in-group gravity's
    const SymGeom G = sym_geometry(M.n, M.wgs, M.want_nbr, M.nshares);
    hipLaunchKernelGGL(block_box_kernel, dim3((G.nb + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, s, L.xs, L.n,
                       G.nb, L.box);
    const int64_t words = (int64_t)G.nb * G.nwords;
    hipLaunchKernelGGL(near_bits_kernel, dim3((unsigned)((words + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, L.box,
                       G.nb, G.nwords, P.H2, L.bits);
    if (P.H2 > 0.0f) launch_near<false, true, true>(M, G, P, s); else launch_near<false, true, false>(M, G, P, s);
    return hipGetLastError();
}

// near field, gravity inside the groups (the partials of the combine)
hipError_t launch_sym_ingroup(const SymLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0 || !L.grav) return hipSuccess;
    if (L.nshares < 1) return hipErrorInvalidValue;
    SymLaunch M = L;
    M.want_nbr = false;  // own group only, no bits
    M.list_prio = false;
    const SymGeom G = sym_geometry(M.n, M.wgs, false, M.nshares);
    if (P.H2 > 0.0f) launch_near<true, false, true>(M, G, P, s); else launch_near<true, false, false>(M, G, P, s);
    return hipGetLastError();
}

bool sym_has_far(const SymLaunch& L) { return L.n > 0 && L.grav && sym_geometry(L.n, L.wgs, false, 1).units > 0; }

template <bool SOFT>
void launch_far_column(const SymLaunch& L, const SymCol& C, const int split, const DevParams& P, hipStream_t s)
{
    hipLaunchKernelGGL((far_column_kernel<SOFT>), dim3(C.wgs), dim3(kFarThreads), 0, s, L.xs, L.n, C.c0, C.c1,
                       C.units, C.wgs, split, L.apart, C.plain ? 1 : 0, L.bpart, P.H2, L.umass);
}

void launch_column_combine(const SymLaunch& L, const SymCol& C, const int split, const bool first, const bool last,
                           const DevParams& P, hipStream_t s)
{
    hipLaunchKernelGGL(sym_column_combine_kernel, dim3((L.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, L.near_part,
                       L.apart, L.bpart, L.accsum, L.n, L.nshares, C.c0, C.c1, C.units, C.wgs, split, C.plain ? 1 : 0,
                       first ? 1 : 0, last ? 1 : 0, L.perm, L.acc, L.finish ? 1 : 0, L.add_acc ? 1 : 0, L.xvel,
                       L.maxvsig, P);
}

// far field: gravity between groups, once per pair (needs only xs and
// umass); in column batches when the partials of the whole would not fit,
// each batch but the last followed by its column combine
hipError_t launch_sym_far(const SymLaunch& L, const DevParams& P, hipStream_t s)
{
    if (!sym_has_far(L)) return hipSuccess;
    const SymGeom G = sym_geometry(L.n, L.wgs, L.want_nbr, L.nshares, L.col_elems);
    if (G.ncols == 1) {
        if (P.H2 > 0.0f)
            hipLaunchKernelGGL((far_kernel<true>), dim3(L.wgs), dim3(kFarThreads), 0, s, L.xs, L.n, G.nsb, G.units,
                               L.wgs, G.split, L.apart, G.plain ? 1 : 0, L.bpart, P.H2, L.umass);
        else
            hipLaunchKernelGGL((far_kernel<false>), dim3(L.wgs), dim3(kFarThreads), 0, s, L.xs, L.n, G.nsb, G.units,
                               L.wgs, G.split, L.apart, G.plain ? 1 : 0, L.bpart, P.H2, L.umass);
        return hipGetLastError();
    }
    const std::vector<int> c = sym_columns(G.nsb, L.col_elems);
    for (int b = 0; b < G.ncols; ++b) {
        const SymCol C = sym_column(G.nsb, G.split, L.wgs, c[b], c[b + 1]);
        if (P.H2 > 0.0f) launch_far_column<true>(L, C, G.split, P, s); else launch_far_column<false>(L, C, G.split, P, s);
        if (b + 1 < G.ncols) launch_column_combine(L, C, G.split, b == 0, false, P, s);
    }
    return hipGetLastError();
}

// combine (after both fields): the acceleration
hipError_t launch_sym_combine(const SymLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0 || !(L.grav || L.finish)) return hipSuccess;
    const SymGeom G = sym_geometry(L.n, L.wgs, L.want_nbr, L.nshares, L.col_elems);
    if (G.ncols > 1 && sym_has_far(L)) {
        // the last column batch's combine, which finishes
        const std::vector<int> c = sym_columns(G.nsb, L.col_elems);
        const SymCol C = sym_column(G.nsb, G.split, L.wgs, c[G.ncols - 1], c[G.ncols]);
        launch_column_combine(L, C, G.split, false, true, P, s);
        return hipGetLastError();
    }
    // threads per slot: about 2^17 threads or more in all
    const int parts = L.n >= 98304 ? 1 : L.n >= 49152 ? 2 : L.n >= 24576 ? 4 : 8;
#define GRASPH_COMBINE(PARTS, PLAIN)                                                                                  \
    hipLaunchKernelGGL((sym_combine_kernel<PARTS, PLAIN>), dim3((L.n + kBlock / PARTS - 1) / (kBlock / PARTS)),        \
                       dim3(kBlock), 0, s, L.near_part, L.apart, L.bpart, L.n, L.nshares, G.nsb, G.units, L.wgs,       \
                       G.split, L.grav ? 1 : 0, L.perm, L.xs, L.acc, L.finish ? 1 : 0, L.add_acc ? 1 : 0, L.vel,       \
                       L.xvel, L.maxvsig, P)
    if (G.plain) {
        switch (parts) {
        case 1: GRASPH_COMBINE(1, true); break;
        case 2: GRASPH_COMBINE(2, true); break;
        case 4: GRASPH_COMBINE(4, true); break;
        default: GRASPH_COMBINE(8, true); break;
        }
    } else {
        GRASPH_COMBINE(1, false);  // more than kFoldsPerFlush units per workgroup: large counts (or few workgroups)
    }
#undef GRASPH_COMBINE
    return hipGetLastError();
}

SymBatchTables sym_batch_tables(const std::vector<int64_t>& offsets, bool want_nbr, int64_t min_n)
{
    SymBatchTables T;
    const int n_inst = (int)offsets.size() - 1;
    int box0 = 0, row0 = 0;
    int64_t bits0 = 0, bpart0 = 0;
    for (int k = 0; k < n_inst; ++k) {
        const int64_t n = offsets[k + 1] - offsets[k];
        const bool sym = n >= min_n && n > 0;
        SymInst S;
        S.off = (int)offsets[k];
        S.n = sym ? (int)n : 0;  // instances left to the one-sided sweep: nothing here
        if (sym)
            for (int64_t s = offsets[k]; s < offsets[k + 1]; s += 256)
                T.blk.push_back(make_int4(k, (int)s, (int)std::min<int64_t>(s + 256, offsets[k + 1]), 0));
        S.nb = (int)((S.n + 63) / 64);
        S.nwords = (S.nb + 31) / 32;
        S.nsb = (int)((S.n + kSB - 1) / kSB);
        S.nrows = (int)rows_with_units(S.nsb);
        S.box0 = box0;
        S.row0 = row0;
        S.bits0 = bits0;
        S.bpart0 = bpart0;
        for (int g = 0; g < S.nrows; ++g) {
            const int64_t u = row_units(g, S.nsb);
            for (int h = 0; h < batch_row_pieces(u); ++h) {
                const int units = (int)(batch_piece_begin(u, h + 1) - batch_piece_begin(u, h));
                T.rows.push_back(SymRow{k, g, h, units});
                T.plain = T.plain && fragments_plain(2 * (int64_t)units);  // two far units per partner group
            }
        }
        T.max_words = std::max(T.max_words, S.nb * S.nwords);
        box0 += S.nb;
        row0 += S.nrows;
        if (want_nbr) bits0 += (int64_t)S.nb * S.nwords;
        bpart0 += 3 * bpart_group_base((S.nsb + kGroupSB - 1) / kGroupSB);
        T.insts.push_back(S);
    }
    std::stable_sort(T.rows.begin(), T.rows.end(), [](const SymRow& a, const SymRow& b) { return a.units > b.units; });
    T.box_elems = 2 * (int64_t)box0;
    T.bits_elems = bits0;
    T.apart_elems = (int64_t)kFragElems * kBatchPieces * row0;
    T.bpart_elems = bpart0;
    return T;
}

hipError_t launch_sym_batch_lists(const SymBatchLaunch& L, hipStream_t s)
{
    if (L.ntotal <= 0 || L.n_blk <= 0 || !L.want_nbr) return hipSuccess;
    hipLaunchKernelGGL(sym_batch_box_kernel, dim3(L.n_blk), dim3(kBlock), 0, s, L.xs, L.blk, L.sinst, L.box);
    if (L.max_words > 0)
        hipLaunchKernelGGL(sym_batch_bits_kernel, dim3((L.max_words + kBlock - 1) / kBlock, L.n_inst), dim3(kBlock),
                           0, s, L.box, L.sinst, L.insts, L.bits);
    if (L.soft) launch_batch_near<false, true, true>(L, s); else launch_batch_near<false, true, false>(L, s);
    return hipGetLastError();
}

hipError_t launch_sym_batch_ingroup(const SymBatchLaunch& L, hipStream_t s)
{
    if (L.ntotal <= 0 || L.n_blk <= 0 || !L.grav) return hipSuccess;
    if (L.soft) launch_batch_near<true, false, true>(L, s); else launch_batch_near<true, false, false>(L, s);
    return hipGetLastError();
}

hipError_t launch_sym_batch_far(const SymBatchLaunch& L, hipStream_t s)
{
    if (!sym_batch_has_far(L)) return hipSuccess;
    hipError_t e;
    if ((e = launch_reset_queue(L.queue, 1, s)) != hipSuccess) return e;
    if (L.soft)
        hipLaunchKernelGGL((sym_batch_far_kernel<true>), dim3(L.wgs), dim3(kFarThreads), 0, s, L.xs, L.sinst, L.insts,
                           L.rows, L.nrows, L.queue, L.apart, L.plain ? 1 : 0, L.bpart, L.umass);
    else
        hipLaunchKernelGGL((sym_batch_far_kernel<false>), dim3(L.wgs), dim3(kFarThreads), 0, s, L.xs, L.sinst, L.insts,
                           L.rows, L.nrows, L.queue, L.apart, L.plain ? 1 : 0, L.bpart, L.umass);
    return hipGetLastError();
}

hipError_t launch_sym_batch_combine(const SymBatchLaunch& L, hipStream_t s)
{
    if (L.ntotal <= 0 || L.n_blk <= 0 || !(L.grav || L.finish)) return hipSuccess;
    if (L.plain)
        hipLaunchKernelGGL(sym_batch_combine_kernel<true>, dim3(L.n_blk), dim3(kBlock), 0, s, L.near_part, L.apart,
                           L.bpart, L.sinst, L.blk, L.grav ? 1 : 0, L.perm, L.xs, L.acc, L.finish ? 1 : 0,
                           L.add_acc ? 1 : 0, L.vel, L.xvel, L.maxvsig, L.prm);
    else
        hipLaunchKernelGGL(sym_batch_combine_kernel<false>, dim3(L.n_blk), dim3(kBlock), 0, s, L.near_part, L.apart,
                           L.bpart, L.sinst, L.blk, L.grav ? 1 : 0, L.perm, L.xs, L.acc, L.finish ? 1 : 0,
                           L.add_acc ? 1 : 0, L.vel, L.xvel, L.maxvsig, L.prm);
    return hipGetLastError();
}

}  // namespace grasph
