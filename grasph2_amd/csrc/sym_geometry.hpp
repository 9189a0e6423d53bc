// sym_geometry.hpp -- how the symmetric sweep (sym_sweep.hip) cuts its work:
// slots in groups of 1024 (one far-field wavefront, 16 slots per lane; a
// super-block is the slots of one far wavefront, here the whole group).  The
// far field's rows are group g against the later groups; one unit is one
// partner super-block, or a piece of one.  Host and device code use these functions, and
// tests/cpp/sym_geometry_check.cpp checks them on the host (every far pair
// exactly once, indices in bounds, every partial that is read was written).
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <vector>

namespace grasph {

// own slots per lane in the far field: the partner-side reduction and the
// partners' DPP moves of a chunk are shared by the lane's slots (16: one
// wavefront owns a group of 1024, 209-235 VGPRs, 2 wavefronts per SIMD; 4 was
// measured 7% slower than 8)
constexpr int kNI = 16;
constexpr int kSB = 64 * kNI;           // slots per super-block (a far wavefront's slots)
constexpr int kGroup = 1024;            // slots per group
constexpr int kGroupSB = kGroup / kSB;  // super-blocks per group (wavefronts per far workgroup)

static_assert(kGroupSB == 1, "the row geometry below takes one far wavefront per group");

// Column batches: the far field may run in batches of partner groups
// [c0, c1) (1 <= c0 < c1 <= nsb), so the partner-side partials, which grow
// with n^2, need only the batch's share.  Row g of a batch holds its
// partner groups [max(g + 1, c0), c1), each cut into Q pieces: rows
// 0 .. c1 - 2 have units.  The unbatched far field is the batch [1, nsb).
__device__ __host__ inline int64_t col_row_units(int64_t g, int64_t c0, int64_t c1, int Q)
{
    const int64_t first = g + 1 > c0 ? g + 1 : c0;
    return first < c1 ? (c1 - first) * Q : 0;
}
// units before row g (g <= c1 - 1): rows below c0 hold c1 - c0 groups
// each, row g' >= c0 holds c1 - 1 - g'
__device__ __host__ inline int64_t col_row_prefix(int64_t g, int64_t c0, int64_t c1, int Q)
{
    const int64_t a = g < c0 ? g : c0, m = g > c0 ? g - c0 : 0;
    return Q * (a * (c1 - c0) + m * (c1 - 1 - c0) - m * (m - 1) / 2);
}
__device__ __host__ inline int64_t col_rows(int64_t c1) { return c1 >= 2 ? c1 - 1 : 0; }
// the row holding unit u of the batch: the largest g with col_row_prefix(g) <= u
__device__ __host__ inline int64_t col_row_of_unit(int64_t u, int64_t c0, int64_t c1, int Q)
{
    int64_t lo = 0, hi = col_rows(c1) - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (col_row_prefix(mid, c0, c1, Q) <= u) lo = mid; else hi = mid - 1;
    }
    return lo;
}
// first partner group of row g in the batch
__device__ __host__ inline int64_t col_first_group(int64_t g, int64_t c0) { return g + 1 > c0 ? g + 1 : c0; }
// first element (of 3 per partner and row) of the batch's partner-side
// partials of group J: the groups c0 .. J - 1 before it take J' rows each
__device__ __host__ inline int64_t col_bpart_base(int64_t J, int64_t c0)
{
    return (int64_t)kGroup * (J * (J - 1) / 2 - c0 * (c0 - 1) / 2);
}

// Column batches of the far field: boundaries c[0] = 1 < c[1] < ... <
// c[last] = nsb, each batch [c[b], c[b + 1]) as wide as keeps its
// partner-side partials within max_elems floats (one group at least).
__host__ inline std::vector<int> sym_columns(int64_t nsb, int64_t max_elems)
{
    std::vector<int> c{1};
    if (nsb <= 1) {
        c.push_back(1);
        return c;
    }
    int64_t c0 = 1;
    while (c0 < nsb) {
        int64_t c1 = c0 + 1;
        while (c1 < nsb && 3 * col_bpart_base(c1 + 1, c0) <= max_elems) ++c1;
        c.push_back((int)c1);
        c0 = c1;
    }
    return c;
}

// far units of row g: each partner super-block of [kGroupSB (g + 1), nsb)
// cut into Q pieces of kSB / Q partners (Q = 1, 2, 4, 8 or 16; more pieces
// give small particle counts enough units to fill the GPU)
__device__ __host__ inline int64_t row_units(int64_t g, int64_t nsb, int Q = 1)
{
    const int64_t u = nsb - (int64_t)kGroupSB * (g + 1);
    return u > 0 ? u * Q : 0;
}
// units before row g (g <= rows_with_units): sum over g' < g of row_units(g')
__device__ __host__ inline int64_t row_prefix(int64_t g, int64_t nsb, int Q = 1)
{
    return Q * (g * nsb - (int64_t)kGroupSB * (g * (g + 1) / 2));
}
// rows with far work: nsb - kGroupSB (g + 1) >= 1
__device__ __host__ inline int64_t rows_with_units(int64_t nsb) { return nsb >= 1 ? (nsb - 1) / kGroupSB : 0; }
// the row holding unit u (0 <= u < row_prefix(rows_with_units(nsb), nsb, Q)): the largest g with row_prefix(g) <= u
__device__ __host__ inline int64_t row_of_unit(int64_t u, int64_t nsb, int Q = 1)
{
    int64_t lo = 0, hi = rows_with_units(nsb) - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) / 2;
        if (row_prefix(mid, nsb, Q) <= u) lo = mid; else hi = mid - 1;
    }
    return lo;
}
// pieces per partner super-block for n particles and `wanted` far
// workgroups (as many as run at once): the kernel lasts as long as its
// fullest share, ceil(U / W) units of kSB / Q partners with W = min(wanted,
// U); the Q of 2, 4, 8, 16 that makes that shortest, the fewest pieces on
// a tie (smaller units cost more per partner)
// a piece has at most 512 partners (a slot's plain FP32 own-side sum takes
// at most that many terms) and at least 64 (4 chunks of 16)
constexpr int kMinSplit = 2, kMaxSplit = 16;
__device__ __host__ inline int far_split(int64_t n, int64_t wanted)
{
    const int64_t nsb = (n + kSB - 1) / kSB;
    const int64_t u1 = row_prefix(rows_with_units(nsb), nsb, 1);
    if (u1 <= 0 || wanted <= 0) return kMinSplit;
    int best = kMinSplit;
    int64_t best_num = 0, best_den = 1;  // makespan as the fraction num / den super-blocks
    for (int Q = kMinSplit; Q <= kMaxSplit; Q *= 2) {
        const int64_t U = u1 * Q, W = wanted < U ? wanted : U;
        const int64_t num = (U + W - 1) / W;
        if (Q == kMinSplit || num * best_den < best_num * Q) {
            best = Q;
            best_num = num;
            best_den = Q;
        }
    }
    return best;
}
// the workgroup whose share [k U / W, (k + 1) U / W) holds unit u
__device__ __host__ inline int64_t wg_of_unit(int64_t u, int64_t U, int64_t W) { return ((u + 1) * W - 1) / U; }
// workgroups of the far kernel: at most the units, so every one has a share and a row's
// fragments come from consecutive workgroups
__host__ inline int far_workgroups(int64_t wanted, int64_t units)
{
    const int64_t w = wanted < units ? wanted : units;
    return (int)(w > 1 ? w : 1);
}
// A batch's far rows go to its wavefronts whole or in pieces: row g of an
// instance (row_units(g, nsb) partner groups) is cut into
// batch_row_pieces(units) pieces of consecutive partner groups, piece h
// starting at batch_piece_begin(units, h); each piece leaves its own-side
// sums as its own fragment (kBatchPieces fragments per row), so the longest
// rows do not set the kernel's length.
constexpr int kBatchPieces = 4;
__device__ __host__ inline int batch_row_pieces(int64_t units) { return units < kBatchPieces ? (int)units : kBatchPieces; }
__device__ __host__ inline int64_t batch_piece_begin(int64_t units, int h)
{
    return h * units / batch_row_pieces(units);
}
// first element of the partner-side partials of group h (the h groups before it contribute)
__device__ __host__ inline int64_t bpart_group_base(int64_t h) { return (int64_t)kGroup * (h * (h - 1) / 2); }

// Own-side fragments: one per (row, wavefront that worked on the row), in
// float4 elements, 2 kGroup of them: for each 64 slots of the group, their
// 64 sums (x y z, w unused), then those sums' 64 corrections (slot s: sums
// at frag_elem(s), corrections 64 further).  A wavefront adds a unit's
// own-side sums to LDS sums and those to the fragment every kFoldsPerFlush
// units and at the end of its stay on the row.  The first flush writes the
// sums (their corrections are exactly zero).  When no wavefront stays longer
// than kFoldsPerFlush units on a row, every fragment is flushed once: it is
// plain, and neither the far field nor the combine touch its corrections;
// otherwise the first flush writes zero corrections too.
constexpr int kFoldsPerFlush = 16;
constexpr int kFragElems = 2 * kGroup;
__device__ __host__ inline int frag_elem(int slot) { return 128 * (slot >> 6) + (slot & 63); }
__device__ __host__ inline bool fragments_plain(int64_t max_units_per_visit)
{
    return max_units_per_visit <= kFoldsPerFlush;
}

}  // namespace grasph
