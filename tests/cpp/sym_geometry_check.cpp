// Host check of the symmetric sweep's work geometry
// (grasph2_amd/csrc/sym_geometry.hpp), no GPU: replays on the host how the
// far kernel's workgroups walk their shares of the rows (as far_work in
// sym_sweep.hip does, with partner super-blocks cut into Q = 1, 2, 4, 8
// or 16 pieces) and how the combine reads the partials back (sym_combine_kernel),
// for many particle counts and workgroup counts, and checks that
//  - every pair of super-blocks in different groups is computed exactly once
//    and no other pair is;
//  - every own-side fragment and partner-side partial index is in bounds of
//    the arrays sym_geometry() sizes, and no partial is written twice;
//  - every fragment and partial the combine reads was written, and every
//    fragment written is read;
//  - every fragment is flushed at least once, and exactly once when the
//    geometry calls the fragments plain (the combine then reads no
//    correction planes).
// Also the column-batched far field (partner groups in batches, each with
// its own partials and fragments), and the batch form: the rows table
// covers each (instance, row) once, and each row's pieces cover its
// partner groups once.
// Prints "ok <cases>" and exits 0, or prints the first failure and exits 1.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <utility>
#include <vector>

#include "../../grasph2_amd/csrc/sym_geometry.hpp"

using namespace grasph;

static int fail(const char* what, int64_t n, int64_t W, int64_t a = -1, int64_t b = -1)
{
    std::printf("FAIL %s n=%lld W=%lld (%lld, %lld)\n", what, (long long)n, (long long)W, (long long)a, (long long)b);
    return 1;
}

static int check(int64_t n, int64_t wanted, int Q)
{
    const int64_t nsb = (n + kSB - 1) / kSB;
    const int64_t G = (nsb + kGroupSB - 1) / kGroupSB;
    const int64_t npos = rows_with_units(nsb);
    const int64_t U = row_prefix(npos, nsb, Q);
    int64_t sum = 0;
    for (int64_t g = 0; g < G; ++g) sum += row_units(g, nsb, Q);
    if (sum != U) return fail("unit total", n, wanted, sum, U);
    if (U == 0) return 0;  // no far field: the near field covers the one group
    const int64_t W = far_workgroups(wanted, U);
    const int64_t apart_elems = (int64_t)kFragElems * (G + wanted);
    const int64_t bpart_elems = 3 * bpart_group_base(G);

    std::vector<char> seen((size_t)(nsb * nsb * Q), 0);
    std::vector<char> bw((size_t)(bpart_elems / 3), 0);
    std::set<std::pair<int64_t, int64_t>> frag;  // (row, workgroup)
    std::map<std::pair<int64_t, int64_t>, int> flushes;  // as far_work flushes its LDS sums
    const bool plain = fragments_plain((U + W - 1) / W);
    for (int64_t k = 0; k < W; ++k) {
        int folds = 0;
        const int64_t u0 = k * U / W, u1 = (k + 1) * U / W;
        if (u0 >= u1) return fail("empty share", n, W, k);
        int64_t g = row_of_unit(u0, nsb, Q), s = u0 - row_prefix(g, nsb, Q);
        if (s < 0 || s >= row_units(g, nsb, Q)) return fail("row search", n, W, u0, g);
        for (int64_t u = u0; u < u1; ++u) {
            if (s == row_units(g, nsb, Q)) {
                frag.insert({g, k});
                if (folds > 0) ++flushes[{g, k}];
                folds = 0;
                ++g;
                s = 0;
                if (g >= npos) return fail("row past the last", n, W, g);
            }
            const int64_t J = kGroupSB * (g + 1) + s / Q, piece = s % Q;
            if (J >= nsb) return fail("partner super-block", n, W, g, J);
            for (int w = 0; w < kGroupSB; ++w) {
                const int64_t sbA = g * kGroupSB + w;
                if (sbA >= nsb) continue;
                char& c = seen[(size_t)((sbA * nsb + J) * Q + piece)];
                if (c) return fail("pair twice", n, W, sbA, J);
                c = 1;
            }
            const int64_t len = kSB / Q, b0 = kSB * J + len * piece;
            for (int64_t B = b0; B < b0 + len && B < n; ++B) {
                const int64_t h = B / kGroup;
                const int64_t x = bpart_group_base(h) + g * kGroup + (B - h * kGroup);
                if (x < 0 || 3 * x + 2 >= bpart_elems) return fail("partial bounds", n, W, g, B);
                if (bw[(size_t)x]) return fail("partial twice", n, W, g, B);
                bw[(size_t)x] = 1;
            }
            ++s;
            if (++folds == kFoldsPerFlush) {
                ++flushes[{g, k}];
                folds = 0;
            }
        }
        frag.insert({g, k});
        if (folds > 0) ++flushes[{g, k}];
    }
    for (const auto& f : frag) {
        const int c = flushes.count(f) ? flushes[f] : 0;
        if (c < 1) return fail("fragment never flushed", n, W, f.first, f.second);
        if (plain && c != 1) return fail("plain fragment flushed twice", n, W, f.first, f.second);
    }
    for (int64_t I = 0; I < nsb; ++I)
        for (int64_t J = 0; J < nsb; ++J)
            for (int p = 0; p < Q; ++p)
                if ((I / kGroupSB < J / kGroupSB) != (seen[(size_t)((I * nsb + J) * Q + p)] != 0))
                    return fail("coverage", n, W, I, J);
    for (const auto& f : frag)
        if ((f.first + f.second + 1) * (int64_t)kFragElems > apart_elems) return fail("fragment bounds", n, W, f.first, f.second);
    // what the combine reads
    size_t reads = 0;
    for (int64_t g = 0; g < G; ++g) {
        if (g < npos) {
            const int64_t k0 = wg_of_unit(row_prefix(g, nsb, Q), U, W);
            const int64_t k1 = wg_of_unit(row_prefix(g + 1, nsb, Q) - 1, U, W);
            for (int64_t k = k0; k <= k1; ++k) {
                if (k * U / W == (k + 1) * U / W) continue;
                if (!frag.count({g, k})) return fail("fragment read, not written", n, W, g, k);
                ++reads;
            }
        }
        for (int64_t t = g * kGroup; t < (g + 1) * kGroup && t < n; ++t)
            for (int64_t gp = 0; gp < g; ++gp)
                if (!bw[(size_t)(bpart_group_base(g) + gp * kGroup + (t - g * kGroup))])
                    return fail("partial read, not written", n, W, gp, t);
    }
    if (reads != frag.size()) return fail("fragment written, not read", n, W, (int64_t)reads, (int64_t)frag.size());
    return 0;
}

// Column batches: the far field of n particles in batches of partner groups
// (sym_columns with max_elems floats of partials each), each replayed as
// far_work does it with its own workgroup count, and the column combine's
// reads after each batch.  Every pair of groups is computed once over all
// batches, and each batch's partials and fragments stay in its arrays.
static int check_columns(int64_t n, int64_t wanted, int Q, int64_t max_elems)
{
    const int64_t nsb = (n + kSB - 1) / kSB, G = nsb;
    const std::vector<int> c = sym_columns(nsb, max_elems);
    if (c.front() != 1 || c.back() != (nsb > 1 ? nsb : 1)) return fail("column ends", n, wanted, c.front(), c.back());
    std::vector<char> seen((size_t)(nsb * nsb * Q), 0);
    const int64_t apart_elems = (int64_t)kFragElems * (G + wanted);
    for (size_t b = 0; b + 1 < c.size(); ++b) {
        const int64_t c0 = c[b], c1 = c[b + 1];
        if (c1 <= c0 && nsb > 1) return fail("empty column", n, wanted, c0, c1);
        const int64_t elems = 3 * col_bpart_base(c1, c0);
        if (elems > max_elems && c1 > c0 + 1) return fail("column partials", n, wanted, c0, c1);
        const int64_t U = col_row_prefix(col_rows(c1), c0, c1, Q);
        int64_t sum = 0;
        for (int64_t g = 0; g < nsb; ++g) sum += col_row_units(g, c0, c1, Q);
        if (sum != U) return fail("column unit total", n, wanted, sum, U);
        if (U == 0) continue;
        const int64_t W = far_workgroups(wanted, U);
        const bool plain = fragments_plain((U + W - 1) / W);
        std::vector<char> bw((size_t)(elems / 3), 0);
        std::set<std::pair<int64_t, int64_t>> frag;
        std::map<std::pair<int64_t, int64_t>, int> flushes;
        for (int64_t k = 0; k < W; ++k) {
            const int64_t u0 = k * U / W, u1 = (k + 1) * U / W;
            if (u0 >= u1) return fail("column empty share", n, W, k);
            int64_t g = col_row_of_unit(u0, c0, c1, Q), s = u0 - col_row_prefix(g, c0, c1, Q);
            if (s < 0 || s >= col_row_units(g, c0, c1, Q)) return fail("column row search", n, W, u0, g);
            int folds = 0;
            for (int64_t u = u0; u < u1; ++u) {
                if (s == col_row_units(g, c0, c1, Q)) {
                    frag.insert({g, k});
                    if (folds > 0) ++flushes[{g, k}];
                    folds = 0;
                    ++g;
                    s = 0;
                    if (g >= col_rows(c1)) return fail("column row past the last", n, W, g);
                }
                const int64_t J = col_first_group(g, c0) + s / Q, piece = s % Q;
                if (J < c0 || J >= c1 || J <= g) return fail("column partner group", n, W, g, J);
                char& sc = seen[(size_t)((g * nsb + J) * Q + piece)];
                if (sc) return fail("column pair twice", n, W, g, J);
                sc = 1;
                const int64_t len = kSB / Q, b0 = kSB * J + len * piece;
                for (int64_t B = b0; B < b0 + len && B < n; ++B) {
                    // as far_unit indexes it: bpart_group_base(J) - bpart_group_base(c0) + g kGroup + local
                    const int64_t x = bpart_group_base(J) - bpart_group_base(c0) + g * kGroup + (B - J * kGroup);
                    if (x < 0 || 3 * x + 2 >= elems) return fail("column partial bounds", n, W, g, B);
                    if (bw[(size_t)x]) return fail("column partial twice", n, W, g, B);
                    bw[(size_t)x] = 1;
                }
                ++s;
                if (++folds == kFoldsPerFlush) {
                    ++flushes[{g, k}];
                    folds = 0;
                }
            }
            frag.insert({g, k});
            if (folds > 0) ++flushes[{g, k}];
        }
        for (const auto& f : frag) {
            const int cnt = flushes.count(f) ? flushes[f] : 0;
            if (cnt < 1 || (plain && cnt != 1)) return fail("column flushes", n, W, f.first, f.second);
            if ((f.first + f.second + 1) * (int64_t)kFragElems > apart_elems) return fail("column fragment bounds", n, W);
        }
        // what sym_column_combine_kernel reads after this batch
        size_t reads = 0;
        for (int64_t g = 0; g < G; ++g) {
            if (g < col_rows(c1)) {
                const int64_t k0 = wg_of_unit(col_row_prefix(g, c0, c1, Q), U, W);
                const int64_t k1 = wg_of_unit(col_row_prefix(g + 1, c0, c1, Q) - 1, U, W);
                for (int64_t k = k0; k <= k1; ++k) {
                    if (!frag.count({g, k})) return fail("column fragment read, not written", n, W, g, k);
                    ++reads;
                }
            }
            if (g >= c0 && g < c1)
                for (int64_t t = g * kGroup; t < (g + 1) * kGroup && t < n; ++t)
                    for (int64_t gp = 0; gp < g; ++gp)
                        if (!bw[(size_t)(col_bpart_base(g, c0) + gp * kGroup + (t - g * kGroup))])
                            return fail("column partial read, not written", n, W, gp, t);
        }
        if (reads != frag.size()) return fail("column fragment written, not read", n, W, (int64_t)reads, (int64_t)frag.size());
    }
    for (int64_t I = 0; I < nsb; ++I)
        for (int64_t J = 0; J < nsb; ++J)
            for (int p = 0; p < Q; ++p)
                if ((I < J) != (seen[(size_t)((I * nsb + J) * Q + p)] != 0)) return fail("column coverage", n, wanted, I, J);
    // one batch is the unbatched far field
    for (int64_t g = 0; g <= rows_with_units(nsb); ++g)
        if (col_row_prefix(g, 1, nsb, Q) != row_prefix(g, nsb, Q)) return fail("column prefix", n, wanted, g);
    return 0;
}

// batch form: one row per (instance, g < rows_with_units), fragments numbered row0 + g
static int check_batch(const std::vector<int64_t>& sizes)
{
    int64_t row0 = 0, bpart0 = 0;
    std::set<std::pair<int64_t, int64_t>> rows;
    for (size_t k = 0; k < sizes.size(); ++k) {
        const int64_t n = sizes[k], nsb = (n + kSB - 1) / kSB, G = (nsb + kGroupSB - 1) / kGroupSB;
        const int64_t nrows = rows_with_units(nsb);
        for (int64_t g = 0; g < nrows; ++g) {
            const int64_t u = row_units(g, nsb);
            if (u <= 0) return fail("batch empty row", n, (int64_t)k, g);
            if (!rows.insert({(int64_t)k, g}).second) return fail("batch row twice", n, (int64_t)k, g);
            // the pieces: consecutive, non-empty, covering the row's partner groups once
            const int P = batch_row_pieces(u);
            if (P < 1 || P > kBatchPieces) return fail("batch pieces", n, (int64_t)k, g);
            int64_t next = 0;
            for (int h = 0; h < P; ++h) {
                const int64_t b = batch_piece_begin(u, h), e = batch_piece_begin(u, h + 1);
                if (b != next || e <= b) return fail("batch piece range", n, (int64_t)k, g, h);
                next = e;
            }
            if (next != u) return fail("batch pieces cover", n, (int64_t)k, g);
        }
        if (row_units(nrows, nsb) != 0) return fail("batch row missing", n, (int64_t)k, nrows);
        row0 += nrows;
        bpart0 += 3 * bpart_group_base(G);
    }
    return bpart0 >= 0 && row0 == (int64_t)rows.size() ? 0 : fail("batch totals", row0, (int64_t)rows.size());
}

int main()
{
    const int64_t ns[] = {1, 63, 64, 255, 256, 257, 1023, 1024, 1025, 2047, 2048, 3000, 4096, 5000, 6000,
                          8191, 16384, 20037, 65536, 100000, 131072};
    const int64_t ws[] = {1, 2, 3, 7, 12, 64, 512, 1024, 100000};
    int cases = 0;
    for (int64_t n : ns)
        for (int64_t w : ws)
            for (int Q : {1, 2, 4, 8, 16}) {
                if (check(n, w, Q)) return 1;
                ++cases;
            }
    srand(7);
    for (int i = 0; i < 60; ++i) {
        const int64_t n = 1 + rand() % 150000, w = 1 + rand() % 2048;
        if (check(n, w, far_split(n, w))) return 1;
        ++cases;
    }
    // the split chosen for a count is the one its workgroup count gives back
    for (int64_t n : ns)
        for (int64_t w : ws) {
            const int Q = far_split(n, w);
            const int64_t nsb = (n + kSB - 1) / kSB;
            const int64_t U = row_prefix(rows_with_units(nsb), nsb, Q);
            if (U > 0 && far_split(n, far_workgroups(w, U)) != Q) return fail("split not stable", n, w, Q);
            ++cases;
        }
    // column batches: from one group per batch to the whole far field
    const int64_t cns[] = {1, 1025, 3000, 20037, 40000, 131072};
    const int64_t cel[] = {1, 3 * 1024 * 3, 3 * 1024 * 20, 3 * 1024 * 300, int64_t(1) << 40};
    for (int64_t n : cns)
        for (int64_t e : cel)
            for (int64_t w : {1, 7, 64, 2048})
                for (int Q : {2, 4, 16}) {
                    if (check_columns(n, w, Q, e)) return 1;
                    ++cases;
                }
    // known answers (partials of batch [c0, c1): 3 kGroup (c1 (c1 - 1) - c0 (c0 - 1)) / 2 floats):
    // 12 GiB holds the whole far field up to 1448 groups (1448 * 1447 / 2 * 3072 =
    // 3218313216 <= 3221225472 floats; 1449 groups need 3222761472); 2048 groups
    // take two batches, [1, 1448) and [1448, 2048); 1 GiB and 1024 groups take
    // seven, the first [1, 418) (418 * 417 / 2 * 3072 = 267734016 <= 268435456 <
    // 269018112 = 419 * 418 / 2 * 3072)
    {
        const int64_t gib12 = (int64_t(12) << 30) / 4, gib1 = (int64_t(1) << 30) / 4;
        if (sym_columns(1448, gib12).size() != 2) return fail("1448 groups in one batch", 1448, 0, 0);
        if (sym_columns(1449, gib12).size() != 3) return fail("1449 groups in two batches", 1449, 0, 0);
        const std::vector<int> two = sym_columns(2048, gib12);
        if (two.size() != 3 || two[1] != 1448) return fail("2048 groups in two batches", 2048, (int64_t)two.size(), 0);
        const std::vector<int> seven = sym_columns(1024, gib1);
        if (seven.size() != 8 || seven[1] != 418) return fail("1024 groups, 1 GiB: seven batches", 1024, (int64_t)seven.size(), 0);
        cases += 4;
    }
    std::vector<int64_t> sizes;
    for (int i = 0; i < 200; ++i) sizes.push_back(rand() % 17000);
    if (check_batch(sizes)) return 1;
    std::printf("ok %d\n", cases + 1);
    return 0;
}
