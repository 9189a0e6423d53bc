// batch.hip -- batched ensembles of independent simulations (gfx950).
//
// A batch packs many instances back to back (see grasph_batch in
// grasph2.h).  Per step:
//   order   per-instance Morton order (eager, every reorder_interval steps)
//   sweep   one persistent launch: every wavefront pulls (instance, tile)
//           items from an atomic counter, longest instances first, so the
//           uneven instances keep all CUs busy to the end
//   passes  density / strain / record / accel and the leapfrog, one launch
//           each over all instances, a block never straddling two instances
// grasph_batch_step captures one step as a HIP graph and replays it.
// grasph_batch_step_loop runs the same steps through the single-instance
// entry points, one instance after the other: the plain launch loop.
// grasph_autotune times the sweep on the batch to choose the particles per
// lane of each size bucket and the persistent launch shape.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace grasph {

struct BatchState {
    // layout of the batch the tables describe
    int n_inst = 0;
    int64_t ntotal = 0;
    std::vector<int64_t> offsets;
    std::vector<grasph_params> params;
    uint32_t flags = 0;
    int tuned_ni[4] = {0, 0, 0, 0};  // i per lane the items table was built with
    int items_ni_max = 1;            // largest ni among the items
    int64_t items_sym_min = -1;      // instances from this size on were left out of the items (-1: none)

    // device tables
    BatchInst* d_insts = nullptr;
    DevParams* d_prm = nullptr;
    int* d_inst_of = nullptr;
    int4* d_items = nullptr;
    int n_items = 0;
    int4* d_blk = nullptr;
    int n_blk = 0;
    int* d_queue = nullptr;
    float* d_ts = nullptr;  // variable step: {current dt, next dt} per instance, zero when the tables are built
    float* d_um = nullptr;  // {uniform, mass} per instance, written before each gravity sweep
    // slot-ordered positions for the sweep (SPH with a Morton order)
    float4* d_xs = nullptr;
    // symmetric sweep (option batch_symmetric): tables and work arrays
    SymInst* d_sinst = nullptr;
    SymRow* d_srows = nullptr;
    int n_srows = 0;
    int4* d_sblk = nullptr;  // 256-slot blocks of the instances in the symmetric sweep
    int n_sblk = 0;
    int64_t sym_min = -1;    // the size threshold the tables were built with (-1: none)
    int sym_max_words = 0;
    bool sym_plain = true;  // SymBatchTables::plain
    float4* d_sbox = nullptr;
    uint32_t* d_sbits = nullptr;
    float4* d_snear = nullptr;
    float4* d_sapart = nullptr;
    float* d_sbpart = nullptr;

    // Morton order
    float* d_bbox = nullptr;
    uint64_t* d_keys = nullptr;
    uint64_t* d_keys_alt = nullptr;
    int* d_vals = nullptr;
    int* d_perm = nullptr;
    void* d_sort_tmp = nullptr;
    size_t sort_bytes = 0;
    int key_bits = 31;
    bool order_valid = false;
    const void* order_src = nullptr;
    int64_t evals_since_order = 0;

    // neighbour lists and pass records
    int* d_nbr = nullptr;
    int* d_cnt = nullptr;
    int64_t nbr_cap = 0;  // slots * kmax allocated
    int kmax = 0;
    SlotRec* d_rec = nullptr;  // slot records of the passes
    int64_t rec_cap = 0;

    // graph of one step
    hipStream_t gs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // the symmetric sweep's far field: its own stream, forked and joined
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_lists = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string graph_key;

    // per-instance contexts of the launch-loop baseline
    std::vector<grasph_ctx*> loop;
};

namespace {

int bucket_of(int64_t n) { return n < 2048 ? 0 : n < 4096 ? 1 : n < 8192 ? 2 : 3; }

void free_ptr(void* p)
{
    if (p) (void)hipFree(p);
}

void drop_graph(BatchState& b)
{
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    b.graph_key.clear();
}

void free_tables(BatchState& b)
{
    void* ps[] = {b.d_insts, b.d_prm, b.d_inst_of, b.d_items, b.d_blk, b.d_queue, b.d_bbox, b.d_keys, b.d_keys_alt,
                  b.d_vals, b.d_perm, b.d_sort_tmp, b.d_ts, b.d_um, b.d_xs, b.d_sinst, b.d_srows, b.d_sbox,
                  b.d_sbits, b.d_snear, b.d_sapart, b.d_sbpart, b.d_sblk};
    for (void* p : ps) free_ptr(p);
    b.d_sinst = nullptr; b.d_srows = nullptr; b.d_sbox = nullptr; b.d_sbits = nullptr; b.d_snear = nullptr;
    b.d_sapart = nullptr; b.d_sbpart = nullptr; b.d_sblk = nullptr;
    b.n_srows = 0;
    b.n_sblk = 0;
    b.sym_min = -1;
    b.sym_max_words = 0;
    b.sym_plain = true;
    b.d_insts = nullptr; b.d_prm = nullptr; b.d_inst_of = nullptr; b.d_items = nullptr; b.d_blk = nullptr;
    b.d_queue = nullptr; b.d_bbox = nullptr; b.d_keys = nullptr; b.d_keys_alt = nullptr; b.d_vals = nullptr;
    b.d_perm = nullptr; b.d_sort_tmp = nullptr; b.d_ts = nullptr; b.d_um = nullptr;
    b.d_xs = nullptr;
    b.sort_bytes = 0;
    b.order_valid = false;
    b.n_items = 0;
    b.n_blk = 0;
}

template <typename T>
int upload(Ctx& c, T*& dst, const std::vector<T>& src, const char* what)
{
    free_ptr(dst);
    dst = nullptr;
    if (hipMalloc((void**)&dst, std::max<size_t>(1, src.size()) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        dst = nullptr;
        return fail(c, GRASPH_E_NOMEM, std::string("device allocation of ") + what + " failed");
    }
    if (!src.empty() && hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        return hip_fail(c, hipGetLastError(), what);
    return 0;
}

// work items of the persistent sweep: instances longest first, tiles of
// 64 * ni slots with ni from the instance's size bucket
int build_items(Ctx& c, BatchState& b)
{
    std::vector<int> order(b.n_inst);
    for (int k = 0; k < b.n_inst; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (b.offsets[x + 1] - b.offsets[x]) > (b.offsets[y + 1] - b.offsets[y]);
    });
    std::vector<int4> items;
    b.items_ni_max = 1;
    const int64_t sym_min = c.batch_symmetric ? c.batch_symmetric_min : -1;
    for (int k : order) {
        const int64_t off = b.offsets[k], n = b.offsets[k + 1] - off;
        if (sym_min >= 0 && n >= sym_min) continue;  // swept by the symmetric sweep
        const int ni = c.batch_ni[bucket_of(n)];
        if (n > 0) b.items_ni_max = std::max(b.items_ni_max, ni);
        for (int64_t s = off; s < off + n; s += 64 * ni) items.push_back(make_int4(k, (int)s, ni, 0));
    }
    b.n_items = (int)items.size();
    b.items_sym_min = sym_min;
    for (int q = 0; q < 4; ++q) b.tuned_ni[q] = c.batch_ni[q];
    drop_graph(b);
    return upload(c, b.d_items, items, "batch work items");
}

int validate(Ctx& c, const grasph_particles* p, const grasph_batch* bt)
{
    int r;
    if ((r = check_particles(c, p))) return r;
    if (!bt || !bt->offsets || !bt->params) return fail(c, GRASPH_E_INVALID, "batch, offsets or params is NULL");
    if (bt->n_instances < 1) return fail(c, GRASPH_E_INVALID, "a batch needs at least one instance");
    if (bt->offsets[0] != 0) return fail(c, GRASPH_E_INVALID, "offsets[0] must be 0");
    for (int k = 0; k < bt->n_instances; ++k)
        if (bt->offsets[k + 1] < bt->offsets[k]) return fail(c, GRASPH_E_INVALID, "offsets must not decrease");
    if (bt->offsets[bt->n_instances] != p->n)
        return fail(c, GRASPH_E_INVALID, "offsets[n_instances] must equal the particle count");
    for (int k = 0; k < bt->n_instances; ++k) {
        if ((r = check_params(c, &bt->params[k]))) return r;
        if (bt->params[k].flags != bt->params[0].flags)
            return fail(c, GRASPH_E_INVALID, "all instances of a batch need the same flags");
    }
    return 0;
}

int ensure_sym(Ctx& c, BatchState& b);

// (Re)build the tables when the batch layout or parameters changed.
int prepare(Ctx& c, const grasph_particles* p, const grasph_batch* bt)
{
    int r;
    if ((r = validate(c, p, bt))) return r;
    if (!c.batch) c.batch = new BatchState;
    BatchState& b = *c.batch;
    const int n_inst = bt->n_instances;
    bool same = b.n_inst == n_inst && b.ntotal == p->n;
    if (same)
        same = std::equal(b.offsets.begin(), b.offsets.end(), bt->offsets) &&
               std::memcmp(b.params.data(), bt->params, n_inst * sizeof(grasph_params)) == 0;
    if (!same) {
        drop_graph(b);
        free_tables(b);
        b.n_inst = n_inst;
        b.ntotal = p->n;
        b.offsets.assign(bt->offsets, bt->offsets + n_inst + 1);
        b.params.assign(bt->params, bt->params + n_inst);
        b.flags = bt->params[0].flags;
        std::vector<BatchInst> insts(n_inst);
        std::vector<DevParams> prm(n_inst);
        std::vector<int> inst_of((size_t)p->n);
        std::vector<int4> blk;
        for (int k = 0; k < n_inst; ++k) {
            prm[k] = make_dev_params(&bt->params[k]);
            const int64_t off = bt->offsets[k], n = bt->offsets[k + 1] - off;
            insts[k] = BatchInst{(int)off, (int)n, prm[k].H2, prm[k].cand};
            for (int64_t i = off; i < off + n; ++i) inst_of[(size_t)i] = k;
            for (int64_t s = off; s < off + n; s += 256)
                blk.push_back(make_int4(k, (int)s, (int)std::min<int64_t>(s + 256, off + n), 0));
        }
        b.n_blk = (int)blk.size();
        if ((r = upload(c, b.d_insts, insts, "batch instance table")) || (r = upload(c, b.d_prm, prm, "batch params")) ||
            (r = upload(c, b.d_inst_of, inst_of, "batch instance map")) || (r = upload(c, b.d_blk, blk, "batch blocks")))
            return r;
        std::vector<int> zero(64, 0);
        if ((r = upload(c, b.d_queue, zero, "batch work counter"))) return r;
        std::vector<float> ts0(2 * (size_t)n_inst, 0.0f);
        if ((r = upload(c, b.d_ts, ts0, "batch time-step state")) ||
            (r = upload(c, b.d_um, ts0, "batch mass records")))
            return r;
        b.key_bits = 30;
        while ((1LL << (b.key_bits - 30)) < n_inst) ++b.key_bits;
        const size_t nt = (size_t)std::max<int64_t>(p->n, 1);
        if (hipMalloc((void**)&b.d_bbox, (size_t)n_inst * 6 * sizeof(float)) != hipSuccess ||
            hipMalloc((void**)&b.d_keys, nt * sizeof(uint64_t)) != hipSuccess ||
            hipMalloc((void**)&b.d_keys_alt, nt * sizeof(uint64_t)) != hipSuccess ||
            hipMalloc((void**)&b.d_vals, nt * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&b.d_perm, nt * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&b.d_xs, nt * sizeof(float4)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, GRASPH_E_NOMEM, "device allocation of the batch order failed");
        }
        b.sort_bytes = batch_order_tmp_bytes((int)p->n, b.key_bits);
        if (hipMalloc(&b.d_sort_tmp, std::max<size_t>(b.sort_bytes, 1)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, GRASPH_E_NOMEM, "device allocation of the batch sort storage failed");
        }
        b.tuned_ni[0] = 0;  // force the items table
    }
    const int64_t want_sym_min = c.batch_symmetric ? c.batch_symmetric_min : -1;
    if (std::memcmp(b.tuned_ni, c.batch_ni, sizeof b.tuned_ni) != 0 || !b.d_items || b.items_sym_min != want_sym_min)
        if ((r = build_items(c, b))) return r;
    // lists and records
    // the capacity set on the context, not one its single-instance calls grew to
    const int64_t need = (int64_t)list_plane(b.ntotal, c.max_neighbours_set);
    if (!b.d_nbr || b.nbr_cap < need || b.kmax != c.max_neighbours_set) {
        free_ptr(b.d_nbr);
        free_ptr(b.d_cnt);
        b.d_nbr = nullptr;
        b.d_cnt = nullptr;
        drop_graph(b);
        if (hipMalloc((void**)&b.d_nbr, (size_t)std::max<int64_t>(need, 1) * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&b.d_cnt, (size_t)std::max<int64_t>(b.ntotal, 1) * sizeof(int)) != hipSuccess) {
            (void)hipGetLastError();
            b.nbr_cap = 0;
            return fail(c, GRASPH_E_NOMEM, "device allocation of the batch neighbour lists failed");
        }
        b.nbr_cap = need;
        b.kmax = c.max_neighbours_set;
    }
    if (b.rec_cap < b.ntotal || !b.d_rec) {
        drop_graph(b);
        int64_t cap = b.rec_cap;
        if ((r = grow(c, b.d_rec, cap, b.ntotal, "batch slot records"))) return r;
        b.rec_cap = cap;
    }
    // the symmetric sweep's tables, here and not at first use: a step may be under graph capture then
    if (c.batch_symmetric && (b.flags & (GRASPH_SPH | GRASPH_SELF_GRAVITY)) && (r = ensure_sym(c, b))) return r;
    if (!b.gs) {
        if (hipStreamCreateWithFlags(&b.gs, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b.ev_in, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b.ev_out, hipEventDisableTiming) != hipSuccess ||
            hipStreamCreateWithFlags(&b.side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b.ev_lists, hipEventDisableTiming) != hipSuccess)
            return hip_fail(c, hipGetLastError(), "batch stream");
    }
    return 0;
}

int ensure_order(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s, bool count_eval)
{
    if (c.reorder_interval <= 0) return 0;
    if (!b.order_valid || b.order_src != (const void*)p->posm || b.evals_since_order >= c.reorder_interval) {
        BatchOrder O;
        O.posm = (const float4*)p->posm;
        O.inst_of = b.d_inst_of;
        O.insts = b.d_insts;
        O.n_inst = b.n_inst;
        O.ntotal = (int)b.ntotal;
        O.bbox = b.d_bbox;
        O.keys = b.d_keys;
        O.keys_alt = b.d_keys_alt;
        O.vals = b.d_vals;
        O.perm = b.d_perm;
        O.sort_tmp = b.d_sort_tmp;
        O.sort_tmp_bytes = b.sort_bytes;
        O.key_bits = b.key_bits;
        ProfScope ps(&c, s, KC_REORDER);
        hipError_t e = launch_batch_order(O, s);
        if (e != hipSuccess) return hip_fail(c, e, "batch Morton order");
        b.order_valid = true;
        b.order_src = p->posm;
        b.evals_since_order = 0;
    }
    if (count_eval) b.evals_since_order++;
    return 0;
}

const int* order_of(const Ctx& c, const BatchState& b) { return c.reorder_interval > 0 ? b.d_perm : nullptr; }

BlockTables tables(const BatchState& b)
{
    BlockTables T;
    T.blk = b.d_blk;
    T.insts = b.d_insts;
    T.prm = b.d_prm;
    T.nblocks = b.n_blk;
    return T;
}

int device_cus(Ctx& c)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        cus = 256;
    }
    return cus;
}

// The slot order of the sweep and the passes: the Morton order for SPH (the
// neighbour work wants it), none for gravity alone.
const int* slot_order(const Ctx& c, const BatchState& b)
{
    return ((b.flags & GRASPH_SPH) && b.order_valid) ? order_of(c, b) : nullptr;
}

// slot-ordered positions (with a slot order) and the passes' slot records
// (positions, velocities and, when no density pass writes them, densities)
int batch_gather(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s)
{
    if (!(b.flags & GRASPH_SPH)) return 0;
    const int* perm = slot_order(c, b);
    const bool dens_pass = !(b.flags & GRASPH_INTEGRATE_DENSITY);
    ProfScope ps(&c, s, KC_REORDER);
    hipError_t e = launch_gather(perm, (int)b.ntotal, (const float4*)p->posm, (const float4*)p->vel,
                                 dens_pass ? nullptr : p->density, perm ? b.d_xs : nullptr, b.d_rec, s);
    if (e != hipSuccess) return hip_fail(c, e, "batched slot-order gather launch");
    return 0;
}

// the persistent sweep; with a slot order, batch_gather must have run for the current positions
int launch_sweep_only(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s, bool mass_done = false)
{
    const uint32_t f = b.flags;
    const DevParams P = make_dev_params(&b.params[0]);
    hipError_t e = launch_reset_queue(b.d_queue, 1, s);
    if (e != hipSuccess) return hip_fail(c, e, "batch work counter reset");
    BatchSweepLaunch L;
    std::memset(&L, 0, sizeof L);
    L.perm = slot_order(c, b);
    L.xs = L.perm ? b.d_xs : (const float4*)p->posm;
    L.items = b.d_items;
    L.n_items = b.n_items;
    L.insts = b.d_insts;
    L.queue = b.d_queue;
    L.acc = (float4*)p->acc;
    L.nbr = b.d_nbr;
    L.nbr_cnt = b.d_cnt;
    L.ntotal = (int)b.ntotal;
    L.kmax = b.kmax;
    L.waves_per_block = c.batch_wpb;
    L.blocks = device_cus(c) * c.batch_bpc;
    L.ni_max = b.items_ni_max;
    L.grav = f & GRASPH_SELF_GRAVITY;
    L.want_nbr = f & GRASPH_SPH;
    bool soft = true;
    for (const auto& q : b.params) soft = soft && q.H > 0.0f;
    L.soft = soft;
    if (L.grav) {
        if (!mass_done) {
            e = launch_mass_check((const float4*)p->posm, b.d_insts, b.n_inst, 0, b.d_um, s);
            if (e != hipSuccess) return hip_fail(c, e, "batch mass check launch");
        }
        L.umass = b.d_um;
    }
    ProfScope ps(&c, s, KC_SWEEP);
    e = launch_batch_sweep(L, P, s);
    if (e != hipSuccess) return hip_fail(c, e, "batched sweep launch");
    return 0;
}

// tables and work arrays of the symmetric sweep, built once per batch layout
int ensure_sym(Ctx& c, BatchState& b)
{
    if (b.d_sinst && b.sym_min == c.batch_symmetric_min) return 0;
    const SymBatchTables T = sym_batch_tables(b.offsets, (b.flags & GRASPH_SPH) != 0, c.batch_symmetric_min);
    int r;
    free_ptr(b.d_sinst); free_ptr(b.d_srows); free_ptr(b.d_sblk);
    b.d_sinst = nullptr; b.d_srows = nullptr; b.d_sblk = nullptr;
    if ((r = upload(c, b.d_sinst, T.insts, "batch symmetric-sweep instances")) ||
        (r = upload(c, b.d_srows, T.rows, "batch symmetric-sweep rows")) ||
        (r = upload(c, b.d_sblk, T.blk, "batch symmetric-sweep blocks")))
        return r;
    b.n_srows = (int)T.rows.size();
    b.n_sblk = (int)T.blk.size();
    b.sym_min = c.batch_symmetric_min;
    b.sym_max_words = T.max_words;
    b.sym_plain = T.plain;
    // fresh arrays for the new tables (upload()-style: free, then allocate)
    void* old[] = {b.d_sbox, b.d_sbits, b.d_snear, b.d_sapart, b.d_sbpart};
    for (void* q : old) free_ptr(q);
    b.d_sbox = nullptr; b.d_sbits = nullptr; b.d_snear = nullptr; b.d_sapart = nullptr; b.d_sbpart = nullptr;
    int64_t cap = 0;
    if ((r = grow(c, b.d_sbox, cap, std::max<int64_t>(T.box_elems, 1), "batch symmetric-sweep boxes"))) return r;
    cap = 0;
    if ((r = grow(c, b.d_sbits, cap, std::max<int64_t>(T.bits_elems, 1), "batch symmetric-sweep near bits"))) return r;
    cap = 0;
    if ((r = grow(c, b.d_snear, cap, std::max<int64_t>(2 * b.ntotal, 1), "batch symmetric-sweep near sums"))) return r;
    cap = 0;
    if ((r = grow(c, b.d_sapart, cap, std::max<int64_t>(T.apart_elems, 1), "batch symmetric-sweep own sums"))) return r;
    cap = 0;
    if ((r = grow(c, b.d_sbpart, cap, std::max<int64_t>(T.bpart_elems, 1), "batch symmetric-sweep partner sums")))
        return r;
    drop_graph(b);
    return 0;
}

// The symmetric sweep over the batch's larger instances, begun (after
// launch_sweep_only, which has then run the mass check: mass_done): the far
// field forked onto b.side (fork; else in order on s), the near field's
// lists on s; sym_batch_end adds its in-group gravity, joins and combines.
// The far field reads only the slot-ordered positions and the mass
// records, written on s before the fork.
int sym_batch_begin(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s, bool mass_done, bool fork,
                    SymBatchLaunch& L)
{
    int r;
    if ((r = ensure_sym(c, b))) return r;
    const uint32_t f = b.flags;
    std::memset(&L, 0, sizeof L);
    L.perm = slot_order(c, b);
    L.xs = L.perm ? b.d_xs : (const float4*)p->posm;
    L.ntotal = (int)b.ntotal;
    L.n_inst = b.n_inst;
    L.grav = f & GRASPH_SELF_GRAVITY;
    L.want_nbr = f & GRASPH_SPH;
    L.finish = false;  // the finishing launch of the batch follows
    bool soft = true;
    for (const auto& q : b.params) soft = soft && q.H > 0.0f;
    L.soft = soft;
    L.sinst = b.d_sinst;
    L.insts = b.d_insts;
    L.prm = b.d_prm;
    L.blk = b.d_sblk;
    L.n_blk = b.n_sblk;
    L.inst_of = b.d_inst_of;
    L.rows = b.d_srows;
    L.nrows = b.n_srows;
    L.max_words = b.sym_max_words;
    L.plain = b.sym_plain;
    L.queue = b.d_queue + 1;
    L.wgs = c.sym_workgroups > 0 ? c.sym_workgroups : sym_default_workgroups(device_cus(c));
    L.acc = (float4*)p->acc;
    L.nbr = b.d_nbr;
    L.nbr_cnt = b.d_cnt;
    L.kmax = b.kmax;
    L.vel = (const float4*)p->vel;
    L.xvel = (float4*)p->xvel;
    L.maxvsig = p->maxvsig;
    L.box = b.d_sbox;
    L.bits = b.d_sbits;
    L.near_part = b.d_snear;
    L.apart = b.d_sapart;
    L.bpart = b.d_sbpart;
    hipError_t e;
    if (L.grav) {
        if (!mass_done) {
            e = launch_mass_check((const float4*)p->posm, b.d_insts, b.n_inst, 0, b.d_um, s);
            if (e != hipSuccess) return hip_fail(c, e, "batch mass check launch");
        }
        L.umass = b.d_um;
    }
    if (sym_batch_has_far(L)) {
        hipStream_t fs = s;
        if (fork) {
            if ((e = hipEventRecord(b.ev_fork, s)) != hipSuccess || (e = hipStreamWaitEvent(b.side, b.ev_fork, 0)))
                return hip_fail(c, e, "batched far-field fork");
            fs = b.side;
        }
        {
            ProfScope ps(&c, fs, KC_FAR);
            if ((e = launch_sym_batch_far(L, fs)) != hipSuccess) return hip_fail(c, e, "batched far-field launch");
        }
    }
    {
        ProfScope ps(&c, s, KC_SWEEP);
        if ((e = launch_sym_batch_lists(L, s)) != hipSuccess) return hip_fail(c, e, "batched near-field list launch");
    }
    if (fork && sym_batch_has_far(L)) {
        // the gravity inside the groups needs only the positions: after the
        // far field on its stream, once the lists are done, so that it shares
        // the vector units with the memory-bound passes on s rather than with
        // the list kernel, which the passes wait for
        if (L.want_nbr && ((e = hipEventRecord(b.ev_lists, s)) != hipSuccess ||
                           (e = hipStreamWaitEvent(b.side, b.ev_lists, 0)) != hipSuccess))
            return hip_fail(c, e, "batched list event");
        {
            ProfScope ps(&c, b.side, KC_SWEEP);
            if ((e = launch_sym_batch_ingroup(L, b.side)) != hipSuccess)
                return hip_fail(c, e, "batched near-field gravity launch");
            L.ingroup_done = true;
        }
        if ((e = hipEventRecord(b.ev_join, b.side)) != hipSuccess) return hip_fail(c, e, "batched far-field join event");
    }
    return 0;
}

int sym_batch_end(Ctx& c, BatchState& b, SymBatchLaunch& L, hipStream_t s, bool forked, bool add_acc)
{
    hipError_t e;
    if (!L.ingroup_done) {
        // the gravity inside the groups last on s: only the combine needs it
        ProfScope ps(&c, s, KC_SWEEP);
        if ((e = launch_sym_batch_ingroup(L, s)) != hipSuccess)
            return hip_fail(c, e, "batched near-field gravity launch");
    }
    if (forked && sym_batch_has_far(L) && (e = hipStreamWaitEvent(s, b.ev_join, 0)) != hipSuccess)
        return hip_fail(c, e, "batched far-field join");
    L.add_acc = add_acc;
    ProfScope ps(&c, s, KC_SWEEP);
    if ((e = launch_sym_batch_combine(L, s)) != hipSuccess) return hip_fail(c, e, "batched symmetric combine launch");
    return 0;
}

// the whole symmetric sweep in order on s (the autotuner times it)
int launch_sym_sweep_batch(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s, bool mass_done)
{
    SymBatchLaunch L;
    int r;
    if ((r = sym_batch_begin(c, b, p, s, mass_done, false, L))) return r;
    return sym_batch_end(c, b, L, s, false, false);
}

int batch_passes(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s);

// derivative passes of all instances (the order must be current), then, with
// the variable step, every instance's next dt := max_timestep as
// calcAcceleration::do_after does (computeDerivatives.h:305-308)
int batch_derivative_launches(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s)
{
    int r = batch_passes(c, b, p, s);
    if (r || !(b.flags & GRASPH_VARIABLE_TIMESTEP)) return r;
    ProfScope ps(&c, s, KC_OTHER);
    hipError_t e = launch_vts_set_next(b.d_ts, b.n_inst, b.d_prm, make_dev_params(&b.params[0]), 1, s);
    if (e != hipSuccess) return hip_fail(c, e, "batched time-step reset launch");
    return 0;
}

int batch_passes(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s)
{
    const uint32_t f = b.flags;
    const bool sph = f & GRASPH_SPH, grav = f & GRASPH_SELF_GRAVITY;
    const bool dens_pass = sph && !(f & GRASPH_INTEGRATE_DENSITY);
    const bool strain_pass = sph && ((f & GRASPH_SOLIDS) || (f & GRASPH_INTEGRATE_DENSITY) || (f & GRASPH_BALSARA_SWITCH));
    const DevParams P = make_dev_params(&b.params[0]);  // flags; values come from the per-instance table
    int r;
    if ((r = batch_gather(c, b, p, s))) return r;
    // the one-sided persistent sweep takes the items (with the symmetric
    // sweep on: the smaller instances), the symmetric sweep the rest; its
    // far field is forked onto a stream of its own and overlaps the rest
    const bool sym = (sph || grav) && c.batch_symmetric;
    SymBatchLaunch SL;
    if (sym && (r = sym_batch_begin(c, b, p, s, false, true, SL))) return r;
    if ((sph || grav) && b.n_items > 0 && (r = launch_sweep_only(c, b, p, s, sym))) return r;

    const int* perm = slot_order(c, b);
    NbrLaunch N;
    std::memset(&N, 0, sizeof N);
    N.posm = (const float4*)p->posm;
    N.vel = (const float4*)p->vel;
    N.xs = perm ? b.d_xs : (const float4*)p->posm;
    N.rec = b.d_rec;
    N.dstress = p->dstress;
    N.balsara_in = p->balsara;
    N.perm = perm;
    N.nbr = b.d_nbr;
    N.nbr_cnt = b.d_cnt;
    N.n = (int)b.ntotal;
    N.kmax = b.kmax;
    N.balsara = p->balsara;
    N.density_dt = p->density_dt;
    N.dstress_dt = p->dstress_dt;
    N.acc = (float4*)p->acc;
    N.xvel = (float4*)p->xvel;
    N.maxvsig = p->maxvsig;
    N.acc_has_gravity = grav;
    N.T = tables(b);
    hipError_t e;
    if (!sph) {
        if (sym && (r = sym_batch_end(c, b, SL, s, true, false))) return r;
        ProfScope ps(&c, s, KC_OTHER);
        e = launch_finish(N, P, s);
        if (e != hipSuccess) return hip_fail(c, e, "batched finish launch");
        return 0;
    }
    if (dens_pass) {
        ProfScope ps(&c, s, KC_DENSITY);
        e = launch_density(N, p->density, P, s);
        if (e != hipSuccess) return hip_fail(c, e, "batched density launch");
    }
    if (strain_pass) {
        ProfScope ps(&c, s, KC_STRAIN);
        e = launch_strain(N, P, s);
        if (e != hipSuccess) return hip_fail(c, e, "batched strain launch");
    } else {
        ProfScope ps(&c, s, KC_OTHER);
        e = launch_prep(N, P, s);
        if (e != hipSuccess) return hip_fail(c, e, "batched record launch");
    }
    // With every instance in the symmetric sweep, the acceleration pass runs
    // while the far field is still going and writes the SPH terms alone; the
    // combine then adds the gravity to them.  Otherwise it adds the SPH
    // terms to the gravity of the sweeps.
    const bool accel_first = sym && b.n_items == 0;
    if (accel_first) {
        N.acc_has_gravity = false;
        ProfScope ps(&c, s, KC_ACCEL);
        e = launch_accel(N, P, s);
        if (e != hipSuccess) return hip_fail(c, e, "batched acceleration launch");
    }
    if (sym && (r = sym_batch_end(c, b, SL, s, true, accel_first && grav))) return r;
    if (accel_first) return 0;
    ProfScope ps(&c, s, KC_ACCEL);
    e = launch_accel(N, P, s);
    if (e != hipSuccess) return hip_fail(c, e, "batched acceleration launch");
    return 0;
}

int batch_integrate_launch(Ctx& c, BatchState& b, const grasph_particles* p, int nfs, hipStream_t s)
{
    IntegLaunch L;
    std::memset(&L, 0, sizeof L);
    L.posm = (float4*)p->posm;
    L.pos64 = p->pos64;
    L.vel = (float4*)p->vel;
    L.acc = (const float4*)p->acc;
    L.xvel = (const float4*)p->xvel;
    L.density = p->density;
    L.density_dt = p->density_dt;
    L.dstress = p->dstress;
    L.dstress_dt = p->dstress_dt;
    L.n = (int)b.ntotal;
    L.not_first_step = nfs;
    L.T = tables(b);
    L.maxvsig = p->maxvsig;
    L.n_instances = b.n_inst;
    L.ts = (b.flags & GRASPH_VARIABLE_TIMESTEP) ? b.d_ts : nullptr;
    ProfScope ps(&c, s, KC_INTEGRATE);
    hipError_t e = launch_integrate_step(L, make_dev_params(&b.params[0]), s);
    if (e != hipSuccess) return hip_fail(c, e, "batched leapfrog launch");
    return 0;
}

grasph_particles sub_view(const grasph_particles* p, int64_t off, int64_t n)
{
    grasph_particles q = *p;
    q.n = n;
    auto o = [&](float* a, int64_t per) { return a ? a + per * off : nullptr; };
    q.posm = o(p->posm, 4);
    q.pos64 = p->pos64 ? p->pos64 + 3 * off : nullptr;
    q.vel = o(p->vel, 4);
    q.acc = o(p->acc, 4);
    q.xvel = o(p->xvel, 4);
    q.density = o(p->density, 1);
    q.density_dt = o(p->density_dt, 1);
    q.balsara = o(p->balsara, 1);
    q.maxvsig = o(p->maxvsig, 1);
    q.dstress = o(p->dstress, 6);
    q.dstress_dt = o(p->dstress_dt, 6);
    return q;
}

std::string graph_key(const Ctx& c, const BatchState& b, const grasph_particles* p)
{
    std::ostringstream k;
    k << p->posm << ' ' << p->pos64 << ' ' << p->vel << ' ' << p->acc << ' ' << p->xvel << ' ' << p->density << ' '
      << p->density_dt << ' ' << p->balsara << ' ' << p->maxvsig << ' ' << p->dstress << ' ' << p->dstress_dt << ' '
      << (c.reorder_interval > 0) << ' ' << c.batch_wpb << ' ' << c.batch_bpc << ' ' << (const void*)b.d_nbr << ' '
      << (const void*)b.d_rec << ' ' << (const void*)b.d_items << ' ' << (const void*)b.d_xs << ' '
      << c.batch_symmetric << ' ' << c.batch_symmetric_min << ' '
      << (const void*)b.d_sinst << ' ' << (const void*)b.d_sbpart << ' ' << (const void*)b.d_sblk;
    return k.str();
}

// time the batched sweep (min of reps), milliseconds: the one-sided
// persistent sweep alone, or (all) with the symmetric sweep of the larger instances
int time_sweep(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s, int reps, double* ms, bool all = false)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return hip_fail(c, hipGetLastError(), "autotune events");
    double best = 1e30;
    int r = 0;
    const bool prof = c.profile;
    c.profile = false;
    for (int k = 0; k < reps + 1 && !r; ++k) {
        (void)hipEventRecord(e0, s);
        if (!all || b.n_items > 0) r = launch_sweep_only(c, b, p, s);
        if (!r && all && c.batch_symmetric) r = launch_sym_sweep_batch(c, b, p, s, b.n_items > 0);
        (void)hipEventRecord(e1, s);
        if (!r && hipEventSynchronize(e1) == hipSuccess) {
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (k > 0) best = std::min(best, (double)t);  // the first run warms up
        }
    }
    c.profile = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return r;
}

// time one whole derivative evaluation of the batch (eager, min of reps
// after a warm-up), milliseconds: the sweeps, the neighbour passes beside
// the far field and the combine, as a step runs them
int time_derivs(Ctx& c, BatchState& b, const grasph_particles* p, hipStream_t s, int reps, double* ms)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return hip_fail(c, hipGetLastError(), "autotune events");
    double best = 1e30;
    int r = 0;
    const bool prof = c.profile;
    c.profile = false;
    for (int k = 0; k < reps + 1 && !r; ++k) {
        (void)hipEventRecord(e0, s);
        r = batch_derivative_launches(c, b, p, s);
        (void)hipEventRecord(e1, s);
        if (!r && hipEventSynchronize(e1) == hipSuccess) {
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            if (k > 0) best = std::min(best, (double)t);  // the first run warms up
        }
    }
    c.profile = prof;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return r;
}

std::string tune_signature(Ctx& c, const BatchState& b)
{
    hipDeviceProp_t prop;
    std::string dev = "unknown";
    if (hipGetDeviceProperties(&prop, c.device) == hipSuccess) dev = prop.gcnArchName;
    int cnt[4] = {0, 0, 0, 0};
    for (int k = 0; k < b.n_inst; ++k) cnt[bucket_of(b.offsets[k + 1] - b.offsets[k])]++;
    std::ostringstream s;
    s << dev << ":cus=" << device_cus(c) << ":buckets=" << cnt[0] << "," << cnt[1] << "," << cnt[2] << "," << cnt[3]
      << ":flags=" << b.flags << ":sym=" << c.batch_symmetric;
    return s.str();
}

}  // namespace

void destroy_batch_state(BatchState* b)
{
    if (!b) return;
    drop_graph(*b);
    free_tables(*b);
    free_ptr(b->d_nbr);
    free_ptr(b->d_cnt);
    free_ptr(b->d_rec);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->gs) (void)hipStreamDestroy(b->gs);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->ev_lists) (void)hipEventDestroy(b->ev_lists);
    if (b->side) (void)hipStreamDestroy(b->side);
    for (grasph_ctx* x : b->loop)
        if (x) grasph_destroy(x);
    delete b;
}

}  // namespace grasph

using grasph::BatchState;
using grasph::Ctx;

extern "C" {

int grasph_batch_compute_derivatives(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if ((r = grasph::prepare(c, p, b))) return r;
    if (p->n == 0) return 0;
    if ((r = grasph::check_derivative_arrays(c, p, c.batch->flags, 0))) return r;
    hipStream_t s = (hipStream_t)stream;
    if ((r = grasph::ensure_order(c, *c.batch, p, s, true))) return r;
    return grasph::batch_derivative_launches(c, *c.batch, p, s);
}

int grasph_batch_integrate(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, int not_first_step,
                           void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if ((r = grasph::prepare(c, p, b))) return r;
    if (p->n == 0) return 0;
    if ((r = grasph::check_integrate_arrays(c, p, c.batch->flags))) return r;
    return grasph::batch_integrate_launch(c, *c.batch, p, not_first_step, (hipStream_t)stream);
}

int grasph_batch_step(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, int n_steps, void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if (n_steps < 0) return grasph::fail(c, GRASPH_E_INVALID, "n_steps must be >= 0");
    if ((r = grasph::prepare(c, p, b))) return r;
    if (p->n == 0 || n_steps == 0) return 0;
    BatchState& B = *c.batch;
    if ((r = grasph::check_derivative_arrays(c, p, B.flags, 0)) || (r = grasph::check_integrate_arrays(c, p, B.flags)))
        return r;
    hipStream_t us = (hipStream_t)stream;
    hipError_t e = hipEventRecord(B.ev_in, us);
    if (e == hipSuccess) e = hipStreamWaitEvent(B.gs, B.ev_in, 0);
    if (e != hipSuccess) return grasph::hip_fail(c, e, "batch stream join");
    for (int step = 0; step < n_steps; ++step) {
        if ((r = grasph::ensure_order(c, B, p, B.gs, true))) return r;
        const std::string key = grasph::graph_key(c, B, p);
        if (!B.exec || key != B.graph_key) {
            grasph::drop_graph(B);
            c.capturing = true;
            e = hipStreamBeginCapture(B.gs, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                c.capturing = false;
                return grasph::hip_fail(c, e, "graph capture begin");
            }
            r = grasph::batch_derivative_launches(c, B, p, B.gs);
            if (!r) r = grasph::batch_integrate_launch(c, B, p, 1, B.gs);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(B.gs, &g);
            c.capturing = false;
            if (r) {
                if (g) (void)hipGraphDestroy(g);
                return r;
            }
            if (e != hipSuccess) return grasph::hip_fail(c, e, "graph capture end");
            B.graph = g;
            e = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                B.exec = nullptr;
                return grasph::hip_fail(c, e, "graph instantiate");
            }
            B.graph_key = key;
        }
        grasph::ProfScope ps(&c, B.gs, grasph::KC_OTHER);
        e = hipGraphLaunch(B.exec, B.gs);
        if (e != hipSuccess) return grasph::hip_fail(c, e, "graph launch");
    }
    e = hipEventRecord(B.ev_out, B.gs);
    if (e == hipSuccess) e = hipStreamWaitEvent(us, B.ev_out, 0);
    if (e != hipSuccess) return grasph::hip_fail(c, e, "batch stream join");
    return 0;
}

int grasph_batch_step_loop(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, int n_steps,
                           void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if (n_steps < 0) return grasph::fail(c, GRASPH_E_INVALID, "n_steps must be >= 0");
    if ((r = grasph::validate(c, p, b))) return r;
    if (!c.batch) c.batch = new BatchState;
    BatchState& B = *c.batch;
    const int n_inst = b->n_instances;
    if ((int)B.loop.size() < n_inst) B.loop.resize(n_inst, nullptr);
    for (int k = 0; k < n_inst; ++k) {
        if (!B.loop[k]) {
            if (grasph_create(c.device, &B.loop[k]) != 0) return grasph::fail(c, GRASPH_E_HIP, grasph_last_error(nullptr));
        }
        Ctx& s = B.loop[k]->c;
        s.max_neighbours = s.max_neighbours_set = c.max_neighbours_set;
        if (s.reorder_interval != c.reorder_interval) {
            s.reorder_interval = c.reorder_interval;
            s.perm_n = 0;
        }
        s.sweep_ni = c.sweep_ni;
        s.sweep_slices = c.sweep_slices;
        s.profile = false;
    }
    for (int step = 0; step < n_steps; ++step)
        for (int k = 0; k < n_inst; ++k) {
            const int64_t off = b->offsets[k], n = b->offsets[k + 1] - off;
            if (n == 0) continue;
            const grasph_particles q = grasph::sub_view(p, off, n);
            if (grasph_compute_derivatives(B.loop[k], &q, &b->params[k], stream) != 0 ||
                grasph_integrate(B.loop[k], &q, &b->params[k], 1, stream) != 0)
                return grasph::fail(c, GRASPH_E_HIP, std::string("instance ") + std::to_string(k) + ": " +
                                                         grasph_last_error(B.loop[k]));
        }
    return 0;
}

int grasph_autotune(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, const char* cache_path,
                    void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if ((r = grasph::prepare(c, p, b))) return r;
    if (p->n == 0) return 0;
    if ((r = grasph::check_derivative_arrays(c, p, c.batch->flags, 0))) return r;
    BatchState& B = *c.batch;
    hipStream_t s = (hipStream_t)stream;
    const std::string sig = grasph::tune_signature(c, B);
    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string key;
            int wpb, bpc, n0, n1, n2, n3;
            int64_t smin;
            if ((ls >> key >> wpb >> bpc >> n0 >> n1 >> n2 >> n3 >> smin) && key == sig && wpb >= 1 && wpb <= 4 &&
                bpc >= 1 && bpc <= 64 && n0 >= 1 && n0 <= 2 && n1 >= 1 && n1 <= 2 && n2 >= 1 && n2 <= 2 && n3 >= 1 &&
                n3 <= 2 && smin >= -1 && smin <= (int64_t(1) << 30) && (smin < 0 || c.batch_symmetric)) {
                c.batch_wpb = wpb;
                c.batch_bpc = bpc;
                c.batch_ni[0] = n0; c.batch_ni[1] = n1; c.batch_ni[2] = n2; c.batch_ni[3] = n3;
                c.batch_symmetric = smin >= 0;
                c.batch_symmetric_min = (int)std::max<int64_t>(smin, 0);
                return grasph::prepare(c, p, b);
            }
        }
    }
    if ((r = grasph::ensure_order(c, B, p, s, false))) return r;
    // the one-sided persistent sweep over every instance for its launch shape
    const bool sym_allowed = c.batch_symmetric && (B.flags & (GRASPH_SPH | GRASPH_SELF_GRAVITY));
    c.batch_symmetric = 0;
    // launch shape with one particle per lane everywhere
    for (int q = 0; q < 4; ++q) c.batch_ni[q] = 1;
    if ((r = grasph::prepare(c, p, b))) return r;
    if ((r = grasph::batch_gather(c, B, p, s))) return r;  // the timed sweeps read the slot-ordered copies
    const int shapes[][2] = {{1, 8}, {1, 16}, {2, 4}, {2, 8}, {4, 2}, {4, 4}, {4, 8}};
    double best = 1e30;
    int bw = 4, bb = 8;
    for (const auto& sh : shapes) {
        c.batch_wpb = sh[0];
        c.batch_bpc = sh[1];
        double t;
        if ((r = grasph::time_sweep(c, B, p, s, 3, &t))) return r;
        if (t < best) {
            best = t;
            bw = sh[0];
            bb = sh[1];
        }
    }
    c.batch_wpb = bw;
    c.batch_bpc = bb;
    // particles per lane, bucket by bucket
    int present[4] = {0, 0, 0, 0};
    for (int k = 0; k < B.n_inst; ++k) present[grasph::bucket_of(B.offsets[k + 1] - B.offsets[k])] = 1;
    for (int q = 0; q < 4; ++q) {
        if (!present[q]) continue;
        c.batch_ni[q] = 2;
        if ((r = grasph::prepare(c, p, b))) return r;
        double t;
        if ((r = grasph::time_sweep(c, B, p, s, 3, &t))) return r;
        if (t < 0.99 * best)
            best = t;
        else
            c.batch_ni[q] = 1;
    }
    // the symmetric sweep from which size bucket on (or for none), timed
    // as whole derivative evaluations: the sweeps alone rank the candidates
    // near 0, 4096 and 8192 within their noise, while the passes beside
    // the far field set the step apart (16.5 against 17.1 ms per step)
    int64_t sym_min = -1;
    if (sym_allowed) {
        const int64_t cands[] = {0, 2048, 4096, 8192, -1};
        double best_all = 1e30;
        int last = -2;
        for (const int64_t cand : cands) {
            int count = 0;  // instances the symmetric sweep takes
            for (int k = 0; k < B.n_inst; ++k) count += cand >= 0 && B.offsets[k + 1] - B.offsets[k] >= cand;
            if (count == last) continue;  // the same split as the candidate before
            last = count;
            c.batch_symmetric = cand >= 0;
            c.batch_symmetric_min = (int)std::max<int64_t>(cand, 0);
            if ((r = grasph::prepare(c, p, b))) return r;
            double t;
            if ((r = grasph::time_derivs(c, B, p, s, 3, &t))) return r;
            if (t < 0.99 * best_all) {
                best_all = t;
                sym_min = cand;
            }
        }
    }
    c.batch_symmetric = sym_min >= 0;
    c.batch_symmetric_min = (int)std::max<int64_t>(sym_min, 0);
    if ((r = grasph::prepare(c, p, b))) return r;
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        out << sig << ' ' << c.batch_wpb << ' ' << c.batch_bpc << ' ' << c.batch_ni[0] << ' ' << c.batch_ni[1] << ' '
            << c.batch_ni[2] << ' ' << c.batch_ni[3] << ' ' << sym_min << '\n';
    }
    return 0;
}

int grasph_get_batch_tuning(const grasph_ctx* ctx, int32_t ni[4], int32_t* waves_per_block, int32_t* blocks_per_cu)
{
    if (!ctx) return GRASPH_E_INVALID;
    const Ctx& c = ctx->c;
    if (ni)
        for (int q = 0; q < 4; ++q) ni[q] = c.batch_ni[q];
    if (waves_per_block) *waves_per_block = c.batch_wpb;
    if (blocks_per_cu) *blocks_per_cu = c.batch_bpc;
    return 0;
}

int grasph_batch_get_current_timesteps(grasph_ctx* ctx, const grasph_batch* b, double* dt)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    if (!b || !b->offsets || !b->params || !dt || b->n_instances < 1)
        return grasph::fail(c, GRASPH_E_INVALID, "batch, its arrays or dt is NULL");
    const int n = b->n_instances;
    if (!(b->params[0].flags & GRASPH_VARIABLE_TIMESTEP)) {
        for (int k = 0; k < n; ++k) dt[k] = (double)b->params[k].fixed_timestep;
        return 0;
    }
    const grasph::BatchState* B = c.batch;
    const bool same = B && B->d_ts && B->n_inst == n && std::equal(B->offsets.begin(), B->offsets.end(), b->offsets) &&
                      std::memcmp(B->params.data(), b->params, n * sizeof(grasph_params)) == 0;
    if (!same) return grasph::fail(c, GRASPH_E_INVALID, "this batch has not been integrated on this context");
    std::vector<float> ts(2 * (size_t)n);
    if (hipSetDevice(c.device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(ts.data(), B->d_ts, ts.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return grasph::hip_fail(c, hipGetLastError(), "reading the batch time steps");
    for (int k = 0; k < n; ++k) dt[k] = (double)ts[2 * (size_t)k];
    return 0;
}

}  // extern "C"
