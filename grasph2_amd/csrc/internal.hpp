// internal.hpp -- library-internal declarations shared by the .hip units.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <vector>

#include "physics.hpp"

namespace grasph {

// Neighbour lists come in tiles of 64 slots: entry k of slot t lies at
// list_entry(t, k, kmax) of its plane, so a wavefront's k-th entries are one
// 256-byte line and a wavefront's whole list is contiguous (a plane laid out
// as [k][slot] puts each entry of a long list in another part of a
// gigabyte-sized array, and the page-table walks made merging and walking
// long lists several times slower).  A plane holds list_plane(n, kmax)
// entries.
__host__ __device__ inline size_t list_entry(int t, int k, int kmax)
{
    return ((size_t)(t >> 6) * (size_t)kmax + (size_t)k) * 64 + (size_t)(t & 63);
}
__host__ __device__ inline size_t list_plane(int64_t n, int kmax) { return (size_t)((n + 63) / 64) * 64 * (size_t)kmax; }

enum KernelClass {
    KC_SWEEP = 0,
    KC_DENSITY = 1,
    KC_STRAIN = 2,
    KC_ACCEL = 3,
    KC_INTEGRATE = 4,
    KC_REORDER = 5,
    KC_OTHER = 6,
    KC_FAR = 7  // far field of the symmetric sweep (the rest of the evaluation runs beside it)
};

// Per-device library state.  Scratch grows on demand and is reused.
struct Ctx {
    int device = 0;
    std::string err;

    // options
    int max_neighbours = 160;      // list capacity of this context's single-instance evaluations (grows)
    int max_neighbours_set = 160;  // the capacity last set: batches and their launch-loop contexts start from it
    // grow max_neighbours after an evaluation whose lists overflowed: its
    // largest count is reduced on the device and copied to pinned host
    // memory; a later call reads it once the copy has landed (no wait)
    int max_neighbours_grow = 1;
    int* list_max_dev = nullptr;
    int* list_max_host = nullptr;
    hipEvent_t list_max_ev = nullptr;
    bool list_max_pending = false;
    int reorder_interval = 8;  // particles move little in 8 steps: the order only affects speed
    int sweep_ni = 1;  // i particles per lane in the single-instance sweep
    int sweep_slices = 0;  // j-split of the single-instance sweep (0: chosen from the occupancy)
    int sweep_symmetric = 2;  // single instance, far pairs once for both sides (sym_sweep.hip): 0 off, 1 on, 2 from 16384 particles
    int sym_workgroups = 0;   // workgroups of its far kernel (0: 2 wavefronts per SIMD)
    int sym_near_shares = 0;  // wavefronts per block of its near field (0: chosen from the occupancy)
    // bytes of partner-side partials one column batch of its far field may take (6 n^2 / 1024 in all)
    int64_t sym_partial_bytes = int64_t(12) << 30;
    int sym_list_priority = 2;  // its near-field lists ahead of the far field: 0 no, 1 yes, 2 below 49152 particles or past 512 list entries
    int batch_symmetric = 1;  // batches: the symmetric sweep for instances of batch_symmetric_min particles or more
    int batch_symmetric_min = 0;
    int cus = 256;         // compute units of the device
    bool profile = false;
    // batched sweep tuning: i per lane per size bucket, persistent launch shape
    int batch_ni[4] = {1, 1, 1, 1};
    int batch_wpb = 4;
    int batch_bpc = 8;

    // variable-step leapfrog state {current dt, next dt}, zero at creation
    float* ts = nullptr;
    // {uniform, mass} of the particles, written before each gravity sweep,
    // then the per-block flags of the check (2 + 256 floats)
    float* umass = nullptr;
    int* mass_flags = nullptr;  // per-256-slot flags of the check when the gather makes them
    int64_t mass_flags_cap = 0;

    // Morton order of the i side (slot t -> particle perm[t])
    int* perm = nullptr;
    int64_t perm_n = 0;
    const void* perm_src = nullptr;  // posm pointer the order was computed for
    int64_t evals_since_reorder = 0;
    uint32_t* keys = nullptr;
    uint32_t* keys_alt = nullptr;
    int* vals = nullptr;
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    float* bbox = nullptr;  // [nblocks][6] partial bounds then final [6]
    int64_t reorder_cap = 0;

    // neighbour lists: one plane per slice (list_plane entries; entry k of
    // slot t at list_entry(t, k, kmax)), counts nbr_cnt[plane * n + slot]
    int* nbr = nullptr;
    int* nbr_cnt = nullptr;
    int64_t nbr_cap_n = 0;
    int nbr_cap_k = 0;
    int64_t last_list_n = 0;  // particles of the last neighbour list built
    int last_list_slices = 1;
    float4* part = nullptr;   // partial sums of a j-split sweep
    int64_t part_cap = 0;
    // work arrays of the symmetric sweep
    float4* sym_box = nullptr;
    int64_t sym_box_cap = 0;
    uint32_t* sym_bits = nullptr;
    int64_t sym_bits_cap = 0;
    float4* sym_near = nullptr;
    int64_t sym_near_cap = 0;
    float4* sym_apart = nullptr;
    int64_t sym_apart_cap = 0;
    float* sym_bpart = nullptr;
    int64_t sym_bpart_cap = 0;
    float4* sym_accsum = nullptr;  // running sums between the far field's column batches
    int64_t sym_accsum_cap = 0;

    // slot-ordered positions for the sweep, and the slot records of the passes
    float4* xs = nullptr;
    int64_t xs_cap = 0;
    struct SlotRec* rec = nullptr;
    int64_t rec_cap = 0;

    // potential and reduction partials
    float* pot = nullptr;
    int64_t pot_cap = 0;
    double* red = nullptr;
    int64_t red_cap = 0;

    // timing
    struct Ev {
        hipEvent_t a, b;
        int cls;
    };
    std::vector<Ev> ev_pending;
    std::vector<hipEvent_t> ev_free;
    double ms[GRASPH_NKERNEL_CLASSES] = {};
    int64_t launches[GRASPH_NKERNEL_CLASSES] = {};
    bool capturing = false;  // no timing events while a graph is being captured

    // beside the symmetric sweep's far field (on the caller's stream), the
    // rest of the evaluation runs on a second stream (default priority,
    // non-blocking), forked from and joined to the caller's
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    struct BatchState* batch = nullptr;

    // The allocation-free entry point runs the derivative path on a
    // stand-in context whose buffers are cut from the device's static
    // scratch array; there no_alloc turns every allocation the path would
    // make into an error.  scratch: that array's device address on this
    // context's device (null, with the reason in scratch_err, when the
    // runtime could not resolve it).
    bool no_alloc = false;
    void* scratch = nullptr;
    std::string scratch_err;
    int noalloc_slices = 0;  // j-slices of the last allocation-free call
};

// Bracket a kernel launch with events when profiling is on.
struct ProfScope {
    Ctx* c;
    hipStream_t s;
    int cls;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(Ctx* ctx, hipStream_t st, int k);
    ~ProfScope();
};

// ---- launchers (return hipError_t of the launch) -------------------------
struct SweepLaunch {
    const float4* posm;  // particles (index order)
    const float4* xs;    // the same in slot order (xs[t] = posm[perm[t]]), or null when perm is null
    const int* perm;
    int n;
    float4* acc;        // gravity (written when grav)
    int* nbr;           // written when want_nbr
    int* nbr_cnt;
    float* pot;         // written when want_pot
    int kmax;
    bool grav, want_nbr, want_pot;
    bool finish;        // apply environmentAcceleration, zero xvel / maxvsig (SPH off)
    const float4* vel;
    float4* xvel;
    float* maxvsig;
    int ni;             // i particles per lane (1 or 2)
    int nslices;        // j-split: slices of slice_len j (lists and counts per slice); the
                        // symmetric sweep: its near field's list shares (planes)
    int near_shares;    // the symmetric sweep: shares of its in-group gravity
    int slice_len;
    float4* part;       // [nslices][n][2] partial sums when nslices > 1
    const float* umass; // {uniform, mass} from launch_mass_check, or null
};
hipError_t launch_sweep(const SweepLaunch& L, const DevParams& P, hipStream_t s);

// The symmetric single-instance sweep (sym_sweep.hip): super-blocks of 256
// slots, groups of 4 super-blocks; gravity inside a group through the
// candidate sweep, between groups once per unordered pair; neighbour lists
// from the blocks of 64 whose boxes come within H.  Sizes of its work
// arrays, in elements.
struct SymGeom {
    int nb, groups, wgs, nwords;
    int64_t nsb, units;   // super-blocks; far work units
    int split;            // pieces per partner super-block in a far unit (1, 2 or 4)
    int64_t box_elems;    // float4: bounding box per block of 64
    int64_t bits_elems;   // uint32: near-block bit matrix (lists wanted)
    int64_t near_elems;   // float4: near-field partial per slot and share
    int64_t apart_elems;  // float4: far own-side fragments (kFragElems each)
    bool plain;           // no far workgroup flushes a fragment twice: no correction planes
    int64_t bpart_elems;  // float: far partner-side partials (of the widest column batch)
    int ncols;            // column batches of the far field (1: the whole far field at once)
};
SymGeom sym_geometry(int64_t n, int wgs, bool want_nbr, int nshares, int64_t col_elems = INT64_MAX);
// one column batch [c0, c1) of the far field's partner groups
struct SymCol {
    int c0, c1;
    int64_t units;
    int wgs;
    bool plain;
};
SymCol sym_column(int64_t nsb, int split, int wanted, int c0, int c1);
int sym_default_workgroups(int cus);  // far-field workgroups for 2 wavefronts per SIMD
struct SymLaunch {
    const float4* xs;   // slot-ordered positions + masses
    const int* perm;    // slot -> particle, or null
    int n;
    bool grav, want_nbr;
    bool finish;        // as SweepLaunch::finish
    bool add_acc;       // acc already holds the SPH terms: the combine adds the gravity to them
    bool list_prio;     // the lists' wavefronts take issue priority over the far field's beside them
    int list_shares;    // its list kernel cut into this many shares (own list planes, merged after)
    int nshares;        // near field cut into this many shares per block (own list planes, merged after)
    float4* acc;
    int* nbr;           // a list per slot and share
    int* nbr_cnt;
    int kmax;
    const float4* vel;
    float4* xvel;
    float* maxvsig;
    const float* umass; // {uniform, mass} (grav)
    int wgs;            // workgroups of the far kernel
    float4* box;
    uint32_t* bits;
    float4* near_part;
    float4* apart;        // own-side fragments (kFragElems each)
    float* bpart;
    int64_t col_elems;    // partner-side partials (floats) a column batch of the far field may take
    float4* accsum;       // running sums between column batches: 2 float4 per slot (column-batched far field)
};
// the symmetric sweep in four launches: the near field's lists and its
// in-group gravity, the far field (no dependence among them: they may run
// on two streams), then the combine
hipError_t launch_sym_lists(const SymLaunch& L, const DevParams& P, hipStream_t s);
hipError_t launch_sym_ingroup(const SymLaunch& L, const DevParams& P, hipStream_t s);
hipError_t launch_sym_far(const SymLaunch& L, const DevParams& P, hipStream_t s);
hipError_t launch_sym_combine(const SymLaunch& L, const DevParams& P, hipStream_t s);
bool sym_has_far(const SymLaunch& L);

// The symmetric sweep over a batch (slot indices are the batch's).
struct BatchInst;
struct SymInst {
    int off, n;         // slots [off, off + n)
    int nb, nwords;     // blocks of 64; words per near-bit row
    int nsb, nrows;     // super-blocks; far rows (groups with later super-blocks)
    int box0, row0;     // its first block box and first far row in the batch's arrays
    int64_t bits0;      // first word of its near-bit matrix
    int64_t bpart0;     // first float of its partner-side partials
};
struct SymRow {
    int inst, g, piece, units;  // piece `piece` of far row g of instance inst: units partner groups
};
struct SymBatchTables {
    std::vector<SymInst> insts;  // every instance; those below min_n have n = 0
    std::vector<SymRow> rows;    // longest first
    std::vector<int4> blk;       // 256-slot blocks (instance, first, end) of the instances in the sweep
    int max_words = 0;
    int64_t box_elems = 0, bits_elems = 0, apart_elems = 0, bpart_elems = 0;
    bool plain = true;  // no row piece flushes its fragment twice
};
// instances of at least min_n particles take the symmetric sweep
SymBatchTables sym_batch_tables(const std::vector<int64_t>& offsets, bool want_nbr, int64_t min_n);
struct SymBatchLaunch {
    const float4* xs;     // slot-ordered positions + masses of the batch
    const int* perm;      // slot -> particle, or null
    int ntotal, n_inst;
    bool grav, want_nbr, finish, soft;
    bool add_acc;         // acc already holds the SPH terms: the combine adds the gravity to them
    const SymInst* sinst;
    const BatchInst* insts;
    const DevParams* prm;
    const int4* blk;      // 256-slot blocks of the instances in the sweep
    int n_blk;
    const int* inst_of;
    const SymRow* rows;
    int nrows;
    int max_words;
    int* queue;           // one work counter
    int wgs;
    bool plain;           // SymBatchTables::plain
    bool ingroup_done;    // the in-group gravity was enqueued after the far field on its stream
    const float* umass;   // {uniform, mass} per instance
    float4* acc;
    int* nbr;
    int* nbr_cnt;
    int kmax;
    const float4* vel;
    float4* xvel;
    float* maxvsig;
    float4* box;
    uint32_t* bits;
    float4* near_part;
    float4* apart;        // own-side fragments (kFragElems each)
    float* bpart;
};
// the batched symmetric sweep in four launches, as the single-instance one
hipError_t launch_sym_batch_lists(const SymBatchLaunch& L, hipStream_t s);
hipError_t launch_sym_batch_ingroup(const SymBatchLaunch& L, hipStream_t s);
hipError_t launch_sym_batch_far(const SymBatchLaunch& L, hipStream_t s);
hipError_t launch_sym_batch_combine(const SymBatchLaunch& L, hipStream_t s);
inline bool sym_batch_has_far(const SymBatchLaunch& L) { return L.ntotal > 0 && L.n_blk > 0 && L.grav && L.nrows > 0; }

// Fold the per-slice lists of a j-split sweep into one list per particle
// (slice 0's arrays; a count above kmax means "scan the instance").
hipError_t launch_merge_lists(int* nbr, int* cnt, int n, int kmax, int nslices, hipStream_t s);
// out[0] = the largest of cnt[0, n) (one workgroup)
hipError_t launch_count_max(const int* cnt, int n, int* out, hipStream_t s);

// Per-slot record of the neighbour passes (64 bytes, slot order): position
// and mass; velocity and density; the acceleration pass's j-side terms
// sigma/rho^2 (xx yy zz xy | xz yz), balsara, density.
struct SlotRec {
    float4 pm, vr, ra, rb;
};

// One instance of a batch, as the device kernels see it.
struct BatchInst {
    int off;     // first particle
    int n;       // particle count
    float H2;    // smoothing length squared
    float cand;  // neighbour-candidate threshold on r^2 + H^2
};

// Block descriptors of the per-particle passes over a batch: block b covers
// slots [blk[b].y, blk[b].z) of instance blk[b].x (never two instances).
struct BlockTables {
    const int4* blk;
    const BatchInst* insts;
    const DevParams* prm;  // per-instance parameters
    int nblocks;           // 0: one instance, the grid covers all slots
};

// Workgroups are dealt to the GPU's 8 XCDs in turn (workgroup b to XCD
// b mod 8), and each XCD has an L2 of its own.  xcd_block gives XCD x the
// x-th contiguous eighth of the nb blocks: consecutive slots are spatial
// neighbours, so a block's neighbour gathers then hit the L2 that its
// neighbouring blocks filled (dealt round-robin, every record of the
// ensemble's strain pass was fetched into several L2s: 11 GB per pass for
// 10 GB of gathers).  A bijection of [0, nb): the work and the results do
// not change, only where they run.
__device__ inline int xcd_block(int b, int nb)
{
    constexpr int kXcds = 8;
    const int x = b % kXcds, i = b / kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + min(x, r) + i;
}

// What one block of a per-particle pass works on.
struct SlotMap {
    int slot0, slot_end, off, n, inst;
    // element c of particle i in an array of 6 (resp. 3) planes per instance
    __device__ size_t plane6(int c, int i) const { return 6 * (size_t)off + (size_t)c * n + (size_t)(i - off); }
    __device__ size_t plane3(int c, int i) const { return 3 * (size_t)off + (size_t)c * n + (size_t)(i - off); }
};

template <bool BATCH>
__device__ inline SlotMap slot_map(const BlockTables& T, int n, int slots_per_block = 256)
{
    SlotMap m;
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    if (BATCH) {
        const int4 b = T.blk[blk];
        const BatchInst I = T.insts[b.x];
        m.inst = b.x;
        m.slot0 = b.y;
        m.slot_end = b.z;
        m.off = I.off;
        m.n = I.n;
    } else {
        m.inst = 0;
        m.slot0 = blk * slots_per_block;
        m.slot_end = n;
        m.off = 0;
        m.n = n;
    }
    return m;
}

struct BatchSweepLaunch {
    const float4* xs;   // particles in slot order (posm itself when perm is null)
    const int* perm;
    const int4* items;  // (instance, first slot, i per lane, 0) per work item, longest instance first
    int n_items;
    const BatchInst* insts;
    int* queue;         // work counter, zero before the launch
    float4* acc;
    int* nbr;
    int* nbr_cnt;
    int ntotal;
    int kmax;
    int blocks;           // persistent grid
    int waves_per_block;  // 1, 2 or 4
    bool grav, want_nbr, soft;
    const float* umass;   // {uniform, mass} per instance, or null
    int ni_max;           // largest ni of the items (1 or 2)
};
hipError_t launch_batch_sweep(const BatchSweepLaunch& L, const DevParams& P, hipStream_t s);
// umass[2k] = 1 when every particle of instance k (insts null: the n
// particles of one instance) has the same mass, umass[2k + 1] = that mass.
// With insts null, umass must hold 2 + 256 floats (per-block flags follow).
hipError_t launch_mass_check(const float4* posm, const BatchInst* insts, int n_inst, int n, float* umass,
                             hipStream_t s);
hipError_t launch_reset_queue(int* queue, int n, hipStream_t s);

// Batched Morton order: per-instance bounding boxes, keys (instance, Z-order
// cell), one radix sort; perm maps each instance's slots into its own range.
struct BatchOrder {
    const float4* posm;
    const int* inst_of;      // instance of every particle
    const BatchInst* insts;
    int n_inst;
    int ntotal;
    float* bbox;             // [n_inst][6]
    uint64_t* keys;
    uint64_t* keys_alt;
    int* vals;
    int* perm;
    void* sort_tmp;
    size_t sort_tmp_bytes;
    int key_bits;
};
size_t batch_order_tmp_bytes(int ntotal, int key_bits);
hipError_t launch_batch_order(const BatchOrder& B, hipStream_t s);

// Neighbour passes.  Slot t is particle perm[t] (perm null: t); inputs of
// the lists' partners come from the slot records, per-particle outputs go
// to particle perm[t] of the caller's arrays, and the lists hold slot
// indices (one list per particle).  posm / vel (index order) serve the
// finish kernel of configurations without SPH.
struct NbrLaunch {
    const float4* posm;
    const float4* vel;
    const float4* xs;  // positions in slot order (posm itself when perm is null)
    SlotRec* rec;
    const float* dstress;
    const float* balsara_in;
    const int* perm;
    const int* nbr;
    const int* nbr_cnt;
    int n;
    int kmax;
    float* balsara;
    float* density_dt;
    float* dstress_dt;
    float4* acc;
    float4* xvel;
    float* maxvsig;
    bool acc_has_gravity;
    int sub;        // lanes per particle in the passes (one instance only): 1 or 4
    BlockTables T;  // batch tables (T.nblocks == 0 for one instance)
};
hipError_t launch_density(const NbrLaunch& L, float* density, const DevParams& P, hipStream_t s);
hipError_t launch_strain(const NbrLaunch& L, const DevParams& P, hipStream_t s);
hipError_t launch_prep(const NbrLaunch& L, const DevParams& P, hipStream_t s);
hipError_t launch_accel(const NbrLaunch& L, const DevParams& P, hipStream_t s);

struct IntegLaunch {
    float4* posm;
    double* pos64;
    float4* vel;
    const float4* acc;
    const float4* xvel;
    float* density;
    const float* density_dt;
    float* dstress;
    const float* dstress_dt;
    int n;
    int not_first_step;
    BlockTables T;  // batch tables (T.nblocks == 0 for one instance)
    // variable-step leapfrog: per instance {current dt, next dt}; null for the fixed step
    float* ts;
    const float* maxvsig;  // read by the velocity criterion
    int n_instances;
};
// Fixed step: fixedTsLeapfrog.  With L.ts: the second kernel of the
// variable step (kick + drift with dt = max(next dt, min_timestep)).
hipError_t launch_integrate(const IntegLaunch& L, const DevParams& P, hipStream_t s);
// First kernel of the variable step: half kick with the current dt and the
// dt criteria folded into the next dt (atomic minimum per block).
hipError_t launch_vts_kick(const IntegLaunch& L, const DevParams& P, hipStream_t s);
// next dt of every instance := initial_timestep (which 0) or max_timestep
// (which 1); per-instance parameters from prm (null: P for one instance)
hipError_t launch_vts_set_next(float* ts, int n_instances, const DevParams* prm, const DevParams& P, int which,
                               hipStream_t s);
// Both stages of one integrate(pb, notFirstStep) call.
hipError_t launch_integrate_step(const IntegLaunch& L, const DevParams& P, hipStream_t s);

struct InvLaunch {
    const float4* posm;
    const double* pos64;
    const float4* vel;
    const float4* acc;
    const float* density;
    const float* dstress;
    const float* pot;  // may be null
    int n;
    double half_kick;
    double* partials;  // [nblocks][GRASPH_NINV] then final [GRASPH_NINV]
    int nblocks;
};
int inv_blocks(int n);
hipError_t launch_invariants(const InvLaunch& L, const grasph_params& prm, hipStream_t s);

// Morton reorder: perm[t] = particle index of the t-th particle along a
// Z-order curve over the bounding box of posm.
size_t reorder_tmp_bytes(int n);
hipError_t launch_reorder(const float4* posm, int n, Ctx& c, hipStream_t s);
// Slot-ordered copies for the sweep and the list passes (perm null: the
// identity): xs[t] = posm[perm[t]] when xs is given, and the slot records'
// position/mass and velocity/density (vel or density null: zeros there).
// slot-ordered copies; with mass_flags (one int per 256 slots) also the
// per-block half of the equal-mass check, which launch_mass_final completes
hipError_t launch_gather(const int* perm, int n, const float4* posm, const float4* vel, const float* density,
                         float4* xs, SlotRec* rec, hipStream_t s, int* mass_flags = nullptr);
hipError_t launch_mass_final(const float4* posm, const int* mass_flags, int nflags, float* umass, hipStream_t s);

// environmentAcceleration and the zero XVEL / MAXVSIG stores of
// calcAcceleration for configurations without SPH (batched form).
hipError_t launch_finish(const NbrLaunch& L, const DevParams& P, hipStream_t s);

// ---- host helpers shared by the C-ABI units -------------------------------
inline int fail(Ctx& c, int code, const std::string& msg)
{
    c.err = msg;
    return code;
}

inline int hip_fail(Ctx& c, hipError_t e, const char* where)
{
    c.err = std::string(where) + ": " + hipGetErrorString(e);
    (void)hipGetLastError();
    return GRASPH_E_HIP;
}

template <typename T>
int grow(Ctx& c, T*& ptr, int64_t& cap, int64_t need, const char* what)
{
    if (need <= cap && ptr) return 0;
    if (c.no_alloc) return fail(c, GRASPH_E_NOMEM, std::string(what) + " would need a device allocation");
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    hipError_t e = hipMalloc((void**)&ptr, (size_t)(need > 0 ? need : 1) * sizeof(T));
    if (e != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        return fail(c, GRASPH_E_NOMEM, std::string("device allocation of ") + what + " failed");
    }
    cap = need;
    return 0;
}

inline int check_params(Ctx& c, const grasph_params* prm)
{
    if (!prm) return fail(c, GRASPH_E_INVALID, "params is NULL");
    if (prm->flags & ~GRASPH_ALL_FLAGS) return fail(c, GRASPH_E_INVALID, "unknown flag bits");
    if (prm->dimension < 1 || prm->dimension > 3) return fail(c, GRASPH_E_INVALID, "dimension must be 1, 2 or 3");
    if (!(prm->H >= 0.0f) || !(prm->H < 3.0e38f)) return fail(c, GRASPH_E_INVALID, "H must be finite and >= 0");
    if ((prm->flags & GRASPH_SPH) && !(prm->H > 0.0f)) return fail(c, GRASPH_E_INVALID, "SPH needs H > 0");
    if (prm->flags & GRASPH_CLOHESSY_WILTSHIRE)
        return fail(c, GRASPH_E_UNSUPPORTED,
                    "CLOHESSY_WILTSHIRE is not supported: src/computeDerivatives.h:37-39 adds a scalar to a "
                    "float3 through an mpUtils operator whose meaning no available source fixes");
    if ((prm->flags & GRASPH_SOLIDS) && (prm->flags & GRASPH_ARTIFICIAL_STRESS))
        return fail(c, GRASPH_E_UNSUPPORTED,
                    "SOLIDS with ARTIFICIAL_STRESS is not supported (the reference does not compile it, "
                    "src/computeDerivatives.h:267)");
    if ((prm->flags & GRASPH_PLASTICITY_MIESE) && (prm->flags & GRASPH_PLASTICITY_MC))
        return fail(c, GRASPH_E_INVALID, "choose up to one plasticity model (settings.h:203-205)");
    if (prm->flags & GRASPH_VARIABLE_TIMESTEP) {
        const float v[5] = {prm->initial_timestep, prm->min_timestep, prm->max_timestep, prm->accel_accuracy,
                            prm->velocity_accuracy};
        for (float x : v)
            if (!(x >= 0.0f) || !(x < 3.0e38f))
                return fail(c, GRASPH_E_INVALID, "variable time-step settings must be finite and >= 0");
    }
    return 0;
}

inline int check_particles(Ctx& c, const grasph_particles* p)
{
    if (!p) return fail(c, GRASPH_E_INVALID, "particles is NULL");
    if (p->n < 0) return fail(c, GRASPH_E_INVALID, "negative particle count");
    if (p->n > 0x7ffffff0LL) return fail(c, GRASPH_E_INVALID, "particle count exceeds 2^31 - 16");
    if (p->n > 0 && !p->posm) return fail(c, GRASPH_E_INVALID, "posm is NULL");
    return 0;
}

// j-split of the single-instance one-sided sweep for n particles at ni per lane
void choose_slices(const Ctx& c, int64_t n, int ni, int* slices, int* slice_len);

// arrays each pass needs, by configuration; 0 or a GRASPH_E_INVALID code
int check_derivative_arrays(Ctx& c, const grasph_particles* p, uint32_t f, int which);
int check_integrate_arrays(Ctx& c, const grasph_particles* p, uint32_t f);

struct BatchState;
void destroy_batch_state(BatchState* b);

}  // namespace grasph

struct grasph_ctx {
    grasph::Ctx c;
};

#define GRASPH_NEED(ptr, name)                                                                                \
    do {                                                                                                      \
        if (!(ptr)) return grasph::fail(c, GRASPH_E_INVALID, std::string(name) + " is NULL but the flags need it"); \
    } while (0)
