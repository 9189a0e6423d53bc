// capi.hip -- the C ABI (include/grasph2.h): argument checks, scratch
// management, pass orchestration and per-kernel timing.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <new>
#include <string>

#include "internal.hpp"

// Work arrays of grasph_compute_derivatives_noalloc: one copy per GPU, at
// its own address on each.  Host code never uses this name as an address:
// scratch_address() resolves it once per device and every kernel argument
// is cut from that address.
static __device__ __align__(256) unsigned char scratch[GRASPH_NOALLOC_SCRATCH_BYTES];

namespace grasph {

static thread_local std::string g_create_err = "no error";

ProfScope::ProfScope(Ctx* ctx, hipStream_t st, int k) : c(ctx), s(st), cls(k)
{
    if (!c || !c->profile || c->capturing) return;
    for (int q = 0; q < 2; ++q) {
        hipEvent_t e = nullptr;
        if (!c->ev_free.empty()) {
            e = c->ev_free.back();
            c->ev_free.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        (q == 0 ? a : b) = e;
    }
    if (a) (void)hipEventRecord(a, s);
}

ProfScope::~ProfScope()
{
    if (!a || !b) {
        if (a) c->ev_free.push_back(a);
        if (b) c->ev_free.push_back(b);
        return;
    }
    (void)hipEventRecord(b, s);
    c->ev_pending.push_back(Ctx::Ev{a, b, cls});
}

// j-split of the single-instance sweep: enough slices for ~16 wavefronts
// per SIMD (measured: the sweep keeps gaining up to there, as more resident
// waves hide the scalar-load and rare-path latency), at most 32 slices of
// at least 512 j each; slice length a multiple of 8.
void choose_slices(const Ctx& c, int64_t n, int ni, int* slices, int* slice_len)
{
    int S = c.sweep_slices;
    if (S <= 0) {
        const int64_t waves = (n + 64 * ni - 1) / (64 * ni);
        const int64_t target = 16 * 4 * (int64_t)c.cus;
        S = (int)std::min<int64_t>(32, std::max<int64_t>(1, (target + waves - 1) / waves));
        S = (int)std::max<int64_t>(1, std::min<int64_t>(S, n / 512));
    }
    int64_t len = (n + S - 1) / S;
    len = (len + 7) / 8 * 8;
    if (len < 8) len = 8;
    *slice_len = (int)len;
    *slices = (int)((n + len - 1) / len);
    if (*slices < 1) *slices = 1;
}

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_dev;       // per device: the array's address, once resolved
std::vector<std::string> g_scratch_err;  // per device: why it could not be

}  // namespace

// The scratch array's address on `device`, which must be current: asked of
// the runtime once per device, then kept.  Null (and err set) on failure.
void* scratch_address(int device, std::string& err)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= device) {
        g_scratch_dev.resize(device + 1, nullptr);
        g_scratch_err.resize(device + 1);
    }
    if (!g_scratch_dev[device] && g_scratch_err[device].empty()) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            g_scratch_err[device] = std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e);
            (void)hipGetLastError();
        } else {
            g_scratch_dev[device] = ptr;
        }
    }
    err = g_scratch_err[device];
    return g_scratch_dev[device];
}

namespace {

int ensure_order(Ctx& c, const grasph_particles* p, hipStream_t s, bool count_eval)
{
    const int64_t n = p->n;
    if (c.reorder_interval <= 0) return 0;
    if (c.reorder_cap < n || !c.perm) {
        if (c.no_alloc) return fail(c, GRASPH_E_NOMEM, "the Morton order would need a device allocation");
        void* bufs[] = {c.perm, c.keys, c.keys_alt, c.vals};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        c.perm = nullptr; c.keys = nullptr; c.keys_alt = nullptr; c.vals = nullptr;
        c.reorder_cap = 0;
        c.perm_n = 0;
        if (hipMalloc((void**)&c.perm, n * sizeof(int)) != hipSuccess ||
            hipMalloc((void**)&c.keys, n * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc((void**)&c.keys_alt, n * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc((void**)&c.vals, n * sizeof(int)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, GRASPH_E_NOMEM, "device allocation of the Morton order failed");
        }
        c.reorder_cap = n;
    }
    if (!c.bbox && c.no_alloc) return fail(c, GRASPH_E_NOMEM, "the bounding box would need a device allocation");
    if (!c.bbox && hipMalloc((void**)&c.bbox, 257 * 6 * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, GRASPH_E_NOMEM, "device allocation of the bounding box failed");
    }
    const bool stale = c.perm_n != n || c.perm_src != (const void*)p->posm ||
                       c.evals_since_reorder >= c.reorder_interval;
    if (stale) {
        const size_t need = reorder_tmp_bytes((int)n);
        if (need > c.sort_tmp_bytes || !c.sort_tmp) {
            if (c.no_alloc) return fail(c, GRASPH_E_NOMEM, "sort storage would need a device allocation");
            if (c.sort_tmp) (void)hipFree(c.sort_tmp);
            c.sort_tmp = nullptr;
            c.sort_tmp_bytes = 0;
            if (hipMalloc(&c.sort_tmp, need ? need : 1) != hipSuccess) {
                (void)hipGetLastError();
                return fail(c, GRASPH_E_NOMEM, "device allocation of sort storage failed");
            }
            c.sort_tmp_bytes = need;
        }
        ProfScope ps(&c, s, KC_REORDER);
        hipError_t e = launch_reorder((const float4*)p->posm, (int)n, c, s);
        if (e != hipSuccess) return hip_fail(c, e, "Morton reorder");
        c.perm_n = n;
        c.perm_src = p->posm;
        c.evals_since_reorder = 0;
    }
    if (count_eval) c.evals_since_reorder++;
    return 0;
}

const int* order_of(Ctx& c) { return c.reorder_interval > 0 ? c.perm : nullptr; }

// The list capacity an earlier evaluation asked for: once the copy of its
// largest neighbour count has landed (queried, never waited for), a count
// above max_neighbours raises it to 1.25 times that count, in steps of 32,
// at most 4096 and at most 16 GiB of list planes.  Overflowing particles
// take the exact all-partner scan meanwhile, so the results do not change.
// whether work enqueued on s is being captured into a graph (then the
// event query and the copy would break the capture)
bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return true;
    }
    return st != hipStreamCaptureStatusNone;
}

void grow_list_capacity(Ctx& c, int64_t n, int slices, hipStream_t s)
{
    if (!c.list_max_pending || c.capturing || stream_capturing(s)) return;
    const hipError_t q = hipEventQuery(c.list_max_ev);
    if (q != hipSuccess) {
        (void)hipGetLastError();  // not ready (or unknown): look again next time
        return;
    }
    c.list_max_pending = false;
    const int m = *c.list_max_host;
    if (!c.max_neighbours_grow || m <= c.max_neighbours) return;
    int64_t k = ((int64_t)m + m / 4 + 31) / 32 * 32;
    const int64_t cap = (int64_t(16) << 30) / (4 * std::max<int64_t>(1, n) * std::max(1, slices));
    k = std::min<int64_t>(std::min<int64_t>(k, 4096), cap / 32 * 32);
    if (k > c.max_neighbours) c.max_neighbours = (int)k;
}

// After the lists of an evaluation are final: their largest count, reduced
// on ws and copied to pinned host memory behind an event (not while the
// caller's stream is being captured into a graph, nor without growth).
int watch_list_counts(Ctx& c, int64_t n, hipStream_t s, hipStream_t ws)
{
    if (!c.max_neighbours_grow || c.no_alloc || c.capturing || stream_capturing(s)) return 0;
    hipError_t e;
    if (!c.list_max_dev && (e = hipMalloc((void**)&c.list_max_dev, sizeof(int))) != hipSuccess)
        return hip_fail(c, e, "neighbour-count maximum");
    if (!c.list_max_host && (e = hipHostMalloc((void**)&c.list_max_host, sizeof(int), hipHostMallocDefault)) != hipSuccess)
        return hip_fail(c, e, "pinned neighbour-count maximum");
    if (!c.list_max_ev && (e = hipEventCreateWithFlags(&c.list_max_ev, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(c, e, "neighbour-count event");
    ProfScope ps(&c, ws, KC_OTHER);
    if ((e = launch_count_max(c.nbr_cnt, (int)n, c.list_max_dev, ws)) != hipSuccess ||
        (e = hipMemcpyAsync(c.list_max_host, c.list_max_dev, sizeof(int), hipMemcpyDeviceToHost, ws)) != hipSuccess ||
        (e = hipEventRecord(c.list_max_ev, ws)) != hipSuccess)
        return hip_fail(c, e, "neighbour-count watch");
    c.list_max_pending = true;
    return 0;
}

// lists for `slices` slices: slices planes of list_plane(n, kmax) entries
// (slots rounded up to tiles of 64) and slices * n counts
int ensure_nbr(Ctx& c, int64_t n, int slices)
{
    const int k = c.max_neighbours;
    const int64_t need_n = (n + 63) / 64 * 64 * slices;
    if (c.nbr && c.nbr_cap_n >= need_n && c.nbr_cap_k >= k) return 0;
    if (c.no_alloc) return fail(c, GRASPH_E_NOMEM, "the neighbour list would need a device allocation");
    if (c.nbr) (void)hipFree(c.nbr);
    if (c.nbr_cnt) (void)hipFree(c.nbr_cnt);
    c.nbr = nullptr; c.nbr_cnt = nullptr; c.nbr_cap_n = 0; c.nbr_cap_k = 0;
    if (hipMalloc((void**)&c.nbr, (size_t)need_n * (size_t)k * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&c.nbr_cnt, (size_t)need_n * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, GRASPH_E_NOMEM, "device allocation of the neighbour list failed");
    }
    c.nbr_cap_n = need_n;
    c.nbr_cap_k = k;
    return 0;
}

SweepLaunch sweep_base(Ctx& c, const grasph_particles* p)
{
    SweepLaunch L;
    std::memset(&L, 0, sizeof L);
    L.posm = (const float4*)p->posm;
    L.perm = order_of(c);
    L.n = (int)p->n;
    L.acc = (float4*)p->acc;
    L.nbr = c.nbr;
    L.nbr_cnt = c.nbr_cnt;
    L.pot = c.pot;
    L.kmax = c.max_neighbours;
    L.vel = (const float4*)p->vel;
    L.xvel = (float4*)p->xvel;
    L.maxvsig = p->maxvsig;
    L.ni = c.sweep_ni;
    return L;
}

NbrLaunch nbr_base(Ctx& c, const grasph_particles* p)
{
    NbrLaunch L;
    std::memset(&L, 0, sizeof L);
    L.posm = (const float4*)p->posm;
    L.vel = (const float4*)p->vel;
    L.rec = c.rec;
    L.dstress = p->dstress;
    L.balsara_in = p->balsara;
    L.perm = order_of(c);
    L.nbr = c.nbr;
    L.nbr_cnt = c.nbr_cnt;
    L.n = (int)p->n;
    // long lists (a capacity grown past 512 after contact): 4 lanes per
    // particle, 4 times the wavefronts to hide the gathers' latency
    L.sub = c.max_neighbours > 512 ? 4 : 1;
    L.kmax = c.max_neighbours;
    L.balsara = p->balsara;
    L.density_dt = p->density_dt;
    L.dstress_dt = p->dstress_dt;
    L.acc = (float4*)p->acc;
    L.xvel = (float4*)p->xvel;
    L.maxvsig = p->maxvsig;
    return L;
}

// The symmetric sweep's default thresholds (option sweep_symmetric = 2),
// measured step times against the one-sided sweep: with neighbour lists
// (collision) 0.319 against 0.322 ms at 16384 particles, 0.269 against
// 0.221 at 8192; gravity alone (Plummer) 0.106 against 0.116 ms at 16384,
// 0.216 against 0.222 at 24576, 0.268 against 0.369 at 32768.
constexpr int64_t kSymAutoMinLists = 16384;
constexpr int64_t kSymAutoMinGravity = 16384;

int sym_wgs(const Ctx& c) { return c.sym_workgroups > 0 ? c.sym_workgroups : sym_default_workgroups(c.cus); }

// shares of the symmetric sweep's near field: ~16 wavefronts per SIMD, at most 16
int sym_near_shares(const Ctx& c, int64_t n)
{
    if (c.sym_near_shares > 0) return c.sym_near_shares;
    const int64_t blocks = (n + 63) / 64;
    const int64_t target = 16 * 4 * (int64_t)c.cus;
    return (int)std::min<int64_t>(16, std::max<int64_t>(1, (target + blocks - 1) / blocks));
}

// The symmetric sweep, begun: buffers; the rest of the evaluation forked
// onto the context's second stream (ws) after the gather and the mass
// check; the far field on s, launched first so that its workgroups are
// placed first and no event stands between it and the work before it; the
// near field's lists on ws.  The far field needs only the slot-ordered
// positions and the mass record; the work enqueued on ws after this (list
// merge, the SPH passes) runs beside it.  sym_sweep_end enqueues the near
// field's in-group gravity on ws, joins ws to s and combines on s.
int sym_sweep_begin(Ctx& c, const SweepLaunch& L, const DevParams& P, hipStream_t s, SymLaunch& S,
                    hipStream_t& ws)
{
    const int64_t col_elems = c.sym_partial_bytes / (int64_t)sizeof(float);
    const SymGeom G = sym_geometry(L.n, sym_wgs(c), L.want_nbr, L.near_shares, col_elems);
    int r;
    if ((r = grow(c, c.sym_box, c.sym_box_cap, G.box_elems, "symmetric sweep boxes")) ||
        (r = grow(c, c.sym_bits, c.sym_bits_cap, G.bits_elems, "symmetric sweep near bits")) ||
        (r = grow(c, c.sym_near, c.sym_near_cap, G.near_elems, "symmetric sweep near sums")) ||
        (r = grow(c, c.sym_apart, c.sym_apart_cap, G.apart_elems, "symmetric sweep own-side sums")) ||
        (r = grow(c, c.sym_bpart, c.sym_bpart_cap, G.bpart_elems, "symmetric sweep partner-side sums")) ||
        (G.ncols > 1 && L.grav &&
         (r = grow(c, c.sym_accsum, c.sym_accsum_cap, 2 * (int64_t)L.n, "symmetric sweep column sums"))))
        return r;
    std::memset(&S, 0, sizeof S);
    S.xs = L.xs ? L.xs : L.posm;
    S.perm = L.perm;
    S.n = L.n;
    S.grav = L.grav;
    S.want_nbr = L.want_nbr;
    S.finish = L.finish;
    S.nshares = L.near_shares;
    S.list_shares = L.nslices;
    S.acc = L.acc;
    S.nbr = L.nbr;
    S.nbr_cnt = L.nbr_cnt;
    S.kmax = L.kmax;
    S.vel = L.vel;
    S.xvel = L.xvel;
    S.maxvsig = L.maxvsig;
    S.umass = L.umass;
    S.wgs = G.wgs;
    S.box = c.sym_box;
    S.bits = c.sym_bits;
    S.near_part = c.sym_near;
    S.apart = c.sym_apart;
    S.bpart = c.sym_bpart;
    S.col_elems = col_elems;
    S.accsum = c.sym_accsum;
    hipError_t e;
    ws = s;
    // With gravity alone, only the in-group gravity (enqueued by
    // sym_sweep_end) runs on the second stream, in the issue cycles the far
    // field's wavefronts leave (Plummer N = 16384: 0.103 -> 0.099 ms per step
    // against both in order on s; launched before the far field it delayed
    // it, 0.101).
    if (sym_has_far(S)) {
        if (c.no_alloc && (!c.side || !c.ev_fork || !c.ev_join))
            return fail(c, GRASPH_E_NOMEM, "the second stream would need creating");
        if (!c.side && (e = hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking)) != hipSuccess)
            return hip_fail(c, e, "second stream");
        if ((!c.ev_fork && (e = hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming)) != hipSuccess) ||
            (!c.ev_join && (e = hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming)) != hipSuccess))
            return hip_fail(c, e, "fork and join events");
        if ((e = hipEventRecord(c.ev_fork, s)) != hipSuccess || (e = hipStreamWaitEvent(c.side, c.ev_fork, 0)) != hipSuccess)
            return hip_fail(c, e, "fork");
        {
            ProfScope ps(&c, s, KC_FAR);
            if ((e = launch_sym_far(S, P, s)) != hipSuccess) return hip_fail(c, e, "far-field launch");
        }
        ws = c.side;
        // Below ~50k particles the far field's window is short next to the
        // lists' work: there the lists take issue priority, so they and the
        // passes after them end inside it (collision N = 32768: 0.435 ->
        // 0.410 ms per step, N = 16384: 0.290 -> 0.283; liquid N = 32768:
        // 0.420 -> 0.387); from 65536 on the side stream ends inside the
        // window without it, and its VALU work would lengthen the far field
        // (collision N = 65536: 0.985 -> 1.008 ms, liquid N = 65536: 0.945
        // -> 1.005, collision N = 131072: 3.79 -> 3.86).  Once the lists have
        // grown long (capacity past 512: bodies in contact), the lists and
        // the passes after them outlast the far field at any size, and the
        // priority shortens the step (collision N = 131072 after impact:
        // 5.87 -> 5.57 ms).
        S.list_prio = c.sym_list_priority == 1 ||
                      (c.sym_list_priority == 2 && (L.n < 49152 || c.max_neighbours > 512));
    }
    ProfScope ps(&c, ws, KC_SWEEP);
    if ((e = launch_sym_lists(S, P, ws)) != hipSuccess) return hip_fail(c, e, "near-field list launch");
    return 0;
}

int sym_sweep_end(Ctx& c, SymLaunch& S, const DevParams& P, hipStream_t s, hipStream_t ws, bool add_acc)
{
    hipError_t e;
    S.add_acc = add_acc;
    {
        // the gravity inside the groups last: only the combine needs it
        ProfScope ps(&c, ws, KC_SWEEP);
        if ((e = launch_sym_ingroup(S, P, ws)) != hipSuccess) return hip_fail(c, e, "near-field gravity launch");
    }
    if (ws != s && ((e = hipEventRecord(c.ev_join, ws)) != hipSuccess || (e = hipStreamWaitEvent(s, c.ev_join, 0))))
        return hip_fail(c, e, "join");
    ProfScope ps(&c, s, KC_SWEEP);
    if ((e = launch_sym_combine(S, P, s)) != hipSuccess) return hip_fail(c, e, "symmetric combine launch");
    return 0;
}

int run_sweep(Ctx& c, const SweepLaunch& L, const DevParams& P, hipStream_t s)
{
    ProfScope ps(&c, s, KC_SWEEP);
    hipError_t e = launch_sweep(L, P, s);
    if (e != hipSuccess) return hip_fail(c, e, "sweep launch");
    return 0;
}

}  // namespace

// which == 0: full computeDerivatives; 1: calcDensity; 2: calcBalsara...; 3: calcAcceleration
int check_derivative_arrays(Ctx& c, const grasph_particles* p, uint32_t f, int which)
{
    const bool sph = f & GRASPH_SPH;
    const bool dens_pass = (which == 0) ? (sph && !(f & GRASPH_INTEGRATE_DENSITY)) : (which == 1);
    const bool strain_pass =
        (which == 0) ? (sph && ((f & GRASPH_SOLIDS) || (f & GRASPH_INTEGRATE_DENSITY) || (f & GRASPH_BALSARA_SWITCH)))
                     : (which == 2);
    const bool accel_pass = (which == 0) || (which == 3);
    if (accel_pass) GRASPH_NEED(p->acc, "acc");
    if (dens_pass) GRASPH_NEED(p->density, "density");
    if (strain_pass) {
        GRASPH_NEED(p->vel, "vel");
        GRASPH_NEED(p->density, "density");
        if (f & GRASPH_SOLIDS) { GRASPH_NEED(p->dstress, "dstress"); }
        if (f & GRASPH_SOLIDS) { GRASPH_NEED(p->dstress_dt, "dstress_dt"); }
        if (f & GRASPH_BALSARA_SWITCH) { GRASPH_NEED(p->balsara, "balsara"); }
        if (f & GRASPH_INTEGRATE_DENSITY) { GRASPH_NEED(p->density_dt, "density_dt"); }
    }
    if (accel_pass && sph) {
        GRASPH_NEED(p->vel, "vel");
        GRASPH_NEED(p->density, "density");
        if (f & GRASPH_SOLIDS) { GRASPH_NEED(p->dstress, "dstress"); }
        if ((f & GRASPH_BALSARA_SWITCH) && which == 3) { GRASPH_NEED(p->balsara, "balsara"); }
        if (f & GRASPH_XSPH) { GRASPH_NEED(p->xvel, "xvel"); }
    }
    return 0;
}

int check_integrate_arrays(Ctx& c, const grasph_particles* p, uint32_t f)
{
    GRASPH_NEED(p->vel, "vel");
    GRASPH_NEED(p->acc, "acc");
    if ((f & GRASPH_SPH) && (f & GRASPH_INTEGRATE_DENSITY)) {
        GRASPH_NEED(p->density, "density");
        GRASPH_NEED(p->density_dt, "density_dt");
    }
    if ((f & GRASPH_SPH) && (f & GRASPH_SOLIDS)) {
        GRASPH_NEED(p->dstress, "dstress");
        GRASPH_NEED(p->dstress_dt, "dstress_dt");
        if (f & GRASPH_PLASTICITY_MC) { GRASPH_NEED(p->density, "density"); }
    }
    if ((f & GRASPH_SPH) && (f & GRASPH_XSPH)) GRASPH_NEED(p->xvel, "xvel");
    if ((f & GRASPH_VARIABLE_TIMESTEP) && (f & GRASPH_VELOCITY_CRITERION)) GRASPH_NEED(p->maxvsig, "maxvsig");
    return 0;
}

namespace {

int derivatives(Ctx& c, const grasph_particles* p, const grasph_params* prm, void* stream, int which)
{
    int r;
    if ((r = check_params(c, prm)) || (r = check_particles(c, p))) return r;
    if (p->n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    const uint32_t f = prm->flags;
    const bool sph = f & GRASPH_SPH, grav = f & GRASPH_SELF_GRAVITY;
    const bool dens_pass = (which == 0) ? (sph && !(f & GRASPH_INTEGRATE_DENSITY)) : (which == 1);
    const bool strain_pass =
        (which == 0) ? (sph && ((f & GRASPH_SOLIDS) || (f & GRASPH_INTEGRATE_DENSITY) || (f & GRASPH_BALSARA_SWITCH)))
                     : (which == 2);
    const bool accel_pass = (which == 0) || (which == 3);
    const bool need_list = (which == 0 || which == 3) ? sph : true;
    if ((which == 1 || which == 2) && !(prm->H > 0.0f)) return fail(c, GRASPH_E_INVALID, "SPH passes need H > 0");

    if ((r = check_derivative_arrays(c, p, f, which))) return r;

    const int64_t n = p->n;
    // the order only matters for the neighbour branch: gravity alone runs in index order
    if (need_list && (r = ensure_order(c, p, s, true))) return r;
    const bool sweep_grav = accel_pass && grav;
    const bool sym_on = c.sweep_symmetric == 1 ||
                        (c.sweep_symmetric == 2 && n >= (need_list ? kSymAutoMinLists : kSymAutoMinGravity));
    const bool sym = sym_on && (sweep_grav || need_list);
    int slices = 1, slice_len = (int)n;
    if ((sweep_grav || need_list) && !sym) choose_slices(c, n, c.sweep_ni, &slices, &slice_len);
    int near_shares = 1;
    if (sym) {
        // the near field's list shares take the slices' list planes; with
        // long lists (after a collision's impact) one share: the merge of
        // shares would move every entry again (1e8 entries: 1.4 ms), and
        // the one list kernel still ends inside the far field's window
        near_shares = sym_near_shares(c, n);
        slices = (int64_t)n * c.max_neighbours >= (int64_t(1) << 26) ? 1 : near_shares;
    }
    if (need_list) {
        grow_list_capacity(c, n, slices, s);
        if ((r = ensure_nbr(c, n, slices))) return r;
        c.last_list_n = n;
        c.last_list_slices = slices;
    }
    if (slices > 1 && sweep_grav && !sym) {
        if ((r = grow(c, c.part, c.part_cap, 2 * n * slices, "sweep partial sums"))) return r;
    }
    const DevParams P = make_dev_params(prm);

    // slot-ordered positions for the sweep and the passes' slot records
    const int* perm = need_list ? order_of(c) : nullptr;
    const float4* xs = (const float4*)p->posm;
    const int gather_blocks = (int)((n + 255) / 256);
    bool mass_flags = false;  // the gather makes the equal-mass check's per-block flags
    if (need_list) {
        const bool reads_vel = strain_pass || (accel_pass && sph);
        const bool gdens = reads_vel && !dens_pass;  // else the density pass fills the records' densities
        mass_flags = sweep_grav;
        if ((perm && (r = grow(c, c.xs, c.xs_cap, n, "slot-ordered positions"))) ||
            (r = grow(c, c.rec, c.rec_cap, n, "slot records")) ||
            (mass_flags && (r = grow(c, c.mass_flags, c.mass_flags_cap, gather_blocks, "mass-check flags"))))
            return r;
        ProfScope ps(&c, s, KC_REORDER);
        hipError_t e = launch_gather(perm, (int)n, (const float4*)p->posm, reads_vel ? (const float4*)p->vel : nullptr,
                                     gdens ? p->density : nullptr, perm ? c.xs : nullptr, c.rec, s,
                                     mass_flags ? c.mass_flags : nullptr);
        if (e != hipSuccess) return hip_fail(c, e, "slot-order gather launch");
        if (perm) xs = c.xs;
    }

    // one all-pairs sweep
    SweepLaunch L = sweep_base(c, p);
    L.perm = perm;
    L.xs = xs;
    L.grav = sweep_grav;
    L.want_nbr = need_list;
    L.want_pot = false;
    L.finish = accel_pass && !sph;
    L.nslices = slices;
    L.near_shares = near_shares;
    L.slice_len = slice_len;
    L.part = c.part;
    if (sweep_grav) {
        // equal masses let the sweep factor the mass out of the gravity sum
        ProfScope ps(&c, s, KC_OTHER);
        hipError_t e = mass_flags ? launch_mass_final(L.posm, c.mass_flags, gather_blocks, c.umass, s)
                                  : launch_mass_check(L.posm, nullptr, 1, L.n, c.umass, s);
        if (e != hipSuccess) return hip_fail(c, e, "mass check launch");
        L.umass = c.umass;
    }
    SymLaunch S;
    hipStream_t ws = s;  // the near field and the passes: the forked stream beside the far field, or s
    if (sym) {
        if ((r = sym_sweep_begin(c, L, P, s, S, ws))) return r;
    } else if ((r = run_sweep(c, L, P, s))) {
        return r;
    }
    if (need_list && slices > 1) {
        // one list per particle for the passes (walking per-slice lists costs them up to 2x)
        ProfScope ps(&c, ws, KC_OTHER);
        hipError_t e = launch_merge_lists(c.nbr, c.nbr_cnt, (int)n, c.max_neighbours, slices, ws);
        if (e != hipSuccess) return hip_fail(c, e, "list merge launch");
        c.last_list_slices = 1;
    }
    if (need_list && (r = watch_list_counts(c, n, s, ws))) return r;

    NbrLaunch N = nbr_base(c, p);
    N.perm = perm;
    N.xs = xs;
    N.rec = c.rec;
    if (dens_pass) {
        ProfScope ps(&c, ws, KC_DENSITY);
        hipError_t e = launch_density(N, p->density, P, ws);
        if (e != hipSuccess) return hip_fail(c, e, "density pass launch");
    }
    if (strain_pass) {
        ProfScope ps(&c, ws, KC_STRAIN);
        hipError_t e = launch_strain(N, P, ws);
        if (e != hipSuccess) return hip_fail(c, e, "strain pass launch");
    }
    if (accel_pass && sph && !strain_pass) {
        ProfScope ps(&c, ws, KC_OTHER);
        hipError_t e = launch_prep(N, P, ws);
        if (e != hipSuccess) return hip_fail(c, e, "record pass launch");
    }
    // With the symmetric sweep, the acceleration pass runs while the far
    // field is still going and writes the SPH terms alone; the combine then
    // adds the gravity to them.  Otherwise it adds the SPH terms to the
    // gravity of the sweep.
    const bool accel_first = sym && accel_pass && sph;
    if (accel_first) {
        N.acc_has_gravity = false;
        ProfScope ps(&c, ws, KC_ACCEL);
        hipError_t e = launch_accel(N, P, ws);
        if (e != hipSuccess) return hip_fail(c, e, "acceleration pass launch");
    }
    if (accel_pass && (f & GRASPH_VARIABLE_TIMESTEP)) {
        // calcAcceleration::do_after sets nextTS = max_timestep
        // (computeDerivatives.h:305-308): the next integrate(pb, true) takes
        // min(max_timestep, this step's criteria), so dt can grow again.
        // Enqueued beside the far field when there is one (ws joins s below).
        hipError_t e = launch_vts_set_next(c.ts, 1, nullptr, P, 1, ws);
        if (e != hipSuccess) return hip_fail(c, e, "time-step reset launch");
    }
    if (sym && (r = sym_sweep_end(c, S, P, s, ws, accel_first && grav))) return r;
    if (accel_pass && sph && !accel_first) {
        N.acc_has_gravity = grav;
        ProfScope ps(&c, s, KC_ACCEL);
        hipError_t e = launch_accel(N, P, s);
        if (e != hipSuccess) return hip_fail(c, e, "acceleration pass launch");
    }
    return 0;
}

// Consecutive 256-byte-aligned pieces of one region; with base null it
// only measures (offsets from 0, no limit).
struct Carve {
    char* base;
    size_t cap;
    size_t used = 0;
    template <typename T>
    T* take(int64_t count)
    {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + (size_t)count * sizeof(T);
        return base && used <= cap ? reinterpret_cast<T*>(base + off) : nullptr;
    }
};

// The work arrays computeDerivatives needs on the one-sided sweep with
// `slices` j-slices, re-sorted every call: cut from `k` into `sh` (its
// capacities set to exactly what the path asks of grow / ensure_*).
void carve_derivative_arrays(Ctx& sh, Carve& k, int64_t n, bool lists, bool grav, int slices)
{
    if (lists) {
        sh.perm = k.take<int>(n);
        sh.keys = k.take<uint32_t>(n);
        sh.keys_alt = k.take<uint32_t>(n);
        sh.vals = k.take<int>(n);
        sh.reorder_cap = n;
        sh.sort_tmp_bytes = reorder_tmp_bytes((int)n);
        sh.sort_tmp = k.take<unsigned char>((int64_t)std::max<size_t>(sh.sort_tmp_bytes, 1));
        sh.bbox = k.take<float>(257 * 6);
        sh.xs = k.take<float4>(n);
        sh.xs_cap = n;
        sh.rec = k.take<SlotRec>(n);
        sh.rec_cap = n;
        if (grav) {
            sh.mass_flags_cap = (n + 255) / 256;
            sh.mass_flags = k.take<int>(sh.mass_flags_cap);
        }
        sh.nbr_cap_n = (n + 63) / 64 * 64 * slices;
        sh.nbr_cap_k = sh.max_neighbours;
        sh.nbr = k.take<int>(sh.nbr_cap_n * sh.max_neighbours);
        sh.nbr_cnt = k.take<int>(sh.nbr_cap_n);
    }
    if (slices > 1 && grav) {
        sh.part_cap = 2 * n * slices;
        sh.part = k.take<float4>(sh.part_cap);
    }
}

}  // namespace

int derivatives_noalloc(Ctx& c, const grasph_particles* p, const grasph_params* prm, void* stream)
{
    int r;
    if ((r = check_params(c, prm)) || (r = check_particles(c, p))) return r;
    if (!c.scratch)
        return fail(c, GRASPH_E_HIP, "no allocation-free scratch on device " + std::to_string(c.device) + ": " +
                                         c.scratch_err);
    if (p->n == 0) return 0;
    const int64_t n = p->n;
    const uint32_t f = prm->flags;
    const bool lists = f & GRASPH_SPH, grav = f & GRASPH_SELF_GRAVITY;

    // a stand-in context: this context's options and state ({current, next}
    // dt, the mass record), the one-sided sweep on one stream, a fresh
    // Morton order, buffers from the scratch array, no allocation allowed
    Ctx sh;
    sh.device = c.device;
    sh.cus = c.cus;
    sh.max_neighbours = c.max_neighbours;
    sh.sweep_ni = c.sweep_ni;
    sh.sweep_symmetric = 0;
    sh.reorder_interval = 1;
    sh.profile = false;
    sh.no_alloc = true;
    sh.ts = c.ts;
    sh.umass = c.umass;

    // the j-split: as many slices as the allocating path would take, fewer
    // until the arrays fit
    sh.sweep_slices = c.sweep_slices;
    int slices = 1, len = (int)n;
    if (grav || lists) choose_slices(sh, n, sh.sweep_ni, &slices, &len);
    size_t need = 0;
    for (;;) {
        Carve m{nullptr, 0};
        carve_derivative_arrays(sh, m, n, lists, grav, slices);
        need = m.used;
        if (need <= (size_t)GRASPH_NOALLOC_SCRATCH_BYTES || slices == 1) break;
        sh.sweep_slices = slices - 1;
        choose_slices(sh, n, sh.sweep_ni, &slices, &len);
    }
    if (need > (size_t)GRASPH_NOALLOC_SCRATCH_BYTES)
        return fail(c, GRASPH_E_INVALID,
                    "allocation-free derivatives of " + std::to_string(n) + " particles need " +
                        std::to_string(need) + " bytes of scratch with these flags and max_neighbours = " +
                        std::to_string(c.max_neighbours) + "; the scratch array holds " +
                        std::to_string((long long)GRASPH_NOALLOC_SCRATCH_BYTES));
    sh.sweep_slices = slices;
    Carve k{static_cast<char*>(c.scratch), (size_t)GRASPH_NOALLOC_SCRATCH_BYTES};
    carve_derivative_arrays(sh, k, n, lists, grav, slices);
    r = derivatives(sh, p, prm, stream, 0);
    if (r) c.err = sh.err;
    c.noalloc_slices = slices;
    return r;
}

}  // namespace grasph

using grasph::Ctx;

extern "C" {

int grasph_version(void) { return GRASPH_VERSION; }

const char* grasph_last_error(const grasph_ctx* ctx)
{
    return ctx ? ctx->c.err.c_str() : grasph::g_create_err.c_str();
}

void grasph_params_default(grasph_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    // settings/settings.h, including its variable-step integrator (:52-62)
    p->flags = GRASPH_SELF_GRAVITY | GRASPH_SPH | GRASPH_INTEGRATE_DENSITY | GRASPH_SOLIDS | GRASPH_PLASTICITY_MC |
               GRASPH_ARTIFICIAL_VISCOSITY | GRASPH_BALSARA_SWITCH | GRASPH_VARIABLE_TIMESTEP |
               GRASPH_ACCELERATION_CRITERION | GRASPH_VELOCITY_CRITERION;
    p->dimension = 3;
    const double count = 16384.0, radius = 0.1, tmass = 1.0;
    const double pradius = radius * std::pow(count, -1.0 / 3.0);  // :97-98
    p->H = (float)(pradius * 2.5);                                // :99
    const double rho0 = tmass / count / (4.0 / 3.0 * pradius * pradius * pradius * M_PI);  // :105
    p->rho0 = (float)rho0;
    p->BULK = 1024.0f;
    p->dBULKdP = 1.0f;
    p->SOUNDSPEED = (float)std::sqrt(1024.0 / rho0);  // :108
    p->shear = 128.0f;
    p->Y = 0.5f;
    p->tanfr = (float)std::tan(45.0 * M_PI / 180.0);  // :121-122
    p->cohesion = 0.0f;
    p->alpha = 1.0f;
    p->mateps = 0.4f;
    p->matexp = 4.0f;
    p->normalsep = p->H * 0.3f;
    p->xsph_factor = 0.5f;
    p->cw_n = (float)std::sqrt(1.0 / 3.0);
    p->constant_acceleration[1] = -1.0f;
    for (int d = 0; d < 3; ++d) {
        p->box_min[d] = -1.2f;
        p->box_max[d] = 1.2f;
    }
    p->box_reflectiveness = 0.5f;
    p->fixed_timestep = 0.002f;     // :50
    p->initial_timestep = 0.002f;   // :56
    p->min_timestep = 0.00002f;     // :57
    p->max_timestep = 0.02f;        // :58
    p->accel_accuracy = 0.01f;      // :60
    p->velocity_accuracy = 0.09f;   // :62
}

int grasph_create(int device, grasph_ctx** out)
{
    if (!out) {
        grasph::g_create_err = "out is NULL";
        return GRASPH_E_INVALID;
    }
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        grasph::g_create_err = std::string("no HIP device: ") + hipGetErrorString(e);
        (void)hipGetLastError();
        return GRASPH_E_HIP;
    }
    if (device < 0 || device >= count) {
        grasph::g_create_err = "device index out of range";
        return GRASPH_E_INVALID;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        grasph::g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return GRASPH_E_HIP;
    }
    grasph_ctx* ctx = new (std::nothrow) grasph_ctx;
    if (!ctx) {
        grasph::g_create_err = "out of host memory";
        return GRASPH_E_NOMEM;
    }
    ctx->c.device = device;
    // the variable-step leapfrog's {current dt, next dt}: zero, as the
    // reference's __device__ currentTS / nextTS start
    if (hipMalloc((void**)&ctx->c.ts, 2 * sizeof(float)) != hipSuccess ||
        hipMemset(ctx->c.ts, 0, 2 * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&ctx->c.umass, (2 + 256) * sizeof(float)) != hipSuccess) {
        (void)hipGetLastError();
        if (ctx->c.ts) (void)hipFree(ctx->c.ts);
        if (ctx->c.umass) (void)hipFree(ctx->c.umass);
        delete ctx;
        grasph::g_create_err = "device allocation of the context state failed";
        return GRASPH_E_NOMEM;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->c.cus = cus;
    // the allocation-free entry point's array on this device (a failure only
    // disables that entry point, which then reports it)
    ctx->c.scratch = grasph::scratch_address(device, ctx->c.scratch_err);
    (void)hipGetLastError();
    ctx->c.err = "no error";
    *out = ctx;
    return 0;
}

int grasph_destroy(grasph_ctx* ctx)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    (void)hipDeviceSynchronize();
    grasph::destroy_batch_state(c.batch);
    c.batch = nullptr;
    void* bufs[] = {c.perm, c.keys,   c.keys_alt, c.vals, c.sort_tmp, c.bbox,
                    c.nbr,  c.nbr_cnt, c.rec,     c.pot,  c.red,      c.part, c.ts, c.umass,
                    c.xs,   c.sym_box, c.sym_bits, c.sym_near, c.sym_apart, c.sym_bpart, c.sym_accsum, c.mass_flags};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (auto& ev : c.ev_pending) {
        (void)hipEventDestroy(ev.a);
        (void)hipEventDestroy(ev.b);
    }
    for (auto e : c.ev_free) (void)hipEventDestroy(e);
    if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
    if (c.ev_join) (void)hipEventDestroy(c.ev_join);
    if (c.list_max_ev) (void)hipEventDestroy(c.list_max_ev);
    if (c.list_max_dev) (void)hipFree(c.list_max_dev);
    if (c.list_max_host) (void)hipHostFree(c.list_max_host);
    if (c.side) (void)hipStreamDestroy(c.side);
    delete ctx;
    return 0;
}

int grasph_set_option(grasph_ctx* ctx, const char* key, int64_t value)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    if (!key) return grasph::fail(c, GRASPH_E_INVALID, "key is NULL");
    const std::string k(key);
    if (k == "max_neighbours") {
        if (value < 1 || value > 4096) return grasph::fail(c, GRASPH_E_INVALID, "max_neighbours must be in [1, 4096]");
        c.max_neighbours = c.max_neighbours_set = (int)value;
    } else if (k == "max_neighbours_grow") {
        if (value != 0 && value != 1) return grasph::fail(c, GRASPH_E_INVALID, "max_neighbours_grow must be 0 or 1");
        c.max_neighbours_grow = (int)value;
    } else if (k == "reorder_interval") {
        if (value < 0) return grasph::fail(c, GRASPH_E_INVALID, "reorder_interval must be >= 0");
        c.reorder_interval = (int)(value > 1000000000 ? 1000000000 : value);
        c.perm_n = 0;
    } else if (k == "sweep_i_per_lane") {
        if (value != 1 && value != 2) return grasph::fail(c, GRASPH_E_INVALID, "sweep_i_per_lane must be 1 or 2");
        c.sweep_ni = (int)value;
    } else if (k == "sweep_slices") {
        if (value < 0 || value > 64) return grasph::fail(c, GRASPH_E_INVALID, "sweep_slices must be in [0, 64]");
        c.sweep_slices = (int)value;
    } else if (k == "sweep_symmetric") {
        if (value < 0 || value > 2) return grasph::fail(c, GRASPH_E_INVALID, "sweep_symmetric must be 0, 1 or 2");
        c.sweep_symmetric = (int)value;
    } else if (k == "sym_partial_bytes") {
        if (value < (int64_t(1) << 20)) return grasph::fail(c, GRASPH_E_INVALID, "sym_partial_bytes must be at least 1 MiB");
        c.sym_partial_bytes = value;
    } else if (k == "sym_workgroups") {
        if (value < 0 || value > 65536) return grasph::fail(c, GRASPH_E_INVALID, "sym_workgroups must be in [0, 65536]");
        c.sym_workgroups = (int)value;
    } else if (k == "batch_symmetric") {
        if (value != 0 && value != 1) return grasph::fail(c, GRASPH_E_INVALID, "batch_symmetric must be 0 or 1");
        c.batch_symmetric = (int)value;
    } else if (k == "batch_symmetric_min") {
        if (value < 0 || value > (int64_t)1 << 30)
            return grasph::fail(c, GRASPH_E_INVALID, "batch_symmetric_min must be in [0, 2^30]");
        c.batch_symmetric_min = (int)value;
    } else if (k == "sym_list_priority") {
        if (value < 0 || value > 2) return grasph::fail(c, GRASPH_E_INVALID, "sym_list_priority must be 0, 1 or 2");
        c.sym_list_priority = (int)value;
    } else if (k == "sym_near_shares") {
        if (value < 0 || value > 16) return grasph::fail(c, GRASPH_E_INVALID, "sym_near_shares must be in [0, 16]");
        c.sym_near_shares = (int)value;
    } else if (k == "batch_waves_per_block") {
        if (value != 1 && value != 2 && value != 4)
            return grasph::fail(c, GRASPH_E_INVALID, "batch_waves_per_block must be 1, 2 or 4");
        c.batch_wpb = (int)value;
    } else if (k == "batch_blocks_per_cu") {
        if (value < 1 || value > 64) return grasph::fail(c, GRASPH_E_INVALID, "batch_blocks_per_cu must be in [1, 64]");
        c.batch_bpc = (int)value;
    } else if (k == "profile") {
        c.profile = value != 0;
    } else {
        return grasph::fail(c, GRASPH_E_INVALID, "unknown option '" + k + "'");
    }
    return 0;
}

int grasph_get_option(const grasph_ctx* ctx, const char* key, int64_t* value)
{
    if (!ctx || !key || !value) return GRASPH_E_INVALID;
    const Ctx& c = ctx->c;
    const std::string k(key);
    if (k == "max_neighbours") *value = c.max_neighbours;
    else if (k == "max_neighbours_grow") *value = c.max_neighbours_grow;
    else if (k == "reorder_interval") *value = c.reorder_interval;
    else if (k == "sweep_i_per_lane") *value = c.sweep_ni;
    else if (k == "sweep_slices") *value = c.sweep_slices;
    else if (k == "sweep_symmetric") *value = c.sweep_symmetric;
    else if (k == "sym_workgroups") *value = c.sym_workgroups;
    else if (k == "sym_partial_bytes") *value = c.sym_partial_bytes;
    else if (k == "batch_symmetric") *value = c.batch_symmetric;
    else if (k == "batch_symmetric_min") *value = c.batch_symmetric_min;
    else if (k == "sym_near_shares") *value = c.sym_near_shares;
    else if (k == "sym_list_priority") *value = c.sym_list_priority;
    else if (k == "batch_waves_per_block") *value = c.batch_wpb;
    else if (k == "batch_blocks_per_cu") *value = c.batch_bpc;
    else if (k == "profile") *value = c.profile ? 1 : 0;
    else if (k == "noalloc_slices") *value = c.noalloc_slices;
    else return GRASPH_E_INVALID;
    return 0;
}

int grasph_compute_derivatives(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm, void* stream)
{
    return ctx ? grasph::derivatives(ctx->c, p, prm, stream, 0) : GRASPH_E_INVALID;
}
int grasph_compute_derivatives_noalloc(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm,
                                       void* stream)
{
    return ctx ? grasph::derivatives_noalloc(ctx->c, p, prm, stream) : GRASPH_E_INVALID;
}
int grasph_calc_density(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm, void* stream)
{
    return ctx ? grasph::derivatives(ctx->c, p, prm, stream, 1) : GRASPH_E_INVALID;
}
int grasph_calc_balsara_density_dt_dstress_dt(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm,
                                              void* stream)
{
    return ctx ? grasph::derivatives(ctx->c, p, prm, stream, 2) : GRASPH_E_INVALID;
}
int grasph_calc_acceleration(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm, void* stream)
{
    return ctx ? grasph::derivatives(ctx->c, p, prm, stream, 3) : GRASPH_E_INVALID;
}

int grasph_integrate(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm, int not_first_step,
                     void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if ((r = grasph::check_params(c, prm)) || (r = grasph::check_particles(c, p))) return r;
    if (p->n == 0) return 0;
    if ((r = grasph::check_integrate_arrays(c, p, prm->flags))) return r;
    grasph::IntegLaunch L;
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
    L.n = (int)p->n;
    L.not_first_step = not_first_step;
    L.maxvsig = p->maxvsig;
    L.n_instances = 1;
    L.ts = (prm->flags & GRASPH_VARIABLE_TIMESTEP) ? c.ts : nullptr;
    hipStream_t s = (hipStream_t)stream;
    grasph::ProfScope ps(&c, s, grasph::KC_INTEGRATE);
    hipError_t e = grasph::launch_integrate_step(L, grasph::make_dev_params(prm), s);
    if (e != hipSuccess) return grasph::hip_fail(c, e, "leapfrog launch");
    return 0;
}

double grasph_get_current_timestep(const grasph_ctx* ctx, const grasph_params* prm)
{
    if (!prm) return 0.0;
    if (!(prm->flags & GRASPH_VARIABLE_TIMESTEP)) return (double)prm->fixed_timestep;  // integration.h:248-249
    if (!ctx || !ctx->c.ts) return 0.0;
    // integration.h:250-253: the device's current dt, after the queued work
    float dt = 0.0f;
    if (hipSetDevice(ctx->c.device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&dt, ctx->c.ts, sizeof dt, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return -1.0;
    }
    return (double)dt;
}

int grasph_reduce_invariants(grasph_ctx* ctx, const grasph_particles* p, const grasph_params* prm,
                             double half_kick_dt, int with_potential, double out[GRASPH_NINV], void* stream)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    int r;
    if ((r = grasph::check_params(c, prm)) || (r = grasph::check_particles(c, p))) return r;
    if (!out) return grasph::fail(c, GRASPH_E_INVALID, "out is NULL");
    for (int k = 0; k < GRASPH_NINV; ++k) out[k] = 0.0;
    if (p->n == 0) return 0;
    GRASPH_NEED(p->vel, "vel");
    if (half_kick_dt != 0.0) GRASPH_NEED(p->acc, "acc");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = p->n;
    if (with_potential) {
        if ((r = grasph::grow(c, c.pot, c.pot_cap, n, "potential"))) return r;
        grasph::SweepLaunch L = grasph::sweep_base(c, p);
        L.perm = nullptr;
        L.pot = c.pot;
        L.want_pot = true;
        if ((r = grasph::run_sweep(c, L, grasph::make_dev_params(prm), s))) return r;
    }
    const int nb = grasph::inv_blocks((int)n);
    if ((r = grasph::grow(c, c.red, c.red_cap, (int64_t)(nb + 1) * GRASPH_NINV, "reduction partials"))) return r;
    grasph::InvLaunch I;
    I.posm = (const float4*)p->posm;
    I.pos64 = p->pos64;
    I.vel = (const float4*)p->vel;
    I.acc = (const float4*)p->acc;
    I.density = p->density;
    I.dstress = p->dstress;
    I.pot = with_potential ? c.pot : nullptr;
    I.n = (int)n;
    I.half_kick = half_kick_dt;
    I.partials = c.red;
    I.nblocks = nb;
    {
        grasph::ProfScope ps(&c, s, grasph::KC_OTHER);
        hipError_t e = grasph::launch_invariants(I, *prm, s);
        if (e != hipSuccess) return grasph::hip_fail(c, e, "invariants launch");
    }
    hipError_t e = hipMemcpyAsync(out, c.red + (size_t)nb * GRASPH_NINV, GRASPH_NINV * sizeof(double),
                                  hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return grasph::hip_fail(c, e, "invariants copy");
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return grasph::hip_fail(c, e, "invariants synchronise");
    return 0;
}

int grasph_get_kernel_times(grasph_ctx* ctx, double ms[GRASPH_NKERNEL_CLASSES],
                            int64_t launches[GRASPH_NKERNEL_CLASSES], int reset)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    for (auto& ev : c.ev_pending) {
        hipError_t e = hipEventSynchronize(ev.b);
        if (e != hipSuccess) return grasph::hip_fail(c, e, "event synchronise");
        float t = 0.0f;
        e = hipEventElapsedTime(&t, ev.a, ev.b);
        if (e != hipSuccess) return grasph::hip_fail(c, e, "event elapsed time");
        c.ms[ev.cls] += t;
        c.launches[ev.cls] += 1;
        c.ev_free.push_back(ev.a);
        c.ev_free.push_back(ev.b);
    }
    c.ev_pending.clear();
    for (int k = 0; k < GRASPH_NKERNEL_CLASSES; ++k) {
        if (ms) ms[k] = c.ms[k];
        if (launches) launches[k] = c.launches[k];
        if (reset) {
            c.ms[k] = 0.0;
            c.launches[k] = 0;
        }
    }
    return 0;
}

int grasph_get_neighbour_stats(grasph_ctx* ctx, int64_t* pairs, int32_t* max_count, int64_t* overflowed)
{
    if (!ctx) return GRASPH_E_INVALID;
    Ctx& c = ctx->c;
    if (pairs) *pairs = 0;
    if (max_count) *max_count = 0;
    if (overflowed) *overflowed = 0;
    if (!c.nbr_cnt || c.last_list_n <= 0) return 0;
    const size_t n = (size_t)c.last_list_n, S = (size_t)c.last_list_slices;
    std::vector<int> cnt(n * S);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(cnt.data(), c.nbr_cnt, cnt.size() * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return grasph::hip_fail(c, e, "neighbour counts copy");
    int64_t tot = 0, over = 0;
    int32_t mx = 0;
    for (size_t t = 0; t < n; ++t) {
        int sum = 0;
        bool o = false;
        for (size_t s = 0; s < S; ++s) {
            const int v = cnt[s * n + t];
            sum += v;
            o = o || v > c.max_neighbours;
        }
        tot += sum;
        if (sum > mx) mx = sum;
        if (o) ++over;
    }
    if (pairs) *pairs = tot;
    if (max_count) *max_count = mx;
    if (overflowed) *overflowed = over;
    return 0;
}

}  // extern "C"
