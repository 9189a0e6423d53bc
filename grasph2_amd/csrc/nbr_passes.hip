// nbr_passes.hip -- the O(N k) neighbour-list passes (gfx950).
//
// density: calcDensity (src/computeDerivatives.h:52-91) over the list.
// strain:  calcBalsaraDensityDTDStressDT (:96-162); also writes the
//          per-particle record {sigma/rho^2, balsara, density} that the
//          acceleration pass reads for its j side, hoisting the per-pair
//          Murnaghan pow and nine divides of :246-257.
// prep:    the same record alone, for configurations that skip the strain
//          pass (:322-324).
// accel:   the SPH part of calcAcceleration (:237-295) plus
//          environmentAcceleration (:300-312), added to the gravity the
//          sweep left in acc with compensated (TwoSum) accumulation.
//
// One particle per lane in Morton order (slot t -> i = perm[t]).  Every
// input a pass reads for a neighbour sits in one 64-byte slot record,
// rec[j] = {x y z m, vx vy vz rho, sigma/rho^2 (6) balsara rho}, in slot
// order: one neighbour costs one contiguous gather, and a particle's
// neighbours' records are close together in memory.  Outputs go to
// particle i of the caller's arrays.  The list is stored in tiles of 64
// slots (list_entry), so a wavefront's loads of its k-th neighbours are
// contiguous, and holds slot indices in increasing order.  A particle whose
// count exceeds the list capacity falls back to a scan over its instance's
// slots with the exact neighbour test, so a full list never loses partners.
//
// Every kernel serves one instance (SlotMap from the kernel arguments) or a
// batch (SlotMap and parameters of the block's instance from tables).
#include "internal.hpp"

// No multiply-add contraction in this file: the pair arithmetic is inlined
// into the list walk and into the overflow scan, and contraction decided per
// call site could round the two differently.  Every product and sum rounds
// as written, as in the FP32 oracle; the passes are bound by their gathers,
// not by arithmetic.
#pragma clang fp contract(off)

namespace grasph {
namespace {

constexpr int kBlock = 256;

struct Tens {  // full 3x3, row-major, as mpu::Mat<float,3,3>
    float m[3][3];
};

__device__ inline Tens sym_to_full(float xx, float yy, float zz, float xy, float xz, float yz)
{
    Tens t;
    t.m[0][0] = xx; t.m[1][1] = yy; t.m[2][2] = zz;
    t.m[0][1] = xy; t.m[1][0] = xy;
    t.m[0][2] = xz; t.m[2][0] = xz;
    t.m[1][2] = yz; t.m[2][1] = yz;
    return t;
}

// src/sph/models.h:250-268, for the unique (d,e) of a symmetric result
__device__ inline float dsdt_entry(const Tens& ed, const Tens& rd, const Tens& S, float shear, int d, int e)
{
    float o = 2.0f * shear * ed.m[d][e];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        if (d == e) o -= 2.0f * shear * ed.m[f][f] / 3.0f;
        o += S.m[d][f] * rd.m[e][f];
        o += S.m[e][f] * rd.m[d][f];
    }
    return o;
}

struct Lists {
    const int* nbr;
    const int* cnt;
};

// Walk slot t's neighbours: the list, or when it overflowed, a scan of the
// instance's slots with the exact test (same partners, same order).
// load(j) fetches what body(j, data) needs; the walk is software-pipelined
// D deep: while partner e is computed, the data of partners e+1 .. e+D-1
// and the indices up to e+2D-1 are in flight (slot u of the rings holds
// the partners e with e mod D = u).  The look-ahead past the end reloads
// the last partner (always in range).  Deeper rings hide more of the gather
// latency where few wavefronts share a SIMD (one instance: 2048 wavefronts
// at 131072 particles), shallower ones keep the registers low where many
// do (batches).
template <int D, int SUB, typename L, typename F>
__device__ __forceinline__ void for_each_neighbour(const float4* __restrict__ xs, const Lists& Ls,
                                                   const size_t nstride, const int t, const int kmax,
                                                   const float4 pi, const SlotMap& m, const float H2, const int sub,
                                                   L&& load, F&& body)
{
    const int cnt = Ls.cnt[t];
    if (cnt <= kmax) {
        // this lane's share of the list: entries sub, sub + SUB, ...
        const int cl = cnt > sub ? (cnt - 1 - sub) / SUB + 1 : 0;
        if (cl <= 0) return;
        const int* list = Ls.nbr + list_entry(t, sub, kmax);
        constexpr size_t kStride = 64 * SUB;
        const int last = cl - 1;
        int jcur[D], jnx[D];
        decltype(load(0)) dd[D];
#pragma unroll
        for (int u = 0; u < D; ++u) jcur[u] = list[kStride * min(u, last)];
#pragma unroll
        for (int u = 0; u < D; ++u) jnx[u] = list[kStride * min(u + D, last)];
#pragma unroll
        for (int u = 0; u < D; ++u) dd[u] = load(jcur[u]);
        for (int k = 0; k < cl; k += D) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int e = k + u;
                if (e < cl) body(jcur[u], dd[u]);
                jcur[u] = jnx[u];
                dd[u] = load(jcur[u]);  // partner e + D (its index came D partners ago)
                jnx[u] = list[kStride * min(e + 2 * D, last)];
            }
        }
    } else {
        // the same share of the partners in slot order: the partner count
        // picks the lane, so the sums split as over the list
        int idx = 0;
        for (int j = m.off; j < m.off + m.n; ++j) {
            const float4 pj = xs[j];
            if (is_neighbour(dist2_exact(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z), H2)) {
                if (SUB == 1 || idx % SUB == sub) body(j, load(j));
                ++idx;
            }
        }
    }
}

// SUB lanes share a particle when its list is long (one instance, SUB = 4):
// lane l of a wavefront serves particle l mod (64 / SUB) of the wavefront's
// 64 / SUB and walks its neighbours sub, sub + SUB, ... (sub = l div (64 /
// SUB)); the lanes' sums are then added in a fixed order (the lane with
// sub = 0 adds sub 1's, then the sum of subs 2 and 3) and that lane writes.
template <int SUB>
__device__ __forceinline__ int pass_sub() { return (int)(threadIdx.x & 63) / (64 / SUB); }
template <int SUB>
__device__ __forceinline__ int pass_slot(const SlotMap& m)
{
    constexpr int kPPW = 64 / SUB;
    return m.slot0 + (int)(threadIdx.x >> 6) * kPPW + (int)(threadIdx.x & 63) % kPPW;
}
template <int SUB>
__device__ __forceinline__ void merge_subs(CSum& a)
{
#pragma unroll
    for (int o = 64 / SUB; o < 64; o <<= 1) {
        const float s = __shfl_xor(a.s, o, 64), c = __shfl_xor(a.c, o, 64);
        a.add(s);
        a.c += c;
    }
}
template <int SUB, typename T>
__device__ __forceinline__ void sum_subs(T& a)
{
#pragma unroll
    for (int o = 64 / SUB; o < 64; o <<= 1) a = a + __shfl_xor(a, o, 64);
}
template <int SUB>
__device__ __forceinline__ void max_subs(float& a)
{
#pragma unroll
    for (int o = 64 / SUB; o < 64; o <<= 1) a = fmaxf(a, __shfl_xor(a, o, 64));
}

struct PV {  // position/mass and velocity/density of a partner
    float4 pm, vr;
};

template <bool BATCH, int SUB>
__global__ __launch_bounds__(kBlock) void density_kernel(const float4* __restrict__ xs, SlotRec* __restrict__ rec,
                                                         const int* __restrict__ perm,
                                                         const Lists Ls,
                                                         const size_t nstride, const int kmax,
                                                         float* __restrict__ density, const BlockTables T,
                                                         const DevParams P0)
{
    // batches, and the long lists of bodies in contact (SUB > 1): issue
    // ahead of the far field's wavefronts that run beside (ensemble of 512:
    // 16.25-16.30 -> 15.98-16.16 ms per step; collision N = 131072 after
    // impact: 5.59-5.62 -> 5.03-5.23)
    if (BATCH || SUB > 1) __builtin_amdgcn_s_setprio(3);
    const SlotMap m = slot_map<BATCH>(T, (int)nstride, kBlock / SUB);
    const int t = pass_slot<SUB>(m), sub = pass_sub<SUB>();
    const bool act = t < m.slot_end;
    if (SUB == 1 && !act) return;
    const int tc = act ? t : m.slot0;  // lanes past the end take part in the lanes' sums only
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    const int i = perm ? perm[tc] : tc;
    const float4 pi = xs[tc];  // positions from the compact slot-ordered array (16 B per partner)
    CSum rho{0.0f, 0.0f};
    if (act)
        for_each_neighbour<SUB == 1 && !BATCH ? 4 : 2, SUB>(xs, Ls, nstride, t, kmax, pi, m, P.H2, sub,
                                                           [&](int j) { return xs[j]; }, [&](int, const float4 pj) {
            rho.add(pj.w * kernel_w_f64(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, P));
        });
    if (SUB > 1) {
        merge_subs<SUB>(rho);
        if (!act || sub != 0) return;
    }
    float r = rho.value();
    if ((P.flags & GRASPH_DEAL_WITH_NO_PARTNERS) && r < 0.001f) r = P.rho0 * 0.01f;  // :85-88
    density[i] = r;
    rec[t].vr.w = r;  // the strain / record passes read densities from the slot records
}

template <bool BATCH, int SUB, bool SOLIDS, bool BALS, bool INTD>
__global__ __launch_bounds__(kBlock) void strain_kernel(
    const float4* __restrict__ xs, SlotRec* __restrict__ rec, const float* __restrict__ dstress,
    const int* __restrict__ perm, const Lists Ls, const size_t nstride, const int kmax, float* __restrict__ balsara,
    float* __restrict__ density_dt, float* __restrict__ dstress_dt, const BlockTables T, const DevParams P0)
{
    // batches, and the long lists of bodies in contact (SUB > 1): issue
    // ahead of the far field's wavefronts that run beside (ensemble of 512:
    // 16.25-16.30 -> 15.98-16.16 ms per step; collision N = 131072 after
    // impact: 5.59-5.62 -> 5.03-5.23)
    if (BATCH || SUB > 1) __builtin_amdgcn_s_setprio(3);
    const SlotMap m = slot_map<BATCH>(T, (int)nstride, kBlock / SUB);
    const int t0 = pass_slot<SUB>(m), sub = pass_sub<SUB>();
    const bool act = t0 < m.slot_end;
    if (SUB == 1 && !act) return;
    const int t = act ? t0 : m.slot0;  // lanes past the end take part in the lanes' sums only
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    const int i = perm ? perm[t] : t;
    const float4 pi = rec[t].pm;
    const float4 vi = rec[t].vr;
    const float rhoi = vi.w;

    // Sums and terms in FP64 from the FP32 inputs: the velocity divergence
    // and curl cancel strongly where bodies meet, and the Balsara factor
    // built from them scales the artificial viscosity of the acceleration
    // pass, where FP32 rounding here showed up amplified.
    double exx = 0, eyy = 0, ezz = 0, exy = 0, exz = 0, eyz = 0;  // strain rate (symmetric)
    double rxy = 0, rxz = 0, ryz = 0;                              // rotation rate (antisymmetric)
    double divv = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;

    if (act)
    for_each_neighbour<SUB == 1 && !BATCH ? 4 : 2, SUB>(xs, Ls, nstride, t, kmax, pi, m, P.H2, sub,
                                                       [&](int j) { return PV{rec[j].pm, rec[j].vr}; },
                                                       [&](int, const PV& q) {
        const float4 pj = q.pm;
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const double s = dw_over_r_f64(rx, ry, rz, P);
        const double gx = s * (double)rx, gy = s * (double)ry, gz = s * (double)rz;
        const float4 vj = q.vr;
        const double vx = (double)vi.x - (double)vj.x, vy = (double)vi.y - (double)vj.y,
                     vz = (double)vi.z - (double)vj.z;
        const double vol = (double)pj.w / (double)vj.w;  // one division per pair, shared by the terms below
        if (SOLIDS) {  // models.h:225-248
            const double tmp = -0.5 * vol;
            exx += tmp * (vx * gx + vx * gx);
            exy += tmp * (vx * gy + vy * gx);
            exz += tmp * (vx * gz + vz * gx);
            eyy += tmp * (vy * gy + vy * gy);
            eyz += tmp * (vy * gz + vz * gy);
            ezz += tmp * (vz * gz + vz * gz);
            rxy += tmp * (vx * gy - vy * gx);
            rxz += tmp * (vx * gz - vz * gx);
            ryz += tmp * (vy * gz - vz * gy);
        } else if (BALS) {  // curl v, computeDerivatives.h:134
            const double f = vol;
            cx += f * (vy * gz - vz * gy);
            cy += f * (vz * gx - vx * gz);
            cz += f * (vx * gy - vy * gx);
        }
        if (BALS || INTD) divv -= vol * (vx * gx + vy * gy + vz * gz);  // :137
    });
    if (SUB > 1) {
        sum_subs<SUB>(exx); sum_subs<SUB>(eyy); sum_subs<SUB>(ezz);
        sum_subs<SUB>(exy); sum_subs<SUB>(exz); sum_subs<SUB>(eyz);
        sum_subs<SUB>(rxy); sum_subs<SUB>(rxz); sum_subs<SUB>(ryz);
        sum_subs<SUB>(divv); sum_subs<SUB>(cx); sum_subs<SUB>(cy); sum_subs<SUB>(cz);
        if (!act || sub != 0) return;
    }

    // do_after, computeDerivatives.h:144-161
    Tens S = sym_to_full(0.f, 0.f, 0.f, 0.f, 0.f, 0.f);
    if (SOLIDS)
        S = sym_to_full(dstress[m.plane6(0, i)], dstress[m.plane6(1, i)], dstress[m.plane6(2, i)],
                        dstress[m.plane6(3, i)], dstress[m.plane6(4, i)], dstress[m.plane6(5, i)]);
    float sdt[6] = {0, 0, 0, 0, 0, 0};
    if (SOLIDS) {
        const Tens ed = sym_to_full((float)exx, (float)eyy, (float)ezz, (float)exy, (float)exz, (float)eyz);
        Tens rd;
        rd.m[0][0] = 0.0f; rd.m[1][1] = 0.0f; rd.m[2][2] = 0.0f;
        rd.m[0][1] = (float)rxy; rd.m[1][0] = -(float)rxy;
        rd.m[0][2] = (float)rxz; rd.m[2][0] = -(float)rxz;
        rd.m[1][2] = (float)ryz; rd.m[2][1] = -(float)ryz;
        sdt[0] = dsdt_entry(ed, rd, S, P.shear, 0, 0);
        sdt[1] = dsdt_entry(ed, rd, S, P.shear, 1, 1);
        sdt[2] = dsdt_entry(ed, rd, S, P.shear, 2, 2);
        sdt[3] = dsdt_entry(ed, rd, S, P.shear, 0, 1);
        sdt[4] = dsdt_entry(ed, rd, S, P.shear, 0, 2);
        sdt[5] = dsdt_entry(ed, rd, S, P.shear, 1, 2);
        // curl from the rotation rate, :150
        cx = -2.0 * ryz;
        cy = 2.0 * rxz;
        cz = -2.0 * rxy;
    }
    float bal = 0.0f;
    if (BALS) bal = balsara_switch_f64(divv, cx, cy, cz, P.c, P.H);
    if (balsara) balsara[i] = bal;
    if (density_dt) density_dt[i] = INTD ? (float)(-(double)rhoi * divv) : 0.0f;
    if (dstress_dt) {
#pragma unroll
        for (int c = 0; c < 6; ++c) dstress_dt[m.plane6(c, i)] = sdt[c];
    }

    // j-side record of the acceleration pass (its do_before, :193-219)
    if (SOLIDS) {
        const float pres = eos_murnaghan(rhoi, P.rho0, P.BULK, P.dBULKdP);
        const float d2 = rhoi * rhoi;
        rec[t].ra = make_float4((S.m[0][0] - pres) / d2, (S.m[1][1] - pres) / d2, (S.m[2][2] - pres) / d2,
                                S.m[0][1] / d2);
        rec[t].rb = make_float4(S.m[0][2] / d2, S.m[1][2] / d2, bal, rhoi);
    } else {
        const float sig = -eos_liquid(rhoi, P.rho0, P.c2);
        rec[t].ra = make_float4(sig / (rhoi * rhoi), 0.0f, 0.0f, 0.0f);
        rec[t].rb = make_float4(0.0f, 0.0f, bal, rhoi);
    }
}

// record only (no strain pass in this configuration)
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void prep_kernel(SlotRec* __restrict__ rec, const float* __restrict__ dstress,
                                                      const float* __restrict__ balsara, const int* __restrict__ perm,
                                                      const size_t nstride, const BlockTables T, const DevParams P0)
{
    const SlotMap m = slot_map<BATCH>(T, (int)nstride);
    const int t = m.slot0 + threadIdx.x;
    if (t >= m.slot_end) return;
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    const int i = perm ? perm[t] : t;
    const float rhoi = rec[t].vr.w;
    const float bal = balsara ? balsara[i] : 0.0f;
    if (P.flags & GRASPH_SOLIDS) {
        const float pres = eos_murnaghan(rhoi, P.rho0, P.BULK, P.dBULKdP);
        const float d2 = rhoi * rhoi;
        rec[t].ra = make_float4((dstress[m.plane6(0, i)] - pres) / d2, (dstress[m.plane6(1, i)] - pres) / d2,
                                (dstress[m.plane6(2, i)] - pres) / d2, dstress[m.plane6(3, i)] / d2);
        rec[t].rb = make_float4(dstress[m.plane6(4, i)] / d2, dstress[m.plane6(5, i)] / d2, bal, rhoi);
    } else {
        const float sig = -eos_liquid(rhoi, P.rho0, P.c2);
        rec[t].ra = make_float4(sig / (rhoi * rhoi), 0.0f, 0.0f, 0.0f);
        rec[t].rb = make_float4(0.0f, 0.0f, bal, rhoi);
    }
}

template <bool BATCH, int SUB, bool SOLIDS, bool BALS, bool AV, bool XSPH, bool ARTS>
__global__ __launch_bounds__(kBlock) void accel_kernel(
    const float4* __restrict__ xs, const SlotRec* __restrict__ rec, const int* __restrict__ perm, const Lists Ls,
    const size_t nstride,
    const int kmax, float4* __restrict__ acc, float4* __restrict__ xvel, float* __restrict__ maxvsig,
    const int acc_has_gravity, const BlockTables T, const DevParams P0)
{
    // batches, and the long lists of bodies in contact (SUB > 1): issue
    // ahead of the far field's wavefronts that run beside (ensemble of 512:
    // 16.25-16.30 -> 15.98-16.16 ms per step; collision N = 131072 after
    // impact: 5.59-5.62 -> 5.03-5.23)
    if (BATCH || SUB > 1) __builtin_amdgcn_s_setprio(3);
    const SlotMap m = slot_map<BATCH>(T, (int)nstride, kBlock / SUB);
    const int t0 = pass_slot<SUB>(m), sub = pass_sub<SUB>();
    const bool act = t0 < m.slot_end;
    if (SUB == 1 && !act) return;
    const int t = act ? t0 : m.slot0;  // lanes past the end take part in the lanes' sums only
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    const int i = perm ? perm[t] : t;
    const float4 pi = rec[t].pm;
    const float4 vi = rec[t].vr;
    const float4 ra = rec[t].ra, rb = rec[t].rb;
    const float rhoi = rb.w, bi = BALS ? rb.z : 1.0f;
    // sigma_i / rho_i^2 and, for liquids with artificial pressure, arts_i
    const float sxx = ra.x, syy = ra.y, szz = ra.z, sxy = ra.w, sxz = rb.x, syz = rb.y;
    const float arts_i = (ARTS && !SOLIDS) ? ((sxx > 0) ? (-P.mateps * sxx) : 0.0f) : 0.0f;

    CSum ax{0.0f, 0.0f}, ay{0.0f, 0.0f}, az{0.0f, 0.0f};
    if (acc_has_gravity && sub == 0) {
        const float4 g = acc[i];
        ax.s = g.x; ay.s = g.y; az.s = g.z;
    }
    float mv = 0.0f;
    float xvx = 0.0f, xvy = 0.0f, xvz = 0.0f;

    if (act)
    for_each_neighbour<SUB == 1 && !BATCH ? 4 : 2, SUB>(xs, Ls, nstride, t, kmax, pi, m, P.H2, sub,
                                                       [&](int j) { return rec[j]; },
                                                       [&](int, const SlotRec& q) {
        const float4 pj = q.pm;
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float r2 = dist2_exact(rx, ry, rz);
        const float r = sqrtf(r2);
        float gx, gy, gz;  // dW/r r_ij: FP64, rounded once (the pair terms cancel where the stresses are large)
        kernel_gradient_f64(rx, ry, rz, P, gx, gy, gz);
        const float4 qa = q.ra, qb = q.rb;
        const float rhoj = qb.w;
        const float mj = pj.w;
        float fx, fy, fz;
        if (SOLIDS) {  // acc += m_j (stress . gradw), stress = sOR_i + sOR_j  (:260, :275)
            const float txx = sxx + qa.x, tyy = syy + qa.y, tzz = szz + qa.z;
            const float txy = sxy + qa.w, txz = sxz + qb.x, tyz = syz + qb.y;
            fx = mj * (txx * gx + txy * gy + txz * gz);
            fy = mj * (txy * gx + tyy * gy + tyz * gz);
            fz = mj * (txz * gx + tyz * gy + tzz * gz);
        } else {
            float st = sxx + qa.x;
            if (ARTS) {  // :263-272 (liquid artificial pressure)
                const float f = powf(w_spline(r, P.H, P.Wpref) / P.Wnorm, P.matexp);
                const float arts_j = (qa.x > 0) ? (-P.mateps * qa.x) : 0.0f;
                st += f * (arts_i + arts_j);
            }
            fx = mj * (st * gx);
            fy = mj * (st * gy);
            fz = mj * (st * gz);
        }
        const float4 vj = q.vr;
        const float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
        if (AV) {  // :277-289
            const float II = artificial_viscosity(BALS ? &mv : nullptr, P.alpha, rhoi, rhoj, vx, vy, vz, rx, ry, rz, r,
                                                  P.c, P.c, bi, BALS ? qb.z : 1.0f);
            const float mII = mj * II;
            fx -= mII * gx;
            fy -= mII * gy;
            fz -= mII * gz;
        }
        ax.add(fx);
        ay.add(fy);
        az.add(fz);
        if (XSPH) {  // :291-294
            const float w = w_spline(r, P.H, P.Wpref);
            const float f = 2.0f * mj / (rhoi + rhoj);
            xvx += f * (vj.x - vi.x) * w;
            xvy += f * (vj.y - vi.y) * w;
            xvz += f * (vj.z - vi.z) * w;
        }
    });
    if (SUB > 1) {
        merge_subs<SUB>(ax); merge_subs<SUB>(ay); merge_subs<SUB>(az);
        if (BALS && AV) max_subs<SUB>(mv);
        if (XSPH) { sum_subs<SUB>(xvx); sum_subs<SUB>(xvy); sum_subs<SUB>(xvz); }
        if (!act || sub != 0) return;
    }

    float fx = ax.value(), fy = ay.value(), fz = az.value();
    environment_acceleration(P, fx, fy, fz);
    acc[i] = make_float4(fx, fy, fz, 0.0f);
    if (maxvsig) maxvsig[i] = mv;
    if (xvel) xvel[i] = make_float4(xvx, xvy, xvz, 0.0f);
}

// Without SPH: environmentAcceleration on the gravity and the zero XVEL /
// MAXVSIG of calcAcceleration's store (computeDerivatives.h:171, 300-312).
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void finish_kernel(const float4* __restrict__ posm,
                                                        const float4* __restrict__ vel, float4* __restrict__ acc,
                                                        float4* __restrict__ xvel, float* __restrict__ maxvsig,
                                                        const int n, const BlockTables T, const DevParams P0)
{
    const SlotMap m = slot_map<BATCH>(T, n);
    const int i = m.slot0 + threadIdx.x;
    if (i >= m.slot_end) return;
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    const bool grav = P.flags & GRASPH_SELF_GRAVITY;
    if (!grav || (P.flags & GRASPH_CONSTANT_ACCELERATION)) {
        float4 a = grav ? acc[i] : make_float4(0.f, 0.f, 0.f, 0.f);  // no sweep ran without gravity
        environment_acceleration(P, a.x, a.y, a.z);
        acc[i] = make_float4(a.x, a.y, a.z, 0.0f);
    }
    if (xvel) xvel[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (maxvsig) maxvsig[i] = 0.0f;
}

inline Lists lists_of(const NbrLaunch& L) { return Lists{L.nbr, L.nbr_cnt}; }

inline int sub_of(const NbrLaunch& L) { return L.T.nblocks > 0 || L.sub <= 1 ? 1 : 4; }
inline dim3 grid_plain(const NbrLaunch& L)  // per-particle kernels without a list walk
{
    return dim3(L.T.nblocks > 0 ? L.T.nblocks : (L.n + kBlock - 1) / kBlock);
}
inline dim3 grid_of(const NbrLaunch& L)
{
    const int per = kBlock / sub_of(L);  // particles per block
    return dim3(L.T.nblocks > 0 ? L.T.nblocks : (L.n + per - 1) / per);
}

template <bool BATCH, int SUB>
hipError_t strain_b(const NbrLaunch& L, const DevParams& P, hipStream_t s)
{
    const bool so = P.flags & GRASPH_SOLIDS, ba = P.flags & GRASPH_BALSARA_SWITCH,
               in = P.flags & GRASPH_INTEGRATE_DENSITY;
#define GRASPH_STRAIN(A, B, C)                                                                                     \
    hipLaunchKernelGGL((strain_kernel<BATCH, SUB, A, B, C>), grid_of(L), dim3(kBlock), 0, s, L.xs, L.rec, L.dstress, \
                       L.perm, lists_of(L), (size_t)L.n, L.kmax, L.balsara, L.density_dt, L.dstress_dt, L.T, P)
    const int key = (so ? 4 : 0) | (ba ? 2 : 0) | (in ? 1 : 0);
    switch (key) {
        case 0: GRASPH_STRAIN(false, false, false); break;
        case 1: GRASPH_STRAIN(false, false, true); break;
        case 2: GRASPH_STRAIN(false, true, false); break;
        case 3: GRASPH_STRAIN(false, true, true); break;
        case 4: GRASPH_STRAIN(true, false, false); break;
        case 5: GRASPH_STRAIN(true, false, true); break;
        case 6: GRASPH_STRAIN(true, true, false); break;
        default: GRASPH_STRAIN(true, true, true); break;
    }
#undef GRASPH_STRAIN
    return hipGetLastError();
}

template <bool BATCH, int SUB>
hipError_t accel_b(const NbrLaunch& L, const DevParams& P, hipStream_t s)
{
    const bool so = P.flags & GRASPH_SOLIDS, ba = P.flags & GRASPH_BALSARA_SWITCH,
               av = P.flags & GRASPH_ARTIFICIAL_VISCOSITY, xs = P.flags & GRASPH_XSPH,
               ar = P.flags & GRASPH_ARTIFICIAL_STRESS;
#define GRASPH_ACCEL(A, B, C, D, E)                                                                                  \
    hipLaunchKernelGGL((accel_kernel<BATCH, SUB, A, B, C, D, E>), grid_of(L), dim3(kBlock), 0, s, L.xs, L.rec,    \
                       L.perm,                                                                                     \
                       lists_of(L), (size_t)L.n, L.kmax, L.acc, L.xvel, L.maxvsig, L.acc_has_gravity ? 1 : 0,  \
                       L.T, P)
    // SOLIDS + ARTIFICIAL_STRESS is rejected before launch (the reference does not compile it).
    const int key = (so ? 16 : 0) | (ba ? 8 : 0) | (av ? 4 : 0) | (xs ? 2 : 0) | (ar ? 1 : 0);
    switch (key) {
        case 0: GRASPH_ACCEL(false, false, false, false, false); break;
        case 1: GRASPH_ACCEL(false, false, false, false, true); break;
        case 2: GRASPH_ACCEL(false, false, false, true, false); break;
        case 3: GRASPH_ACCEL(false, false, false, true, true); break;
        case 4: GRASPH_ACCEL(false, false, true, false, false); break;
        case 5: GRASPH_ACCEL(false, false, true, false, true); break;
        case 6: GRASPH_ACCEL(false, false, true, true, false); break;
        case 7: GRASPH_ACCEL(false, false, true, true, true); break;
        case 8: GRASPH_ACCEL(false, true, false, false, false); break;
        case 9: GRASPH_ACCEL(false, true, false, false, true); break;
        case 10: GRASPH_ACCEL(false, true, false, true, false); break;
        case 11: GRASPH_ACCEL(false, true, false, true, true); break;
        case 12: GRASPH_ACCEL(false, true, true, false, false); break;
        case 13: GRASPH_ACCEL(false, true, true, false, true); break;
        case 14: GRASPH_ACCEL(false, true, true, true, false); break;
        case 15: GRASPH_ACCEL(false, true, true, true, true); break;
        case 16: GRASPH_ACCEL(true, false, false, false, false); break;
        case 18: GRASPH_ACCEL(true, false, false, true, false); break;
        case 20: GRASPH_ACCEL(true, false, true, false, false); break;
        case 22: GRASPH_ACCEL(true, false, true, true, false); break;
        case 24: GRASPH_ACCEL(true, true, false, false, false); break;
        case 26: GRASPH_ACCEL(true, true, false, true, false); break;
        case 28: GRASPH_ACCEL(true, true, true, false, false); break;
        case 30: GRASPH_ACCEL(true, true, true, true, false); break;
        default: return hipErrorInvalidValue;
    }
#undef GRASPH_ACCEL
    return hipGetLastError();
}

}  // namespace

hipError_t launch_density(const NbrLaunch& L, float* density, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (L.T.nblocks > 0)
        hipLaunchKernelGGL((density_kernel<true, 1>), grid_of(L), dim3(kBlock), 0, s, L.xs, L.rec, L.perm, lists_of(L),
                           (size_t)L.n, L.kmax, density, L.T, P);
    else if (sub_of(L) == 4)
        hipLaunchKernelGGL((density_kernel<false, 4>), grid_of(L), dim3(kBlock), 0, s, L.xs, L.rec, L.perm,
                           lists_of(L), (size_t)L.n, L.kmax, density, L.T, P);
    else
        hipLaunchKernelGGL((density_kernel<false, 1>), grid_of(L), dim3(kBlock), 0, s, L.xs, L.rec, L.perm,
                           lists_of(L), (size_t)L.n, L.kmax, density, L.T, P);
    return hipGetLastError();
}

hipError_t launch_strain(const NbrLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (L.T.nblocks > 0) return strain_b<true, 1>(L, P, s);
    return sub_of(L) == 4 ? strain_b<false, 4>(L, P, s) : strain_b<false, 1>(L, P, s);
}

hipError_t launch_prep(const NbrLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (L.T.nblocks > 0)
        hipLaunchKernelGGL(prep_kernel<true>, grid_plain(L), dim3(kBlock), 0, s, L.rec, L.dstress, L.balsara_in, L.perm,
                           (size_t)L.n, L.T, P);
    else
        hipLaunchKernelGGL(prep_kernel<false>, grid_plain(L), dim3(kBlock), 0, s, L.rec, L.dstress, L.balsara_in,
                           L.perm, (size_t)L.n, L.T, P);
    return hipGetLastError();
}

hipError_t launch_accel(const NbrLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (L.T.nblocks > 0) return accel_b<true, 1>(L, P, s);
    return sub_of(L) == 4 ? accel_b<false, 4>(L, P, s) : accel_b<false, 1>(L, P, s);
}

hipError_t launch_finish(const NbrLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (L.T.nblocks > 0)
        hipLaunchKernelGGL(finish_kernel<true>, grid_plain(L), dim3(kBlock), 0, s, L.posm, L.vel, L.acc, L.xvel,
                           L.maxvsig, L.n, L.T, P);
    else
        hipLaunchKernelGGL(finish_kernel<false>, grid_plain(L), dim3(kBlock), 0, s, L.posm, L.vel, L.acc, L.xvel,
                           L.maxvsig, L.n, L.T, P);
    return hipGetLastError();
}

}  // namespace grasph
