/*
 * grasph_oracle.cpp -- CPU oracle for the GraSPH2 derivative + leapfrog path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the reference's
 * host all-pairs semantics (src/particles/algorithms.h:233-256: a fresh job per
 * particle i, attributes not in load_type value-initialised to zero, a
 * sequential j = 0..N-1 loop, then do_after and a store of store_type).  It is
 * the checker of the HIP path and the timed CPU baseline; nothing in the
 * product links or calls it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.
 *
 * Precision: every function exists for T = float (the reference's default
 * SINGLE_PRECISION build, settings/precisionSettings.h:16) and T = double
 * (its DOUBLE_PRECISION build, src/types.h:52-63), which is the FP64 oracle.
 * Built with -ffp-contract=off so every operation rounds exactly as written.
 *
 * Data layout is the C ABI's (include/grasph2.h) in precision T:
 *   posm T[n][4] (x,y,z,m); vel/acc/xvel T[n][4]; density, density_dt,
 *   balsara, maxvsig T[n]; dstress, dstress_dt T[6][n] planes xx|yy|zz|xy|xz|yz.
 * Internally tensors are full 3x3 row-major matrices as in the reference
 * (mpu::Mat<T,3,3>); a symmetric plane set expands to the full matrix on load
 * and the upper triangle is stored.
 *
 * Pinning (what checks each function; see tests/test_oracle_pins.py):
 *   kernel W, dW, prefactors     closed forms: normalisation integral,
 *                                continuity at q=1/2 and q=1, dW = dW/dr
 *   eos liquid / murnaghan       closed forms, special values
 *   artificialViscosity          receding pair -> 0; head-on pair by hand
 *   balsaraSwitch                pure compression -> ~1, pure rotation -> ~0
 *   plasticity (von Mises)       hand-computed J2 / scaling example
 *   calcDensity                  lattice sum + self term ~ rho0 (W normalised,
 *                                self term excluded); isolated pair by hand
 *   calcBalsara...DStressDT      linear velocity field v = L x on a lattice:
 *                                edot ~ sym(L), rotation ~ skew(L) (transpose
 *                                and sign pinned by an asymmetric L); rigid
 *                                rotation: dS/dt = W S - S W (co-rotation)
 *   calcAcceleration gravity     two-body closed form; force = -grad of the
 *                                pair potential (finite differences);
 *                                sum m a = 0
 *   calcAcceleration SPH         uniform-pressure interior -> ~0; linear
 *                                pressure field -> a ~ -grad p / rho; sum m a = 0
 *   artificial pressure          a pair by hand: f = (W(r)/W(normalsep))^matexp,
 *                                the tension-only gate, f = 1 at normalsep
 *   XSPH                         a pair by hand, 2 m_j/(rho_i+rho_j)(v_j-v_i)W;
 *                                uniform flow -> 0; sum m xvel = 0; the drift
 *   density rate                 v = L x lattice: rho_dot = -rho tr L
 *   Balsara (liquid curl)        v = L x lattice: |div| / (|div|+|curl|+eps),
 *                                rigid rotation -> 0, expansion -> 1
 *   fixedTsLeapfrog              circular two-body orbit: energy and angular
 *                                momentum bounded over many periods, period
 *                                matches Kepler; half-kick first step
 *   Mohr-Coulomb yield           leapfrog by hand: Y = tan(phi) p + c in
 *                                compression, at rest density and in tension
 *   variable step                criteria by hand; the acceleration pass's
 *                                reset of the next dt; two-body main loop
 *                                dt = min(max_timestep, criteria); constant
 *                                acceleration integrated exactly
 *   invariants / potential       two-body closed forms
 * Parity unpinned (no reference test, fixture or buildable reference covers
 * it, and the mathematics fixes only its form): the mpUtils operator
 * semantics (component-wise float3 ops, Mat(T) diagonal, Mat*vec by rows;
 * immaterial for the symmetric tensors used here).  CLOHESSY_WILTSHIRE is
 * rejected (validate() returns -3): src/computeDerivatives.h:37-39 adds a
 * scalar to an f3_t through an mpUtils operator no available source defines.
 */
#include "../include/grasph2.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

/* particle arrays in the C-ABI layout, precision given by the entry point suffix */
extern "C" {

typedef struct oracle_particles_t {
    int64_t n;
    void *posm, *vel, *acc, *xvel, *density, *density_dt, *balsara, *maxvsig, *dstress, *dstress_dt;
} oracle_particles_t;

}

namespace {

template <typename T> struct V3 { T x, y, z; };
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline V3<T>& operator+=(V3<T>& a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <typename T> inline V3<T>& operator-=(V3<T>& a, V3<T> b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }

template <typename T> struct M3 {
    T m[3][3];
};
template <typename T> inline M3<T> mzero() { M3<T> r; std::memset(&r, 0, sizeof r); return r; }

/* ---- settings-derived constants (src/computeDerivatives.h:26-28) -------- */
template <typename T> struct Consts {
    T H, H2, Wpref, dWpref, rho0, BULK, dBULKdP, c, shear, Y, tanfr, cohesion, alpha;
    T mateps, matexp, normalsep, xsph, dt;
    V3<T> const_acc, bmin, bmax;
    T refl;
    T ts_init, ts_min, ts_max, acc_acc, vel_acc;  // settings.h:56-62
    int dim;
    uint32_t f;
    bool has(uint32_t b) const { return (f & b) != 0; }
};

/* kernel.h:41-60 splinePrefactor, :68-87 dsplinePrefactor */
template <typename T> T splinePrefactor(int dim, T h)
{
    if (dim == 1) return T(4.0) / (T(3.0) * h);
    if (dim == 2) return T(40.0) / (T(7.0) * T(M_PI) * h * h);
    return T(8.0) / (T(M_PI) * h * h * h);
}
template <typename T> T dsplinePrefactor(int dim, T h)
{
    if (dim == 1) return T(24.0) / (T(3.0) * h * h);
    if (dim == 2) return T(240.0) / (T(7.0) * T(M_PI) * h * h * h);
    return T(48.0) / (T(M_PI) * h * h * h * h);
}

template <typename T> Consts<T> make_consts(const grasph_params* p)
{
    Consts<T> c;
    c.f = p->flags;
    c.dim = p->dimension;
    c.H = T(p->H);
    c.H2 = c.H * c.H;
    c.Wpref = splinePrefactor<T>(c.dim, c.H);
    c.dWpref = dsplinePrefactor<T>(c.dim, c.H);
    c.rho0 = T(p->rho0); c.BULK = T(p->BULK); c.dBULKdP = T(p->dBULKdP);
    c.c = T(p->SOUNDSPEED); c.shear = T(p->shear); c.Y = T(p->Y);
    c.tanfr = T(p->tanfr); c.cohesion = T(p->cohesion); c.alpha = T(p->alpha);
    c.mateps = T(p->mateps); c.matexp = T(p->matexp); c.normalsep = T(p->normalsep);
    c.xsph = T(p->xsph_factor); c.dt = T(p->fixed_timestep);
    c.const_acc = {T(p->constant_acceleration[0]), T(p->constant_acceleration[1]), T(p->constant_acceleration[2])};
    c.bmin = {T(p->box_min[0]), T(p->box_min[1]), T(p->box_min[2])};
    c.bmax = {T(p->box_max[0]), T(p->box_max[1]), T(p->box_max[2])};
    c.refl = T(p->box_reflectiveness);
    c.ts_init = T(p->initial_timestep); c.ts_min = T(p->min_timestep); c.ts_max = T(p->max_timestep);
    c.acc_acc = T(p->accel_accuracy); c.vel_acc = T(p->velocity_accuracy);
    return c;
}

/* ---- src/sph/kernel.h:130-136 Wspline, :169-175 dWspline ---------------- */
template <typename T> inline T Wspline(T r, T h, T factor)
{
    T q = r / h;
    return factor * ((q <= T(1.0)) ? ((q < T(0.5)) ? (T(6.0) * q * q * q - T(6.0) * q * q + T(1.0))
                                                   : T(2.0) * (T(1.0) - q) * (T(1.0) - q) * (T(1.0) - q))
                                   : T(0.0));
}
template <typename T> inline T dWspline(T r, T h, T factor)
{
    T q = r / h;
    return factor * ((q <= T(1.0)) ? ((q < T(0.5)) ? (T(3.0) * q * q - T(2.0) * q)
                                                   : T(-1.0) * (T(1.0) - q) * (T(1.0) - q))
                                   : T(0.0));
}

/* ---- src/sph/eos.h:88-96 ------------------------------------------------ */
template <typename T> inline T eos_liquid(T rho, T rho0, T c2) { return c2 * (rho - rho0); }
template <typename T> inline T eos_murnaghan(T rho, T rho0, T K0, T dK0)
{
    return K0 / dK0 * (std::pow(rho / rho0, dK0) - T(1.0));
}

/* ---- src/sph/models.h:158-185 artificialViscosity ----------------------- */
template <typename T>
inline T artificialViscosity(T* maxVsig, T alpha, T di, T dj, V3<T> vij, V3<T> rij, T r, T ci, T cj, T bi, T bj)
{
    const T wij = dot(rij, vij) / r;
    T II = T(0.0);
    if (wij < T(0.0)) {
        const T vsig = T(ci + cj - T(3.0) * wij);
        const T rhoij = (di + dj) * T(0.5);
        II = T(-0.25) * (bi + bj) * alpha * wij * vsig / rhoij;
        if (maxVsig) *maxVsig = std::max(*maxVsig, vsig);
    }
    return II;
}

/* models.h:187-192 */
template <typename T> inline T balsaraSwitch(T divv, V3<T> curlv, T c, T h)
{
    const T absdiv = std::fabs(divv);
    const T abscurl = length(curlv);
    return absdiv / (absdiv + abscurl + T(0.0001) * c / h);
}

/* models.h:194-205 (von Mises: factor Y^2/(3 J2), not its square root) */
template <typename T> inline void plasticity(M3<T>& S, T Y)
{
    T J2 = T(0.0);
    for (int e = 0; e < 9; ++e) J2 += S.m[e / 3][e % 3] * S.m[e / 3][e % 3];
    J2 *= T(0.5);
    const T f = Y * Y / (T(3.0) * J2);
    if (f < T(1))
        for (int e = 0; e < 9; ++e) S.m[e / 3][e % 3] *= f;
}

/* models.h:207-223 */
template <typename T> inline T mohrCoulombYieldStress(T tanfa, T pressure, T coh) { return tanfa * pressure + coh; }
template <typename T> inline M3<T> artificialStress(T mateps, const M3<T>& s)
{
    M3<T> a;
    for (int e = 0; e < 9; ++e) {
        T v = s.m[e / 3][e % 3];
        a.m[e / 3][e % 3] = (v > 0) ? (-mateps * v) : T(0.0);
    }
    return a;
}
template <typename T> inline T artificialPressure(T mateps, T s) { return (s > 0) ? (-mateps * s) : T(0.0); }

/* models.h:225-248 */
template <typename T>
inline void addStrainRateAndRotationRate(M3<T>& ed, M3<T>& rd, T mj, T dj, V3<T> v, V3<T> g)
{
    T tmp = T(-0.5) * mj / dj;
    const T vv[3] = {v.x, v.y, v.z};
    const T gg[3] = {g.x, g.y, g.z};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) ed.m[a][b] += tmp * (vv[a] * gg[b] + vv[b] * gg[a]);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) rd.m[a][b] += tmp * (vv[a] * gg[b] - vv[b] * gg[a]);
}

/* models.h:250-268 */
template <typename T> inline M3<T> dstress_dt(const M3<T>& ed, const M3<T>& rd, const M3<T>& S, T shear)
{
    M3<T> o;
    for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
            o.m[d][e] = T(2.0) * shear * ed.m[d][e];
            for (int f = 0; f < 3; f++) {
                if (d == e) o.m[d][e] -= T(2.0) * shear * ed.m[f][f] / T(3.0);
                o.m[d][e] += S.m[d][f] * rd.m[e][f];
                o.m[d][e] += S.m[e][f] * rd.m[d][f];
            }
        }
    return o;
}

/* ---- particle access in the C-ABI layout -------------------------------- */
template <typename T> struct Parts {
    int64_t n;
    T *posm, *vel, *acc, *xvel, *density, *density_dt, *balsara, *maxvsig, *dstress, *dstress_dt;
    V3<T> pos(int64_t i) const { return {posm[4 * i], posm[4 * i + 1], posm[4 * i + 2]}; }
    T mass(int64_t i) const { return posm[4 * i + 3]; }
    static V3<T> v4(const T* a, int64_t i) { return a ? V3<T>{a[4 * i], a[4 * i + 1], a[4 * i + 2]} : V3<T>{0, 0, 0}; }
    static void s4(T* a, int64_t i, V3<T> v)
    {
        if (!a) return;
        a[4 * i] = v.x; a[4 * i + 1] = v.y; a[4 * i + 2] = v.z; a[4 * i + 3] = T(0);
    }
    static T s1(const T* a, int64_t i) { return a ? a[i] : T(0); }
    M3<T> tens(const T* a, int64_t i) const
    {
        M3<T> r = mzero<T>();
        if (!a) return r;
        r.m[0][0] = a[0 * n + i]; r.m[1][1] = a[1 * n + i]; r.m[2][2] = a[2 * n + i];
        r.m[0][1] = r.m[1][0] = a[3 * n + i];
        r.m[0][2] = r.m[2][0] = a[4 * n + i];
        r.m[1][2] = r.m[2][1] = a[5 * n + i];
        return r;
    }
    void store_tens(T* a, int64_t i, const M3<T>& s) const
    {
        if (!a) return;
        a[0 * n + i] = s.m[0][0]; a[1 * n + i] = s.m[1][1]; a[2 * n + i] = s.m[2][2];
        a[3 * n + i] = s.m[0][1]; a[4 * n + i] = s.m[0][2]; a[5 * n + i] = s.m[1][2];
    }
};

/* index list helper: sample == NULL means every particle */
struct Range {
    const int64_t* s;
    int64_t cnt;
    int64_t at(int64_t k) const { return s ? s[k] : k; }
};

/* ---- calcDensity: src/computeDerivatives.h:52-91 ------------------------ */
template <typename T> void calc_density(const Consts<T>& C, const Parts<T>& P, Range R)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < R.cnt; ++k) {
        const int64_t i = R.at(k);
        const V3<T> pi = P.pos(i);
        T density = T(0); /* DENSITY is not in load_type: value-initialised */
        for (int64_t j = 0; j < P.n; ++j) {
            const V3<T> rij = pi - P.pos(j);
            const T r2 = dot(rij, rij);
            if (r2 <= C.H2 && r2 > 0) {
                const T r = std::sqrt(r2);
                const T w = Wspline(r, C.H, C.Wpref);
                density += P.mass(j) * w;
            }
        }
        if (C.has(GRASPH_DEAL_WITH_NO_PARTNERS))
            if (density < T(0.001)) density = C.rho0 * T(0.01);
        P.density[i] = density;
    }
}

/* ---- calcBalsaraDensityDTDStressDT: src/computeDerivatives.h:96-162 ----- */
template <typename T> void calc_balsara(const Consts<T>& C, const Parts<T>& P, Range R)
{
    const bool solids = C.has(GRASPH_SOLIDS), bals = C.has(GRASPH_BALSARA_SWITCH),
               intd = C.has(GRASPH_INTEGRATE_DENSITY);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < R.cnt; ++k) {
        const int64_t i = R.at(k);
        const V3<T> pi = P.pos(i);
        const V3<T> vi = P.v4(P.vel, i);
        const T rhoi = P.density[i];
        M3<T> edot = mzero<T>(), rdot = mzero<T>();
        T divv = T(0);
        V3<T> curlv{0, 0, 0};
        for (int64_t j = 0; j < P.n; ++j) {
            const V3<T> rij = pi - P.pos(j);
            const T r2 = dot(rij, rij);
            if (r2 <= C.H2 && r2 > 0) {
                const T r = std::sqrt(r2);
                const T dw = dWspline(r, C.H, C.dWpref);
                const V3<T> gradw = (dw / r) * rij;
                const V3<T> vij = vi - P.v4(P.vel, j);
                const T mj = P.mass(j), rhoj = P.density[j];
                if (solids)
                    addStrainRateAndRotationRate(edot, rdot, mj, rhoj, vij, gradw);
                else if (bals)
                    curlv += mj / rhoj * cross(vij, gradw);
                if (bals || intd) divv -= (mj / rhoj) * dot(vij, gradw);
            }
        }
        /* do_after; store_type attributes not set stay zero */
        M3<T> sdt = mzero<T>();
        T balsara = T(0), rho_dt = T(0);
        if (solids) {
            sdt = dstress_dt(edot, rdot, P.tens(P.dstress, i), C.shear);
            curlv = V3<T>{T(-2) * rdot.m[1][2], T(-2) * rdot.m[2][0], T(-2) * rdot.m[0][1]};
        }
        if (intd) rho_dt = -rhoi * divv;
        if (bals) balsara = balsaraSwitch(divv, curlv, C.c, C.H);
        if (P.balsara) P.balsara[i] = balsara;
        if (P.density_dt) P.density_dt[i] = rho_dt;
        P.store_tens(P.dstress_dt, i, sdt);
    }
}

/* computeDerivatives.h:33-47 environmentAcceleration: the constant
 * acceleration (:42-44).  CLOHESSY_WILTSHIRE (:36-40) is rejected by
 * validate(): it adds a scalar to an f3_t through an mpUtils operator
 * (v0.12.0, not vendored) whose meaning no available source fixes. */
template <typename T> inline void environmentAcceleration(const Consts<T>& C, V3<T>& acc)
{
    if (C.has(GRASPH_CONSTANT_ACCELERATION)) acc += C.const_acc;
}

/* ---- calcAcceleration: src/computeDerivatives.h:167-313 ----------------- */
template <typename T> void calc_acceleration(const Consts<T>& C, const Parts<T>& P, Range R)
{
    const bool grav = C.has(GRASPH_SELF_GRAVITY), sph = C.has(GRASPH_SPH), solids = C.has(GRASPH_SOLIDS),
               bals = C.has(GRASPH_BALSARA_SWITCH), av = C.has(GRASPH_ARTIFICIAL_VISCOSITY),
               arts = C.has(GRASPH_ARTIFICIAL_STRESS), xsph = C.has(GRASPH_XSPH);
    const T c2 = C.c * C.c;
    const T Wnorm = arts ? Wspline(C.normalsep, C.H, C.Wpref) : T(1);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < R.cnt; ++k) {
        const int64_t i = R.at(k);
        const V3<T> pi = P.pos(i);
        const V3<T> vi = P.v4(P.vel, i);
        const T rhoi = P.s1(P.density, i), bi = P.s1(P.balsara, i);
        V3<T> acc{0, 0, 0}, xvel{0, 0, 0};
        T maxvsig = T(0);
        /* do_before */
        M3<T> sOR_i = mzero<T>(), arts_i = mzero<T>();
        T sOR_il = 0, arts_il = 0;
        if (sph) {
            if (solids) {
                M3<T> sig = P.tens(P.dstress, i);
                const T pres = eos_murnaghan(rhoi, C.rho0, C.BULK, C.dBULKdP);
                sig.m[0][0] -= pres; sig.m[1][1] -= pres; sig.m[2][2] -= pres;
                for (int e = 0; e < 9; ++e) sOR_i.m[e / 3][e % 3] = sig.m[e / 3][e % 3] / (rhoi * rhoi);
                if (arts) arts_i = artificialStress(C.mateps, sOR_i);
            } else {
                const T sig = -eos_liquid(rhoi, C.rho0, c2);
                sOR_il = sig / (rhoi * rhoi);
                if (arts) arts_il = artificialPressure(C.mateps, sOR_il);
            }
        }
        for (int64_t j = 0; j < P.n; ++j) {
            const V3<T> rij = pi - P.pos(j);
            const T r2 = dot(rij, rij);
            if (!(r2 > 0)) continue;
            const T mj = P.mass(j);
            if (grav) {
                const T distSqr = r2 + C.H2;
                const T invDist = T(1) / std::sqrt(distSqr);
                const T invDistCube = invDist * invDist * invDist;
                acc -= rij * mj * invDistCube;
            }
            if (sph && r2 <= C.H2) {
                const T r = std::sqrt(r2);
                const T dw = dWspline(r, C.H, C.dWpref);
                const V3<T> gradw = (dw / r) * rij;
                const T rhoj = P.density[j];
                const V3<T> vj = P.v4(P.vel, j);
                const V3<T> vij = vi - vj;
                V3<T> sg;
                if (solids) {
                    M3<T> sig = P.tens(P.dstress, j);
                    const T pres = eos_murnaghan(rhoj, C.rho0, C.BULK, C.dBULKdP);
                    sig.m[0][0] -= pres; sig.m[1][1] -= pres; sig.m[2][2] -= pres;
                    M3<T> st;
                    for (int e = 0; e < 9; ++e)
                        st.m[e / 3][e % 3] = sOR_i.m[e / 3][e % 3] + sig.m[e / 3][e % 3] / (rhoj * rhoj);
                    (void)arts; /* SOLIDS + ARTIFICIAL_STRESS is rejected (does not compile upstream) */
                    sg = {st.m[0][0] * gradw.x + st.m[0][1] * gradw.y + st.m[0][2] * gradw.z,
                          st.m[1][0] * gradw.x + st.m[1][1] * gradw.y + st.m[1][2] * gradw.z,
                          st.m[2][0] * gradw.x + st.m[2][1] * gradw.y + st.m[2][2] * gradw.z};
                } else {
                    const T sig = -eos_liquid(rhoj, C.rho0, c2);
                    const T sOR_j = sig / (rhoj * rhoj);
                    T st = sOR_il + sOR_j;
                    if (arts) {
                        const T f = std::pow(Wspline(r, C.H, C.Wpref) / Wnorm, C.matexp);
                        st += f * (arts_il + artificialPressure(C.mateps, sOR_j));
                    }
                    sg = st * gradw;
                }
                acc += mj * sg;
                if (av) {
                    const T bj = bals ? P.s1(P.balsara, j) : T(1);
                    const T II = artificialViscosity(bals ? &maxvsig : (T*)nullptr, C.alpha, rhoi, rhoj, vij, rij, r,
                                                     C.c, C.c, bals ? bi : T(1), bj);
                    acc -= mj * II * gradw;
                }
                if (xsph) xvel += T(2) * mj / (rhoi + rhoj) * (vj - vi) * Wspline(r, C.H, C.Wpref);
            }
        }
        environmentAcceleration(C, acc);
        P.s4(P.acc, i, acc);
        P.s4(P.xvel, i, xvel);
        if (P.maxvsig) P.maxvsig[i] = maxvsig;
    }
}

/* ---- integration.h:24-125 (and the kick-drift of :168-217) --------------- */
/* kick v += a (dt kick_factor), drift, density and stress update, plasticity,
 * box boundary: fixedTsLeapfrog::do_for_each (:89-125) with kick_factor 1
 * (not first step) or 1/2, variableTsLeapfrog_useNewTimestep (:175-216)
 * with kick_factor 1/2 -- the two bodies are the same statements. */
template <typename T> void leapfrog_kick_drift(const Consts<T>& C, const Parts<T>& P, const T dt, const T kick_factor)
{
    const bool sph = C.has(GRASPH_SPH), xsph = C.has(GRASPH_XSPH), intd = C.has(GRASPH_INTEGRATE_DENSITY),
               solids = C.has(GRASPH_SOLIDS), box = C.has(GRASPH_SIMPLE_BOX_BOUNDARY);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P.n; ++i) {
        V3<T> pos = P.pos(i), vel = P.v4(P.vel, i);
        const V3<T> acc = P.v4(P.acc, i);
        vel = vel + acc * (dt * kick_factor);
        if (xsph && sph)
            pos = pos + (vel + C.xsph * P.v4(P.xvel, i)) * dt;
        else
            pos = pos + vel * dt;
        if (sph) {
            if (intd) {
                T rho = P.density[i] + P.density_dt[i] * dt;
                if (rho < T(0.0)) rho = T(0.0);
                P.density[i] = rho;
            }
            if (solids) {
                M3<T> S = P.tens(P.dstress, i);
                const M3<T> Sd = P.tens(P.dstress_dt, i);
                for (int e = 0; e < 9; ++e) S.m[e / 3][e % 3] += Sd.m[e / 3][e % 3] * dt;
                if (C.has(GRASPH_PLASTICITY_MC))
                    plasticity(S, mohrCoulombYieldStress(C.tanfr, eos_murnaghan(P.density[i], C.rho0, C.BULK, C.dBULKdP),
                                                         C.cohesion));
                else if (C.has(GRASPH_PLASTICITY_MIESE))
                    plasticity(S, C.Y);
                P.store_tens(P.dstress, i, S);
            }
        }
        if (box) { /* integration.h:33-73 */
            const T vm = T(1.0) + C.refl;
            T* pp[3] = {&pos.x, &pos.y, &pos.z};
            T* vv[3] = {&vel.x, &vel.y, &vel.z};
            const T mx[3] = {C.bmax.x, C.bmax.y, C.bmax.z}, mn[3] = {C.bmin.x, C.bmin.y, C.bmin.z};
            const int nd = (C.dim == 3) ? 3 : 2;
            for (int d = 0; d < nd; ++d) {
                if (*pp[d] > mx[d]) { *pp[d] = mx[d]; *vv[d] -= vm * *vv[d]; }
                else if (*pp[d] < mn[d]) { *pp[d] = mn[d]; *vv[d] -= vm * *vv[d]; }
            }
        }
        P.posm[4 * i] = pos.x; P.posm[4 * i + 1] = pos.y; P.posm[4 * i + 2] = pos.z;
        P.s4(P.vel, i, vel);
    }
}

template <typename T> void fixed_ts_leapfrog(const Consts<T>& C, const Parts<T>& P, int not_first_step)
{
    leapfrog_kick_drift(C, P, C.dt, (not_first_step) ? T(1.0) : T(0.5));  // integration.h:95
}

/* variableTsLeapfrog_getNewTimestep (integration.h:130-165): half kick with
 * the current dt ts[0], then ts[1] = min(ts[1], max_timestep, criteria of
 * every particle) -- the block minima and the atomicMin of the reference
 * give exactly this minimum.  Criteria: sqrt(accel_accuracy H / |a|) and
 * velocity_accuracy H / maxvsig (:149-154). */
template <typename T> void vts_get_new_timestep(const Consts<T>& C, const Parts<T>& P, T* ts)
{
    const T dt = ts[0];
    const T accH = C.acc_acc * C.H, velH = C.vel_acc * C.H;
    T mindt = C.ts_max;
    for (int64_t i = 0; i < P.n; ++i) {
        V3<T> vel = P.v4(P.vel, i);
        const V3<T> acc = P.v4(P.acc, i);
        vel = vel + acc * (dt * T(0.5));
        P.s4(P.vel, i, vel);
        if (C.has(GRASPH_ACCELERATION_CRITERION)) mindt = std::fmin(mindt, std::sqrt(accH / length(acc)));
        if (C.has(GRASPH_VELOCITY_CRITERION)) mindt = std::fmin(mindt, velH / P.maxvsig[i]);
    }
    if (P.n > 0) ts[1] = std::fmin(ts[1], mindt);
}

/* variableTsLeapfrog_useNewTimestep (integration.h:167-217) */
template <typename T> void vts_use_new_timestep(const Consts<T>& C, const Parts<T>& P, T* ts)
{
    const T dt = (ts[1] < C.ts_min) ? C.ts_min : ts[1];
    ts[0] = dt;
    leapfrog_kick_drift(C, P, dt, T(0.5));
}

/* integrate(pb, notFirstStep) with VARIABLE_TIMESTEP_LEAPFROG (:225-243) */
template <typename T> void variable_ts_leapfrog(const Consts<T>& C, const Parts<T>& P, int not_first_step, T* ts)
{
    if (not_first_step)
        vts_get_new_timestep(C, P, ts);
    else
        ts[1] = C.ts_init;
    vts_use_new_timestep(C, P, ts);
}

template <typename T> int validate(const grasph_params* p)
{
    if (!p) return -1;
    if (p->dimension < 1 || p->dimension > 3) return -1;
    if ((p->flags & GRASPH_SOLIDS) && (p->flags & GRASPH_ARTIFICIAL_STRESS)) return -3;
    if (p->flags & GRASPH_CLOHESSY_WILTSHIRE) return -3;
    if ((p->flags & GRASPH_PLASTICITY_MIESE) && (p->flags & GRASPH_PLASTICITY_MC)) return -1;
    return 0;
}

} // namespace

namespace {
template <typename T> Parts<T> parts_from(const oracle_particles_t* p)
{
    Parts<T> r;
    r.n = p->n;
    r.posm = (T*)p->posm; r.vel = (T*)p->vel; r.acc = (T*)p->acc; r.xvel = (T*)p->xvel;
    r.density = (T*)p->density; r.density_dt = (T*)p->density_dt; r.balsara = (T*)p->balsara;
    r.maxvsig = (T*)p->maxvsig; r.dstress = (T*)p->dstress; r.dstress_dt = (T*)p->dstress_dt;
    return r;
}

/* ts: the variable step's {currentTS, nextTS} (integration.h:127-128), or
 * NULL.  With VARIABLE_TIMESTEP_LEAPFROG, calcAcceleration::do_after sets
 * nextTS = max_timestep (computeDerivatives.h:305-308). */
template <typename T> int pass(int which, const grasph_params* prm, const oracle_particles_t* p, const int64_t* s,
                               int64_t ns, T* ts)
{
    int v = validate<T>(prm);
    if (v) return v;
    // the SPH passes read and write the SPH arrays, which a state of
    // gravity alone does not have: an error, not a write through null
    if (which < 2 && (!p->density || !p->vel)) return -1;
    if (which == 1 && (prm->flags & GRASPH_SOLIDS) && (!p->dstress || !p->dstress_dt)) return -1;
    Consts<T> C = make_consts<T>(prm);
    Parts<T> P = parts_from<T>(p);
    Range R{s, s ? ns : P.n};
    if (which == 0) calc_density(C, P, R);
    else if (which == 1) calc_balsara(C, P, R);
    else {
        calc_acceleration(C, P, R);
        if (ts && C.has(GRASPH_VARIABLE_TIMESTEP)) ts[1] = C.ts_max;
    }
    return 0;
}

/* computeDerivatives.h:315-328 */
template <typename T> int derivs(const grasph_params* prm, const oracle_particles_t* p, T* ts)
{
    int v = validate<T>(prm);
    if (v) return v;
    const uint32_t f = prm->flags;
    if (f & GRASPH_SPH) {
        if (!(f & GRASPH_INTEGRATE_DENSITY)) pass<T>(0, prm, p, nullptr, 0, ts);
        if ((f & GRASPH_SOLIDS) || (f & GRASPH_INTEGRATE_DENSITY) || (f & GRASPH_BALSARA_SWITCH))
            pass<T>(1, prm, p, nullptr, 0, ts);
    }
    pass<T>(2, prm, p, nullptr, 0, ts);
    return 0;
}

/* FP64 pair potential phi_i = -sum_{j: r2>0} m_j / sqrt(r2 + H^2) */
template <typename T> void potential(const grasph_params* prm, const oracle_particles_t* p, double* phi)
{
    Parts<T> P = parts_from<T>(p);
    const double H2 = double(prm->H) * double(prm->H);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < P.n; ++i) {
        double s = 0;
        const double xi = P.posm[4 * i], yi = P.posm[4 * i + 1], zi = P.posm[4 * i + 2];
        for (int64_t j = 0; j < P.n; ++j) {
            const double dx = xi - P.posm[4 * j], dy = yi - P.posm[4 * j + 1], dz = zi - P.posm[4 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0) s -= double(P.posm[4 * j + 3]) / std::sqrt(r2 + H2);
        }
        phi[i] = s;
    }
}

/* internal energy per unit mass of the EOS, u(rho) = int_{rho0}^{rho} p/rho'^2 */
inline double u_eos(const grasph_params* prm, double rho)
{
    const double r0 = prm->rho0;
    if (!(rho > 0)) return 0;
    if (prm->flags & GRASPH_SOLIDS) {
        const double K = prm->BULK, n = prm->dBULKdP;
        if (std::fabs(n - 1.0) < 1e-12) return K * (std::log(rho / r0) / r0 + 1.0 / rho - 1.0 / r0);
        return K / n * ((std::pow(rho, n - 1) - std::pow(r0, n - 1)) / ((n - 1) * std::pow(r0, n)) + 1.0 / rho - 1.0 / r0);
    }
    const double c2 = double(prm->SOUNDSPEED) * double(prm->SOUNDSPEED);
    return c2 * (std::log(rho / r0) + r0 / rho - 1.0);
}

template <typename T>
void invariants(const grasph_params* prm, const oracle_particles_t* p, const double* pos64, double hk,
                const double* phi, double* out)
{
    Parts<T> P = parts_from<T>(p);
    double ek = 0, ep = 0, ei = 0, px = 0, py = 0, pz = 0, lx = 0, ly = 0, lz = 0, m = 0;
    for (int64_t i = 0; i < P.n; ++i) {
        const double mi = P.posm[4 * i + 3];
        double vx = P.vel[4 * i], vy = P.vel[4 * i + 1], vz = P.vel[4 * i + 2];
        if (P.acc && hk != 0) {
            vx += double(P.acc[4 * i]) * hk; vy += double(P.acc[4 * i + 1]) * hk; vz += double(P.acc[4 * i + 2]) * hk;
        }
        double x = P.posm[4 * i], y = P.posm[4 * i + 1], z = P.posm[4 * i + 2];
        if (pos64) { x = pos64[i]; y = pos64[P.n + i]; z = pos64[2 * P.n + i]; }
        ek += 0.5 * mi * (vx * vx + vy * vy + vz * vz);
        px += mi * vx; py += mi * vy; pz += mi * vz;
        lx += mi * (y * vz - z * vy); ly += mi * (z * vx - x * vz); lz += mi * (x * vy - y * vx);
        m += mi;
        if (phi) ep += 0.5 * mi * phi[i];
        if ((prm->flags & GRASPH_SPH) && P.density) {
            const double rho = P.density[i];
            ei += mi * u_eos(prm, rho);
            if ((prm->flags & GRASPH_SOLIDS) && P.dstress && rho > 0) {
                const int64_t n = P.n;
                double ss = 0;
                for (int c = 0; c < 6; ++c) {
                    const double s = P.dstress[c * n + i];
                    ss += (c < 3 ? 1.0 : 2.0) * s * s;
                }
                ei += mi * ss / (4.0 * double(prm->shear) * rho);
            }
        }
    }
    out[GRASPH_INV_EKIN] = ek; out[GRASPH_INV_EPOT] = ep; out[GRASPH_INV_EINT] = ei;
    out[GRASPH_INV_PX] = px; out[GRASPH_INV_PY] = py; out[GRASPH_INV_PZ] = pz;
    out[GRASPH_INV_LX] = lx; out[GRASPH_INV_LY] = ly; out[GRASPH_INV_LZ] = lz;
    out[GRASPH_INV_MASS] = m;
}
} // namespace

extern "C" {

void oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
    (void)n;
}
int oracle_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* passes on a sample of i (NULL = all); j always runs over all particles;
 * ts = {currentTS, nextTS} in the precision of the suffix or NULL */
int oracle_calc_density_f32(const grasph_params* prm, const oracle_particles_t* p, const int64_t* s, int64_t ns)
{ return pass<float>(0, prm, p, s, ns, nullptr); }
int oracle_calc_density_f64(const grasph_params* prm, const oracle_particles_t* p, const int64_t* s, int64_t ns)
{ return pass<double>(0, prm, p, s, ns, nullptr); }
int oracle_calc_balsara_f32(const grasph_params* prm, const oracle_particles_t* p, const int64_t* s, int64_t ns)
{ return pass<float>(1, prm, p, s, ns, nullptr); }
int oracle_calc_balsara_f64(const grasph_params* prm, const oracle_particles_t* p, const int64_t* s, int64_t ns)
{ return pass<double>(1, prm, p, s, ns, nullptr); }
int oracle_calc_acceleration_f32(const grasph_params* prm, const oracle_particles_t* p, const int64_t* s, int64_t ns,
                                 float* ts)
{ return pass<float>(2, prm, p, s, ns, ts); }
int oracle_calc_acceleration_f64(const grasph_params* prm, const oracle_particles_t* p, const int64_t* s, int64_t ns,
                                 double* ts)
{ return pass<double>(2, prm, p, s, ns, ts); }

int oracle_compute_derivatives_f32(const grasph_params* prm, const oracle_particles_t* p, float* ts)
{ return derivs<float>(prm, p, ts); }
int oracle_compute_derivatives_f64(const grasph_params* prm, const oracle_particles_t* p, double* ts)
{ return derivs<double>(prm, p, ts); }

int oracle_fixed_ts_leapfrog_f32(const grasph_params* prm, const oracle_particles_t* p, int nfs)
{
    int v = validate<float>(prm);
    if (v) return v;
    fixed_ts_leapfrog(make_consts<float>(prm), parts_from<float>(p), nfs);
    return 0;
}
int oracle_fixed_ts_leapfrog_f64(const grasph_params* prm, const oracle_particles_t* p, int nfs)
{
    int v = validate<double>(prm);
    if (v) return v;
    fixed_ts_leapfrog(make_consts<double>(prm), parts_from<double>(p), nfs);
    return 0;
}
/* ts = {currentTS, nextTS} in the precision of the suffix, updated in place */
int oracle_variable_ts_leapfrog_f32(const grasph_params* prm, const oracle_particles_t* p, int nfs, float* ts)
{
    int v = validate<float>(prm);
    if (v) return v;
    variable_ts_leapfrog(make_consts<float>(prm), parts_from<float>(p), nfs, ts);
    return 0;
}
int oracle_variable_ts_leapfrog_f64(const grasph_params* prm, const oracle_particles_t* p, int nfs, double* ts)
{
    int v = validate<double>(prm);
    if (v) return v;
    variable_ts_leapfrog(make_consts<double>(prm), parts_from<double>(p), nfs, ts);
    return 0;
}

void oracle_potential_f32(const grasph_params* prm, const oracle_particles_t* p, double* phi) { potential<float>(prm, p, phi); }
void oracle_potential_f64(const grasph_params* prm, const oracle_particles_t* p, double* phi) { potential<double>(prm, p, phi); }
void oracle_invariants_f32(const grasph_params* prm, const oracle_particles_t* p, const double* pos64, double hk,
                           const double* phi, double* out)
{ invariants<float>(prm, p, pos64, hk, phi, out); }
void oracle_invariants_f64(const grasph_params* prm, const oracle_particles_t* p, const double* pos64, double hk,
                           const double* phi, double* out)
{ invariants<double>(prm, p, pos64, hk, phi, out); }

/* scalar physics functions, for the closed-form pins */
double oracle_wspline_f64(double r, double h, int dim) { return Wspline<double>(r, h, splinePrefactor<double>(dim, h)); }
double oracle_dwspline_f64(double r, double h, int dim) { return dWspline<double>(r, h, dsplinePrefactor<double>(dim, h)); }
float oracle_wspline_f32(float r, float h, int dim) { return Wspline<float>(r, h, splinePrefactor<float>(dim, h)); }
float oracle_dwspline_f32(float r, float h, int dim) { return dWspline<float>(r, h, dsplinePrefactor<float>(dim, h)); }
double oracle_eos_liquid_f64(double rho, double rho0, double c2) { return eos_liquid<double>(rho, rho0, c2); }
double oracle_eos_murnaghan_f64(double rho, double rho0, double K0, double dK0) { return eos_murnaghan<double>(rho, rho0, K0, dK0); }
double oracle_artificial_viscosity_f64(double* maxvsig, double alpha, double di, double dj, const double* vij,
                                       const double* rij, double r, double ci, double cj, double bi, double bj)
{
    return artificialViscosity<double>(maxvsig, alpha, di, dj, V3<double>{vij[0], vij[1], vij[2]},
                                       V3<double>{rij[0], rij[1], rij[2]}, r, ci, cj, bi, bj);
}
double oracle_balsara_switch_f64(double divv, const double* curl, double c, double h)
{ return balsaraSwitch<double>(divv, V3<double>{curl[0], curl[1], curl[2]}, c, h); }
void oracle_plasticity_f64(double* S9, double Y)
{
    M3<double> S;
    std::memcpy(S.m, S9, sizeof(S.m));
    plasticity(S, Y);
    std::memcpy(S9, S.m, sizeof(S.m));
}
void oracle_dstress_dt_f64(const double* ed9, const double* rd9, const double* S9, double shear, double* out9)
{
    M3<double> e, r, s;
    std::memcpy(e.m, ed9, sizeof(e.m)); std::memcpy(r.m, rd9, sizeof(r.m)); std::memcpy(s.m, S9, sizeof(s.m));
    M3<double> o = dstress_dt(e, r, s, shear);
    std::memcpy(out9, o.m, sizeof(o.m));
}
void oracle_strain_rotation_f64(double* ed9, double* rd9, double mj, double dj, const double* v, const double* g)
{
    M3<double> e, r;
    std::memcpy(e.m, ed9, sizeof(e.m)); std::memcpy(r.m, rd9, sizeof(r.m));
    addStrainRateAndRotationRate(e, r, mj, dj, V3<double>{v[0], v[1], v[2]}, V3<double>{g[0], g[1], g[2]});
    std::memcpy(ed9, e.m, sizeof(e.m)); std::memcpy(rd9, r.m, sizeof(r.m));
}

} // extern "C"
