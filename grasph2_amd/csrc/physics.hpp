// physics.hpp -- device-side physics of the GraSPH2 hot path (gfx950).
//
// Each function states the reference lines whose arithmetic it follows.  The
// operation order of the reference is kept (so FP32 results track the
// reference's to a few ulp); hipcc may still contract a*b+c into one FMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/grasph2.h"

namespace grasph {

// Runtime form of the compile-time constants of settings/settings.h and
// src/computeDerivatives.h:26-28, all in FP32 as the reference computes them.
struct DevParams {
    uint32_t flags;
    int32_t dim;
    float H, H2, Wpref, dWpref;
    float rho0, BULK, dBULKdP, c, c2, shear, Y, tanfr, cohesion, alpha;
    float mateps, matexp, normalsep, Wnorm, xsph, dt;
    float cax, cay, caz;
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz, refl;
    float cand;  // d2 = r2 + H2 threshold that admits every pair with r2 <= H2
    // variable-step leapfrog, settings/settings.h:56-62
    float ts_init, ts_min, ts_max, acc_acc, vel_acc;
    // FP64 constants of the passes' kernel and gradient: 1/H, the W and dW
    // prefactors and the dW prefactor over H (SPH only, H > 0)
    double invH64, Wpref64, dWpref64, dWprefH64;
};

// Everything below rounds each product and sum as written, wherever it is
// inlined: no multiply-add contraction (HIP contracts by default, and a
// contraction decided per call site would let two walks over the same pairs
// -- the neighbour list and its overflow scan -- round differently).  The
// FP32 oracle is built the same way.  The default resumes at the end of
// this header.
#pragma clang fp contract(off)

// src/sph/kernel.h:41-60
__host__ __device__ inline float spline_prefactor(int dim, float h)
{
    if (dim == 1) return 4.0f / (3.0f * h);
    if (dim == 2) return 40.0f / (7.0f * float(M_PI) * h * h);
    return 8.0f / (float(M_PI) * h * h * h);
}
// src/sph/kernel.h:68-87
__host__ __device__ inline float dspline_prefactor(int dim, float h)
{
    if (dim == 1) return 24.0f / (3.0f * h * h);
    if (dim == 2) return 240.0f / (7.0f * float(M_PI) * h * h * h);
    return 48.0f / (float(M_PI) * h * h * h * h);
}

// src/sph/kernel.h:130-136 (cubic B-spline, support radius h)
__host__ __device__ inline float w_spline(float r, float h, float factor)
{
    const float q = r / h;
    return factor * ((q <= 1.0f) ? ((q < 0.5f) ? (6.0f * q * q * q - 6.0f * q * q + 1.0f)
                                               : 2.0f * (1.0f - q) * (1.0f - q) * (1.0f - q))
                                 : 0.0f);
}
// src/sph/kernel.h:169-175
__host__ __device__ inline float dw_spline(float r, float h, float factor)
{
    const float q = r / h;
    return factor * ((q <= 1.0f) ? ((q < 0.5f) ? (3.0f * q * q - 2.0f * q) : -1.0f * (1.0f - q) * (1.0f - q))
                                 : 0.0f);
}

// The kernel gradient g = dW(r)/r * r_ij (src/sph/kernel.h:169-175 over r)
// of a pair at the FP32 separation (rx, ry, rz), evaluated in FP64 and
// rounded once per component.  Where a particle's pair terms cancel (the
// compressed zone of a contact), the half-dozen FP32 roundings of this
// factor, amplified by the cancellation, dominated the acceleration error
// against FP64; rounded once they leave a fifth of it.  For q < 1/2 the
// factor is prefactor (3 q - 2) / H (q / r = 1 / H), for q >= 1/2
// -prefactor (1 - q)^2 / r; both pieces meet at q = 1/2.
__device__ inline double dw_over_r_f64(float rx, float ry, float rz, const DevParams& P)
{
    const double x = rx, y = ry, z = rz;
    const double r2 = (x * x + y * y) + z * z;
    const double ir = rsqrt(r2);
    const double q = r2 * ir * P.invH64;
    const double om = 1.0 - q;
    return (q < 0.5) ? P.dWprefH64 * (3.0 * q - 2.0) : -P.dWpref64 * (om * om) * ir;
}
__device__ inline void kernel_gradient_f64(float rx, float ry, float rz, const DevParams& P, float& gx, float& gy,
                                           float& gz)
{
    const double s = dw_over_r_f64(rx, ry, rz, P);
    gx = (float)(s * (double)rx);
    gy = (float)(s * (double)ry);
    gz = (float)(s * (double)rz);
}
// W(r) (src/sph/kernel.h:130-136) in FP64 at the FP32 separation, rounded
// once: the density sum then carries about half an ulp per particle, which
// the pressures of a strongly compressed zone would otherwise amplify
__device__ inline float kernel_w_f64(float rx, float ry, float rz, const DevParams& P)
{
    const double x = rx, y = ry, z = rz;
    const double q = sqrt((x * x + y * y) + z * z) * P.invH64;
    const double om = 1.0 - q;
    const double w = (q < 0.5) ? (6.0 * q * q * q - 6.0 * q * q + 1.0) : 2.0 * (om * om * om);
    return (float)(P.Wpref64 * w);
}

// src/sph/eos.h:88-96
__device__ inline float eos_liquid(float rho, float rho0, float c2) { return c2 * (rho - rho0); }
__device__ inline float eos_murnaghan(float rho, float rho0, float K0, float dK0)
{
    return K0 / dK0 * (powf(rho / rho0, dK0) - 1.0f);
}

// Exact FP32 squared distance in the reference's dot() order
// (x*x + y*y) + z*z with every product and sum rounded (no contraction),
// so the neighbour test r2 <= H2 && r2 > 0 decides exactly as the oracle.
// (HIP's __fmul_rn / __fadd_rn are plain operators that a contracting caller
// may fuse; the contract(off) region above keeps these unfused.)
__device__ inline float dist2_exact(float rx, float ry, float rz)
{
    return (rx * rx + ry * ry) + rz * rz;
}

// a * b and a + b rounded as written, never fused into a multiply-add
__device__ inline float mul_exact(float a, float b) { return a * b; }
__device__ inline float add_exact(float a, float b) { return a + b; }

// Neighbour predicate of src/computeDerivatives.h:72,122,238:  0 < r2 <= H2.
// As unsigned bit patterns of non-negative floats: r2 in (0, H2]  <=>
// bits(r2) - 1 < bits(H2); +0 wraps to 0xffffffff and NaN lies above +inf.
__device__ inline bool is_neighbour(float r2, float H2)
{
    return (__float_as_uint(r2) - 1u) < __float_as_uint(H2);
}

// TwoSum (Knuth): s + e == a + b exactly.  Used for the compensated
// (Kahan-Babuska) accumulation of acceleration, density and potential.
struct CSum {
    float s, c;
    __device__ inline void add(float x)
    {
        const float t = s + x;
        const float bp = t - s;
        const float e = (s - (t - bp)) + (x - bp);
        s = t;
        c += e;
    }
    __device__ inline float value() const { return s + c; }
};

// src/sph/models.h:172-185 (with maxVsig) / :158-170 (without)
__device__ inline float artificial_viscosity(float* maxvsig, float alpha, float di, float dj, float vx, float vy,
                                             float vz, float rx, float ry, float rz, float r, float ci, float cj,
                                             float bi, float bj)
{
    const float wij = (rx * vx + ry * vy + rz * vz) / r;
    float II = 0.0f;
    if (wij < 0.0f) {
        const float vsig = ci + cj - 3.0f * wij;
        const float rhoij = (di + dj) * 0.5f;
        II = -0.25f * (bi + bj) * alpha * wij * vsig / rhoij;
        if (maxvsig) *maxvsig = fmaxf(*maxvsig, vsig);
    }
    return II;
}

// src/sph/models.h:187-192
__device__ inline float balsara_switch(float divv, float cx, float cy, float cz, float c, float h)
{
    const float absdiv = fabsf(divv);
    const float abscurl = sqrtf(cx * cx + cy * cy + cz * cz);
    return absdiv / (absdiv + abscurl + 0.0001f * c / h);
}

// the same from FP64 sums, rounded once
__device__ inline float balsara_switch_f64(double divv, double cx, double cy, double cz, float c, float h)
{
    const double absdiv = fabs(divv);
    const double abscurl = sqrt(cx * cx + cy * cy + cz * cz);
    return (float)(absdiv / (absdiv + abscurl + 0.0001 * (double)c / (double)h));
}

// Symmetric tensor, unique components xx yy zz xy xz yz.
struct Sym6 {
    float xx, yy, zz, xy, xz, yz;
};

// src/sph/models.h:194-205 von Mises (the factor is Y^2/(3 J2), applied as is).
// J2 sums the nine entries in the reference's row-major order.
__device__ inline void plasticity(Sym6& S, float Y)
{
    float J2 = 0.0f;
    J2 += S.xx * S.xx; J2 += S.xy * S.xy; J2 += S.xz * S.xz;
    J2 += S.xy * S.xy; J2 += S.yy * S.yy; J2 += S.yz * S.yz;
    J2 += S.xz * S.xz; J2 += S.yz * S.yz; J2 += S.zz * S.zz;
    J2 *= 0.5f;
    const float f = Y * Y / (3.0f * J2);
    if (f < 1.0f) {
        S.xx *= f; S.yy *= f; S.zz *= f; S.xy *= f; S.xz *= f; S.yz *= f;
    }
}

__host__ inline DevParams make_dev_params(const grasph_params* p)
{
    DevParams d;
    d.flags = p->flags;
    d.dim = p->dimension;
    d.H = p->H;
    d.H2 = p->H * p->H;
    d.Wpref = spline_prefactor(d.dim, d.H);
    d.dWpref = dspline_prefactor(d.dim, d.H);
    d.rho0 = p->rho0; d.BULK = p->BULK; d.dBULKdP = p->dBULKdP;
    d.c = p->SOUNDSPEED; d.c2 = p->SOUNDSPEED * p->SOUNDSPEED;
    d.shear = p->shear; d.Y = p->Y; d.tanfr = p->tanfr; d.cohesion = p->cohesion; d.alpha = p->alpha;
    d.mateps = p->mateps; d.matexp = p->matexp; d.normalsep = p->normalsep;
    d.Wnorm = w_spline(p->normalsep, d.H, d.Wpref);
    d.xsph = p->xsph_factor; d.dt = p->fixed_timestep;
    d.cax = p->constant_acceleration[0]; d.cay = p->constant_acceleration[1]; d.caz = p->constant_acceleration[2];
    d.bminx = p->box_min[0]; d.bminy = p->box_min[1]; d.bminz = p->box_min[2];
    d.bmaxx = p->box_max[0]; d.bmaxy = p->box_max[1]; d.bmaxz = p->box_max[2];
    d.refl = p->box_reflectiveness;
    d.ts_init = p->initial_timestep; d.ts_min = p->min_timestep; d.ts_max = p->max_timestep;
    d.acc_acc = p->accel_accuracy; d.vel_acc = p->velocity_accuracy;
    // d2 = fma-chain(r2) + H2 carries at most a few ulp of relative error, so
    // pairs with r2 <= H2 all satisfy d2 <= 2 H2 (1 + 2^-18).
    d.cand = 2.0f * d.H2 * (1.0f + 1.0f / 262144.0f);
    const double h = p->H;
    d.invH64 = h > 0.0 ? 1.0 / h : 0.0;
    d.Wpref64 = d.dim == 1 ? 4.0 / (3.0 * h) : d.dim == 2 ? 40.0 / (7.0 * M_PI * h * h) : 8.0 / (M_PI * h * h * h);
    d.dWpref64 = d.dim == 1 ? 24.0 / (3.0 * h * h) : d.dim == 2 ? 240.0 / (7.0 * M_PI * h * h * h)
                                                                : 48.0 / (M_PI * h * h * h * h);
    d.dWprefH64 = h > 0.0 ? d.dWpref64 / h : 0.0;
    return d;
}

// src/computeDerivatives.h:33-47 environmentAcceleration: the constant
// acceleration (:42-44).  CLOHESSY_WILTSHIRE is rejected by the parameter
// check: its expression adds a scalar to a float3 through an mpUtils
// operator that no available source defines.
__device__ inline void environment_acceleration(const DevParams& P, float& ax, float& ay, float& az)
{
    if (P.flags & GRASPH_CONSTANT_ACCELERATION) {
        ax += P.cax; ay += P.cay; az += P.caz;
    }
}

#pragma clang fp contract(fast)

}  // namespace grasph
