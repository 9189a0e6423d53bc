// integrate.hip -- leapfrog kick-drift with plasticity (fixed and variable
// step), and the FP64 global invariants (gfx950).
//
// leapfrog: fixedTsLeapfrog::do_for_each (src/integration.h:82-125) with
// doPlasticity (:24-31) and handleBoundaryConditions (:33-73); one particle
// per lane, every array read and written once, coalesced (HBM-bound).  With
// FP64 master positions the drift runs in FP64 and the FP32 mirror is
// rewritten from it; without them it drifts the FP32 positions as the
// reference does.  One kernel body serves one instance or a batch (block
// descriptors and per-instance parameters from tables).
//
// Variable step (src/integration.h:127-217): the same kernel, with dt read
// on the device from the instance's {current dt, next dt} pair and a half
// kick, does variableTsLeapfrog_useNewTimestep; vts_kick_kernel does
// variableTsLeapfrog_getNewTimestep (half kick with the current dt, then the
// minimum of the dt criteria into next dt with one atomic minimum per block
// -- non-negative floats order like their bit patterns).  No host round trip.
#include "internal.hpp"

namespace grasph {
namespace {

constexpr int kBlock = 256;

// one axis of handleBoundaryConditions (integration.h:33-73): put a particle
// past a wall on the wall and reflect that velocity component
template <typename T>
__device__ __forceinline__ void box_axis(T& x, float& v, const float lo, const float hi, const float vm)
{
    if (x > (T)hi) {
        x = (T)hi;
        v -= vm * v;
    } else if (x < (T)lo) {
        x = (T)lo;
        v -= vm * v;
    }
}

// dt of the second variable-step kernel, integration.h:170
__device__ __forceinline__ float vts_dt(const float* t, const DevParams& P)
{
    const float next = t[1];
    return (next < P.ts_min) ? P.ts_min : next;
}

template <bool BATCH>
__global__ __launch_bounds__(kBlock) void leapfrog_kernel(float4* __restrict__ posm, double* __restrict__ pos64,
                                                          float4* __restrict__ vel, const float4* __restrict__ acc,
                                                          const float4* __restrict__ xvel,
                                                          float* __restrict__ density,
                                                          const float* __restrict__ density_dt,
                                                          float* __restrict__ dstress,
                                                          const float* __restrict__ dstress_dt, const int n,
                                                          const int not_first_step, const BlockTables T,
                                                          const DevParams P0, float* __restrict__ ts)
{
    const SlotMap m = slot_map<BATCH>(T, n);
    const int i = m.slot0 + threadIdx.x;  // the integrator walks particles in index order
    if (i >= m.slot_end) return;
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    float dt = P.dt;
    float kick_factor = (not_first_step) ? 1.0f : 0.5f;  // integration.h:95
    if (ts) {
        float* t = ts + 2 * m.inst;
        dt = vts_dt(t, P);
        kick_factor = 0.5f;  // integration.h:182
        if (i == m.off) t[0] = dt;  // currentTS = dt (:171-172); every block only reads t[1]
    }
    const bool sph = P.flags & GRASPH_SPH;

    const float4 a = acc[i];
    float4 v = vel[i];
    const float kick = dt * kick_factor;
    v.x = v.x + a.x * kick;
    v.y = v.y + a.y * kick;
    v.z = v.z + a.z * kick;

    // integration.h:98-102
    float ux = v.x, uy = v.y, uz = v.z;
    if ((P.flags & GRASPH_XSPH) && sph) {
        const float4 xv = xvel[i];
        ux = v.x + P.xsph * xv.x;
        uy = v.y + P.xsph * xv.y;
        uz = v.z + P.xsph * xv.z;
    }
    // with FP64 masters the FP32 mirror is only written (x y z; the mass in
    // w stays), so posm is not read
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    double X = 0, Y = 0, Z = 0;
    if (pos64) {
        X = pos64[m.plane3(0, i)] + (double)ux * (double)dt;
        Y = pos64[m.plane3(1, i)] + (double)uy * (double)dt;
        Z = pos64[m.plane3(2, i)] + (double)uz * (double)dt;
        p.x = (float)X; p.y = (float)Y; p.z = (float)Z;
    } else {
        p = posm[i];
        p.x = p.x + ux * dt;
        p.y = p.y + uy * dt;
        p.z = p.z + uz * dt;
    }

    float rho = 0.0f;
    if (sph) {
        if (P.flags & GRASPH_INTEGRATE_DENSITY) {  // :105-110
            rho = density[i] + density_dt[i] * dt;
            if (rho < 0.0f) rho = 0.0f;
            density[i] = rho;
        } else if (P.flags & GRASPH_PLASTICITY_MC) {
            rho = density[i];
        }
        if (P.flags & GRASPH_SOLIDS) {  // :111-117
            Sym6 S;
            S.xx = dstress[m.plane6(0, i)] + dstress_dt[m.plane6(0, i)] * dt;
            S.yy = dstress[m.plane6(1, i)] + dstress_dt[m.plane6(1, i)] * dt;
            S.zz = dstress[m.plane6(2, i)] + dstress_dt[m.plane6(2, i)] * dt;
            S.xy = dstress[m.plane6(3, i)] + dstress_dt[m.plane6(3, i)] * dt;
            S.xz = dstress[m.plane6(4, i)] + dstress_dt[m.plane6(4, i)] * dt;
            S.yz = dstress[m.plane6(5, i)] + dstress_dt[m.plane6(5, i)] * dt;
            if (P.flags & GRASPH_PLASTICITY_MC)
                plasticity(S, P.tanfr * eos_murnaghan(rho, P.rho0, P.BULK, P.dBULKdP) + P.cohesion);
            else if (P.flags & GRASPH_PLASTICITY_MIESE)
                plasticity(S, P.Y);
            dstress[m.plane6(0, i)] = S.xx; dstress[m.plane6(1, i)] = S.yy; dstress[m.plane6(2, i)] = S.zz;
            dstress[m.plane6(3, i)] = S.xy; dstress[m.plane6(4, i)] = S.xz; dstress[m.plane6(5, i)] = S.yz;
        }
    }

    if (P.flags & GRASPH_SIMPLE_BOX_BOUNDARY) {  // :33-73, on the FP64 masters when present
        const float vm = 1.0f + P.refl;
        if (pos64) {
            box_axis(X, v.x, P.bminx, P.bmaxx, vm);
            box_axis(Y, v.y, P.bminy, P.bmaxy, vm);
            if (P.dim == 3) box_axis(Z, v.z, P.bminz, P.bmaxz, vm);
            p.x = (float)X; p.y = (float)Y; p.z = (float)Z;
        } else {
            box_axis(p.x, v.x, P.bminx, P.bmaxx, vm);
            box_axis(p.y, v.y, P.bminy, P.bmaxy, vm);
            if (P.dim == 3) box_axis(p.z, v.z, P.bminz, P.bmaxz, vm);
        }
    }
    if (pos64) {
        pos64[m.plane3(0, i)] = X; pos64[m.plane3(1, i)] = Y; pos64[m.plane3(2, i)] = Z;
    }
    if (pos64)
        *reinterpret_cast<float3*>(posm + i) = make_float3(p.x, p.y, p.z);
    else
        posm[i] = p;
    vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

// variableTsLeapfrog_getNewTimestep (src/integration.h:130-165).  The
// criteria are rounded as written (no contraction), so dt matches an FP32
// evaluation of the same formulas bit for bit.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void vts_kick_kernel(float4* __restrict__ vel, const float4* __restrict__ acc,
                                                          const float* __restrict__ maxvsig, const int n,
                                                          const BlockTables T, const DevParams P0,
                                                          float* __restrict__ ts)
{
    __shared__ float red[kBlock / 64];
    const SlotMap m = slot_map<BATCH>(T, n);
    const int i = m.slot0 + threadIdx.x;
    const DevParams P = BATCH ? T.prm[m.inst] : P0;
    float* t = ts + 2 * m.inst;
    float mindt = P.ts_max;  // :134
    if (i < m.slot_end) {
        const float dt = t[0];
        const float4 a = acc[i];
        float4 v = vel[i];
        const float h = dt * 0.5f;  // :144
        v.x = v.x + a.x * h;
        v.y = v.y + a.y * h;
        v.z = v.z + a.z * h;
        vel[i] = make_float4(v.x, v.y, v.z, 0.0f);
        if (P.flags & GRASPH_ACCELERATION_CRITERION) {  // :149-151
            const float la = __fsqrt_rn(dist2_exact(a.x, a.y, a.z));  // length(acc), products and sums unfused
            mindt = fminf(mindt, __fsqrt_rn(__fdiv_rn(mul_exact(P.acc_acc, P.H), la)));
        }
        if (P.flags & GRASPH_VELOCITY_CRITERION)  // :152-154
            mindt = fminf(mindt, __fdiv_rn(mul_exact(P.vel_acc, P.H), maxvsig[i]));
    }
    // block minimum, then one atomic minimum on the bit pattern (:158-168)
    for (int off = 32; off >= 1; off >>= 1) mindt = fminf(mindt, __shfl_xor(mindt, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mindt;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int w = 1; w < kBlock / 64; ++w) r = fminf(r, red[w]);
        atomicMin((int*)&t[1], __float_as_int(r));
    }
}

// next dt := initial_timestep (integrate(pb, false), integration.h:233-234)
// or max_timestep (the reset at the end of every acceleration pass,
// computeDerivatives.h:305-308), one thread per instance
__global__ void vts_set_next_kernel(float* __restrict__ ts, const int ninst, const DevParams* __restrict__ prm,
                                    const DevParams P0, const int which)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ninst) return;
    const DevParams& P = prm ? prm[k] : P0;
    ts[2 * k + 1] = which == 0 ? P.ts_init : P.ts_max;
}

// internal energy per unit mass of the EOS, u(rho) = int_{rho0}^{rho} p / rho'^2 drho'
__device__ inline double u_eos(uint32_t flags, double rho, double r0, double K, double n, double c2)
{
    if (!(rho > 0)) return 0.0;
    if (flags & GRASPH_SOLIDS) {
        if (fabs(n - 1.0) < 1e-12) return K * (log(rho / r0) / r0 + 1.0 / rho - 1.0 / r0);
        return K / n * ((pow(rho, n - 1) - pow(r0, n - 1)) / ((n - 1) * pow(r0, n)) + 1.0 / rho - 1.0 / r0);
    }
    return c2 * (log(rho / r0) + r0 / rho - 1.0);
}

struct InvArgs {
    uint32_t flags;
    double rho0, K, n, c2, shear, half_kick;
};

__global__ __launch_bounds__(kBlock) void invariants_kernel(const float4* __restrict__ posm,
                                                            const double* __restrict__ pos64,
                                                            const float4* __restrict__ vel,
                                                            const float4* __restrict__ acc,
                                                            const float* __restrict__ density,
                                                            const float* __restrict__ dstress,
                                                            const float* __restrict__ pot, const int n,
                                                            double* __restrict__ partials, const InvArgs A)
{
    double v[GRASPH_NINV];
#pragma unroll
    for (int k = 0; k < GRASPH_NINV; ++k) v[k] = 0.0;
    const size_t ns = (size_t)n;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float4 p = posm[i];
        const double m = p.w;
        const float4 u = vel[i];
        double vx = u.x, vy = u.y, vz = u.z;
        if (acc && A.half_kick != 0.0) {
            const float4 a = acc[i];
            vx += (double)a.x * A.half_kick; vy += (double)a.y * A.half_kick; vz += (double)a.z * A.half_kick;
        }
        double x = p.x, y = p.y, z = p.z;
        if (pos64) { x = pos64[i]; y = pos64[ns + i]; z = pos64[2 * ns + i]; }
        v[GRASPH_INV_EKIN] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
        if (pot) v[GRASPH_INV_EPOT] += 0.5 * m * (double)pot[i];
        if ((A.flags & GRASPH_SPH) && density) {
            const double rho = density[i];
            double e = m * u_eos(A.flags, rho, A.rho0, A.K, A.n, A.c2);
            if ((A.flags & GRASPH_SOLIDS) && dstress && rho > 0) {
                double ss = 0;
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const double s = dstress[c * ns + i];
                    ss += (c < 3 ? 1.0 : 2.0) * s * s;
                }
                e += m * ss / (4.0 * A.shear * rho);
            }
            v[GRASPH_INV_EINT] += e;
        }
        v[GRASPH_INV_PX] += m * vx; v[GRASPH_INV_PY] += m * vy; v[GRASPH_INV_PZ] += m * vz;
        v[GRASPH_INV_LX] += m * (y * vz - z * vy);
        v[GRASPH_INV_LY] += m * (z * vx - x * vz);
        v[GRASPH_INV_LZ] += m * (x * vy - y * vx);
        v[GRASPH_INV_MASS] += m;
    }
    // fixed-order block reduction: wave shuffle, then the block's waves in order
    __shared__ double sh[kBlock / 64][GRASPH_NINV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < GRASPH_NINV; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[wave][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < GRASPH_NINV) {
        double x = 0;
        for (int w = 0; w < kBlock / 64; ++w) x += sh[w][threadIdx.x];
        partials[(size_t)blockIdx.x * GRASPH_NINV + threadIdx.x] = x;
    }
}

__global__ void invariants_final(double* __restrict__ partials, const int nblocks)
{
    const int k = threadIdx.x;
    if (k >= GRASPH_NINV) return;
    double x = 0;
    for (int b = 0; b < nblocks; ++b) x += partials[(size_t)b * GRASPH_NINV + k];
    partials[(size_t)nblocks * GRASPH_NINV + k] = x;
}

}  // namespace

hipError_t launch_integrate(const IntegLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (L.T.nblocks > 0)
        hipLaunchKernelGGL(leapfrog_kernel<true>, dim3(L.T.nblocks), dim3(kBlock), 0, s, L.posm, L.pos64, L.vel, L.acc,
                           L.xvel, L.density, L.density_dt, L.dstress, L.dstress_dt, L.n, L.not_first_step, L.T, P,
                           L.ts);
    else
        hipLaunchKernelGGL(leapfrog_kernel<false>, dim3((L.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, L.posm,
                           L.pos64, L.vel, L.acc, L.xvel, L.density, L.density_dt, L.dstress, L.dstress_dt, L.n,
                           L.not_first_step, L.T, P, L.ts);
    return hipGetLastError();
}

hipError_t launch_vts_kick(const IntegLaunch& L, const DevParams& P, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    if (!L.ts) return hipErrorInvalidValue;
    if (L.T.nblocks > 0)
        hipLaunchKernelGGL(vts_kick_kernel<true>, dim3(L.T.nblocks), dim3(kBlock), 0, s, L.vel, L.acc, L.maxvsig,
                           L.n, L.T, P, L.ts);
    else
        hipLaunchKernelGGL(vts_kick_kernel<false>, dim3((L.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, L.vel,
                           L.acc, L.maxvsig, L.n, L.T, P, L.ts);
    return hipGetLastError();
}

hipError_t launch_vts_set_next(float* ts, int n_instances, const DevParams* prm, const DevParams& P, int which,
                               hipStream_t s)
{
    if (!ts || n_instances <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(vts_set_next_kernel, dim3((n_instances + 255) / 256), dim3(256), 0, s, ts, n_instances, prm, P,
                       which);
    return hipGetLastError();
}

// integrate(pb, notFirstStep), src/integration.h:219-244
hipError_t launch_integrate_step(const IntegLaunch& L, const DevParams& P, hipStream_t s)
{
    if (!L.ts) return launch_integrate(L, P, s);
    hipError_t e;
    if (L.not_first_step) {
        if ((e = launch_vts_kick(L, P, s)) != hipSuccess) return e;
    } else if ((e = launch_vts_set_next(L.ts, L.n_instances, L.T.nblocks > 0 ? L.T.prm : nullptr, P, 0, s)) !=
               hipSuccess) {
        return e;
    }
    return launch_integrate(L, P, s);
}

int inv_blocks(int n)
{
    int b = (n + kBlock - 1) / kBlock;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return b;
}

hipError_t launch_invariants(const InvLaunch& L, const grasph_params& prm, hipStream_t s)
{
    InvArgs A;
    A.flags = prm.flags;
    A.rho0 = prm.rho0; A.K = prm.BULK; A.n = prm.dBULKdP;
    A.c2 = (double)prm.SOUNDSPEED * (double)prm.SOUNDSPEED;
    A.shear = prm.shear;
    A.half_kick = L.half_kick;
    hipLaunchKernelGGL(invariants_kernel, dim3(L.nblocks), dim3(kBlock), 0, s, L.posm, L.pos64, L.vel, L.acc,
                       L.density, L.dstress, L.pot, L.n, L.partials, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(invariants_final, dim3(1), dim3(64), 0, s, L.partials, L.nblocks);
    return hipGetLastError();
}

}  // namespace grasph
