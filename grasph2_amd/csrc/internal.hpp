// internal.hpp -- library-internal declarations shared by the .hip units.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <vector>

#include "physics.hpp"

namespace grasph {

enum KernelClass { KC_SWEEP = 0, KC_STRAIN = 1, KC_ACCEL = 2, KC_INTEGRATE = 3, KC_REORDER = 4, KC_OTHER = 5 };

// Per-device library state.  Scratch grows on demand and is reused.
struct Ctx {
    int device = 0;
    std::string err;

    // options
    int max_neighbours = 160;
    int reorder_interval = 1;
    bool profile = false;

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

    // neighbour list: nbr[k * nbr_stride + slot], nbr_cnt[slot]
    int* nbr = nullptr;
    int* nbr_cnt = nullptr;
    int64_t nbr_cap_n = 0;
    int nbr_cap_k = 0;
    int64_t last_list_n = 0;  // particles of the last neighbour list built

    // pass-C per-particle record: sigma/rho^2 (6), balsara, density
    float4* jrec = nullptr;
    int64_t jrec_cap = 0;

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
    double ms[GRASPH_NKERNEL_CLASSES] = {0, 0, 0, 0, 0, 0};
    int64_t launches[GRASPH_NKERNEL_CLASSES] = {0, 0, 0, 0, 0, 0};
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
    const float4* posm;
    const int* perm;
    int n;
    float4* acc;        // gravity (written when grav)
    float* density;     // written when dens
    int* nbr;           // written when nbr
    int* nbr_cnt;
    float* pot;         // written when pot
    int kmax;
    bool grav, dens, want_nbr, want_pot;
    bool finish;        // apply environmentAcceleration, zero xvel / maxvsig (SPH off)
    const float4* vel;
    float4* xvel;
    float* maxvsig;
};
hipError_t launch_sweep(const SweepLaunch& L, const DevParams& P, hipStream_t s);

struct NbrLaunch {
    const float4* posm;
    const float4* vel;
    const float* density;
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
    float4* jrec;
    float4* acc;
    float4* xvel;
    float* maxvsig;
    bool acc_has_gravity;
};
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
};
hipError_t launch_integrate(const IntegLaunch& L, const DevParams& P, hipStream_t s);

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

}  // namespace grasph
