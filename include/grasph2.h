/*
 * grasph2.h -- C ABI of the MI355X-native GraSPH2 derivative + leapfrog path.
 *
 * Plain C: pointers, sizes and integer status codes only; no HIP or C++ types.
 * Every entry point returns 0 on success and a negative GRASPH_E_* code on
 * failure; grasph_last_error() then describes the failure.  All device work is
 * enqueued asynchronously on the stream passed in (a hipStream_t given as
 * void*, NULL = the null stream); the library never synchronises the device
 * unless an entry point says so.
 *
 * Reference interface replaced (paths relative to the GraSPH2 source tree):
 *   computeDerivatives(pb)            src/computeDerivatives.h:315-328
 *   do_for_each_pair_fast<calcDensity>                       :52-91, 320
 *   do_for_each_pair_fast<calcBalsaraDensityDTDStressDT>     :96-162, 323
 *   do_for_each_pair_fast<calcAcceleration>                  :167-313, 327
 *   integrate(pb, notFirstStep)       src/integration.h:219-244
 *   do_for_each<fixedTsLeapfrog>      src/integration.h:82-125, 224
 *   getCurrentTimestep()              src/integration.h:246-255
 *   compile-time physics switches     settings/settings.h:31-171
 *   DeviceParticleBuffer<DEV_...>     settings/settings.h:184,
 *                                     src/particles/particle_buffer_impl.h:45-243
 */
#ifndef GRASPH2_H
#define GRASPH2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRASPH_VERSION 10100 /* major*10000 + minor*100 + patch */

/* ---- status codes ------------------------------------------------------ */
#define GRASPH_OK 0
#define GRASPH_E_INVALID (-1)     /* bad argument: null pointer, bad size, bad flag combination */
#define GRASPH_E_HIP (-2)         /* a HIP runtime call failed */
#define GRASPH_E_UNSUPPORTED (-3) /* combination the reference does not build, or whose meaning it leaves open */
#define GRASPH_E_NOMEM (-4)       /* device allocation of library scratch failed */

/* ---- physics switches (settings/settings.h) ---------------------------- */
#define GRASPH_SELF_GRAVITY          (1u << 0)  /* ENABLE_SELF_GRAVITY       settings.h:34  */
#define GRASPH_SPH                   (1u << 1)  /* ENABLE_SPH                settings.h:37  */
#define GRASPH_INTEGRATE_DENSITY     (1u << 2)  /* INTEGRATE_DENSITY         settings.h:40  */
#define GRASPH_DEAL_WITH_NO_PARTNERS (1u << 3)  /* DEAL_WITH_NO_PARTNERS     settings.h:41  */
#define GRASPH_SOLIDS                (1u << 4)  /* SOLIDS                    settings.h:42  */
#define GRASPH_PLASTICITY_MIESE      (1u << 5)  /* PLASTICITY_MIESE          settings.h:116 */
#define GRASPH_PLASTICITY_MC         (1u << 6)  /* PLASTICITY_MC             settings.h:120 */
#define GRASPH_CLOHESSY_WILTSHIRE    (1u << 7)  /* CLOHESSY_WILTSHIRE        settings.h:136:
                                                   GRASPH_E_UNSUPPORTED (computeDerivatives.h:37-39
                                                   adds a scalar to a float3 through an mpUtils
                                                   operator whose meaning no source here fixes) */
#define GRASPH_CONSTANT_ACCELERATION (1u << 8)  /* CONSTANT_ACCELERATION     settings.h:141 */
#define GRASPH_SIMPLE_BOX_BOUNDARY   (1u << 9)  /* SIMPLE_BOX_BOUNDARY       settings.h:146 */
#define GRASPH_ARTIFICIAL_VISCOSITY  (1u << 10) /* ARTIFICIAL_VISCOSITY      settings.h:155 */
#define GRASPH_BALSARA_SWITCH        (1u << 11) /* BALSARA_SWITCH            settings.h:159 */
#define GRASPH_ARTIFICIAL_STRESS     (1u << 12) /* ARTIFICIAL_STRESS         settings.h:163 */
#define GRASPH_XSPH                  (1u << 13) /* XSPH                      settings.h:170 */
#define GRASPH_VARIABLE_TIMESTEP     (1u << 14) /* VARIABLE_TIMESTEP_LEAPFROG settings.h:55 (else FIXED_...) */
#define GRASPH_ACCELERATION_CRITERION (1u << 15) /* ACCELERATION_CRITERION   settings.h:59  */
#define GRASPH_VELOCITY_CRITERION    (1u << 16) /* VELOCITY_CRITERION        settings.h:61  */
#define GRASPH_ALL_FLAGS             ((1u << 17) - 1u)

/*
 * Runtime replacement of the compile-time constants of settings/settings.h.
 * Gravity uses G = 1 and the SPH smoothing length H as softening
 * (src/computeDerivatives.h:231).
 */
typedef struct grasph_params {
    uint32_t flags;          /* GRASPH_* switches above */
    int32_t dimension;       /* 1, 2 or 3                       settings.h:31  */
    float H;                 /* smoothing length                settings.h:99  */
    float rho0;              /* rest density                    settings.h:105 */
    float BULK;              /* bulk modulus K0                 settings.h:106 */
    float dBULKdP;           /* dK0/dp (Murnaghan exponent)     settings.h:107 */
    float SOUNDSPEED;        /* speed of sound c                settings.h:108 */
    float shear;             /* shear modulus mu                settings.h:111 */
    float Y;                 /* von Mises yield stress          settings.h:117 */
    float tanfr;             /* tan(friction angle)             settings.h:122 */
    float cohesion;          /* Mohr-Coulomb cohesion           settings.h:123 */
    float alpha;             /* artificial viscosity strength   settings.h:156 */
    float mateps;            /* artificial stress epsilon       settings.h:164 */
    float matexp;            /* artificial stress exponent      settings.h:165 */
    float normalsep;         /* artificial stress separation    settings.h:166 */
    float xsph_factor;       /* XSPH factor                     settings.h:171 */
    float cw_n;              /* Clohessy-Wiltshire n (unused)   settings.h:137 */
    float constant_acceleration[3];   /*                        settings.h:142 */
    float box_min[3];        /*                                 settings.h:147 */
    float box_max[3];        /*                                 settings.h:148 */
    float box_reflectiveness;/*                                 settings.h:149 */
    float fixed_timestep;    /* dt of the fixed-step leapfrog   settings.h:50  */
    float initial_timestep;  /* variable step: first dt         settings.h:56  */
    float min_timestep;      /*   smallest dt                   settings.h:57  */
    float max_timestep;      /*   largest dt                    settings.h:58  */
    float accel_accuracy;    /*   dt <= sqrt(accel_accuracy H / |a|)   :60   */
    float velocity_accuracy; /*   dt <= velocity_accuracy H / maxvsig  :62   */
} grasph_params;

/* Fill *p with the defaults of settings/settings.h: self-gravity, SPH,
 * integrated density, solids, Mohr-Coulomb plasticity, artificial viscosity,
 * the Balsara switch and the variable-step leapfrog with both time-step
 * criteria. */
void grasph_params_default(grasph_params* p);

/*
 * Device particle arrays ("DeviceParticleBuffer<DEV_POSM, DEV_VEL, ...>").
 * Caller-owned device memory; the library never frees it.  Layout:
 *   posm       float[n][4]   x, y, z, mass            (DEV_POSM)
 *   pos64      double[3][n]  FP64 master position planes x|y|z, or NULL.
 *              When set, the integrator drifts in FP64 and writes the FP32
 *              mirror posm.xyz from it; when NULL it drifts posm in FP32.
 *   vel, acc, xvel  float[n][4]  xyz + unused pad     (DEV_VEL/ACC/XVEL)
 *   density, density_dt, balsara, maxvsig  float[n]   (DEV_DENSITY ...)
 *   dstress, dstress_dt  float[6][n]  symmetric tensor planes
 *              xx|yy|zz|xy|xz|yz                      (DEV_DSTRESS[_DT])
 * Arrays a configuration does not read or write may be NULL; an entry point
 * returns GRASPH_E_INVALID when one it needs is NULL.
 */
typedef struct grasph_particles {
    int64_t n;
    float* posm;
    double* pos64;
    float* vel;
    float* acc;
    float* xvel;
    float* density;
    float* density_dt;
    float* balsara;
    float* maxvsig;
    float* dstress;
    float* dstress_dt;
} grasph_particles;

typedef struct grasph_ctx grasph_ctx;

/* library identity */
int grasph_version(void);
/* Message of the last failure on ctx (or of the last failed grasph_create
 * when ctx is NULL).  Never NULL. */
const char* grasph_last_error(const grasph_ctx* ctx);

/* Per-device context: owns scratch (neighbour lists, Morton order, partial
 * sums, timing events).  Not thread-safe; use one context per host thread. */
int grasph_create(int device, grasph_ctx** out);
int grasph_destroy(grasph_ctx* ctx);

/* Tuning knobs.  Keys: "max_neighbours" (list capacity per particle,
 * default 160; a particle with more neighbours takes an exact scan of all
 * partners, with the same results; above 512, four lanes share each
 * particle's neighbours in the single-instance passes, which splits their
 * sums four ways: equal to FP32 rounding), "max_neighbours_grow" (1 = after an
 * single-instance evaluation whose lists overflowed, a later one raises
 * max_neighbours to 1.25 times the largest count seen, in steps of 32, up
 * to 4096 and 16 GiB of lists; the count travels to the host behind an
 * event and is read without waiting, and not during graph capture; the
 * default; 0 = fixed; batches keep the capacity last set),
 * "reorder_interval"
 * (Morton re-sort every k derivative evaluations, default 8; 0 = never;
 * the order changes only the order of compensated sums, not which pairs
 * are found; a new particle count or position array re-sorts at once,
 * and an array allocated at a freed one's address counts as the same
 * particles until the next re-sort: slower, never wrong),
 * "profile" (1 = record HIP events around each kernel), "sweep_i_per_lane"
 * (1 or 2), "sweep_slices" (j-split of the sweep, 0 = automatic),
 * "batch_waves_per_block", "batch_blocks_per_cu" (batched sweep launch
 * shape), "sweep_symmetric" (single
 * instance: 1 = gravity between groups of 1024 particles evaluated once per
 * pair for both sides, 0 = the one-sided sweep, 2 = the symmetric sweep from
 * 16384 particles on, the default; results agree to FP32 rounding, the
 * neighbour lists are identical), "sym_workgroups" (workgroups of the
 * symmetric sweep's far-field kernel, single instances and batches, 0 = two
 * wavefronts per SIMD), "sym_partial_bytes" (single instance: device memory
 * the symmetric far field's partner-side partials may take, at least 1 MiB,
 * default 12 GiB; they need 6 n^2 / 1024 bytes in all, and past this the far
 * field runs in column batches of partner groups, each with its share of
 * the partials, the running sums kept between batches in 32 B per particle;
 * results agree with the unbatched far field to FP32 rounding),
 * "sym_near_shares" (wavefronts per 64 particles in its near field, 1-16,
 * 0 = chosen from the occupancy), "sym_list_priority" (its near-field list
 * kernel takes issue priority over the far field running beside it: 1 = yes,
 * 0 = no, 2 = below 49152 particles or once max_neighbours has grown past
 * 512, the default), "batch_symmetric" (1 = batches use the
 * symmetric sweep for their instances of at least "batch_symmetric_min"
 * particles (default 0), the smaller ones the one-sided persistent sweep,
 * the default; gravity then equals the single-instance path's to FP32
 * rounding, the neighbour sums bit for bit; 0 = every instance takes the
 * one-sided persistent sweep, whose results equal the single-instance
 * path's bit for bit when that path is not j-split). */
int grasph_set_option(grasph_ctx* ctx, const char* key, int64_t value);
/* The current value of an option of grasph_set_option (the autotuner sets
 * some of them), or the read-only "noalloc_slices" (the j-split the last
 * grasph_compute_derivatives_noalloc call on this context took);
 * GRASPH_E_INVALID for an unknown key. */
int grasph_get_option(const grasph_ctx* ctx, const char* key, int64_t* value);

/* ---- the hot path ------------------------------------------------------- */

/* computeDerivatives(pb)                       src/computeDerivatives.h:315-328
 * Runs the passes the flags select, in the reference's order:
 * density summation (:320), strain rate / Balsara / density rate (:323),
 * acceleration (:327).  Fused here into one all-pairs sweep (gravity,
 * density, neighbour list) plus two neighbour-list passes.
 * With GRASPH_VARIABLE_TIMESTEP the acceleration pass (here and in
 * grasph_calc_acceleration) sets the context's next dt to max_timestep, as
 * calcAcceleration::do_after sets nextTS (:305-308). */
int grasph_compute_derivatives(grasph_ctx* ctx, const grasph_particles* p,
                               const grasph_params* prm, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same passes and results as
 * grasph_compute_derivatives with options sweep_symmetric = 0 and
 * reorder_interval = 1 (and sweep_slices as chosen below), computed
 * without any device allocation: every work array (Morton order and sort
 * storage, slot-ordered positions and records, neighbour lists, the
 * j-split's partial sums) is cut from one static device array of
 * GRASPH_NOALLOC_SCRATCH_BYTES per GPU, held from library load until the
 * process ends.  Runs on `stream` alone (no second stream), re-sorts the
 * Morton order on every call and records no timing events.  The j-split
 * ("sweep_slices", automatic when 0) is reduced until the arrays fit; when
 * even one slice does not fit, returns GRASPH_E_INVALID and says how many
 * bytes the call needs.  With SPH and the default 160 neighbours per list a
 * particle needs ~760 bytes (about 44000 particles); gravity alone needs
 * none at one slice.  Every call on one GPU shares that GPU's array: calls
 * that overlap in time on one GPU (several host threads, or streams that do
 * not order them) race and give wrong results.  Nothing in the array needs
 * to start at zero: each call writes every element before reading it. */
#define GRASPH_NOALLOC_SCRATCH_BYTES (32LL * 1024 * 1024)
int grasph_compute_derivatives_noalloc(grasph_ctx* ctx, const grasph_particles* p,
                                       const grasph_params* prm, void* stream);

/* The three passes individually, each standalone (each runs its own sweep):
 *   do_for_each_pair_fast<calcDensity>                   computeDerivatives.h:52-91
 *   do_for_each_pair_fast<calcBalsaraDensityDTDStressDT> computeDerivatives.h:96-162
 *   do_for_each_pair_fast<calcAcceleration>              computeDerivatives.h:167-313 */
int grasph_calc_density(grasph_ctx* ctx, const grasph_particles* p,
                        const grasph_params* prm, void* stream);
int grasph_calc_balsara_density_dt_dstress_dt(grasph_ctx* ctx, const grasph_particles* p,
                                              const grasph_params* prm, void* stream);
int grasph_calc_acceleration(grasph_ctx* ctx, const grasph_particles* p,
                             const grasph_params* prm, void* stream);

/* integrate(pb, notFirstStep)                 src/integration.h:219-244
 * FIXED_TIMESTEP_LEAPFROG: fixedTsLeapfrog, src/integration.h:82-125.
 * GRASPH_VARIABLE_TIMESTEP: the two kernels of VARIABLE_TIMESTEP_LEAPFROG,
 * src/integration.h:127-217.  With not_first_step, a half kick with the
 * current dt that also folds the particles' dt criteria into the next dt
 * (an atomic minimum; the acceleration pass of the derivative call before
 * it set the next dt to max_timestep, so dt = min(max_timestep, criteria)
 * and grows again when the criteria loosen); without it the next dt is set
 * to initial_timestep.  Then a half kick and the drift with
 * dt = max(next dt, min_timestep), which becomes the current dt.  The two
 * values live in the context (zero at creation, like the reference's
 * __device__ currentTS / nextTS), so the calls stay asynchronous. */
int grasph_integrate(grasph_ctx* ctx, const grasph_particles* p,
                     const grasph_params* prm, int not_first_step, void* stream);

/* getCurrentTimestep()                         src/integration.h:246-255
 * fixed_timestep, or with GRASPH_VARIABLE_TIMESTEP the context's current dt
 * (waits for the device, as the reference's cudaMemcpyFromSymbol does). */
double grasph_get_current_timestep(const grasph_ctx* ctx, const grasph_params* prm);

/* ---- batched ensembles (no counterpart in the reference) ---------------
 * Many independent simulations advanced together.  Instance k owns the
 * particles [offsets[k], offsets[k+1]) of one grasph_particles whose n is
 * offsets[n_instances]; each array is the concatenation of the instances'
 * own single-instance layouts (a planar array of P planes holds instance k's
 * planes at P * offsets[k], plane stride offsets[k+1] - offsets[k]).  So an
 * instance's sub-view is a pointer offset and every single-instance entry
 * point applies to it unchanged.  params[k] may differ per instance except
 * for the flags.  offsets and params are host arrays. */
typedef struct grasph_batch {
    int32_t n_instances;
    const int64_t* offsets;      /* [n_instances + 1], offsets[0] == 0, non-decreasing */
    const grasph_params* params; /* [n_instances] */
} grasph_batch;

/* computeDerivatives for every instance: one persistent all-pairs launch
 * over a queue of (instance, tile) items, longest instances first, then one
 * launch per neighbour pass over all instances.  With
 * GRASPH_VARIABLE_TIMESTEP every instance's next dt := its max_timestep
 * afterwards (computeDerivatives.h:305-308). */
int grasph_batch_compute_derivatives(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b,
                                     void* stream);
/* integrate(pb, notFirstStep) for every instance, one launch (two with the
 * variable step, whose per-instance dt state lives with the batch's tables:
 * zero when a batch layout or its params are first seen). */
int grasph_batch_integrate(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, int not_first_step,
                           void* stream);
/* getCurrentTimestep() of every instance of the batch last prepared with
 * exactly these offsets and params: dt[k] (waits for the device). */
int grasph_batch_get_current_timesteps(grasph_ctx* ctx, const grasph_batch* b, double* dt);
/* n_steps of (computeDerivatives, integrate(true)) for every instance; each
 * step is one replay of a captured HIP graph. */
int grasph_batch_step(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, int n_steps, void* stream);
/* The same steps as the plain per-instance launch loop (single-instance
 * entry points on each instance's sub-view), for comparison. */
int grasph_batch_step_loop(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, int n_steps,
                           void* stream);
/* Pick, by timing the batched sweep on this batch, the particles per lane of
 * each size bucket (N < 2048, < 4096, < 8192, >= 8192) and the wavefronts
 * per block and blocks per CU of the persistent launch.  With cache_path,
 * reads a previous result for this device and bucket set from that file or
 * appends the new one.  Synchronises `stream`. */
int grasph_autotune(grasph_ctx* ctx, const grasph_particles* p, const grasph_batch* b, const char* cache_path,
                    void* stream);
/* The tuning in use: ni[4] per bucket, waves per block, blocks per CU. */
int grasph_get_batch_tuning(const grasph_ctx* ctx, int32_t ni[4], int32_t* waves_per_block, int32_t* blocks_per_cu);

/* ---- diagnostics (no counterpart in the reference) --------------------- */

#define GRASPH_INV_EKIN 0   /* 1/2 sum m |v + a*half_kick_dt|^2             */
#define GRASPH_INV_EPOT 1   /* -1/2 sum_i sum_{j!=i} m_i m_j / sqrt(r^2+H^2) */
#define GRASPH_INV_EINT 2   /* elastic + compressional energy (SPH only)     */
#define GRASPH_INV_PX 3     /* linear momentum sum m v                      */
#define GRASPH_INV_PY 4
#define GRASPH_INV_PZ 5
#define GRASPH_INV_LX 6     /* angular momentum sum m x cross v             */
#define GRASPH_INV_LY 7
#define GRASPH_INV_LZ 8
#define GRASPH_INV_MASS 9   /* total mass                                   */
#define GRASPH_NINV 10

/* FP64 global invariants.  half_kick_dt synchronises the half-step leapfrog
 * velocities to the current time (pass dt/2 after computeDerivatives, 0 to
 * use the stored velocities).  Runs its own gravity-potential sweep when
 * with_potential is nonzero.  Synchronises `stream` before returning. */
int grasph_reduce_invariants(grasph_ctx* ctx, const grasph_particles* p,
                             const grasph_params* prm, double half_kick_dt,
                             int with_potential, double out[GRASPH_NINV], void* stream);

/* Per-kernel timing recorded while option "profile" is 1: accumulated
 * milliseconds and launch counts (recorded intervals) per kernel class, in
 * the order sweep, density, strain, accel, integrate, reorder, other, far.
 * "far" is the far-field kernel of the symmetric sweep, which runs on the
 * caller's stream while the near field, the list merge and the SPH passes
 * run beside it on the context's second stream; "sweep" holds the rest of
 * the all-pairs sweep (with the symmetric sweep three intervals per
 * evaluation: near-field lists, in-group gravity, combine).  Synchronises the recorded events.  Resets the counters when
 * reset != 0. */
#define GRASPH_NKERNEL_CLASSES 8
int grasph_get_kernel_times(grasph_ctx* ctx, double ms[GRASPH_NKERNEL_CLASSES],
                            int64_t launches[GRASPH_NKERNEL_CLASSES], int reset);

/* Neighbour statistics of the last neighbour list this context built:
 * ordered neighbour pairs (0 < r^2 <= H^2), the largest per-particle count,
 * and how many particles exceeded "max_neighbours" (those take the all-j
 * fallback).  Synchronises the device. */
int grasph_get_neighbour_stats(grasph_ctx* ctx, int64_t* pairs, int32_t* max_count, int64_t* overflowed);

#ifdef __cplusplus
}
#endif

#endif /* GRASPH2_H */
