// api_roundtrip.cpp -- drives the C++ template API (include/grasph2/particles.hpp)
// the way src/main.cu does: host buffer -> device buffer, then
// computeDerivatives(pb); integrate(pb, true) for a number of steps, then back
// to the host.  Input and output are raw little-endian files written and read
// by tests/test_gpu_cpp_api.py:
//   in:  grasph_params, int64 n, int32 steps, then posm f32[n][4], vel f32[n][4],
//        density f32[n], dstress f32[6][n]
//   out: posm, vel, acc f32[n][4], density, balsara, maxvsig f32[n],
//        dstress, dstress_dt f32[6][n]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "grasph2/particles.hpp"

using HostT = HostParticleBuffer<HOST_POSM, HOST_VEL, HOST_ACC, HOST_BALSARA, HOST_XVEL, HOST_DENSITY,
                                 HOST_DENSITY_DT, HOST_DSTRESS, HOST_DSTRESS_DT, HOST_MAXVSIG>;
using DevT = DeviceParticleBuffer<DEV_POSM, DEV_POS64, DEV_VEL, DEV_ACC, DEV_BALSARA, DEV_XVEL, DEV_DENSITY,
                                  DEV_DENSITY_DT, DEV_DSTRESS, DEV_DSTRESS_DT, DEV_MAXVSIG>;

static void rd(FILE* f, void* p, size_t b)
{
    if (fread(p, 1, b, f) != b) { fprintf(stderr, "short read\n"); exit(2); }
}
static void wr(FILE* f, const void* p, size_t b) { fwrite(p, 1, b, f); }

int main(int argc, char** argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
    FILE* fi = fopen(argv[1], "rb");
    if (!fi) return 2;
    grasph_params prm;
    int64_t n;
    int32_t steps;
    rd(fi, &prm, sizeof prm);
    rd(fi, &n, sizeof n);
    rd(fi, &steps, sizeof steps);
    std::vector<float> posm(4 * n), vel(4 * n), dens(n), ds(6 * n);
    rd(fi, posm.data(), 16 * n); rd(fi, vel.data(), 16 * n); rd(fi, dens.data(), 4 * n); rd(fi, ds.data(), 24 * n);
    fclose(fi);

    grasph::settings() = prm;
    HostT hpb(n);
    for (int64_t i = 0; i < n; ++i) {  // through the particle interface, as an initial-condition generator would
        Particle<POS, MASS, VEL, DENSITY, DSTRESS> p;
        p.pos = {posm[4 * i], posm[4 * i + 1], posm[4 * i + 2]};
        p.mass = posm[4 * i + 3];
        p.vel = {vel[4 * i], vel[4 * i + 1], vel[4 * i + 2]};
        p.density = dens[i];
        m3_t s;
        s[0][0] = ds[i]; s[1][1] = ds[n + i]; s[2][2] = ds[2 * n + i];
        s[0][1] = s[1][0] = ds[3 * n + i];
        s[0][2] = s[2][0] = ds[4 * n + i];
        s[1][2] = s[2][1] = ds[5 * n + i];
        p.dstress = s;
        hpb.storeParticle(i, p);
    }
    int rc = 0;
    {
    DevT pb(n);
    pb = hpb;
    try {
        for (int s = 0; s < steps; ++s) {
            computeDerivatives(pb);
            integrate(pb, true);
        }
        if (steps == 0) computeDerivatives(pb);
        grasph::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        rc = 1;
    }
    if (rc == 0) hpb = pb;
    }  // device buffer freed here
    if (rc != 0) {
        grasph::shutdown();
        return rc;
    }
    grasph_particles v = hpb.view();
    FILE* fo = fopen(argv[2], "wb");
    if (!fo) return 2;
    wr(fo, v.posm, 16 * n); wr(fo, v.vel, 16 * n); wr(fo, v.acc, 16 * n);
    wr(fo, v.density, 4 * n); wr(fo, v.balsara, 4 * n); wr(fo, v.maxvsig, 4 * n);
    wr(fo, v.dstress, 24 * n); wr(fo, v.dstress_dt, 24 * n);
    fclose(fo);
    // a loaded particle round-trips through the interface
    auto p = hpb.loadParticle<POS, ACC>(0);
    printf("timestep %g particle0 acc %g %g %g\n", (double)getCurrentTimestep(), p.acc.x, p.acc.y, p.acc.z);
    grasph::shutdown();
    return 0;
}
