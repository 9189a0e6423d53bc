// ic_device_roundtrip.cpp -- an end-to-end run through the C++ layer as
// src/main.cu:191-266 drives the reference: initial conditions from the
// InitGenerator and particle sources (include/grasph2/initial_conditions.hpp),
// uploaded into a DeviceParticleBuffer, stepped on the GPU with
// computeDerivatives(pb); integrate(pb, true), then stored with
// ResultStorageManager::printToFile(DeviceParticleBuffer) (the device-side
// duplicate, downloaded and written by the worker thread) and read back
// through ps::TextFile.
//
// usage: ic_device_roundtrip <dir> <params.bin> <plummer|collision> <n> <steps>
//   params.bin  a grasph_params (the settings of the run)
// writes to <dir>:
//   ic.bin      the generated initial conditions, float32 rows x y z m vx vy vz rho
//   after.bin   the state after the steps, downloaded: posm f32[n][4],
//               vel f32[n][4], acc f32[n][4], density f32[n], pos64 f64[3][n]
//   snap_*.tsv  the snapshot printToFile wrote after the steps
// Exit status 1 when the snapshot file, read back through ps::TextFile, does
// not match the downloaded state to its printed precision.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include <grasph2/initial_conditions.hpp>
#include <grasph2/storage.hpp>

using IcBuf = HostParticleBuffer<HOST_POSM, HOST_VEL, HOST_DENSITY>;
using DevT = DeviceParticleBuffer<DEV_POSM, DEV_POS64, DEV_VEL, DEV_ACC, DEV_BALSARA, DEV_XVEL, DEV_DENSITY,
                                  DEV_DENSITY_DT, DEV_DSTRESS, DEV_DSTRESS_DT, DEV_MAXVSIG>;
using OutBuf = HostParticleBuffer<HOST_POSM, HOST_POS64, HOST_VEL, HOST_ACC, HOST_DENSITY>;

static int fail(const std::string& what)
{
    std::fprintf(stderr, "check failed: %s\n", what.c_str());
    return 1;
}

static bool write_file(const std::string& name, const void* p, size_t bytes)
{
    std::ofstream f(name, std::ios::binary);
    f.write(static_cast<const char*>(p), (std::streamsize)bytes);
    return (bool)f;
}

static int run(const std::string& dir, const std::string& kind, size_t n, int steps)
{
    // initial conditions (main.cu:191-228)
    InitGenerator<IcBuf> gen;
    if (kind == "plummer") {
        gen.addParticles(ps::PlummerSphere(n, 1.0f, 10.0f, 1.0f, 1.0f, 41).addRandomPlummerVelocity(1.0f, 42), true,
                         true);
    } else {
        // two uniform spheres of radius 1 with a gap of 0.1, closing at 0.5
        // along an axis 30 degrees off the centre line
        const size_t na = n / 2;
        const f1_t c = std::cos(0.5235988f), s = std::sin(0.5235988f);
        gen.addParticles(ps::UniformSphere(na, 1.0f, (f1_t)na, 1.0f, 43)
                             .move(f3_t{-1.05f, 0.0f, 0.0f})
                             .addLinearVelocity(f3_t{0.25f * c, 0.25f * s, 0.0f}),
                         true, true);
        gen.addParticles(ps::UniformSphere(n - na, 1.0f, (f1_t)(n - na), 1.0f, 44)
                             .move(f3_t{1.05f, 0.0f, 0.0f})
                             .addLinearVelocity(f3_t{-0.25f * c, -0.25f * s, 0.0f}),
                         true, true);
    }
    if (gen.numberOfParticles() != n) return fail("particle count");
    const IcBuf ic = gen.generate();
    std::vector<float> rows;
    rows.reserve(8 * n);
    for (size_t i = 0; i < n; ++i) {
        const auto p = ic.loadParticle<POS, MASS, VEL, DENSITY>(i);
        const float r[8] = {p.pos.x, p.pos.y, p.pos.z, p.mass, p.vel.x, p.vel.y, p.vel.z, p.density};
        rows.insert(rows.end(), r, r + 8);
    }
    if (!write_file(dir + "ic.bin", rows.data(), rows.size() * sizeof(float))) return fail("write ic.bin");

    // upload, step, store (main.cu:228, 256-257, 266)
    DevT pb(n);
    pb = ic;
    double t = 0.0;
    for (int k = 0; k < steps; ++k) {
        computeDerivatives(pb);
        integrate(pb, true);
        t += getCurrentTimestep();
    }
    grasph::ResultStorageManager<> storage(dir, "snap_", 4);
    storage.printToFile(pb, (f1_t)t);
    OutBuf out;
    out = pb;  // the state the snapshot must hold
    // the simulation goes on at once; the snapshot keeps the state above
    computeDerivatives(pb);
    integrate(pb, true);
    storage.flush();
    const std::string snap = storage.filename((f1_t)t);
    const grasph_particles v = out.view();
    std::vector<unsigned char> blob;
    auto put = [&](const void* p, size_t b) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        blob.insert(blob.end(), c, c + b);
    };
    put(v.posm, 16 * n);
    put(v.vel, 16 * n);
    put(v.acc, 16 * n);
    put(v.density, 4 * n);
    put(v.pos64, 24 * n);
    if (!write_file(dir + "after.bin", blob.data(), blob.size())) return fail("write after.bin");

    // the snapshot read back (TextFile.h:69-165): x y z m vx vy vz rho
    ps::TextFile<Particle<POS, MASS, VEL, DENSITY>> src(snap);
    if (src.getNumParticles() != n) return fail("snapshot line count of " + snap);
    InitGenerator<IcBuf> back;
    back.addParticles(src);
    const IcBuf read = back.generate();
    for (size_t i = 0; i < n; ++i) {
        const auto a = read.loadParticle<POS, MASS, VEL, DENSITY>(i);
        const auto b = out.loadParticle<POS, MASS, VEL, DENSITY>(i);
        const float ga[8] = {a.pos.x, a.pos.y, a.pos.z, a.mass, a.vel.x, a.vel.y, a.vel.z, a.density};
        // the snapshot prints the FP32 mirror (posm), 7 decimals
        const float gb[8] = {v.posm[4 * i], v.posm[4 * i + 1], v.posm[4 * i + 2], b.mass, b.vel.x, b.vel.y, b.vel.z,
                             b.density};
        for (int c = 0; c < 8; ++c)
            if (!(std::fabs(ga[c] - gb[c]) <= 5.1e-8f + 1.2e-7f * std::fabs(gb[c])))
                return fail("snapshot value " + std::to_string(c) + " of particle " + std::to_string(i) + ": " +
                            std::to_string(ga[c]) + " vs " + std::to_string(gb[c]));
    }
    std::printf("snapshot %s: %zu particles match, t = %.9g\n", snap.c_str(), n, t);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s dir params.bin plummer|collision n steps\n", argv[0]);
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/";
    {
        std::ifstream f(argv[2], std::ios::binary);
        if (!f.read(reinterpret_cast<char*>(&grasph::settings()), sizeof(grasph_params))) return fail("params");
    }
    int rc = 0;
    try {
        rc = run(dir, argv[3], (size_t)std::atol(argv[4]), std::atoi(argv[5]));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        rc = 1;
    }
    grasph::shutdown();
    return rc;
}
