// Initial conditions through the C++ layer (include/grasph2/initial_conditions.hpp),
// host only (no GPU).  Writes what the generators made as raw little-endian
// float32 rows of 8 values (x y z m vx vy vz rho) for the Python checks.
//
// usage: ic_generate <dir>
//   uniform.bin   UniformSphere(20000, R 2, M 3, rho0 5, seed 11)
//   plummer.bin   PlummerSphere(20000, a 1, cut-off 10, M 1, rho0 1, seed 12)
//                   .addRandomPlummerVelocity(G 1, seed 13)
//   chain_in.tsv  64 particles (x y z m vx vy vz) built below, as a snapshot
//   chain.bin     chain_in.tsv through loadParticlesFromFile, modifiers
//                   move, rotate, addAngularVelocity, addLinearVelocity,
//                   setConstant<DENSITY>, with both centre-of-mass corrections
//   chain_desired.txt  the source's desired COM and COM velocity (6 values)
//   noise.bin / plain.bin  UniformSphere(1000, seed 5) with and without
//                   addPositionalNoise(0.01, 6).addRandomVelocity(0.2, 7)
//   two.bin       UniformSphere(100) then PlummerSphere(50).move(3, 0, 0)
//   stdout        "generate_ms <ms>" for 1,048,576 particles of a rotating
//                 uniform sphere with both corrections
// Exit status 1 on a failed in-program check (csv read == tsv read,
// determinism of seeded sources, HDF5 rejection).
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <grasph2/initial_conditions.hpp>
#include <grasph2/storage.hpp>

using Buf = HostParticleBuffer<HOST_POSM, HOST_VEL, HOST_DENSITY>;

static std::vector<float> rows(const Buf& b)
{
    std::vector<float> out;
    out.reserve(b.size() * 8);
    for (size_t i = 0; i < b.size(); ++i) {
        const auto p = b.loadParticle<POS, MASS, VEL, DENSITY>(i);
        const float r[8] = {p.pos.x, p.pos.y, p.pos.z, p.mass, p.vel.x, p.vel.y, p.vel.z, p.density};
        out.insert(out.end(), r, r + 8);
    }
    return out;
}

static void dump(const std::string& name, const Buf& b)
{
    const auto v = rows(b);
    std::ofstream f(name, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
    if (!f) throw std::runtime_error("could not write " + name);
}

static int fail(const char* what)
{
    std::fprintf(stderr, "check failed: %s\n", what);
    return 1;
}

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s dir\n", argv[0]);
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/";

    {
        InitGenerator<Buf> g;
        g.addParticles(ps::UniformSphere(20000, 2.0f, 3.0f, 5.0f, 11));
        dump(dir + "uniform.bin", g.generate());
    }
    {
        InitGenerator<Buf> g;
        g.addParticles(ps::PlummerSphere(20000, 1.0f, 10.0f, 1.0f, 1.0f, 12).addRandomPlummerVelocity(1.0f, 13));
        dump(dir + "plummer.bin", g.generate());
    }

    // modifier chain on a file source
    {
        HostParticleBuffer<HOST_POSM, HOST_VEL> in(64);
        for (size_t i = 0; i < 64; ++i) {
            Particle<POS, MASS, VEL> p;
            const float f = (float)i;
            p.pos = {0.03f * f - 1.0f, 0.5f - 0.011f * f * f / 64.0f, 0.2f * (float)(i % 7) - 0.6f};
            p.mass = 0.5f + 0.01f * (float)(i % 13);
            p.vel = {0.1f * (float)(i % 5), -0.05f * f / 64.0f, 0.02f};
            in.storeParticle(i, p);
        }
        grasph::writeTextSnapshot(dir + "chain_in.tsv", in);
        // the same rows comma-separated
        {
            std::ifstream t(dir + "chain_in.tsv");
            std::string text((std::istreambuf_iterator<char>(t)), std::istreambuf_iterator<char>());
            for (auto& c : text)
                if (c == '\t') c = ',';
            std::ofstream(dir + "chain_in.csv") << text;
        }
        auto chain = [](auto& src) -> auto& {
            return src.move(f3_t{0.5f, -1.0f, 2.0f})
                .rotate(f3_t{1.0f, 2.0f, 3.0f}, 0.7f)
                .addAngularVelocity(f3_t{0.1f, -0.2f, 0.3f})
                .addLinearVelocity(f3_t{1.0f, 0.0f, -1.0f})
                .template setConstant<DENSITY>(2.5f);
        };
        ps::TextFile<Particle<POS, MASS, VEL>> src(dir + "chain_in.tsv");
        chain(src);
        {
            std::ofstream d(dir + "chain_desired.txt");
            d.precision(9);
            const f3_t c = src.getDesiredCOM(), v = src.getDesiredVelOfCOM();
            d << c.x << " " << c.y << " " << c.z << " " << v.x << " " << v.y << " " << v.z << "\n";
        }
        InitGenerator<Buf> g;
        g.addParticles(src, true, true);
        const Buf out = g.generate();
        dump(dir + "chain.bin", out);

        // loadParticlesFromFile picks the separator by file ending
        InitGenerator<Buf> gt, gc;
        loadParticlesFromFile<Particle<POS, MASS, VEL>>(dir + "chain_in.tsv", gt);
        loadParticlesFromFile<Particle<POS, MASS, VEL>>(dir + "chain_in.csv", gc);
        if (gt.numberOfParticles() != 64 || gc.numberOfParticles() != 64) return fail("file particle count");
        if (rows(gt.generate()) != rows(gc.generate())) return fail("csv read differs from tsv read");
        bool threw = false;
        try {
            loadParticlesFromFile<Particle<POS, MASS, VEL>>(dir + "x.h5", gt);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        if (!threw) return fail("hdf5 input accepted");
    }

    // random modifiers: same seeds give the same particles
    {
        InitGenerator<Buf> a, b, c;
        a.addParticles(ps::UniformSphere(1000, 1.0f, 1.0f, 1.0f, 5).addPositionalNoise(0.01f, 6).addRandomVelocity(0.2f, 7));
        b.addParticles(ps::UniformSphere(1000, 1.0f, 1.0f, 1.0f, 5).addPositionalNoise(0.01f, 6).addRandomVelocity(0.2f, 7));
        c.addParticles(ps::UniformSphere(1000, 1.0f, 1.0f, 1.0f, 5));
        const Buf na = a.generate();
        if (rows(na) != rows(b.generate())) return fail("seeded sources not deterministic");
        dump(dir + "noise.bin", na);
        dump(dir + "plain.bin", c.generate());
    }

    // two sources, in order
    {
        InitGenerator<Buf> g;
        g.addParticles(ps::UniformSphere(100, 1.0f, 2.0f, 3.0f, 21));
        g.addParticles(ps::PlummerSphere(50, 0.5f, 2.0f, 4.0f, 6.0f, 22).move(f3_t{3.0f, 0.0f, 0.0f}));
        if (g.numberOfParticles() != 150) return fail("two-source count");
        dump(dir + "two.bin", g.generate());
    }

    // generation rate, the setup of a rotating uniform sphere
    {
        const auto t0 = std::chrono::steady_clock::now();
        InitGenerator<Buf> g;
        g.addParticles(ps::UniformSphere(1048576, 1.0f, 1.0f, 1.0f, 31).addAngularVelocity(f3_t{0, 0, 0.3f}), true, true);
        const Buf big = g.generate();
        const auto t1 = std::chrono::steady_clock::now();
        if (big.size() != 1048576) return fail("big size");
        std::printf("generate_ms %.1f\n", std::chrono::duration<double, std::milli>(t1 - t0).count());
    }
    return 0;
}
