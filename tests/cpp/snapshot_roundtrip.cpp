// Snapshot files through the C++ layer (include/grasph2/storage.hpp), host
// only (no GPU): writes a HostParticleBuffer in GraSPH2's text format, reads
// it back with ps::TextFile, and queues one through ResultStorageManager.
//
// usage: snapshot_roundtrip <dir>
//   <dir>/direct.tsv   writeTextSnapshot of the buffer built below
//   <dir>/reread.tsv   the same buffer after a read through ps::TextFile
//   <dir>/snap_*.tsv   one snapshot through ResultStorageManager at t = 0.25
//   stdout: the manager's file name
#include <cstdio>
#include <iostream>

#include <grasph2/storage.hpp>

using Buf = HostParticleBuffer<HOST_POSM, HOST_VEL, HOST_DENSITY, HOST_DSTRESS, HOST_MAXVSIG>;
using Row = Particle<POS, MASS, VEL, DENSITY, DSTRESS, MAXVSIG>;

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s dir\n", argv[0]);
        return 2;
    }
    const std::string dir = std::string(argv[1]) + "/";
    const size_t n = 5;
    Buf h(n);
    for (size_t i = 0; i < n; ++i) {
        Row p;
        const float f = (float)i;
        p.pos = {0.1f * f - 0.25f, 1.0f / 3.0f + f, -2.5e-3f * f};
        p.mass = 1.0f / (1.0f + f);
        p.vel = {f, -f * 0.5f, 1e-7f * f};
        p.density = 1000.0f + f;
        m3_t s;
        s(0) = 0.5f; s(4) = -0.25f * f; s(8) = 0.125f;
        s(1) = s(3) = 0.01f * f;  // symmetric: the storage keeps the upper triangle
        s(2) = s(6) = -0.02f;
        s(5) = s(7) = 3.0f;
        p.dstress = s;
        p.max_vsig = 2.0f * f;
        h.storeParticle(i, p);
    }
    grasph::writeTextSnapshot(dir + "direct.tsv", h);

    ps::TextFile<Row> src(dir + "direct.tsv");
    if (src.size() != n) {
        std::fprintf(stderr, "read %zu particles, expected %zu\n", src.size(), n);
        return 1;
    }
    Buf back;
    src.load(back);
    grasph::writeTextSnapshot(dir + "reread.tsv", back);

    grasph::ResultStorageManager<Buf> store(dir, "snap_", 4);
    store.printToFile(h, 0.25f);
    store.flush();
    std::cout << store.filename(0.25f) << "\n";
    return 0;
}
