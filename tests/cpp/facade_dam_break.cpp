// Scene code in the reference's style (cf. src/tests/test_pcisph3.cpp:1785-1856)
// against the C++ facade: builder -> SphParticleSet3 -> grid -> PciSphSolver3
// -> Advance.  Prints one JSON line with what the parity test compares.
#include <cstdio>
#include <cstdlib>

#include "bubbles/bubbles.h"

int main(int argc, char **argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 3;
    const Float spacing = 0.01;
    const Float h = 2.0 * spacing;
    const Float ts = spacing * std::pow(2.0, 1.0 / 3.0);  // BCC-calibrated target spacing
    try {
        Shape *container = MakeBox(Transform(), vec3f(0.5, 0.25, 0.5), true);
        ParticleSetBuilder3 builder;
        Bounds3f b = container->GetBounds();
        const vec3f origin(b.pMin.x + ts + 0.0, b.pMin.y + ts + 0.0, b.pMin.z + ts + 0.0);
        for (int z = 0; z < 16; z++)
            for (int y = 0; y < 8; y++)
                for (int x = 0; x < 16; x++)
                    builder.AddParticle(vec3f(x * spacing + origin.x, y * spacing + origin.y,
                                              z * spacing + origin.z));
        Grid3 *grid = UtilBuildGridForDomain(container->GetBounds(), h, 1.0);
        ColliderSetBuilder3 cb;
        cb.AddCollider3(container);
        ColliderSet3 *colliders = cb.GetColliderSet();
        SphParticleSet3 *sphSet = SphParticleSet3FromBuilder(&builder);
        PciSphSolver3 solver;
        solver.Initialize(DefaultSphSolverData3());
        solver.SetColliders(colliders);
        solver.Setup(1000.0, ts, h / ts, grid, sphSet);
        solver.loopMode = BBX_MODE_TEXTBOOK;
        solver.maxIterations = 8;
        int substeps = 0;
        for (int f = 0; f < frames; f++) {
            solver.Advance(1.0 / 240.0);
            substeps += solver.lastSubsteps;
        }
        ParticleSet3 *pSet = sphSet->GetParticleSet();
        double sy = 0, sd = 0;
        for (int i = 0; i < pSet->GetParticleCount(); i++) {
            sy += pSet->GetParticlePosition(i).y;
            sd += pSet->GetParticleDensity(i);
        }
        // frame dump in the reference's serializer format, read back
        // through the loader into a builder
        int reread = -1;
        if (argc > 2) {
            std::remove(argv[2]);
            SaveSphParticleSet(solver.GetSphSolverData(), argv[2],
                               SERIALIZER_POSITION | SERIALIZER_VELOCITY | SERIALIZER_DENSITY);
            ParticleSetBuilder3 back;
            int flags = 0;
            reread = SerializerLoadSphDataSet3(&back, argv[2], flags);
            if (flags != (SERIALIZER_POSITION | SERIALIZER_VELOCITY | SERIALIZER_DENSITY)) reread = -2;
        }
        Bucket *bk = pSet->GetParticleBucket(0);
        vec3ui dims = grid->GetIndexCount();
        std::printf("{\"n\": %d, \"substeps\": %d, \"mass\": %.17g, \"delta_denom\": %.17g, "
                    "\"grid\": [%u, %u, %u], \"sum_y\": %.17g, \"sum_density\": %.17g, "
                    "\"bucket0\": [",
                    pSet->GetParticleCount(), substeps, pSet->GetMass(), solver.ComputeDeltaDenom(),
                    dims.x, dims.y, dims.z, sy, sd);
        for (int k = 0; k < bk->Count(); k++) std::printf("%s%d", k ? ", " : "", bk->Get(k));
        std::printf("], \"hash0\": %u, \"reread\": %d}\n",
                    grid->GetLinearHashedPosition(pSet->GetParticlePosition(0)), reread);
    } catch (const bubbles_error &e) {
        std::fprintf(stderr, "bubbles error: %s\n", e.what());
        return 3;
    }
    return 0;
}
