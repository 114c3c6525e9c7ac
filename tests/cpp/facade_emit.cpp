// Continuous emission in the reference's style (cf. the emitter scene of
// src/tests/test_pcisph_extra.cpp:544-620): a ContinuousParticleSetBuilder3
// seeded with one block of particles, its cells mapped with MapGrid, and
// MapGridEmit after every frame refilling the mapped cells wherever the fluid
// has left them.  Prints one JSON line (particle count after every frame);
// with a path argument also writes the final positions (doubles, id order).
#include <cstdio>
#include <cstdlib>

#include "bubbles/bubbles.h"

int main(int argc, char **argv) {
    const int frames = argc > 1 ? std::atoi(argv[1]) : 16;
    const int maxParticles = argc > 2 ? std::atoi(argv[2]) : 576;
    const Float spacing = 0.01;
    const Float h = 2.0 * spacing;
    const Float ts = spacing * std::pow(2.0, 1.0 / 3.0);
    try {
        Shape *container = MakeBox(Transform(), vec3f(0.5, 0.25, 0.5), true);
        Grid3 *grid = UtilBuildGridForDomain(container->GetBounds(), h, 1.0);
        ColliderSetBuilder3 cb;
        cb.AddCollider3(container);

        auto velocityField = [&](const vec3f &p) -> vec3f { return vec3f(0.5 * p.x, -1.0, 0.0); };

        ContinuousParticleSetBuilder3 pBuilder(maxParticles);
        pBuilder.SetKernelRadius(h);
        // the emitter: a 6x4x6 lattice block near the top of the tank
        const vec3f corner(-0.03, 0.05, -0.03);
        for (int z = 0; z < 6; z++)
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 6; x++) {
                    vec3f p(corner.x + x * spacing, corner.y + y * spacing, corner.z + z * spacing);
                    pBuilder.AddParticle(p, velocityField(p));
                }
        pBuilder.Commit();

        SphParticleSet3 *sphSet = SphParticleSet3FromContinuousBuilder(&pBuilder);
        ResetAndDistribute(grid, sphSet->GetParticleSet());
        pBuilder.MapGrid(grid);

        PciSphSolver3 solver;
        solver.Initialize(DefaultSphSolverData3());
        solver.SetColliders(cb.GetColliderSet());
        solver.Setup(1000.0, ts, h / ts, grid, sphSet);
        solver.loopMode = BBX_MODE_TEXTBOOK;
        solver.maxIterations = 8;

        ParticleSet3 *pSet = sphSet->GetParticleSet();
        std::vector<int> counts;
        for (int f = 0; f < frames; f++) {
            solver.Advance(1.0 / 240.0);
            // odd frames through MapGridEmitToOther with the builder as its own
            // target (the same refill, without the early exit once warned)
            if (f % 2 == 0)
                pBuilder.MapGridEmit(velocityField, spacing);  // may print the capacity warning
            else
                pBuilder.MapGridEmitToOther(&pBuilder, velocityField, spacing);
            counts.push_back(pSet->GetParticleCount());
        }
        std::printf("\n{\"mapped\": %zu, \"counts\": [", pBuilder.mappedCellSet.size());
        for (size_t f = 0; f < counts.size(); f++) std::printf("%s%d", f ? ", " : "", counts[f]);
        std::printf("], \"warned\": %d}\n", pBuilder.warned ? 1 : 0);
        if (argc > 3) {
            const int n = pSet->GetParticleCount();
            std::vector<double> out(3 * (size_t)n);
            for (int i = 0; i < n; i++) {
                vec3f p = pSet->GetParticlePosition(i);
                out[3 * (size_t)i] = p.x;
                out[3 * (size_t)i + 1] = p.y;
                out[3 * (size_t)i + 2] = p.z;
            }
            FILE *fp = std::fopen(argv[3], "wb");
            if (!fp || std::fwrite(out.data(), sizeof(double), out.size(), fp) != out.size()) {
                std::fprintf(stderr, "cannot write %s\n", argv[3]);
                return 4;
            }
            std::fclose(fp);
        }
    } catch (const bubbles_error &e) {
        std::fprintf(stderr, "bubbles error: %s\n", e.what());
        return 3;
    }
    return 0;
}
