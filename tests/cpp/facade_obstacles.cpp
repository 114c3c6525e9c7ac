// Obstacles in the reference's style (cf. the moving-box tank scene of
// src/tests/test_pcisph_extra.cpp:544-620): a sphere, a rotating box and an
// OBJ mesh inside the container; the box turns 5 degrees per frame
// (Shape::Update) and the sphere is switched off after eight frames
// (ColliderSet3::SetActive).  Prints one JSON line; writes the final
// positions (doubles, id order) to the given path.
#include <cstdio>
#include <cstdlib>

#include "bubbles/bubbles.h"

int main(int argc, char **argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: facade_obstacles <frames> <mesh.obj> <out.bin>\n");
        return 2;
    }
    const int frames = std::atoi(argv[1]);
    const Float spacing = 0.01;
    const Float h = 2.0 * spacing;
    const Float ts = spacing * std::pow(2.0, 1.0 / 3.0);
    try {
        Shape *container = MakeBox(Transform(), vec3f(0.5, 0.4, 0.5), true);
        Shape *ball = MakeSphere(Translate(0.0, -0.12, 0.0), 0.05);
        Shape *paddle = MakeBox(Translate(-0.06, -0.15, -0.06) * RotateZ(25), vec3f(0.1, 0.03, 0.1));
        Shape *block = MakeMesh(argv[2], Transform());

        ParticleSetBuilder3 builder;
        const vec3f origin(-0.08, -0.02, -0.08);
        for (int z = 0; z < 16; z++)
            for (int y = 0; y < 12; y++)
                for (int x = 0; x < 16; x++)
                    builder.AddParticle(vec3f(x * spacing + origin.x, y * spacing + origin.y, z * spacing + origin.z),
                                        vec3f(0.0, -1.5, 0.0));
        Grid3 *grid = UtilBuildGridForDomain(container->GetBounds(), h, 1.0);
        ColliderSetBuilder3 cb;
        cb.AddCollider3(ball);
        cb.AddCollider3(paddle, 0.1);
        cb.AddCollider3(block);
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
            if (f == 8) colliders->SetActive(0, false);
            paddle->Update(Translate(-0.06, -0.15, -0.06) * RotateZ(25 + 5 * f));
            solver.Advance(1.0 / 240.0);
            substeps += solver.lastSubsteps;
        }
        ParticleSet3 *pSet = sphSet->GetParticleSet();
        const int n = pSet->GetParticleCount();
        std::vector<double> out(3 * (size_t)n);
        double miny = 1e300;
        for (int i = 0; i < n; i++) {
            vec3f p = pSet->GetParticlePosition(i);
            out[3 * (size_t)i] = p.x;
            out[3 * (size_t)i + 1] = p.y;
            out[3 * (size_t)i + 2] = p.z;
            miny = std::fmin(miny, p.y);
        }
        FILE *fp = std::fopen(argv[3], "wb");
        if (!fp || std::fwrite(out.data(), sizeof(double), out.size(), fp) != out.size()) {
            std::fprintf(stderr, "cannot write %s\n", argv[3]);
            return 4;
        }
        std::fclose(fp);
        Bounds3f mb = block->GetBounds();
        std::printf("{\"n\": %d, \"substeps\": %d, \"min_y\": %.17g, \"mesh_triangles\": %d, "
                    "\"mesh_bounds\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g]}\n",
                    n, substeps, miny, block->mesh->nTriangles, mb.pMin.x, mb.pMin.y, mb.pMin.z, mb.pMax.x,
                    mb.pMax.y, mb.pMax.z);
    } catch (const bubbles_error &e) {
        std::fprintf(stderr, "bubbles error: %s\n", e.what());
        return 3;
    }
    return 0;
}
