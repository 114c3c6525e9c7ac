// LoadObj and MakeMesh of the C++ facade on the host (no device call):
// prints the parsed vertices, the triangles and the mesh bounds as JSON.
#include <cstdio>

#include "bubbles/bubbles.h"

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    try {
        ParsedMesh *mesh = LoadObj(argv[1]);
        std::printf("{\"name\": \"%s\", \"vertices\": %d, \"triangles\": [", mesh->name, mesh->nVertices);
        for (int t = 0; t < mesh->nTriangles; t++)
            std::printf("%s[%d, %d, %d]", t ? ", " : "", mesh->indices[3 * (size_t)t],
                        mesh->indices[3 * (size_t)t + 1], mesh->indices[3 * (size_t)t + 2]);
        Shape *s = MakeMesh(mesh, Translate(1.0, 2.0, 3.0) * Scale(2.0));
        Bounds3f b = s->GetBounds();
        std::printf("], \"bounds\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g]}\n", b.pMin.x, b.pMin.y, b.pMin.z,
                    b.pMax.x, b.pMax.y, b.pMax.z);
    } catch (const bubbles_error &e) {
        std::printf("{\"error\": \"%s\"}\n", e.what());
        return 3;
    }
    return 0;
}
