/*
 * bubbles/bubbles.h -- C++ facade of the MI355X SPH / PCISPH hot path with
 * the reference's own class and function names, so scene code written
 * against Bubbles' solver API compiles against this library.
 *
 * Mirrors (reference file:line):
 *   vec3f / vec3ui / Float / Bounds3f ......... src/core/geometry.h:69,497-650,1463-1615
 *   Bucket ..................................... src/core/particle.h:22-60
 *   ParticleSet3, SphParticleSet3 ............. src/core/particle.h:154-401,508-609
 *   ParticleSetBuilder3, SphParticleSet3FromBuilder  src/core/particle.h:611-710, src/core/particle.cpp:21-27
 *   Grid3, UtilBuildGridForDomain .............. src/core/grid.h:197-600, src/core/util.cpp:269-287
 *   Transform, Translate/Scale/RotateX|Y|Z ..... src/core/transform.h
 *   Shape, MakeBox, MakeSphere, MakeMesh ....... src/core/shape.h:125-290, src/shapes/box.cpp:6-10,
 *                                                src/shapes/sphere.cpp:5-9, src/shapes/bvh.cpp:51-61
 *   ParsedMesh, LoadObj, GenerateShapeSDF ...... src/core/shape.h:12-22, src/third/obj_loader.cpp:434-441,
 *                                                src/core/shape.cpp:478-505
 *   Collider3, ColliderSetBuilder3, ColliderSet3  src/core/collider.h, src/core/collider.cpp:300-353
 *   SphSolverData3, DefaultSphSolverData3 ...... src/core/sph_solver.h:31-50, src/solvers/sph_solver3.cpp:124-151
 *   PciSphSolver3 .............................. src/core/pcisph_solver.h:56-83
 *   SphSolver3 ................................. src/core/sph_solver.h:76-96
 *
 * Differences from the reference, by design:
 *   - data lives in device memory owned by the solver; per-particle
 *     accessors download lazily (cached until the next step) and setters
 *     upload immediately;
 *   - errors throw bubbles_error instead of printing and exiting;
 *   - colliders are boxes, spheres and meshes (meshes through their SDF);
 *   - PciSphSolver3 exposes the predict-correct loop settings (the reference
 *     hard-codes maxIt 5, tol 0.01 and, in effect, one iteration).
 * The device path is FP32 (+ Kahan); Float stays double at this API.
 */
#ifndef BUBBLES_FACADE_H
#define BUBBLES_FACADE_H

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../bubbles_amd.h"

typedef double Float;

struct bubbles_error : std::runtime_error {
    explicit bubbles_error(const std::string &m) : std::runtime_error(m) {}
};

inline void bbx_throw_if(int rc, const char *what) {
    if (rc != BBX_OK) throw bubbles_error(std::string(what) + ": " + bbx_last_error());
}

struct vec3f {
    Float x, y, z;
    vec3f() : x(0), y(0), z(0) {}
    explicit vec3f(Float a) : x(a), y(a), z(a) {}
    vec3f(Float a, Float b, Float c) : x(a), y(b), z(c) {}
    Float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    Float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    vec3f operator+(const vec3f &v) const { return vec3f(x + v.x, y + v.y, z + v.z); }
    vec3f operator-(const vec3f &v) const { return vec3f(x - v.x, y - v.y, z - v.z); }
    vec3f operator*(Float s) const { return vec3f(x * s, y * s, z * s); }
    Float LengthSquared() const { return x * x + y * y + z * z; }
    Float Length() const { return std::sqrt(LengthSquared()); }
};
inline vec3f operator*(Float s, const vec3f &v) { return v * s; }

struct vec3ui {
    unsigned int x, y, z;
    vec3ui() : x(0), y(0), z(0) {}
    vec3ui(unsigned a, unsigned b, unsigned c) : x(a), y(b), z(c) {}
    unsigned operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct Bounds3f {
    vec3f pMin, pMax;
    Bounds3f() {}
    Bounds3f(const vec3f &a, const vec3f &b)
        : pMin(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)),
          pMax(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)) {}
    Float ExtentOn(int i) const { return std::fabs(pMax[i] - pMin[i]); }
    vec3f Center() const { return (pMin + pMax) * 0.5; }
};

/* Transform (src/core/transform.h): a 4x4 affine matrix and its inverse */
struct Transform {
    Float m[4][4], mInv[4][4];
    Transform() {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) m[i][j] = mInv[i][j] = i == j ? 1.0 : 0.0;
    }
    Transform(const Float a[4][4], const Float ai[4][4]) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                m[i][j] = a[i][j];
                mInv[i][j] = ai[i][j];
            }
    }
    Transform operator*(const Transform &t) const {
        Float r[4][4], ri[4][4];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                r[i][j] = ri[i][j] = 0;
                for (int k = 0; k < 4; k++) {
                    r[i][j] += m[i][k] * t.m[k][j];
                    ri[i][j] += t.mInv[i][k] * mInv[k][j];
                }
            }
        return Transform(r, ri);
    }
    vec3f Point(const vec3f &p) const {
        return vec3f(m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3],
                     m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3],
                     m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3]);
    }
    /* the rows 0..2 of m and mInv, as bbx_collider wants them */
    void Rows(double *to_world, double *to_object) const {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 4; j++) {
                to_world[4 * i + j] = m[i][j];
                to_object[4 * i + j] = mInv[i][j];
            }
    }
    bool IsTranslation() const {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                if (m[i][j] != (i == j ? 1.0 : 0.0)) return false;
        return true;
    }
};
inline Transform Inverse(const Transform &t) { return Transform(t.mInv, t.m); }
inline Transform Translate(Float x, Float y, Float z) {
    Transform t;
    t.m[0][3] = x; t.m[1][3] = y; t.m[2][3] = z;
    t.mInv[0][3] = -x; t.mInv[1][3] = -y; t.mInv[2][3] = -z;
    return t;
}
inline Transform Translate(const vec3f &v) { return Translate(v.x, v.y, v.z); }
inline Transform Scale(Float x, Float y, Float z) {
    Transform t;
    t.m[0][0] = x; t.m[1][1] = y; t.m[2][2] = z;
    t.mInv[0][0] = 1 / x; t.mInv[1][1] = 1 / y; t.mInv[2][2] = 1 / z;
    return t;
}
inline Transform Scale(Float s) { return Scale(s, s, s); }
namespace bubbles_detail {
inline Transform rotation(int a, int b, Float degrees) {
    const Float th = degrees * (3.14159265358979323846 / 180.0), c = std::cos(th), s = std::sin(th);
    Transform t;
    t.m[a][a] = c; t.m[a][b] = -s; t.m[b][a] = s; t.m[b][b] = c;
    t.mInv[a][a] = c; t.mInv[a][b] = s; t.mInv[b][a] = -s; t.mInv[b][b] = c;
    return t;
}
}  // namespace bubbles_detail
inline Transform RotateX(Float deg) { return bubbles_detail::rotation(1, 2, deg); }
inline Transform RotateY(Float deg) { return bubbles_detail::rotation(2, 0, deg); }
inline Transform RotateZ(Float deg) { return bubbles_detail::rotation(0, 1, deg); }

/* ParsedMesh (src/core/shape.h:12-22): vertices and triangle indices */
struct ParsedMesh {
    std::vector<vec3f> p;
    std::vector<int> indices;  // 3 per triangle
    int nTriangles = 0, nVertices = 0;
    char name[256] = {0};
};

/* LoadObj (src/third/obj_loader.cpp:434-441): the first object's vertices
 * ("v x y z") and faces ("f a b c ...", fan-triangulated, 1-based or
 * negative indices, "a/b/c" tokens use a) of a Wavefront OBJ file. */
inline ParsedMesh *LoadObj(const char *path) {
    FILE *fp = std::fopen(path, "r");
    if (!fp) throw bubbles_error(std::string("LoadObj: cannot open ") + path);
    ParsedMesh *mesh = new ParsedMesh();
    char line[4096];
    while (std::fgets(line, sizeof(line), fp)) {
        const char *c = line;
        while (*c == ' ' || *c == '\t') c++;
        if (c[0] == 'v' && (c[1] == ' ' || c[1] == '\t')) {
            double x = 0, y = 0, z = 0;
            if (std::sscanf(c + 2, "%lf %lf %lf", &x, &y, &z) == 3) mesh->p.push_back(vec3f(x, y, z));
        } else if (c[0] == 'f' && (c[1] == ' ' || c[1] == '\t')) {
            std::vector<int> face;
            const char *q = c + 2;
            while (*q) {
                while (*q == ' ' || *q == '\t') q++;
                if (!*q || *q == '\n' || *q == '\r') break;
                char *end = nullptr;
                long v = std::strtol(q, &end, 10);
                if (end == q) break;
                const int nv = (int)mesh->p.size();
                face.push_back(v < 0 ? nv + (int)v : (int)v - 1);
                q = end;
                while (*q && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') q++;
            }
            for (size_t k = 2; k < face.size(); k++) {
                mesh->indices.push_back(face[0]);
                mesh->indices.push_back(face[k - 1]);
                mesh->indices.push_back(face[k]);
            }
        }
    }
    std::fclose(fp);
    mesh->nVertices = (int)mesh->p.size();
    mesh->nTriangles = (int)mesh->indices.size() / 3;
    const char *slash = std::strrchr(path, '/');
    std::snprintf(mesh->name, sizeof(mesh->name), "%s", slash ? slash + 1 : path);
    for (int i : mesh->indices)
        if (i < 0 || i >= mesh->nVertices) {
            delete mesh;
            throw bubbles_error(std::string("LoadObj: face index out of range in ") + path);
        }
    return mesh;
}

enum ShapeType { ShapeBox = 0, ShapeSphere = 1, ShapeMesh = 2 };

/* Shape (src/core/shape.h:125-271): box, sphere or mesh */
struct Shape {
    ShapeType type = ShapeBox;
    Transform ObjectToWorld, WorldToObject;
    bool reverseOrientation = false;
    Float sizex = 0, sizey = 0, sizez = 0, radius = 0;
    vec3f linearVelocity, angularVelocity;
    ParsedMesh *mesh = nullptr;
    bbx_sdf *grid = nullptr;  // mesh SDF on the device
    Bounds3f meshBounds;
    /* GetBounds (src/core/shape.cpp:158-180): transformed local bounds */
    Bounds3f GetBounds() const {
        if (type == ShapeMesh) return meshBounds;
        const vec3f h = type == ShapeBox ? vec3f(sizex / 2.0, sizey / 2.0, sizez / 2.0) : vec3f(radius);
        Bounds3f b(ObjectToWorld.Point(vec3f(-h.x, -h.y, -h.z)), ObjectToWorld.Point(vec3f(-h.x, -h.y, -h.z)));
        for (int c = 0; c < 8; c++) {
            const vec3f q = ObjectToWorld.Point(vec3f(c & 1 ? h.x : -h.x, c & 2 ? h.y : -h.y, c & 4 ? h.z : -h.z));
            b = Bounds3f(vec3f(std::fmin(b.pMin.x, q.x), std::fmin(b.pMin.y, q.y), std::fmin(b.pMin.z, q.z)),
                         vec3f(std::fmax(b.pMax.x, q.x), std::fmax(b.pMax.y, q.y), std::fmax(b.pMax.z, q.z)));
        }
        return b;
    }
    /* Update (src/core/shape.cpp:294-298): a new placement for the next step */
    void Update(const Transform &toWorld) {
        ObjectToWorld = toWorld;
        WorldToObject = Inverse(toWorld);
    }
    void SetVelocities(const vec3f &vel, const vec3f &angular) {
        linearVelocity = vel;
        angularVelocity = angular;
    }
};

/* MakeBox (src/shapes/box.cpp:6-10) */
inline Shape *MakeBox(const Transform &toWorld, const vec3f &size, bool reverseOrientation = false) {
    Shape *s = new Shape();
    s->type = ShapeBox;
    s->Update(toWorld);
    s->sizex = size.x; s->sizey = size.y; s->sizez = size.z;
    s->reverseOrientation = reverseOrientation;
    return s;
}

/* MakeSphere (src/shapes/sphere.cpp:5-9) */
inline Shape *MakeSphere(const Transform &toWorld, Float radius, bool reverseOrientation = false) {
    Shape *s = new Shape();
    s->type = ShapeSphere;
    s->Update(toWorld);
    s->radius = radius;
    s->reverseOrientation = reverseOrientation;
    return s;
}

/* MakeMesh (src/shapes/bvh.cpp:51-61): the transform is applied to the
 * vertices, the shape itself stays at the identity */
inline Shape *MakeMesh(ParsedMesh *mesh, const Transform &toWorld, bool reverseOrientation = false) {
    if (!mesh || mesh->nTriangles < 1) throw bubbles_error("MakeMesh: empty mesh");
    for (auto &v : mesh->p) v = toWorld.Point(v);
    Shape *s = new Shape();
    s->type = ShapeMesh;
    s->mesh = mesh;
    s->reverseOrientation = reverseOrientation;
    vec3f lo(1e300), hi(-1e300);
    for (int i : mesh->indices) {
        const vec3f &v = mesh->p[(size_t)i];
        for (int k = 0; k < 3; k++) {
            lo[k] = std::fmin(lo[k], v[k]);
            hi[k] = std::fmax(hi[k], v[k]);
        }
    }
    s->meshBounds = Bounds3f(lo, hi);
    return s;
}
inline Shape *MakeMesh(const char *path, const Transform &toWorld, bool reverseOrientation = false) {
    return MakeMesh(LoadObj(path), toWorld, reverseOrientation);
}

/* GenerateShapeSDF (src/core/shape.cpp:478-505) for meshes, on the device */
inline void GenerateShapeSDF(Shape *shape, Float dx = 0.01, Float margin = 0.1) {
    if (!shape || shape->type != ShapeMesh || !shape->mesh)
        throw bubbles_error("GenerateShapeSDF: only mesh shapes carry an SDF here");
    std::vector<double> v(3 * shape->mesh->p.size());
    for (size_t i = 0; i < shape->mesh->p.size(); i++)
        for (int k = 0; k < 3; k++) v[3 * i + k] = shape->mesh->p[i][k];
    if (shape->grid) bbx_sdf_destroy(shape->grid);
    shape->grid = nullptr;
    bbx_throw_if(bbx_sdf_from_mesh(v.data(), shape->mesh->nVertices, shape->mesh->indices.data(),
                                   shape->mesh->nTriangles, dx, margin, &shape->grid),
                 "GenerateShapeSDF");
}

/* Collider3 (src/core/collider.h:62-80) */
struct Collider3 {
    Shape *shape = nullptr;
    Float frictionCoefficient = 0;
    bool isActive = true;
    void SetActive(bool active) { isActive = active; }
    bool IsActive() const { return isActive; }
};

/* MakeCollider3 (src/core/collider.cpp:345-353): a mesh gets its SDF */
inline Collider3 *MakeCollider3(Shape *shape, Float frictionCoef = 0) {
    if (!shape) throw bubbles_error("MakeCollider3: null shape");
    if (shape->type == ShapeMesh && !shape->grid) GenerateShapeSDF(shape);
    Collider3 *c = new Collider3();
    c->shape = shape;
    c->frictionCoefficient = frictionCoef;
    return c;
}

struct ColliderSet3 {
    std::vector<Collider3 *> colliders;
    int nColiders = 0;
    void SetActive(int which, bool active) {
        if (which >= 0 && which < (int)colliders.size()) colliders[(size_t)which]->SetActive(active);
    }
    bool IsActive(int which) const {
        return which >= 0 && which < (int)colliders.size() && colliders[(size_t)which]->IsActive();
    }
};

struct ColliderSetBuilder3 {
    std::vector<Collider3 *> colliders;
    void AddCollider3(Collider3 *collider) { colliders.push_back(collider); }
    void AddCollider3(Shape *shape, Float frictionCoef = 0) { colliders.push_back(MakeCollider3(shape, frictionCoef)); }
    ColliderSet3 *GetColliderSet() {
        if (colliders.empty()) throw bubbles_error("No colliders given for {GetColliderSet}");
        ColliderSet3 *c = new ColliderSet3();
        c->colliders = colliders;
        c->nColiders = (int)colliders.size();
        return c;
    }
};

/* Grid3 geometry (Build + hash + 27-cell lists) */
class Grid3 {
  public:
    bbx_params prm{};
    bbx_grid_desc desc{};
    unsigned int GetCellCount() const { return desc.total; }
    vec3ui GetIndexCount() const { return vec3ui(desc.usizes[0], desc.usizes[1], desc.usizes[2]); }
    vec3f GetCellSize() const { return vec3f(desc.cells_len[0], desc.cells_len[1], desc.cells_len[2]); }
    Bounds3f GetBounds() const {
        return Bounds3f(vec3f(desc.bounds_min[0], desc.bounds_min[1], desc.bounds_min[2]),
                        vec3f(desc.bounds_max[0], desc.bounds_max[1], desc.bounds_max[2]));
    }
    int GetDimensions() const { return 3; }
    vec3ui GetHashedPosition(const vec3f &p) const {
        double q[3] = {p.x, p.y, p.z};
        int id, u[3];
        bbx_throw_if(bbx_grid_hash_desc(&desc, q, 1, &id, u), "GetHashedPosition");
        if (id < 0) throw bubbles_error("GetHashedPosition: point outside the grid");
        return vec3ui((unsigned)u[0], (unsigned)u[1], (unsigned)u[2]);
    }
    unsigned int GetLinearHashedPosition(const vec3f &p) const {
        double q[3] = {p.x, p.y, p.z};
        int id;
        bbx_throw_if(bbx_grid_hash_desc(&desc, q, 1, &id, nullptr), "GetLinearHashedPosition");
        if (id < 0) throw bubbles_error("GetLinearHashedPosition: point outside the grid");
        return (unsigned)id;
    }
    vec3ui GetCellIndex(unsigned int i) const {
        unsigned plane = desc.usizes[0] * desc.usizes[1];
        unsigned z = i / plane, rem = i - z * plane;
        return vec3ui(rem % desc.usizes[0], rem / desc.usizes[0], z);
    }
    /* GetNeighborsOf (src/core/grid.h:502-510): fills an internal list */
    int GetNeighborsOf(int id, int **neighbors) {
        vec3ui u = GetCellIndex((unsigned)id);
        int c = 0;
        const int nx = (int)desc.usizes[0], ny = (int)desc.usizes[1], nz = (int)desc.usizes[2];
        for (int j = (int)u.y - 1; j <= (int)u.y + 1; j++) {
            if (j < 0 || j >= ny) continue;
            for (int i = (int)u.x - 1; i <= (int)u.x + 1; i++) {
                if (i < 0 || i >= nx) continue;
                for (int k = (int)u.z - 1; k <= (int)u.z + 1; k++) {
                    if (k < 0 || k >= nz) continue;
                    list_[c++] = i + j * nx + k * nx * ny;
                }
            }
        }
        if (neighbors) *neighbors = list_;
        return c;
    }

  private:
    int list_[27];
};

/* UtilBuildGridForDomain (src/core/util.cpp:269-287) */
inline Grid3 *UtilBuildGridForDomain(Bounds3f domain, Float spacing, Float spacingScale = 2.0) {
    Grid3 *g = new Grid3();
    bbx_throw_if(bbx_default_params(&g->prm), "UtilBuildGridForDomain");
    vec3f c = domain.Center();
    for (int k = 0; k < 3; k++) {
        g->prm.tank_center[k] = c[k];
        g->prm.tank_size[k] = domain.ExtentOn(k);
    }
    g->prm.grid_spacing = spacing;
    g->prm.grid_scale = spacingScale;
    bbx_throw_if(bbx_grid_describe(&g->prm, &g->desc), "UtilBuildGridForDomain");
    return g;
}

struct Bucket {
    std::vector<int> ids;
    int Count() const { return (int)ids.size(); }
    int Get(int where) const { return where < (int)ids.size() ? ids[where] : -1; }
};

class PciSphSolver3;
class SphSolver3;

/* ParticleSet3: host copy until a solver owns the particles, then lazy
 * downloads from the device. */
class ParticleSet3 {
  public:
    std::vector<double> pos, vel;
    int count = 0;
    Float mass = 1e-3, radius = 1e-3;
    bbx_solver *dev = nullptr;

    int GetParticleCount() const { return count; }
    Float GetMass() const { return mass; }
    Float GetRadius() const { return radius; }
    vec3f GetParticlePosition(int i) { return vfield(BBX_FIELD_POSITION, i); }
    vec3f GetParticleVelocity(int i) { return vfield(BBX_FIELD_VELOCITY, i); }
    vec3f GetParticleForce(int i) { return vfield(BBX_FIELD_FORCE, i); }
    Float GetParticleDensity(int i) { return sfield(BBX_FIELD_DENSITY, i); }
    Float GetParticlePressure(int i) { return sfield(BBX_FIELD_PRESSURE, i); }
    void SetParticlePosition(int i, const vec3f &p) { setv(BBX_FIELD_POSITION, i, p); }
    void SetParticleVelocity(int i, const vec3f &v) { setv(BBX_FIELD_VELOCITY, i, v); }
    Bucket *GetParticleBucket(int i) {
        if (!dev) throw bubbles_error("GetParticleBucket: particles are not in a solver yet");
        if (buckets_.empty()) {
            std::vector<int> counts(count), ids((size_t)count * BBX_BUCKET_SIZE);
            bbx_throw_if(bbx_neighbors_get(dev, counts.data(), ids.data()), "GetParticleBucket");
            buckets_.resize(count);
            for (int p = 0; p < count; p++)
                buckets_[p].ids.assign(ids.begin() + (size_t)p * BBX_BUCKET_SIZE,
                                       ids.begin() + (size_t)p * BBX_BUCKET_SIZE + counts[p]);
        }
        return &buckets_[i];
    }
    /* drop cached downloads (called after every device step) */
    void Invalidate() {
        for (auto &c : cache_) c.clear();
        buckets_.clear();
    }

  private:
    std::vector<double> cache_[10];
    std::vector<Bucket> buckets_;
    const std::vector<double> &fetch(int field, int width) {
        auto &c = cache_[field];
        if (c.empty()) {
            c.resize((size_t)count * width);
            bbx_throw_if(bbx_solver_download(dev, field, c.data()), "download");
        }
        return c;
    }
    vec3f vfield(int field, int i) {
        if (i < 0 || i >= count) throw bubbles_error("particle id out of range");
        if (!dev) {
            if (field == BBX_FIELD_POSITION) return vec3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            if (field == BBX_FIELD_VELOCITY) return vec3f(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
            return vec3f(0);
        }
        const auto &c = fetch(field, 3);
        return vec3f(c[3 * (size_t)i], c[3 * (size_t)i + 1], c[3 * (size_t)i + 2]);
    }
    Float sfield(int field, int i) {
        if (i < 0 || i >= count) throw bubbles_error("particle id out of range");
        if (!dev) return 0;
        return fetch(field, 1)[i];
    }
    void setv(int field, int i, const vec3f &v) {
        if (i < 0 || i >= count) throw bubbles_error("particle id out of range");
        std::vector<double> &host = field == BBX_FIELD_POSITION ? pos : vel;
        if (dev) {
            host = fetch(field, 3);
            host[3 * (size_t)i] = v.x;
            host[3 * (size_t)i + 1] = v.y;
            host[3 * (size_t)i + 2] = v.z;
            bbx_throw_if(bbx_solver_upload(dev, field, host.data()), "upload");
            Invalidate();
        } else {
            host[3 * (size_t)i] = v.x;
            host[3 * (size_t)i + 1] = v.y;
            host[3 * (size_t)i + 2] = v.z;
        }
    }
};

class SphParticleSet3 {
  public:
    ParticleSet3 *particleSet = nullptr;
    Float targetDensity = 1000.0, targetSpacing = 0.1, kernelRadiusOverTargetSpacing = 2.0;
    Float kernelRadius = 0.2;
    ParticleSet3 *GetParticleSet() { return particleSet; }
    Float GetKernelRadius() const { return kernelRadius; }
    Float GetTargetDensity() const { return targetDensity; }
    Float GetTargetSpacing() const { return targetSpacing; }
    void SetTargetDensity(Float d) { targetDensity = d; }
    void SetTargetSpacing(Float s) {
        targetSpacing = s;
        kernelRadius = kernelRadiusOverTargetSpacing * s;
        if (particleSet) particleSet->radius = s;
    }
    void SetRelativeKernelRadius(Float r) {
        kernelRadiusOverTargetSpacing = r;
        kernelRadius = r * targetSpacing;
    }
};

class ParticleSetBuilder3 {
  public:
    std::vector<vec3f> positions, velocities;
    int AddParticle(const vec3f &pos, const vec3f &vel = vec3f(0), const vec3f & = vec3f(0)) {
        positions.push_back(pos);
        velocities.push_back(vel);
        return 1;
    }
    int GetParticleCount() const { return (int)positions.size(); }
    void SetVelocityForAll(const vec3f &v) {
        for (auto &x : velocities) x = v;
    }
    ParticleSet3 *MakeParticleSet() {
        if (positions.empty()) throw bubbles_error("No particles in builder");
        ParticleSet3 *p = new ParticleSet3();
        p->count = (int)positions.size();
        p->pos.resize(3 * positions.size());
        p->vel.resize(3 * positions.size());
        for (size_t i = 0; i < positions.size(); i++)
            for (int k = 0; k < 3; k++) {
                p->pos[3 * i + k] = positions[i][k];
                p->vel[3 * i + k] = velocities[i][k];
            }
        positions.clear();
        velocities.clear();
        return p;
    }
};

inline SphParticleSet3 *SphParticleSet3FromBuilder(ParticleSetBuilder3 *builder) {
    if (!builder) throw bubbles_error("Invalid builder pointer for SphParticleSet3FromBuilder");
    SphParticleSet3 *s = new SphParticleSet3();
    s->particleSet = builder->MakeParticleSet();
    return s;
}

namespace bubbles_detail {
/* the chains of a full distribution: every particle in the chain of its
 * hashed cell, in ascending id (DistributeByParticle,
 * src/core/grid.h:390-407); points outside the grid are in none */
inline void hashed_chains(const Grid3 *g, const std::vector<double> &pos, int n, std::vector<int> &start,
                          std::vector<int> &members) {
    const unsigned total = g->GetCellCount();
    start.assign((size_t)total + 1, 0);
    std::vector<int> id((size_t)n);
    if (n) bbx_throw_if(bbx_grid_hash_desc(&g->desc, pos.data(), n, id.data(), nullptr), "chains");
    for (int i = 0; i < n; i++)
        if (id[i] >= 0) start[(size_t)id[i] + 1]++;
    for (unsigned c = 0; c < total; c++) start[c + 1] += start[c];
    members.assign((size_t)start[total], -1);
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (int i = 0; i < n; i++)
        if (id[i] >= 0) members[(size_t)fill[id[i]]++] = i;
}

inline void require_domain(bbx_solver *dev, const Grid3 *g) {
    bbx_info info;
    bbx_throw_if(bbx_solver_info(dev, &info), "chains");
    if (info.grid.total != g->GetCellCount()) throw bubbles_error("grid is not the solver's domain");
}

/* Cell chains of a particle set as CSR (cell_start total+1, members in
 * chain order) plus its positions by id: the solver's chains once a solver
 * owns the particles, else those of a full distribution (what Setup does). */
inline void chain_snapshot(ParticleSet3 *ps, const Grid3 *g, std::vector<int> &start,
                           std::vector<int> &members, std::vector<double> &pos) {
    const int n = ps->GetParticleCount();
    if (!ps->dev) {
        pos = ps->pos;
        hashed_chains(g, pos, n, start, members);
        return;
    }
    require_domain(ps->dev, g);
    const unsigned total = g->GetCellCount();
    start.assign((size_t)total + 1, 0);
    pos.assign(3 * (size_t)n, 0.0);
    members.assign((size_t)std::max(n, 1), -1);
    if (n) bbx_throw_if(bbx_solver_download(ps->dev, BBX_FIELD_POSITION, pos.data()), "chains");
    bbx_throw_if(bbx_grid_get_chains(ps->dev, start.data(), members.data()), "chains");
    members.resize((size_t)start[total]);
}
}  // namespace bubbles_detail

/* ResetAndDistribute (src/core/grid.h:688-697): the chains of a full
 * distribution.  Before a solver owns the particles there is no chain state
 * to reset: Setup distributes. */
inline void ResetAndDistribute(Grid3 *grid, ParticleSet3 *pSet) {
    if (!grid || !pSet || !pSet->dev) return;
    bubbles_detail::require_domain(pSet->dev, grid);
    const int n = pSet->GetParticleCount();
    std::vector<double> pos(3 * (size_t)n);
    if (n) bbx_throw_if(bbx_solver_download(pSet->dev, BBX_FIELD_POSITION, pos.data()), "ResetAndDistribute");
    std::vector<int> start, members;
    bubbles_detail::hashed_chains(grid, pos, n, start, members);
    if (members.empty()) members.push_back(-1);
    bbx_throw_if(bbx_grid_set_chains(pSet->dev, start.data(), members.data()), "ResetAndDistribute");
    pSet->Invalidate();
}

static const int kDefaultMaxParticles = 1000000;

/* ContinuousParticleSetBuilder3 (src/core/grid.h:1245-1450): a particle set
 * that keeps growing while a solver steps it.  AddParticle queues up to
 * maxNumOfParticles; Commit hands the queue to the set: to the device when a
 * solver owns it (bbx_solver_append: ids n.., each new particle at the tail
 * of its cell's chain, DistributeByParticleList), else to the host arrays
 * Setup will upload.  MapGrid records, for every cell a particle hashes into,
 * the positions of that cell's chain; MapGridEmit adds those positions back
 * wherever no particle of the cell's current chain lies closer than d. */
class ContinuousParticleSetBuilder3 {
  public:
    std::vector<vec3f> positions, velocities, forces;
    ParticleSet3 *particleSet;
    Float kernelRadius = 2.0 * 0.02;
    int maxNumOfParticles;
    int reemitions = 0;
    bool reemitOnce;
    bool warned = false;
    Grid3 *mappedDomain = nullptr;
    std::set<unsigned int> mappedCellSet;
    std::map<unsigned int, std::vector<vec3f>> mappedPositions;

    explicit ContinuousParticleSetBuilder3(int maxParticles = kDefaultMaxParticles, bool reemitOnce_ = false)
        : particleSet(new ParticleSet3()), maxNumOfParticles(std::max(1, maxParticles)), reemitOnce(reemitOnce_) {}

    void SetKernelRadius(Float radius) { kernelRadius = radius; }
    ParticleSet3 *GetParticleSet() { return particleSet; }
    int GetParticleCount() { return particleSet->GetParticleCount(); }

    int AddParticle(const vec3f &pos, const vec3f &vel = vec3f(0), const vec3f &force = vec3f(0)) {
        const int total = particleSet->GetParticleCount() + (int)positions.size();
        if (total + 1 <= maxNumOfParticles) {
            positions.push_back(pos);
            velocities.push_back(vel);
            forces.push_back(force);
            return 1;
        }
        if (!warned) {
            std::printf("\nReached maximum builder capacity\n");
            warned = true;
        }
        return 0;
    }

    void Commit() {
        if (positions.empty()) return;
        const int k = (int)positions.size();
        std::vector<double> p(3 * (size_t)k), v(3 * (size_t)k);
        for (int i = 0; i < k; i++)
            for (int a = 0; a < 3; a++) {
                p[3 * (size_t)i + a] = positions[(size_t)i][a];
                v[3 * (size_t)i + a] = velocities[(size_t)i][a];
            }
        ParticleSet3 *ps = particleSet;
        if (ps->dev) {
            // the reference sizes the set for maxNumOfParticles up front
            int cap = 0;
            bbx_throw_if(bbx_solver_capacity(ps->dev, &cap), "Commit");
            if (cap < maxNumOfParticles) bbx_throw_if(bbx_solver_reserve(ps->dev, maxNumOfParticles), "Commit");
            bbx_throw_if(bbx_solver_append(ps->dev, p.data(), v.data(), k), "Commit");
            ps->Invalidate();
        } else {
            ps->pos.insert(ps->pos.end(), p.begin(), p.end());
            ps->vel.insert(ps->vel.end(), v.begin(), v.end());
        }
        ps->count += k;
        positions.clear();
        velocities.clear();
        forces.clear();
    }

    void MapGrid(Grid3 *grid) {
        if (!grid) throw bubbles_error("MapGrid: null grid");
        mappedDomain = grid;
        std::vector<int> start, members;
        std::vector<double> pos;
        bubbles_detail::chain_snapshot(particleSet, grid, start, members, pos);
        const int n = particleSet->GetParticleCount();
        for (int i = 0; i < n; i++) {
            const unsigned h =
                grid->GetLinearHashedPosition(vec3f(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]));
            mappedCellSet.insert(h);
            if (start[h + 1] == start[h])
                throw bubbles_error("Called for MapGrid but grid is inconsistent, missing first distribution?");
            std::vector<vec3f> cellPos;
            for (int k = start[h]; k < start[h + 1]; k++) {
                const size_t j = (size_t)members[(size_t)k];
                cellPos.push_back(vec3f(pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]));
            }
            mappedPositions[h] = cellPos;
        }
    }

    /* MapGridEmitToOther (src/core/grid.h:1323-1364): this builder's mapped
     * positions refill the same cells of the other builder's grid.  As in
     * the reference, the other grid's chain members are looked up by id in
     * this builder's particle set. */
    void MapGridEmitToOther(ContinuousParticleSetBuilder3 *other,
                            const std::function<vec3f(const vec3f &)> &velocity, Float d = 0.02) {
        if (reemitions > 0 && reemitOnce) return;
        if (!other || !other->mappedDomain) throw bubbles_error("MapGridEmitToOther: the other builder has no mapped grid");
        std::vector<int> start, members;
        std::vector<double> otherPos, lookup;
        bubbles_detail::chain_snapshot(other->particleSet, other->mappedDomain, start, members, otherPos);
        const int n = particleSet->GetParticleCount();
        if (particleSet->dev) {
            lookup.assign(3 * (size_t)n, 0.0);
            if (n) bbx_throw_if(bbx_solver_download(particleSet->dev, BBX_FIELD_POSITION, lookup.data()), "MapGridEmitToOther");
        } else {
            lookup = particleSet->pos;
        }
        if (refill(other, start, members, lookup, velocity, d)) reemitions += 1;
    }

    void MapGridEmit(const std::function<vec3f(const vec3f &)> &velocity, Float d = 0.02) {
        if (reemitions > 0 && reemitOnce) return;
        if (warned) return;
        if (!mappedDomain) throw bubbles_error("MapGridEmit: call MapGrid first");
        std::vector<int> start, members;
        std::vector<double> pos;
        bubbles_detail::chain_snapshot(particleSet, mappedDomain, start, members, pos);
        if (refill(this, start, members, pos, velocity, d)) reemitions += 1;
    }

  private:
    /* Whether p keeps at least distance d from every particle of one chain
     * (members[first .. first + len), positions looked up by id). */
    static bool clear_of_chain(const vec3f &p, const std::vector<int> &members, int first, int len,
                               const std::vector<double> &lookup, Float d) {
        for (int k = first; k < first + len; k++) {
            const size_t j = (size_t)members[(size_t)k];
            if (3 * j + 2 >= lookup.size()) throw bubbles_error("MapGridEmit: chain id beyond the looked-up set");
            if ((vec3f(lookup[3 * j], lookup[3 * j + 1], lookup[3 * j + 2]) - p).Length() < d) return false;
        }
        return true;
    }

    /* The refill rule of MapGridEmit and MapGridEmitToOther (src/core/grid.h:
     * 1323-1395).  The mapped cells are visited in ascending id.  A cell
     * whose chain in the target grid already holds a bucket's worth
     * (BBX_BUCKET_SIZE) of particles is passed over; otherwise its first
     * min(BBX_BUCKET_SIZE - chain length, mapped count) positions are tried
     * in mapping order, and each one clear of the chain (as it stood when
     * the call began) is added to the target.  When the target's
     * AddParticle refuses one (capacity), what was added is committed and
     * the call ends: returns false.  Otherwise anything added is committed
     * at the end. */
    bool refill(ContinuousParticleSetBuilder3 *target, const std::vector<int> &start,
                const std::vector<int> &members, const std::vector<double> &lookup,
                const std::function<vec3f(const vec3f &)> &velocity, Float d) {
        int added = 0;
        for (unsigned h : mappedCellSet) {
            const int first = start[h], len = start[h + 1] - start[h];
            if (len >= BBX_BUCKET_SIZE) continue;
            const std::vector<vec3f> &cand = mappedPositions[h];
            const size_t tries = std::min((size_t)(BBX_BUCKET_SIZE - len), cand.size());
            for (size_t t = 0; t < tries; t++) {
                if (!clear_of_chain(cand[t], members, first, len, lookup, d)) continue;
                if (!target->AddParticle(cand[t], velocity(cand[t]))) {
                    target->Commit();
                    return false;
                }
                added++;
            }
        }
        if (added > 0) target->Commit();
        return true;
    }
};

/* SphParticleSet3FromContinuousBuilder (src/core/particle.cpp:29-35): the
 * builder keeps the particle set and grows it on every Commit. */
inline SphParticleSet3 *SphParticleSet3FromContinuousBuilder(ContinuousParticleSetBuilder3 *builder) {
    if (!builder) throw bubbles_error("Invalid builder pointer for SphParticleSet3FromContinuousBuilder");
    SphParticleSet3 *s = new SphParticleSet3();
    s->particleSet = builder->GetParticleSet();
    return s;
}

struct SphSolverData3 {
    SphParticleSet3 *sphpSet = nullptr;
    Grid3 *domain = nullptr;
    ColliderSet3 *collider = nullptr;
    Float dragCoefficient = 0.0001, eosExponent = 7.0, negativePressureScale = 0.0;
    Float viscosity = 0.04, pseudoViscosity = 10.0, soundSpeed = 100.0, timestepScale = 1.0;
    vec3f gravity = vec3f(0.0, (double)(-9.8f), 0.0);
};

/* DefaultSphSolverData3 (src/solvers/sph_solver3.cpp:124-151) */
inline SphSolverData3 *DefaultSphSolverData3(bool with_gravity = true) {
    SphSolverData3 *d = new SphSolverData3();
    if (!with_gravity) d->gravity = vec3f(0);
    return d;
}

namespace bubbles_detail {
/* Create the device solver from the SphSolverData3 + Setup arguments. */
inline bbx_solver *make_device(int kind, SphSolverData3 *d, Float targetDensity, Float targetSpacing,
                               Float relativeRadius, Grid3 *domain, SphParticleSet3 *pSet) {
    if (!d || !domain || !pSet || !pSet->particleSet)
        throw bubbles_error("Invalid solver initialization");
    bbx_params p = domain->prm;
    p.target_density = targetDensity;
    p.target_spacing = targetSpacing;
    p.relative_radius = relativeRadius;
    p.eos_exponent = d->eosExponent;
    p.negative_pressure_scale = d->negativePressureScale;
    p.viscosity = d->viscosity;
    p.pseudo_viscosity = d->pseudoViscosity;
    p.sound_speed = d->soundSpeed;
    p.drag_coefficient = d->dragCoefficient;
    for (int k = 0; k < 3; k++) p.gravity[k] = d->gravity[k];
    ParticleSet3 *ps = pSet->particleSet;
    bbx_solver *s = nullptr;
    bbx_throw_if(bbx_solver_create(kind, &p, ps->pos.data(), ps->vel.data(), ps->count, &s), "Setup");
    bbx_throw_if(bbx_solver_setup(s), "Setup");
    bbx_info info;
    bbx_throw_if(bbx_solver_info(s, &info), "Setup");
    ps->dev = s;
    ps->mass = info.mass;
    ps->radius = info.radius;
    pSet->targetDensity = targetDensity;
    pSet->targetSpacing = targetSpacing;
    pSet->kernelRadiusOverTargetSpacing = relativeRadius;
    pSet->kernelRadius = info.kernel_radius;
    d->domain = domain;
    d->sphpSet = pSet;
    return s;
}

/* true when the set is the grid's domain as one reversed, frictionless box:
 * the solver's built-in tank, which needs no collider list */
inline bool is_plain_tank(const ColliderSet3 *c, const Grid3 *g) {
    if (!c || !g || c->colliders.size() != 1) return false;
    const Collider3 *k = c->colliders[0];
    const Shape *s = k->shape;
    if (!k->isActive || k->frictionCoefficient != 0 || !s || s->type != ShapeBox || !s->reverseOrientation ||
        !s->ObjectToWorld.IsTranslation())
        return false;
    const Float size[3] = {s->sizex, s->sizey, s->sizez};
    for (int i = 0; i < 3; i++)
        if (s->ObjectToWorld.m[i][3] != g->prm.tank_center[i] || size[i] != g->prm.tank_size[i]) return false;
    return true;
}

/* ColliderSet3 -> the device's collider list (ColliderSet3::ResolveCollision
 * semantics, src/core/collider.cpp:238-264); called at Setup and before
 * every step, so Shape::Update, SetVelocities and SetActive take effect */
inline void push_colliders(bbx_solver *dev, const ColliderSet3 *c, const Grid3 *g, bool &listInUse,
                           std::vector<bbx_collider> &last) {
    if (!c || is_plain_tank(c, g)) {
        if (listInUse) bbx_throw_if(bbx_solver_set_colliders(dev, nullptr, 0), "SetColliders");
        listInUse = false;
        last.clear();
        return;
    }
    std::vector<bbx_collider> list(c->colliders.size());
    for (size_t i = 0; i < list.size(); i++) {
        const Collider3 *k = c->colliders[i];
        const Shape *s = k->shape;
        bbx_collider &b = list[i];
        std::memset(&b, 0, sizeof(b));  // padding included, for the comparison below
        b.kind = s->type == ShapeBox ? BBX_SHAPE_BOX : (s->type == ShapeSphere ? BBX_SHAPE_SPHERE : BBX_SHAPE_MESH);
        b.reverse_orientation = s->reverseOrientation ? 1 : 0;
        b.active = k->isActive ? 1 : 0;
        b.friction = k->frictionCoefficient;
        s->ObjectToWorld.Rows(b.to_world, b.to_object);
        b.size[0] = s->sizex; b.size[1] = s->sizey; b.size[2] = s->sizez;
        b.radius = s->radius;
        for (int d = 0; d < 3; d++) {
            b.linear_velocity[d] = s->linearVelocity[d];
            b.angular_velocity[d] = s->angularVelocity[d];
            b.bounds_min[d] = s->meshBounds.pMin[d];
            b.bounds_max[d] = s->meshBounds.pMax[d];
        }
        if (s->type == ShapeMesh) {
            if (!s->grid) throw bubbles_error("SetColliders: mesh collider without an SDF");
            b.sdf = s->grid;
        }
    }
    // unchanged since the last upload: nothing to do (the upload waits for
    // the solver's stream)
    if (listInUse && list.size() == last.size() &&
        std::memcmp(list.data(), last.data(), sizeof(bbx_collider) * list.size()) == 0)
        return;
    bbx_throw_if(bbx_solver_set_colliders(dev, list.data(), (int)list.size()), "SetColliders");
    listInUse = true;
    last = list;
}
}  // namespace bubbles_detail

/* PciSphSolver3 (src/core/pcisph_solver.h:56-83) */
class PciSphSolver3 {
  public:
    /* loop settings; the reference's shipped loop is REFERENCE_EXACT, maxIt 5, tol 0.01 */
    int maxIterations = 5;
    Float maxErrorDensity = 0.01;
    int loopMode = BBX_MODE_REFERENCE_EXACT;
    int warmStart = BBX_WARM_NONE;

    PciSphSolver3() {}
    ~PciSphSolver3() {
        if (dev_) bbx_solver_destroy(dev_);
    }
    void Initialize(SphSolverData3 *data) { data_ = data; }
    void Setup(Float targetDensity, Float targetSpacing, Float relativeRadius, Grid3 *domain,
               SphParticleSet3 *pSet) {
        if (!data_) throw bubbles_error("Invalid call to {PciSphSolver3::Setup}");
        dev_ = bubbles_detail::make_device(BBX_PCISPH, data_, targetDensity, targetSpacing, relativeRadius,
                                           domain, pSet);
        colliders();
        bbx_info info;
        bbx_throw_if(bbx_solver_info(dev_, &info), "Setup");
        deltaDenom = info.delta_denom;
        Float m = info.mass / targetDensity;
        massOverTargetDensitySquared = m * m;
    }
    void SetColliders(ColliderSet3 *colliders) {
        if (!data_) throw bubbles_error("Invalid solverData for {SetColliders}");
        data_->collider = colliders;
        if (dev_) this->colliders();
    }
    ColliderSet3 *GetColliders() { return data_ ? data_->collider : nullptr; }
    void SetViscosityCoefficient(Float v) {
        if (!data_) throw bubbles_error("Invalid solverData for {SetViscosityCoefficient}");
        if (dev_) throw bubbles_error("SetViscosityCoefficient must precede Setup");
        data_->viscosity = v > 0 ? v : 0;
    }
    SphSolverData3 *GetSphSolverData() { return data_; }
    SphParticleSet3 *GetSphParticleSet() { return data_ ? data_->sphpSet : nullptr; }
    int GetParticleCount() { return GetSphParticleSet()->GetParticleSet()->GetParticleCount(); }
    /* PciSphSolver3::Advance (src/solvers/pcisph_solver3.cpp:67-89) */
    void Advance(Float timeIntervalInSeconds) {
        require();
        colliders();
        int n = 0;
        bbx_throw_if(bbx_pcisph_advance(dev_, timeIntervalInSeconds, maxIterations, maxErrorDensity,
                                        loopMode, warmStart, &n),
                     "Advance");
        lastSubsteps = n;
        GetSphParticleSet()->GetParticleSet()->Invalidate();
    }
    /* one fixed-dt substep (AdvanceTimeStep); returns its statistics */
    bbx_stats AdvanceSubstep(Float dt) {
        require();
        colliders();
        bbx_stats st;
        bbx_throw_if(bbx_pcisph_substep(dev_, dt, maxIterations, maxErrorDensity, loopMode, warmStart, &st),
                     "AdvanceSubstep");
        stepInterval = st.gpu_ms;
        GetSphParticleSet()->GetParticleSet()->Invalidate();
        return st;
    }
    Float GetAdvanceTime() const { return stepInterval; }
    Float ComputeBeta(Float dt) const { return 2.0 * massOverTargetDensitySquared * dt * dt; }
    Float ComputeDelta(Float dt) const {
        return std::fabs(deltaDenom) > 0 ? -1 / (ComputeBeta(dt) * deltaDenom) : 0;
    }
    Float ComputeDeltaDenom() const { return deltaDenom; }
    bbx_solver *Device() { return dev_; }

    Float massOverTargetDensitySquared = 0, deltaDenom = 0, stepInterval = 0;
    int lastSubsteps = 0;

  private:
    SphSolverData3 *data_ = nullptr;
    bbx_solver *dev_ = nullptr;
    bool listInUse_ = false;
    void require() {
        if (!dev_) throw bubbles_error("PciSphSolver3: call Setup first");
    }
    std::vector<bbx_collider> lastList_;
    void colliders() { bubbles_detail::push_colliders(dev_, data_->collider, data_->domain, listInUse_, lastList_); }
};

/* SphSolver3 (src/core/sph_solver.h:76-96), race-free two-phase integration */
class SphSolver3 {
  public:
    SphSolver3() {}
    ~SphSolver3() {
        if (dev_) bbx_solver_destroy(dev_);
    }
    void Initialize(SphSolverData3 *data) { data_ = data; }
    void Setup(Float targetDensity, Float targetSpacing, Float relativeRadius, Grid3 *domain,
               SphParticleSet3 *pSet) {
        if (!data_) throw bubbles_error("Invalid SphSolver3 initialization");
        dev_ = bubbles_detail::make_device(BBX_SPH, data_, targetDensity, targetSpacing, relativeRadius,
                                           domain, pSet);
        bubbles_detail::push_colliders(dev_, data_->collider, data_->domain, listInUse_, lastList_);
    }
    void SetColliders(ColliderSet3 *c) {
        if (!data_) throw bubbles_error("Invalid solverData for {SetColliders}");
        data_->collider = c;
        if (dev_) bubbles_detail::push_colliders(dev_, data_->collider, data_->domain, listInUse_, lastList_);
    }
    ColliderSet3 *GetColliders() { return data_ ? data_->collider : nullptr; }
    SphSolverData3 *GetSphSolverData() { return data_; }
    void SetViscosityCoefficient(Float v) {
        if (dev_) throw bubbles_error("SetViscosityCoefficient must precede Setup");
        data_->viscosity = v > 0 ? v : 0;
    }
    void SetPseudoViscosityCoefficient(Float v) {
        if (dev_) throw bubbles_error("SetPseudoViscosityCoefficient must precede Setup");
        data_->pseudoViscosity = v > 0 ? v : 0;
    }
    Float GetKernelRadius() { return data_->sphpSet->GetKernelRadius(); }
    SphParticleSet3 *GetSphParticleSet() { return data_ ? data_->sphpSet : nullptr; }
    void Advance(Float timeIntervalInSeconds) {
        if (!dev_) throw bubbles_error("SphSolver3: call Setup first");
        bubbles_detail::push_colliders(dev_, data_->collider, data_->domain, listInUse_, lastList_);
        int n = 0;
        bbx_throw_if(bbx_sph_advance(dev_, timeIntervalInSeconds, &n), "Advance");
        lastSubsteps = n;
        data_->sphpSet->GetParticleSet()->Invalidate();
    }
    bbx_solver *Device() { return dev_; }
    int lastSubsteps = 0;

  private:
    SphSolverData3 *data_ = nullptr;
    bbx_solver *dev_ = nullptr;
    bool listInUse_ = false;
    std::vector<bbx_collider> lastList_;
};

/* ---- serializer (src/third/serializer.h, src/third/serializer.cpp) ---- */
#ifndef SERIALIZER_POSITION
#define SERIALIZER_POSITION BBX_SER_POSITION
#define SERIALIZER_VELOCITY BBX_SER_VELOCITY
#define SERIALIZER_DENSITY BBX_SER_DENSITY
#define SERIALIZER_BOUNDARY BBX_SER_BOUNDARY
#define SERIALIZER_NORMAL BBX_SER_NORMAL
#define SERIALIZER_MASS BBX_SER_MASS
#define SERIALIZER_LAYERS BBX_SER_LAYERS
#define SERIALIZER_XYZ BBX_SER_XYZ
#define SERIALIZER_RULE_BOUNDARY_EXCLUSIVE BBX_SER_BOUNDARY_EXCLUSIVE
#endif

/* SaveSphParticleSet (src/third/serializer.cpp:875-911): appends the
 * particles of the solver that owns data's particle set (device data in id
 * order) to `filename` as one FluidBegin ... FluidEnd block. */
inline void SaveSphParticleSet(SphSolverData3 *data, const char *filename, int flags) {
    if (!data || !data->sphpSet || !data->sphpSet->GetParticleSet() || !data->sphpSet->GetParticleSet()->dev)
        throw bubbles_error("SaveSphParticleSet: the particle set is not in a solver");
    bbx_throw_if(bbx_solver_save(data->sphpSet->GetParticleSet()->dev, filename, flags), "SaveSphParticleSet");
}

/* SerializerLoadSphDataSet3 (src/third/serializer.cpp:562-577): adds the
 * particles of the file's first fluid block to the builder; flags takes the
 * block's "Format".  Returns the particle count. */
inline int SerializerLoadSphDataSet3(ParticleSetBuilder3 *builder, const char *filename, int &flags,
                                     std::vector<int> *boundary = nullptr) {
    if (!builder) throw bubbles_error("SerializerLoadSphDataSet3: null builder");
    int n = 0, fl = flags;
    bbx_throw_if(bbx_serializer_read(filename, -1, &n, &fl, nullptr, nullptr, nullptr, nullptr, nullptr),
                 "SerializerLoadSphDataSet3");
    std::vector<double> p(3 * (size_t)n), v(3 * (size_t)n);
    std::vector<int> b((size_t)n);
    fl = flags;
    bbx_throw_if(bbx_serializer_read(filename, n, &n, &fl, p.data(), v.data(), nullptr, nullptr, b.data()),
                 "SerializerLoadSphDataSet3");
    flags = fl;
    for (int i = 0; i < n; i++) {
        builder->AddParticle(vec3f(p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]),
                             vec3f(v[3 * (size_t)i], v[3 * (size_t)i + 1], v[3 * (size_t)i + 2]));
        if (boundary) boundary->push_back(b[(size_t)i]);
    }
    return n;
}

#endif
