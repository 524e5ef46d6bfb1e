// Scene::intersect / Scene::occluded of the C++ drop-in (reference
// scene/scene.cpp:107-165) on a scene of spheres, Triangle shapes, an added
// TriangleMesh and a smooth, rotated and scaled MeshInstance. Writes the rays
// and one record per ray to argv[1] as text; tests/test_dropin_intersect.py
// checks them against the oracle's FP64 traversal of the same scene.
//
// Output lines: "ray ox oy oz dx dy dz", then "hit t object px py pz nx ny nz
// sx sy sz front u v occluded" or "miss occluded".
#include "albedo_dropin.hpp"

#include <cmath>
#include <cstdio>
#include <memory>

using namespace albedo;

namespace {

// A unit octahedron subdivided once and pushed onto the unit sphere: 32 faces.
TriangleMesh ball()
{
    std::vector<Point3r> p = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    std::vector<TriangleMesh::Face> f = {{0, 2, 4}, {2, 1, 4}, {1, 3, 4}, {3, 0, 4},
                                         {2, 0, 5}, {1, 2, 5}, {3, 1, 5}, {0, 3, 5}};
    std::vector<TriangleMesh::Face> out;
    auto mid = [&](std::size_t a, std::size_t b) {
        Point3r m{(p[a].x + p[b].x) / 2, (p[a].y + p[b].y) / 2, (p[a].z + p[b].z) / 2};
        const Real l = std::sqrt(m.x * m.x + m.y * m.y + m.z * m.z);
        p.push_back({m.x / l, m.y / l, m.z / l});
        return p.size() - 1;
    };
    for (const auto& t : f) {
        const std::size_t a = mid(t[0], t[1]), b = mid(t[1], t[2]), c = mid(t[2], t[0]);
        out.push_back({t[0], a, c});
        out.push_back({a, t[1], b});
        out.push_back({c, b, t[2]});
        out.push_back({a, b, c});
    }
    return TriangleMesh{p, out};
}

uint64_t state = 0x853c49e6748fea9bULL;
Real uniform()
{
    state = state * 6364136223846793005ULL + 1442695040888963407ULL;
    return Real(state >> 11) * (1.0 / 9007199254740992.0);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    Scene scene;
    const Material grey{DiffuseMaterial{ConstantTexture{{0.6, 0.6, 0.6}}}};
    scene.add(Sphere{{0.0, -1e5, 0.0}, 1e5}, grey);
    scene.add(Sphere{{-1.5, 0.6, 0.0}, 0.6}, grey);
    scene.add(Triangle{{0.5, 0.0, -1.0}, {1.5, 0.0, -1.0}, {1.0, 1.2, -1.2}}, grey);
    // Two faces and a zero-area one: objects 3 and 4 in the reference's count.
    scene.add(TriangleMesh{{{-0.5, 0.0, 1.0}, {0.5, 0.0, 1.0}, {0.0, 1.0, 1.1}, {1.0, 1.0, 1.1}, {2.0, 2.0, 1.2}},
                           {{0, 1, 2}, {1, 3, 2}, {0, 0, 4}}},
              grey);
    auto geometry = std::make_shared<const MeshGeometry>(ball(), MeshShading::Smooth);
    scene.add(MeshInstance{geometry, Transform::translation({0.4, 0.9, 0.2}) *
                                         Transform::rotation({1, 1, 0}, 35) * Transform::scaling({0.5, 0.7, 0.6})},
              grey);
    scene.add(Triangle{{-2.5, 0.0, 2.0}, {-1.5, 0.0, 2.0}, {-2.0, 1.0, 2.0}, {Point2r{0.1, 0.2}, Point2r{0.9, 0.1},
                                                                            Point2r{0.5, 0.8}}},
              grey);
    if (scene.shapeCount() != 7) return 3;

    std::vector<Ray> rays;
    for (int i = 0; i < 4000; ++i) {
        const Real th = 2 * 3.14159265358979323846 * uniform(), z = 0.2 + 0.8 * uniform();
        const Real rr = std::sqrt(1 - z * z);
        const Point3r o{5 * rr * std::cos(th), 0.2 + 3 * z, 5 * rr * std::sin(th)};
        const Point3r target{-2.6 + 4.4 * uniform(), 1.6 * uniform(), -1.4 + 3.6 * uniform()};
        rays.push_back({o, {target.x - o.x, target.y - o.y, target.z - o.z}});
    }
    const Interval range{0.0, std::numeric_limits<Real>::infinity()};
    const auto hits = scene.intersect(rays, range);
    const auto occ = scene.occluded(rays, range);
    // The one-ray forms agree with the batches.
    for (std::size_t i = 0; i < rays.size(); i += 397) {
        const auto one = scene.intersect(rays[i], range);
        if (bool(one) != bool(hits[i]) || (one && (one->t != hits[i]->t || one->object != hits[i]->object))) return 4;
        if (scene.occluded(rays[i], range) != (occ[i] != 0)) return 5;
    }
    // An empty range hits nothing.
    if (scene.intersect(rays[0], Interval{2.0, 1.0}) || scene.occluded(rays[0], Interval{}) ) return 6;
    FILE* f = std::fopen(argv[1], "w");
    if (!f) return 1;
    for (std::size_t i = 0; i < rays.size(); ++i) {
        const Ray& r = rays[i];
        std::fprintf(f, "ray %.17g %.17g %.17g %.17g %.17g %.17g\n", r.origin.x, r.origin.y, r.origin.z,
                     r.direction.x, r.direction.y, r.direction.z);
        if (const auto& h = hits[i]) {
            std::fprintf(f, "hit %.17g %zu %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d %.17g %.17g %d\n",
                         h->t, h->object, h->point.x, h->point.y, h->point.z, h->normal.x, h->normal.y,
                         h->normal.z, h->shading_normal.x, h->shading_normal.y, h->shading_normal.z,
                         h->front_face ? 1 : 0, h->uv.x, h->uv.y, int(occ[i]));
        } else {
            std::fprintf(f, "miss %d\n", int(occ[i]));
        }
    }
    std::fclose(f);
    std::printf("rays %zu sky %.2f\n", rays.size(), scene.sky().zenith.r);
    return 0;
}
