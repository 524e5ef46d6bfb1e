// Exercises the C++ drop-in API the way the reference's own callers use it
// (apps/albedo/main.cpp:85-106): build a Scene, a PerspectiveCamera and a
// RenderJob, call renderImage into a PixelBuffer. Prints the image mean, checks
// the reference's argument errors, the default estimator and renderImage's
// accumulation, and writes the first image's per-pixel means (FP32 RGB) to
// argv[2] when given. Exit status 0 on success.
#include "albedo_dropin.hpp"

#include <cstdio>
#include <cstring>
#include <memory>

using namespace albedo;

namespace {

ab_camera abiCamera(const CameraSettings& s)
{
    ab_camera c{};
    const Real f[3] = {s.look_from.x, s.look_from.y, s.look_from.z}, a[3] = {s.look_at.x, s.look_at.y, s.look_at.z},
               u[3] = {s.up.x, s.up.y, s.up.z};
    for (int k = 0; k < 3; ++k) {
        c.look_from[k] = f[k];
        c.look_at[k] = a[k];
        c.up[k] = u[k];
    }
    c.vfov_degrees = s.vertical_fov_degrees;
    c.aspect = s.aspect_ratio;
    return c;
}

bool sameImage(const PixelBuffer& a, const PixelBuffer& b)
{
    for (int y = 0; y < a.height(); ++y)
        for (int x = 0; x < a.width(); ++x) {
            const Rgb p = a.average({x, y}), q = b.average({x, y});
            if (p.r != q.r || p.g != q.g || p.b != q.b) return false;
        }
    return true;
}

} // namespace

int main(int argc, char** argv)
{
    const int width = argc > 1 ? std::atoi(argv[1]) : 64;
    const int height = width * 9 / 16;
    Scene scene;
    const Material white{DiffuseMaterial{ConstantTexture{{0.73, 0.73, 0.73}}}};
    scene.add(Sphere{{0.0, -1e5, 0.0}, 1e5}, white);
    scene.add(Sphere{{0.0, 6.0 + 1e5, 0.0}, 1e5}, white);
    scene.add(Sphere{{0.0, 0.0, -6.0 - 1e5}, 1e5}, white);
    scene.add(Sphere{{-3.0 - 1e5, 0.0, 0.0}, 1e5}, DiffuseMaterial{ConstantTexture{{0.65, 0.05, 0.05}}});
    scene.add(Sphere{{3.0 + 1e5, 0.0, 0.0}, 1e5}, DiffuseMaterial{ConstantTexture{{0.12, 0.45, 0.15}}});
    scene.add(Sphere{{-0.9, 1.0, -1.0}, 1.0}, DielectricMaterial{1.5, 0.0});
    scene.add(Sphere{{1.6, 0.8, -2.2}, 0.8}, ConductorMaterial{{0.156, 0.424, 1.383}, {3.602, 2.472, 1.915}, 0.25});
    scene.add(Sphere{{-2.2, 0.5, -2.8}, 0.5},
              MixMaterial{ConductorMaterial{{0.155, 0.117, 0.138}, {4.83, 3.12, 2.15}, 0.0},
                          DiffuseMaterial{ConstantTexture{{0.9, 0.9, 0.85}}}, PerlinPattern{3.0, 6.0}, 0.5});
    // A two-triangle checker floor patch as a mesh instance.
    TriangleMesh quad{{{-1, 0.01, 1}, {1, 0.01, 1}, {1, 0.01, -1}, {-1, 0.01, -1}},
                      {{0, 1, 2}, {0, 2, 3}},
                      {{-1, 1}, {1, 1}, {1, -1}, {-1, -1}},
                      {{0, 1, 2}, {0, 2, 3}}};
    auto geometry = std::make_shared<const MeshGeometry>(quad, MeshShading::Flat);
    scene.add(MeshInstance{geometry, Transform::translation({0.5, 0.0, 1.0}) * Transform::rotation({0, 1, 0}, 30)},
              DiffuseMaterial{CheckerTexture{{0.73, 0.73, 0.73}, {0.3, 0.3, 0.3}, 1.25}});
    scene.addAreaLight(Sphere{{0.0, 5.6, -1.2}, 1.1}, {9.0, 7.6, 5.8});
    scene.addLight(PointLight{{0.8, 4.2, 3.2}, {1.2, 2.0, 3.6}});
    scene.setSky(SkyLight::uniform({0.05, 0.06, 0.08}));

    const PerspectiveCamera camera{CameraSettings{{0.0, 2.4, 7.8}, {0.0, 1.9, -0.5}, {0, 1, 0}, 38.0,
                                                  double(width) / height}};
    PixelBuffer image{width, height};
    renderImage(RenderJob{camera, scene, 4, 7, PathEstimator{8, 3}, availableThreadCount()}, image);
    double sum = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const Rgb c = image.average({x, y});
            sum += c.r + c.g + c.b;
        }
    std::printf("mean %.6f\n", sum / (3.0 * width * height));
    if (argc > 2) {
        std::vector<float> out(std::size_t(width) * height * 3);
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const Rgb c = image.average({x, y});
                float* o = &out[(std::size_t(y) * width + x) * 3];
                o[0] = float(c.r);
                o[1] = float(c.g);
                o[2] = float(c.b);
            }
        FILE* f = std::fopen(argv[2], "wb");
        if (!f || std::fwrite(out.data(), sizeof(float), out.size(), f) != out.size()) return 1;
        std::fclose(f);
    }

    // renderImage adds samples (reference render_image.hpp:37-48): two seeds
    // rendered into one buffer average like the merged sums of the two renders.
    PixelBuffer twice{width, height};
    renderImage(RenderJob{camera, scene, 2, 11, PathEstimator{8, 3}, 1}, twice);
    renderImage(RenderJob{camera, scene, 2, 12, PathEstimator{8, 3}, 1}, twice);
    const std::size_t npx = std::size_t(width) * height;
    std::vector<uint64_t> first(npx * 4, 0), second(npx * 4, 0);
    const ab_camera cam = abiCamera(camera.settings());
    ab_render_params rp{};
    rp.width = uint32_t(width);
    rp.height = uint32_t(height);
    rp.samples_per_axis = 2;
    rp.max_depth = 8;
    rp.roulette_start = 3;
    rp.estimator = AB_ESTIMATOR_PATH;
    rp.seed = 11;
    if (ab_render_sums(scene.device(), &cam, &rp, first.data(), nullptr) != AB_OK) return 1;
    rp.seed = 12;
    if (ab_render_sums(scene.device(), &cam, &rp, second.data(), nullptr) != AB_OK) return 1;
    std::vector<float> merged(npx * 3);
    if (ab_sums_merge(first.data(), second.data(), npx) != AB_OK ||
        ab_sums_resolve(first.data(), npx, merged.data()) != AB_OK)
        return 1;
    int mismatches = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const Rgb c = twice.average({x, y});
            const float* m = &merged[(std::size_t(y) * width + x) * 3];
            const float got[3] = {float(c.r), float(c.g), float(c.b)};
            if (std::memcmp(got, m, sizeof(got)) != 0 || first[(std::size_t(y) * width + x) * 4 + 3] != 8) ++mismatches;
        }
    std::printf("accumulate_mismatches %d\n", mismatches);
    // addSample keeps the reference's sums and counts.
    PixelBuffer manual{2, 1};
    manual.addSample({0, 0}, Rgb{1.0, 2.0, 3.0});
    manual.addSample({0, 0}, Rgb{3.0, -2.0, 0.5});
    const Rgb avg = manual.average({0, 0}), none = manual.average({1, 0});
    const bool samples_ok = avg.r == 2.0 && avg.g == 0.0 && avg.b == 1.75 && none.r == 0 && none.g == 0 && none.b == 0;
    std::printf("add_sample_ok %d\n", samples_ok ? 1 : 0);
    // A default RenderJob scatters only, as the reference's does (render_image.hpp:28).
    PixelBuffer by_default{width, height}, scatter_only{width, height};
    renderImage(RenderJob{camera, scene, 2, 5}, by_default);
    renderImage(RenderJob{camera, scene, 2, 5, ScatterOnlyEstimator{}}, scatter_only);
    PixelBuffer path{width, height};
    renderImage(RenderJob{camera, scene, 2, 5, PathEstimator{}}, path);
    std::printf("default_is_scatter_only %d\n", sameImage(by_default, scatter_only) && !sameImage(by_default, path));
    int size_errors = 0;
    for (const auto& wh : {std::pair<int, int>{0, 4}, {4, -1}, {1 << 15, 1 << 14}})
        try {
            PixelBuffer bad{wh.first, wh.second};
        } catch (const std::invalid_argument&) {
            ++size_errors;
        }
    std::printf("size_errors %d\n", size_errors);
    if (!PixelBuffer::isValidSize(1 << 14, 1 << 14) || PixelBuffer::isValidSize(1 << 14, (1 << 14) + 1)) return 1;

    // The other estimators of the reference's RadianceEstimator variant.
    PixelBuffer normals{width, height};
    renderImage(RenderJob{camera, scene, 1, 7, NormalsEstimator{}, 1}, normals);
    PixelBuffer scatter{width, height};
    renderImage(RenderJob{camera, scene, 2, 7, ScatterOnlyEstimator{8}, 1}, scatter);
    double nmin = 1, nmax = 0, ssum = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const Rgb c = normals.average({x, y});
            nmin = std::min({nmin, double(c.r), double(c.g), double(c.b)});
            nmax = std::max({nmax, double(c.r), double(c.g), double(c.b)});
            const Rgb d = scatter.average({x, y});
            ssum += d.r + d.g + d.b;
        }
    std::printf("normals in [%.3f, %.3f], scatter-only mean %.6f\n", nmin, nmax, ssum / (3.0 * width * height));
    if (nmin < 0 || nmax > 1 || !(ssum > 0)) return 1;

    // World-space triangles (reference geometry/triangle.hpp, scene/scene.cpp:33-60):
    // a tent of four faces, one of them without area, added as a mesh, as
    // Triangle shapes and as an identity MeshInstance renders the same image.
    const TriangleMesh tent{{{-1, 0.02, 0}, {0, 1.2, -0.4}, {1, 0.02, 0}, {0, 0.02, -1.5}, {2, 0.02, 0}},
                            {{0, 1, 2}, {0, 2, 4}, {2, 1, 3}, {3, 1, 0}}};
    const Material red{DiffuseMaterial{CheckerTexture{{0.8, 0.2, 0.2}, {0.2, 0.2, 0.8}, 3.0}}};
    Scene as_mesh, as_triangles, as_instance;
    for (Scene* s : {&as_mesh, &as_triangles, &as_instance}) {
        s->add(Sphere{{0.0, -1e4, 0.0}, 1e4}, white);
        s->addLight(PointLight{{0.8, 4.2, 3.2}, {6.0, 6.0, 6.0}});
        s->setSky(SkyLight::uniform({0.2, 0.25, 0.3}));
    }
    as_mesh.add(tent, red);
    for (const auto& f : tent.faces()) {
        const auto& P = tent.positions();
        if (formsTriangle(P[f[0]], P[f[1]], P[f[2]])) as_triangles.add(Triangle{P[f[0]], P[f[1]], P[f[2]]}, red);
    }
    as_instance.add(MeshInstance{std::make_shared<const MeshGeometry>(tent, MeshShading::Flat), Transform{}}, red);
    for (Scene* s : {&as_mesh, &as_triangles, &as_instance}) s->add(Sphere{{0.6, 0.5, 0.8}, 0.4}, white);
    PixelBuffer img_mesh{width, height}, img_tris{width, height}, img_inst{width, height};
    const PerspectiveCamera tent_camera{CameraSettings{{0.5, 2.0, 4.0}, {0.0, 0.4, -0.4}, {0, 1, 0}, 45.0,
                                                       double(width) / height}};
    renderImage(RenderJob{tent_camera, as_mesh, 3, 9, PathEstimator{8, 3}, 1}, img_mesh);
    renderImage(RenderJob{tent_camera, as_triangles, 3, 9, PathEstimator{8, 3}, 1}, img_tris);
    renderImage(RenderJob{tent_camera, as_instance, 3, 9, PathEstimator{8, 3}, 1}, img_inst);
    const bool forms_equal = sameImage(img_mesh, img_tris) && sameImage(img_mesh, img_inst) &&
                             !sameImage(img_mesh, by_default);
    const bool counts_ok = as_mesh.shapeCount() == 5 && as_triangles.shapeCount() == 5 && as_instance.shapeCount() == 3;
    std::printf("triangle_forms_equal %d\nshape_counts_ok %d\n", forms_equal ? 1 : 0, counts_ok ? 1 : 0);
    int triangle_errors = 0;
    const Point3r o{0, 0, 0}, x{1, 0, 0}, y{0, 1, 0};
    const Real nan = std::nan("");
    for (const std::function<void()>& bad : std::vector<std::function<void()>>{
             [&] { Triangle t{o, x, Point3r{2, 0, 0}}; },
             [&] { Triangle t{o, x, Point3r{0, nan, 0}}; },
             [&] { Triangle t{o, x, y, {Point2r{0, 0}, Point2r{1, nan}, Point2r{1, 1}}}; },
             [&] { TriangleMesh m{{o, x, Point3r{nan, 0, 0}}, {{0, 1, 2}}}; }})
        try {
            bad();
        } catch (const std::invalid_argument&) {
            ++triangle_errors;
        }
    Scene empty_mesh;
    empty_mesh.add(TriangleMesh{{o, x, Point3r{2, 0, 0}}, {{0, 1, 2}}}, red);
    const Triangle unit{o, x, y};
    const bool triangle_ok = empty_mesh.shapeCount() == 0 && unit.normal().z == 1.0 &&
                             unit.textureCoordinates()[2].x == 1.0 && unit.textureCoordinates()[2].y == 1.0;
    std::printf("triangle_errors %d\ntriangle_ok %d\n", triangle_errors, triangle_ok ? 1 : 0);

    int errors_seen = 0;
    try {
        renderImage(RenderJob{camera, scene, 0, 7, PathEstimator{}, 1}, image);
    } catch (const std::invalid_argument&) {
        ++errors_seen;
    }
    try {
        renderImage(RenderJob{camera, scene, 2, 7, PathEstimator{}, 0}, image);
    } catch (const std::invalid_argument&) {
        ++errors_seen;
    }
    try {
        Sphere bad{{0, 0, 0}, -1.0};
        (void)bad;
    } catch (const std::invalid_argument&) {
        ++errors_seen;
    }
    std::printf("errors %d\n", errors_seen);
    return (errors_seen == 3 && sum > 0 && mismatches == 0 && samples_ok && size_errors == 3 && forms_equal &&
            counts_ok && triangle_errors == 4 && triangle_ok)
               ? 0
               : 1;
}
