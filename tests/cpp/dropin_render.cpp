// Exercises the C++ drop-in API the way the reference's own callers use it
// (apps/albedo/main.cpp:85-106): build a Scene, a PerspectiveCamera and a
// RenderJob, call renderImage into a PixelBuffer. Prints the image mean and
// checks the reference's argument errors. Exit status 0 on success.
#include "albedo_dropin.hpp"

#include <cstdio>
#include <memory>

using namespace albedo;

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
    return (errors_seen == 3 && sum > 0) ? 0 : 1;
}
