// Reads a scene with the drop-in readScene and prints what it found, or the
// error, in a line format the Python tests parse. No GPU work, except in the
// render form.
//   scene_check FILE                 read FILE (meshes resolved beside it)
//   scene_check - [BASE] [KEY=VALUE]  read stdin; BASE enables meshes; limits
//   scene_check FILE --render W H SPA DEPTH SEED OUT
//                                    also render it with renderImage and
//                                    PathEstimator{DEPTH, 3} at the image's aspect
//                                    ratio, writing the per-pixel means (FP32 RGB)
//                                    to OUT
#include "scene_file.hpp"

#include <cstdio>
#include <iostream>
#include <sstream>

using namespace albedo;

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    if (argc == 9 && std::string(argv[2]) == "--render") {
        try {
            SceneDescription d = readScene(std::filesystem::path(argv[1]));
            const int w = std::atoi(argv[3]), h = std::atoi(argv[4]), spa = std::atoi(argv[5]);
            d.camera.aspect_ratio = double(w) / h;
            const PerspectiveCamera camera{d.camera};
            PixelBuffer image{w, h};
            renderImage(RenderJob{camera, d.scene, spa, std::strtoull(argv[7], nullptr, 10),
                                  PathEstimator{std::atoi(argv[6]), 3}},
                        image);
            std::vector<float> out(std::size_t(w) * h * 3);
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    const Rgb c = image.average({x, y});
                    float* o = &out[(std::size_t(y) * w + x) * 3];
                    o[0] = float(c.r);
                    o[1] = float(c.g);
                    o[2] = float(c.b);
                }
            FILE* f = std::fopen(argv[8], "wb");
            if (!f || std::fwrite(out.data(), sizeof(float), out.size(), f) != out.size()) return 1;
            std::fclose(f);
            std::printf("rendered objects=%zu\n", d.scene.shapeCount());
        } catch (const std::exception& e) {
            std::printf("failure %s\n", e.what());
            return 1;
        }
        return 0;
    }
    SceneLimits limits;
    std::optional<std::filesystem::path> base;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        const auto eq = a.find('=');
        if (eq == std::string::npos) {
            base = a;
            continue;
        }
        const std::string k = a.substr(0, eq);
        const std::size_t v = std::stoull(a.substr(eq + 1));
        if (k == "max_line_length") limits.max_line_length = v;
        else if (k == "max_objects") limits.max_objects = v;
        else if (k == "max_materials") limits.max_materials = v;
        else if (k == "max_meshes") limits.max_meshes = v;
        else if (k == "max_instance_operations") limits.max_instance_operations = v;
        else if (k == "max_ground_points") limits.max_ground_points = v;
        else if (k == "max_vertices") limits.mesh_limits.max_vertices = v;
        else if (k == "max_triangles") limits.mesh_limits.max_triangles = v;
    }
    try {
        std::optional<SceneDescription> d;
        if (std::string(argv[1]) == "-") {
            std::stringstream buf;
            buf << std::cin.rdbuf();
            std::istringstream in(buf.str());
            if (base) d.emplace(readScene(in, *base, limits));
            else d.emplace(readScene(in, limits));
        } else {
            d.emplace(readScene(std::filesystem::path(argv[1]), limits));
        }
        const CameraSettings& c = d->camera;
        std::printf("ok objects=%zu from=%.17g,%.17g,%.17g at=%.17g,%.17g,%.17g up=%.17g,%.17g,%.17g fov=%.17g\n",
                    d->scene.shapeCount(), c.look_from.x, c.look_from.y, c.look_from.z, c.look_at.x, c.look_at.y,
                    c.look_at.z, c.up.x, c.up.y, c.up.z, c.vertical_fov_degrees);
    } catch (const SceneParseError& e) {
        std::printf("error line=%zu %s\n", e.line(), e.what());
        return 1;
    } catch (const std::exception& e) {
        std::printf("failure %s\n", e.what());
        return 1;
    }
    return 0;
}
