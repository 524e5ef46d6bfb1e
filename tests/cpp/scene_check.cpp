// Reads a scene with the drop-in readScene and prints what it found, or the
// error, in a line format the Python tests parse. No GPU work.
//   scene_check FILE                 read FILE (meshes resolved beside it)
//   scene_check - [BASE] [KEY=VALUE]  read stdin; BASE enables meshes; limits
#include "scene_file.hpp"

#include <cstdio>
#include <iostream>
#include <sstream>

using namespace albedo;

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
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
