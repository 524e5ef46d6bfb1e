// albedo <scene> <output.ppm> [width]: the reference's command-line renderer
// (apps/albedo/main.cpp) on the MI355X library: reads the scene file, renders
// width x round(width * 9 / 16) pixels at 16 x 16 samples per pixel with the
// default path estimator, writes a binary PPM.
#include "scene_file.hpp"

#include <charconv>
#include <cstdlib>
#include <iostream>

namespace {

constexpr int kDefaultWidth = 960, kMinWidth = 16, kMaxWidth = 8192, kSamplesPerAxis = 16;
constexpr const char* kUsage = "usage: albedo <scene> <output.ppm> [width]\n"
                               "  scene: a scene file, such as assets/showcase.scene\n"
                               "  width: image width in pixels, 16 to 8192 (default 960)\n";

std::optional<int> parseWidth(std::string_view text)
{
    int w{};
    const auto [end, ec] = std::from_chars(text.data(), text.data() + text.size(), w);
    if (ec != std::errc{} || end != text.data() + text.size() || w < kMinWidth || w > kMaxWidth) return std::nullopt;
    return w;
}

void render(const std::filesystem::path& scene_path, const std::filesystem::path& out, int width)
{
    std::error_code ec;
    if (std::filesystem::equivalent(scene_path, out, ec))
        throw std::runtime_error(out.string() + ": the output would overwrite the scene file");
    const int height = std::max(1, int(std::lround(width / (16.0 / 9.0))));
    albedo::SceneDescription d = [&] {
        try {
            return albedo::readScene(scene_path);
        } catch (const albedo::SceneParseError& e) {
            throw std::runtime_error(scene_path.string() + ": " + e.what());
        }
    }();
    d.camera.aspect_ratio = double(width) / height;
    std::optional<albedo::PerspectiveCamera> camera;
    try {
        camera.emplace(d.camera);
    } catch (const std::invalid_argument& e) {
        throw std::runtime_error(scene_path.string() + ": the camera cannot render this image: " + e.what());
    }
    albedo::PixelBuffer image{width, height};
    albedo::renderImage(albedo::RenderJob{*camera, d.scene, kSamplesPerAxis, 0, albedo::PathEstimator{},
                                          albedo::availableThreadCount()},
                        image);
    albedo::writePpm(out, image);
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 3 || argc > 4) {
        std::cerr << kUsage;
        return EXIT_FAILURE;
    }
    const std::optional<int> width = argc == 4 ? parseWidth(argv[3]) : std::optional<int>(kDefaultWidth);
    if (!width) {
        std::cerr << "albedo: invalid width '" << argv[3] << "'\n" << kUsage;
        return EXIT_FAILURE;
    }
    try {
        render(argv[1], argv[2], *width);
        std::cout << "albedo: wrote " << argv[2] << "\n";
    } catch (const std::exception& e) {
        std::cerr << "albedo: " << e.what() << "\n";
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
