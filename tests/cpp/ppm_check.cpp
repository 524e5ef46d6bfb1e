// Output-format checks of the drop-in encodeSrgb / toByte / writePpm against the
// reference's cases (tests/color/srgb_test.cpp:20-80, tests/image/ppm_test.cpp:20-110).
// Prints one line per failed check and "checks N failures F". No GPU work.
#include "scene_file.hpp"

#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>

using namespace albedo;

static int checks = 0, failures = 0;

static void expect(bool ok, const char* what)
{
    ++checks;
    if (!ok) {
        ++failures;
        std::printf("FAIL %s\n", what);
    }
}

static void setPixel(PixelBuffer& image, int x, int y, float r, float g, float b)
{
    image.addSample({x, y}, Rgb{r, g, b});
}

static PixelBuffer twoPixelImage()
{
    PixelBuffer image{2, 1};
    setPixel(image, 0, 0, 1.0f, 0.0f, 0.0f);
    setPixel(image, 1, 0, 0.5f, 0.5f, 0.5f);
    return image;
}

static std::string readBinary(const std::filesystem::path& path)
{
    std::ifstream file{path, std::ios::binary};
    return {std::istreambuf_iterator<char>{file}, std::istreambuf_iterator<char>{}};
}

int main(int argc, char** argv)
{
    const double nan = std::numeric_limits<double>::quiet_NaN(), inf = std::numeric_limits<double>::infinity();
    // sRGB transfer function and byte quantisation.
    expect(std::abs(encodeSrgb(0.0)) < 1e-12 && std::abs(encodeSrgb(1.0) - 1.0) < 1e-12, "end points");
    expect(std::abs(encodeSrgb(0.001) - 0.01292) < 1e-12, "linear segment");
    expect(std::abs(encodeSrgb(0.0031308) - 0.04045) < 1e-6, "breakpoint, linear side");
    expect(std::abs(encodeSrgb(0.0031308 * (1 + 1e-9)) - 0.04045) < 1e-6, "breakpoint, power side");
    expect(std::abs(encodeSrgb(0.5) - 0.735357) < 1e-6, "mid grey");
    expect(std::abs(encodeSrgb(-0.5) - 12.92 * -0.5) < 1e-12, "negative on the linear segment");
    expect(encodeSrgb(2.0) > 1.0, "above one stays above one");
    expect(toByte(encodeSrgb(nan)) == 0, "NaN is byte 0");
    expect(toByte(0.0) == 0 && toByte(1.0) == 255, "unit range to full bytes");
    expect(toByte(encodeSrgb(0.5)) == 188 && toByte(0.5) == 128, "rounds to nearest");
    expect(toByte(-0.5) == 0 && toByte(7.0) == 255 && toByte(inf) == 255 && toByte(-inf) == 0, "clamps");
    // PPM bytes.
    const std::string two_pixel_ppm{"P6\n2 1\n255\n\xFF\x00\x00\xBC\xBC\xBC", 11 + 6};
    {
        std::ostringstream out{std::ios::binary};
        writePpm(out, twoPixelImage());
        expect(out.str() == two_pixel_ppm, "header and pixel bytes");
    }
    {
        PixelBuffer image{1, 2};
        setPixel(image, 0, 0, 1.0f, 1.0f, 1.0f);
        std::ostringstream out{std::ios::binary};
        writePpm(out, image);
        expect(out.str() == std::string("P6\n1 2\n255\n\xFF\xFF\xFF\x00\x00\x00", 11 + 6), "rows top to bottom");
    }
    {
        PixelBuffer image{1, 1};
        setPixel(image, 0, 0, 4.0f, 1.0f, -2.0f);
        std::ostringstream out{std::ios::binary};
        writePpm(out, image);
        expect(out.str() == std::string("P6\n1 1\n255\n\xFF\xFF\x00", 11 + 3), "clamps brighter than white");
    }
    {
        std::ostringstream out;
        out.setstate(std::ios::badbit);
        bool threw = false;
        try {
            writePpm(out, twoPixelImage());
        } catch (const std::runtime_error&) {
            threw = true;
        }
        expect(threw, "failed stream throws");
    }
    if (argc > 1) {
        const std::filesystem::path dir{argv[1]};
        const auto file = dir / "out.ppm";
        {
            std::ofstream existing{file, std::ios::binary};
            existing << std::string(64, 'x');
        }
        writePpm(file, twoPixelImage());
        expect(readBinary(file) == two_pixel_ppm, "overwrites a longer file with identical bytes");
        bool threw = false;
        std::string message;
        try {
            writePpm(dir / "missing" / "image.ppm", twoPixelImage());
        } catch (const std::runtime_error& e) {
            threw = true;
            message = e.what();
        }
        expect(threw && message.find("image.ppm") != std::string::npos, "open failure names the path");
    }
    std::printf("checks %d failures %d\n", checks, failures);
    return failures == 0 ? 0 : 1;
}
