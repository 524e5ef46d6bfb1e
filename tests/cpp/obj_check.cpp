// Reads an OBJ mesh with the drop-in readObj and prints what it found, or the
// error, in a line format tests/test_obj.py parses. No GPU work.
//   obj_check [KEY=VALUE...]   read stdin; limits max_line_length, max_vertices,
//                              max_texture_coordinates, max_triangles
//   obj_check --failed         a stream already in a failed state
//   obj_check --read-error     a stream whose buffer fails after its text
//   obj_check --huge-limit     a line length limit no buffer can hold
//   obj_check --file PATH      read a file by path
#include "scene_file.hpp"

#include <cstdio>
#include <iostream>
#include <limits>
#include <sstream>

using namespace albedo;

namespace {

// Serves its text, then fails the next read as a disk error would.
class FailingBuffer : public std::streambuf {
public:
    explicit FailingBuffer(std::string text) : text_(std::move(text)) { setg(&text_[0], &text_[0], &text_[0] + text_.size()); }

protected:
    int_type underflow() override { throw std::runtime_error("read failure"); }

private:
    std::string text_;
};

// Faces whose triangle has a representable area: the ones a mesh geometry keeps
// (reference geometry/triangle_mesh.cpp:76-90).
std::size_t facesWithArea(const TriangleMesh& m)
{
    std::size_t n = 0;
    for (const auto& f : m.faces()) {
        const Point3r &a = m.positions()[f[0]], &b = m.positions()[f[1]], &c = m.positions()[f[2]];
        const Real e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, e2[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
        const Real x = e1[1] * e2[2] - e1[2] * e2[1], y = e1[2] * e2[0] - e1[0] * e2[2], z = e1[0] * e2[1] - e1[1] * e2[0];
        n += std::isnormal(x * x + y * y + z * z) ? 1 : 0;
    }
    return n;
}

void print(const TriangleMesh& m)
{
    std::printf("ok positions=%zu faces=%zu uv=%zu uvfaces=%zu triangles=%zu", m.positions().size(), m.faces().size(),
                m.textureCoordinates().size(), m.textureFaces().size(), facesWithArea(m));
    std::printf(" f=");
    for (const auto& f : m.faces()) std::printf("%zu,%zu,%zu;", std::size_t(f[0]), std::size_t(f[1]), std::size_t(f[2]));
    std::printf(" tf=");
    for (const auto& f : m.textureFaces())
        std::printf("%zu,%zu,%zu;", std::size_t(f[0]), std::size_t(f[1]), std::size_t(f[2]));
    std::printf(" p=");
    for (const auto& p : m.positions()) std::printf("%.17g,%.17g,%.17g;", p.x, p.y, p.z);
    std::printf(" t=");
    for (const auto& t : m.textureCoordinates()) std::printf("%.17g,%.17g;", t.x, t.y);
    std::printf("\n");
}

} // namespace

int main(int argc, char** argv)
{
    ObjLimits limits;
    std::string mode, file;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a.rfind("--", 0) == 0) {
            mode = a;
            if (a == "--file" && i + 1 < argc) file = argv[++i];
            continue;
        }
        const auto eq = a.find('=');
        const std::string k = a.substr(0, eq);
        const std::size_t v = std::stoull(a.substr(eq + 1));
        if (k == "max_line_length") limits.max_line_length = v;
        else if (k == "max_vertices") limits.max_vertices = v;
        else if (k == "max_texture_coordinates") limits.max_texture_coordinates = v;
        else if (k == "max_triangles") limits.max_triangles = v;
    }
    std::stringstream buf;
    if (mode != "--file") buf << std::cin.rdbuf();
    try {
        if (mode == "--file") {
            print(readObj(std::filesystem::path(file), limits));
        } else if (mode == "--failed") {
            std::istringstream in(buf.str());
            in.setstate(std::ios::failbit);
            print(readObj(in, limits));
        } else if (mode == "--read-error") {
            FailingBuffer fb(buf.str());
            std::istream in(&fb);
            print(readObj(in, limits));
        } else if (mode == "--huge-limit") {
            std::istringstream in(buf.str());
            limits.max_line_length = std::numeric_limits<std::size_t>::max();
            print(readObj(in, limits));
        } else {
            std::istringstream in(buf.str());
            print(readObj(in, limits));
        }
    } catch (const ObjParseError& e) {
        std::printf("error line=%zu %s\n", e.line(), e.what());
        return 1;
    } catch (const std::invalid_argument& e) {
        std::printf("invalid %s\n", e.what());
        return 1;
    } catch (const std::exception& e) {
        std::printf("failure %s\n", e.what());
        return 1;
    }
    return 0;
}
