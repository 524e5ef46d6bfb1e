// The host scene build (albedo_amd/csrc/scene_build.cpp) on the CPU, meant to be
// compiled with AddressSanitizer and UBSan: two meshes (smooth with uv, flat
// without; a zero-area face), spheres (one of radius 1e5), scaled and moved
// instances, the three light kinds; every valid build flag combination; then
// descriptions the build must reject. Prints one line per build and "ok" when
// every valid one builds and every invalid one is rejected.
#include "scene_build.hpp"
#include <cmath>
#include <cstdio>
#include <vector>
using namespace ab;
int main()
{
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    std::vector<float> uv;
    const int n = 40;
    for (int j = 0; j <= n; ++j)
        for (int i = 0; i <= n; ++i) {
            pos.push_back(float(i) / n);
            pos.push_back(0.2f * std::sin(6.0f * i / n) * std::cos(5.0f * j / n));
            pos.push_back(float(j) / n);
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            uint32_t a = j * (n + 1) + i, b = a + 1, c = a + n + 1, d = c + 1;
            uint32_t f[6] = {a, b, d, a, d, c};
            for (uint32_t k : f) idx.push_back(k);
            for (int k = 0; k < 12; ++k) uv.push_back(0.1f * k);
        }
    idx.push_back(0); idx.push_back(0); idx.push_back(1); // a zero-area face
    for (int k = 0; k < 6; ++k) uv.push_back(0.0f);
    const uint32_t nf = uint32_t(idx.size() / 3);
    ab_mesh_desc meshes[2] = {{pos.data(), uint32_t(pos.size() / 3), idx.data(), nf, uv.data(), 1},
                              {pos.data(), uint32_t(pos.size() / 3), idx.data(), nf, nullptr, 0}};
    std::vector<ab_object_desc> objs;
    for (int k = 0; k < 12; ++k) {
        ab_object_desc o{};
        o.material = k % 3;
        o.light = -1;
        if (k % 3 == 0) {
            o.kind = AB_OBJ_SPHERE;
            o.center[0] = k; o.center[1] = 1; o.center[2] = -k;
            o.radius = k == 0 ? 1e5 : 0.5;
        } else {
            o.kind = AB_OBJ_INSTANCE;
            o.mesh = k % 2;
            const double s = 1.0 + 0.1 * k;
            double fw[12] = {s, 0, 0, double(k), 0, s, 0, 0, 0, 0, s, double(-k)};
            double inv[12] = {1 / s, 0, 0, -k / s, 0, 1 / s, 0, 0, 0, 0, 1 / s, k / s};
            for (int q = 0; q < 12; ++q) { o.to_world[q] = fw[q]; o.to_object[q] = inv[q]; }
        }
        objs.push_back(o);
    }
    objs[3].light = 0;
    ab_material_desc mats[3] = {};
    mats[1].kind = AB_MAT_CONDUCTOR;
    mats[1].eta[0] = mats[1].eta[1] = mats[1].eta[2] = 0.2f;
    mats[1].k[0] = mats[1].k[1] = mats[1].k[2] = 3.0f;
    mats[2].kind = AB_MAT_DIELECTRIC;
    mats[2].ior = 1.5f;
    ab_light_desc lights[3] = {};
    lights[0].kind = AB_LIGHT_AREA; lights[0].object = 3;
    lights[1].kind = AB_LIGHT_POINT; lights[1].position[1] = 5;
    lights[2].kind = AB_LIGHT_SKY;
    int failures = 0;
    for (uint32_t flags : {0u, 1u, 2u, 4u, 6u, 5u}) {
        ab_scene_desc d{meshes, 2, objs.data(), uint32_t(objs.size()), mats, 3, lights, 3, flags};
        HostScene hs;
        std::string err;
        const ab_status st = buildHostScene(d, hs, err);
        std::printf("flags %u: status %d nodes %zu tris %zu depth %u/%u split %d/%d %s\n", flags, int(st),
                    hs.nodes.size(), hs.tris.size(), hs.tlas_depth, hs.max_blas_depth, hs.sphere_root, hs.mesh_root,
                    err.c_str());
        if (st != AB_OK) ++failures;
    }
    // Invalid: a face past the positions, a wide + device combination, an unknown flag.
    std::vector<uint32_t> bad = idx;
    bad[4] = 1u << 20;
    ab_mesh_desc badm[1] = {{pos.data(), uint32_t(pos.size() / 3), bad.data(), nf, nullptr, 0}};
    ab_object_desc inst = objs[1];
    inst.mesh = 0;
    for (uint32_t flags : {0u, 3u, 8u}) {
        ab_scene_desc d{flags ? meshes : badm, 1, &inst, 1, mats, 3, lights + 1, 1, flags};
        HostScene hs;
        std::string err;
        const ab_status st = buildHostScene(d, hs, err);
        std::printf("invalid (flags %u): status %d: %s\n", flags, int(st), err.c_str());
        if (st == AB_OK) ++failures;
    }
    std::printf("%s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
