// Host-side scene build: validation, BVH construction and flattening into the
// device layouts of layout.hpp. Runs on the CPU once per scene.
#pragma once

#include "../../include/albedo_gpu.h"
#include "layout.hpp"

#include <cstdint>
#include <string>
#include <vector>

namespace ab {

// A mesh whose BLAS the device builds (AB_BUILD_DEVICE_BLAS): its kept
// triangles sit in HostScene::tris at tri_base in kept order until then.
struct DeviceMesh {
    uint32_t mesh;
    uint32_t tri_base;
    uint32_t count;
};

// Everything the device needs for one scene, in host memory.
struct HostScene {
    std::vector<DevNode> nodes;
    std::vector<DevTri> tris;
    std::vector<uint32_t> leafobj;
    std::vector<DevObject> objects;
    std::vector<double> spheres; // 4 per object
    std::vector<DevTriShading> shading;
    std::vector<DevMaterial> materials;
    std::vector<DevLight> lights;
    int32_t tlas_root{kNoChild};
    // Four-wide form of the same trees (DevScene::wide_root): its TLAS root and
    // the most stack entries its TLAS and its deepest BLAS need.
    int32_t wide_tlas_root{kNoChild};
    // With both spheres and instances: TLASes over each kind alone (DevScene::
    // sphere_root, mesh_root) and their levels.
    int32_t sphere_root{kNoChild};
    int32_t mesh_root{kNoChild};
    uint32_t sphere_depth{0};
    uint32_t mesh_depth{0};
    uint32_t wide_tlas_stack{0};
    uint32_t wide_blas_stack{0};
    int32_t sky_light{-1};
    uint32_t tlas_depth{0};
    uint32_t max_blas_depth{0};
    uint64_t tlas_nodes{0};
    uint64_t blas_nodes{0};
    uint64_t hash{0};
    // Meshes left to the device build; their instances' blas_root is kNoChild
    // until the build places their nodes, and max_blas_depth excludes them.
    std::vector<DeviceMesh> device_meshes;
};

// Validates `desc` as the reference constructors do and builds the scene.
// Returns AB_OK, or an error code with a message in `error`.
ab_status buildHostScene(const ab_scene_desc& desc, HostScene& out, std::string& error);

// 64-bit content hash of a scene description (every byte that affects rendering).
uint64_t hashSceneDesc(const ab_scene_desc& desc);

} // namespace ab
