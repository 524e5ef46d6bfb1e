// Device data layouts shared by the host builder and the HIP kernels.
//
// HBM layout of one scene (all arrays 16-byte aligned, structure-of-arrays where
// a kernel streams them):
//   nodes     BVH2 nodes, 64 B each. A node stores the boxes of BOTH children, so
//             one 64-byte fetch decides both (Aila & Laine 2009 layout); TLAS and
//             every BLAS live in one array.
//   tris      pre-gathered triangles in BLAS leaf order, 48 B (3 x float4):
//             p0.xyz + kept-triangle index, p1.xyz, p2.xyz. FP32.
//   leafobj   TLAS leaf entries -> object index.
//   objects   per object: kind, material, light, BLAS root, FP32 3x4 maps.
//   spheres   per object: FP64 center and radius (double4).
//   shading   per triangle (kept order): corner normals and uv (smooth meshes).
#pragma once

#include <cstdint>

namespace ab {

// Child reference: >= 0 is an interior node index; < 0 is a leaf encoded as
// ~((first << 3) | count) with count in 1..4 (leaves hold at most 4 primitives,
// reference accel/bvh.cpp:33). kNoChild marks an absent second child.
constexpr int32_t kNoChild = INT32_MIN;
constexpr int kLeafCountBits = 3;

inline constexpr int32_t encodeLeaf(uint32_t first, uint32_t count)
{
    return ~int32_t((first << kLeafCountBits) | count);
}

struct alignas(16) DevNode {
    // child 0 box: lo = (a[0], a[2], b[0]), hi = (a[1], a[3], b[1])
    // child 1 box: lo = (b[2], c[0], c[2]), hi = (b[3], c[1], c[3])
    float a[4];
    float b[4];
    float c[4];
    int32_t child[2];
    int32_t pad[2];
};
static_assert(sizeof(DevNode) == 64, "node must be one 64-byte fetch");

struct alignas(16) DevTri {
    float p0[3];
    uint32_t prim; // index among the mesh's kept triangles
    float p1[3];
    uint32_t pad1;
    float p2[3];
    uint32_t pad2;
};
static_assert(sizeof(DevTri) == 48, "triangle must be three float4");

enum DevObjectKind : uint32_t { kObjSphere = 0, kObjInstance = 1 };

struct alignas(16) DevObject {
    uint32_t kind;
    uint32_t material;
    int32_t light;
    int32_t blas_root;   // child reference of the mesh's BLAS root (instances)
    uint32_t mesh;
    uint32_t tri_shading_base; // first DevTriShading of the mesh
    uint32_t smooth;
    uint32_t tri_count;  // kept triangles of the mesh (instances): the bound of a hit's prim id
    float to_object[12]; // row-major 3x4, FP32
    float to_world[12];  // row-major 3x4, FP32
    int32_t wide_root;   // child reference of the mesh's four-wide BLAS root (instances), or kNoChild
    uint32_t pad[3];
};

// Per kept triangle data needed only by shading (not by traversal).
struct alignas(16) DevTriShading {
    float p0[3], p1[3], p2[3]; // vertices (object space)
    float n0[3], n1[3], n2[3]; // corner normals (object space); face normal if flat
    float fn[3];               // unit face normal (object space)
    float uv[6];
    float pad[2];
};
static_assert(sizeof(DevTriShading) == 116 + 12, "shading record size");

struct alignas(16) DevMaterial {
    uint32_t kind;     // ab_material_kind
    uint32_t tex_kind; // diffuse texture kind
    uint32_t first, second;
    float a[4];        // texture a / conductor eta
    float b[4];        // texture b / conductor k
    float frequency;   // checker frequency / perlin scale
    float veins;
    float roughness;
    float ior;
    float mix_scale, mix_veins, mix_threshold, pad;
};

struct alignas(16) DevLight {
    uint32_t kind;
    uint32_t object;
    uint32_t pad[2];
    double position[3]; // point light (FP64 kept for the cone math's center too)
    double radius;      // area light radius
    float a[4];         // intensity / radiance / nadir
    float b[4];         // zenith
};

// Pointers of one resident scene, passed by value to kernels.
// Nodes of one scene's BVHs (TLAS and every BLAS together): their bytes must
// fit a 32-bit offset.
constexpr uint32_t kMaxNodes = 1u << 26;
static_assert(sizeof(DevNode) * uint64_t(kMaxNodes) <= (uint64_t(1) << 32), "node offsets fit 32 bits");
// Triangles of one scene: the three copies' bytes must fit a 32-bit offset.
constexpr uint32_t kMaxTriangles = uint32_t((uint64_t(1) << 32) / (3 * 48));

struct DevScene {
    const DevNode* nodes;
    // Every triangle three times, in three copies of tri_count triangles: copy c
    // holds the vertices' coordinates in the order (c + 1, c + 2, c) mod 3, the
    // sheared-space order of a ray whose direction is largest along axis c.
    const DevTri* tris;
    const uint32_t* leafobj;
    const DevObject* objects;
    const double* spheres; // 4 doubles per object
    const DevTriShading* shading;
    const DevMaterial* materials;
    const DevLight* lights;
    int32_t tlas_root;
    uint32_t num_objects; // the bound of a hit's object id
    uint32_t num_lights;
    int32_t sky_light;
    uint32_t stack_depth; // most traversal-stack entries a query holds (traceRay and RayQuery)
    uint32_t has_spheres; // any object is a sphere (else ray queries skip the FP64 test)
    uint32_t has_meshes;  // any object is a mesh instance (the path tracer then traces with lane refill)
    uint32_t tri_count;   // triangles per copy in tris
    // The same trees four wide (wide nodes are two consecutive DevNodes, children
    // 0-1 then 2-3, at an even index): the TLAS root, or kNoChild when the scene
    // has no four-wide form, and the most stack entries a four-wide query holds.
    int32_t wide_root;
    uint32_t wide_stack;
    // Scenes with both spheres and mesh instances also hold a TLAS over the
    // spheres alone and one over the instances alone (kNoChild otherwise), and
    // the most stack entries a query over each holds (the instances' with their
    // BLASes). The path tracer traces such a scene's rays in two passes: the
    // spheres one ray per lane, then the instances with lane refill, seeded with
    // the nearest sphere hit.
    int32_t sphere_root;
    int32_t mesh_root;
    uint32_t sphere_stack;
    uint32_t mesh_stack;
};

} // namespace ab
