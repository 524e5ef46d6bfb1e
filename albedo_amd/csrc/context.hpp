// Library-internal state behind the C ABI handles.
#pragma once

#include "../../include/albedo_gpu.h"
#include "layout.hpp"
#include "scene_build.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>

struct ab_scene {
    ab_context* ctx{nullptr};
    std::atomic<int> refs{1};
    uint64_t hash{0};
    void* arena{nullptr};
    uint64_t bytes{0};
    ab::DevScene dev{};
    uint32_t tlas_depth{0}, blas_depth{0};
    uint64_t tlas_nodes{0}, blas_nodes{0}, triangles{0};
    uint32_t num_lights{0};
};

namespace ab {
struct Server;
class Engine;
}

struct ab_context {
    int device{0};
    hipStream_t stream{nullptr};
    std::mutex mu;                           // guards the scene cache and config
    std::map<uint64_t, ab_scene*> resident;  // scene cache keyed by content hash
    ab_launch_config config{256, 1u << 21, 0, 1};
    std::chrono::steady_clock::time_point t0{std::chrono::steady_clock::now()};
    ab::Server* server{nullptr};             // serving loop, when started
    std::mutex render_mu;                    // serialises ab_render / ab_render_batch
    ab::Engine* engine{nullptr};             // engine of ab_render / ab_render_batch
    ~ab_context();
};

namespace ab {

// Records `msg` as the calling thread's last error and returns `code`.
ab_status fail(ab_status code, const std::string& msg);
// Maps a HIP error to AB_DEVICE_ERROR / AB_OUT_OF_MEMORY with a message.
ab_status hipFail(hipError_t e, const char* what);
double msSince(const ab_context* ctx);
// Finds the no-allocation renderer's device array on `device` (current), once
// per device; ab_render_static reports a failure.
void staticScratchSetup(int device);

} // namespace ab
