// Serving loop (filled in with the batched engine).
#include "context.hpp"
#include "engine.hpp"

namespace ab {
struct Server {};
} // namespace ab

ab_context::~ab_context()
{
    delete server;
    delete engine;
    for (auto& kv : resident) {
        (void)hipFree(kv.second->arena);
        delete kv.second;
    }
    if (stream) (void)hipStreamDestroy(stream);
}
