// Serving loop (filled in with the batched engine).
#include "context.hpp"

namespace ab {
struct Server {};
} // namespace ab

ab_context::~ab_context()
{
    delete server;
    for (auto& kv : resident) {
        (void)hipFree(kv.second->arena);
        delete kv.second;
    }
    if (stream) (void)hipStreamDestroy(stream);
}
