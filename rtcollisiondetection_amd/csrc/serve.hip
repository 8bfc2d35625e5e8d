// Serving loop (continuous batching over resident scenes).
#include "ctx.hpp"

namespace rtcd {
void serve_shutdown(SweepBuffers&) {}
}  // namespace rtcd
