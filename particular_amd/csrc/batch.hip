// Placeholder translation unit; filled in by the subsystem that owns BatchState.
#include "pnb_internal.hpp"

namespace pnb {
struct BatchState {};
void batch_free(BatchState* s) { delete s; }
}  // namespace pnb
