// Placeholder translation unit; filled in by the subsystem that owns MlpState.
#include "pnb_internal.hpp"

namespace pnb {
struct MlpState {};
void mlp_free(MlpState* s) { delete s; }
}  // namespace pnb
