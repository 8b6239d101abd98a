// Placeholder translation unit; filled in by the subsystem that owns BhState.
#include "pnb_internal.hpp"

namespace pnb {
struct BhState {};
void bh_free(BhState* s) { delete s; }
}  // namespace pnb
