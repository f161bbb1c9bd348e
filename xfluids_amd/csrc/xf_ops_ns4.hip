// Kernel instances for 4 species: sweeps (every direction, scheme,
// precision form and tile shape), states and chemistry, and their dispatch
// table.  One translation unit per species count so the build runs them in
// parallel.
#include "xf_context.h"

namespace xf {

Ops ops_ns4() { return Launch<4, true>::ops(); }

} // namespace xf
