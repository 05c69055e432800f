// Approximate minimum degree ordering (Ordering::AMD, src/Ordering.cpp:411-525
// -> amdbar.F).  Not restated yet: requesting it raises the same logic_error
// an unsupported ordering string raises in the reference.
#include "analysis.hpp"

namespace spk {

void order_amd(const Graph& g, Ordering& ord) {
  (void)g;
  (void)ord;
  throw std::logic_error("AMD ordering is not available in this build");
}

}  // namespace spk
