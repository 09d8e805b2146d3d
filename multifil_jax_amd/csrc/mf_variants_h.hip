// Kernel variants, group h: 32 node slots and 16 heads per thread at 1024
// threads (10x10), placement level 2, fixed spacing only.  The
// dynamic-spacing forms are not built: their one run (96 instances of
// 10x10) stopped responding and the cause has not been found, so such a
// launch is refused on the host (pick) instead.
#include "mf_kernel.h"

namespace mf {
const Variant* variants_h(int& n) {
  static const Variant v[] = {
      make_variant<1024, 32, 16, false, false, 2>(), make_variant<1024, 32, 16, true, false, 2>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
