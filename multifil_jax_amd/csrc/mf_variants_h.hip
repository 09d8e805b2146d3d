// Kernel variants, group h: 32 node slots and 16 heads per thread at 1024
// threads (10x10), placement level 2.
#include "mf_kernel.h"

namespace mf {
const Variant* variants_h(int& n) {
  static const Variant v[] = {
      make_variant<1024, 32, 16, false, false, 2>(), make_variant<1024, 32, 16, true, false, 2>(),
      make_variant<1024, 32, 16, false, true, 2>(),  make_variant<1024, 32, 16, true, true, 2>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
