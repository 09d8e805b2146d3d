// Kernel variants, group g: 24 node slots and 10 heads per thread at 1024
// threads (8x8, Drosophila 4x6), placement level 2.
#include "mf_kernel.h"

namespace mf {
const Variant* variants_g(int& n) {
  static const Variant v[] = {
      make_variant<1024, 24, 10, false, false, 2>(), make_variant<1024, 24, 10, true, false, 2>(),
      make_variant<1024, 24, 10, false, true, 2>(),  make_variant<1024, 24, 10, true, true, 2>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
