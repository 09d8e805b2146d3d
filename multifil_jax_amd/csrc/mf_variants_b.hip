// Kernel variants, group b (the host concatenates the groups a..e in
// order and picks the first variant that fits a launch).
#include "mf_kernel.h"

namespace mf {
const Variant* variants_b(int& n) {
  static const Variant v[] = {
      make_variant<1024, 5, 3, false, true>(),
      make_variant<1024, 5, 3, true, true>(),
      make_variant<1024, 8, 4, false, true>(),
      make_variant<1024, 8, 4, true, true>(),
      make_variant<1024, 5, 3, false>(),
      make_variant<1024, 5, 3, true>(),
      make_variant<1024, 8, 4, false>(),
      make_variant<1024, 8, 4, true>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
