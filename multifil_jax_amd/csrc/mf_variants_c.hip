// Kernel variants, group c (the host concatenates the groups a..e in
// order and picks the first variant that fits a launch).
#include "mf_kernel.h"

namespace mf {
const Variant* variants_c(int& n) {
  static const Variant v[] = {
      make_variant<512, 4, 2, false>(),
      make_variant<512, 4, 2, true>(),
      make_variant<512, 8, 8, false>(),
      make_variant<512, 8, 8, true>(),
      make_variant<256, 8, 4, false>(),
      make_variant<256, 8, 4, true>(),
      make_variant<256, 16, 16, false>(),
      make_variant<256, 16, 16, true>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
