// Kernel variants, group a (the host concatenates the groups a..e in
// order and picks the first variant that fits a launch).
#include "mf_kernel.h"

namespace mf {
const Variant* variants_a(int& n) {
  static const Variant v[] = {
      make_variant<1024, 2, 1, false>(),
      make_variant<1024, 2, 1, true>(),
      make_variant<768, 2, 1, false>(),
      make_variant<768, 2, 1, true>(),
      make_variant<768, 2, 1, false, true>(),
      make_variant<768, 2, 1, true, true>(),
      make_variant<1024, 2, 1, false, true>(),
      make_variant<1024, 2, 1, true, true>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
