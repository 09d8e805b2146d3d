// Kernel variants, group e (the host concatenates the groups a..e in
// order and picks the first variant that fits a launch).
#include "mf_kernel.h"

namespace mf {
const Variant* variants_e(int& n) {
  static const Variant v[] = {
      make_variant<1024, 16, 8, false, false, true>(),
      make_variant<1024, 16, 8, true, false, true>(),
      make_variant<1024, 16, 8, false, true, true>(),
      make_variant<1024, 16, 8, true, true, true>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
