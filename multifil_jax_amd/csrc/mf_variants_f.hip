// Kernel variants, group f: lattices above 6x6 (8x8, 10x10 and the larger
// flight-muscle lattices), with the positions, per-head site / state and TM
// states in the device workspace too (placement level 2).
#include "mf_kernel.h"

namespace mf {
const Variant* variants_f(int& n) {
  static const Variant v[] = {
      make_variant<1024, 16, 8, false, false, 2>(), make_variant<1024, 16, 8, true, false, 2>(),
      make_variant<1024, 16, 8, false, true, 2>(),  make_variant<1024, 16, 8, true, true, 2>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
