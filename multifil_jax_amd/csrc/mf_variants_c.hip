// Kernel variants, group c (the host concatenates the groups a..h in
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
      // dynamic spacing at 512 threads: 2 waves per SIMD, 256 VGPRs; a 2x2
      // dynamic step takes 155 us here against 178 at 768 threads.  The 2x2
      // lattice's 12 chains need 3 node slots per lane at 8 waves (a thick
      // and a thin chain on waves 0-3, one thin chain on the others).
      make_variant<512, 3, 2, false, true>(),
      make_variant<512, 3, 2, true, true>(),
      make_variant<512, 4, 2, false, true>(),
      make_variant<512, 4, 2, true, true>(),
  };
  n = (int)(sizeof(v) / sizeof(v[0]));
  return v;
}
}  // namespace mf
