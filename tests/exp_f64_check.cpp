// Host-side check of mf::exp_f64 and mf::atan2_pos_f64 (the step kernel's
// double exp and crossbridge angle) against the C library: for exp the
// largest ulp distance over 2e6 pseudo-random points of [-708, 709.7] and
// [-100, 100] (normal results), the largest absolute error over the
// subnormal range, and the special values; for atan2 the largest ulp
// distance over 2e6 pairs y in [0.5, 40], x in [-60, 60] and the axis cases.
// Prints two lines:
// worst_ulp x_at_worst subnormal_abs_err inf_at_710 nan_in_nan zero_below_745 one_at_zero
// atan2_worst_ulp y_at_worst x_at_worst exact_axis_cases
#include "mf_common.h"
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

static int64_t ulps(double a, double b) {
  int64_t ia, ib;
  std::memcpy(&ia, &a, 8);
  std::memcpy(&ib, &b, 8);
  if (ia < 0) ia = INT64_MIN - ia;
  if (ib < 0) ib = INT64_MIN - ib;
  return ia > ib ? ia - ib : ib - ia;
}

int main() {
  int64_t worst = 0;
  double wx = 0.0;
  uint64_t s = 88172645463325252ull;  // xorshift64
  for (int i = 0; i < 2000000; ++i) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    const double u = (double)(s >> 11) / 9007199254740992.0;
    const double x = i < 1000000 ? -708.0 + u * (709.7 + 708.0) : -100.0 + 200.0 * u;
    const int64_t d = ulps(std::exp(x), mf::exp_f64(x));
    if (d > worst) { worst = d; wx = x; }
  }
  double sub = 0.0;
  for (double x = -745.2; x < -708.0; x += 0.001) sub = std::fmax(sub, std::fabs(std::exp(x) - mf::exp_f64(x)));
  std::printf("%lld %.17g %.6g %d %d %d %d\n", (long long)worst, wx, sub, (int)std::isinf(mf::exp_f64(710.0)),
              (int)std::isnan(mf::exp_f64(NAN)), (int)(mf::exp_f64(-800.0) == 0.0), (int)(mf::exp_f64(0.0) == 1.0));
  int64_t aw = 0;
  double ay = 0.0, axw = 0.0;
  for (int i = 0; i < 2000000; ++i) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    const double u = (double)(s >> 11) / 9007199254740992.0;
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    const double v = (double)(s >> 11) / 9007199254740992.0;
    const double y = 0.5 + 39.5 * u, x = -60.0 + 120.0 * v;
    const int64_t d = ulps(std::atan2(y, x), mf::atan2_pos_f64(y, x));
    if (d > aw) { aw = d; ay = y; axw = x; }
  }
  const int axis = (mf::atan2_pos_f64(14.0, 0.0) == std::atan2(14.0, 0.0)) +
                   (mf::atan2_pos_f64(14.0, 14.0) == std::atan2(14.0, 14.0)) +
                   (mf::atan2_pos_f64(14.0, -14.0) == std::atan2(14.0, -14.0));
  std::printf("%lld %.17g %.17g %d\n", (long long)aw, ay, axw, axis);
  return 0;
}
